/*
 * nwgr_graphgen.h — synthetic graph generators and the host CSR builder.
 *
 * Host-only C ABI (no HIP dependency), built into libnwgr_graphgen.so.
 *
 * The reference ships no R-MAT or Erdos-Renyi generator: its drivers read
 * MatrixMarket or "NW" binary files (bench/common.hpp:82-101).  These
 * generators produce the synthetic inputs of the PageRank workloads; the CSR
 * builder reproduces the structure the reference builds from an edge list:
 *
 *   nwgr_csr_build   replaces  index_adjacency<1>(edge_list)
 *                              (adjacency.hpp:90-97 -> build.hpp:209-265,
 *                               fill_directed<1>), plus the optional
 *                              uniq (build.hpp:425-435) and
 *                              remove_self_loops (build.hpp:437-443),
 *                              plus build_degrees (bench/common.hpp:109-120)
 *                              for the out-degree array.
 *
 * Result layout ("in-CSR", the reference's adjacency<1>): row v lists the
 * sources u of the edges u->v, sorted ascending (the reference leaves row
 * order unspecified after an unstable sort, build.hpp:262; sorted rows are
 * one of the orders it may produce).  offsets has n+1 entries, offsets[n]==m.
 * outdeg[u] = number of edges leaving u (= number of rows u appears in).
 *
 * Random numbers: the k-th output (k = 0, 1, ...) of SplitMix64 seeded with
 * `key` is mix64(key + (k+1)*0x9e3779b97f4a7c15), Vigna's SplitMix64.  Being
 * counter based, every edge can be generated independently and in parallel,
 * and the generation can be replayed anywhere.
 *
 *   R-MAT (scale S, edge factor F): N = 2^S, F*N draws.  Draw e uses outputs
 *   e*W .. e*W+W-1, W = ceil(S/2); level l (0 = most significant bit) takes
 *   the high 32 bits of output e*W + l/2 when l is even, the low 32 bits when
 *   l is odd, as an unsigned r.  With T1 = round(a*2^32), T2 = round((a+b)*2^32),
 *   T3 = round((a+b+c)*2^32): r < T1 -> quadrant a (no bits), r < T2 -> b
 *   (destination bit), r < T3 -> c (source bit), else d (both bits).  The edge
 *   is source -> destination.  No vertex permutation.
 *
 *   Erdos-Renyi G(n, m) with replacement: draw e uses output e; source =
 *   (hi32 * n) >> 32, destination = (lo32 * n) >> 32.
 *
 *   Both then drop self-loops and duplicate edges.
 */
#ifndef NWGR_GRAPHGEN_H
#define NWGR_GRAPHGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host-resident CSR owned by the library. */
typedef struct nwgr_host_csr nwgr_host_csr;

enum {
  NWGR_CSR_DEDUP = 1,            /* drop duplicate (u, v) pairs          */
  NWGR_CSR_DROP_SELF_LOOPS = 2,  /* drop u == v                           */
};

/* Build an in-CSR from an edge list (src[i] -> dst[i]), 0 <= ids < n.
 * Returns 0 on success, non-zero (and a message via nwgr_gen_last_error)
 * on invalid input or allocation failure. nthreads <= 0 picks the
 * hardware concurrency. */
int nwgr_csr_build(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                   uint32_t flags, int nthreads, nwgr_host_csr** out);

/* R-MAT generator (see the header comment), self-loops and duplicates
 * removed. scale in [1, 31], edge_factor >= 1, (edge_factor << scale) < 2^32. */
int nwgr_gen_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed, double a, double b,
                  double c, int nthreads, nwgr_host_csr** out);

/* Erdos-Renyi G(n, m) drawn with replacement, self-loops and duplicates
 * removed. m_draws < 2^32. */
int nwgr_gen_er(uint32_t n, uint64_t m_draws, uint64_t seed, int nthreads, nwgr_host_csr** out);

/* R-MAT on a vertex count that need not be a power of two: scale
 * S = ceil(log2 n); candidate draws c = 0, 1, 2, ... (same bit recipe as
 * nwgr_gen_rmat) are accepted when both endpoints are < n, until
 * `draws` candidates were accepted; then self-loops and duplicates are
 * removed.  Single-threaded (many such graphs are generated in parallel by
 * nwgr_gen_rmat_many). */
int nwgr_gen_rmat_n(uint32_t n, uint64_t draws, uint64_t seed, double a, double b, double c,
                    nwgr_host_csr** out);

/* Many independent R-MAT graphs: graph g has ns[g] vertices, ef*ns[g]
 * accepted draws and key SplitMix64(seed) output g.  Results are stored
 * block-diagonally in ONE csr of sum(ns) vertices (graph g's vertex i is
 * global vertex vbase[g] + i, vbase = exclusive scan of ns). */
int nwgr_gen_rmat_many(uint32_t count, const uint32_t* ns, uint32_t ef, uint64_t seed, double a,
                       double b, double c, int nthreads, nwgr_host_csr** out);

uint32_t nwgr_host_csr_n(const nwgr_host_csr* g);
uint64_t nwgr_host_csr_m(const nwgr_host_csr* g);
const uint32_t* nwgr_host_csr_offsets(const nwgr_host_csr* g); /* n+1 */
const uint32_t* nwgr_host_csr_indices(const nwgr_host_csr* g); /* m   */
const uint32_t* nwgr_host_csr_outdeg(const nwgr_host_csr* g);  /* n   */
void nwgr_host_csr_free(nwgr_host_csr* g);

/* Transpose an in-CSR into the out-CSR (row u lists destinations v of
 * u->v, ascending) and the permutation perm[k] = position in the in-CSR
 * of the k-th out-CSR entry.  Arrays are caller-allocated: out_offsets
 * n+1, out_indices m, perm m. */
int nwgr_csr_transpose(uint32_t n, uint64_t m, const uint32_t* offsets, const uint32_t* indices,
                       uint32_t* out_offsets, uint32_t* out_indices, uint32_t* perm, int nthreads);

/* SplitMix64 output k for key (exposed for tests). */
uint64_t nwgr_splitmix64_at(uint64_t key, uint64_t k);

const char* nwgr_gen_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NWGR_GRAPHGEN_H */
