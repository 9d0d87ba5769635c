/*
 * nwgr_io.h -- graph ingestion for the PageRank path (host C ABI, built into
 * libnwgr_graphgen.so; no HIP dependency).
 *
 * Reads the inputs the reference's drivers read (bench/common.hpp:82-101,
 * load_graph): MatrixMarket coordinate files (read_mm, io/mmio.hpp:171-217,
 * mm_fill :95-128) and the "NW" binary edge list (index_edge_list::serialize
 * / deserialize, edge_list.hpp:190-233), plus the "NW" binary adjacency
 * (indexed_struct_of_arrays::serialize / deserialize,
 * containers/compressed.hpp:303-339).  The MatrixMarket reader splits the
 * entry section into per-thread ranges at line boundaries and parses them in
 * parallel (the reference reads it line by line with a stringstream, which
 * its driver calls "slow", common.hpp:94).
 *
 * Semantics kept from the reference:
 *  - entry "i j [v]" (1-based) is the edge (i-1) -> (j-1); in a "symmetric"
 *    file an off-diagonal entry also gives (j-1) -> (i-1), right after it
 *    (mm_fill for directed edge lists, mmio.hpp:95-128);
 *  - the vertex count is max(n0 of the size line, largest id + 1)
 *    (edge_list(n0) + push_back, edge_list.hpp:143-154);
 *  - values are read for non-pattern files and returned as doubles (the
 *    weighted solve may use them); a pattern file has none.
 * Binary formats are byte-for-byte those of the reference on a little-endian
 * LP64 host with uint32 vertex ids (util/defaults.hpp:23-24): magic string
 * with its terminating NUL, directedness as a 4-byte enum, size_t counts.
 */
#ifndef NWGR_IO_H
#define NWGR_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host edge list owned by the library. */
typedef struct nwgr_edge_list nwgr_edge_list;

enum { NWGR_UNDIRECTED = 0, NWGR_DIRECTED = 1 }; /* the reference's directedness enum */

/* MatrixMarket coordinate file -> directed edge list (as read_mm<directed>).
 * Supported headers: "%%MatrixMarket matrix coordinate {pattern|real|integer|double}
 * {general|symmetric}".  nthreads <= 0 picks the hardware concurrency.
 * Returns 0, or non-zero with a message from nwgr_gen_last_error(). */
int nwgr_read_mm(const char* path, int nthreads, nwgr_edge_list** out);

/* Same with the reference's other reader forms: directedness
 * NWGR_UNDIRECTED reads every entry once (read_mm<undirected>, mm_fill
 * :130-170), NWGR_DIRECTED as nwgr_read_mm; flags NWGR_MM_INT_ATTRIBUTE also
 * keeps an int attribute per edge as read_mm<..., int> reads it (the sign and
 * digits the value field starts with, 1 in a pattern file), which
 * bench/ProcessEdgelist.cpp:37-50 writes out. */
enum { NWGR_MM_INT_ATTRIBUTE = 1 };
int nwgr_read_mm_as(const char* path, int nthreads, int directedness, int flags, nwgr_edge_list** out);

/* "NW Graph index_edge_list" binary (unipartite) -> edge list.  The first two
 * arrays are the ids; a third array of 4-byte values is kept as the int
 * attribute, further arrays are skipped (the reference reads the arrays its
 * own type has, containers/soa.hpp:302-309). */
int nwgr_read_nw_edge_list(const char* path, nwgr_edge_list** out);

/* Edge list -> "NW Graph index_edge_list" binary (unipartite, no
 * attributes); directedness NWGR_DIRECTED or NWGR_UNDIRECTED. */
int nwgr_write_nw_edge_list(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                            int directedness);
/* Same with an int attribute array after the ids (attr[m], or NULL for none):
 * the layout of edge_list<directedness, int>::serialize, the output of
 * bench/ProcessEdgelist.cpp. */
int nwgr_write_nw_edge_list_attr(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                                 const int32_t* attr, int directedness);

/* In-CSR (offsets n+1, indices m) -> "NW GRAPH indexed_struct_of_arrays"
 * binary, and back (the arrays are returned in a host CSR freed with
 * nwgr_host_csr_free, include/nwgr_graphgen.h). */
typedef struct nwgr_host_csr nwgr_host_csr;
int nwgr_write_nw_adjacency(const char* path, uint32_t n, uint64_t m, const uint32_t* offsets,
                            const uint32_t* indices);
int nwgr_read_nw_adjacency(const char* path, nwgr_host_csr** out);

uint32_t nwgr_edge_list_n(const nwgr_edge_list* e);
uint64_t nwgr_edge_list_m(const nwgr_edge_list* e);
int nwgr_edge_list_directedness(const nwgr_edge_list* e);
const uint32_t* nwgr_edge_list_src(const nwgr_edge_list* e);
const uint32_t* nwgr_edge_list_dst(const nwgr_edge_list* e);
const double* nwgr_edge_list_values(const nwgr_edge_list* e); /* NULL when the file has none */
const int32_t* nwgr_edge_list_int_values(const nwgr_edge_list* e); /* NULL unless kept (see above) */
void nwgr_edge_list_free(nwgr_edge_list* e);

#ifdef __cplusplus
}
#endif

#endif /* NWGR_IO_H */
