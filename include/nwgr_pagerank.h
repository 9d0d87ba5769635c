/*
 * nwgr_pagerank.h -- C ABI of the MI355X PageRank library (libnwgr_pagerank.so).
 *
 * Drop-in boundary for NWGraph's PageRank hot path.  The reference entry
 * point is the C++20 template
 *
 *   template <adjacency_list_graph Graph, typename Real>
 *   void nw::graph::page_rank(const Graph& graph,
 *                             const std::vector<vertex_id_type>& degrees,
 *                             std::vector<Real>& page_rank, Real damping_factor,
 *                             Real threshold, size_t max_iters, size_t num_threads);
 *   (include/nwgraph/algorithms/page_rank.hpp:89-91, driven by
 *    bench/pr.cpp:137-139 with Graph = adjacency<1>, Real = float)
 *
 * which the C++ header include/nwgr/page_rank.hpp re-declares on top of this
 * ABI.  Plain pointers and sizes only; no HIP, torch or C++ types cross it
 * (streams are passed as void*, which is what hipStream_t is).
 *
 * Graph input ("in-CSR", the reference's adjacency<1>): offsets[n+1] with
 * offsets[0] = 0 and offsets[n] = m (uint32, as default_index_t,
 * util/defaults.hpp:23-24), indices[m] uint32 = the in-neighbours of each row
 * (compressed.hpp:67-69, 179-186).  degrees[n] = the per-vertex divisor of
 * the contribution, the out-degree when built by build_degrees
 * (bench/common.hpp:109-120); NULL means "count out-degrees from indices".
 *
 * Iteration semantics: Jacobi sweeps as page_rank_v10
 * (experimental/algorithms/page_rank.hpp:450-505): from r_0 = 1/n (or init),
 *   r'_i = (1-d)*t_i + d*( sum_{j in row i} r_j/deg_j + D*u_i ),
 *   D = sum_{deg_j = 0} r_j,  u = 0 (LEAK, the reference), 1/n (UNIFORM) or t
 *   (TELEPORT);  error = sum_i |r'_i - r_i| accumulated in double;
 * stop after the first sweep with error < tol, or after max_iters sweeps
 * (page_rank.hpp:121-147).  Ranks are FP32 with FP32 accumulation.  Not
 * converging within max_iters is not an error (as in the reference): the
 * call returns NWGR_OK with result.converged = 0.
 *
 * Errors: every function returns an nwgr_status; on failure a message is
 * available from nwgr_last_error() (thread-local).  The reference has no
 * error reporting on this path (page_rank returns void).
 *
 * Threading: a graph handle is used by one host thread at a time; distinct
 * handles may be used concurrently.  num_threads of the reference has no
 * counterpart (it is unused there as well, page_rank.hpp:91).
 */
#ifndef NWGR_PAGERANK_H
#define NWGR_PAGERANK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nwgr_status {
  NWGR_OK = 0,
  NWGR_EINVAL = 1,   /* invalid argument or malformed CSR          */
  NWGR_ENOMEM = 2,   /* device or host allocation failed            */
  NWGR_EDEVICE = 3,  /* HIP runtime error / no usable device        */
  NWGR_ENOTSUP = 4,  /* feature not supported for this handle       */
} nwgr_status;

typedef enum nwgr_dangling {
  NWGR_DANGLING_LEAK = 0,     /* reference behaviour: mass of deg-0 vertices is lost */
  NWGR_DANGLING_UNIFORM = 1,  /* redistributed as 1/n                                */
  NWGR_DANGLING_TELEPORT = 2, /* redistributed along the teleport vector             */
} nwgr_dangling;

typedef struct nwgr_pr_graph nwgr_pr_graph; /* device-resident graph, opaque */

typedef struct nwgr_pr_params {
  double damping;       /* d, the reference's damping_factor (0.85)                   */
  double tol;           /* L1 threshold, the reference's threshold                    */
  uint32_t max_iters;   /* sweep cap, the reference's max_iters                       */
  uint32_t dangling;    /* nwgr_dangling                                              */
  const float* teleport;/* n entries in caller vertex order, or NULL (= 1/n).         */
                        /* Host pointer for nwgr_pr_run_f32, device pointer for the   */
                        /* _device entry points.                                      */
  const float* init;    /* initial ranks (n), same pointer rules, or NULL (= 1/n)     */
  uint32_t chunk;       /* sweeps per captured hipGraph launch; 0 = default           */
  uint32_t trace;       /* 1: print the reference's per-sweep trace table to stdout   */
} nwgr_pr_params;

typedef struct nwgr_pr_result {
  uint32_t iters;     /* sweeps performed                                              */
  uint32_t converged; /* 1 if the last sweep's error < tol                             */
  double error;       /* L1 change of the last sweep                                   */
  double ms;          /* device time of the call (init + sweeps + output), hipEvents   */
} nwgr_pr_result;

/* Fills the reference defaults of bench/pr.cpp: damping 0.85, tol 1e-4
 * (pr.cpp:27), max_iters 20 (pr.cpp:26), LEAK, no teleport/init. */
void nwgr_pr_params_default(nwgr_pr_params* p);

/* Checks p as every solve does before it starts -- damping in [0, 1], tol not
 * NaN, a known dangling mode (TELEPORT with a teleport vector) -- and, taking
 * p->teleport and p->init as HOST arrays of n entries, that every entry is
 * finite and non-negative.  NWGR_EINVAL otherwise.  The solves check device
 * vectors the same way on the device; the reference checks nothing
 * (page_rank.hpp:89-91 takes only damping and threshold). */
int nwgr_pr_params_check(const nwgr_pr_params* p, uint32_t n);

/* Upload an in-CSR (host arrays) to `device` and prepare it for sweeps:
 * relabelling, the tiled layout the default tuning's sweeps use, the work
 * plan, all built here so that the first solve costs what later ones do.
 * Replaces the graph/degrees arguments of page_rank (page_rank.hpp:89-91):
 * the caller keeps ownership of its arrays; the handle owns device memory.
 * Validates offsets (monotone, offsets[0]=0, offsets[n]=m) and indices < n.
 * n = 0 (with m = 0) gives an empty handle that holds no device memory and
 * needs no device; every solve on it returns NWGR_OK at once, as the
 * reference returns for N = 0 (page_rank.hpp:92-147): result.iters = 1
 * (max_iters when tol <= 0, 0 when max_iters = 0), error 0.
 * n < 2^31 (the sorts over vertices count in int) and m <= 2^32 - 2^28: the
 * reference's uint32 offsets (util/defaults.hpp:23-24) with room for the
 * tiled layout's padding in 32-bit ordinals.  Counts above one sort pass
 * (2^30 items, nwgr_pr_set_sort_batch) are sorted and scanned in passes.
 * Weighted solves (nwgr_wpr_*) need m < 2^31. */
int nwgr_pr_graph_create(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                         const uint32_t* degrees, int device, nwgr_pr_graph** out);

/* Same from device arrays (e.g. torch tensors) on `stream`.  The arrays are
 * read during the call only.  Indices are range-checked on the device. */
int nwgr_pr_graph_create_device(const uint32_t* d_offsets, const uint32_t* d_indices, uint32_t n,
                                uint64_t m, const uint32_t* d_degrees, int device, void* stream,
                                nwgr_pr_graph** out);

/* Items per device sort or scan call when a graph is built (default 2^30,
 * the range [2^16, 2^30]); larger counts go in passes of at most this many,
 * with results identical to a single call.  Process-wide; a small value makes
 * the passes testable on small graphs.  No reference counterpart (the
 * reference sorts on the host, build.hpp:209-265). */
int nwgr_pr_set_sort_batch(uint64_t items);

/* Destroys a handle; a handle of the resident-graph cache (below) is also
 * dropped from the cache. */
void nwgr_pr_graph_destroy(nwgr_pr_graph* g);

/* Resident-graph cache for repeated drop-in calls.  The reference's
 * page_rank runs on a graph the caller built once (page_rank.hpp:92-119 sets
 * up only the rank vectors) and pr.exe times that call over trials on one
 * graph (bench/pr.cpp:94-95); nw::graph::page_rank (include/nwgr/page_rank.hpp)
 * and nwgr_amd.page_rank get their device graph from here instead of
 * creating and destroying one per call.
 * nwgr_pr_graph_cached returns the handle made for an earlier call with the
 * same offsets, indices and degrees addresses, n, m and device whose
 * contents still match -- the degrees word for word, offsets and indices at
 * 4096 fixed positions and their last word -- (*hit = 1), else creates one
 * as nwgr_pr_graph_create does (*hit = 0; `hit` may be NULL).  The cache owns
 * its handles: keep one only until the next cache call (the least recently
 * used is destroyed when a new graph would pass the capacity, default 2), or
 * destroy it with nwgr_pr_graph_destroy.  A graph edited in place so that
 * the compared words stay equal is not noticed: clear the cache after such
 * an edit.  The cache functions are serialised by a lock; a handle itself is
 * used by one thread at a time. */
int nwgr_pr_graph_cached(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                         const uint32_t* degrees, int device, nwgr_pr_graph** out, uint32_t* hit);
/* One whole drop-in call on the resident graph of the arrays: the solve
 * nw::graph::page_rank performs (page_rank.hpp:89-91; NWGR_SOLVE_F32 as
 * nwgr_pr_run_f32 into float r_out, NWGR_SOLVE_F64 as nwgr_pr_run_f64 into
 * double r_out) or the residual push of page_rank_v13
 * (experimental/algorithms/page_rank.hpp:632-634; NWGR_SOLVE_DELTA_F32 as
 * nwgr_pr_delta_run_f32), with host vectors.  The graph is found or made as
 * by nwgr_pr_graph_cached, except that a key match is confirmed while the
 * solve runs: if the contents changed, that result is discarded, the graph
 * rebuilt and the solve run again.  n = 0 writes nothing and returns
 * NWGR_OK.  Holds the cache lock for the whole call.  *hit as for
 * nwgr_pr_graph_cached (`hit` may be NULL). */
#define NWGR_SOLVE_F32 0u
#define NWGR_SOLVE_F64 1u
#define NWGR_SOLVE_DELTA_F32 2u
int nwgr_pr_cached_run(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                       const uint32_t* degrees, int device, uint32_t solver, const nwgr_pr_params* p, void* r_out,
                       nwgr_pr_result* res, uint32_t* hit);
/* Destroys every cached handle. */
void nwgr_pr_cache_clear(void);
/* Most handles kept (1 .. 64); the least recently used beyond it are destroyed. */
int nwgr_pr_cache_set_capacity(uint32_t entries);
/* Lookups that found / did not find a resident graph, and handles held. */
void nwgr_pr_cache_stats(uint64_t* hits, uint64_t* misses, uint32_t* entries);

uint32_t nwgr_pr_graph_n(const nwgr_pr_graph* g);
uint64_t nwgr_pr_graph_m(const nwgr_pr_graph* g);
uint64_t nwgr_pr_graph_device_bytes(const nwgr_pr_graph* g);
/* Time (ms) the handle spent building its device layout (relabel, bins). */
double nwgr_pr_graph_setup_ms(const nwgr_pr_graph* g);

/* Plain PageRank; the ranks are written to the host array r_out[n] in the
 * caller's vertex order.  Replaces the body of page_rank
 * (page_rank.hpp:92-147); r_out plays the role of the page_rank vector.
 * A teleport or initial vector with a NaN, infinite or negative entry gives
 * NWGR_EINVAL.  The handle keeps the device buffers of this form, so a
 * repeated call allocates no device memory. */
int nwgr_pr_run_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* r_out, nwgr_pr_result* res);

/* Same with the output in device memory (d_r_out[n]) and all work enqueued
 * on `stream` (NULL = the null stream).  Synchronises `stream` before
 * returning, since the sweep count is decided on the device. */
int nwgr_pr_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* d_r_out,
                           nwgr_pr_result* res, void* stream);

/* Block Gauss-Seidel PageRank: the in-place sweep of the reference's
 * page_rank (v11, page_rank.hpp:124-139: each row's new rank and
 * contribution replace the old ones within the sweep) made deterministic.
 * The rows, in the handle's relabelled order (nwgr_pr_gs_phases), are cut
 * into `phases` consecutive phases of about equal sweep cost (8 per entry,
 * 2.9 per fragment -- about min(entries, tiles holding sources) + entries /
 * 256 -- 80 per row with entries, 1 per source row without, 0 per isolated
 * vertex; or the measured cuts of nwgr_pr_share_balance); inside a
 * phase every row gathers the contributions as they stood when the phase
 * began, and a phase's results replace the old ones before the next phase
 * gathers.  phases = 1 is the Jacobi solve of nwgr_pr_run_f32; more phases
 * take fewer sweeps to a tolerance (one row per phase would be the serial
 * v11 sweep) and launch two kernels per phase.  Dangling mass leaks (LEAK
 * only, as in the reference); teleport and init as for nwgr_pr_run_f32;
 * error = the L1 change of the whole sweep, stop rule and result as the
 * Jacobi solve.  phases in [1, 64]. */
int nwgr_pr_gs_run_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t phases, float* r_out,
                       nwgr_pr_result* res);
int nwgr_pr_gs_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t phases, float* d_r_out,
                              nwgr_pr_result* res, void* stream);
/* One GPU's share of a row-partitioned solve (no reference counterpart: the
 * single-GPU view of the multi-GPU design).  The rows are cut into `shares`
 * consecutive ranges as the phases of the block Gauss-Seidel solve; the
 * sweeps update the rows of share `share` alone, on a tiled layout that
 * holds only that share's entries and stores only its rows, against the
 * contributions of the other rows held at their initial values -- what one
 * GPU computes between the all-gathers of its slice.  Stop rule on the
 * share's own L1 change; d_r_out receives all n ranks (the other rows at
 * their initial values).  LEAK dangling; shares in [1, 64]. */
int nwgr_pr_share_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t shares, uint32_t share,
                                 float* d_r_out, nwgr_pr_result* res, void* stream);
/* Measured balance of the row shares (no reference counterpart: the
 * partition a multi-GPU solve would take).  Each of the `shares` shares is
 * swept alone; the rows are re-cut with every row of a share costing its
 * share's measured-to-mean sweep time (clamped to [1/2, 2]) times more, and
 * the layouts rebuilt, `rounds` times (<= 8); the cuts whose slowest share
 * measured shortest are kept -- for the shares and for the block Gauss-Seidel
 * solve with phases = shares, whose results follow its phase cuts -- until
 * the phases are rebuilt (another phase count, another tile width).
 * *us_max (may be NULL): that share's device time per sweep (us). */
int nwgr_pr_share_balance(nwgr_pr_graph* g, uint32_t shares, uint32_t rounds, double* us_max, void* stream);
/* The sweep order of the block Gauss-Seidel solve: order[i] (n entries,
 * caller ids; may be NULL) is the i-th row visited -- in-degree
 * non-increasing, ties by ascending id, the isolated vertices last -- and
 * cuts[0 .. phases] (may be NULL) the positions where the phases begin
 * (cuts[phases] = n): the cuts of the phases' layouts when they are built
 * for this phase count (balanced ones after nwgr_pr_share_balance), else the
 * model's. */
int nwgr_pr_gs_phases(const nwgr_pr_graph* g, uint32_t phases, uint32_t* order, uint32_t* cuts);

/* FP64 PageRank: the iteration and stop rule of nwgr_pr_run_f32 with double
 * ranks, contributions and sums and the integer degrees converted exactly --
 * the reference's Real = double (page_rank.hpp:89-91; its PageRank test runs
 * in double, test/page_rank_test.cpp:66,93).  r_out: n doubles in caller
 * order (host array; device array for the _device form, on `stream`).
 * Teleport and initial vectors are float arrays as for the FP32 solve (host
 * arrays here, device arrays for the _device form), widened to double.  The
 * sweeps run on the degree-binned work plan of the handle. */
int nwgr_pr_run_f64(nwgr_pr_graph* g, const nwgr_pr_params* p, double* r_out, nwgr_pr_result* res);
int nwgr_pr_run_device_f64(nwgr_pr_graph* g, const nwgr_pr_params* p, double* d_r_out, nwgr_pr_result* res,
                           void* stream);

/* Residual ("delta") PageRank.  Replaces page_rank_v13
 * (experimental/algorithms/page_rank.hpp:632-698, pr.exe version 13,
 * bench/pr.cpp:146) and the serial page_rank_v3 (:566-630), which returns
 * the same ranks: a residual it keeps is at or below the threshold and never
 * pushes.  Ranks start at 0, residuals at 1 - d; each iteration pushes every
 * residual above p->tol (the reference's `threshold`) into its vertex's
 * rank and sends residual * d / degree along its out-edges; the sum a vertex
 * receives is its next residual.  Stops after the first iteration that pushes
 * from no vertex of non-zero degree, or after p->max_iters iterations.  The
 * ranks are not normalised (divide by n for the LEAK PageRank they
 * approximate).  p->damping, p->tol, p->max_iters, p->chunk and p->trace
 * are used; a teleport, an initial vector or a dangling mode other than LEAK
 * is rejected.  The sweeps run on the engine the handle's tuning selects.
 * res->iters = iterations, res->error = pushes of the last one,
 * res->converged = 1 when that is 0; nwgr_pr_last_errors gives the pushes of
 * every iteration. */
int nwgr_pr_delta_run_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* r_out, nwgr_pr_result* res);
int nwgr_pr_delta_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* d_r_out,
                                 nwgr_pr_result* res, void* stream);

/* ---------------------------------------------------------------------------
 * Differentiable weighted PageRank (no reference counterpart: the weighted
 * form of the same Jacobi iteration, for gradient-based callers such as the
 * torch autograd bridge in nwgr_amd/autograd.py).
 *
 * Edge k of the caller's in-CSR (row v, neighbour u = indices[k]) is u -> v
 * with weight w[k] > 0.  P_uv = w_uv / s_u, s_u = sum of u's out-edge weights;
 *   r_v = (1-d) t_v + d (sum_{u->v} P_uv r_u + D q_v),  D = sum_{s_u=0} r_u,
 * q = 0 (LEAK), 1/n (UNIFORM), t (TELEPORT); t = p->teleport (device
 * pointer, caller order) or 1/n when NULL.  All vectors are device arrays in
 * the caller's vertex / edge order; the transposed CSR is built on the device
 * on first use and kept by the handle.
 * ------------------------------------------------------------------------- */

/* Forward fixed point; ranks to d_r_out[n].  Stop rule as the plain solve. */
int nwgr_wpr_forward_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, const float* d_w, float* d_r_out,
                                nwgr_pr_result* res, void* stream);

/* Adjoint of L(r) at the forward fixed point d_r (as returned by the
 * forward with the same w, p): given d_grad_r = dL/dr, writes
 * d_grad_w[m] = dL/dw (caller edge order) and d_grad_t[n] = dL/dt.
 * The adjoint iteration lambda = d (P lambda + 1_dangling (q . lambda)) + g
 * stops when the L1 change of lambda is below p->tol, an ABSOLUTE threshold
 * (scale it by ||dL/dr||_1 for a relative one), or after p->max_iters sweeps. */
int nwgr_wpr_adjoint_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, const float* d_w, const float* d_r,
                                const float* d_grad_r, float* d_grad_w, float* d_grad_t, nwgr_pr_result* res,
                                void* stream);

/* ---------------------------------------------------------------------------
 * Work decomposition of the sweeps (no reference counterpart: the knobs of
 * the degree-binned device layout, set by an autotuner or by hand).
 * ------------------------------------------------------------------------- */
typedef struct nwgr_pr_tuning {
  uint32_t lds_cache;          /* vertices (the hottest after the degree relabelling) whose
                                  contributions every sweep workgroup stages in LDS; 0 = none */
  uint32_t hub_threshold;      /* rows with more entries are split into chunks */
  uint32_t hub_chunk;          /* entries per hub chunk */
  uint32_t unit_edges;         /* target entries per work unit of the other rows */
  uint32_t zero_rows_per_unit; /* rows per work unit among rows without entries */
  uint32_t blocks_per_cu;      /* sweep workgroups (of 1024 threads) per compute unit */
  uint32_t chunk;              /* sweeps per captured hipGraph launch */
  uint32_t engine;             /* NWGR_ENGINE_BINNED, NWGR_ENGINE_TILED or NWGR_ENGINE_AUTO */
  uint32_t tile_bits;          /* tiled engine: source tile width 2^tile_bits, 13 .. 15 */
  uint32_t tile_fill_cost;     /* tiled engine: plain chunks (1024 entries) one tile fill is
                                  charged when the chunks are cut into workgroup ranges */
  uint32_t tile_combine;       /* tiled engine, second pass: NWGR_COMBINE_GATHER,
                                  NWGR_COMBINE_ACCUMULATE or NWGR_COMBINE_AUTO */
} nwgr_pr_tuning;

/* Second pass of the tiled engine.  GATHER: every row reads its partial sums
 * (one per source tile it has entries in) through a row-ordered slot list,
 * tiles ascending.  ACCUMULATE: rows are cut into blocks of at most 16384;
 * a workgroup streams the block's partials tile by tile (each tile's partials
 * of a block are contiguous) into a double-precision LDS accumulator per row,
 * then applies the row epilogue: no scattered reads, for graphs whose rows
 * have about one entry per source tile (low locality, e.g. Erdos-Renyi).
 * The double sum of a row's float partials is exact, hence independent of
 * the order the atomics land in, whenever the partials lie within a factor
 * of about 2^19 of each other.  AUTO: ACCUMULATE when the CSR has fewer than
 * 1.5 entries per fragment (run of one row inside one tile), else GATHER. */
enum { NWGR_COMBINE_GATHER = 0, NWGR_COMBINE_ACCUMULATE = 1, NWGR_COMBINE_AUTO = 2 };

/* Sweep engines.  BINNED: pull gathers over degree-binned rows, the hottest
 * contributions staged in LDS.  TILED: source tiles staged in LDS, per-tile
 * partial sums added by a second pass.  AUTO (default): TILED for a CSR whose
 * entries average at least two per (row, source tile) run or with more than
 * 2^21 vertices, else BINNED. */
enum { NWGR_ENGINE_BINNED = 0, NWGR_ENGINE_TILED = 1, NWGR_ENGINE_AUTO = 2 };

int nwgr_pr_graph_get_tuning(const nwgr_pr_graph* g, nwgr_pr_tuning* t);
/* Validates and applies `t`: re-plans the work units of the handle (cheap,
 * host side; the relabelled CSR is kept). */
int nwgr_pr_graph_set_tuning(nwgr_pr_graph* g, const nwgr_pr_tuning* t);

/* Per-sweep L1 errors of the last run (up to `cap` entries, and of at most
 * its first 4096 sweeps: the device history is that long, whatever
 * max_iters is); returns the number written.  Backs the reference's trace
 * table (page_rank.hpp:42-52). */
uint32_t nwgr_pr_last_errors(const nwgr_pr_graph* g, double* errs, uint32_t cap);

/* Device timing of the last run's sweep kernels: summed event time of the
 * captured sweep chunks (*sweep_ms), sweep-kernel launches (*launches: the
 * sweeps plus the launch that found the stop condition, plus any launched
 * past it) and chunks.  Backs the per-sweep time of the trace table
 * (page_rank.hpp:123,142).  Any output pointer may be NULL. */
int nwgr_pr_last_profile(const nwgr_pr_graph* g, double* sweep_ms, uint32_t* launches, uint32_t* chunks);

/* Device time (ms) of every sweep of the last run, when it was traced
 * (params.trace = 1 captures one sweep per hipGraph launch and times each,
 * the per-iteration time of the reference's trace, page_rank.hpp:123,142);
 * returns the number written (up to `cap`), 0 after an untraced run. */
uint32_t nwgr_pr_last_sweep_times(const nwgr_pr_graph* g, double* ms, uint32_t cap);

/* Batched personalized PageRank: k (1 .. 256) columns solved together, column
 * j with the one-hot teleport at vertex seeds[j] (caller ids; duplicates
 * allowed).  Per column: the iteration of the plain solve with teleport =
 * e_seed -- init 1/n, params->dangling LEAK, UNIFORM or TELEPORT (to the
 * seed); params->teleport, init and chunk are not used.  All columns sweep
 * together until every column's L1 change is below params->tol (or
 * max_iters); result->error is the largest column change of the last sweep.
 * R_out is n x k row-major (R_out[v * k + j]) in caller vertex order.  The
 * _device form takes device pointers and a hipStream_t (NULL = the handle's
 * stream).  No reference counterpart: the reference solves one teleport
 * vector per call (page_rank.hpp:89-148). */
int nwgr_ppr_batch_f32(nwgr_pr_graph* g, const nwgr_pr_params* params, uint32_t k, const uint32_t* seeds,
                       float* R_out, nwgr_pr_result* result);
int nwgr_ppr_batch_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* params, uint32_t k, const uint32_t* d_seeds,
                              float* d_R_out, nwgr_pr_result* result, void* stream);
/* Per-column L1 change of the last sweep of the last batched solve (up to
 * `cap` entries); returns the number written. */
uint32_t nwgr_ppr_last_column_errors(const nwgr_pr_graph* g, double* errs, uint32_t cap);

/* ---- in-CSR construction on the device ----------------------------------
 * From the edge list d_src[i] -> d_dst[i] (device arrays, ids < n): row v of
 * the result lists the sources of the edges into v in ascending order (the
 * reference's adjacency<1>, index_adjacency<1> adjacency.hpp:90-97 ->
 * fill_directed<1> build.hpp:209-265), with duplicates dropped under
 * NWGR_BUILD_DEDUP (uniq, build.hpp:425-435) and u == v dropped under
 * NWGR_BUILD_DROP_SELF_LOOPS (remove_self_loops, build.hpp:437-443).
 * d_offsets: n + 1 entries; d_indices: room for m entries; d_out_degree
 * (NULL = not wanted): n entries, the out-degrees of the result
 * (build_degrees, bench/common.hpp:109-120).  *m_out receives the entries
 * kept.  Ids >= n give NWGR_EINVAL.  m < 2^31. */
enum { NWGR_BUILD_DEDUP = 1, NWGR_BUILD_DROP_SELF_LOOPS = 2 };
int nwgr_csr_build_device(const uint32_t* d_src, const uint32_t* d_dst, uint64_t m, uint32_t n, uint32_t flags,
                          uint32_t* d_offsets, uint32_t* d_indices, uint32_t* d_out_degree, uint64_t* m_out,
                          void* stream);
/* The same build followed by nwgr_pr_graph_create_device on its result, with
 * the out-degrees as divisors. */
int nwgr_pr_graph_create_from_edges_device(const uint32_t* d_src, const uint32_t* d_dst, uint64_t m, uint32_t n,
                                           uint32_t flags, int device, void* stream, nwgr_pr_graph** out);

/* ---- many small graphs in one launch ------------------------------------
 * `count` independent graphs stored block-diagonally in one in-CSR over n
 * vertices: graph g owns vertices [vbase[g], vbase[g+1]) (vbase has count + 1
 * entries, vbase[0] = 0, vbase[count] = n, at most 65536 vertices per graph)
 * and every entry of its rows is one of its own vertices (else NWGR_EINVAL).
 * `degrees` (NULL = out-degrees) are the divisors, as in nwgr_pr_graph_create.
 * A run solves every graph on its own -- the plain iteration with n_g in
 * place of n: init 1/n_g (or params->init), teleport (1-d)/n_g (or the
 * caller's n-vector params->teleport, one distribution per graph), dangling
 * LEAK / UNIFORM / TELEPORT, stop on the graph's own L1 change < tol or
 * max_iters -- all in one persistent launch.  result->iters is the largest
 * sweep count, result->error the largest final change, result->converged is
 * set when every graph converged.  No reference counterpart: the reference
 * solves one graph per call (page_rank.hpp:89-148). */
typedef struct nwgr_multi_graph nwgr_multi_graph;
int nwgr_multi_create(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                      const uint32_t* vbase, uint32_t count, const uint32_t* degrees, int device,
                      nwgr_multi_graph** out);
void nwgr_multi_destroy(nwgr_multi_graph* mg);
/* Host r_out (n floats) and host params->teleport / init. */
int nwgr_multi_run_f32(nwgr_multi_graph* mg, const nwgr_pr_params* params, float* r_out, nwgr_pr_result* result);
/* Device r_out, teleport and init; stream NULL = the handle's stream. */
int nwgr_multi_run_device_f32(nwgr_multi_graph* mg, const nwgr_pr_params* params, float* d_r_out,
                              nwgr_pr_result* result, void* stream);
/* Sweeps and final L1 change of every graph of the last run (count entries
 * each; either pointer may be NULL). */
int nwgr_multi_graph_results(const nwgr_multi_graph* mg, uint32_t* iters, double* errors);
/* Decomposition: graphs with at most lds_max_n vertices (<= 36864) gather
 * from an LDS copy of their contributions; rows with more than
 * min(heavy_min, 512) entries are reduced by a wavefront each, shorter rows
 * by 2 to 32 lanes (16 entries per lane and pass) or one lane (at most 16
 * entries).  Default 36864, 512. */
int nwgr_multi_get_tuning(const nwgr_multi_graph* mg, uint32_t* lds_max_n, uint32_t* heavy_min);
int nwgr_multi_set_tuning(nwgr_multi_graph* mg, uint32_t lds_max_n, uint32_t heavy_min);
uint32_t nwgr_multi_n(const nwgr_multi_graph* mg);
uint32_t nwgr_multi_count(const nwgr_multi_graph* mg);
double nwgr_multi_setup_ms(const nwgr_multi_graph* mg);

/* ---- verification: top-k ranked vertices --------------------------------
 * d_ids[0 .. k) = the k vertices of largest rank, by non-increasing rank and
 * ascending vertex id among equal ranks; d_vals (may be NULL) their ranks.
 * Replaces the verification print of bench/pr.cpp:47-52 (print_n_ranks over
 * proxysort, util/proxysort.hpp:36-65, whose tie order is unspecified).
 * The _device form takes device arrays and a hipStream_t; nwgr_top_k_f32
 * takes host arrays and runs on `device`.  n < 2^31, k <= n. */
int nwgr_top_k_device_f32(const float* d_ranks, uint32_t n, uint32_t k, uint32_t* d_ids, float* d_vals,
                          void* stream);
int nwgr_top_k_f32(const float* ranks, uint32_t n, uint32_t k, uint32_t* ids, float* vals, int device);

/* nwgr_top_k_device_f32 without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The sort's scratch (keys, ids, the radix sort's temporary storage) is a
 * fixed 68 MiB array in the library's device memory, one per GPU, held until
 * the process ends; n <= 4194304 (2^22), else NWGR_EINVAL.  Same results as
 * nwgr_top_k_device_f32, on the current device, enqueued on `stream` without
 * synchronising it.  Every call on one GPU uses that GPU's one scratch: calls
 * that overlap on one GPU (other streams, other threads) race; calls on one
 * stream are ordered. */
int nwgr_top_k_device_f32_noalloc(const float* d_ranks, uint32_t n, uint32_t k, uint32_t* d_ids, float* d_vals,
                                  void* stream);

/* Engine the last solve on `g` ran: NWGR_ENGINE_BINNED, NWGR_ENGINE_TILED
 * (second pass GATHER), 2 (tiled, second pass ACCUMULATE) or 3 (tiled block
 * Gauss-Seidel phases); -1 before the first solve or for a null handle. */
int nwgr_pr_last_engine(const nwgr_pr_graph* g);

/* Thread-local message of the last failure on this thread ("" if none). */
const char* nwgr_last_error(void);

/* Number of HIP devices visible (0 when none); never fails. */
int nwgr_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* NWGR_PAGERANK_H */
