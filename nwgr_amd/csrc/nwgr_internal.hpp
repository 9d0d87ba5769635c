// Internal declarations shared by the HIP translation units of libnwgr_pagerank.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nwgr_pagerank.h"

namespace nwgr {

// ---------------------------------------------------------------------------
// errors: internal code throws Error; the C boundary converts it to a status.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define NWGR_HIP(call)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      int code_ = (e_ == hipErrorOutOfMemory) ? NWGR_ENOMEM : NWGR_EDEVICE;                   \
      throw ::nwgr::Error(code_, std::string(#call) + ": " + hipGetErrorString(e_) + " (" +    \
                                     __FILE__ + ":" + std::to_string(__LINE__) + ")");        \
    }                                                                                         \
  } while (0)

#define NWGR_REQUIRE(cond, msg)                                   \
  do {                                                            \
    if (!(cond)) throw ::nwgr::Error(NWGR_EINVAL, (msg));          \
  } while (0)

// Converts exceptions to a status at the ABI boundary.
template <class F>
int guarded(F&& f) {
  try {
    f();
    set_last_error("");
    return NWGR_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return NWGR_ENOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return NWGR_EDEVICE;
  }
}

// ---------------------------------------------------------------------------
// device memory owned by a handle
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void alloc(size_t n) {
    reset();
    if (n == 0) n = 16;
    NWGR_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// Work decomposition of one sweep.
//
// Rows (destination vertices) are relabelled so that in-degree is
// non-increasing with the new id.  The rows then fall into contiguous
// segments; each segment is processed by groups of G lanes per row
// (G = 1 .. 64), or, for hub rows, by whole workgroups over fixed-size edge
// chunks.  A "unit" is the work one workgroup takes from the queue at a time.
constexpr int kBlock = 256;      // threads per workgroup of the auxiliary kernels
constexpr int kWave = 64;        // one work unit is processed by one wavefront
constexpr int kSweepThreads = 1024;  // sweep workgroup: one per CU, 16 wavefronts
// Workgroups of a launch are dealt to the XCDs (each with its own L2) round
// robin: workgroup b runs on XCD b % kXcds.
constexpr uint32_t kXcds = 8;
// Contributions of vertices [0, kLdsCache) -- the hottest ones after the
// degree relabelling -- are staged in LDS by every sweep workgroup (144 KiB
// of the 160 KiB per CU).
constexpr uint32_t kLdsCache = 36864;
constexpr int kMaxSegs = 16;
// Edge arrays (indices, weights) carry this many extra entries so that the
// batched gather loops may load past the last edge of a row without a bounds
// branch; vertex arrays gathered from carry one extra entry (index n) that
// stays 0, the sentinel for loads past a row's end.
constexpr uint32_t kEdgePad = 4096;

struct Seg {
  uint32_t row_lo, row_hi;    // rows [row_lo, row_hi)
  uint32_t G;                 // lanes per row (power of two); 0 = hub chunks
  uint32_t rows_per_unit;     // rows a unit covers (multiple of P * kBlock / G)
  uint32_t unit_lo, unit_hi;  // global unit ids [unit_lo, unit_hi)
  uint32_t U;                 // edges per lane loaded in one batch (0: rows without entries)
  uint32_t P;                 // rows per lane group processed together
};

struct SegTable {
  Seg s[kMaxSegs];
  uint32_t nseg;
  uint32_t total_units;
};

// One chunk of a hub row.  Chunks of one row are consecutive units.
struct HubUnit {
  uint32_t row;
  uint32_t e_lo, e_hi;    // edge range of this chunk
  uint32_t first;         // index of the row's first chunk in the hub unit list
  uint32_t nchunks;       // chunks of this row
  uint32_t hub;           // hub row ordinal (ticket slot)
};

// Device-side solver state.  Cross-workgroup values are exchanged across
// kernel boundaries (ping-pong on the sweep parity), except the work queue
// counters and hub tickets, which use read-modify-write atomics only.
struct State {
  uint32_t ctr[2];    // ctr[p]: sweeps completed when the sweep of parity p ran
  uint32_t work[2];   // work queue counters, one per parity
  uint32_t done;      // 1 once converged or capped
  uint32_t iters;     // sweeps performed (valid once done)
  double err;         // error of the last sweep (valid once done)
  double dang;        // dangling mass (or adjoint Q) of the final iterate (valid once done)
  double dsum;        // tiled engine: dangling mass (or Q) of the sweep's input iterate
  double last_e;      // error of the latest completed sweep (the decay estimate past the history)
  float rq[2];        // quiet rows (no entries): their common rank in the iterate of parity p
  // Output permutation written by the gather pass of the stopping launch
  // (null: a launch of its own writes the output).  Rows from out_quiet_from
  // on hold the quiet rank; fin_done is set by the first tile pass after the
  // stop, so the later launches of the chunks in flight leave `out` alone.
  float* out;
  uint32_t out_quiet_from;
  uint32_t fin_done;
};

// Per-sweep errors kept on the device for the trace (the first kHistCap
// sweeps of a solve); the solve itself needs O(1) memory for any max_iters.
constexpr uint32_t kHistCap = 4096;

// Host-mapped mirror the solver publishes into: progress every sweep, the
// result when it stops.
struct HostFlag {
  volatile uint32_t done;
  volatile uint32_t iters;
  volatile double err;
  volatile uint32_t progress;   // sweeps completed as of the latest launch
  volatile uint32_t pad;
  volatile double last_err;     // error of sweep `progress`
  volatile double prev_err;     // error of sweep `progress - 1` (0 if none)
  volatile double first_err;    // error of the initial iterate (kDelta: pushes of iteration 1)
  volatile uint32_t finished;   // the output permutation enqueued before the stop was seen has run
  volatile uint32_t pad2;
};

// Timing of the last solve, from events around each captured chunk.
struct SolveProfile {
  double chunk_ms = 0;      // summed device time of the sweep chunks
  uint32_t launches = 0;    // sweep kernel launches (sweeps + the stopping launch + idle ones)
  uint32_t chunks = 0;
  std::vector<double> launch_ms;  // traced solves (one launch per chunk): device time of launch 1, 2, ...
};

// Engine choice.  Auto takes the tiled engine for a CSR whose entries form
// on average at least kTiledMinRun entries per (row, source tile) run ...
constexpr uint32_t kEngineBinned = 0, kEngineTiled = 1, kEngineAuto = 2;
constexpr uint32_t kCombineGather = 0, kCombineAccumulate = 1, kCombineAuto = 2;
constexpr double kTiledMinRun = 2.0;
// ... or for any CSR whose gathered vector outgrows one XCD's 4 MiB L2: the
// degree-binned gathers then come from the MALL a cache line per entry, while
// the tiled engine streams (measured on G(2^25, 16 * 2^25): 5.0 ms per sweep
// tiled against 8.6 ms binned).
constexpr uint32_t kTiledMinN = 1u << 21;

struct Tuning {
  uint32_t lds_cache = kLdsCache;  // hot vertices staged in LDS per sweep workgroup
  uint32_t hub_threshold = 4096;   // rows with in-degree > this are hubs
  uint32_t hub_chunk = 2048;       // edges per hub unit (one wavefront)
  uint32_t unit_edges = 2048;      // target edges per non-hub unit (one wavefront)
  uint32_t zero_rows_per_unit = 2048;
  uint32_t blocks_per_cu = 1;
  uint32_t chunk = 32;             // sweeps per captured graph (R-MAT s22: 766 against 751 GTEPS at 16)
  uint32_t engine = kEngineAuto;   // kEngineBinned, kEngineTiled or kEngineAuto
  uint32_t tile_bits = 15;         // tiled engine: log2 of the source tile width (13 .. 15)
  uint32_t tile_fill_cost = 40;    // tiled engine: chunk-equivalents charged per tile fill
  uint32_t tile_combine = kCombineAuto;  // tiled engine second pass: gather, accumulate or auto
};

// Work decomposition of one CSR orientation: segments over row positions and
// the hub chunk list.  `rows` maps a position to a row when the rows are not
// already stored in non-increasing degree order (the transposed CSR).
struct Plan {
  SegTable segs{};
  DevBuf hub_units;  // HubUnit[]
  DevBuf hub_part;   // u32[] (float bits) hub chunk partial sums
  DevBuf hub_ticket; // u32[] per hub row
  DevBuf rows;       // u32[n] position -> row, or empty (identity)
  uint32_t n_hub_units = 0, n_hubs = 0;
};

// ---------------------------------------------------------------------------
// Tiled layout of a CSR (the "tiled" engine).
//
// Source vertices are cut into tiles of T = 2^bits consecutive ids (the
// degree relabelling puts the hottest sources in the first tiles).  Edges are
// stored tile-major -- ordered by (tile, row, source) -- as 16-bit local ids
// whose top bit marks the first edge of a fragment: a run of edges of one row
// inside one tile, additionally cut at every group boundary (kGroupEdges
// consecutive entries: the share of 16 lanes, one DPP row).  A chunk is
// kChunkEdges consecutive entries (one wavefront's share, kLaneEdges per
// lane); every tile's run is padded to whole chunks, the padding forming
// dummy fragments.
//
// A sweep has two passes.  The tile pass gives every workgroup a fixed,
// contiguous range of chunks (cut at tile boundaries into tasks); per task it
// copies the tile's contributions into LDS and gathers from there.  Each lane
// reduces its entries; runs crossing lanes are joined by a segmented scan
// over the 16 lanes of a DPP row (no LDS traffic), and every fragment's sum
// is stored at the fragment's ordinal (the chunk's first ordinal,
// `chunk_sp0`, plus the fragment starts before it in the chunk), so a
// chunk's partials are written contiguously.  The combine pass adds each row's partials in tile order
// (`rslot` lists them row by row) and applies the row epilogue.
constexpr uint32_t kMaxTileBits = 15;          // local ids use 15 bits of a u16
constexpr uint32_t kMaxTile = 1u << kMaxTileBits;
constexpr uint32_t kLaneEdges = 16;
constexpr uint32_t kGroupEdges = 16 * kLaneEdges;
constexpr uint32_t kChunkEdges = 64 * kLaneEdges;
constexpr uint32_t kWaveRowMin = 32;           // rows with more partials are combined by a wavefront
// The gather form of the second pass works through items, each workgroup its
// own list: rows with more than kBigRowMin partials one each (the whole
// workgroup adds them), windows of consecutive rows -- at most kWindowRows
// rows (kRowsPerThread per thread) and kCombineStage partials, staged in LDS
// in one step -- and blocks of kQuietRows quiet rows.
constexpr uint32_t kCombineThreads = 256;
constexpr uint32_t kRowsPerThread = 4;
constexpr uint32_t kWindowRows = kCombineThreads * kRowsPerThread;
constexpr uint32_t kCombineStage = 2048;
constexpr uint32_t kBigRowMin = 2048;
constexpr uint32_t kQuietRows = 16 * kCombineThreads;
static_assert(kBigRowMin <= kCombineStage, "a row that is not big fits one window");

// Fragment sums of one chunk pass out of the tile pass through a per-wavefront
// LDS window of this many slots; a chunk with more fragments takes one pass
// per window.
constexpr uint32_t kTileWindow = 496;

// Storage order inside a chunk.  The layout is built in logical positions
// (entry i of lane l of chunk c at c*1024 + 16 l + i); the local ids and the
// tiled weights are stored permuted so that every 16-byte load of a
// wavefront reads one contiguous 1 KiB: ids as two halves (entries 0-7 of
// every lane, then 8-15), weights as four quarters (entries 4q .. 4q+3).
__host__ __device__ constexpr uint64_t eloc_pos(uint64_t P) {
  const uint32_t r = static_cast<uint32_t>(P % kChunkEdges), lane = r / kLaneEdges, i = r % kLaneEdges;
  return P - r + (i / 8) * (kChunkEdges / 2) + lane * 8 + i % 8;
}
__host__ __device__ constexpr uint64_t wgt_pos(uint64_t P) {
  const uint32_t r = static_cast<uint32_t>(P % kChunkEdges), lane = r / kLaneEdges, i = r % kLaneEdges;
  return P - r + (i / 4) * (kChunkEdges / 4) + lane * 4 + i % 4;
}
static_assert(eloc_pos(16 + 9) == 512 + 8 + 1 && wgt_pos(16 + 9) == 512 + 4 + 1 && eloc_pos(1024 + 7) == 1024 + 7,
              "chunk storage order");

constexpr uint32_t kNoTile = 0xffffffffu;

struct TiledTask {
  uint32_t tile;
  uint32_t chunk_lo, chunk_hi;  // chunks [chunk_lo, chunk_hi) of this tile
};

struct Tiled {
  bool built = false;
  uint32_t row_lo = 0, row_hi = 0;  // the rows whose entries the layout holds; the second pass writes only these
  bool quiet_tail = true;        // its second pass also refreshes the quiet rows (rows [nz, n) without entries)
  uint32_t bits = 0;             // tile width 2^bits the layout was built with
  uint32_t ntiles = 0, nchunks = 0;
  uint32_t nsub = 0;             // fragments including the dummy ones
  uint32_t nblocks = 0;          // workgroups of the tile pass the tasks were planned for
  uint32_t fill_cost = 0;        // tile fill cost the tasks were planned with
  uint32_t ntasks = 0;
  uint32_t nbig = 0;             // rows with more than kBigRowMin partials
  DevBuf eloc;       // u16[nchunks * kChunkEdges]
  DevBuf epos;       // u32[nchunks * kChunkEdges] CSR edge index (padding: m)
  DevBuf w;          // f32[nchunks * kChunkEdges] weights in tiled order (weighted kinds)
  DevBuf chunk_sp0;  // u32[nchunks] ordinal of the fragment holding each chunk's first entry
  DevBuf pair_ptr;   // u32[n + 1] each row's range in rslot
  DevBuf rslot;      // u32[nsub] fragment ordinals, row by row, tiles ascending
  DevBuf partial;    // f32[nsub] one partial sum per fragment ordinal
  DevBuf tasks;      // TiledTask[ntasks]
  DevBuf wg_task;    // u32[nblocks + 1] tasks [wg_task[b], wg_task[b + 1]) of workgroup b
  DevBuf wg_tile0;   // u32[nblocks] tile of workgroup b's first task (kNoTile: it has none)
  // The cut on the host: every chunk's cost (the fitted model, then scaled by
  // measured workgroup times), the tasks and each workgroup's first task; and
  // the workgroups' measured tile-pass times (10 ns ticks, kCalibSweeps
  // sweeps by done-sweep count modulo kCalibSweeps, then workgroup)
  std::vector<double> ccost;
  std::vector<TiledTask> tasks_h;
  std::vector<uint32_t> wg_h;
  DevBuf wg_time;
  uint32_t calib_rounds = 0;  // measured re-cuts kept
  // Big rows are left out of the windows of the gather pass: lpp and lslot are
  // pair_ptr and rslot over the other rows only (a big row's range empty), and
  // heavy marks the big rows (a bit per row).  Without big rows (light_split
  // false) the pass reads pair_ptr and rslot.
  bool light_split = false;
  DevBuf lpp;        // u32[n + 1]
  DevBuf lslot;      // u32[fragments of the other rows]
  DevBuf heavy;      // u32[(n + 31) / 32]
  // Windows of the gather pass (host): row boundaries win_h[0] = row_lo < ...
  // < win_h[nwin] = row_hi; no window crosses nz (so a solve whose quiet rows
  // are kept as a scalar takes the windows that start before nz).
  uint32_t nwin = 0;
  std::vector<uint32_t> win_h;
  // Item lists of the gather pass for [0] a sweep that stores every row and
  // [1] one that keeps the quiet rows as the scalar: items are numbered big
  // rows, then windows, then quiet-row blocks; workgroup b takes
  // items[ptr[b] .. ptr[b + 1]), each with its descriptor (GatherDesc),
  // planned for gplan_blocks workgroups.
  struct GatherPlan {
    uint32_t nwin = 0, nqb = 0;
    DevBuf ptr, items, desc;
  };
  GatherPlan gplan[2];
  uint32_t gplan_blocks = 0;
  // host: estimated cost of each big row and window; big rows with their
  // pair_ptr bounds; each window boundary's position in lpp
  std::vector<float> big_cost, win_cost;
  std::vector<uint32_t> big_h, big_pp, win_b;
  std::vector<uint32_t> tile_chunk0;  // host: first chunk of each tile (ntiles + 1)
  // accumulate form of the second pass (built on first use): rows cut into
  // blocks of at most kAccRows rows and about kAccFrags fragments; inside a
  // tile, a block's fragments are consecutive ordinals (rows ascend)
  bool acc_built = false;
  bool acc_on = false;   // the second pass accumulates (else it gathers through rslot)
  uint32_t nab = 0;      // row blocks
  uint32_t acc_rows = 0; // rows per block at most (kAccRows; the kernel also takes kAccRows / 2)
  DevBuf ablo;           // u32[nab + 1] first row of each block
  DevBuf aseg;           // u32[(nab + 1) * ntiles] first ordinal of block b's fragments in tile t
  DevBuf afrow;          // u16[nsub] row of each fragment minus its block's first row
};

constexpr uint32_t kAccRows = 16384;       // rows per accumulate block (double accumulators in LDS)
constexpr uint32_t kAccFrags = 1u << 18;   // fragments per accumulate block (cut earlier when exceeded)

// Kinds of fixed-point solve sharing the sweep machinery.
enum SolveKind : uint32_t {
  kPlain = 0,     // r' = (1-d) t + d (sum c_j + D u),  c = r / deg
  kWeighted = 1,  // same with per-edge weights, c = r / s (s = out-weight sums)
  kAdjoint = 2,   // lambda' = d (s > 0 ? (sum w lambda)/s : Q) + g   on the transpose
  kDelta = 3,     // residual push: res = sum of deltas; rank += res, delta = res d / deg if res > thr
  kNumKinds = 4,
  kCoef = 4,      // one pass on the transpose: gradient coefficients from the final lambda
};
// Captured-graph slot of the block Gauss-Seidel solve (plain kind, phases).
constexpr uint32_t kExecGs = kNumKinds;
constexpr uint32_t kExecSlots = kNumKinds + 1;
// Most phases of a block Gauss-Seidel sweep.
constexpr uint32_t kMaxGsPhases = 64;

// Per-edge weights enter the sweep of these kinds.
constexpr bool weighted_kind(uint32_t kind) { return kind == kWeighted || kind == kAdjoint || kind == kCoef; }

// What one fixed-point solve iterates over.
struct SolveSpec {
  uint32_t kind = kPlain;
  const Plan* plan = nullptr;
  const uint32_t* offs = nullptr;  // CSR the sweep gathers over
  const uint32_t* idx = nullptr;
  const float* w = nullptr;        // per-edge weights in that CSR's order (weighted kinds)
  const float* div = nullptr;      // per-row divisor: degree (plain) or out-weight sum
  const float* gsrc = nullptr;     // adjoint source term, new labels
  const float* init = nullptr;     // forward initial ranks in caller order, or null
  const float* tel_src = nullptr;  // teleport in caller order (permuted at init), or null
  const Tiled* tiled = nullptr;    // tiled engine layout, or null for the degree-binned engine
  const float* tw = nullptr;       // weights in the tiled order (weighted kinds, tiled engine)
  bool quiet = false;              // rows [g.nz, n) kept as the scalar State::rq (plain kind, tiled)
  float thr = 0.f;                 // push threshold (kDelta)
  // block Gauss-Seidel: one tiled layout per phase (rows of consecutive
  // ranges), swept in order over one contribution vector updated in place
  const std::vector<Tiled>* phases = nullptr;
  int only = -1;                   // >= 0: sweep that one phase alone (one GPU's row share)
};

// Device state of the batched personalized solve (ppr.hip).
struct PprState;
struct PprDeleter {
  void operator()(PprState* p) const;
};

struct Graph {
  int device = 0;
  bool cached = false;            // owned by the resident-graph cache (cache.hip)
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t num_cu = 0;
  uint32_t grid = 0;              // sweep workgroups
  double setup_ms = 0;
  Tuning tune;

  // relabelled layout
  DevBuf perm;      // u32[n]  new -> original id
  DevBuf iperm;     // u32[n]  original -> new id
  DevBuf offs;      // u32[n+1]
  DevBuf idx;       // u32[m]  new source ids
  DevBuf epos;      // u32[m]  new edge position -> original edge position
  DevBuf degf;      // f32[n]  divisor of the contribution (caller degrees), new order
  DevBuf degu;      // u32[n]  the same divisors as given (the FP64 solve)
  Plan fwd;         // plan of the in-CSR (rows already in degree order)
  std::vector<uint32_t> in_deg_h;  // host copy of in-degrees (new order)
  std::vector<uint32_t> offs_h;    // host copy of offs

  // transposed (out-)CSR in the new labels, built on first use
  bool has_t = false;
  DevBuf t_offs;    // u32[n+1]
  DevBuf t_idx;     // u32[m]  destinations
  DevBuf t_pos;     // u32[m]  out-CSR entry -> in-CSR (new order) position
  Plan adj;         // plan of the out-CSR (row list by out-degree)
  std::vector<uint32_t> t_deg_h, t_rows_h, t_offs_h;  // host copies for re-planning
  Tiled tfwd, tadj;  // tiled layouts of the in-CSR and of the transposed CSR (built on first use)
  int auto_tiled[2] = {-1, -1};  // engine auto choice per CSR (fwd, adj): -1 not measured yet
  uint32_t auto_bits = 0;        // tile width the auto choices were measured for
  uint32_t grid_c = 0;  // workgroups of the tiled engine's second pass (the resident count of the gather form)
  uint32_t part_stride = 0;  // partial pairs per parity in `partials` (at least every producer's count)

  // weighted / adjoint work arrays
  DevBuf w_in, w_out;  // f32[m] weights in in-CSR (new) and out-CSR order
  DevBuf s;            // f32[n] out-weight sums (new labels)
  DevBuf gsrc;         // f32[n] adjoint source term (new labels)
  DevBuf rsave;        // f32[n] forward ranks (new labels) for the gradient
  DevBuf coef_a, coef_b; // f32[n] per-source gradient coefficients

  // Quiet rows: the rows without entries, [nz, n) after the degree relabelling.
  // Under a uniform teleport with LEAK or UNIFORM dangling all of them hold one
  // common rank, so the tiled plain solve keeps it as a scalar (State::rq) and
  // never stores their ranks (their contributions rq / degree are stored).
  uint32_t nz = 0;
  // rows from n_act on are isolated vertices (no entries, no out-edges): no
  // sweep reads their contributions, and a solve that keeps the quiet rows
  // as a scalar stores none for them (n_act >= nz)
  uint32_t n_act = 0;
  uint32_t nq_dangling = 0;  // quiet rows whose divisor is 0 (their rank is dangling mass)
  bool quiet_ready = false;  // nq_dangling counted

  // iteration buffers
  DevBuf r[2], c[2];
  DevBuf tel;       // f32[n] permuted teleport (when given)
  DevBuf partials;  // double[2][grid][2]
  DevBuf state;     // State
  uint32_t hist_cap = 0;  // per-sweep errors kept in hist_h
  DevBuf out_buf;   // f32[n] ranks of a host-output solve before the copy back
  DevBuf tel_h, init_h;  // f32[n] device copies of a host-output solve's teleport / initial vector
  DevBuf vflag;     // u32 result of the device check of the caller's vectors
  HostFlag* hflag = nullptr;   // pinned, mapped
  double* hist_h = nullptr;    // pinned, mapped: the sweeps write their errors here directly
  HostFlag* hflag_dev = nullptr;
  double* hist_dev = nullptr;  // hist_h as the device addresses it
  void* pin_out = nullptr;     // pinned staging of host-output results (copy_out)
  size_t pin_bytes = 0;
  // timing events the solves reuse (one solve at a time per handle)
  std::vector<hipEvent_t> events;

  // captured chunk graphs: [kind or kExecGs][parity of the first sweep][length]
  static constexpr uint32_t kMaxChunk = 32;
  hipGraphExec_t exec[kExecSlots][2][kMaxChunk + 1] = {};
  uint64_t exec_key[kExecSlots] = {};
  // sweeps of the latest solve per slot, with its graphs' key (first-chunk length of the next solve)
  uint64_t hint_key[kExecSlots] = {};
  uint32_t hint_iters[kExecSlots] = {};
  // the tile passes record their workgroups' times (Tiled::wg_time)
  bool calib_on = false;

  // block Gauss-Seidel (built on first use): the phases' layouts, their row
  // cuts (positions in the relabelled order, phases + 1 entries) and the
  // per-phase (error, dangling) pairs [2 parities][gs_stride][2], phase q's
  // at the sum of the earlier phases' counts
  std::vector<Tiled> gs;
  std::vector<uint32_t> gs_cuts;
  uint32_t gs_bits = 0;
  DevBuf gs_part;
  uint32_t gs_stride = 0;  // pairs per parity in gs_part
  hipStream_t own_stream = nullptr;

  std::vector<double> last_errs;
  SolveProfile last_profile;
  int last_engine = -1;

  // FP64 solve (fp64.hip; built on first use): ranks and contributions with
  // the zero sentinel at index n, the permuted teleport, hub chunk partials
  DevBuf r64[2], c64[2], tel64, hub64;

  // batched personalized solve (built on first use)
  std::unique_ptr<PprState, PprDeleter> ppr;
  uint32_t ppr_cols = 0;

  void drop_graphs(int kind = -1) {
    for (int k = 0; k < static_cast<int>(kExecSlots); ++k) {
      if (kind >= 0 && k != kind) continue;
      for (auto& row : exec[k])
        for (auto& e : row)
          if (e) {
            (void)hipGraphExecDestroy(e);
            e = nullptr;
          }
      exec_key[k] = 0;
    }
  }

  ~Graph();
};

inline uint64_t bytes_of(const Tiled& t) {
  uint64_t b = 0;
  for (const DevBuf* d : {&t.eloc, &t.epos, &t.w, &t.chunk_sp0, &t.pair_ptr, &t.rslot, &t.partial, &t.tasks,
                          &t.wg_task, &t.wg_tile0, &t.gplan[0].ptr, &t.gplan[0].items,
                          &t.gplan[1].ptr, &t.gplan[1].items, &t.gplan[0].desc, &t.gplan[1].desc, &t.ablo, &t.aseg, &t.afrow, &t.lpp,
                          &t.lslot, &t.heavy, &t.wg_time})
    b += d->bytes;
  return b;
}

inline uint64_t bytes_of(const Graph& g) {
  uint64_t b = bytes_of(g.tfwd) + bytes_of(g.tadj) + g.gs_part.bytes;
  for (const Tiled& t : g.gs) b += bytes_of(t);
  for (const DevBuf* d : {&g.perm, &g.iperm, &g.offs, &g.idx, &g.epos, &g.degf, &g.fwd.hub_units,
                          &g.fwd.hub_part, &g.fwd.hub_ticket, &g.fwd.rows, &g.t_offs, &g.t_idx, &g.t_pos,
                          &g.adj.hub_units, &g.adj.hub_part, &g.adj.hub_ticket, &g.adj.rows, &g.w_in,
                          &g.w_out, &g.s, &g.gsrc, &g.rsave, &g.coef_a, &g.coef_b, &g.r[0], &g.r[1], &g.c[0],
                          &g.c[1], &g.tel, &g.partials, &g.state, &g.degu, &g.out_buf, &g.tel_h,
                          &g.init_h, &g.r64[0], &g.r64[1], &g.c64[0], &g.c64[1], &g.tel64, &g.hub64})
    b += d->bytes;
  return b;
}

// Items per hipCUB sort / scan call (batched.hip); larger counts are cut into
// passes of at most this many.
uint64_t sort_batch();
void set_sort_batch(uint64_t items);
// Stable sort of keys[0, count) (each < nbuckets, below 2^end_bit) as a
// placement: out_val[pos] = original index and out_key[pos] = key (out_key may
// be null when count exceeds one pass), with bucket_start[key] the first
// position of the key's run (the exclusive prefix of the key counts).
void stable_sort_place(const uint32_t* keys, uint64_t count, int end_bit, uint32_t nbuckets,
                       const uint32_t* bucket_start, uint32_t* out_key, uint32_t* out_val, hipStream_t s);
// out = exclusive prefix sum of in[0, count) (u32), in passes
void exclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t count, hipStream_t s);
// Most entries a graph may hold: below 2^32 with room for the tiled layout's
// padding (at most 1023 entries per source tile, and at most 2^18 tiles of
// 2^13 sources for n < 2^31) in 32-bit ordinals.
constexpr uint64_t kMaxGraphEntries = (uint64_t{1} << 32) - (uint64_t{1} << 28);

// resident-graph cache (cache.hip): the handle of a graph the caller passed
// before (same arrays and contents), or a new one the cache then owns
Graph* cache_acquire(const uint32_t* offs, const uint32_t* idx, uint32_t n, uint64_t m, const uint32_t* deg,
                     int device, bool& hit);
// drop `g` from the cache without destroying it; false if it is not cached
bool cache_forget(Graph* g);
// run(g) on the resident graph of the arrays, their contents compared while
// it runs (run again on a rebuilt graph if they changed)
void cache_run(const uint32_t* offs, const uint32_t* idx, uint32_t n, uint64_t m, const uint32_t* deg, int device,
               const std::function<void(Graph*)>& run, bool& hit);

// device-to-host copy of a result through the handle's pinned staging buffer
// (graph.hip); returns once `dst` holds it
void copy_out(Graph& g, void* dst, const void* src, size_t bytes, hipStream_t s);

// graph construction (graph.hip)
Graph* create_graph(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                    const uint32_t* degrees, int device, hipStream_t stream, bool inputs_on_device);
// segments and hub chunks for rows given in non-increasing degree order;
// `row_of` maps a position to a row (identity when empty)
void build_plan(Graph& g, Plan& plan, const std::vector<uint32_t>& deg_by_pos,
                const std::vector<uint32_t>& offs_h, const std::vector<uint32_t>& row_of, hipStream_t s);
// transposed CSR + its plan (graph.hip); no-op once built
void ensure_transpose(Graph& g, hipStream_t s);
// throws NWGR_EINVAL for a tuning out of range (graph.hip)
void validate_tuning(const Tuning& t);
// validates g.tune, rebuilds the plans and the sweep grid (graph.hip)
void replan(Graph& g, hipStream_t s);
// tiled layout of a CSR over n rows with m entries, all in rows [row_lo,
// row_hi) (tiled.hip); no-op once built
void build_tiled(Graph& g, Tiled& t, const uint32_t* offs, const uint32_t* idx, uint64_t m, uint32_t row_lo,
                 uint32_t row_hi, hipStream_t s);
// phase cuts of block Gauss-Seidel: `phases` consecutive ranges of the
// relabelled rows of about equal modelled sweep cost (positions, phases + 1);
// with old_cuts and scale, every row of old range k costs scale[k] times more
std::vector<uint32_t> gs_cut_rows(const Graph& g, uint32_t phases, const std::vector<uint32_t>* old_cuts = nullptr,
                                  const std::vector<double>* scale = nullptr);
// the phases' layouts for the row cuts `cuts` (phases + 1 entries), replacing
// any built before (tiled.hip)
void build_gs(Graph& g, const std::vector<uint32_t>& cuts, hipStream_t s);
// the phases' layouts for `phases` phases (tiled.hip); no-op once built
void ensure_gs(Graph& g, uint32_t phases, hipStream_t s);
// row blocks and segment table of the accumulate form of the second pass
// (tiled.hip); no-op once built
void build_accum(Graph& g, Tiled& t, hipStream_t s);
// cut the chunks of `t` into per-workgroup task ranges (tiled.hip)
void plan_tasks(Graph& g, Tiled& t, hipStream_t s);
// sweeps whose tile-pass workgroup times one calibration solve records
constexpr uint32_t kCalibSweeps = 8;
// scale the chunk costs of `t` by the measured-to-modelled time of the
// workgroup that added them (us per workgroup of the current cut) and cut
// again (tiled.hip); returns false when there is nothing to scale
bool recut_tasks(Graph& g, Tiled& t, const std::vector<double>& wg_us, hipStream_t s);
// measured re-cuts of the default layout's tile pass (pagerank.hip)
void calibrate_tile_pass(Graph& g, hipStream_t s);

// workgroups of the tile pass for the current tile width
uint32_t tile_blocks(const Graph& g);
// workgroups of the tile pass of layout t (fewer for a row-range layout)
uint32_t layout_blocks(const Graph& g, const Tiled& t);
// the tiled layout the engine choice selects for the in-CSR (adj = false) or
// the transposed CSR (adj = true), built on first use; null = degree-binned
Tiled* select_tiled(Graph& g, bool adj, hipStream_t s);
// entries (>= 1) of `weights` (CSR order, padded with zeros) gathered into the
// tiled order of `t` (t.w)
void tiled_weights(Graph& g, Tiled& t, const float* weights, hipStream_t s);
// elements of an iteration vector (r, c): n + 1 rounded up to whole maximal tiles
inline size_t iter_len(uint32_t n) {
  return ((static_cast<size_t>(n) + 1 + kMaxTile - 1) / kMaxTile) * kMaxTile;
}

// solves (pagerank.hip, weighted.hip)
// Runs the chunked fixed-point iteration; the final iterate is left in
// g.r[res.iters & 1] (new labels).
// What an item of the gather pass needs before its first vector load, read
// with its item number: a window's rows [x, y) and their partials [z, w) in
// lpp order; a big row x with its partials [y, z) in pair_ptr order; a
// quiet-row block's index x.
struct GatherDesc {
  uint32_t x, y, z, w;
};

// The gather plan a sweep of this kind takes (1: quiet rows kept as the
// scalar, which only the plain kind does).
inline uint32_t gather_plan_index(const Tiled& t, uint32_t kind, bool quiet) {
  (void)t;
  return kind == kPlain && quiet ? 1u : 0u;
}
// (Error, dangling) pairs one sweep of the solve writes (one per workgroup of
// each second pass): the count the next sweep's stop test adds.
uint32_t solve_pairs(const Graph& g, const SolveSpec& sp);
// Workgroups of the gather pass resident on one CU.
uint32_t combine_blocks_per_cu();
// Item lists of the gather pass (tiled.hip), for g.grid_c workgroups.
void plan_gather(Graph& g, Tiled& t, hipStream_t s);

// The output a solve ends with: its ranks in caller order (quiet rows from
// quiet_from on hold the quiet scalar).  solve() enqueues it as soon as the
// launches issued are expected to include the stopping one (it runs only if
// the solve has then stopped), and again after the stop when that was early.
struct SolveFinish {
  float* out;
  uint32_t quiet_from;
};
void solve(Graph& g, const SolveSpec& spec, const nwgr_pr_params& p, nwgr_pr_result& res, hipStream_t s,
           const SolveFinish* fin = nullptr);
void run_plain(Graph& g, const nwgr_pr_params& p, float* d_out, nwgr_pr_result& res, hipStream_t s);
// block Gauss-Seidel solve with `phases` phases (LEAK dangling); with
// only >= 0, the sweeps update that phase's rows alone (a row share)
void run_gs(Graph& g, const nwgr_pr_params& p, uint32_t phases, float* d_out, nwgr_pr_result& res, hipStream_t s,
            int only = -1);
// FP64 Jacobi solve (fp64.hip); d_out: n doubles in caller order
void run_plain64(Graph& g, const nwgr_pr_params& p, double* d_out, nwgr_pr_result& res, hipStream_t s);
// residual ("delta") solve, page_rank_v13; p.tol is the push threshold and
// res.iters counts iterations (pushes), res.error the last one's push count
void run_delta(Graph& g, const nwgr_pr_params& p, float* d_out, nwgr_pr_result& res, hipStream_t s);
// one gradient-coefficient pass over the transpose (kind kCoef)
void run_coef_pass(Graph& g, const SolveSpec& spec, const nwgr_pr_params& p, const float* lam, hipStream_t s);
// k personalized columns with one-hot teleports at d_seeds (caller ids);
// d_out is n x k row-major in caller order (ppr.hip)
void run_ppr(Graph& g, const nwgr_pr_params& p, uint32_t k, const uint32_t* d_seeds, float* d_out,
             nwgr_pr_result& res, hipStream_t s);
// count the quiet rows whose divisor is 0 (once per handle)
void prepare_quiet(Graph& g, hipStream_t s);
// out[v] = src[iperm[v]] * scale (new labels -> caller order)
void unpermute(Graph& g, const float* src, float* out, float scale, hipStream_t s);
void validate_params(const nwgr_pr_params& p);
// validate_params plus, when `host_vectors`, the teleport and initial vectors
// (host arrays of n entries): every entry finite and non-negative
void check_params(const nwgr_pr_params& p, uint32_t n, bool host_vectors);
// the same check of device vectors, by one pass on the device (d_flag: one
// u32 of scratch); synchronises `s` when there is a vector to check
void check_vectors(const nwgr_pr_params& p, uint32_t n, uint32_t* d_flag, hipStream_t s);
// result of a solve on a graph without vertices
void empty_result(const nwgr_pr_params& p, nwgr_pr_result& res);
void empty_delta_result(const nwgr_pr_params& p, nwgr_pr_result& res);

struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    NWGR_HIP(hipGetDevice(&prev));
    if (prev != dev) NWGR_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

}  // namespace nwgr
