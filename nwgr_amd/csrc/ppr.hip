// Batched personalized PageRank: k teleport columns over one graph, solved
// together as one block iteration R' = d (P^T R + dangling) + (1 - d) E,
// where column j of E is the one-hot vector of seed j.
//
// The rank block is stored row-major in the relabelled vertex order, K4
// floats per vertex (k rounded up to a multiple of 4, at most 256), so the
// gather of one in-neighbour is one coalesced row read: lane l of a
// wavefront holds columns 4l .. 4l+3 as a float4.  A sweep is three
// launches: k_ppr_sweep (rows, and chunks of hub rows, taken from a work
// queue), k_ppr_hub (hub rows: chunk partials in order + epilogue) and
// k_ppr_reduce (per-column L1 change and dangling mass, convergence).  The
// queue counter is reset by k_ppr_reduce, after every sweep kernel of the
// sweep has finished, so no fences are needed.
//
// Semantics per column are those of the plain solve with teleport = e_seed
// (page_rank_v10, experimental/algorithms/page_rank.hpp:450-505): init 1/n,
// dangling LEAK / UNIFORM / TELEPORT (to the seed), stop once every column's
// L1 change is below tol or at max_iters.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nwgr_internal.hpp"

namespace nwgr {

namespace {

constexpr uint32_t kPprMaxCols = 256;
constexpr uint32_t kPprHub = 2048;       // rows with more entries are split into chunks
constexpr uint32_t kPprChunk = 2048;     // entries per hub chunk
constexpr uint32_t kPprUnitEdges = 512;  // entries per unit of ordinary rows
constexpr uint32_t kPprUnitRows = 64;    // rows per unit at most
constexpr uint32_t kPprBatch = 8;        // in-neighbours loaded together
constexpr int kPB = 256;                 // threads per workgroup
constexpr uint32_t kNoHub = 0xffffffffu;

struct PprUnit {
  uint32_t row_lo, row_hi;  // rows [row_lo, row_hi), or the hub row
  uint32_t e_lo, e_hi;      // hub chunk: entries [e_lo, e_hi)
  uint32_t chunk;           // hub chunk ordinal, or kNoHub
};

struct PprHubRow {
  uint32_t row;
  uint32_t chunk0, nchunks;
};

struct PprArgs {
  const uint32_t* __restrict__ offs;
  const uint32_t* __restrict__ idx;
  const float* __restrict__ inv_deg;  // n + 1 entries, 0 for dangling rows and the sentinel n
  const float* __restrict__ R_in;     // (n + 1) x K4, row n zero
  float* __restrict__ R_out;
  const uint32_t* __restrict__ seeds;  // K4 new ids (n for padding columns)
  const double* __restrict__ dang_in;  // K4: dangling mass of R_in per column
  const PprUnit* __restrict__ units;
  const PprHubRow* __restrict__ hubs;
  float* __restrict__ hub_part;        // nchunks x K4
  double* __restrict__ part;           // [block][2][K4] err, dangling of R_out
  State* st;
  uint32_t* queue;
  uint32_t nunits, nhubs;
  uint32_t K4, n, mode;
  float d_hi, d_lo, one_minus_d, un;
};

constexpr uint32_t kLeak = 0, kUniform = 1, kTeleport = 2;

__device__ __forceinline__ float4 fma4(float4 x, float s, float4 acc) {
  return make_float4(fmaf(x.x, s, acc.x), fmaf(x.y, s, acc.y), fmaf(x.z, s, acc.z), fmaf(x.w, s, acc.w));
}

// Sum over entries [b, e) of R_in[u] * inv_deg[u], kPprBatch entries at a
// time (loads past e read the zero sentinel row n).
__device__ __forceinline__ float4 gather_rows(const PprArgs& a, uint32_t b, uint32_t e, uint32_t col) {
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t k = b; k < e; k += kPprBatch) {
    uint32_t u[kPprBatch];
    float s[kPprBatch];
    float4 x[kPprBatch];
#pragma unroll
    for (uint32_t j = 0; j < kPprBatch; ++j) {
      const uint32_t v = __builtin_amdgcn_readfirstlane(a.idx[k + j]);
      u[j] = k + j < e ? v : a.n;
    }
#pragma unroll
    for (uint32_t j = 0; j < kPprBatch; ++j) s[j] = a.inv_deg[u[j]];
#pragma unroll
    for (uint32_t j = 0; j < kPprBatch; ++j)
      x[j] = *reinterpret_cast<const float4*>(a.R_in + static_cast<size_t>(u[j]) * a.K4 + col);
#pragma unroll
    for (uint32_t j = 0; j < kPprBatch; ++j) acc = fma4(x[j], s[j], acc);
  }
  return acc;
}

// Row epilogue for the lane's four columns; accumulates the L1 change and
// the dangling mass of the new iterate.
__device__ __forceinline__ void ppr_finish(const PprArgs& a, uint32_t row, float4 z, uint32_t col,
                                           const uint4 seed, const float4 D, double (&err)[4], double (&dg)[4]) {
  const float4 old = *reinterpret_cast<const float4*>(a.R_in + static_cast<size_t>(row) * a.K4 + col);
  const bool dangling = a.inv_deg[row] == 0.f;
  const float zs[4] = {z.x, z.y, z.z, z.w};
  const float Ds[4] = {D.x, D.y, D.z, D.w};
  const uint32_t ss[4] = {seed.x, seed.y, seed.z, seed.w};
  const float os[4] = {old.x, old.y, old.z, old.w};
  float rn[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bool is_seed = row == ss[c];
    const float u = a.mode == kUniform ? a.un : (a.mode == kTeleport && is_seed ? 1.f : 0.f);
    const float base = is_seed ? a.one_minus_d : 0.f;
    const float s = fmaf(Ds[c], u, zs[c]);
    rn[c] = fmaf(a.d_hi, s, fmaf(a.d_lo, s, base));
    err[c] += fabs(static_cast<double>(rn[c]) - static_cast<double>(os[c]));
    if (dangling) dg[c] += static_cast<double>(rn[c]);
  }
  *reinterpret_cast<float4*>(a.R_out + static_cast<size_t>(row) * a.K4 + col) =
      make_float4(rn[0], rn[1], rn[2], rn[3]);
}

// Per-column (err, dangling) of the workgroup into part[blockIdx][2][K4].
__device__ __forceinline__ void ppr_block_partials(const PprArgs& a, double (&err)[4], double (&dg)[4],
                                                   uint32_t col, bool active) {
  __shared__ double s_acc[kPB / 64][64][8];
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    s_acc[w][l][c] = err[c];
    s_acc[w][l][4 + c] = dg[c];
  }
  __syncthreads();
  if (w == 0 && active) {
    double e4[4] = {0, 0, 0, 0}, d4[4] = {0, 0, 0, 0};
    for (uint32_t q = 0; q < kPB / 64; ++q)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        e4[c] += s_acc[q][l][c];
        d4[c] += s_acc[q][l][4 + c];
      }
    double* p = a.part + static_cast<size_t>(blockIdx.x) * 2 * a.K4;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      p[col + c] = e4[c];
      p[a.K4 + col + c] = d4[c];
    }
  }
}

__global__ __launch_bounds__(kPB) void k_ppr_sweep(PprArgs a) {
  __shared__ uint32_t s_done;
  if (threadIdx.x == 0) s_done = a.st->done;
  __syncthreads();
  if (s_done) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t col = lane * 4;
  const bool active = col < a.K4;
  const uint32_t cc = active ? col : 0;
  const uint4 seed = *reinterpret_cast<const uint4*>(a.seeds + cc);
  const float4 D = make_float4(static_cast<float>(a.dang_in[cc]), static_cast<float>(a.dang_in[cc + 1]),
                               static_cast<float>(a.dang_in[cc + 2]), static_cast<float>(a.dang_in[cc + 3]));
  double err[4] = {0, 0, 0, 0}, dg[4] = {0, 0, 0, 0};
  for (;;) {
    uint32_t ui = 0;
    if (lane == 0) ui = __hip_atomic_fetch_add(a.queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ui = __builtin_amdgcn_readfirstlane(ui);  // lane 0 took the unit
    if (ui >= a.nunits) break;
    const PprUnit u = a.units[ui];
    if (u.chunk != kNoHub) {
      const float4 z = gather_rows(a, u.e_lo, u.e_hi, cc);
      if (active) *reinterpret_cast<float4*>(a.hub_part + static_cast<size_t>(u.chunk) * a.K4 + col) = z;
      continue;
    }
    for (uint32_t row = u.row_lo; row < u.row_hi; ++row) {
      const uint32_t b = a.offs[row], e = a.offs[row + 1];
      const float4 z = gather_rows(a, b, e, cc);
      if (active) ppr_finish(a, row, z, col, seed, D, err, dg);
    }
  }
  ppr_block_partials(a, err, dg, col, active);
}

// Hub rows: one wavefront each, chunk partials in chunk order.
__global__ __launch_bounds__(kPB) void k_ppr_hub(PprArgs a, double* __restrict__ part_hub) {
  __shared__ uint32_t s_done;
  if (threadIdx.x == 0) s_done = a.st->done;
  __syncthreads();
  if (s_done) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t col = lane * 4;
  const bool active = col < a.K4;
  const uint32_t cc = active ? col : 0;
  const uint4 seed = *reinterpret_cast<const uint4*>(a.seeds + cc);
  const float4 D = make_float4(static_cast<float>(a.dang_in[cc]), static_cast<float>(a.dang_in[cc + 1]),
                               static_cast<float>(a.dang_in[cc + 2]), static_cast<float>(a.dang_in[cc + 3]));
  double err[4] = {0, 0, 0, 0}, dg[4] = {0, 0, 0, 0};
  const uint32_t nw = gridDim.x * (kPB / 64);
  for (uint32_t h = blockIdx.x * (kPB / 64) + threadIdx.x / 64; h < a.nhubs; h += nw) {
    const PprHubRow hr = a.hubs[h];
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t q = 0; q < hr.nchunks; ++q) {
      const float4 x = *reinterpret_cast<const float4*>(a.hub_part + static_cast<size_t>(hr.chunk0 + q) * a.K4 + cc);
      z.x += x.x;
      z.y += x.y;
      z.z += x.z;
      z.w += x.w;
    }
    if (active) ppr_finish(a, hr.row, z, col, seed, D, err, dg);
  }
  PprArgs b = a;
  b.part = part_hub;
  ppr_block_partials(b, err, dg, col, active);
}

struct PprReduceArgs {
  const double* part;      // [nb][2][K4]
  const double* part_hub;  // [nh][2][K4]
  double* dang_out;        // K4
  double* colerr;          // K4: L1 change of the last sweep per column
  State* st;
  HostFlag* hflag;
  double* hist;
  uint32_t* queue;
  uint32_t hist_cap, nb, nh, K4, k, max_iters;
  double tol;
};

// Per-column totals of a sweep's block partials.  Four groups of kPB threads
// each add a quarter of the blocks (then of the hub blocks) for their column,
// eight blocks' loads in flight per thread; group 0 adds the four group sums
// in group order, so the totals do not depend on the schedule.
constexpr uint32_t kReduceGroups = 4;

__global__ __launch_bounds__(kPB * kReduceGroups) void k_ppr_reduce(PprReduceArgs a) {
  __shared__ double s_max[kPB / 64];
  __shared__ double s_sum[2][kReduceGroups][kPB];
  if (a.st->done) return;
  const uint32_t gq = threadIdx.x / kPB;  // group
  const uint32_t t = threadIdx.x % kPB;   // column
  auto range_sum = [&](const double* part, uint32_t nblocks, double& e, double& D) {
    const uint32_t per = (nblocks + kReduceGroups - 1) / kReduceGroups;
    const uint32_t b0 = min(nblocks, gq * per), b1 = min(nblocks, b0 + per);
    uint32_t b = b0;
    for (; b + 8 <= b1; b += 8) {
      double ev[8], dv[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        ev[j] = part[(static_cast<size_t>(b + j) * 2) * a.K4 + t];
        dv[j] = part[(static_cast<size_t>(b + j) * 2 + 1) * a.K4 + t];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        e += ev[j];
        D += dv[j];
      }
    }
    for (; b < b1; ++b) {
      e += part[(static_cast<size_t>(b) * 2) * a.K4 + t];
      D += part[(static_cast<size_t>(b) * 2 + 1) * a.K4 + t];
    }
  };
  double e = 0, D = 0;
  if (t < a.K4) {
    range_sum(a.part, a.nb, e, D);
    range_sum(a.part_hub, a.nh, e, D);
  }
  s_sum[0][gq][t] = e;
  s_sum[1][gq][t] = D;
  __syncthreads();
  // every thread stays to the last barrier; group 0 finishes the columns
  double m = 0.0;
  if (gq == 0) {
    e = 0;
    D = 0;
#pragma unroll
    for (uint32_t q = 0; q < kReduceGroups; ++q) {
      e += s_sum[0][q][t];
      D += s_sum[1][q][t];
    }
    if (t < a.K4) {
      a.dang_out[t] = D;
      a.colerr[t] = e;
    }
    m = t < a.k ? e : 0.0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (gq == 0 && (t & 63) == 0) s_max[t / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = 0;
    for (int i = 0; i < kPB / 64; ++i) mx = fmax(mx, s_max[i]);
    const uint32_t sweeps = a.st->ctr[0] + 1;
    a.st->ctr[0] = sweeps;
    if (sweeps - 1 < a.hist_cap) a.hist[sweeps - 1] = mx;
    a.st->err = mx;
    a.hflag->progress = sweeps;
    a.hflag->last_err = mx;
    *a.queue = 0;
    if (mx < a.tol || sweeps >= a.max_iters) {
      a.st->done = 1;
      a.st->iters = sweeps;
      a.hflag->iters = sweeps;
      a.hflag->err = mx;
      a.hflag->done = 1;
    }
  }
}

__global__ void k_ppr_init(float* __restrict__ R0, float* __restrict__ R1, uint32_t n, uint32_t K4, float un) {
  const uint64_t total = static_cast<uint64_t>(n + 1) * K4;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    R0[i] = i < static_cast<uint64_t>(n) * K4 ? un : 0.f;
    R1[i] = 0.f;
  }
}

__global__ void k_inv_deg(const float* __restrict__ degf, uint32_t n, float* __restrict__ inv) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x)
    inv[i] = (i < n && degf[i] != 0.f) ? 1.0f / degf[i] : 0.f;
}

__global__ void k_ppr_seeds(const uint32_t* __restrict__ seeds_caller, const uint32_t* __restrict__ iperm,
                            uint32_t k, uint32_t K4, uint32_t n, uint32_t* __restrict__ seeds_new) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < K4) seeds_new[j] = j < k ? iperm[seeds_caller[j]] : n;
}

// out[v * k + j] = R[iperm[v]][j]
__global__ void k_ppr_unpermute(const float* __restrict__ R, const uint32_t* __restrict__ iperm, uint32_t n,
                                uint32_t k, uint32_t K4, float* __restrict__ out) {
  const uint64_t total = static_cast<uint64_t>(n) * k;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = i / k, j = i - v * k;
    out[i] = R[static_cast<uint64_t>(iperm[v]) * K4 + j];
  }
}

inline uint32_t blocks_for(uint64_t work) {
  return static_cast<uint32_t>(std::clamp<uint64_t>((work + kPB - 1) / kPB, 1, 8192));
}

}  // namespace

// Device buffers of the batched solve, sized for K4 columns.
struct PprState {
  uint32_t K4 = 0;
  DevBuf R[2], inv_deg, seeds, dang[2], colerr, units, hubs, hub_part, part, part_hub, queue;
  uint32_t nunits = 0, nhubs = 0, nchunks = 0, nb = 0, nh = 0, ndangling = 0;
  bool planned = false;
};

namespace {

void ppr_plan(Graph& g, PprState& ps, hipStream_t s) {
  if (ps.planned) return;
  std::vector<PprUnit> units;
  std::vector<PprHubRow> hubs;
  uint32_t chunks = 0;
  const uint32_t n = g.n;
  uint32_t r = 0;
  while (r < n) {
    const uint32_t deg = g.in_deg_h[r];
    if (deg > kPprHub) {
      const uint32_t b = g.offs_h[r], e = g.offs_h[r + 1];
      PprHubRow h{r, chunks, 0};
      for (uint32_t c = b; c < e; c += kPprChunk) {
        units.push_back({r, r + 1, c, std::min(c + kPprChunk, e), chunks++});
        ++h.nchunks;
      }
      hubs.push_back(h);
      ++r;
      continue;
    }
    uint32_t hi = r, edges = 0;
    while (hi < n && hi - r < kPprUnitRows && g.in_deg_h[hi] <= kPprHub && (hi == r || edges + g.in_deg_h[hi] <= kPprUnitEdges)) {
      edges += g.in_deg_h[hi];
      ++hi;
    }
    units.push_back({r, hi, 0, 0, kNoHub});
    r = hi;
  }
  ps.nunits = static_cast<uint32_t>(units.size());
  ps.nhubs = static_cast<uint32_t>(hubs.size());
  ps.nchunks = chunks;
  ps.units.alloc(sizeof(PprUnit) * std::max<size_t>(units.size(), 1));
  ps.hubs.alloc(sizeof(PprHubRow) * std::max<size_t>(hubs.size(), 1));
  if (!units.empty())
    NWGR_HIP(hipMemcpyAsync(ps.units.p, units.data(), sizeof(PprUnit) * units.size(), hipMemcpyHostToDevice, s));
  if (!hubs.empty())
    NWGR_HIP(hipMemcpyAsync(ps.hubs.p, hubs.data(), sizeof(PprHubRow) * hubs.size(), hipMemcpyHostToDevice, s));
  ps.inv_deg.alloc(sizeof(float) * (static_cast<size_t>(n) + 1));
  k_inv_deg<<<blocks_for(n + 1), kPB, 0, s>>>(g.degf.as<float>(), n, ps.inv_deg.as<float>());
  NWGR_HIP(hipGetLastError());
  std::vector<float> inv(n);
  NWGR_HIP(hipMemcpyAsync(inv.data(), ps.inv_deg.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  ps.ndangling = static_cast<uint32_t>(std::count(inv.begin(), inv.end(), 0.f));
  ps.nb = g.num_cu * 4;
  ps.nh = std::max<uint32_t>(1, std::min<uint32_t>(g.num_cu, (ps.nhubs + 3) / 4));
  ps.queue.alloc(sizeof(uint32_t));
  ps.planned = true;
}

void ppr_alloc(Graph& g, PprState& ps, uint32_t K4) {
  if (ps.K4 == K4) return;
  const size_t rows = static_cast<size_t>(g.n) + 1;
  for (auto& b : ps.R) b.alloc(sizeof(float) * rows * K4);
  ps.seeds.alloc(sizeof(uint32_t) * K4);
  for (auto& b : ps.dang) b.alloc(sizeof(double) * K4);
  ps.colerr.alloc(sizeof(double) * K4);
  ps.hub_part.alloc(sizeof(float) * std::max<size_t>(ps.nchunks, 1) * K4);
  ps.part.alloc(sizeof(double) * ps.nb * 2 * K4);
  ps.part_hub.alloc(sizeof(double) * ps.nh * 2 * K4);
  NWGR_HIP(hipMemset(ps.part_hub.p, 0, ps.part_hub.bytes));  // stays 0 without hub rows
  ps.K4 = K4;
}

}  // namespace

void run_ppr(Graph& g, const nwgr_pr_params& p, uint32_t k, const uint32_t* d_seeds, float* d_out,
             nwgr_pr_result& res, hipStream_t s) {
  {
    // the seeds are the teleport vectors: TELEPORT dangling needs no params->teleport
    nwgr_pr_params q = p;
    q.teleport = d_seeds ? reinterpret_cast<const float*>(d_seeds) : nullptr;
    if (q.dangling != NWGR_DANGLING_TELEPORT) q.teleport = p.teleport;
    validate_params(q);
  }
  NWGR_REQUIRE(k >= 1 && k <= kPprMaxCols, "k must be in [1, 256]");
  NWGR_REQUIRE(d_seeds && d_out, "null argument");
  {
    std::vector<uint32_t> hs(k);
    NWGR_HIP(hipMemcpyAsync(hs.data(), d_seeds, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, s));
    NWGR_HIP(hipStreamSynchronize(s));
    for (uint32_t v : hs) NWGR_REQUIRE(v < g.n, "seed out of range");
  }
  if (!g.ppr) g.ppr.reset(new PprState());
  PprState& ps = *g.ppr;
  ppr_plan(g, ps, s);
  const uint32_t K4 = (k + 3) & ~3u;
  ppr_alloc(g, ps, K4);
  const uint32_t n = g.n;

  hipEvent_t ev0, ev1;
  NWGR_HIP(hipEventCreate(&ev0));
  NWGR_HIP(hipEventCreate(&ev1));
  std::memset(const_cast<HostFlag*>(g.hflag), 0, sizeof(HostFlag));
  NWGR_HIP(hipEventRecord(ev0, s));
  k_ppr_seeds<<<1, kPB, 0, s>>>(d_seeds, g.iperm.as<uint32_t>(), k, K4, n, ps.seeds.as<uint32_t>());
  k_ppr_init<<<blocks_for(static_cast<uint64_t>(n + 1) * K4), kPB, 0, s>>>(ps.R[0].as<float>(), ps.R[1].as<float>(),
                                                                           n, K4, static_cast<float>(1.0 / n));
  NWGR_HIP(hipGetLastError());
  {
    // dangling mass of the uniform start, per column
    std::vector<double> d0(K4, static_cast<double>(ps.ndangling) * static_cast<double>(static_cast<float>(1.0 / n)));
    NWGR_HIP(hipMemcpyAsync(ps.dang[0].p, d0.data(), sizeof(double) * K4, hipMemcpyHostToDevice, s));
  }
  NWGR_HIP(hipMemsetAsync(g.state.p, 0, sizeof(State), s));
  NWGR_HIP(hipMemsetAsync(ps.queue.p, 0, sizeof(uint32_t), s));

  PprArgs a{};
  a.offs = g.offs.as<uint32_t>();
  a.idx = g.idx.as<uint32_t>();
  a.inv_deg = ps.inv_deg.as<float>();
  a.seeds = ps.seeds.as<uint32_t>();
  a.units = ps.units.as<PprUnit>();
  a.hubs = ps.hubs.as<PprHubRow>();
  a.hub_part = ps.hub_part.as<float>();
  a.part = ps.part.as<double>();
  a.st = g.state.as<State>();
  a.queue = ps.queue.as<uint32_t>();
  a.nunits = ps.nunits;
  a.nhubs = ps.nhubs;
  a.K4 = K4;
  a.n = n;
  a.mode = p.dangling;
  const double d = p.damping;
  a.d_hi = static_cast<float>(d);
  a.d_lo = static_cast<float>(d - static_cast<double>(a.d_hi));
  a.one_minus_d = static_cast<float>(1.0 - d);
  a.un = static_cast<float>(1.0 / n);
  PprReduceArgs ra{};
  ra.part = ps.part.as<double>();
  ra.part_hub = ps.part_hub.as<double>();
  ra.colerr = ps.colerr.as<double>();
  ra.st = g.state.as<State>();
  ra.hflag = g.hflag_dev;
  ra.hist = g.hist_dev;
  ra.queue = ps.queue.as<uint32_t>();
  ra.hist_cap = g.hist_cap;
  ra.nb = ps.nb;
  ra.nh = ps.nh;
  ra.K4 = K4;
  ra.k = k;
  ra.max_iters = p.max_iters;
  ra.tol = p.tol;

  // sweeps in batches; launches after the stop return at once
  const uint32_t kBatchSweeps = 4;
  std::vector<hipEvent_t> marks;
  uint32_t launched = 0;
  bool stop = p.max_iters == 0;
  size_t checked = 0;
  while (!stop && launched < p.max_iters) {
    for (uint32_t i = 0; i < kBatchSweeps && launched < p.max_iters; ++i, ++launched) {
      const uint32_t par = (launched + 1) & 1;  // sweep launched+1 writes R[par]
      a.R_in = ps.R[par ^ 1].as<float>();
      a.R_out = ps.R[par].as<float>();
      a.dang_in = ps.dang[par ^ 1].as<double>();
      ra.dang_out = ps.dang[par].as<double>();
      k_ppr_sweep<<<ps.nb, kPB, 0, s>>>(a);
      if (ps.nhubs) k_ppr_hub<<<ps.nh, kPB, 0, s>>>(a, ps.part_hub.as<double>());
      k_ppr_reduce<<<1, kPB * kReduceGroups, 0, s>>>(ra);
    }
    NWGR_HIP(hipGetLastError());
    hipEvent_t e;
    NWGR_HIP(hipEventCreate(&e));
    NWGR_HIP(hipEventRecord(e, s));
    marks.push_back(e);
    // keep one batch in flight: wait for the previous batch, then look
    if (marks.size() >= 2) {
      NWGR_HIP(hipEventSynchronize(marks[marks.size() - 2]));
      checked = marks.size() - 1;
      if (g.hflag->done) stop = true;
    }
  }
  (void)checked;
  NWGR_HIP(hipEventRecord(ev1, s));
  NWGR_HIP(hipEventSynchronize(ev1));
  for (auto e : marks) (void)hipEventDestroy(e);
  State st{};
  NWGR_HIP(hipMemcpy(&st, g.state.p, sizeof(State), hipMemcpyDeviceToHost));
  const uint32_t iters = p.max_iters == 0 ? 0 : st.iters;
  const float* fin = ps.R[iters & 1].as<float>();
  k_ppr_unpermute<<<blocks_for(static_cast<uint64_t>(n) * k), kPB, 0, s>>>(fin, g.iperm.as<uint32_t>(), n, k, K4,
                                                                          d_out);
  NWGR_HIP(hipGetLastError());
  NWGR_HIP(hipStreamSynchronize(s));
  float ms = 0;
  NWGR_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  res.iters = iters;
  res.error = p.max_iters == 0 ? 0.0 : st.err;
  res.converged = (iters > 0 && res.error < p.tol) ? 1u : 0u;
  res.ms = ms;
  g.last_errs.assign(g.hist_h, g.hist_h + std::min(iters, g.hist_cap));  // written by the sweeps (mapped)
  g.ppr_cols = k;
}

void ppr_column_errors(const Graph& g, double* out, uint32_t cap, uint32_t& written) {
  written = 0;
  if (!g.ppr || !g.ppr->colerr.p) return;
  written = std::min(cap, g.ppr_cols);
  if (written) NWGR_HIP(hipMemcpy(out, g.ppr->colerr.p, sizeof(double) * written, hipMemcpyDeviceToHost));
}

void PprDeleter::operator()(PprState* p) const { delete p; }

}  // namespace nwgr

struct nwgr_pr_graph : nwgr::Graph {};

extern "C" {

int nwgr_ppr_batch_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t k, const uint32_t* d_seeds,
                              float* d_R_out, nwgr_pr_result* res, void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    NWGR_REQUIRE(g->n > 0, "a graph without vertices has no seed vertices");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_ppr(*g, *p, k, d_seeds, d_R_out, *res, s);
  });
}

int nwgr_ppr_batch_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t k, const uint32_t* seeds, float* R_out,
                       nwgr_pr_result* res) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(g && p && res && seeds && R_out, "null argument");
    NWGR_REQUIRE(k >= 1 && k <= nwgr::kPprMaxCols, "k must be in [1, 256]");
    NWGR_REQUIRE(g->n > 0, "a graph without vertices has no seed vertices");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = g->own_stream;
    nwgr::DevBuf ds, dr;
    ds.alloc(sizeof(uint32_t) * k);
    dr.alloc(sizeof(float) * static_cast<size_t>(g->n) * k);
    NWGR_HIP(hipMemcpyAsync(ds.p, seeds, sizeof(uint32_t) * k, hipMemcpyHostToDevice, s));
    nwgr::run_ppr(*g, *p, k, ds.as<uint32_t>(), dr.as<float>(), *res, s);
    NWGR_HIP(hipMemcpy(R_out, dr.p, sizeof(float) * static_cast<size_t>(g->n) * k, hipMemcpyDeviceToHost));
  });
}

uint32_t nwgr_ppr_last_column_errors(const nwgr_pr_graph* g, double* errs, uint32_t cap) {
  if (!g || !errs) return 0;
  uint32_t w = 0;
  try {
    nwgr::ppr_column_errors(*g, errs, cap, w);
  } catch (...) {
    return 0;
  }
  return w;
}

}  // extern "C"
