// PageRank sweeps on gfx950 and the C ABI of libnwgr_pagerank.
//
// One sweep = one launch of k_sweep: a pull gather-SpMV over the in-CSR
// (page_rank.hpp:127-136, per-edge gather :129-131) with the teleport,
// dangling-mass and L1-residual terms fused into the row epilogue.  Jacobi
// semantics as page_rank_v10 (experimental/algorithms/page_rank.hpp:450-505).
//
// Work distribution: workgroups take units from a queue (one RMW atomic per
// unit); rows are pre-sorted by in-degree so each unit covers rows of one
// degree band, processed by G lanes per row (G = 1..64), or a chunk of one hub
// row processed by the whole workgroup.  Hub chunks are combined in chunk
// order by the last chunk to finish (RMW atomics only, no cache flushes).
//
// Cross-workgroup totals (L1 error, dangling mass) never meet inside a kernel:
// every workgroup writes its partial sums, and every workgroup of the NEXT
// sweep re-reduces them in the same fixed order, so all workgroups take the
// same convergence decision without fences.  The sweep that finds the
// previous sweep converged (or the cap reached) records the result and exits;
// later launches see `done` and exit at once.  Sweeps are captured in chunks
// of K launches in a hipGraph; the host reads a mapped flag after each chunk.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "nwgr_internal.hpp"

namespace nwgr {

namespace {
thread_local std::string t_last_error;
}

void set_last_error(const std::string& msg) { t_last_error = msg; }

namespace {

constexpr uint32_t kDanglingLeak = 0, kDanglingUniform = 1, kDanglingTeleport = 2;

struct SweepArgs {
  const uint32_t* __restrict__ offs;
  const uint32_t* __restrict__ idx;
  const float* __restrict__ degf;
  const float* __restrict__ tel;     // permuted teleport or null (uniform)
  const float* __restrict__ r_in;
  const float* __restrict__ c_in;
  float* __restrict__ r_out;
  float* __restrict__ c_out;
  const double* __restrict__ part_in;   // [grid][2] of the previous sweep (or init)
  double* __restrict__ part_out;        // [grid][2] of this sweep
  State* st;
  HostFlag* hflag;
  double* hist;
  uint32_t hist_cap;
  const HubUnit* __restrict__ hub_units;
  uint32_t* hub_part;
  uint32_t* hub_ticket;
  SegTable segs;
  uint32_t p;          // parity of this sweep
  uint32_t max_iters;
  uint32_t dangling;
  float d_hi, d_lo;    // damping as an unevaluated float sum d_hi + d_lo
  float base_u;        // (1-d)/n as float (uniform teleport)
  float one_minus_d;   // 1-d as float (explicit teleport)
  float un;            // 1/n as float
  double tol;
};

struct InitArgs {
  const uint32_t* __restrict__ perm;
  const float* __restrict__ degf;
  const float* __restrict__ init;      // caller order or null
  const float* __restrict__ tel_src;   // caller order or null
  float* __restrict__ tel;             // permuted copy (when tel_src)
  float* __restrict__ r0;
  float* __restrict__ c0;
  double* __restrict__ part0;
  State* st;
  uint32_t n;
  float un;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum of (a, b) over the workgroup; every thread receives the totals.
// Deterministic: fixed butterfly inside each wave, waves added in order.
__device__ __forceinline__ void block_sum2(double& a, double& b, double (*s)[kBlock / 64]) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    s[0][w] = a;
    s[1][w] = b;
  }
  __syncthreads();
  double ta = 0, tb = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) {
    ta += s[0][i];
    tb += s[1][i];
  }
  a = ta;
  b = tb;
}

__device__ __forceinline__ uint32_t ld_stream(const uint32_t* p) { return __builtin_nontemporal_load(p); }

// Row epilogue: teleport + dangling + damping, residual, next contribution.
__device__ __forceinline__ void finish_row(const SweepArgs& a, uint32_t row, float z, float Df, double& err,
                                           double& dang) {
  const float rold = a.r_in[row];
  float base, t;
  if (a.tel) {
    t = a.tel[row];
    base = a.one_minus_d * t;
  } else {
    t = a.un;
    base = a.base_u;
  }
  const float u = a.dangling == kDanglingUniform ? a.un : (a.dangling == kDanglingTeleport ? t : 0.f);
  const float s = fmaf(Df, u, z);
  const float rn = fmaf(a.d_hi, s, fmaf(a.d_lo, s, base));
  a.r_out[row] = rn;
  const float dg = a.degf[row];
  a.c_out[row] = rn / dg;  // IEEE division; +inf/NaN for dg == 0 is never gathered
  err += fabs(static_cast<double>(rn) - static_cast<double>(rold));
  if (dg == 0.f) dang += static_cast<double>(rn);
}

__global__ __launch_bounds__(kBlock) void k_init(InitArgs a) {
  __shared__ double s_red[2][kBlock / 64];
  double dang = 0, unused = 0;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const uint32_t src = a.perm[i];
    const float r = a.init ? a.init[src] : a.un;
    const float dg = a.degf[i];
    a.r0[i] = r;
    a.c0[i] = r / dg;
    if (dg == 0.f) dang += static_cast<double>(r);
    if (a.tel_src) a.tel[i] = a.tel_src[src];
  }
  block_sum2(dang, unused, s_red);
  if (threadIdx.x == 0) {
    a.part0[2 * blockIdx.x] = INFINITY;  // "previous error" seen by sweep 1
    a.part0[2 * blockIdx.x + 1] = dang;
    if (blockIdx.x == 0) {
      a.st->ctr[0] = 0;
      a.st->ctr[1] = 0;
      a.st->work[0] = 0;
      a.st->work[1] = 0;
      a.st->done = 0;
      a.st->iters = 0;
      a.st->err = INFINITY;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_sweep(SweepArgs a) {
  __shared__ double s_red[2][kBlock / 64];
  __shared__ float s_hub[kBlock / 64];
  __shared__ uint32_t s_unit[2];
  __shared__ uint32_t s_done;
  const uint32_t tid = threadIdx.x;
  // one decision per workgroup: `done` may be written by workgroup 0 of this
  // very launch (only when every workgroup stops anyway)
  if (tid == 0) s_done = a.st->done;
  __syncthreads();
  if (s_done) return;

  // totals of the previous sweep, identical in every workgroup
  double E = 0, D = 0;
  for (uint32_t b = tid; b < gridDim.x; b += kBlock) {
    E += a.part_in[2 * b];
    D += a.part_in[2 * b + 1];
  }
  block_sum2(E, D, s_red);
  const uint32_t done_sweeps = a.st->ctr[a.p ^ 1];
  if ((done_sweeps > 0 && E < a.tol) || done_sweeps >= a.max_iters) {
    if (blockIdx.x == 0 && tid == 0) {
      a.st->done = 1;
      a.st->iters = done_sweeps;
      a.st->err = E;
      if (done_sweeps > 0 && done_sweeps - 1 < a.hist_cap) a.hist[done_sweeps - 1] = E;
      a.hflag->iters = done_sweeps;
      a.hflag->err = E;
      a.hflag->done = 1;
    }
    return;
  }
  if (blockIdx.x == 0 && tid == 0) {
    a.st->ctr[a.p] = done_sweeps + 1;
    __hip_atomic_exchange(&a.st->work[a.p ^ 1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (done_sweeps > 0 && done_sweeps - 1 < a.hist_cap) a.hist[done_sweeps - 1] = E;
    // progress for the host's chunk planning
    a.hflag->last_err = E;
    a.hflag->progress = done_sweeps;
  }

  const float Df = static_cast<float>(D);
  double err = 0, dang = 0;
  uint32_t* queue = &a.st->work[a.p];
  const uint32_t total = a.segs.total_units;
  if (tid == 0) s_unit[0] = __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  uint32_t sel = 0;
  uint32_t cur = s_unit[0];
  while (cur < total) {
    if (tid == 0)
      s_unit[sel ^ 1] = __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // segment of this unit (unrolled: constant indices keep the table in SGPRs)
    Seg sg = a.segs.s[0];
#pragma unroll
    for (int i = 1; i < kMaxSegs; ++i)
      if (i < static_cast<int>(a.segs.nseg) && cur >= a.segs.s[i].unit_lo) sg = a.segs.s[i];
    const uint32_t ul = cur - sg.unit_lo;
    if (sg.G == 0) {
      // one chunk of a hub row, whole workgroup
      const HubUnit hu = a.hub_units[ul];
      float z = 0.f;
#pragma unroll 4
      for (uint32_t k = hu.e_lo + tid; k < hu.e_hi; k += kBlock) z += a.c_in[ld_stream(a.idx + k)];
      z = wave_sum(z);
      if ((tid & 63) == 0) s_hub[tid >> 6] = z;
      __syncthreads();
      if (tid == 0) {
        float zs = 0.f;
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) zs += s_hub[i];
        uint32_t old = __hip_atomic_exchange(&a.hub_part[ul], __float_as_uint(zs), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        // the partial must be performed before the ticket is taken
        asm volatile("s_waitcnt vmcnt(0)" : : "v"(old) : "memory");
        const uint32_t t = __hip_atomic_fetch_add(&a.hub_ticket[hu.hub], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        if (t == hu.nchunks - 1) {
          float zz = 0.f;
          for (uint32_t c = 0; c < hu.nchunks; ++c)
            zz += __uint_as_float(__hip_atomic_fetch_or(&a.hub_part[hu.first + c], 0u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT));
          __hip_atomic_exchange(&a.hub_ticket[hu.hub], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          finish_row(a, hu.row, zz, Df, err, dang);
        }
      }
    } else {
      const uint32_t G = sg.G;
      const uint32_t lg = __builtin_ctz(G);
      const uint32_t lane = tid & (G - 1);
      const uint32_t row_end = min(sg.row_lo + (ul + 1) * sg.rows_per_unit, sg.row_hi);
      for (uint32_t row = sg.row_lo + ul * sg.rows_per_unit + (tid >> lg); row < row_end;
           row += (kBlock >> lg)) {
        const uint32_t b = a.offs[row], e = a.offs[row + 1];
        float z = 0.f;
#pragma unroll 4
        for (uint32_t k = b + lane; k < e; k += G) z += a.c_in[ld_stream(a.idx + k)];
        for (uint32_t o = G >> 1; o > 0; o >>= 1) z += __shfl_xor(z, o);
        if (lane == 0) finish_row(a, row, z, Df, err, dang);
      }
    }
    __syncthreads();
    sel ^= 1;
    cur = s_unit[sel];
  }
  block_sum2(err, dang, s_red);
  if (tid == 0) {
    a.part_out[2 * blockIdx.x] = err;
    a.part_out[2 * blockIdx.x + 1] = dang;
  }
}

// out[v] = r_final[iperm[v]]; the final iterate is r[iters & 1].
__global__ __launch_bounds__(kBlock) void k_unpermute(const uint32_t* __restrict__ iperm,
                                                      const float* __restrict__ r0,
                                                      const float* __restrict__ r1, const State* st,
                                                      uint32_t n, float* __restrict__ out) {
  const float* r = (st->iters & 1) ? r1 : r0;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) out[v] = r[iperm[v]];
}

uint64_t fnv_mix(uint64_t h, uint64_t v) {
  h ^= v;
  return h * 0x100000001b3ULL;
}

SweepArgs sweep_args(Graph& g, const nwgr_pr_params& p, uint32_t parity, bool has_tel) {
  SweepArgs a{};
  a.offs = g.offs.as<uint32_t>();
  a.idx = g.idx.as<uint32_t>();
  a.degf = g.degf.as<float>();
  a.tel = has_tel ? g.tel.as<float>() : nullptr;
  a.r_in = g.r[parity ^ 1].as<float>();
  a.c_in = g.c[parity ^ 1].as<float>();
  a.r_out = g.r[parity].as<float>();
  a.c_out = g.c[parity].as<float>();
  double* part = g.partials.as<double>();
  a.part_in = part + 2 * g.grid * (parity ^ 1);
  a.part_out = part + 2 * g.grid * parity;
  a.st = g.state.as<State>();
  a.hflag = g.hflag_dev;
  a.hist = g.hist.as<double>();
  a.hist_cap = g.hist_cap;
  a.hub_units = g.hub_units.as<HubUnit>();
  a.hub_part = g.hub_part.as<uint32_t>();
  a.hub_ticket = g.hub_ticket.as<uint32_t>();
  a.segs = g.segs;
  a.p = parity;
  a.max_iters = p.max_iters;
  a.dangling = p.dangling;
  const double d = p.damping;
  a.d_hi = static_cast<float>(d);
  a.d_lo = static_cast<float>(d - static_cast<double>(a.d_hi));
  a.base_u = static_cast<float>((1.0 - d) / g.n);
  a.one_minus_d = static_cast<float>(1.0 - d);
  a.un = static_cast<float>(1.0 / g.n);
  a.tol = p.tol;
  return a;
}

void validate_params(const nwgr_pr_params& p) {
  NWGR_REQUIRE(std::isfinite(p.damping) && p.damping >= 0.0 && p.damping <= 1.0,
               "damping must be in [0, 1]");
  NWGR_REQUIRE(!std::isnan(p.tol), "tol is NaN");
  NWGR_REQUIRE(p.dangling <= kDanglingTeleport, "unknown dangling mode");
  NWGR_REQUIRE(p.dangling != kDanglingTeleport || p.teleport != nullptr,
               "dangling = TELEPORT needs a teleport vector");
}

}  // namespace

// Captured chunk of `len` consecutive sweep launches whose first launch has
// parity `parity` (sweep j has parity j & 1).
hipGraphExec_t chunk_graph(Graph& g, const nwgr_pr_params& p, bool has_tel, uint32_t parity, uint32_t len,
                           hipStream_t s) {
  hipGraphExec_t& e = g.exec[parity][len];
  if (e) return e;
  NWGR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (uint32_t i = 0; i < len; ++i) {
    SweepArgs a = sweep_args(g, p, parity ^ (i & 1), has_tel);
    k_sweep<<<g.grid, kBlock, 0, s>>>(a);
  }
  hipGraph_t graph;
  NWGR_HIP(hipStreamEndCapture(s, &graph));
  hipError_t err = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  NWGR_HIP(err);
  return e;
}

void run_plain(Graph& g, const nwgr_pr_params& p, float* d_out, nwgr_pr_result& res, hipStream_t s) {
  validate_params(p);
  const uint32_t K = std::clamp<uint32_t>(p.chunk ? p.chunk : g.tune.chunk, 1, Graph::kMaxChunk);
  if (g.hist_cap < p.max_iters || !g.hist.p) {
    g.hist_cap = std::max<uint32_t>(p.max_iters, 1);
    g.hist.alloc(sizeof(double) * g.hist_cap);
    g.drop_graphs();
  }
  const bool has_tel = p.teleport != nullptr;
  if (has_tel && !g.tel.p) {
    g.tel.alloc(sizeof(float) * g.n);
    g.drop_graphs();
  }

  // the captured graphs bake in the parameters and the stream
  uint64_t key = 0xcbf29ce484222325ULL, bits;
  double dd = p.damping, tt = p.tol;
  std::memcpy(&bits, &dd, 8);
  key = fnv_mix(key, bits);
  std::memcpy(&bits, &tt, 8);
  key = fnv_mix(key, bits);
  key = fnv_mix(key, p.max_iters);
  key = fnv_mix(key, p.dangling);
  key = fnv_mix(key, has_tel);
  key = fnv_mix(key, reinterpret_cast<uintptr_t>(s));
  if (g.exec_key != key) {
    g.drop_graphs();
    g.exec_key = key;
  }

  struct Events {
    std::vector<hipEvent_t> v;
    ~Events() {
      for (auto e : v) (void)hipEventDestroy(e);
    }
    hipEvent_t make(bool timing = true) {
      hipEvent_t e;
      NWGR_HIP(hipEventCreateWithFlags(&e, timing ? hipEventDefault : hipEventDisableTiming));
      v.push_back(e);
      return e;
    }
  } evs;
  hipEvent_t ev0 = evs.make(), ev1 = evs.make();

  std::memset(const_cast<HostFlag*>(g.hflag), 0, sizeof(HostFlag));
  NWGR_HIP(hipEventRecord(ev0, s));
  InitArgs ia{};
  ia.perm = g.perm.as<uint32_t>();
  ia.degf = g.degf.as<float>();
  ia.init = p.init;
  ia.tel_src = p.teleport;
  ia.tel = has_tel ? g.tel.as<float>() : nullptr;
  ia.r0 = g.r[0].as<float>();
  ia.c0 = g.c[0].as<float>();
  ia.part0 = g.partials.as<double>();
  ia.st = g.state.as<State>();
  ia.n = g.n;
  ia.un = static_cast<float>(1.0 / g.n);
  k_init<<<g.grid, kBlock, 0, s>>>(ia);
  NWGR_HIP(hipGetLastError());

  // Chunks of captured sweeps, at most two in flight.  Launch j performs
  // sweep j, or stops if sweep j-1 converged / the cap was reached, so the
  // solve needs launches 1 .. (converging sweep + 1) <= max_iters + 1.  After
  // the first chunk, chunk lengths follow the observed geometric decay of
  // the error so that the stopping launch closes the last chunk.
  struct Chunk {
    hipEvent_t a, b;
    uint32_t len;
  };
  std::vector<Chunk> pending;
  size_t head = 0;
  const uint64_t limit = static_cast<uint64_t>(p.max_iters) + 1;  // last launch index needed
  uint64_t next = 1;                                                // next launch index
  SolveProfile prof;
  uint32_t p1 = 0, p2 = 0;
  double e1 = 0, e2 = 0;  // two (progress, error) observations
  auto launch = [&](uint32_t len) {
    Chunk c{evs.make(), evs.make(), len};
    hipGraphExec_t e = chunk_graph(g, p, has_tel, static_cast<uint32_t>(next & 1), len, s);
    NWGR_HIP(hipEventRecord(c.a, s));
    NWGR_HIP(hipGraphLaunch(e, s));
    NWGR_HIP(hipEventRecord(c.b, s));
    pending.push_back(c);
    next += len;
    prof.launches += len;
    prof.chunks += 1;
  };
  auto plan = [&]() -> uint64_t {
    // launches still useful beyond `next - 1`
    if (next > limit) return 0;
    uint64_t want = K;
    if (p2 > p1 && e1 > 0 && e2 > 0 && e2 < e1 && std::isfinite(p.tol) && p.tol > 0) {
      const double rho = std::pow(e2 / e1, 1.0 / (p2 - p1));
      double more = e2 < p.tol ? 0.0 : std::ceil(std::log(p.tol / e2) / std::log(rho));
      if (!std::isfinite(more) || more > 1e6) more = 1e6;
      const uint64_t stop_launch = p2 + static_cast<uint64_t>(more) + 1;  // launch that detects it
      want = stop_launch >= next ? stop_launch - next + 1 : 0;
    }
    want = std::min<uint64_t>(want, K);
    want = std::min<uint64_t>(want, limit - next + 1);
    return want;
  };
  launch(static_cast<uint32_t>(std::min<uint64_t>(K, limit)));
  bool done = false;
  while (!done) {
    if (pending.size() - head < 2) {
      uint64_t len = plan();
      if (len > 0) {
        launch(static_cast<uint32_t>(len));
        continue;
      }
    }
    if (head == pending.size()) {
      // nothing in flight and nothing planned: extend by one launch
      if (next > limit) throw Error(NWGR_EDEVICE, "solver passed its launch limit without stopping");
      launch(1);
      continue;
    }
    NWGR_HIP(hipEventSynchronize(pending[head].b));
    float cms = 0;
    NWGR_HIP(hipEventElapsedTime(&cms, pending[head].a, pending[head].b));
    prof.chunk_ms += cms;
    ++head;
    if (g.hflag->done) {
      done = true;
      break;
    }
    const uint32_t pr = g.hflag->progress;
    const double le = g.hflag->last_err;
    if (pr > p2) {
      p1 = p2;
      e1 = e2;
      p2 = pr;
      e2 = le;
    }
  }
  // drain chunks still in flight (their launches see `done` and return)
  for (; head < pending.size(); ++head) {
    NWGR_HIP(hipEventSynchronize(pending[head].b));
    float cms = 0;
    NWGR_HIP(hipEventElapsedTime(&cms, pending[head].a, pending[head].b));
    prof.chunk_ms += cms;
  }
  k_unpermute<<<std::min<uint32_t>((g.n + kBlock - 1) / kBlock, 8192), kBlock, 0, s>>>(
      g.iperm.as<uint32_t>(), g.r[0].as<float>(), g.r[1].as<float>(), g.state.as<State>(), g.n, d_out);
  NWGR_HIP(hipGetLastError());
  NWGR_HIP(hipEventRecord(ev1, s));
  NWGR_HIP(hipEventSynchronize(ev1));
  float ms = 0;
  NWGR_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  res.iters = g.hflag->iters;
  res.error = g.hflag->err;
  res.converged = (res.iters > 0 && res.error < p.tol) ? 1u : 0u;
  res.ms = ms;
  g.last_profile = prof;
  g.last_errs.assign(std::min(res.iters, g.hist_cap), 0.0);
  if (!g.last_errs.empty())
    NWGR_HIP(hipMemcpy(g.last_errs.data(), g.hist.p, sizeof(double) * g.last_errs.size(), hipMemcpyDeviceToHost));
  if (p.trace) {
    // the reference's trace table (page_rank.hpp:42-52, :109, :142); the
    // per-sweep time is the sweep chunks' device time divided evenly
    std::printf("%-5s%-20s%-20s%-20s\n", "iter", "error", "time", "outgoing");
    for (uint32_t i = 0; i < g.last_errs.size(); ++i)
      std::printf("%-5u%-20.10f%-20.10f%-20.10f\n", i, g.last_errs[i], prof.chunk_ms / 1e3 / prof.launches, 0.0);
  }
}

}  // namespace nwgr

using nwgr::guarded;

struct nwgr_pr_graph : nwgr::Graph {};

extern "C" {

void nwgr_pr_params_default(nwgr_pr_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->damping = 0.85;
  p->tol = 1e-4;
  p->max_iters = 20;
  p->dangling = NWGR_DANGLING_LEAK;
}

const char* nwgr_last_error(void) { return nwgr::t_last_error.c_str(); }

int nwgr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int nwgr_pr_graph_create(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                         const uint32_t* degrees, int device, nwgr_pr_graph** out) {
  return guarded([&] {
    NWGR_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    nwgr::Graph* g = nwgr::create_graph(offsets, indices, n, m, degrees, device, nullptr, false);
    *out = static_cast<nwgr_pr_graph*>(g);
  });
}

int nwgr_pr_graph_create_device(const uint32_t* d_offsets, const uint32_t* d_indices, uint32_t n, uint64_t m,
                                const uint32_t* d_degrees, int device, void* stream, nwgr_pr_graph** out) {
  return guarded([&] {
    NWGR_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    nwgr::Graph* g = nwgr::create_graph(d_offsets, d_indices, n, m, d_degrees, device,
                                        static_cast<hipStream_t>(stream), true);
    *out = static_cast<nwgr_pr_graph*>(g);
  });
}

void nwgr_pr_graph_destroy(nwgr_pr_graph* g) {
  if (!g) return;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(g->device);
  delete static_cast<nwgr::Graph*>(g);
  (void)hipSetDevice(prev);
}

uint32_t nwgr_pr_graph_n(const nwgr_pr_graph* g) { return g ? g->n : 0; }
uint64_t nwgr_pr_graph_m(const nwgr_pr_graph* g) { return g ? g->m : 0; }
uint64_t nwgr_pr_graph_device_bytes(const nwgr_pr_graph* g) { return g ? nwgr::bytes_of(*g) : 0; }
double nwgr_pr_graph_setup_ms(const nwgr_pr_graph* g) { return g ? g->setup_ms : 0.0; }

int nwgr_pr_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* d_r_out, nwgr_pr_result* res,
                           void* stream) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && d_r_out && res, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_plain(*g, *p, d_r_out, *res, s);
  });
}

int nwgr_pr_run_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* r_out, nwgr_pr_result* res) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && r_out && res, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = g->own_stream;
    const size_t nb = sizeof(float) * g->n;
    nwgr::DevBuf out, tel, init;
    out.alloc(nb);
    nwgr_pr_params q = *p;
    if (p->teleport) {
      tel.alloc(nb);
      NWGR_HIP(hipMemcpyAsync(tel.p, p->teleport, nb, hipMemcpyHostToDevice, s));
      q.teleport = tel.as<float>();
    }
    if (p->init) {
      init.alloc(nb);
      NWGR_HIP(hipMemcpyAsync(init.p, p->init, nb, hipMemcpyHostToDevice, s));
      q.init = init.as<float>();
    }
    nwgr::run_plain(*g, q, out.as<float>(), *res, s);
    NWGR_HIP(hipMemcpyAsync(r_out, out.p, nb, hipMemcpyDeviceToHost, s));
    NWGR_HIP(hipStreamSynchronize(s));
  });
}

int nwgr_pr_last_profile(const nwgr_pr_graph* g, double* sweep_ms, uint32_t* launches, uint32_t* chunks) {
  return guarded([&] {
    NWGR_REQUIRE(g != nullptr, "null graph");
    if (sweep_ms) *sweep_ms = g->last_profile.chunk_ms;
    if (launches) *launches = g->last_profile.launches;
    if (chunks) *chunks = g->last_profile.chunks;
  });
}

uint32_t nwgr_pr_last_errors(const nwgr_pr_graph* g, double* errs, uint32_t cap) {
  if (!g || !errs) return 0;
  uint32_t k = static_cast<uint32_t>(std::min<size_t>(cap, g->last_errs.size()));
  std::memcpy(errs, g->last_errs.data(), sizeof(double) * k);
  return k;
}

}  // extern "C"
