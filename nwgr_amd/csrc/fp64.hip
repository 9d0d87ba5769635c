// FP64 PageRank on the device: the Jacobi iteration of page_rank_v10 /
// page_rank_v1 (experimental/algorithms/page_rank.hpp:137-174, 450-505) with
// double ranks, contributions and sums -- the reference's Real = double
// (its PageRank test computes in double, test/page_rank_test.cpp:66,93).
//
// It runs on the handle's relabelled in-CSR and degree-binned work plan: hub
// rows in chunks of one wavefront each, added in chunk order by the chunk
// that finishes last; the other rows by groups of G lanes (G = 64 .. 1 from
// the row's degree band), one unit of rows per wavefront from a work queue.
// Divisors are the caller's integer degrees, converted exactly.  The stop
// protocol is the FP32 engine's: every workgroup writes its (error, dangling)
// pair; every workgroup of the next launch adds all pairs in one fixed order,
// so all take the same decision; a launch after the stop returns at once.
// Bound: the latency of the random 8-byte gathers (no LDS staging).

#include <algorithm>
#include <cmath>
#include <cstring>

#include "nwgr_internal.hpp"

namespace nwgr {
namespace {

constexpr uint32_t kT64 = 1024;  // sweep workgroup: one per CU, 16 wavefronts
constexpr uint32_t kLeak = 0, kUniform = 1, kTeleport = 2;

struct Args64 {
  const uint32_t* __restrict__ offs;
  const uint32_t* __restrict__ idx;
  const uint32_t* __restrict__ deg;   // divisors, new order
  const double* __restrict__ tel;     // permuted teleport, or null (uniform)
  const double* __restrict__ r_in;
  const double* __restrict__ c_in;
  double* __restrict__ r_out;
  double* __restrict__ c_out;
  const double* __restrict__ part_in;  // [grid][2] of the previous launch (or init)
  double* __restrict__ part_out;
  State* st;
  HostFlag* hflag;
  double* hist;
  uint32_t hist_cap;
  const HubUnit* __restrict__ hub_units;
  double* hub_part;
  uint32_t* hub_ticket;
  SegTable segs;
  uint32_t p, n, max_iters, dangling;
  double d, base_u, one_minus_d, un, tol;
};

struct Init64 {
  const uint32_t* __restrict__ perm;
  const uint32_t* __restrict__ deg;
  const float* __restrict__ init;     // caller order, or null
  const float* __restrict__ tel_src;  // caller order, or null
  double* __restrict__ tel;
  double* __restrict__ r0;
  double* __restrict__ c0;
  double* part0;
  State* st;
  uint32_t n;
  double un;
};

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__global__ __launch_bounds__(kT64) void k_init64(Init64 a) {
  __shared__ double s_red[kT64 / 64];
  double dang = 0;
  for (uint32_t i = blockIdx.x * kT64 + threadIdx.x; i < a.n; i += gridDim.x * kT64) {
    const uint32_t src = a.perm[i];
    if (a.tel_src) a.tel[i] = static_cast<double>(a.tel_src[src]);
    const double r = a.init ? static_cast<double>(a.init[src]) : a.un;
    const uint32_t dg = a.deg[i];
    a.r0[i] = r;
    a.c0[i] = r / static_cast<double>(dg);  // +inf for degree 0, never gathered
    if (dg == 0) dang += r;
  }
  dang = wave_sum64(dang);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x / 64] = dang;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (uint32_t w = 0; w < kT64 / 64; ++w) t += s_red[w];
    a.part0[2 * blockIdx.x] = INFINITY;  // no "previous error" before sweep 1
    a.part0[2 * blockIdx.x + 1] = t;
    if (blockIdx.x == 0) {
      a.st->ctr[0] = a.st->ctr[1] = 0;
      a.st->work[0] = a.st->work[1] = 0;
      a.st->done = 0;
      a.st->iters = 0;
      a.st->err = INFINITY;
      a.st->last_e = 0;
    }
  }
}

// Row epilogue in double: r' = (1-d) t + d (z + D u), as orc_pr_jacobi_f64.
__device__ __forceinline__ void finish64(const Args64& a, uint32_t row, double z, double D, double& err,
                                         double& dang) {
  const double t = a.tel ? a.tel[row] : a.un;
  const double base = a.tel ? a.one_minus_d * t : a.base_u;
  const double u = a.dangling == kUniform ? a.un : (a.dangling == kTeleport ? t : 0.0);
  const double rn = base + a.d * (z + D * u);
  const uint32_t dg = a.deg[row];
  a.r_out[row] = rn;
  a.c_out[row] = rn / static_cast<double>(dg);
  err += fabs(rn - a.r_in[row]);
  if (dg == 0) dang += rn;
}

__global__ __launch_bounds__(kT64) void k_sweep64(Args64 a) {
  __shared__ double s_red[2][kT64 / 64];
  __shared__ uint32_t s_done;
  const uint32_t tid = threadIdx.x, wl = tid & 63, wv = tid / 64;
  if (tid == 0) s_done = a.st->done;
  __syncthreads();
  if (s_done) return;
  // totals of the previous sweep, the same in every workgroup
  double E = 0, D = 0;
  for (uint32_t b = tid; b < gridDim.x; b += kT64) {
    E += a.part_in[2 * b];
    D += a.part_in[2 * b + 1];
  }
  E = wave_sum64(E);
  D = wave_sum64(D);
  if (wl == 0) {
    s_red[0][wv] = E;
    s_red[1][wv] = D;
  }
  __syncthreads();
  E = 0;
  D = 0;
#pragma unroll
  for (uint32_t i = 0; i < kT64 / 64; ++i) {
    E += s_red[0][i];
    D += s_red[1][i];
  }
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
    a.hflag->last_err = E;
    a.hflag->prev_err = done_sweeps >= 2 ? a.st->last_e : 0.0;
    a.st->last_e = E;
    a.hflag->progress = done_sweeps;
  }
  __syncthreads();  // every wave is past its reads of s_red

  uint32_t* queue = &a.st->work[a.p];
  double err = 0, dang = 0;
  const uint32_t total = a.segs.total_units;
  uint32_t cur = 0;
  if (wl == 0) cur = __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  cur = __builtin_amdgcn_readfirstlane(cur);
  while (cur < total) {
    Seg sg = a.segs.s[0];
#pragma unroll
    for (int i = 1; i < kMaxSegs; ++i)
      if (i < static_cast<int>(a.segs.nseg) && cur >= a.segs.s[i].unit_lo) sg = a.segs.s[i];
    const uint32_t ul = cur - sg.unit_lo;
    if (sg.G == 0) {
      // one chunk of a hub row, whole wavefront
      const HubUnit hu = a.hub_units[ul];
      double z = 0;
      for (uint32_t k = hu.e_lo + wl; k < hu.e_hi; k += 64) z += a.c_in[a.idx[k]];
      z = wave_sum64(z);
      if (wl == 0) {
        if (hu.nchunks == 1) {
          finish64(a, hu.row, z, D, err, dang);
        } else {
          __hip_atomic_store(&a.hub_part[ul], z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          // release: the partial is visible to whichever chunk finishes last
          const uint32_t t = __hip_atomic_fetch_add(&a.hub_ticket[hu.hub], 1u, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_AGENT);
          if (t == hu.nchunks - 1) {
            double zz = 0;
            for (uint32_t c = 0; c < hu.nchunks; ++c)
              zz += __hip_atomic_load(&a.hub_part[hu.first + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.hub_ticket[hu.hub], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            finish64(a, hu.row, zz, D, err, dang);
          }
        }
      }
    } else {
      // rows of one degree band, G lanes per row
      const uint32_t lg = __builtin_ctz(sg.G), lane = wl & (sg.G - 1);
      const uint32_t r_lo = sg.row_lo + ul * sg.rows_per_unit;
      const uint32_t r_hi = min(r_lo + sg.rows_per_unit, sg.row_hi);
      for (uint32_t row0 = r_lo; row0 < r_hi; row0 += 64u >> lg) {
        const uint32_t row = row0 + (wl >> lg);
        const bool valid = row < r_hi;
        double z = 0;
        if (valid) {
          const uint32_t e = a.offs[row + 1];
          for (uint32_t k = a.offs[row] + lane; k < e; k += sg.G) z += a.c_in[a.idx[k]];
        }
        for (uint32_t o = sg.G >> 1; o > 0; o >>= 1) z += __shfl_xor(z, o);
        if (valid && lane == 0) finish64(a, row, z, D, err, dang);
      }
    }
    uint32_t nxt = 0;
    if (wl == 0) nxt = __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    cur = __builtin_amdgcn_readfirstlane(nxt);
  }
  err = wave_sum64(err);
  dang = wave_sum64(dang);
  if (wl == 0) {
    s_red[0][wv] = err;
    s_red[1][wv] = dang;
  }
  __syncthreads();
  if (tid == 0) {
    double te = 0, td = 0;
#pragma unroll
    for (uint32_t i = 0; i < kT64 / 64; ++i) {
      te += s_red[0][i];
      td += s_red[1][i];
    }
    a.part_out[2 * blockIdx.x] = te;
    a.part_out[2 * blockIdx.x + 1] = td;
  }
}

__global__ __launch_bounds__(256) void k_unpermute64(const uint32_t* __restrict__ iperm, const double* __restrict__ r,
                                                     uint32_t n, double* __restrict__ out) {
  for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) out[v] = r[iperm[v]];
}

void ensure(DevBuf& b, size_t bytes) {
  if (b.bytes < bytes) b.alloc(bytes);
}

}  // namespace

void run_plain64(Graph& g, const nwgr_pr_params& p, double* d_out, nwgr_pr_result& res, hipStream_t s) {
  validate_params(p);
  if (!g.vflag.p) g.vflag.alloc(sizeof(uint32_t));
  check_vectors(p, g.n, g.vflag.as<uint32_t>(), s);
  NWGR_REQUIRE(!g.fwd.rows.p, "the forward plan lists rows in position order");
  const uint32_t n = g.n;
  for (int b = 0; b < 2; ++b) {
    ensure(g.r64[b], sizeof(double) * n);
    ensure(g.c64[b], sizeof(double) * n);
  }
  if (p.teleport) ensure(g.tel64, sizeof(double) * n);
  ensure(g.hub64, sizeof(double) * std::max<uint32_t>(g.fwd.n_hub_units, 1));
  const uint32_t grid = g.grid;  // the part slots hold at least this many pairs
  std::memset(const_cast<HostFlag*>(g.hflag), 0, sizeof(HostFlag));
  hipEvent_t e0, e1;
  NWGR_HIP(hipEventCreate(&e0));
  NWGR_HIP(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } eg{e0, e1};
  NWGR_HIP(hipEventRecord(e0, s));
  double* part = g.partials.as<double>();
  Init64 ia{};
  ia.perm = g.perm.as<uint32_t>();
  ia.deg = g.degu.as<uint32_t>();
  ia.init = p.init;
  ia.tel_src = p.teleport;
  ia.tel = p.teleport ? g.tel64.as<double>() : nullptr;
  ia.r0 = g.r64[0].as<double>();
  ia.c0 = g.c64[0].as<double>();
  ia.part0 = part;
  ia.st = g.state.as<State>();
  ia.n = n;
  ia.un = 1.0 / n;
  k_init64<<<grid, kT64, 0, s>>>(ia);
  NWGR_HIP(hipGetLastError());

  Args64 a{};
  a.offs = g.offs.as<uint32_t>();
  a.idx = g.idx.as<uint32_t>();
  a.deg = g.degu.as<uint32_t>();
  a.tel = ia.tel;
  a.st = ia.st;
  a.hflag = g.hflag_dev;
  a.hist = g.hist_dev;
  a.hist_cap = g.hist_cap;
  a.hub_units = g.fwd.hub_units.as<HubUnit>();
  a.hub_part = g.hub64.as<double>();
  a.hub_ticket = g.fwd.hub_ticket.as<uint32_t>();
  a.segs = g.fwd.segs;
  a.n = n;
  a.max_iters = p.max_iters;
  a.dangling = p.dangling;
  a.d = p.damping;
  a.one_minus_d = 1.0 - p.damping;
  a.un = 1.0 / n;
  a.base_u = (1.0 - p.damping) / n;
  a.tol = p.tol;
  // launch j performs sweep j (parity j & 1) or finds the stop condition, so
  // launches 1 .. max_iters + 1 always suffice; chunks of K, then a check
  const uint32_t K = std::clamp<uint32_t>(p.chunk ? p.chunk : 8, 1, Graph::kMaxChunk);
  const uint64_t limit = static_cast<uint64_t>(p.max_iters) + 1;
  uint64_t next = 1;
  while (!g.hflag->done) {
    if (next > limit) throw Error(NWGR_EDEVICE, "FP64 solve passed its launch limit without stopping");
    const uint64_t len = std::min<uint64_t>(K, limit - next + 1);
    for (uint64_t i = 0; i < len; ++i, ++next) {
      a.p = static_cast<uint32_t>(next & 1);
      a.r_in = g.r64[a.p ^ 1].as<double>();
      a.c_in = g.c64[a.p ^ 1].as<double>();
      a.r_out = g.r64[a.p].as<double>();
      a.c_out = g.c64[a.p].as<double>();
      a.part_in = part + 2 * g.part_stride * (a.p ^ 1);
      a.part_out = part + 2 * g.part_stride * a.p;
      k_sweep64<<<grid, kT64, 0, s>>>(a);
    }
    NWGR_HIP(hipGetLastError());
    NWGR_HIP(hipStreamSynchronize(s));
  }
  res.iters = g.hflag->iters;
  res.error = g.hflag->err;
  res.converged = (res.iters > 0 && res.error < p.tol) ? 1u : 0u;
  k_unpermute64<<<std::min<uint32_t>((n + 255) / 256, 8192), 256, 0, s>>>(
      g.iperm.as<uint32_t>(), g.r64[res.iters & 1].as<double>(), n, d_out);
  NWGR_HIP(hipGetLastError());
  NWGR_HIP(hipEventRecord(e1, s));
  NWGR_HIP(hipEventSynchronize(e1));
  float ms = 0;
  NWGR_HIP(hipEventElapsedTime(&ms, e0, e1));
  res.ms = ms;
  g.last_engine = 0;
  g.last_profile = SolveProfile{};
  g.last_errs.assign(g.hist_h, g.hist_h + std::min(res.iters, g.hist_cap));  // written by the sweeps (mapped)
}

}  // namespace nwgr

struct nwgr_pr_graph : nwgr::Graph {};

extern "C" {

int nwgr_pr_run_device_f64(nwgr_pr_graph* g, const nwgr_pr_params* p, double* d_r_out, nwgr_pr_result* res,
                           void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(d_r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_plain64(*g, *p, d_r_out, *res, s);
  });
}

int nwgr_pr_run_f64(nwgr_pr_graph* g, const nwgr_pr_params* p, double* r_out, nwgr_pr_result* res) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = g->own_stream;
    const size_t nb = sizeof(float) * g->n;
    nwgr_pr_params q = *p;
    auto staged = [&](nwgr::DevBuf& b) -> float* {
      if (b.bytes < nb) b.alloc(nb);
      return b.as<float>();
    };
    if (p->teleport) {
      q.teleport = staged(g->tel_h);
      NWGR_HIP(hipMemcpyAsync(const_cast<float*>(q.teleport), p->teleport, nb, hipMemcpyHostToDevice, s));
    }
    if (p->init) {
      q.init = staged(g->init_h);
      NWGR_HIP(hipMemcpyAsync(const_cast<float*>(q.init), p->init, nb, hipMemcpyHostToDevice, s));
    }
    // the double ranks go through the handle's output staging buffer
    if (g->out_buf.bytes < 2 * nb) g->out_buf.alloc(2 * nb);
    double* out = static_cast<double*>(g->out_buf.p);
    nwgr::run_plain64(*g, q, out, *res, s);
    nwgr::copy_out(*g, r_out, out, 2 * nb, s);
  });
}

}  // extern "C"
