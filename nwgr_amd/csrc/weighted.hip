// Differentiable weighted PageRank: forward fixed point and its adjoint.
//
// No reference counterpart: this is the weighted form of the same Jacobi
// iteration (experimental/algorithms/page_rank.hpp:450-505) with edge
// weights w (in the caller's in-CSR edge order) and teleport t:
//   P_uv = w_uv / s_u,  s_u = sum of u's out-edge weights,
//   r_v  = (1-d) t_v + d (sum_{u->v} P_uv r_u + D q_v),  D = sum_{s_u = 0} r_u,
//   q = 0 (LEAK), 1/n (UNIFORM), t (TELEPORT).
// For L(r*) with g = dL/dr, the adjoint lambda solves
//   lambda_u = d (s_u > 0 ? sum_v P_uv lambda_v : q . lambda) + g_u,
// a gather over the OUT-edges of u (the transposed CSR, built on the device),
// and
//   dL/dw_uv = d r_u (lambda_v - (P lambda)_u) / s_u,
//   dL/dt_v  = ((1-d) + [TELEPORT] d D) lambda_v.
// Only the converged forward ranks are needed; nothing per sweep is stored.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "nwgr_internal.hpp"

namespace nwgr {

namespace {

constexpr int kT = 256;

inline uint32_t grid_for(uint64_t work) {
  uint64_t g = (work + kT - 1) / kT;
  return static_cast<uint32_t>(std::clamp<uint64_t>(g, 1, 65535u * 4));
}

// dst[k] = src[map[k]]
__global__ void k_gather(const float* __restrict__ src, const uint32_t* __restrict__ map, uint64_t m,
                         float* __restrict__ dst) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride)
    dst[k] = src[map[k]];
}

// Per-edge gradient over the in-CSR (new order): edge k of row v with
// source u = idx[k] gets coef_a[u] * lambda[v] - coef_b[u], written at the
// caller's edge position epos[k].  Units of the forward plan, visited
// statically (no reductions).
struct GradArgs {
  const uint32_t* __restrict__ offs;
  const uint32_t* __restrict__ idx;
  const uint32_t* __restrict__ epos;
  const float* __restrict__ lam;
  const float* __restrict__ ca;
  const float* __restrict__ cb;
  const HubUnit* __restrict__ hub_units;
  float* __restrict__ dw;
  SegTable segs;
};

__global__ __launch_bounds__(kBlock) void k_edge_grad(GradArgs a) {
  // one wavefront per unit of the forward plan
  constexpr uint32_t kWavesPerBlock = kBlock / kWave;
  const uint32_t wl = threadIdx.x & (kWave - 1);
  for (uint32_t cur = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; cur < a.segs.total_units;
       cur += gridDim.x * kWavesPerBlock) {
    Seg sg = a.segs.s[0];
#pragma unroll
    for (int i = 1; i < kMaxSegs; ++i)
      if (i < static_cast<int>(a.segs.nseg) && cur >= a.segs.s[i].unit_lo) sg = a.segs.s[i];
    const uint32_t ul = cur - sg.unit_lo;
    if (sg.G == 0) {
      const HubUnit hu = a.hub_units[ul];
      const float lv = a.lam[hu.row];
      for (uint32_t k = hu.e_lo + wl; k < hu.e_hi; k += kWave) {
        const uint32_t u = a.idx[k];
        a.dw[a.epos[k]] = fmaf(a.ca[u], lv, -a.cb[u]);
      }
    } else {
      const uint32_t G = sg.G;
      const uint32_t lg = __builtin_ctz(G);
      const uint32_t lane = wl & (G - 1);
      const uint32_t end = min(sg.row_lo + (ul + 1) * sg.rows_per_unit, sg.row_hi);
      for (uint32_t row = sg.row_lo + ul * sg.rows_per_unit + (wl >> lg); row < end; row += (kWave >> lg)) {
        const uint32_t b = a.offs[row], e = a.offs[row + 1];
        const float lv = a.lam[row];
        for (uint32_t k = b + lane; k < e; k += G) {
          const uint32_t u = a.idx[k];
          a.dw[a.epos[k]] = fmaf(a.ca[u], lv, -a.cb[u]);
        }
      }
    }
  }
}

// partial sums of r over rows with s == 0 (dangling mass), one per block
__global__ __launch_bounds__(kT) void k_dangling_partial(const float* __restrict__ r, const float* __restrict__ s,
                                                         uint32_t n, double* part) {
  __shared__ double sh[kT / 64];
  double acc = 0;
  for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT)
    if (!(s[i] > 0.f)) acc += static_cast<double>(r[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < kT / 64; ++i) t += sh[i];
    part[blockIdx.x] = t;
  }
}

void prepare_weights(Graph& g, const float* d_w, hipStream_t s) {
  ensure_transpose(g, s);
  const uint64_t m = g.m;
  if (!g.w_in.p) {
    g.w_in.alloc(sizeof(float) * (m + kEdgePad));
    g.w_out.alloc(sizeof(float) * (m + kEdgePad));
    NWGR_HIP(hipMemsetAsync(g.w_in.p, 0, g.w_in.bytes, s));
    NWGR_HIP(hipMemsetAsync(g.w_out.p, 0, g.w_out.bytes, s));
    g.s.alloc(sizeof(float) * g.n);
  }
  if (m) {
    k_gather<<<grid_for(m), kT, 0, s>>>(d_w, g.epos.as<uint32_t>(), m, g.w_in.as<float>());
    NWGR_HIP(hipGetLastError());
    k_gather<<<grid_for(m), kT, 0, s>>>(g.w_in.as<float>(), g.t_pos.as<uint32_t>(), m, g.w_out.as<float>());
    NWGR_HIP(hipGetLastError());
  }
  // s_u = sum over the out-row u (deterministic segmented reduction)
  size_t tmp = 0;
  const uint32_t* ob = g.t_offs.as<uint32_t>();
  NWGR_HIP(hipcub::DeviceSegmentedReduce::Sum(nullptr, tmp, g.w_out.as<float>(), g.s.as<float>(),
                                              static_cast<int>(g.n), ob, ob + 1, s));
  DevBuf t;
  t.alloc(tmp);
  NWGR_HIP(hipcub::DeviceSegmentedReduce::Sum(t.p, tmp, g.w_out.as<float>(), g.s.as<float>(),
                                              static_cast<int>(g.n), ob, ob + 1, s));
}

SolveSpec forward_spec(Graph& g, const nwgr_pr_params& p) {
  SolveSpec sp;
  sp.kind = kWeighted;
  sp.plan = &g.fwd;
  sp.offs = g.offs.as<uint32_t>();
  sp.idx = g.idx.as<uint32_t>();
  sp.w = g.w_in.as<float>();
  sp.div = g.s.as<float>();
  sp.init = p.init;
  sp.tel_src = p.teleport;
  return sp;
}

// Engine choice for one orientation; the tiled engine reads the weights in
// its own order.
void attach_tiled(Graph& g, SolveSpec& sp, bool adj, const float* weights, hipStream_t s) {
  Tiled* t = select_tiled(g, adj, s);
  if (!t) return;
  tiled_weights(g, *t, weights, s);
  sp.tiled = t;
  sp.tw = t->w.as<float>();
}

SolveSpec adjoint_spec(Graph& g, const nwgr_pr_params& p) {
  SolveSpec sp;
  sp.kind = kAdjoint;
  sp.plan = &g.adj;
  sp.offs = g.t_offs.as<uint32_t>();
  sp.idx = g.t_idx.as<uint32_t>();
  sp.w = g.w_out.as<float>();
  sp.div = g.s.as<float>();
  sp.gsrc = g.gsrc.as<float>();
  sp.tel_src = p.teleport;
  return sp;
}

}  // namespace

void run_weighted_forward(Graph& g, const nwgr_pr_params& p, const float* d_w, float* d_out, nwgr_pr_result& res,
                          hipStream_t s) {
  validate_params(p);
  NWGR_REQUIRE(d_w != nullptr || g.m == 0, "weights are null");
  if (!g.vflag.p) g.vflag.alloc(sizeof(uint32_t));
  check_vectors(p, g.n, g.vflag.as<uint32_t>(), s);
  prepare_weights(g, d_w, s);
  SolveSpec sp = forward_spec(g, p);
  attach_tiled(g, sp, false, g.w_in.as<float>(), s);
  solve(g, sp, p, res, s);
  const float* fin = g.r[res.iters & 1].as<float>();
  unpermute(g, fin, d_out, 1.0f, s);
  NWGR_HIP(hipStreamSynchronize(s));
}

void run_weighted_adjoint(Graph& g, const nwgr_pr_params& p, const float* d_w, const float* d_r,
                          const float* d_grad_r, float* d_dw, float* d_dt, nwgr_pr_result& res, hipStream_t s) {
  validate_params(p);
  NWGR_REQUIRE(d_r && d_grad_r && d_dt && (d_dw || g.m == 0) && (d_w || g.m == 0), "null argument");
  if (!g.vflag.p) g.vflag.alloc(sizeof(uint32_t));
  check_vectors(p, g.n, g.vflag.as<uint32_t>(), s);
  prepare_weights(g, d_w, s);
  const uint32_t n = g.n;
  if (!g.gsrc.p) {
    g.gsrc.alloc(sizeof(float) * n);
    g.rsave.alloc(sizeof(float) * n);
    g.coef_a.alloc(sizeof(float) * n);
    g.coef_b.alloc(sizeof(float) * n);
    g.drop_graphs(kAdjoint);
  }
  k_gather<<<grid_for(n), kT, 0, s>>>(d_grad_r, g.perm.as<uint32_t>(), n, g.gsrc.as<float>());
  NWGR_HIP(hipGetLastError());
  k_gather<<<grid_for(n), kT, 0, s>>>(d_r, g.perm.as<uint32_t>(), n, g.rsave.as<float>());
  NWGR_HIP(hipGetLastError());
  SolveSpec sa = adjoint_spec(g, p);
  attach_tiled(g, sa, true, g.w_out.as<float>(), s);
  solve(g, sa, p, res, s);
  const float* lam = g.r[res.iters & 1].as<float>();
  // gradient coefficients from the converged lambda, then per-edge gradients
  run_coef_pass(g, sa, p, lam, s);
  if (g.m) {
    GradArgs ga{};
    ga.offs = g.offs.as<uint32_t>();
    ga.idx = g.idx.as<uint32_t>();
    ga.epos = g.epos.as<uint32_t>();
    ga.lam = lam;
    ga.ca = g.coef_a.as<float>();
    ga.cb = g.coef_b.as<float>();
    ga.hub_units = g.fwd.hub_units.as<HubUnit>();
    ga.dw = d_dw;
    ga.segs = g.fwd.segs;
    const uint32_t waves = kBlock / kWave;
    k_edge_grad<<<std::clamp<uint32_t>((g.fwd.segs.total_units + waves - 1) / waves, 1, 8192), kBlock, 0, s>>>(ga);
    NWGR_HIP(hipGetLastError());
  }
  // dL/dt = ((1-d) + [TELEPORT] d D) lambda
  double D = 0;
  if (p.dangling == NWGR_DANGLING_TELEPORT) {
    const uint32_t nb = std::min<uint32_t>(grid_for(n), 1024);
    DevBuf part;
    part.alloc(sizeof(double) * nb);
    k_dangling_partial<<<nb, kT, 0, s>>>(g.rsave.as<float>(), g.s.as<float>(), n, part.as<double>());
    NWGR_HIP(hipGetLastError());
    std::vector<double> ph(nb);
    NWGR_HIP(hipMemcpyAsync(ph.data(), part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    NWGR_HIP(hipStreamSynchronize(s));
    for (double x : ph) D += x;
  }
  const float scale = static_cast<float>((1.0 - p.damping) + (p.dangling == NWGR_DANGLING_TELEPORT ? p.damping * D : 0.0));
  unpermute(g, lam, d_dt, scale, s);
  NWGR_HIP(hipStreamSynchronize(s));
}

}  // namespace nwgr

struct nwgr_pr_graph : nwgr::Graph {};

extern "C" {

int nwgr_wpr_forward_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, const float* d_w, float* d_r_out,
                                nwgr_pr_result* res, void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(d_r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_weighted_forward(*g, *p, d_w, d_r_out, *res, s);
  });
}

int nwgr_wpr_adjoint_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, const float* d_w, const float* d_r,
                                const float* d_grad_r, float* d_grad_w, float* d_grad_t, nwgr_pr_result* res,
                                void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_weighted_adjoint(*g, *p, d_w, d_r, d_grad_r, d_grad_w, d_grad_t, *res, s);
  });
}

}  // extern "C"
