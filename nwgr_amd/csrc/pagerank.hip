// PageRank sweeps on gfx950 and the C ABI of the plain solve.
//
// One sweep = one launch of k_sweep: a pull gather-SpMV over the in-CSR
// (page_rank.hpp:127-136, per-edge gather :129-131) with the teleport,
// dangling-mass and L1-residual terms fused into the row epilogue.  Jacobi
// semantics as page_rank_v10 (experimental/algorithms/page_rank.hpp:450-505).
// The same kernel, instantiated per solve kind, runs the weighted forward and
// the adjoint (on the transposed CSR) of the differentiable form.
//
// Work distribution: workgroups take units from a queue (one RMW atomic per
// unit); rows are visited in non-increasing degree order so each unit covers
// rows of one degree band, processed by G lanes per row (G = 1..64), or a
// chunk of one hub row processed by the whole workgroup.  Hub chunks are
// combined in chunk order by the last chunk to finish (RMW atomics only, no
// cache flushes).
//
// Cross-workgroup totals (L1 error, dangling mass) never meet inside a kernel:
// every workgroup writes its partial sums, and every workgroup of the NEXT
// sweep re-reduces them in the same fixed order, so all workgroups take the
// same convergence decision without fences.  The launch that finds the
// previous sweep converged (or the cap reached) records the result and
// exits; launches after it return at once.  Sweeps are captured in hipGraph
// chunks; the host sizes the chunks from the observed error decay.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

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
  const float* __restrict__ w;       // per-edge weights (weighted kinds)
  const uint32_t* __restrict__ rows; // position -> row, or null
  const float* __restrict__ div;     // degree or out-weight sum per row
  const float* __restrict__ tel;     // permuted teleport or null (uniform)
  const float* __restrict__ gsrc;    // adjoint source term
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
  uint32_t n;          // vertices; also the zero sentinel index of the gathers
  uint32_t lds_k;      // entries of the gathered vector staged in LDS (<= n + 1)
  uint32_t max_iters;
  uint32_t dangling;
  float d_hi, d_lo;    // damping as an unevaluated float sum d_hi + d_lo
  float base_u;        // (1-d)/n as float (uniform teleport)
  float one_minus_d;   // 1-d as float (explicit teleport)
  float un;            // 1/n as float
  uint32_t nrows;      // rows the sweep stores: n, or nz with quiet rows
  uint32_t nq_dangling;  // quiet rows with divisor 0
  uint32_t quiet;      // 1: rows [nrows, n) are quiet (common rank State::rq)
  uint32_t q_end;      // quiet rows from here on are isolated: their contributions are not stored
  float thr;           // push threshold (kDelta)
  double tol;
  // gradient coefficients (kCoef): a = d r / s, b = a * (P lambda)
  const float* __restrict__ rsave;
  float* __restrict__ coef_a;
  float* __restrict__ coef_b;
};

struct InitArgs {
  const uint32_t* __restrict__ perm;
  const float* __restrict__ div;
  const float* __restrict__ init;      // caller order or null
  const float* __restrict__ tel_src;   // caller order or null
  const float* __restrict__ gsrc;      // adjoint: lambda_0 (new labels)
  float* __restrict__ tel;             // permuted copy (when tel_src)
  float* __restrict__ r0;
  float* __restrict__ r1;              // or null: the other iterate buffer gets the initial ranks too
  float* __restrict__ c0;
  double* __restrict__ part0;
  State* st;
  uint32_t n;
  uint32_t adjoint;
  uint32_t delta;      // residual solve: iteration 1's push of the residual res0
  uint32_t dangling;
  float un;
  float thr, d, res0;  // push threshold, damping, initial residual 1 - d (kDelta)
  uint32_t npairs;     // pairs sweep 1 adds: this grid's, then zeros
  uint32_t quiet_from; // rows from here on are quiet (their ranks are not stored)
  float* out;          // caller-order output the stopping gather pass writes (null: none)
  uint32_t c_from;     // contributions stored below this row (the rest are isolated vertices')
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
template <int NT = kBlock>
__device__ __forceinline__ void block_sum2(double& a, double& b, double (*s)[NT / 64]) {
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
  for (int i = 0; i < NT / 64; ++i) {
    ta += s[0][i];
    tb += s[1][i];
  }
  a = ta;
  b = tb;
}

__device__ __forceinline__ uint32_t ld_stream(const uint32_t* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ld_stream(const float* p) { return __builtin_nontemporal_load(p); }

// Contribution of vertex j: from the LDS copy of the hot prefix when j < K,
// else from global memory.
__device__ __forceinline__ float contrib(const float* __restrict__ lds, uint32_t K, const float* __restrict__ c,
                                         uint32_t j) {
  return j < K ? lds[j] : c[j];
}

// Sum over edges k = k0, k0 + step, ... < e of c[idx[k]] (times w[k]).
// Edges are taken U at a time with no bounds branch inside a batch: the U
// index loads are issued together, then the U gathers, so each lane keeps U
// loads in flight.  Loads past `e` read padding (edge arrays carry kEdgePad
// extra entries) and are redirected to the sentinel vertex n, whose entry
// is 0.  The summation order is fixed: batch by batch, in k order.
template <int U, bool W>
__device__ __forceinline__ float gather_sum(const uint32_t* __restrict__ idx, const float* __restrict__ w,
                                            const float* __restrict__ lds, uint32_t K, const float* __restrict__ c,
                                            uint32_t k0, uint32_t e, uint32_t step, uint32_t sentinel) {
  float z = 0.f;
  for (uint32_t k = k0; k < e; k += step * U) {
    uint32_t j[U];
    float wv[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t kk = k + q * step;
      const uint32_t jj = ld_stream(idx + kk);
      j[q] = kk < e ? jj : sentinel;
      if constexpr (W) wv[q] = ld_stream(w + kk);
    }
    float v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) v[q] = contrib(lds, K, c, j[q]);
#pragma unroll
    for (int q = 0; q < U; ++q) z += W ? wv[q] * v[q] : v[q];
  }
  return z;
}

// Per-row operands of the epilogue, loaded before the row's gathers so that
// their latency overlaps the gathers'.
struct RowIn {
  float old;  // previous iterate (rank or lambda)
  float div;  // degree / out-weight sum
  float t;    // teleport entry (forward, adjoint with TELEPORT)
  float g;    // adjoint source term, or the forward rank (coef)
};

template <uint32_t KIND>
__device__ __forceinline__ RowIn load_row(const SweepArgs& a, uint32_t row) {
  RowIn in;
  in.div = a.div[row];
  in.old = KIND == kCoef ? 0.f : a.r_in[row];
  in.t = a.tel ? a.tel[row] : a.un;
  if constexpr (KIND == kAdjoint) in.g = a.gsrc[row];
  else if constexpr (KIND == kCoef) in.g = a.rsave[row];
  else in.g = 0.f;
  return in;
}

// Row epilogue.
//  forward kinds: teleport + dangling + damping, residual, next contribution
//  adjoint: lambda' = d (s > 0 ? z/s : Q) + g, residual, Q partial of lambda'
//  delta: residual from the row's deltas, then the next push
//  coef: gradient coefficients from P lambda = z/s
template <uint32_t KIND>
__device__ __forceinline__ void finish_row(const SweepArgs& a, uint32_t row, float z, float Df, const RowIn& in,
                                           double& err, double& dang) {
  if constexpr (KIND == kPlain || KIND == kWeighted) {
    const float base = a.tel ? a.one_minus_d * in.t : a.base_u;
    const float u = a.dangling == kDanglingUniform ? a.un : (a.dangling == kDanglingTeleport ? in.t : 0.f);
    const float s = fmaf(Df, u, z);
    const float rn = fmaf(a.d_hi, s, fmaf(a.d_lo, s, base));
    a.r_out[row] = rn;
    a.c_out[row] = rn / in.div;  // IEEE division; +inf/NaN for div == 0 is never gathered
    err += fabs(static_cast<double>(rn) - static_cast<double>(in.old));
    if (in.div == 0.f) dang += static_cast<double>(rn);
  } else if constexpr (KIND == kAdjoint) {
    const float x = in.div > 0.f ? z / in.div : Df;
    const float ln = fmaf(a.d_hi, x, fmaf(a.d_lo, x, in.g));
    a.r_out[row] = ln;
    err += fabs(static_cast<double>(ln) - static_cast<double>(in.old));
    if (a.dangling == kDanglingUniform)
      dang += static_cast<double>(ln) * static_cast<double>(a.un);
    else if (a.dangling == kDanglingTeleport)
      dang += static_cast<double>(ln) * static_cast<double>(in.t);
  } else if constexpr (KIND == kDelta) {
    // z is the row's sum of deltas (none negative), the residual of phase 2
    // (experimental/algorithms/page_rank.hpp:677-690); phase 1 of the next
    // iteration (:661-674) pushes it when above the threshold.  `err` counts
    // the pushing rows of non-zero degree (the reference's `changed`).
    const float res = z > 0.f ? z : 0.f;
    const bool push = res > a.thr;
    const bool send = push && in.div > 0.f;
    a.r_out[row] = push ? in.old + res : in.old;
    a.c_out[row] = send ? res * a.d_hi / in.div : 0.f;
    if (send) err += 1.0;
  } else {  // kCoef
    float ca = 0.f, cb = 0.f;
    if (in.div > 0.f) {
      ca = (a.d_hi + a.d_lo) * in.g / in.div;
      cb = ca * (z / in.div);
    }
    a.coef_a[row] = ca;
    a.coef_b[row] = cb;
  }
}

// Rows of one lane-group segment for one wavefront: G lanes per row, P rows
// per lane group in flight together, U entries per lane per batch.  Row
// bounds and epilogue operands of all P rows are loaded first, then the index
// batches, then the gathers; rows longer than U*G entries continue in a loop.
template <uint32_t KIND, int U, int P>
__device__ __forceinline__ void process_rows(const SweepArgs& a, const float* __restrict__ lds, uint32_t K,
                                             const Seg& sg, uint32_t ul, float Df, double& err, double& dang) {
  constexpr bool kW = weighted_kind(KIND);
  const uint32_t wl = threadIdx.x & (kWave - 1);
  const uint32_t G = sg.G;
  const uint32_t lg = __builtin_ctz(G);
  const uint32_t lane = wl & (G - 1);
  const uint32_t per_pass = kWave >> lg;
  const uint32_t pos_end = min(sg.row_lo + (ul + 1) * sg.rows_per_unit, sg.row_hi);
  for (uint32_t pos0 = sg.row_lo + ul * sg.rows_per_unit + (wl >> lg); pos0 < pos_end; pos0 += P * per_pass) {
    uint32_t row[P], b[P], e[P];
    bool valid[P];
    RowIn in[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const uint32_t pos = pos0 + p * per_pass;
      valid[p] = pos < pos_end;
      const uint32_t pc = valid[p] ? pos : pos_end - 1;
      row[p] = a.rows ? a.rows[pc] : pc;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if constexpr (U > 0) {
        b[p] = a.offs[row[p]];
        e[p] = valid[p] ? a.offs[row[p] + 1] : b[p];
      }
      in[p] = load_row<KIND>(a, row[p]);
    }
    float z[P];
#pragma unroll
    for (int p = 0; p < P; ++p) z[p] = 0.f;
    if constexpr (U > 0) {
      uint32_t j[P][U];
      float wv[P][U];
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < U; ++q) {
          const uint32_t kk = b[p] + lane + q * G;
          const uint32_t jj = ld_stream(a.idx + kk);
          j[p][q] = kk < e[p] ? jj : a.n;
          if constexpr (kW) wv[p][q] = ld_stream(a.w + kk);
        }
      float v[P][U];
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < U; ++q) v[p][q] = contrib(lds, K, a.c_in, j[p][q]);
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < U; ++q) z[p] += kW ? wv[p][q] * v[p][q] : v[p][q];
      // the rest of rows longer than one batch
#pragma unroll
      for (int p = 0; p < P; ++p)
        if (b[p] + U * G < e[p])
          z[p] += gather_sum<U, kW>(a.idx, a.w, lds, K, a.c_in, b[p] + lane + U * G, e[p], G, a.n);
#pragma unroll
      for (int p = 0; p < P; ++p)
        for (uint32_t o = G >> 1; o > 0; o >>= 1) z[p] += __shfl_xor(z[p], o);
    }
    if (lane == 0) {
#pragma unroll
      for (int p = 0; p < P; ++p)
        if (valid[p]) finish_row<KIND>(a, row[p], z[p], Df, in[p], err, dang);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_init(InitArgs a) {
  __shared__ double s_red[2][kBlock / 64];
  double dang = 0, sends = 0;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    if (a.delta) {
      // iteration 1, phase 1: every vertex holds the residual 1 - d
      // (experimental/algorithms/page_rank.hpp:645-651) and pushes it when
      // it exceeds the threshold; ranks start at 0
      const float dg = a.div[i];
      const bool push = a.res0 > a.thr;
      const bool send = push && dg > 0.f;
      a.r0[i] = push ? a.res0 : 0.f;
      a.c0[i] = send ? a.res0 * a.d / dg : 0.f;
      if (send) sends += 1.0;
      continue;
    }
    // (the caller-order index only when a caller vector is read)
    const uint32_t src = a.tel_src || a.init ? a.perm[i] : 0u;
    float t = a.un;
    if (a.tel_src) {
      t = a.tel_src[src];
      a.tel[i] = t;
    }
    if (a.adjoint) {
      const float l = a.gsrc[i];
      a.r0[i] = l;
      if (a.dangling == kDanglingUniform)
        dang += static_cast<double>(l) * static_cast<double>(a.un);
      else if (a.dangling == kDanglingTeleport)
        dang += static_cast<double>(l) * static_cast<double>(t);
    } else {
      const float r = a.init ? a.init[src] : a.un;
      const float dg = a.div[i];
      if (i < a.quiet_from) {  // quiet rows' ranks live in the scalar State::rq
        a.r0[i] = r;
        if (a.r1) a.r1[i] = r;
      }
      if (i < a.c_from) a.c0[i] = r / dg;
      if (dg == 0.f) dang += static_cast<double>(r);
    }
  }
  block_sum2(dang, sends, s_red);
  // the pairs past the grid's, up to the count sweep 1 adds, hold nothing
  for (uint32_t b = gridDim.x + blockIdx.x * kBlock + threadIdx.x; b < a.npairs; b += stride) {
    a.part0[2 * static_cast<size_t>(b)] = 0.0;
    a.part0[2 * static_cast<size_t>(b) + 1] = 0.0;
  }
  if (threadIdx.x == 0) {
    // "previous error" seen by sweep 1: none, or the pushes of iteration 1
    a.part0[2 * blockIdx.x] = a.delta ? sends : INFINITY;
    a.part0[2 * blockIdx.x + 1] = dang;
    if (blockIdx.x == 0) {
      a.st->ctr[0] = 0;
      a.st->ctr[1] = 0;
      a.st->work[0] = 0;
      a.st->work[1] = 0;
      a.st->done = 0;
      a.st->iters = 0;
      a.st->err = INFINITY;
      a.st->dang = 0;
      a.st->last_e = 0;
      a.st->rq[0] = a.un;  // quiet rows of the initial iterate (used without an init vector only)
      a.st->rq[1] = a.un;
      a.st->out = a.out;
      a.st->out_quiet_from = a.quiet_from;
      a.st->fin_done = 0;
    }
  }
}

// One sweep.  One workgroup of kSweepThreads per CU: it stages the hot
// prefix of the gathered vector in LDS, then every wavefront takes units from
// the queue on its own until the queue is empty.
template <uint32_t KIND>
__global__ __launch_bounds__(kSweepThreads) void k_sweep(SweepArgs a) {
  constexpr bool kW = weighted_kind(KIND);
  constexpr int kNW = kSweepThreads / kWave;
  extern __shared__ float s_cache[];  // lds_k floats (dynamic)
  __shared__ double s_red[2][kNW];
  __shared__ uint32_t s_done;
  const uint32_t tid = threadIdx.x;
  const uint32_t wl = tid & (kWave - 1);
  const uint32_t wv = tid / kWave;
  float Df = 0.f;
  uint32_t* queue;
  if constexpr (KIND != kCoef) {
    // one decision per workgroup: `done` may be written by workgroup 0 of
    // this very launch (only when every workgroup stops anyway)
    if (tid == 0) s_done = a.st->done;
    __syncthreads();
    if (s_done) return;

    // totals of the previous sweep, identical in every workgroup: a fixed
    // per-lane order, a fixed butterfly, waves added in order
    double E = 0, D = 0;
    for (uint32_t b = tid; b < gridDim.x; b += kSweepThreads) {
      E += a.part_in[2 * b];
      D += a.part_in[2 * b + 1];
    }
    E = wave_sum(E);
    D = wave_sum(D);
    if (wl == 0) {
      s_red[0][wv] = E;
      s_red[1][wv] = D;
    }
    __syncthreads();
    E = 0;
    D = 0;
#pragma unroll
    for (int i = 0; i < kNW; ++i) {
      E += s_red[0][i];
      D += s_red[1][i];
    }
    const uint32_t done_sweeps = a.st->ctr[a.p ^ 1];
    // the residual solve may stop before its first sweep (iteration 1
    // pushed nothing); the others always sweep once
    if (((done_sweeps > 0 || KIND == kDelta) && E < a.tol) || done_sweeps >= a.max_iters) {
      if (blockIdx.x == 0 && tid == 0) {
        a.st->done = 1;
        a.st->iters = done_sweeps;
        a.st->err = E;
        a.st->dang = D;
        if (done_sweeps > 0 && done_sweeps - 1 < a.hist_cap) a.hist[done_sweeps - 1] = E;
        a.hflag->iters = done_sweeps;
        a.hflag->err = E;
        if (done_sweeps == 0) a.hflag->first_err = E;
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
      a.hflag->prev_err = done_sweeps >= 2 ? a.st->last_e : 0.0;
      a.st->last_e = E;
      a.hflag->progress = done_sweeps;
      if (done_sweeps == 0) a.hflag->first_err = E;
    }
    Df = static_cast<float>(D);
    queue = &a.st->work[a.p];
  } else {
    queue = &a.st->work[0];  // zeroed by k_reset_queue before the pass
  }

  // stage the hot prefix [0, K) of the gathered vector (entry n is the zero
  // sentinel, so K <= n + 1 covers it when the graph is small)
  const uint32_t K = a.lds_k;
  for (uint32_t i = tid; i < K; i += kSweepThreads) s_cache[i] = a.c_in[i];
  __syncthreads();

  double err = 0, dang = 0;
  const uint32_t total = a.segs.total_units;
  uint32_t cur = 0, nxt = 0;
  if (wl == 0) cur = __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  cur = __builtin_amdgcn_readfirstlane(cur);
  while (cur < total) {
    // take the next unit now; its latency overlaps this unit's work
    if (wl == 0) nxt = __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // segment of this unit (unrolled: constant indices keep the table in SGPRs)
    Seg sg = a.segs.s[0];
#pragma unroll
    for (int i = 1; i < kMaxSegs; ++i)
      if (i < static_cast<int>(a.segs.nseg) && cur >= a.segs.s[i].unit_lo) sg = a.segs.s[i];
    const uint32_t ul = cur - sg.unit_lo;
    if (sg.G == 0) {
      // one chunk of a hub row, whole wavefront
      const HubUnit hu = a.hub_units[ul];
      float z = gather_sum<8, kW>(a.idx, a.w, s_cache, K, a.c_in, hu.e_lo + wl, hu.e_hi, kWave, a.n);
      z = wave_sum(z);
      if (wl == 0) {
        if (hu.nchunks == 1) {
          finish_row<KIND>(a, hu.row, z, Df, load_row<KIND>(a, hu.row), err, dang);
        } else {
          uint32_t old = __hip_atomic_exchange(&a.hub_part[ul], __float_as_uint(z), __ATOMIC_RELAXED,
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
            finish_row<KIND>(a, hu.row, zz, Df, load_row<KIND>(a, hu.row), err, dang);
          }
        }
      }
    } else if (sg.U == 0) {
      process_rows<KIND, 0, 4>(a, s_cache, K, sg, ul, Df, err, dang);
    } else if (sg.U == 8) {
      process_rows<KIND, 8, 1>(a, s_cache, K, sg, ul, Df, err, dang);
    } else if (sg.P == 2) {
      process_rows<KIND, 2, 2>(a, s_cache, K, sg, ul, Df, err, dang);
    } else {
      process_rows<KIND, 2, 4>(a, s_cache, K, sg, ul, Df, err, dang);
    }
    cur = __builtin_amdgcn_readfirstlane(nxt);  // lane 0 took it; every lane is active here
  }
  if constexpr (KIND != kCoef) {
    err = wave_sum(err);
    dang = wave_sum(dang);
    __syncthreads();  // every wave is past its last read of s_red
    if (wl == 0) {
      s_red[0][wv] = err;
      s_red[1][wv] = dang;
    }
    __syncthreads();
    if (tid == 0) {
      double te = 0, td = 0;
#pragma unroll
      for (int i = 0; i < kNW; ++i) {
        te += s_red[0][i];
        td += s_red[1][i];
      }
      a.part_out[2 * blockIdx.x] = te;
      a.part_out[2 * blockIdx.x + 1] = td;
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled engine.  Per sweep: k_tile reduces every fragment (a run of one row's
// entries inside one source tile and one lane's share of a chunk) to a
// partial sum from an LDS copy of the tile's contributions; k_combine adds
// each row's partials (tiles ascending) and applies the row epilogue.

struct TileArgs {
  const uint16_t* __restrict__ eloc;
  const float* __restrict__ w;  // weights in tiled order (weighted kinds)
  const uint32_t* __restrict__ chunk_sp0;  // each chunk's first ordinal
  const TiledTask* __restrict__ tasks;
  const uint32_t* __restrict__ wg_task;
  const uint32_t* __restrict__ wg_tile0;  // tile of each workgroup's first task (kNoTile: no task)
  const float* __restrict__ c_in;  // padded to whole tiles
  float* __restrict__ partial;
  const double* __restrict__ part_in;
  State* st;
  HostFlag* hflag;
  double* hist;
  uint32_t hist_cap;
  uint32_t nparts;  // partial pairs written by the previous combine pass (or init)
  uint32_t bits;    // tile width 2^bits
  uint32_t p;
  uint32_t max_iters;
  uint32_t check;   // 1: this launch takes the stop decision (the first tile pass of a sweep); 0: it only
                    // returns when the sweep's first pass stopped the solve (later block Gauss-Seidel phases)
  double tol;
  // calibration (null: off): each workgroup's time, 10 ns ticks, at
  // [(done sweeps % kCalibSweeps) * gridDim.x + workgroup]
  uint32_t* __restrict__ wg_time;
};

struct CombineArgs {
  SweepArgs s;  // epilogue operands (r_in, r_out, c_out, div, tel, ...)
  const uint32_t* __restrict__ pair_ptr;
  const uint32_t* __restrict__ rslot;
  const float* __restrict__ partial;
  // the windows' row ranges and slots (pair_ptr and rslot without the big
  // rows), and the big-row bits (null: no big rows)
  const uint32_t* __restrict__ lpp;
  const uint32_t* __restrict__ lslot;
  const uint32_t* __restrict__ heavy;
  // items: big rows [0, nbig), windows [nbig, nbig + nwin), quiet-row blocks
  // after them; workgroup b takes items[ptr[b] .. ptr[b + 1])
  const uint32_t* __restrict__ ptr;
  const uint32_t* __restrict__ items;
  const GatherDesc* __restrict__ desc;  // per list entry
  uint32_t nbig, nwin, nqb;
  // the output permutation (State::out): caller order -> row, and the iterates
  const uint32_t* __restrict__ iperm;
  const float* rp0;
  const float* rp1;
};

// The gather pass of the stopping launch writes the caller-order output
// out[v] = r[iperm[v]], r the final iterate r_pair[st->iters & 1] (rows from
// State::out_quiet_from on hold its quiet rank), when State::out is set and
// the first tile pass after the stop has not yet marked it written.  NT
// threads per workgroup, every workgroup of the grid.
template <uint32_t NT>
__device__ __forceinline__ void finish_output(const CombineArgs& a) {
  const State* st = a.s.st;
  float* out = st->out;
  if (!out || st->fin_done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.s.hflag->finished = 1;
  const uint32_t it = st->iters;
  const float* __restrict__ r = (it & 1) ? a.rp1 : a.rp0;
  const float rq = st->rq[it & 1];
  const uint32_t qf = st->out_quiet_from, n = a.s.n;
  constexpr uint32_t U = 4;  // caller positions in flight per thread
  const uint32_t stride = gridDim.x * NT;
  for (uint32_t v0 = blockIdx.x * NT + threadIdx.x; v0 < n; v0 += U * stride) {
    uint32_t i[U];
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) i[j] = v0 + j * stride < n ? a.iperm[v0 + j * stride] : qf;
    float x[U];
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) x[j] = i[j] < qf ? r[i[j]] : rq;
#pragma unroll
    for (uint32_t j = 0; j < U; ++j)
      if (v0 + j * stride < n) out[v0 + j * stride] = x[j];
  }
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// The loads of one chunk for one lane: its kLaneEdges local ids (and
// weights), plus the ordinal of the chunk's first fragment (the same in every
// lane; the lanes' own first ordinals follow from the start bits).  Issued
// ahead of use (see k_tile).
struct ChunkIn {
  u32x4 e0, e1;
  f32x4 w0, w1, w2, w3;
  uint32_t base;
};

template <bool W>
__device__ __forceinline__ ChunkIn load_chunk(const TileArgs& a, uint32_t c, uint32_t wl) {
  ChunkIn in;
  // chunk storage order (eloc_pos, wgt_pos): each load below reads 1 KiB
  // contiguous over the wavefront
  const uint64_t cb = static_cast<uint64_t>(c) * kChunkEdges;
  in.e0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(a.eloc + cb + wl * 8));
  in.e1 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(a.eloc + cb + kChunkEdges / 2 + wl * 8));
  if constexpr (W) {
    in.w0 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(a.w + cb + wl * 4));
    in.w1 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(a.w + cb + kChunkEdges / 4 + wl * 4));
    in.w2 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(a.w + cb + kChunkEdges / 2 + wl * 4));
    in.w3 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(a.w + cb + 3 * kChunkEdges / 4 + wl * 4));
  }
  in.base = a.chunk_sp0[c];  // one address for the wave
  return in;
}

// DPP moves inside rows of 16 lanes; lanes whose source falls outside the
// row read 0.
constexpr int kDppRowShr = 0x110;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xf, 0xf, false));
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Lane bit of a wave-uniform mask, as a lane predicate held in SGPRs.
__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// Last lane of each DPP row: every group ends there.
constexpr uint64_t kRowLastLanes = 0x8000800080008000ull;

// Per-wavefront LDS window through which a chunk's fragment sums leave in
// coalesced stores.  With a 2^15-entry tile the 16 windows fill the rest of
// the 160 KiB of LDS; a chunk with more fragments passes through it in turns.
constexpr uint32_t kStageW = kTileWindow;
// After the windows, one slot per lane that takes the stores of running sums
// that end no fragment, so every entry's store is unconditional (no lane-mask
// switch per entry).  Never read.
constexpr uint32_t kJunk = kWave;
// Dynamic LDS of k_tile for a tile of 2^bits contributions.
constexpr size_t tile_lds_bytes(uint32_t bits) {
  return sizeof(float) * ((size_t{1} << bits) + kSweepThreads / kWave * kStageW + kJunk);
}
static_assert(tile_lds_bytes(kMaxTileBits) <= 160 * 1024, "k_tile: LDS of one CU");

// Orders one wavefront's LDS accesses around it (no workgroup barrier).  The
// LDS unit executes a wavefront's DS instructions in issue order, so only the
// compiler must be kept from moving accesses across; a memory fence here would
// also wait for the wavefront's outstanding global loads and stores (one
// counter for both on this target), draining the chunk prefetch.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  __asm__ volatile("" ::: "memory");
}

// LDS accesses at absolute LDS byte addresses (32-bit pointers of the LDS
// address space), so a precomputed address costs no further add.
typedef __attribute__((address_space(3))) float lds_f32;

__device__ __forceinline__ float lds_ld(uint32_t b) { return *(const lds_f32*)static_cast<uintptr_t>(b); }

__device__ __forceinline__ void lds_st(uint32_t b, float v) { *(lds_f32*)static_cast<uintptr_t>(b) = v; }

// Word 3 of a raw buffer resource on this target (32-bit elements, no format
// conversion).
constexpr int kRawBufferWord3 = 0x00020000;

// One window of fragment sums held in registers on its way out: lane l holds
// slots l, l + 64, ..., l + 448 of partial[base, base + nw).
struct WindowOut {
  float o[8];
  uint32_t base;
  uint32_t nw;
};

// Reads the window at LDS address stage_b into registers.
__device__ __forceinline__ void read_window(WindowOut& out, uint32_t stage_b, uint32_t wl) {
  static_assert(kStageW <= 8 * kWave, "read_window: eight slots per lane");
  // (opaque per call, so the per-slot offsets fold into the instructions'
  // immediate offsets instead of living in registers across the chunk loop)
  uint32_t l4 = wl * 4u;
  __asm__("" : "+v"(l4));
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) out.o[j] = lds_ld(stage_b + l4 + j * kWave * 4u);
  out.o[7] = lds_ld(stage_b + min(7 * kWave * 4u + l4, (kStageW - 1) * 4u));
}

// Stores a window to partial[base, base + nw) with a fixed count of stores:
// raw buffer stores bounded to nw elements, so the lanes past nw are dropped
// by the bounds check instead of branched around (nw = 0 stores nothing).
__device__ __forceinline__ void store_window(float* partial, const WindowOut& out, uint32_t wl) {
  // (base and nw are wave-uniform; the compiler is told so, or it would
  // build the resource per lane and loop over its distinct values)
  const uint32_t base = __builtin_amdgcn_readfirstlane(out.base);
  const uint32_t nw = __builtin_amdgcn_readfirstlane(out.nw);
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(partial + base, 0, static_cast<int>(nw * 4u), kRawBufferWord3);
  uint32_t l4 = wl * 4u;
  __asm__("" : "+v"(l4));
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j)
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(out.o[j]), rs, static_cast<int>(l4 + j * kWave * 4u), 0, 0);
}

// Waits for every outstanding global load and store of the wavefront.  Loads
// and stores share one counter and may complete out of order, so a wait for
// one load after a store is a wait for both: the chunk loop waits once per
// batch, before it issues the next batch's loads and the previous batch's
// stores, so the wait covers only operations issued a whole batch earlier.
__device__ __forceinline__ void wait_global() {
  // s_waitcnt vmcnt(0), expcnt and lgkmcnt at their maxima (no wait)
  __builtin_amdgcn_s_waitcnt(0x0f70);
}

// Carry of a fragment that runs into a lane from the lanes before it in the
// lane's DPP row: an inclusive segmented scan of the lanes' open tail sums
// (`seen`: a fragment starts in the lane), shifted by one lane (0 in the
// row's first lane).
__device__ __forceinline__ float row_carry(float tail, bool seen) {
  float tv = tail;
  uint32_t tf = seen ? 1u : 0u;
#define NWGR_SCAN_STEP(D)                                 \
  {                                                       \
    const float ov = dpp_f32<kDppRowShr | (D)>(tv);       \
    const uint32_t of = dpp_u32<kDppRowShr | (D)>(tf);    \
    tv = tf ? tv : tv + ov;                               \
    tf |= of;                                             \
  }
  NWGR_SCAN_STEP(1)
  NWGR_SCAN_STEP(2)
  NWGR_SCAN_STEP(4)
  NWGR_SCAN_STEP(8)
#undef NWGR_SCAN_STEP
  return dpp_f32<kDppRowShr | 1>(tv);
}

// Inclusive prefix sum of x over the 64 lanes: Hillis-Steele steps inside
// each 16-lane DPP row, then the totals of the rows before added from lanes
// 15, 31 and 47.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += dpp_u32<kDppRowShr | 1>(x);
  x += dpp_u32<kDppRowShr | 2>(x);
  x += dpp_u32<kDppRowShr | 4>(x);
  x += dpp_u32<kDppRowShr | 8>(x);
  const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), 15));
  const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), 31));
  const uint32_t r2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), 47));
  return x + (lane_bit(0xffffffffffff0000ull) ? r0 : 0u) + (lane_bit(0xffffffff00000000ull) ? r1 : 0u) +
         (lane_bit(0xffff000000000000ull) ? r2 : 0u);
}

// One lane's share of a chunk: 16 entries, gathered from the LDS tile at LDS
// address lds0.
//
// Every per-entry predicate is a wave mask in SGPRs, taken once by a ballot
// (entry i starts a fragment: bit 15 of its id), so the per-entry work is a
// handful of VALU instructions: the tile address, the running fragment sum
// and the fragment's slot are selects and adds on those masks.  Each entry's
// running sum is stored: at its fragment's slot in the wavefront's LDS window
// when the entry ends the fragment, else at the lane's junk slot, so no store
// needs a lane-mask switch.  The window leaves in coalesced stores.  A
// fragment that runs in from the previous lane is completed after a segmented
// scan over the lane's DPP row (groups are cut at row boundaries): the carry
// of the lanes before is added to its in-lane sum in the window, once per lane.
//
// Summation order, fixed: entries of a fragment in order within each lane;
// a fragment spanning lanes adds the lanes' sums left to right.
template <bool W>
__device__ __forceinline__ void tile_lane(const TileArgs& a, uint32_t lds0, uint32_t stage_b, uint32_t junk_b,
                                          uint32_t mask, const ChunkIn& in, uint32_t wl, WindowOut& out) {
  static_assert(kLaneEdges == 16 && kGroupEdges == 16 * kLaneEdges, "tile_lane: 16 entries, 16-lane rows");
  const uint32_t words[8] = {in.e0.x, in.e0.y, in.e0.z, in.e0.w, in.e1.x, in.e1.y, in.e1.z, in.e1.w};
  const float wt[16] = {in.w0.x, in.w0.y, in.w0.z, in.w0.w, in.w1.x, in.w1.y, in.w1.z, in.w1.w,
                        in.w2.x, in.w2.y, in.w2.z, in.w2.w, in.w3.x, in.w3.y, in.w3.z, in.w3.w};
  float v[16];
  uint64_t sm[16];  // lanes whose entry i starts a fragment
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t w = words[k];
    v[2 * k] = lds_ld(lds0 + ((w & mask) << 2));
    v[2 * k + 1] = lds_ld(lds0 + (((w >> 16) & mask) << 2));
    sm[2 * k] = __builtin_amdgcn_ballot_w64((w & 0x8000u) != 0);
    sm[2 * k + 1] = __builtin_amdgcn_ballot_w64(static_cast<int32_t>(w) < 0);
  }
  if constexpr (W) {
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] *= wt[i];
  }
  // the chunk's first ordinal (lane 0, entry 0 starts a fragment); each
  // lane's count of fragment starts and their inclusive prefix over the wave:
  // the chunk's fragment count, and the lane's first fragment's ordinal
  const uint32_t base = __builtin_amdgcn_readfirstlane(in.base);
  uint32_t nst = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) nst += static_cast<uint32_t>(__builtin_popcount(words[k] & 0x80008000u));
  const uint32_t incl = wave_incl_scan(nst);
  const uint32_t cnt = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
  uint64_t seen_m = sm[0];  // lanes where a fragment starts
#pragma unroll
  for (int i = 1; i < 16; ++i) seen_m |= sm[i];
  // entry 15 ends a fragment when the next lane's entry 0 starts one, or at a row end
  const uint64_t end15_m = (sm[0] >> 1) | kRowLastLanes;
  // lanes where the fragment running in from the previous lane ends
  const bool fix = lane_bit(~sm[0] & (seen_m | end15_m));
  const bool seen = lane_bit(seen_m);
  // LDS address of the slot of entry 0's fragment: the starts before this
  // lane, plus one if entry 0 starts a fragment, minus one
  uint32_t q0 = stage_b + (incl - nst) * 4u + ((words[0] >> 13) & 4u) - 4u;
  // (opaque to the compiler: otherwise it re-derives every entry's address
  // from sp0 and base)
  __asm__("" : "+v"(q0));
  if (cnt <= kStageW) {
    uint32_t qa = q0;
    float acc = v[0];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i > 0) {
        qa += lane_bit(sm[i]) ? 4u : 0u;
        acc = lane_bit(sm[i]) ? v[i] : acc + v[i];
      }
      lds_st(lane_bit(i < 15 ? sm[i + 1] : end15_m) ? qa : junk_b, acc);
    }
    const float carry = row_carry(acc, seen);
    if (fix) lds_st(q0, lds_ld(q0) + carry);
    wave_sync();
    read_window(out, stage_b, wl);
    out.base = base;
    out.nw = cnt;
    wave_sync();
  } else {
    // more fragments than the window holds: one pass per window, each
    // keeping the sums whose slots fall inside it; all windows but the last
    // are stored at once
    // (each entry's running sum and slot are formed once, in registers)
    float acc[16];
    uint32_t qs[16];
    acc[0] = v[0];
    qs[0] = q0;
#pragma unroll
    for (int i = 1; i < 16; ++i) {
      qs[i] = qs[i - 1] + (lane_bit(sm[i]) ? 4u : 0u);
      acc[i] = lane_bit(sm[i]) ? v[i] : acc[i - 1] + v[i];
    }
    const float tail = row_carry(acc[15], seen);
    for (uint32_t w0 = 0;; w0 += kStageW) {
      const uint32_t lo = stage_b + w0 * 4u, hi = lo + kStageW * 4u;  // slot addresses of this window
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const bool end = lane_bit(i < 15 ? sm[i + 1] : end15_m);
        lds_st(end && qs[i] >= lo && qs[i] < hi ? qs[i] - w0 * 4u : junk_b, acc[i]);
      }
      if (fix && q0 >= lo && q0 < hi) lds_st(q0 - w0 * 4u, lds_ld(q0 - w0 * 4u) + tail);
      wave_sync();
      read_window(out, stage_b, wl);
      out.base = base + w0;
      out.nw = min(kStageW, cnt - w0);
      wave_sync();
      if (w0 + kStageW >= cnt) break;
      store_window(a.partial, out, wl);
    }
  }
}

// The chunks of one task taken by one wavefront (lo + wv, lo + wv + 16, ...),
// NB per batch, double-buffered.  Each batch starts with one wait for all of
// the wavefront's global operations (its own loads, issued a batch earlier,
// and the previous batch's stores), then issues the next batch's loads --
// unconditionally, with clamped chunk ids -- and the previous batch's stores
// of fragment sums, which waited in registers, and then reduces its chunks.
// A batch past the end of the range is reduced by a branch that leaves the
// loop; its missing chunks store nothing.
template <bool W, int NB>
__device__ __forceinline__ void run_chunks(const TileArgs& a, uint32_t lds0, uint32_t stage_b, uint32_t junk_b,
                                           uint32_t mask, uint32_t first, uint32_t hi, uint32_t wl) {
  constexpr uint32_t S = kSweepThreads / kWave;
  if (first >= hi) return;
  const uint32_t last = hi - 1;
  ChunkIn A[NB], B[NB];
  WindowOut P[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) P[j] = WindowOut{{}, 0u, 0u};
#pragma unroll
  for (int j = 0; j < NB; ++j) A[j] = load_chunk<W>(a, min(first + j * S, last), wl);
  for (uint32_t c = first;;) {
    wait_global();
#pragma unroll
    for (int j = 0; j < NB; ++j) B[j] = load_chunk<W>(a, min(c + (NB + j) * S, last), wl);
#pragma unroll
    for (int j = 0; j < NB; ++j) store_window(a.partial, P[j], wl);
    if (c + (NB - 1) * S >= hi) {
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        P[j].nw = 0;
        if (c + j * S < hi) tile_lane<W>(a, lds0, stage_b, junk_b, mask, A[j], wl, P[j]);
      }
      break;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) tile_lane<W>(a, lds0, stage_b, junk_b, mask, A[j], wl, P[j]);
    c += NB * S;
    if (c >= hi) break;
    wait_global();
#pragma unroll
    for (int j = 0; j < NB; ++j) A[j] = load_chunk<W>(a, min(c + (NB + j) * S, last), wl);
#pragma unroll
    for (int j = 0; j < NB; ++j) store_window(a.partial, P[j], wl);
    if (c + (NB - 1) * S >= hi) {
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        P[j].nw = 0;
        if (c + j * S < hi) tile_lane<W>(a, lds0, stage_b, junk_b, mask, B[j], wl, P[j]);
      }
      break;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) tile_lane<W>(a, lds0, stage_b, junk_b, mask, B[j], wl, P[j]);
    c += NB * S;
    if (c >= hi) break;
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) store_window(a.partial, P[j], wl);
}

// Tile copy into LDS: all J loads of a thread are issued before its LDS
// stores.  The source vector is padded to whole tiles, so no bounds checks.
template <int J>
__device__ __forceinline__ void fill_tile(const float4* __restrict__ src, float4* __restrict__ dst, uint32_t tid) {
  float4 v[J];
#pragma unroll
  for (int j = 0; j < J; ++j) v[j] = src[tid + j * kSweepThreads];
#pragma unroll
  for (int j = 0; j < J; ++j) dst[tid + j * kSweepThreads] = v[j];
}

template <uint32_t KIND>
__global__ __launch_bounds__(kSweepThreads) void k_tile(TileArgs a) {
  constexpr bool kW = weighted_kind(KIND);
  constexpr int kNW = kSweepThreads / kWave;
  // All LDS is dynamic, so the tile starts at address 0 (its gathers need no
  // base) and is 16-byte aligned for the float4 fill: 2^bits floats of tile,
  // then kStageW floats per wavefront, then the kJunk slots.  The convergence
  // reduction runs before the first tile and borrows the start of the stage
  // area.
  extern __shared__ __align__(16) float s_tile[];
  double(*s_red)[kNW] = reinterpret_cast<double(*)[kNW]>(s_tile + (1u << a.bits));
  uint32_t* s_flag = reinterpret_cast<uint32_t*>(s_red + 2);
  const uint32_t tid = threadIdx.x;
  const uint32_t wl = tid & (kWave - 1);
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid / kWave);
  // consecutive task ranges (mostly the same tile) go to workgroups of one
  // XCD, so a tile's contributions are fetched into that XCD's L2 once
  const uint32_t G = gridDim.x;
  const uint32_t vb = G % kXcds == 0 ? (blockIdx.x % kXcds) * (G / kXcds) + blockIdx.x / kXcds : blockIdx.x;
  // the task range, the stop flag and the previous sweep's pairs are loaded
  // together: one memory round trip before the first tile, not three
  const uint32_t t_lo = a.wg_task[vb], t_hi = a.wg_task[vb + 1];
  // the first tile's contributions are loaded now, in flight during the stop
  // test, and written to LDS after it
  const uint32_t mask = (1u << a.bits) - 1;
  const uint32_t nq = (mask + 1) / 4;  // float4 per tile
  constexpr uint32_t kJ = kMaxTile / 4 / kSweepThreads;
  const uint32_t tile0 = a.wg_tile0[vb];
  const uint64_t t_start = a.wg_time ? wall_clock64() : 0;
  uint32_t swept = 0;  // sweeps done before this one (calibration record)
  float4 f0[kJ];
  if (tile0 != kNoTile) {
    const float4* src = reinterpret_cast<const float4*>(a.c_in + (static_cast<size_t>(tile0) << a.bits));
#pragma unroll
    for (uint32_t j = 0; j < kJ; ++j)
      f0[j] = tid + j * kSweepThreads < nq ? src[tid + j * kSweepThreads] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (KIND != kCoef && !a.check) {
    // a later phase of a block Gauss-Seidel sweep: the sweep's first tile
    // pass took the stop decision
    if (tid == 0) *s_flag = a.st->done;
    __syncthreads();
    if (*s_flag) return;
  } else if constexpr (KIND != kCoef) {
    // totals of the previous sweep, identical in every workgroup: a fixed
    // per-thread order (four pairs in flight), a fixed butterfly, waves in order
    double E = 0, D = 0;
    const double2* __restrict__ pin = reinterpret_cast<const double2*>(a.part_in);
    for (uint32_t b0 = tid; b0 < a.nparts; b0 += 4 * kSweepThreads) {
      double2 v[4];
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) v[j] = b0 + j * kSweepThreads < a.nparts ? pin[b0 + j * kSweepThreads]
                                                                                : make_double2(0.0, 0.0);
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        E += v[j].x;
        D += v[j].y;
      }
    }
    if (tid == 0) {
      const uint32_t dn = a.st->done;
      *s_flag = dn;
      // a launch past the stop: the stopping launch's gather pass has written
      // the output (workgroup 0 of the stopping launch reads 0 here: it sets
      // `done` itself, later)
      if (dn && blockIdx.x == 0) a.st->fin_done = 1;
    }
    __syncthreads();
    if (*s_flag) return;
    E = wave_sum(E);
    D = wave_sum(D);
    if (wl == 0) {
      s_red[0][wv] = E;
      s_red[1][wv] = D;
    }
    __syncthreads();
    E = 0;
    D = 0;
#pragma unroll
    for (int i = 0; i < kNW; ++i) {
      E += s_red[0][i];
      D += s_red[1][i];
    }
    const uint32_t done_sweeps = a.st->ctr[a.p ^ 1];
    // the residual solve may stop before its first sweep (iteration 1
    // pushed nothing); the others always sweep once
    if (((done_sweeps > 0 || KIND == kDelta) && E < a.tol) || done_sweeps >= a.max_iters) {
      if (blockIdx.x == 0 && tid == 0) {
        a.st->done = 1;
        a.st->iters = done_sweeps;
        a.st->err = E;
        a.st->dang = D;
        if (done_sweeps > 0 && done_sweeps - 1 < a.hist_cap) a.hist[done_sweeps - 1] = E;
        a.hflag->iters = done_sweeps;
        a.hflag->err = E;
        if (done_sweeps == 0) a.hflag->first_err = E;
        a.hflag->done = 1;
      }
      return;
    }
    swept = done_sweeps;
    if (blockIdx.x == 0 && tid == 0) {
      a.st->ctr[a.p] = done_sweeps + 1;
      a.st->dsum = D;  // read by this sweep's combine pass
      if (done_sweeps > 0 && done_sweeps - 1 < a.hist_cap) a.hist[done_sweeps - 1] = E;
      a.hflag->last_err = E;
      a.hflag->prev_err = done_sweeps >= 2 ? a.st->last_e : 0.0;
      a.st->last_e = E;
      a.hflag->progress = done_sweeps;
      if (done_sweeps == 0) a.hflag->first_err = E;
    }
  }

  // LDS address of the tile (held in a register: a constant LDS base would be
  // added again at every access)
  uint32_t lds0 = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lds_f32*)s_tile));
  __asm__("" : "+s"(lds0));
  if (tile0 != kNoTile) {
    float4* dst = reinterpret_cast<float4*>(s_tile);
#pragma unroll
    for (uint32_t j = 0; j < kJ; ++j)
      if (tid + j * kSweepThreads < nq) dst[tid + j * kSweepThreads] = f0[j];
  }
  for (uint32_t ti = t_lo; ti < t_hi; ++ti) {
    const TiledTask task = a.tasks[ti];
    if (ti > t_lo) {
      __syncthreads();  // every wave is done with the previous tile
      const float4* src = reinterpret_cast<const float4*>(a.c_in + (static_cast<size_t>(task.tile) << a.bits));
      float4* dst = reinterpret_cast<float4*>(s_tile);
      if (nq == kMaxTile / 4)
        fill_tile<kMaxTile / 4 / kSweepThreads>(src, dst, tid);
      else if (nq == kMaxTile / 8)
        fill_tile<kMaxTile / 8 / kSweepThreads>(src, dst, tid);
      else
        fill_tile<kMaxTile / 16 / kSweepThreads>(src, dst, tid);
    }
    __syncthreads();
    // one chunk per batch: two per batch held twice the registers and ran
    // slower (R-MAT s22: 61.0 against 58.5 us per launch)
    run_chunks<kW, 1>(a, lds0, lds0 + (mask + 1 + wv * kStageW) * 4u, lds0 + (mask + 1 + kNW * kStageW + wl) * 4u,
                      mask, task.chunk_lo + wv, task.chunk_hi, wl);
  }
  if (a.wg_time) {
    __syncthreads();  // every wave is done
    if (tid == 0) a.wg_time[(swept % kCalibSweeps) * G + vb] = static_cast<uint32_t>(wall_clock64() - t_start);
  }
}

// The quiet rows' common rank (plain kind, rows [nrows, n) without entries):
// the row epilogue with z = 0 (the same operations, so the value a stored row
// without entries gets), computed alike by every workgroup of NT threads.
// Their contributions are stored (one streaming pass); their ranks are not.
// Workgroup 0 adds their residual and dangling mass as count x value.
template <uint32_t KIND, uint32_t NT>
__device__ __forceinline__ void quiet_rows(const SweepArgs& a, float Df, double& err, double& dang) {
  if constexpr (KIND == kPlain) {
    if (a.quiet) {
      const float u = a.dangling == kDanglingUniform ? a.un : 0.f;
      const float sq = fmaf(Df, u, 0.f);
      const float rn = fmaf(a.d_hi, sq, fmaf(a.d_lo, sq, a.base_u));
      for (uint32_t v = a.nrows + blockIdx.x * NT + threadIdx.x; v < a.q_end; v += gridDim.x * NT)
        a.c_out[v] = rn / a.div[v];
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        const float ro = a.st->rq[a.p ^ 1];
        a.st->rq[a.p] = rn;
        err += static_cast<double>(a.n - a.nrows) * fabs(static_cast<double>(rn) - static_cast<double>(ro));
        dang += static_cast<double>(a.nq_dangling) * static_cast<double>(rn);
      }
    }
  }
}

// Gather form of the second pass.  One wave of resident workgroups, each
// working through its own list of items (big rows, windows, quiet-row
// blocks), dealt on the host by estimated cost so that the lists end
// together.  A fixed list per workgroup keeps every (error, dangling) pair --
// one per workgroup -- the same from run to run.
constexpr uint32_t kCombineWaves = kCombineThreads / kWave;

// Sum over the 64 lanes without LDS traffic: an inclusive scan inside each
// 16-lane DPP row, then the four row totals (lanes 15, 31, 47, 63) added in a
// fixed order.  Every lane receives the total.
__device__ __forceinline__ float wave_total(float v) {
  v += dpp_f32<kDppRowShr | 1>(v);
  v += dpp_f32<kDppRowShr | 2>(v);
  v += dpp_f32<kDppRowShr | 4>(v);
  v += dpp_f32<kDppRowShr | 8>(v);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 47));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  return (r0 + r1) + (r2 + r3);
}
// wavefront-summed rows of one window (more than kWaveRowMin partials each)
constexpr uint32_t kMaxWaveRows = kCombineStage / (kWaveRowMin + 1) + 1;

// Workgroup barrier for LDS only: waits for the wavefront's LDS operations,
// not its global stores (a __syncthreads fence would hold every wavefront
// until its rank and contribution stores were performed), and keeps the
// compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt at their maxima (no wait)
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}

struct CombineLds {
  float buf[kCombineStage];  // the window's partials, row by row
  uint2 hl[kMaxWaveRows];    // wave rows: (first partial - window's first, count)
  float z[kMaxWaveRows];     // sums of the wave rows, in list order
  float zw[kCombineWaves];   // a big row's per-wavefront sums
  double red[2][kCombineWaves];
  uint32_t nh, flag;
};

// A row with more than kBigRowMin partials: every thread adds a strided share
// (eight loads in flight), wavefronts by butterfly, then their sums in order.
template <uint32_t KIND>
__device__ __forceinline__ void combine_big(const CombineArgs& a, const GatherDesc& d, float Df, CombineLds& L,
                                            double& err, double& dang) {
  constexpr uint32_t U = 8;
  const uint32_t tid = threadIdx.x, wl = tid & (kWave - 1), wv = tid / kWave;
  const uint32_t row = d.x, b = d.y, e = d.z;
  float z = 0.f;
  for (uint32_t k0 = b + tid; k0 < e; k0 += U * kCombineThreads) {
    uint32_t sl[U];
    float v[U];
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) {
      const uint32_t k = k0 + j * kCombineThreads;
      sl[j] = k < e ? a.rslot[k] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) v[j] = k0 + j * kCombineThreads < e ? a.partial[sl[j]] : 0.f;
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) z += v[j];
  }
  z = wave_total(z);
  if (wl == 0) L.zw[wv] = z;
  lds_barrier();
  if (tid == 0) {
    float t = 0.f;
#pragma unroll
    for (uint32_t i = 0; i < kCombineWaves; ++i) t += L.zw[i];
    finish_row<KIND>(a.s, row, t, Df, load_row<KIND>(a.s, row), err, dang);
  }
}

// One window (rows [win[k], win[k + 1]), kRowsPerThread per thread: rows
// w0 + tid + j * kCombineThreads): its partials (at most kCombineStage, big
// rows left out) are staged in LDS in one step -- all index loads of a thread
// issued together, then all gathers, in two halves -- then a thread adds each
// of its rows of at most kWaveRowMin partials in tile order, a wavefront each
// longer row (lanes strided, then a DPP total).
template <uint32_t KIND>
__device__ __forceinline__ void combine_window(const CombineArgs& a, const GatherDesc& d, float Df, CombineLds& L,
                                               double& err, double& dang) {
  constexpr uint32_t U = kCombineStage / kCombineThreads;
  constexpr uint32_t R = kRowsPerThread;
  const uint32_t tid = threadIdx.x, wl = tid & (kWave - 1), wv = tid / kWave;
  const uint32_t w0 = d.x, w1 = d.y, b0 = d.z, e0 = d.w;
  const uint32_t* __restrict__ lp = a.lpp;
  uint32_t b[R], e[R];
  uint32_t wave_m = 0, light_m = 0;  // bit j: row j of this thread is a wave row / a light row
  uint32_t hj[R];                     // a wave row's place in the list
#pragma unroll
  for (uint32_t j = 0; j < R; ++j) {
    const uint32_t row = w0 + tid + j * kCombineThreads;
    b[j] = e[j] = 0;
    if (row < w1) {
      b[j] = lp[row];
      e[j] = lp[row + 1];
    }
  }
  // two halves of U / 2 loads per thread: the second half's index loads are
  // in flight while the first half's gathers are
  static_assert(U % 2 == 0, "combine_window: two halves");
  {
    constexpr uint32_t H = U / 2;
    uint32_t sa[H], sb[H];
    float va[H], vb[H];
#pragma unroll
    for (uint32_t j = 0; j < H; ++j) {
      const uint32_t k = b0 + tid + j * kCombineThreads;
      sa[j] = k < e0 ? a.lslot[k] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < H; ++j) {
      const uint32_t k = b0 + tid + (H + j) * kCombineThreads;
      sb[j] = k < e0 ? a.lslot[k] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < H; ++j) va[j] = b0 + tid + j * kCombineThreads < e0 ? a.partial[sa[j]] : 0.f;
#pragma unroll
    for (uint32_t j = 0; j < H; ++j)
      vb[j] = b0 + tid + (H + j) * kCombineThreads < e0 ? a.partial[sb[j]] : 0.f;
#pragma unroll
    for (uint32_t j = 0; j < H; ++j)
      if (b0 + tid + j * kCombineThreads < e0) L.buf[tid + j * kCombineThreads] = va[j];
#pragma unroll
    for (uint32_t j = 0; j < H; ++j)
      if (b0 + tid + (H + j) * kCombineThreads < e0) L.buf[tid + (H + j) * kCombineThreads] = vb[j];
  }
  // the rows' epilogue operands, in flight while the rows are added up
  RowIn in[R];
#pragma unroll
  for (uint32_t j = 0; j < R; ++j) {
    const uint32_t row = w0 + tid + j * kCombineThreads;
    in[j] = row < w1 ? load_row<KIND>(a.s, row) : RowIn{};
  }
#pragma unroll
  for (uint32_t j = 0; j < R; ++j) {
    const uint32_t row = w0 + tid + j * kCombineThreads;
    if (row >= w1 || (a.heavy && ((a.heavy[row >> 5] >> (row & 31)) & 1u))) continue;  // big rows: items of their own
    if (e[j] - b[j] > kWaveRowMin) {
      wave_m |= 1u << j;
      hj[j] = atomicAdd(&L.nh, 1u);  // list order is free: each row is summed by one wavefront
      L.hl[hj[j]] = make_uint2(b[j] - b0, e[j] - b[j]);
    } else {
      light_m |= 1u << j;
    }
  }
  lds_barrier();
#pragma unroll
  for (uint32_t j = 0; j < R; ++j) {
    if (!((light_m >> j) & 1u)) continue;
    float z = 0.f;
#pragma unroll 8
    for (uint32_t k = b[j]; k < e[j]; ++k) z += L.buf[k - b0];
    finish_row<KIND>(a.s, w0 + tid + j * kCombineThreads, z, Df, in[j], err, dang);
  }
  const uint32_t nh = L.nh;
  for (uint32_t h = wv; h < nh; h += kCombineWaves) {
    const uint2 x = L.hl[h];
    float z = 0.f;
    for (uint32_t k = wl; k < x.y; k += kWave) z += L.buf[x.x + k];
    z = wave_total(z);
    if (wl == 0) L.z[h] = z;
  }
  lds_barrier();
#pragma unroll
  for (uint32_t j = 0; j < R; ++j)
    if ((wave_m >> j) & 1u) finish_row<KIND>(a.s, w0 + tid + j * kCombineThreads, L.z[hj[j]], Df, in[j], err, dang);
}

// Quiet-row block q (plain kind, rows [nrows, n) without entries): the row
// epilogue with z = 0 -- the same operations, so the value a stored row
// without entries gets -- stored as contributions only; block 0 adds their
// residual and dangling mass as count x value and records the common rank.
// With the dangling mass leaked the common rank is (1 - d)/n in every sweep
// (z = 0, no dangling term), so once both contribution buffers hold it
// (sweeps 1 and 2) the stores are skipped.
template <uint32_t KIND>
__device__ __forceinline__ void combine_quiet(const CombineArgs& a, uint32_t q, float Df, double& err,
                                              double& dang) {
  if constexpr (KIND == kPlain) {
    const SweepArgs& s = a.s;
    const float u = s.dangling == kDanglingUniform ? s.un : 0.f;
    const float sq = fmaf(Df, u, 0.f);
    const float rn = fmaf(s.d_hi, sq, fmaf(s.d_lo, sq, s.base_u));
    const uint32_t lo = s.nrows + q * kQuietRows,
                   hi = s.dangling != kDanglingLeak || s.st->ctr[s.p] <= 2 ? max(lo, min(s.q_end, lo + kQuietRows))
                                                                          : lo;
    constexpr uint32_t U = kQuietRows / kCombineThreads;  // one round of loads
    for (uint32_t v0 = lo + threadIdx.x; v0 < hi; v0 += U * kCombineThreads) {
      float dv[U];
#pragma unroll
      for (uint32_t j = 0; j < U; ++j) dv[j] = v0 + j * kCombineThreads < hi ? s.div[v0 + j * kCombineThreads] : 1.f;
#pragma unroll
      for (uint32_t j = 0; j < U; ++j)
        if (v0 + j * kCombineThreads < hi) s.c_out[v0 + j * kCombineThreads] = rn / dv[j];
    }
    if (q == 0 && threadIdx.x == 0) {
      const float ro = s.st->rq[s.p ^ 1];
      s.st->rq[s.p] = rn;
      err += static_cast<double>(s.n - s.nrows) * fabs(static_cast<double>(rn) - static_cast<double>(ro));
      dang += static_cast<double>(s.nq_dangling) * static_cast<double>(rn);
    }
  }
}

template <uint32_t KIND>
__global__ __launch_bounds__(kCombineThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_combine(CombineArgs a) {
  __shared__ CombineLds L;
  const uint32_t tid = threadIdx.x, wl = tid & (kWave - 1), wv = tid / kWave;
  if (tid == 0) {
    L.flag = KIND != kCoef ? a.s.st->done : 0u;
    L.nh = 0;
  }
  __syncthreads();
  if (L.flag) {
    if constexpr (KIND != kCoef) finish_output<kCombineThreads>(a);
    return;
  }
  const float Df = KIND != kCoef ? static_cast<float>(a.s.st->dsum) : 0.f;
  double err = 0, dang = 0;
  const uint32_t i0 = a.ptr[blockIdx.x], i1 = a.ptr[blockIdx.x + 1];
  // each entry's item number and descriptor are read one entry ahead
  uint32_t item = i0 < i1 ? a.items[i0] : 0u;
  GatherDesc d = i0 < i1 ? a.desc[i0] : GatherDesc{};
  for (uint32_t i = i0; i < i1; ++i) {
    const uint32_t nj = min(i + 1, i1 - 1);
    const uint32_t nitem = a.items[nj];
    const GatherDesc nd = a.desc[nj];
    if (item < a.nbig) {
      combine_big<KIND>(a, d, Df, L, err, dang);
    } else if (item - a.nbig < a.nwin) {
      combine_window<KIND>(a, d, Df, L, err, dang);
    } else if (item - a.nbig - a.nwin < a.nqb) {
      combine_quiet<KIND>(a, d.x, Df, err, dang);
    }
    item = nitem;
    d = nd;
    lds_barrier();  // every thread is done with the item's LDS
    if (tid == 0) L.nh = 0;
    lds_barrier();
  }
  if constexpr (KIND != kCoef) {
    err = wave_sum(err);
    dang = wave_sum(dang);
    if (wl == 0) {
      L.red[0][wv] = err;
      L.red[1][wv] = dang;
    }
    lds_barrier();
    if (tid == 0) {
      double te = 0, td = 0;
#pragma unroll
      for (uint32_t k = 0; k < kCombineWaves; ++k) {
        te += L.red[0][k];
        td += L.red[1][k];
      }
      a.s.part_out[2 * blockIdx.x] = te;
      a.s.part_out[2 * blockIdx.x + 1] = td;
    }
  }
}

// Accumulate form of the second pass: one workgroup per block of at most
// kAccRows consecutive rows at a time.  Inside every source tile a block's
// fragments are consecutive ordinals, so the workgroup streams its partials
// tile by tile (a wavefront per tile, four loads in flight per lane) into a
// double accumulator per row in LDS, then applies the row epilogue to the
// block's rows.  No scattered reads: the partials and their 16-bit block-local
// rows are read once, in order.
struct AccArgs {
  const uint32_t* __restrict__ blo;   // [nab + 1] first row of each block
  const uint32_t* __restrict__ seg;   // [(nab + 1) * ntiles] first ordinal of block b in tile t
  const uint16_t* __restrict__ frow;  // [nsub] row of each fragment in its block
  uint32_t nab, ntiles;
  uint32_t cap;                       // rows per block at most (accumulators before the spare slots)
};

// Adds one block's partials into the LDS accumulators.  Wavefront wv takes
// tiles wv, wv + kNW, ...: in groups of 64, lane l loads the segment bounds of
// the group's l-th tile, and the wavefront walks them P tiles at a time
// (bounds broadcast with readlane), U loads per lane from each segment, all
// issued before the first add.  Lanes past a segment's end read entry 0 and
// add it to their own spare slot after the block's rows (`spare` + lane), so
// every load and add is unconditional and no two lanes share a spare slot.
template <int P, int U>
__device__ __forceinline__ void acc_tiles(const float* __restrict__ partial, const uint16_t* __restrict__ frow,
                                          double* s_acc, const uint32_t* __restrict__ s0p,
                                          const uint32_t* __restrict__ s1p, uint32_t ntiles, uint32_t wv, uint32_t wl,
                                          uint32_t spare) {
  constexpr uint32_t kNW = kSweepThreads / kWave;
  for (uint32_t g0 = 0; g0 + wv < ntiles; g0 += kNW * kWave) {
    const uint32_t tl = g0 + wv + kNW * wl;
    const uint32_t lb0 = tl < ntiles ? s0p[tl] : 0u, lb1 = tl < ntiles ? s1p[tl] : 0u;
    const uint32_t ng = min(static_cast<uint32_t>(kWave), (ntiles - g0 - wv + kNW - 1) / kNW);
    for (uint32_t i = 0; i < ng; i += P) {
      uint32_t b0[P], ln[P];
      uint32_t len = 0;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const uint32_t idx = min(i + p, static_cast<uint32_t>(kWave - 1));
        b0[p] = __builtin_amdgcn_readlane(lb0, idx);
        const uint32_t e = __builtin_amdgcn_readlane(lb1, idx);
        ln[p] = i + p < ng ? e - b0[p] : 0u;
        len = max(len, ln[p]);
      }
      for (uint32_t o = wl; o < len; o += U * kWave) {
        float v[P * U];
        uint32_t r[P * U];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
          for (int j = 0; j < U; ++j) {
            const uint32_t oj = o + j * kWave;
            const uint32_t k = oj < ln[p] ? b0[p] + oj : 0u;
            v[p * U + j] = partial[k];
            r[p * U + j] = oj < ln[p] ? static_cast<uint32_t>(frow[k]) : spare + wl;
          }
#pragma unroll
        for (int x = 0; x < P * U; ++x) atomicAdd(&s_acc[r[x]], static_cast<double>(v[x]));
      }
    }
  }
}

template <uint32_t KIND>
__global__ __launch_bounds__(kSweepThreads) void k_accum(CombineArgs a, AccArgs q) {
  constexpr uint32_t kNW = kSweepThreads / kWave;
  extern __shared__ double s_acc[];  // q.cap accumulators + kWave spare slots (dynamic)
  __shared__ double s_red[2][kNW];
  __shared__ uint32_t s_flag;
  const uint32_t tid = threadIdx.x;
  const uint32_t wl = tid & (kWave - 1);
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid / kWave);
  float Df = 0.f;
  if constexpr (KIND != kCoef) {
    if (tid == 0) s_flag = a.s.st->done;
    __syncthreads();
    if (s_flag) {
      finish_output<kSweepThreads>(a);
      return;
    }
    Df = static_cast<float>(a.s.st->dsum);
  }
  double err = 0, dang = 0;
  const uint32_t nrows = a.s.nrows;  // quiet rows [nrows, n) are not stored
  for (uint32_t b = blockIdx.x; b < q.nab; b += gridDim.x) {
    const uint32_t lo = q.blo[b];
    if (lo >= nrows) break;  // blocks ascend
    const uint32_t cnt = min(q.blo[b + 1], nrows) - lo;
    for (uint32_t i = tid; i < cnt; i += kSweepThreads) s_acc[i] = 0.0;
    __syncthreads();
    const uint32_t* __restrict__ s0p = q.seg + static_cast<size_t>(b) * q.ntiles;
    const uint32_t* __restrict__ s1p = s0p + q.ntiles;
    if (q.cap > kAccRows / 2)
      acc_tiles<2, 4>(a.partial, q.frow, s_acc, s0p, s1p, q.ntiles, wv, wl, q.cap);
    else
      acc_tiles<4, 2>(a.partial, q.frow, s_acc, s0p, s1p, q.ntiles, wv, wl, q.cap);
    __syncthreads();
    for (uint32_t i = tid; i < cnt; i += kSweepThreads) {
      const uint32_t row = lo + i;
      const RowIn in = load_row<KIND>(a.s, row);
      finish_row<KIND>(a.s, row, static_cast<float>(s_acc[i]), Df, in, err, dang);
    }
    __syncthreads();  // the next block clears the accumulators
  }
  quiet_rows<KIND, kSweepThreads>(a.s, Df, err, dang);
  if constexpr (KIND != kCoef) {
    err = wave_sum(err);
    dang = wave_sum(dang);
    if (wl == 0) {
      s_red[0][wv] = err;
      s_red[1][wv] = dang;
    }
    __syncthreads();
    if (tid == 0) {
      double te = 0, td = 0;
#pragma unroll
      for (uint32_t i = 0; i < kNW; ++i) {
        te += s_red[0][i];
        td += s_red[1][i];
      }
      a.s.part_out[2 * blockIdx.x] = te;
      a.s.part_out[2 * blockIdx.x + 1] = td;
    }
  }
}

__global__ void k_reset_queue(State* st) { st->work[0] = 0; }

// out[v] = r[iperm[v]] * scale.  With `st`, r is the final iterate
// r_pair[st->iters & 1], and rows from `quiet_from` on hold its quiet rank.
// With `hf`: only once the solve has stopped (else nothing), and reports that
// it ran in the host flag.
__global__ __launch_bounds__(kBlock) void k_unpermute(const uint32_t* __restrict__ iperm,
                                                      const float* __restrict__ r0, const float* __restrict__ r1,
                                                      const State* st, uint32_t n, uint32_t quiet_from, float scale,
                                                      float* __restrict__ out, HostFlag* hf) {
  if (hf && !st->done) return;
  if (hf && blockIdx.x == 0 && threadIdx.x == 0) hf->finished = 1;
  const float* r = (st && (st->iters & 1)) ? r1 : r0;
  const float rq = st ? st->rq[st->iters & 1] : 0.f;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const uint32_t i = iperm[v];
    out[v] = (i < quiet_from ? r[i] : rq) * scale;
  }
}

// Flags `bit` in *bad when an entry of v is NaN, infinite or negative.
__global__ __launch_bounds__(kBlock) void k_check_vec(const float* __restrict__ v, uint32_t n, uint32_t bit,
                                                      uint32_t* bad) {
  bool ok = true;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) ok &= (v[i] >= 0.f) & (v[i] <= FLT_MAX);
  if (!ok) atomicOr(bad, bit);
}

// Quiet rows: the count of zero divisors among them.
__global__ __launch_bounds__(kBlock) void k_quiet_setup(const float* __restrict__ degf, uint32_t nz, uint32_t n,
                                                        uint32_t* zeros) {
  uint32_t z = 0;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t v = nz + blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) z += degf[v] == 0.f ? 1u : 0u;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o);
  if ((threadIdx.x & (kWave - 1)) == 0 && z) atomicAdd(zeros, z);
}

uint64_t fnv_mix(uint64_t h, uint64_t v) {
  h ^= v;
  return h * 0x100000001b3ULL;
}

SweepArgs sweep_args(Graph& g, const SolveSpec& sp, const nwgr_pr_params& p, uint32_t parity) {
  SweepArgs a{};
  a.offs = sp.offs;
  a.idx = sp.idx;
  a.w = sp.w;
  a.rows = sp.plan->rows.p ? sp.plan->rows.as<uint32_t>() : nullptr;
  a.div = sp.div;
  a.tel = sp.tel_src ? g.tel.as<float>() : nullptr;
  a.gsrc = sp.gsrc;
  a.r_in = g.r[parity ^ 1].as<float>();
  a.r_out = g.r[parity].as<float>();
  if (sp.kind == kAdjoint) {
    a.c_in = a.r_in;  // the adjoint gathers lambda itself
    a.c_out = nullptr;
  } else {
    a.c_in = g.c[parity ^ 1].as<float>();
    a.c_out = g.c[parity].as<float>();
  }
  if (sp.phases) {
    // block Gauss-Seidel: contributions updated in place, phase by phase;
    // pairs in blocks of grid_c per phase
    a.c_in = a.c_out = g.c[0].as<float>();
    const size_t stride = 2 * static_cast<size_t>(g.gs_stride);
    a.part_in = g.gs_part.as<double>() + stride * (parity ^ 1);
    a.part_out = g.gs_part.as<double>() + stride * parity;
  } else {
    double* part = g.partials.as<double>();
    a.part_in = part + 2 * g.part_stride * (parity ^ 1);
    a.part_out = part + 2 * g.part_stride * parity;
  }
  a.st = g.state.as<State>();
  a.hflag = g.hflag_dev;
  a.hist = g.hist_dev;
  a.hist_cap = g.hist_cap;
  a.hub_units = sp.plan->hub_units.as<HubUnit>();
  a.hub_part = sp.plan->hub_part.as<uint32_t>();
  a.hub_ticket = sp.plan->hub_ticket.as<uint32_t>();
  a.segs = sp.plan->segs;
  a.p = parity;
  a.n = g.n;
  a.lds_k = std::min<uint32_t>(g.tune.lds_cache, g.n + 1);
  a.max_iters = p.max_iters;
  a.dangling = p.dangling;
  const double d = p.damping;
  a.d_hi = static_cast<float>(d);
  a.d_lo = static_cast<float>(d - static_cast<double>(a.d_hi));
  a.base_u = static_cast<float>((1.0 - d) / g.n);
  a.one_minus_d = static_cast<float>(1.0 - d);
  a.un = static_cast<float>(1.0 / g.n);
  a.quiet = sp.quiet ? 1u : 0u;
  a.nrows = sp.quiet ? g.nz : g.n;
  a.q_end = sp.quiet ? g.n_act : g.n;
  a.nq_dangling = g.nq_dangling;
  a.thr = sp.thr;
  a.tol = p.tol;
  a.rsave = g.rsave.as<float>();
  a.coef_a = g.coef_a.as<float>();
  a.coef_b = g.coef_b.as<float>();
  return a;
}

void launch_sweep(uint32_t kind, uint32_t grid, const SweepArgs& a, hipStream_t s) {
  // the LDS cache is dynamic shared memory; allow up to the full 144 KiB once
  static const bool attrs = [] {
    const int bytes = static_cast<int>(kLdsCache * sizeof(float));
    for (const void* f : {reinterpret_cast<const void*>(&k_sweep<kPlain>),
                          reinterpret_cast<const void*>(&k_sweep<kWeighted>),
                          reinterpret_cast<const void*>(&k_sweep<kAdjoint>),
                          reinterpret_cast<const void*>(&k_sweep<kDelta>),
                          reinterpret_cast<const void*>(&k_sweep<kCoef>)})
      NWGR_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    return true;
  }();
  (void)attrs;
  const size_t shm = sizeof(float) * a.lds_k;
  switch (kind) {
    case kPlain: k_sweep<kPlain><<<grid, kSweepThreads, shm, s>>>(a); break;
    case kWeighted: k_sweep<kWeighted><<<grid, kSweepThreads, shm, s>>>(a); break;
    case kAdjoint: k_sweep<kAdjoint><<<grid, kSweepThreads, shm, s>>>(a); break;
    case kDelta: k_sweep<kDelta><<<grid, kSweepThreads, shm, s>>>(a); break;
    default: k_sweep<kCoef><<<grid, kSweepThreads, shm, s>>>(a); break;
  }
}

}  // namespace

uint32_t solve_pairs(const Graph& g, const SolveSpec& sp) {
  if (sp.phases) return sp.only >= 0 ? g.grid_c : g.grid_c * static_cast<uint32_t>(sp.phases->size());
  return sp.tiled ? g.grid_c : g.grid;
}

uint32_t combine_blocks_per_cu() {
  static const uint32_t v = [] {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_combine<kPlain>),
                                                     kCombineThreads, 0) != hipSuccess ||
        nb < 1)
      nb = 2;
    return static_cast<uint32_t>(nb);
  }();
  return v;
}

namespace {

// One tiled pass pair over layout `t`: k_tile over the workgroup task ranges,
// then the second pass (k_combine, or k_accum in the accumulate form).
// `check`: the tile pass takes the stop decision from the `nparts` pairs
// the previous sweep wrote.
void launch_layout(Graph& g, uint32_t kind, const Tiled& t, const float* tw, const SweepArgs& sa, uint32_t check,
                   uint32_t nparts, hipStream_t s) {
  static const bool attrs = [] {
    const int bytes = static_cast<int>(tile_lds_bytes(kMaxTileBits));
    for (const void* f : {reinterpret_cast<const void*>(&k_tile<kPlain>),
                          reinterpret_cast<const void*>(&k_tile<kWeighted>),
                          reinterpret_cast<const void*>(&k_tile<kAdjoint>),
                          reinterpret_cast<const void*>(&k_tile<kDelta>),
                          reinterpret_cast<const void*>(&k_tile<kCoef>)})
      NWGR_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    return true;
  }();
  (void)attrs;
  TileArgs ta{};
  ta.eloc = t.eloc.as<uint16_t>();
  ta.w = tw;
  ta.chunk_sp0 = t.chunk_sp0.as<uint32_t>();
  ta.tasks = t.tasks.as<TiledTask>();
  ta.wg_task = t.wg_task.as<uint32_t>();
  ta.wg_tile0 = t.wg_tile0.as<uint32_t>();
  ta.c_in = sa.c_in;
  ta.partial = t.partial.as<float>();
  ta.part_in = sa.part_in;
  ta.st = sa.st;
  ta.hflag = sa.hflag;
  ta.hist = sa.hist;
  ta.hist_cap = sa.hist_cap;
  ta.nparts = nparts;
  ta.bits = t.bits;
  ta.p = sa.p;
  ta.max_iters = sa.max_iters;
  ta.check = check;
  ta.tol = sa.tol;
  ta.wg_time = g.calib_on && t.wg_time.bytes >= sizeof(uint32_t) * kCalibSweeps * t.nblocks
                   ? t.wg_time.as<uint32_t>() : nullptr;
  CombineArgs ca{};
  ca.s = sa;
  ca.pair_ptr = t.pair_ptr.as<uint32_t>();
  ca.rslot = t.rslot.as<uint32_t>();
  ca.partial = t.partial.as<float>();
  ca.lpp = t.light_split ? t.lpp.as<uint32_t>() : ca.pair_ptr;
  ca.lslot = t.light_split ? t.lslot.as<uint32_t>() : ca.rslot;
  ca.heavy = t.light_split ? t.heavy.as<uint32_t>() : nullptr;
  NWGR_REQUIRE(t.gplan_blocks == g.grid_c, "gather plan made for another grid");
  const Tiled::GatherPlan& gp = t.gplan[gather_plan_index(t, kind, sa.quiet != 0)];
  ca.ptr = gp.ptr.as<uint32_t>();
  ca.items = gp.items.as<uint32_t>();
  ca.desc = gp.desc.as<GatherDesc>();
  ca.nbig = t.nbig;
  ca.nwin = gp.nwin;
  ca.nqb = kind == kPlain ? gp.nqb : 0u;
  ca.iperm = g.iperm.as<uint32_t>();
  ca.rp0 = g.r[0].as<float>();
  ca.rp1 = g.r[1].as<float>();
  const size_t shm = tile_lds_bytes(t.bits);
  const uint32_t nb = t.nblocks;
  if (t.acc_on) {
    static const bool acc_attrs = [] {
      const int bytes = static_cast<int>((kAccRows + kWave) * sizeof(double));
      for (const void* f : {reinterpret_cast<const void*>(&k_accum<kPlain>),
                            reinterpret_cast<const void*>(&k_accum<kWeighted>),
                            reinterpret_cast<const void*>(&k_accum<kAdjoint>),
                            reinterpret_cast<const void*>(&k_accum<kDelta>),
                            reinterpret_cast<const void*>(&k_accum<kCoef>)})
        NWGR_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
      return true;
    }();
    (void)acc_attrs;
    AccArgs q{t.ablo.as<uint32_t>(), t.aseg.as<uint32_t>(), t.afrow.as<uint16_t>(), t.nab, t.ntiles, t.acc_rows};
    const size_t ashm = sizeof(double) * (t.acc_rows + kWave);
    switch (kind) {
      case kPlain:
        k_tile<kPlain><<<nb, kSweepThreads, shm, s>>>(ta);
        k_accum<kPlain><<<g.grid_c, kSweepThreads, ashm, s>>>(ca, q);
        break;
      case kWeighted:
        k_tile<kWeighted><<<nb, kSweepThreads, shm, s>>>(ta);
        k_accum<kWeighted><<<g.grid_c, kSweepThreads, ashm, s>>>(ca, q);
        break;
      case kAdjoint:
        k_tile<kAdjoint><<<nb, kSweepThreads, shm, s>>>(ta);
        k_accum<kAdjoint><<<g.grid_c, kSweepThreads, ashm, s>>>(ca, q);
        break;
      case kDelta:
        k_tile<kDelta><<<nb, kSweepThreads, shm, s>>>(ta);
        k_accum<kDelta><<<g.grid_c, kSweepThreads, ashm, s>>>(ca, q);
        break;
      default:
        k_tile<kCoef><<<nb, kSweepThreads, shm, s>>>(ta);
        k_accum<kCoef><<<g.grid_c, kSweepThreads, ashm, s>>>(ca, q);
        break;
    }
    NWGR_HIP(hipGetLastError());
    return;
  }
  switch (kind) {
    case kPlain:
      k_tile<kPlain><<<nb, kSweepThreads, shm, s>>>(ta);
      k_combine<kPlain><<<g.grid_c, kCombineThreads, 0, s>>>(ca);
      break;
    case kWeighted:
      k_tile<kWeighted><<<nb, kSweepThreads, shm, s>>>(ta);
      k_combine<kWeighted><<<g.grid_c, kCombineThreads, 0, s>>>(ca);
      break;
    case kAdjoint:
      k_tile<kAdjoint><<<nb, kSweepThreads, shm, s>>>(ta);
      k_combine<kAdjoint><<<g.grid_c, kCombineThreads, 0, s>>>(ca);
      break;
    case kDelta:
      k_tile<kDelta><<<nb, kSweepThreads, shm, s>>>(ta);
      k_combine<kDelta><<<g.grid_c, kCombineThreads, 0, s>>>(ca);
      break;
    default:
      k_tile<kCoef><<<nb, kSweepThreads, shm, s>>>(ta);
      k_combine<kCoef><<<g.grid_c, kCombineThreads, 0, s>>>(ca);
      break;
  }
  NWGR_HIP(hipGetLastError());
}

// One tiled sweep of the whole-graph layout.
void launch_tiled(Graph& g, const SolveSpec& sp, const SweepArgs& sa, hipStream_t s) {
  launch_layout(g, sp.kind, *sp.tiled, sp.tw, sa, 1, g.grid_c, s);
}

// One block Gauss-Seidel sweep: the phases' pass pairs in order over the one
// contribution vector; phase q writes its (error, dangling) pairs to its own
// block, and the next sweep's first tile pass adds all of them.
void launch_gs(Graph& g, const SolveSpec& sp, const SweepArgs& sa, hipStream_t s) {
  const uint32_t P = static_cast<uint32_t>(sp.phases->size());
  const uint32_t total = solve_pairs(g, sp);
  if (sp.only >= 0) {  // one row share alone: its pairs are the sweep's
    launch_layout(g, sp.kind, (*sp.phases)[sp.only], nullptr, sa, 1u, total, s);
    return;
  }
  for (uint32_t q = 0; q < P; ++q) {
    SweepArgs aq = sa;
    aq.part_out = sa.part_out + 2 * static_cast<size_t>(g.grid_c) * q;
    launch_layout(g, sp.kind, (*sp.phases)[q], nullptr, aq, q == 0 ? 1u : 0u, total, s);
  }
}

// Captured chunk of `len` consecutive sweep launches whose first launch has
// parity `parity` (sweep j has parity j & 1).
hipGraphExec_t chunk_graph(Graph& g, const SolveSpec& sp, const nwgr_pr_params& p, uint32_t parity, uint32_t len,
                           hipStream_t s) {
  hipGraphExec_t& e = g.exec[sp.phases ? kExecGs : sp.kind][parity][len];
  if (e) return e;
  NWGR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (uint32_t i = 0; i < len; ++i) {
    const SweepArgs a = sweep_args(g, sp, p, parity ^ (i & 1));
    if (sp.phases)
      launch_gs(g, sp, a, s);
    else if (sp.tiled)
      launch_tiled(g, sp, a, s);
    else
      launch_sweep(sp.kind, g.grid, a, s);
  }
  hipGraph_t graph;
  NWGR_HIP(hipStreamEndCapture(s, &graph));
  hipError_t err = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  NWGR_HIP(err);
  return e;
}

// The handle's events, taken in turn by one solve (created once and kept by
// the handle, up to kKeptEvents: creating and destroying them per solve cost
// host time between solves); a traced solve's further ones are its own.
struct Events {
  static constexpr size_t kKeptEvents = 64;
  Graph& g;
  size_t next = 0;
  std::vector<hipEvent_t> own;
  ~Events() {
    for (hipEvent_t e : own) (void)hipEventDestroy(e);
  }
  hipEvent_t make() {
    if (next < g.events.size()) return g.events[next++];
    hipEvent_t e;
    NWGR_HIP(hipEventCreate(&e));
    if (g.events.size() < kKeptEvents) {
      g.events.push_back(e);
      ++next;
    } else {
      own.push_back(e);
    }
    return e;
  }
};

}  // namespace

void validate_params(const nwgr_pr_params& p) {
  NWGR_REQUIRE(std::isfinite(p.damping) && p.damping >= 0.0 && p.damping <= 1.0, "damping must be in [0, 1]");
  NWGR_REQUIRE(!std::isnan(p.tol), "tol is NaN");
  NWGR_REQUIRE(p.dangling <= kDanglingTeleport, "unknown dangling mode");
  NWGR_REQUIRE(p.dangling != kDanglingTeleport || p.teleport != nullptr,
               "dangling = TELEPORT needs a teleport vector");
}

void check_params(const nwgr_pr_params& p, uint32_t n, bool host_vectors) {
  validate_params(p);
  if (!host_vectors) return;
  for (const float* v : {p.teleport, p.init}) {
    if (!v) continue;
    // NaN fails both comparisons, +inf the second
    bool ok = true;
    for (uint32_t i = 0; i < n; ++i) ok &= (v[i] >= 0.f) & (v[i] <= FLT_MAX);
    NWGR_REQUIRE(ok, v == p.teleport ? "teleport has a NaN, infinite or negative entry"
                                     : "init has a NaN, infinite or negative entry");
  }
}

void check_vectors(const nwgr_pr_params& p, uint32_t n, uint32_t* d_flag, hipStream_t s) {
  if (n == 0 || (!p.teleport && !p.init)) return;
  NWGR_HIP(hipMemsetAsync(d_flag, 0, sizeof(uint32_t), s));
  const uint32_t blocks = std::min<uint32_t>((n + kBlock - 1) / kBlock, 4096);
  if (p.teleport) k_check_vec<<<blocks, kBlock, 0, s>>>(p.teleport, n, 1u, d_flag);
  if (p.init) k_check_vec<<<blocks, kBlock, 0, s>>>(p.init, n, 2u, d_flag);
  NWGR_HIP(hipGetLastError());
  uint32_t bad = 0;
  NWGR_HIP(hipMemcpyAsync(&bad, d_flag, sizeof(bad), hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  NWGR_REQUIRE(!(bad & 1u), "teleport has a NaN, infinite or negative entry");
  NWGR_REQUIRE(!(bad & 2u), "init has a NaN, infinite or negative entry");
}

void empty_delta_result(const nwgr_pr_params& p, nwgr_pr_result& res) {
  // an empty graph: iteration 1 pushes from no vertex, so the residual solve
  // stops there (experimental/algorithms/page_rank.hpp:693-696)
  validate_params(p);
  NWGR_REQUIRE(!p.teleport && !p.init && p.dangling == kDanglingLeak,
               "the residual solve takes no teleport, initial vector or dangling mode");
  res = nwgr_pr_result{};
  res.iters = std::min<uint32_t>(p.max_iters, 1);
  res.converged = res.iters > 0 ? 1u : 0u;
}

void empty_result(const nwgr_pr_params& p, nwgr_pr_result& res) {
  // an empty graph: every sweep changes nothing (error 0), so the first one
  // converges when tol > 0 (page_rank.hpp:121-147 over an empty range)
  res = nwgr_pr_result{};
  res.iters = p.tol > 0 ? std::min<uint32_t>(p.max_iters, 1) : p.max_iters;
  res.error = 0.0;
  res.converged = (res.iters > 0 && p.tol > 0) ? 1u : 0u;
}

void finish_launch(Graph& g, const SolveFinish& f, bool cond, hipStream_t s) {
  k_unpermute<<<std::min<uint32_t>((g.n + kBlock - 1) / kBlock, 8192), kBlock, 0, s>>>(
      g.iperm.as<uint32_t>(), g.r[0].as<float>(), g.r[1].as<float>(), g.state.as<State>(), g.n, f.quiet_from, 1.0f,
      f.out, cond ? g.hflag_dev : nullptr);
  NWGR_HIP(hipGetLastError());
}

void solve(Graph& g, const SolveSpec& sp, const nwgr_pr_params& p, nwgr_pr_result& res, hipStream_t s,
           const SolveFinish* fin) {
  // a traced solve captures one sweep per chunk, so the events around the
  // chunks time every sweep on its own (the reference's per-iteration time,
  // page_rank.hpp:123,142)
  const bool per_sweep = p.trace != 0;
  const uint32_t K = per_sweep ? 1u : std::clamp<uint32_t>(p.chunk ? p.chunk : g.tune.chunk, 1, Graph::kMaxChunk);
  const bool has_tel = sp.tel_src != nullptr;
  if (has_tel && !g.tel.p) {
    g.tel.alloc(sizeof(float) * g.n);
    g.drop_graphs();
  }
  // (error, dangling) pairs of one sweep; init writes as many
  const uint32_t npairs = solve_pairs(g, sp);
  if (sp.phases) {
    if (g.gs_stride < npairs || !g.gs_part.p) {
      g.gs_stride = npairs;
      g.gs_part.alloc(sizeof(double) * 2 * 2 * static_cast<size_t>(npairs));
      g.drop_graphs();
    }
  } else if (g.part_stride < npairs || !g.partials.p) {
    g.part_stride = std::max(npairs, g.part_stride);
    g.partials.alloc(sizeof(double) * 2 * 2 * static_cast<size_t>(g.part_stride));
    g.drop_graphs();
  }

  // the captured graphs bake in the arguments and the stream
  uint64_t key = 0xcbf29ce484222325ULL, bits;
  double dd = p.damping, tt = p.tol;
  std::memcpy(&bits, &dd, 8);
  key = fnv_mix(key, bits);
  std::memcpy(&bits, &tt, 8);
  key = fnv_mix(key, bits);
  key = fnv_mix(key, p.max_iters);
  key = fnv_mix(key, p.dangling);
  key = fnv_mix(key, has_tel);
  key = fnv_mix(key, sp.quiet);
  uint32_t thr_bits;
  std::memcpy(&thr_bits, &sp.thr, 4);
  key = fnv_mix(key, thr_bits);
  key = fnv_mix(key, reinterpret_cast<uintptr_t>(s));
  for (const void* ptr : {static_cast<const void*>(sp.offs), static_cast<const void*>(sp.idx),
                          static_cast<const void*>(sp.w), static_cast<const void*>(sp.div),
                          static_cast<const void*>(sp.gsrc), static_cast<const void*>(sp.plan),
                          static_cast<const void*>(sp.tiled), static_cast<const void*>(sp.tw),
                          sp.phases ? static_cast<const void*>(sp.phases->data()) : nullptr,
                          static_cast<const void*>(g.gs_part.p), static_cast<const void*>(g.partials.p)})
    key = fnv_mix(key, reinterpret_cast<uintptr_t>(ptr));
  key = fnv_mix(key, g.part_stride);
  key = fnv_mix(key, g.gs_stride);
  key = fnv_mix(key, sp.phases ? sp.phases->size() : 0);
  key = fnv_mix(key, static_cast<uint64_t>(sp.only + 1));
  key = fnv_mix(key, g.calib_on);
  const uint32_t slot = sp.phases ? kExecGs : sp.kind;
  if (g.exec_key[slot] != key) {
    g.drop_graphs(static_cast<int>(slot));
    g.exec_key[slot] = key;
  }

  Events evs{g};
  hipEvent_t ev0 = evs.make(), ev1 = evs.make();
  std::memset(const_cast<HostFlag*>(g.hflag), 0, sizeof(HostFlag));
  NWGR_HIP(hipEventRecord(ev0, s));
  InitArgs ia{};
  ia.perm = g.perm.as<uint32_t>();
  ia.div = sp.div;
  ia.init = sp.init;
  ia.tel_src = sp.tel_src;
  ia.gsrc = sp.gsrc;
  ia.tel = has_tel ? g.tel.as<float>() : nullptr;
  ia.r0 = g.r[0].as<float>();
  ia.c0 = g.c[0].as<float>();
  ia.part0 = sp.phases ? g.gs_part.as<double>() : g.partials.as<double>();
  // a row share leaves the other rows as they start, in both iterates
  ia.r1 = sp.only >= 0 ? g.r[1].as<float>() : nullptr;
  ia.st = g.state.as<State>();
  ia.n = g.n;
  ia.adjoint = sp.kind == kAdjoint;
  ia.delta = sp.kind == kDelta;
  ia.dangling = p.dangling;
  ia.un = static_cast<float>(1.0 / g.n);
  ia.thr = sp.thr;
  ia.d = static_cast<float>(p.damping);
  ia.res0 = static_cast<float>(1.0 - static_cast<double>(ia.d));
  ia.npairs = npairs;
  ia.quiet_from = sp.quiet ? g.nz : g.n;
  // the stopping launch's gather pass writes the output (tiled sweeps without
  // phases); otherwise a launch of its own does
  const bool fused_fin = fin && sp.tiled && !sp.phases;
  ia.c_from = sp.quiet ? g.n_act : g.n;
  ia.out = fused_fin ? fin->out : nullptr;
  k_init<<<std::min(sp.tiled ? g.grid_c : g.grid, npairs), kBlock, 0, s>>>(ia);
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
  uint32_t pobs = 0;      // latest observed progress (sweeps completed)
  double eobs = 0, rho = 0;  // its error and the per-sweep decay ratio
  auto launch = [&](uint32_t len) {
    Chunk c{evs.make(), evs.make(), len};
    hipGraphExec_t e = chunk_graph(g, sp, p, static_cast<uint32_t>(next & 1), len, s);
    NWGR_HIP(hipEventRecord(c.a, s));
    NWGR_HIP(hipGraphLaunch(e, s));
    NWGR_HIP(hipEventRecord(c.b, s));
    pending.push_back(c);
    next += len;
    prof.launches += len;
    prof.chunks += 1;
  };
  // Before the decay rate is known: four sweeps, or -- when the handle's
  // previous solve had the same captured arguments and the same uniform start
  // (no teleport or initial vector) -- that solve's sweeps plus the stopping
  // launch.  Launch planning only: every launch past the stop returns at once.
  const bool hinted = !p.init && !p.teleport && !per_sweep && g.hint_key[slot] == key && g.hint_iters[slot] > 0;
  const uint64_t hint_stop = hinted ? static_cast<uint64_t>(g.hint_iters[slot]) + 1 : 0;
  const uint32_t K0 = hinted ? static_cast<uint32_t>(std::min<uint64_t>(K, hint_stop)) : std::min<uint32_t>(K, 4);
  // launch that is expected to detect convergence (0: no estimate yet)
  auto predicted_stop = [&]() -> uint64_t {
    if (pobs == 0 && hint_stop) return hint_stop;
    if (pobs > 0 && eobs > 0 && rho > 0 && rho < 1 && std::isfinite(p.tol) && p.tol > 0) {
      double more = eobs < p.tol ? 0.0 : std::ceil(std::log(p.tol / eobs) / std::log(rho));
      if (!std::isfinite(more) || more > 1e6) more = 1e6;
      return pobs + static_cast<uint64_t>(more) + 1;
    }
    return 0;
  };
  auto plan = [&]() -> uint64_t {
    // launches still useful beyond `next - 1`
    if (next > limit) return 0;
    uint64_t want = K0;
    if (const uint64_t stop_launch = predicted_stop()) {
      want = stop_launch >= next ? stop_launch - next + 1 : 0;
    } else if (pobs > 0) {
      want = K;  // no usable decay estimate (stagnating or tol <= 0): full chunks
    }
    want = std::min<uint64_t>(want, K);
    want = std::min<uint64_t>(want, limit - next + 1);
    return want;
  };
  auto retire = [&](const Chunk& c) {
    NWGR_HIP(hipEventSynchronize(c.b));
    float cms = 0;
    NWGR_HIP(hipEventElapsedTime(&cms, c.a, c.b));
    prof.chunk_ms += cms;
    if (per_sweep) prof.launch_ms.push_back(cms);
  };
  // progress as the device publishes it, sweep by sweep, into host-mapped
  // memory (read while chunks run, so later chunks are sized before the
  // earlier ones end)
  auto observe = [&]() {
    const uint32_t pr = g.hflag->progress;
    if (pr > pobs) {
      pobs = pr;
      eobs = g.hflag->last_err;
      const double pe = g.hflag->prev_err;
      rho = (pr >= 2 && pe > 0 && eobs < pe) ? eobs / pe : 0.0;
    }
  };
  // the output permutation, enqueued once the launches issued are expected
  // to include the one that detects convergence: it does nothing unless the
  // solve has stopped by the time it runs, and reports in the host flag
  bool fin_enqueued = false;
  auto maybe_finish = [&]() {
    if (!fin || fused_fin || fin_enqueued) return;
    const uint64_t stop_launch = predicted_stop();
    if (next > limit || (stop_launch && next > stop_launch)) {
      finish_launch(g, *fin, true, s);
      fin_enqueued = true;
    }
  };
  launch(static_cast<uint32_t>(std::min<uint64_t>(K0, limit)));
  maybe_finish();  // a first chunk sized from the previous solve covers the stop
  for (;;) {
    if (pending.size() - head < 2) {
      // the second chunk waits, while the first runs, for the decay of its
      // sweeps 2 and 3 (or for the first chunk's end); after a hinted first
      // chunk nothing more is planned until it ends
      if (!hinted && prof.chunks == 1 && head < pending.size() && g.hflag->progress < 3 && !g.hflag->done &&
          hipEventQuery(pending[head].b) == hipErrorNotReady) {
        std::this_thread::yield();
        continue;
      }
      observe();
      const uint64_t len = plan();
      if (len > 0) {
        launch(static_cast<uint32_t>(len));
        maybe_finish();
        continue;
      }
    }
    if (head == pending.size()) {
      // nothing in flight and nothing planned: extend by one launch
      if (next > limit) throw Error(NWGR_EDEVICE, "solver passed its launch limit without stopping");
      launch(1);
      continue;
    }
    retire(pending[head++]);
    if (g.hflag->done) break;
    observe();
  }
  // chunks still in flight see `done` and return at once; the per-sweep
  // errors come back with the same synchronisation
  for (; head < pending.size(); ++head) retire(pending[head]);
  const uint32_t iters_done = g.hflag->iters;
  const uint32_t nhist = std::min(iters_done, g.hist_cap);  // written into mapped memory by the sweeps
  NWGR_HIP(hipEventRecord(ev1, s));
  NWGR_HIP(hipEventSynchronize(ev1));
  if (fin && !g.hflag->finished) {
    // the early output permutation ran before the stop (or was not issued)
    finish_launch(g, *fin, false, s);
    NWGR_HIP(hipEventRecord(ev1, s));
    NWGR_HIP(hipEventSynchronize(ev1));
  }
  float ms = 0;
  NWGR_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  res.iters = g.hflag->iters;
  res.error = g.hflag->err;
  res.converged = (res.iters > 0 && res.error < p.tol) ? 1u : 0u;
  res.ms = ms;
  g.last_profile = prof;
  g.last_engine = !sp.tiled ? 0 : (sp.tiled->acc_on ? 2 : 1);
  g.last_errs.assign(g.hist_h, g.hist_h + nhist);
  g.hint_key[slot] = key;
  g.hint_iters[slot] = res.iters;
  if (p.trace) {
    // the reference's trace table (page_rank.hpp:42-52, :109, :142), with
    // every sweep's own device time (seconds)
    std::printf("%-5s%-20s%-20s%-20s\n", "iter", "error", "time", "outgoing");
    for (uint32_t i = 0; i < g.last_errs.size(); ++i) {
      const double t = i < prof.launch_ms.size() ? prof.launch_ms[i] / 1e3 : prof.chunk_ms / 1e3 / prof.launches;
      std::printf("%-5u%-20.10f%-20.10f%-20.10f\n", i, g.last_errs[i], t, 0.0);
    }
    std::fflush(stdout);
  }
}

void run_coef_pass(Graph& g, const SolveSpec& sp, const nwgr_pr_params& p, const float* lam, hipStream_t s) {
  SweepArgs a = sweep_args(g, sp, p, 0);
  a.c_in = lam;
  a.r_in = lam;
  k_reset_queue<<<1, 1, 0, s>>>(g.state.as<State>());
  NWGR_HIP(hipGetLastError());
  if (sp.tiled) {
    SolveSpec sc = sp;
    sc.kind = kCoef;
    launch_tiled(g, sc, a, s);
  } else {
    launch_sweep(kCoef, g.grid, a, s);
  }
  NWGR_HIP(hipGetLastError());
}

void prepare_quiet(Graph& g, hipStream_t s) {
  if (g.quiet_ready || g.nz >= g.n) return;
  if (!g.vflag.p) g.vflag.alloc(sizeof(uint32_t));
  NWGR_HIP(hipMemsetAsync(g.vflag.p, 0, sizeof(uint32_t), s));
  k_quiet_setup<<<std::min<uint32_t>((g.n - g.nz + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(
      g.degf.as<float>(), g.nz, g.n, g.vflag.as<uint32_t>());
  NWGR_HIP(hipGetLastError());
  NWGR_HIP(hipMemcpyAsync(&g.nq_dangling, g.vflag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  g.quiet_ready = true;
}

void unpermute(Graph& g, const float* src, float* out, float scale, hipStream_t s) {
  k_unpermute<<<std::min<uint32_t>((g.n + kBlock - 1) / kBlock, 8192), kBlock, 0, s>>>(
      g.iperm.as<uint32_t>(), src, src, nullptr, g.n, g.n, scale, out, nullptr);
  NWGR_HIP(hipGetLastError());
}

void run_plain(Graph& g, const nwgr_pr_params& p, float* d_out, nwgr_pr_result& res, hipStream_t s) {
  validate_params(p);
  if (!g.vflag.p) g.vflag.alloc(sizeof(uint32_t));
  check_vectors(p, g.n, g.vflag.as<uint32_t>(), s);
  SolveSpec sp;
  sp.kind = kPlain;
  sp.plan = &g.fwd;
  sp.offs = g.offs.as<uint32_t>();
  sp.idx = g.idx.as<uint32_t>();
  sp.div = g.degf.as<float>();
  sp.init = p.init;
  sp.tel_src = p.teleport;
  sp.tiled = select_tiled(g, false, s);
  // quiet rows share one rank when the teleport is uniform, the initial
  // iterate uniform and the dangling mass leaked or spread uniformly
  sp.quiet = sp.tiled && !p.teleport && !p.init && p.dangling != kDanglingTeleport && g.nz < g.n;
  if (sp.quiet) prepare_quiet(g, s);
  const SolveFinish fin{d_out, sp.quiet ? g.nz : g.n};
  solve(g, sp, p, res, s, &fin);
}

// Measured re-cuts of the default layout's tile pass.  The cut model leaves
// workgroups whose times differ by several microseconds, the same workgroups
// sweep after sweep (R-MAT s22: slowest 4.5 us past the mean, sweep-to-sweep
// correlation of the workgroups' times 0.9 to 0.96).  A calibration solve
// (uniform start, UNIFORM dangling, tol 0) records every workgroup's
// tile-pass time over kCalibSweeps sweeps; recut_tasks scales the chunk costs
// by them and cuts again.  kCalibRounds rounds; the cut whose slowest
// workgroup measured shortest is kept.  The cut decides only which workgroup
// adds which chunk: every partial sum goes to the same slot, so results do not
// change.  Layouts with fewer than kCalibMinChunks chunks per workgroup keep
// the model's cut.
void calibrate_tile_pass(Graph& g, hipStream_t s) {
  constexpr uint32_t kCalibRounds = 3;
  constexpr uint32_t kCalibMinChunks = 64;
  Tiled* t = select_tiled(g, false, s);
  if (!t || !t->built || t->nblocks == 0 || t->nchunks < kCalibMinChunks * t->nblocks) return;
  const uint32_t G = t->nblocks;
  t->wg_time.alloc(sizeof(uint32_t) * kCalibSweeps * G);
  nwgr_pr_params p{};
  p.damping = 0.85;
  p.tol = 0.0;
  p.max_iters = kCalibSweeps + 1;  // sweeps 2 .. kCalibSweeps + 1 are recorded
  p.dangling = kDanglingUniform;
  SolveSpec sp;
  sp.kind = kPlain;
  sp.plan = &g.fwd;
  sp.offs = g.offs.as<uint32_t>();
  sp.idx = g.idx.as<uint32_t>();
  sp.div = g.degf.as<float>();
  sp.tiled = t;
  sp.quiet = g.nz < g.n;
  if (sp.quiet) prepare_quiet(g, s);
  std::vector<uint32_t> ticks(static_cast<size_t>(kCalibSweeps) * G);
  std::vector<double> us(G), best_cost;
  double best = std::numeric_limits<double>::infinity();
  g.calib_on = true;
  for (uint32_t round = 0;; ++round) {
    NWGR_HIP(hipMemsetAsync(t->wg_time.p, 0, t->wg_time.bytes, s));
    nwgr_pr_result res{};
    solve(g, sp, p, res, s);
    NWGR_HIP(hipMemcpyAsync(ticks.data(), t->wg_time.p, t->wg_time.bytes, hipMemcpyDeviceToHost, s));
    NWGR_HIP(hipStreamSynchronize(s));
    double worst = 0;
    for (uint32_t b = 0; b < G; ++b) {
      double sum = 0;
      uint32_t cnt = 0;
      for (uint32_t k = 0; k < kCalibSweeps; ++k)
        if (const uint32_t v = ticks[static_cast<size_t>(k) * G + b]) {
          sum += v;
          cnt += 1;
        }
      us[b] = cnt ? sum / cnt / 100.0 : 0.0;  // 100 MHz wall clock
      worst = std::max(worst, us[b]);
    }
    if (worst < best) {
      best = worst;
      best_cost = t->ccost;
    }
    if (round == kCalibRounds || !recut_tasks(g, *t, us, s)) break;
  }
  g.calib_on = false;
  if (!best_cost.empty() && best_cost != t->ccost) {
    t->ccost = std::move(best_cost);
    plan_tasks(g, *t, s);
  }
  g.drop_graphs();
  // the calibration solves are not the caller's: no last solve to report
  g.last_errs.clear();
  g.last_profile = SolveProfile{};
  g.last_engine = -1;
  for (uint32_t k = 0; k < kExecSlots; ++k) g.hint_key[k] = 0;
}

void run_gs(Graph& g, const nwgr_pr_params& p, uint32_t phases, float* d_out, nwgr_pr_result& res, hipStream_t s,
            int only) {
  validate_params(p);
  NWGR_REQUIRE(p.dangling == kDanglingLeak,
               "block Gauss-Seidel takes LEAK dangling (the reference's in-place sweep redistributes none)");
  NWGR_REQUIRE(phases >= 1 && phases <= kMaxGsPhases, "phases must be in [1, 64]");
  if (!g.vflag.p) g.vflag.alloc(sizeof(uint32_t));
  check_vectors(p, g.n, g.vflag.as<uint32_t>(), s);
  ensure_gs(g, phases, s);
  SolveSpec sp;
  sp.kind = kPlain;
  sp.plan = &g.fwd;
  sp.offs = g.offs.as<uint32_t>();
  sp.idx = g.idx.as<uint32_t>();
  sp.div = g.degf.as<float>();
  sp.init = p.init;
  sp.tel_src = p.teleport;
  NWGR_REQUIRE(only < static_cast<int>(phases), "share index out of range");
  sp.phases = &g.gs;
  sp.only = only;
  sp.tiled = &g.gs[only >= 0 ? only : 0];
  // quiet rows are kept as the scalar only where the last phase refreshes them
  sp.quiet = !p.teleport && !p.init && g.nz < g.n && (only < 0 || only + 1 == static_cast<int>(phases));
  if (sp.quiet) prepare_quiet(g, s);
  const SolveFinish fin{d_out, sp.quiet ? g.nz : g.n};
  solve(g, sp, p, res, s, &fin);
  g.last_engine = 3;
}

// Measured balance of the row shares (the phases with phases = shares): each
// share swept alone (LEAK, tol 0, kBalanceSweeps sweeps, the second of two
// solves timed), then the rows re-cut with every row of share k costing its
// share's measured-to-mean sweep time (clamped to [1/2, 2]) times more, the
// layouts rebuilt and measured again; `rounds` re-cuts, the cuts with the
// shortest slowest share kept.  Returns that share's time per sweep (us).
// The cut model leaves shares 2.5 us apart at R-MAT s22 in a way no linear
// fit of their entries, fragments and rows explains.
double balance_shares(Graph& g, uint32_t shares, uint32_t rounds, hipStream_t s) {
  constexpr uint32_t kBalanceSweeps = 10;
  ensure_gs(g, shares, s);
  nwgr_pr_params p{};
  p.damping = 0.85;
  p.tol = 0.0;
  p.max_iters = kBalanceSweeps;
  p.dangling = kDanglingLeak;
  auto measure = [&](std::vector<double>& t) {
    t.assign(shares, 0.0);
    for (uint32_t q = 0; q < shares; ++q) {
      SolveSpec sp;
      sp.kind = kPlain;
      sp.plan = &g.fwd;
      sp.offs = g.offs.as<uint32_t>();
      sp.idx = g.idx.as<uint32_t>();
      sp.div = g.degf.as<float>();
      sp.phases = &g.gs;
      sp.only = static_cast<int>(q);
      sp.tiled = &g.gs[q];
      sp.quiet = g.nz < g.n && q + 1 == shares;
      if (sp.quiet) prepare_quiet(g, s);
      for (int rep = 0; rep < 2; ++rep) {
        nwgr_pr_result res{};
        solve(g, sp, p, res, s);
        t[q] = g.last_profile.chunk_ms * 1e3 / std::max(1u, res.iters);
      }
    }
    return *std::max_element(t.begin(), t.end());
  };
  std::vector<double> t;
  double best = measure(t);
  std::vector<uint32_t> best_cuts = g.gs_cuts;
  for (uint32_t round = 0; round < rounds; ++round) {
    double mean = 0;
    for (double x : t) mean += x / shares;
    std::vector<double> f(shares);
    for (uint32_t q = 0; q < shares; ++q) f[q] = std::clamp(t[q] / mean, 0.5, 2.0);
    const std::vector<uint32_t> old = g.gs_cuts;
    build_gs(g, gs_cut_rows(g, shares, &old, &f), s);
    const double worst = measure(t);
    if (worst < best) {
      best = worst;
      best_cuts = g.gs_cuts;
    }
  }
  // a rebuild drops the sweeps captured for the other cuts
  if (best_cuts != g.gs_cuts) build_gs(g, best_cuts, s);
  // the balancing solves are not the caller's: no last solve to report, no
  // sweep count to start the next share or phase solve from
  g.last_errs.clear();
  g.last_profile = SolveProfile{};
  g.last_engine = -1;
  g.hint_key[kExecGs] = 0;
  return best;
}

void run_delta(Graph& g, const nwgr_pr_params& p, float* d_out, nwgr_pr_result& res, hipStream_t s) {
  validate_params(p);
  NWGR_REQUIRE(!p.teleport && !p.init && p.dangling == kDanglingLeak,
               "the residual solve takes no teleport, initial vector or dangling mode");
  if (p.max_iters == 0) {
    // no iteration: the ranks keep their initial value 0
    NWGR_HIP(hipMemsetAsync(d_out, 0, sizeof(float) * g.n, s));
    NWGR_HIP(hipStreamSynchronize(s));
    res = nwgr_pr_result{};
    g.last_errs.clear();
    return;
  }
  SolveSpec sp;
  sp.kind = kDelta;
  sp.plan = &g.fwd;
  sp.offs = g.offs.as<uint32_t>();
  sp.idx = g.idx.as<uint32_t>();
  sp.div = g.degf.as<float>();
  sp.tiled = select_tiled(g, false, s);
  sp.thr = static_cast<float>(p.tol);
  // iteration 1's push runs in the init kernel, each sweep is phase 2 of one
  // iteration fused with phase 1 of the next; a sweep's "error" is its push
  // count, so the solve stops after the first sweep that pushed nothing
  nwgr_pr_params q = p;
  q.max_iters = p.max_iters - 1;
  q.tol = 0.5;
  const SolveFinish fin{d_out, g.n};
  solve(g, sp, q, res, s, &fin);
  // per-iteration push counts: iteration 1 (init), then one per sweep
  const double first = g.hflag->first_err;
  g.last_errs.insert(g.last_errs.begin(), first);
  if (res.iters == 0) res.error = first;
  res.iters += 1;
  res.converged = res.error < 0.5 ? 1u : 0u;
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

int nwgr_pr_params_check(const nwgr_pr_params* p, uint32_t n) {
  return guarded([&] {
    NWGR_REQUIRE(p != nullptr, "null argument");
    nwgr::check_params(*p, n, true);
  });
}

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
  if (g->cached) (void)nwgr::cache_forget(g);
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(g->device);
  delete static_cast<nwgr::Graph*>(g);
  (void)hipSetDevice(prev);
}

uint32_t nwgr_pr_graph_n(const nwgr_pr_graph* g) { return g ? g->n : 0; }
int nwgr_pr_last_engine(const nwgr_pr_graph* g) { return g ? g->last_engine : -1; }
uint64_t nwgr_pr_graph_m(const nwgr_pr_graph* g) { return g ? g->m : 0; }
uint64_t nwgr_pr_graph_device_bytes(const nwgr_pr_graph* g) { return g ? nwgr::bytes_of(*g) : 0; }
double nwgr_pr_graph_setup_ms(const nwgr_pr_graph* g) { return g ? g->setup_ms : 0.0; }

int nwgr_pr_graph_get_tuning(const nwgr_pr_graph* g, nwgr_pr_tuning* t) {
  return guarded([&] {
    NWGR_REQUIRE(g && t, "null argument");
    const nwgr::Tuning& u = g->tune;
    t->lds_cache = u.lds_cache;
    t->hub_threshold = u.hub_threshold;
    t->hub_chunk = u.hub_chunk;
    t->unit_edges = u.unit_edges;
    t->zero_rows_per_unit = u.zero_rows_per_unit;
    t->blocks_per_cu = u.blocks_per_cu;
    t->chunk = u.chunk;
    t->engine = u.engine;
    t->tile_bits = u.tile_bits;
    t->tile_fill_cost = u.tile_fill_cost;
    t->tile_combine = u.tile_combine;
  });
}

int nwgr_pr_graph_set_tuning(nwgr_pr_graph* g, const nwgr_pr_tuning* t) {
  return guarded([&] {
    NWGR_REQUIRE(g && t, "null argument");
    const nwgr::Tuning old = g->tune;
    g->tune.lds_cache = t->lds_cache;
    g->tune.hub_threshold = t->hub_threshold;
    g->tune.hub_chunk = t->hub_chunk;
    g->tune.unit_edges = t->unit_edges;
    g->tune.zero_rows_per_unit = t->zero_rows_per_unit;
    g->tune.blocks_per_cu = t->blocks_per_cu;
    g->tune.chunk = t->chunk;
    g->tune.engine = t->engine;
    g->tune.tile_bits = t->tile_bits;
    g->tune.tile_fill_cost = t->tile_fill_cost;
    g->tune.tile_combine = t->tile_combine;
    if (g->n == 0) {  // no layout to re-plan
      try {
        nwgr::validate_tuning(g->tune);
      } catch (...) {
        g->tune = old;
        throw;
      }
      return;
    }
    nwgr::DeviceGuard dg(g->device);
    try {
      nwgr::replan(*g, g->own_stream);
    } catch (...) {
      g->tune = old;
      nwgr::replan(*g, g->own_stream);
      throw;
    }
  });
}

int nwgr_pr_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* d_r_out, nwgr_pr_result* res,
                           void* stream) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::check_params(*p, g->n, false);
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(d_r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_plain(*g, *p, d_r_out, *res, s);
  });
}

int nwgr_pr_run_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* r_out, nwgr_pr_result* res) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = g->own_stream;
    const size_t nb = sizeof(float) * g->n;
    // the handle keeps its staging buffers, so repeated calls allocate nothing
    auto staged = [&](nwgr::DevBuf& b) -> float* {
      if (b.bytes < nb) b.alloc(nb);
      return b.as<float>();
    };
    float* out = staged(g->out_buf);
    nwgr_pr_params q = *p;
    if (p->teleport) {
      q.teleport = staged(g->tel_h);
      NWGR_HIP(hipMemcpyAsync(const_cast<float*>(q.teleport), p->teleport, nb, hipMemcpyHostToDevice, s));
    }
    if (p->init) {
      q.init = staged(g->init_h);
      NWGR_HIP(hipMemcpyAsync(const_cast<float*>(q.init), p->init, nb, hipMemcpyHostToDevice, s));
    }
    nwgr::run_plain(*g, q, out, *res, s);
    nwgr::copy_out(*g, r_out, out, nb, s);
  });
}

int nwgr_pr_gs_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t phases, float* d_r_out,
                              nwgr_pr_result* res, void* stream) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    NWGR_REQUIRE(phases >= 1 && phases <= nwgr::kMaxGsPhases, "phases must be in [1, 64]");
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(d_r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_gs(*g, *p, phases, d_r_out, *res, s);
  });
}

int nwgr_pr_gs_run_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t phases, float* r_out,
                       nwgr_pr_result* res) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    NWGR_REQUIRE(phases >= 1 && phases <= nwgr::kMaxGsPhases, "phases must be in [1, 64]");
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = g->own_stream;
    const size_t nb = sizeof(float) * g->n;
    auto staged = [&](nwgr::DevBuf& b) -> float* {
      if (b.bytes < nb) b.alloc(nb);
      return b.as<float>();
    };
    float* out = staged(g->out_buf);
    nwgr_pr_params q = *p;
    if (p->teleport) {
      q.teleport = staged(g->tel_h);
      NWGR_HIP(hipMemcpyAsync(const_cast<float*>(q.teleport), p->teleport, nb, hipMemcpyHostToDevice, s));
    }
    if (p->init) {
      q.init = staged(g->init_h);
      NWGR_HIP(hipMemcpyAsync(const_cast<float*>(q.init), p->init, nb, hipMemcpyHostToDevice, s));
    }
    nwgr::run_gs(*g, q, phases, out, *res, s);
    nwgr::copy_out(*g, r_out, out, nb, s);
  });
}

int nwgr_pr_share_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, uint32_t shares, uint32_t share,
                                 float* d_r_out, nwgr_pr_result* res, void* stream) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    NWGR_REQUIRE(shares >= 1 && shares <= nwgr::kMaxGsPhases, "shares must be in [1, 64]");
    NWGR_REQUIRE(share < shares, "share must be < shares");
    if (g->n == 0) return nwgr::empty_result(*p, *res);
    NWGR_REQUIRE(d_r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_gs(*g, *p, shares, d_r_out, *res, s, static_cast<int>(share));
  });
}

int nwgr_pr_share_balance(nwgr_pr_graph* g, uint32_t shares, uint32_t rounds, double* us_max, void* stream) {
  return guarded([&] {
    NWGR_REQUIRE(g != nullptr, "null argument");
    NWGR_REQUIRE(shares >= 1 && shares <= nwgr::kMaxGsPhases, "shares must be in [1, 64]");
    NWGR_REQUIRE(rounds <= 8, "rounds must be <= 8");
    if (us_max) *us_max = 0.0;
    if (g->n == 0) return;
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    const double w = nwgr::balance_shares(*g, shares, rounds, s);
    if (us_max) *us_max = w;
  });
}

int nwgr_pr_gs_phases(const nwgr_pr_graph* g, uint32_t phases, uint32_t* order, uint32_t* cuts) {
  return guarded([&] {
    NWGR_REQUIRE(g != nullptr, "null argument");
    NWGR_REQUIRE(phases >= 1 && phases <= nwgr::kMaxGsPhases, "phases must be in [1, 64]");
    if (cuts) {
      if (g->n == 0) {
        std::fill(cuts, cuts + phases + 1, 0u);
      } else if (g->gs.size() == phases && g->gs_cuts.size() == phases + 1) {
        std::copy(g->gs_cuts.begin(), g->gs_cuts.end(), cuts);  // the layouts' cuts (measured ones after a balance)
      } else {
        const std::vector<uint32_t> c = nwgr::gs_cut_rows(*g, phases);
        std::copy(c.begin(), c.end(), cuts);
      }
    }
    if (order && g->n) {
      nwgr::DeviceGuard dg(g->device);
      NWGR_HIP(hipMemcpy(order, g->perm.p, sizeof(uint32_t) * g->n, hipMemcpyDeviceToHost));
    }
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

uint32_t nwgr_pr_last_sweep_times(const nwgr_pr_graph* g, double* ms, uint32_t cap) {
  if (!g || !ms) return 0;
  const size_t k = std::min<size_t>({cap, g->last_profile.launch_ms.size(), g->last_errs.size()});
  std::copy(g->last_profile.launch_ms.begin(), g->last_profile.launch_ms.begin() + k, ms);
  return static_cast<uint32_t>(k);
}

int nwgr_pr_delta_run_device_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* d_r_out, nwgr_pr_result* res,
                                 void* stream) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    if (g->n == 0) return nwgr::empty_delta_result(*p, *res);
    NWGR_REQUIRE(d_r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    nwgr::run_delta(*g, *p, d_r_out, *res, s);
  });
}

int nwgr_pr_delta_run_f32(nwgr_pr_graph* g, const nwgr_pr_params* p, float* r_out, nwgr_pr_result* res) {
  return guarded([&] {
    NWGR_REQUIRE(g && p && res, "null argument");
    nwgr::validate_params(*p);
    if (g->n == 0) return nwgr::empty_delta_result(*p, *res);
    NWGR_REQUIRE(r_out != nullptr, "null argument");
    nwgr::DeviceGuard dg(g->device);
    hipStream_t s = g->own_stream;
    const size_t nb = sizeof(float) * g->n;
    if (g->out_buf.bytes < nb) g->out_buf.alloc(nb);
    nwgr::run_delta(*g, *p, g->out_buf.as<float>(), *res, s);
    nwgr::copy_out(*g, r_out, g->out_buf.p, nb, s);
  });
}

uint32_t nwgr_pr_last_errors(const nwgr_pr_graph* g, double* errs, uint32_t cap) {
  if (!g || !errs) return 0;
  uint32_t k = static_cast<uint32_t>(std::min<size_t>(cap, g->last_errs.size()));
  std::memcpy(errs, g->last_errs.data(), sizeof(double) * k);
  return k;
}

}  // extern "C"
