// Many independent small graphs solved in one persistent launch.
//
// The graphs are stored block-diagonally in one in-CSR: graph g owns the
// vertices [vbase[g], vbase[g+1]) and only edges among them (at most 65536
// vertices per graph, so entries are kept as 16-bit local ids).  Each
// workgroup of k_multi takes a graph from a queue (largest first), runs its
// whole Jacobi iteration (page_rank_v10, experimental/algorithms/
// page_rank.hpp:450-505, per graph: init 1/n_g, teleport (1-d)/n_g or the
// caller's vector, dangling LEAK / UNIFORM / TELEPORT, stop on the graph's
// own L1 change) with workgroup barriers between sweeps, then takes the next
// graph.  The contributions of a graph's first `lds_max_n` vertices (all of
// a smaller graph) are mirrored in LDS for the gathers.  Cross-workgroup
// communication is the queue counter alone (one RMW atomic per graph).
//
// Inside every graph the vertices are relabelled by non-increasing in-degree
// (stable), so rows of about one length are neighbours: a row gets a group
// of lanes sized to its length (a wavefront for rows of more than
// `heavy_min` entries, down to one lane for rows of at most 16), and the
// lanes of a group read consecutive entries of the row.  The caller's
// vectors (teleport, initial ranks, output) are read and written through the
// permutation.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "nwgr_internal.hpp"

namespace nwgr {

namespace {

constexpr int kMT = 1024;                 // threads per workgroup
constexpr uint32_t kMultiMaxN = 65536;    // vertices per graph (16-bit local ids)
constexpr uint32_t kMultiLdsMax = 36864;  // contributions mirrored in LDS at most (144 KiB)
constexpr uint32_t kDLeak = 0, kDUniform = 1, kDTeleport = 2;
constexpr int kU = 16;              // entries a lane loads per pass over its row
constexpr int kClasses = 7;         // row classes: class c gives a row 64 >> c lanes

struct MultiArgs {
  const uint32_t* __restrict__ offs;   // n + 1 (global)
  const uint16_t* __restrict__ lidx;   // m + padding, local ids
  const uint32_t* __restrict__ vbase;  // count + 1
  const uint32_t* __restrict__ perm;   // n: internal vertex -> caller vertex (inside its graph)
  const float* __restrict__ degf;      // n: divisor of the contribution
  const uint32_t* __restrict__ order;  // processing order of the graphs
  const float* __restrict__ tel;       // n or null
  const float* __restrict__ init;      // n or null
  float* r[2];
  float* c[2];
  float* __restrict__ out;             // n
  uint32_t* __restrict__ iters_out;    // count
  double* __restrict__ err_out;        // count
  uint32_t* queue;
  uint32_t m;  // entries; lidx[m .. m + 64) is zero padding
  uint32_t count, lds_max_n, heavy_min, max_iters, dangling;
  double tol;
  float d_hi, d_lo, one_minus_d;
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ float wsumf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// (a, b) summed over the workgroup in a fixed order; every thread gets them.
__device__ __forceinline__ void bsum2(double& a, double& b, double (*s)[kMT / 64]) {
  a = wsum(a);
  b = wsum(b);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    s[0][w] = a;
    s[1][w] = b;
  }
  __syncthreads();
  double ta = 0, tb = 0;
#pragma unroll
  for (int i = 0; i < kMT / 64; ++i) {
    ta += s[0][i];
    tb += s[1][i];
  }
  a = ta;
  b = tb;
}

// Where a graph's gathered contributions live: all in the LDS mirror (the
// graph has at most lds_max_n vertices), its first lds_max_n vertices there
// and the rest in global memory (after the in-degree relabelling the first
// vertices are the most gathered ones), or all in global memory.
// A graph of at most lds_max_n / 2 vertices keeps two LDS copies (this
// sweep's and the next), so its contributions never go to global memory.
constexpr int kMirrorAll = 0, kMirrorPrefix = 1, kMirrorNone = 2, kMirrorDouble = 3;

template <int M>
__device__ __forceinline__ float cget(const float* __restrict__ cl, const float* __restrict__ cg, uint32_t K,
                                      uint32_t j) {
  if constexpr (M == kMirrorAll || M == kMirrorDouble) return cl[j];
  else if constexpr (M == kMirrorPrefix) return j < K ? cl[j] : cg[j];
  else return cg[j];
}

// One graph, start to finish, by the whole workgroup.
// Row classes of a graph (its rows are in non-increasing degree order):
// class 0 takes the rows of more than min(heavy_min, kU * 32) entries, a
// wavefront each; class c >= 1 the rows of at most min(heavy_min,
// kU * (64 >> c)) entries, down to the next class's bound, 64 >> c lanes each
// (at most kU entries per lane and pass).  band[c] = first row of class c,
// band[kClasses] = n_g.
__device__ __forceinline__ void class_bounds(const MultiArgs& a, uint32_t v0, uint32_t ng, uint32_t tid,
                                             uint32_t* s_band) {
  if (tid + 1 < kClasses) {
    const uint32_t c = tid + 1;
    const uint32_t thr = min(a.heavy_min, static_cast<uint32_t>(kU) * (64u >> c));
    uint32_t lo = 0, hi = ng;  // first row with at most thr entries
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.offs[v0 + mid + 1] - a.offs[v0 + mid] <= thr) hi = mid; else lo = mid + 1;
    }
    s_band[c] = lo;
  } else if (tid + 1 == kClasses) {
    s_band[0] = 0;
    s_band[kClasses] = ng;
  }
}

template <int M>
__device__ void solve_graph(const MultiArgs& a, uint32_t g, float* cl, double (*s_red)[kMT / 64],
                            uint32_t* s_band) {
  const uint32_t tid = threadIdx.x, wl = tid & 63, wv = tid >> 6;
  const uint32_t v0 = a.vbase[g], ng = a.vbase[g + 1] - v0;
  // vertices [0, K) are mirrored in LDS
  const uint32_t K = (M == kMirrorAll || M == kMirrorDouble) ? ng : (M == kMirrorPrefix ? a.lds_max_n : 0u);
  const uint32_t half = a.lds_max_n / 2;  // second LDS copy (kMirrorDouble)
  const float un = 1.0f / static_cast<float>(ng);
  const float base_u = static_cast<float>((1.0 - (static_cast<double>(a.d_hi) + a.d_lo)) / ng);
  // start: r = init or 1/n_g, contributions, dangling mass
  double D = 0, unused = 0;
  for (uint32_t v = tid; v < ng; v += kMT) {
    const float r = a.init ? a.init[a.perm[v0 + v]] : un;
    const float dg = a.degf[v0 + v];
    a.r[0][v0 + v] = r;
    const float c = r / dg;
    if constexpr (M != kMirrorDouble) a.c[0][v0 + v] = c;
    if (v < K) cl[v] = c;
    if (dg == 0.f) D += static_cast<double>(r);
  }
  class_bounds(a, v0, ng, tid, s_band);
  bsum2(D, unused, s_red);  // barrier: LDS, the start vectors and the classes are complete
  uint32_t band[kClasses + 1], ucum[kClasses + 1];  // class bounds, units before each class
#pragma unroll
  for (int k = 0; k <= kClasses; ++k) band[k] = __builtin_amdgcn_readfirstlane(s_band[k]);
  ucum[0] = 0;
#pragma unroll
  for (int k = 0; k < kClasses; ++k) ucum[k + 1] = ucum[k] + ((band[k + 1] - band[k] + (1u << k) - 1u) >> k);
  const uint32_t nunits = ucum[kClasses];
  uint32_t p = 0, it = 0;
  double E = 0;
  while (it < a.max_iters) {
    const float* __restrict__ cg = a.c[p] + v0;
    const float* __restrict__ rin = a.r[p] + v0;
    float* __restrict__ rout = a.r[p ^ 1] + v0;
    float* __restrict__ cout = a.c[p ^ 1] + v0;
    // LDS copy gathered in this sweep, and (kMirrorDouble) the one written
    const float* __restrict__ lcur = M == kMirrorDouble ? cl + (p ? half : 0u) : cl;
    float* __restrict__ lnext = cl + (p ? 0u : half);
    const float Df = static_cast<float>(D);
    double err = 0, dang = 0;
    // row epilogue: the new rank (returned), its contribution (returned, and
    // kept in LDS for kMirrorDouble), residual and dangling mass
    auto epilogue = [&](uint32_t v, float z, float t, float dg, float ro, float& c) {
      const float base = a.tel ? a.one_minus_d * t : base_u;
      const float u = a.dangling == kDUniform ? un : (a.dangling == kDTeleport ? t : 0.f);
      const float s = fmaf(Df, u, z);
      const float rn = fmaf(a.d_hi, s, fmaf(a.d_lo, s, base));
      c = rn / dg;
      if constexpr (M == kMirrorDouble) lnext[v] = c;
      err += fabs(static_cast<double>(rn) - static_cast<double>(ro));
      if (dg == 0.f) dang += static_cast<double>(rn);
      return rn;
    };
    auto finish = [&](uint32_t v, float z, float t, float dg, float ro) {
      float c;
      rout[v] = epilogue(v, z, t, dg, ro, c);
      if constexpr (M != kMirrorDouble) cout[v] = c;
    };
    // Rows by class: a class-c unit is 2^c consecutive rows with 64 >> c
    // lanes each, and wavefront wv takes the units wv, wv + 16, ... of all
    // classes in turn.  The lanes of a row load its entries j, j + G, ...
    // (kU of them per pass, no bounds branch: loads past the row's end read
    // the padding of lidx and are masked to 0), so the lanes of a load read
    // consecutive entries; a butterfly over the row's lanes adds the lane
    // sums, and the row's first lane finishes the row.  Two-stage pipeline:
    // the bounds and epilogue operands of the unit after next and the first
    // pass of entries of the next unit are loaded while a unit is gathered.
    const uint32_t pad = a.m;  // an index into the padding of lidx
    struct Unit {
      uint32_t rr, b, e, lg, j;
      float t, dg, ro;
      bool fin;  // the lane finishes row rr
    };
    auto unit_at = [&](uint32_t u) {
      Unit o;
      uint32_t c = 0, first = 0, rlo = 0, rhi = 0;
#pragma unroll
      for (int k = 0; k < kClasses; ++k) {
        if (u >= ucum[k]) {
          c = k;
          first = ucum[k];
          rlo = band[k];
          rhi = band[k + 1];
        }
      }
      o.lg = 6u - c;  // log2 of the lanes per row
      o.j = wl & ((1u << o.lg) - 1u);
      const uint32_t row = rlo + ((u - first) << c) + (wl >> o.lg);
      // past the last unit every lane takes row 0 and no entries
      const bool ok = u < nunits && row < rhi;
      o.rr = ok ? row : (u < nunits ? rlo : 0u);
      o.b = a.offs[v0 + o.rr];
      o.e = ok ? a.offs[v0 + o.rr + 1] : o.b;
      o.t = a.tel ? a.tel[a.perm[v0 + o.rr]] : un;
      o.dg = a.degf[v0 + o.rr];
      o.ro = rin[o.rr];
      o.fin = ok && o.j == 0;
      return o;
    };
    auto first_pass = [&](const Unit& o, uint32_t (&jj)[kU]) {
      const uint32_t G = 1u << o.lg;
#pragma unroll
      for (int q = 0; q < kU; ++q) {
        const uint32_t k = o.b + o.j + q * G;
        jj[q] = a.lidx[k < o.e ? k : pad];
      }
    };
    Unit cu = unit_at(wv), nu = unit_at(wv + kMT / 64);
    uint32_t jc[kU];
    first_pass(cu, jc);
    for (uint32_t u = wv; u < nunits; u += kMT / 64) {
      const Unit nn = unit_at(u + 2 * (kMT / 64));
      uint32_t jn[kU];
      first_pass(nu, jn);
      const uint32_t G = 1u << cu.lg;
      float z = 0.f;
#pragma unroll
      for (int q = 0; q < kU; ++q) {
        const float x = cget<M>(lcur, cg, K, jc[q]);
        z += cu.b + cu.j + q * G < cu.e ? x : 0.f;
      }
      for (uint32_t k = cu.b + cu.j + G * kU; k < cu.e; k += G * kU) {
        uint32_t jj[kU];
#pragma unroll
        for (int q = 0; q < kU; ++q) jj[q] = a.lidx[k + q * G < cu.e ? k + q * G : pad];
#pragma unroll
        for (int q = 0; q < kU; ++q) {
          const float x = cget<M>(lcur, cg, K, jj[q]);
          z += k + q * G < cu.e ? x : 0.f;
        }
      }
      for (uint32_t o = 1; o < G; o <<= 1) z += __shfl_xor(z, static_cast<int>(o));
      if (cu.fin) finish(cu.rr, z, cu.t, cu.dg, cu.ro);
      cu = nu;
      nu = nn;
#pragma unroll
      for (int q = 0; q < kU; ++q) jc[q] = jn[q];
    }
    bsum2(err, dang, s_red);  // barrier: every gather of this sweep is done
    if constexpr (M == kMirrorAll || M == kMirrorPrefix) {
      for (uint32_t v = tid; v < K; v += kMT) cl[v] = cout[v];
      __syncthreads();
    }
    p ^= 1;
    ++it;
    E = err;
    D = dang;
    if (E < a.tol) break;
  }
  const float* fin = a.r[p] + v0;
  for (uint32_t v = tid; v < ng; v += kMT) a.out[a.perm[v0 + v]] = fin[v];
  if (tid == 0) {
    a.iters_out[g] = it;
    a.err_out[g] = E;
  }
}

__global__ __launch_bounds__(kMT) void k_multi(MultiArgs a) {
  extern __shared__ float cl[];
  __shared__ double s_red[2][kMT / 64];
  __shared__ uint32_t s_g;
  __shared__ uint32_t s_band[kClasses + 1];
  for (;;) {
    if (threadIdx.x == 0) {
      const uint32_t q = __hip_atomic_fetch_add(a.queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_g = q < a.count ? a.order[q] : 0xffffffffu;
    }
    __syncthreads();
    const uint32_t g = s_g;
    __syncthreads();
    if (g == 0xffffffffu) return;
    const uint32_t ng = a.vbase[g + 1] - a.vbase[g];
    if (ng <= a.lds_max_n / 2)
      solve_graph<kMirrorDouble>(a, g, cl, s_red, s_band);
    else if (ng <= a.lds_max_n)
      solve_graph<kMirrorAll>(a, g, cl, s_red, s_band);
    else if (a.lds_max_n > 0)
      solve_graph<kMirrorPrefix>(a, g, cl, s_red, s_band);
    else
      solve_graph<kMirrorNone>(a, g, cl, s_red, s_band);
    __syncthreads();
  }
}

// graph id of every vertex
__global__ void k_vertex_graph(const uint32_t* __restrict__ vbase, uint32_t count, uint32_t* __restrict__ gid) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x)
    for (uint32_t v = vbase[g]; v < vbase[g + 1]; ++v) gid[v] = g;
}

// Relabelled in-CSR: internal row R holds the entries of caller row perm[R],
// in their order, as local ids of the relabelled graph; a flag for an entry
// whose source lies outside its row's graph.
__global__ void k_localize(const uint32_t* __restrict__ offs_in, const uint32_t* __restrict__ idx,
                           const uint32_t* __restrict__ offs_new, const uint32_t* __restrict__ perm,
                           const uint32_t* __restrict__ iperm, const uint32_t* __restrict__ vbase,
                           const uint32_t* __restrict__ gid, uint32_t n, uint16_t* __restrict__ lidx,
                           uint32_t* __restrict__ bad) {
  for (uint32_t R = blockIdx.x * blockDim.x + threadIdx.x; R < n; R += gridDim.x * blockDim.x) {
    const uint32_t g = gid[R];
    const uint32_t lo = vbase[g], hi = vbase[g + 1];
    const uint32_t o = perm[R];
    const uint32_t b = offs_in[o], e = offs_in[o + 1];
    uint32_t k = offs_new[R];
    for (uint32_t q = b; q < e; ++q, ++k) {
      const uint32_t u = idx[q];
      if (u < lo || u >= hi) {
        atomicOr(bad, 1u);
        lidx[k] = 0;
      } else {
        lidx[k] = static_cast<uint16_t>(iperm[u] - lo);
      }
    }
  }
}

// divisor of internal vertex R: the caller's degree of perm[R]
__global__ void k_degf_perm(const uint32_t* __restrict__ deg, const uint32_t* __restrict__ perm, uint32_t n,
                            float* __restrict__ out) {
  for (uint32_t R = blockIdx.x * blockDim.x + threadIdx.x; R < n; R += gridDim.x * blockDim.x)
    out[R] = static_cast<float>(deg[perm[R]]);
}

__global__ void k_out_degree_u32(const uint32_t* __restrict__ idx, uint64_t m, uint32_t* __restrict__ deg) {
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m;
       k += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    atomicAdd(&deg[idx[k]], 1u);
}

inline uint32_t nblk(uint64_t work, int t = 256) {
  return static_cast<uint32_t>(std::clamp<uint64_t>((work + t - 1) / t, 1, 65535u * 4));
}

}  // namespace

struct MultiGraph {
  int device = 0;
  uint32_t n = 0, count = 0;
  uint64_t m = 0;
  uint32_t num_cu = 0;
  uint32_t lds_max_n = kMultiLdsMax;
  uint32_t heavy_min = 512;
  DevBuf offs, lidx, vbase, perm, degf, order, r[2], c[2], iters, errs, queue;
  std::vector<uint32_t> offs_h, vbase_h;
  // per-graph results of the last run, copied into pinned host memory
  uint32_t* iters_h = nullptr;
  double* errs_h = nullptr;
  bool has_run = false;
  hipStream_t stream = nullptr;
  double setup_ms = 0;

  ~MultiGraph() {
    if (iters_h) (void)hipHostFree(iters_h);
    if (errs_h) (void)hipHostFree(errs_h);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {


MultiGraph* multi_create(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                         const uint32_t* vbase, uint32_t count, const uint32_t* degrees, int device) {
  NWGR_REQUIRE(offsets && vbase, "null argument");
  NWGR_REQUIRE(m == 0 || indices, "indices is null");
  NWGR_REQUIRE(count >= 1, "count must be >= 1");
  // (a row's lanes form entry positions up to 4 * 64 past its end)
  NWGR_REQUIRE(m < 0xffffffffULL - 256, "m must fit the uint32 offsets with room for the row passes");
  NWGR_REQUIRE(offsets[0] == 0 && offsets[n] == m, "offsets must start at 0 and end at m");
  NWGR_REQUIRE(vbase[0] == 0 && vbase[count] == n, "vbase must start at 0 and end at n");
  for (uint32_t g = 0; g < count; ++g) {
    NWGR_REQUIRE(vbase[g + 1] > vbase[g], "every graph needs at least one vertex");
    NWGR_REQUIRE(vbase[g + 1] - vbase[g] <= kMultiMaxN, "a graph has more than 65536 vertices");
  }
  for (uint32_t v = 0; v < n; ++v) NWGR_REQUIRE(offsets[v + 1] >= offsets[v], "offsets must be non-decreasing");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(NWGR_EDEVICE, "no HIP device available");
  NWGR_REQUIRE(device >= 0 && device < ndev, "device ordinal out of range");
  DeviceGuard dg(device);
  const auto t0 = std::chrono::steady_clock::now();
  auto mg = std::make_unique<MultiGraph>();
  mg->device = device;
  mg->n = n;
  mg->m = m;
  mg->count = count;
  hipDeviceProp_t prop;
  NWGR_HIP(hipGetDeviceProperties(&prop, device));
  mg->num_cu = static_cast<uint32_t>(prop.multiProcessorCount);
  NWGR_HIP(hipStreamCreateWithFlags(&mg->stream, hipStreamNonBlocking));
  hipStream_t s = mg->stream;
  // per graph, rows by non-increasing in-degree (stable): perm[new] = old
  std::vector<uint32_t> perm(n), iperm(n);
  {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < hw; ++t)
      th.emplace_back([&, t] {
        for (uint32_t g = t; g < count; g += hw) {
          const uint32_t lo = vbase[g], hi = vbase[g + 1];
          std::iota(perm.begin() + lo, perm.begin() + hi, lo);
          std::stable_sort(perm.begin() + lo, perm.begin() + hi, [&](uint32_t x, uint32_t y) {
            return offsets[x + 1] - offsets[x] > offsets[y + 1] - offsets[y];
          });
          for (uint32_t R = lo; R < hi; ++R) iperm[perm[R]] = R;
        }
      });
    for (auto& x : th) x.join();
  }
  mg->offs_h.assign(n + 1, 0);
  for (uint32_t R = 0; R < n; ++R) mg->offs_h[R + 1] = mg->offs_h[R] + (offsets[perm[R] + 1] - offsets[perm[R]]);
  mg->vbase_h.assign(vbase, vbase + count + 1);
  mg->offs.alloc(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  mg->vbase.alloc(sizeof(uint32_t) * (count + 1));
  mg->perm.alloc(sizeof(uint32_t) * n);
  NWGR_HIP(hipMemcpyAsync(mg->offs.p, mg->offs_h.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, s));
  NWGR_HIP(hipMemcpyAsync(mg->vbase.p, vbase, sizeof(uint32_t) * (count + 1), hipMemcpyHostToDevice, s));
  NWGR_HIP(hipMemcpyAsync(mg->perm.p, perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
  DevBuf d_idx, d_offs, d_iperm, gid, bad;
  d_idx.alloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  d_offs.alloc(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  d_iperm.alloc(sizeof(uint32_t) * n);
  if (m) NWGR_HIP(hipMemcpyAsync(d_idx.p, indices, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s));
  NWGR_HIP(hipMemcpyAsync(d_offs.p, offsets, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, s));
  NWGR_HIP(hipMemcpyAsync(d_iperm.p, iperm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
  mg->lidx.alloc(sizeof(uint16_t) * (m + 64));
  NWGR_HIP(hipMemsetAsync(mg->lidx.p, 0, mg->lidx.bytes, s));
  gid.alloc(sizeof(uint32_t) * n);
  bad.alloc(sizeof(uint32_t));
  NWGR_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
  k_vertex_graph<<<nblk(count), 256, 0, s>>>(mg->vbase.as<uint32_t>(), count, gid.as<uint32_t>());
  k_localize<<<nblk(n), 256, 0, s>>>(d_offs.as<uint32_t>(), d_idx.as<uint32_t>(), mg->offs.as<uint32_t>(),
                                     mg->perm.as<uint32_t>(), d_iperm.as<uint32_t>(), mg->vbase.as<uint32_t>(),
                                     gid.as<uint32_t>(), n, mg->lidx.as<uint16_t>(), bad.as<uint32_t>());
  NWGR_HIP(hipGetLastError());
  // divisors: the caller's degrees, or out-degrees
  mg->degf.alloc(sizeof(float) * n);
  DevBuf deg;
  deg.alloc(sizeof(uint32_t) * n);
  if (degrees) {
    NWGR_HIP(hipMemcpyAsync(deg.p, degrees, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
  } else {
    NWGR_HIP(hipMemsetAsync(deg.p, 0, deg.bytes, s));
    if (m) k_out_degree_u32<<<nblk(m), 256, 0, s>>>(d_idx.as<uint32_t>(), m, deg.as<uint32_t>());
  }
  k_degf_perm<<<nblk(n), 256, 0, s>>>(deg.as<uint32_t>(), mg->perm.as<uint32_t>(), n, mg->degf.as<float>());
  NWGR_HIP(hipGetLastError());
  uint32_t badh = 0;
  NWGR_HIP(hipMemcpyAsync(&badh, bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  NWGR_REQUIRE(badh == 0, "an edge joins two different graphs");
  // largest graphs first (by entries, then vertices)
  std::vector<uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    const uint64_t mx = offsets[vbase[x + 1]] - offsets[vbase[x]], my = offsets[vbase[y + 1]] - offsets[vbase[y]];
    return mx != my ? mx > my : (vbase[x + 1] - vbase[x]) > (vbase[y + 1] - vbase[y]);
  });
  mg->order.alloc(sizeof(uint32_t) * count);
  NWGR_HIP(hipMemcpyAsync(mg->order.p, order.data(), sizeof(uint32_t) * count, hipMemcpyHostToDevice, s));
  for (int b = 0; b < 2; ++b) {
    mg->r[b].alloc(sizeof(float) * n);
    mg->c[b].alloc(sizeof(float) * n);
  }
  mg->iters.alloc(sizeof(uint32_t) * count);
  mg->errs.alloc(sizeof(double) * count);
  mg->queue.alloc(sizeof(uint32_t));
  mg->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return mg.release();
}

void multi_run(MultiGraph& mg, const nwgr_pr_params& p, float* d_out, nwgr_pr_result& res, hipStream_t s) {
  validate_params(p);
  {
    DevBuf flag;
    flag.alloc(sizeof(uint32_t));
    check_vectors(p, mg.n, flag.as<uint32_t>(), s);
  }
  static const bool attrs = [] {
    NWGR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_multi), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(kMultiLdsMax * sizeof(float))));
    return true;
  }();
  (void)attrs;
  MultiArgs a{};
  a.offs = mg.offs.as<uint32_t>();
  a.lidx = mg.lidx.as<uint16_t>();
  a.vbase = mg.vbase.as<uint32_t>();
  a.perm = mg.perm.as<uint32_t>();
  a.degf = mg.degf.as<float>();
  a.order = mg.order.as<uint32_t>();
  a.tel = p.teleport;
  a.init = p.init;
  for (int b = 0; b < 2; ++b) {
    a.r[b] = mg.r[b].as<float>();
    a.c[b] = mg.c[b].as<float>();
  }
  a.out = d_out;
  a.iters_out = mg.iters.as<uint32_t>();
  a.err_out = mg.errs.as<double>();
  a.queue = mg.queue.as<uint32_t>();
  a.m = static_cast<uint32_t>(mg.m);
  a.count = mg.count;
  a.lds_max_n = mg.lds_max_n;
  a.heavy_min = mg.heavy_min;
  a.max_iters = p.max_iters;
  a.dangling = p.dangling;
  a.tol = p.tol;
  const double d = p.damping;
  a.d_hi = static_cast<float>(d);
  a.d_lo = static_cast<float>(d - static_cast<double>(a.d_hi));
  a.one_minus_d = static_cast<float>(1.0 - d);
  hipEvent_t e0, e1;
  NWGR_HIP(hipEventCreate(&e0));
  NWGR_HIP(hipEventCreate(&e1));
  NWGR_HIP(hipMemsetAsync(mg.queue.p, 0, sizeof(uint32_t), s));
  NWGR_HIP(hipEventRecord(e0, s));
  const size_t shm = sizeof(float) * mg.lds_max_n;
  // one persistent workgroup per CU: a 1024-thread workgroup at this kernel's
  // register count (over 64 VGPRs) takes all of a CU's wavefront slots, so a
  // second one could not run beside it even when two LDS mirrors would fit
  // (capping the kernel at 64 VGPRs spills; 512-thread workgroups two per CU
  // measured 317 against 380 GTEPS on 2,048 R-MAT graphs)
  k_multi<<<std::min<uint32_t>(mg.num_cu, mg.count), kMT, shm, s>>>(a);
  const hipError_t le = hipGetLastError();
  NWGR_HIP(hipEventRecord(e1, s));
  NWGR_HIP(hipEventSynchronize(e1));
  NWGR_HIP(le);
  float ms = 0;
  NWGR_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!mg.iters_h) NWGR_HIP(hipHostMalloc(reinterpret_cast<void**>(&mg.iters_h), sizeof(uint32_t) * mg.count));
  if (!mg.errs_h) NWGR_HIP(hipHostMalloc(reinterpret_cast<void**>(&mg.errs_h), sizeof(double) * mg.count));
  NWGR_HIP(hipMemcpyAsync(mg.iters_h, mg.iters.p, sizeof(uint32_t) * mg.count, hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipMemcpyAsync(mg.errs_h, mg.errs.p, sizeof(double) * mg.count, hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  mg.has_run = true;
  uint32_t mx = 0;
  double emax = 0;
  bool all = true;
  for (uint32_t g = 0; g < mg.count; ++g) {
    mx = std::max(mx, mg.iters_h[g]);
    emax = std::max(emax, mg.errs_h[g]);
    all = all && mg.iters_h[g] > 0 && mg.errs_h[g] < p.tol;
  }
  res.iters = mx;
  res.error = emax;
  res.converged = all ? 1u : 0u;
  res.ms = ms;
}

}  // namespace

}  // namespace nwgr

struct nwgr_multi_graph : nwgr::MultiGraph {};

extern "C" {

int nwgr_multi_create(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                      const uint32_t* vbase, uint32_t count, const uint32_t* degrees, int device,
                      nwgr_multi_graph** out) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    *out = static_cast<nwgr_multi_graph*>(nwgr::multi_create(offsets, indices, n, m, vbase, count, degrees, device));
  });
}

void nwgr_multi_destroy(nwgr_multi_graph* mg) {
  if (!mg) return;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(mg->device);
  delete static_cast<nwgr::MultiGraph*>(mg);
  (void)hipSetDevice(prev);
}

int nwgr_multi_run_device_f32(nwgr_multi_graph* mg, const nwgr_pr_params* p, float* d_r_out, nwgr_pr_result* res,
                              void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(mg && p && d_r_out && res, "null argument");
    nwgr::DeviceGuard dg(mg->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : mg->stream;
    nwgr::multi_run(*mg, *p, d_r_out, *res, s);
  });
}

int nwgr_multi_run_f32(nwgr_multi_graph* mg, const nwgr_pr_params* p, float* r_out, nwgr_pr_result* res) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(mg && p && r_out && res, "null argument");
    nwgr::DeviceGuard dg(mg->device);
    hipStream_t s = mg->stream;
    nwgr::DevBuf out, tel, ini;
    out.alloc(sizeof(float) * mg->n);
    nwgr_pr_params q = *p;
    if (p->teleport) {
      tel.alloc(sizeof(float) * mg->n);
      NWGR_HIP(hipMemcpyAsync(tel.p, p->teleport, sizeof(float) * mg->n, hipMemcpyHostToDevice, s));
      q.teleport = tel.as<float>();
    }
    if (p->init) {
      ini.alloc(sizeof(float) * mg->n);
      NWGR_HIP(hipMemcpyAsync(ini.p, p->init, sizeof(float) * mg->n, hipMemcpyHostToDevice, s));
      q.init = ini.as<float>();
    }
    nwgr::multi_run(*mg, q, out.as<float>(), *res, s);
    NWGR_HIP(hipMemcpy(r_out, out.p, sizeof(float) * mg->n, hipMemcpyDeviceToHost));
  });
}

int nwgr_multi_graph_results(const nwgr_multi_graph* mg, uint32_t* iters, double* errors) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(mg != nullptr, "null argument");
    NWGR_REQUIRE(mg->has_run, "no run yet");
    if (iters) std::copy(mg->iters_h, mg->iters_h + mg->count, iters);
    if (errors) std::copy(mg->errs_h, mg->errs_h + mg->count, errors);
  });
}

int nwgr_multi_get_tuning(const nwgr_multi_graph* mg, uint32_t* lds_max_n, uint32_t* heavy_min) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(mg != nullptr, "null argument");
    if (lds_max_n) *lds_max_n = mg->lds_max_n;
    if (heavy_min) *heavy_min = mg->heavy_min;
  });
}

int nwgr_multi_set_tuning(nwgr_multi_graph* mg, uint32_t lds_max_n, uint32_t heavy_min) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(mg != nullptr, "null argument");
    NWGR_REQUIRE(lds_max_n <= nwgr::kMultiLdsMax, "lds_max_n must be <= 36864");
    NWGR_REQUIRE(heavy_min >= 1, "heavy_min must be >= 1");
    nwgr::DeviceGuard dg(mg->device);
    mg->lds_max_n = lds_max_n;
    mg->heavy_min = heavy_min;
  });
}

uint32_t nwgr_multi_n(const nwgr_multi_graph* mg) { return mg ? mg->n : 0; }
uint32_t nwgr_multi_count(const nwgr_multi_graph* mg) { return mg ? mg->count : 0; }
double nwgr_multi_setup_ms(const nwgr_multi_graph* mg) { return mg ? mg->setup_ms : 0.0; }

}  // extern "C"
