// Synthetic graph generators and the host in-CSR builder (host C++20, std::thread).
//
// Semantics follow the reference's CSR construction: fill_directed<1>
// (build.hpp:209-265) buckets every edge u->v into row v, uniq
// (build.hpp:425-435) and remove_self_loops (build.hpp:437-443) are applied
// when requested, and out-degrees are counted as build_degrees does
// (bench/common.hpp:109-120).  Rows come out sorted ascending.  The R-MAT and
// Erdos-Renyi recipes have no reference counterpart; they are specified in
// include/nwgr_graphgen.h.

#include "nwgr_graphgen.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "host_internal.hpp"

namespace {
thread_local std::string g_err;
}  // namespace

namespace nwgr_host {
int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

int pick_threads(int nthreads) {
  if (nthreads > 0) return nthreads;
  unsigned hc = std::thread::hardware_concurrency();
  return hc ? static_cast<int>(std::min(hc, 64u)) : 1;
}
}  // namespace nwgr_host

namespace {

using nwgr_host::fail;
using nwgr_host::pick_threads;

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// k-th output (0-based) of SplitMix64 seeded with key.
inline uint64_t sm64_at(uint64_t key, uint64_t k) { return mix64(key + (k + 1) * kGolden); }

// Static split of [0, n) over T threads; fn(lo, hi).
template <class F>
void par_range(uint64_t n, int T, F&& fn) {
  if (T <= 1 || n < 4096) {
    fn(uint64_t{0}, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T);
  std::exception_ptr err;
  std::atomic<bool> failed{false};
  for (int t = 0; t < T; ++t) {
    uint64_t lo = n * t / T, hi = n * (t + 1) / T;
    th.emplace_back([&, lo, hi] {
      try {
        fn(lo, hi);
      } catch (...) {
        if (!failed.exchange(true)) err = std::current_exception();
      }
    });
  }
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
}

// Dynamic chunks of [0, n); fn(lo, hi).
template <class F>
void par_dynamic(uint64_t n, uint64_t chunk, int T, F&& fn) {
  if (T <= 1 || n <= chunk) {
    fn(uint64_t{0}, n);
    return;
  }
  std::atomic<uint64_t> next{0};
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 0; t < T; ++t) {
    th.emplace_back([&] {
      for (;;) {
        uint64_t lo = next.fetch_add(chunk, std::memory_order_relaxed);
        if (lo >= n) break;
        fn(lo, std::min(n, lo + chunk));
      }
    });
  }
  for (auto& x : th) x.join();
}

struct RmatThresholds {
  uint32_t t1, t2, t3;
};

uint32_t to_u32_threshold(double p) {
  double x = std::nearbyint(p * 4294967296.0);
  if (x <= 0) return 0;
  if (x >= 4294967295.0) return 0xffffffffu;
  return static_cast<uint32_t>(x);
}

RmatThresholds rmat_thresholds(double a, double b, double c) {
  return {to_u32_threshold(a), to_u32_threshold(a + b), to_u32_threshold(a + b + c)};
}

// Draw number `e` of an R-MAT at `scale` bits: returns (src, dst).
inline void rmat_draw(uint64_t key, uint64_t e, uint32_t scale, const RmatThresholds& th, uint32_t& u,
                      uint32_t& v) {
  const uint64_t W = (scale + 1) / 2;
  uint32_t su = 0, sv = 0;
  uint64_t word = 0;
  for (uint32_t l = 0; l < scale; ++l) {
    if ((l & 1) == 0) word = sm64_at(key, e * W + l / 2);
    uint32_t r = (l & 1) ? static_cast<uint32_t>(word) : static_cast<uint32_t>(word >> 32);
    uint32_t bit = 1u << (scale - 1 - l);
    if (r < th.t1) {
    } else if (r < th.t2) {
      sv |= bit;
    } else if (r < th.t3) {
      su |= bit;
    } else {
      su |= bit;
      sv |= bit;
    }
  }
  u = su;
  v = sv;
}

inline void er_draw(uint64_t key, uint64_t e, uint32_t n, uint32_t& u, uint32_t& v) {
  uint64_t w = sm64_at(key, e);
  u = static_cast<uint32_t>(((w >> 32) * static_cast<uint64_t>(n)) >> 32);
  v = static_cast<uint32_t>(((w & 0xffffffffULL) * static_cast<uint64_t>(n)) >> 32);
}

// Two-pass bucket build.  `emit(lo, hi, sink)` must call sink(u, v) for each
// draw in [lo, hi) and produce the same draws on every call.
template <class Emit>
int build_from_draws(uint32_t n, uint64_t draws, uint32_t flags, int T, Emit&& emit,
                     nwgr_host_csr** out) {
  const bool drop_loops = flags & NWGR_CSR_DROP_SELF_LOOPS;
  const bool dedup = flags & NWGR_CSR_DEDUP;
  auto g = std::make_unique<nwgr_host_csr>();
  g->n = n;
  std::vector<uint32_t> cnt(n, 0);
  // pass 1: count entries per destination row
  par_range(draws, T, [&](uint64_t lo, uint64_t hi) {
    emit(lo, hi, [&](uint32_t u, uint32_t v) {
      if (drop_loops && u == v) return;
      std::atomic_ref<uint32_t>(cnt[v]).fetch_add(1, std::memory_order_relaxed);
    });
  });
  std::vector<uint64_t> start(static_cast<size_t>(n) + 1, 0);
  for (uint32_t v = 0; v < n; ++v) start[v + 1] = start[v] + cnt[v];
  const uint64_t total = start[n];
  std::vector<uint32_t> tmp(total);
  std::vector<uint64_t> cursor(start.begin(), start.end() - 1);
  // pass 2: scatter sources into their rows (order fixed up by the sort below)
  par_range(draws, T, [&](uint64_t lo, uint64_t hi) {
    emit(lo, hi, [&](uint32_t u, uint32_t v) {
      if (drop_loops && u == v) return;
      uint64_t p = std::atomic_ref<uint64_t>(cursor[v]).fetch_add(1, std::memory_order_relaxed);
      tmp[p] = u;
    });
  });
  cursor.clear();
  cursor.shrink_to_fit();
  // per row: sort, optionally unique
  par_dynamic(n, 4096, T, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t v = lo; v < hi; ++v) {
      uint32_t* b = tmp.data() + start[v];
      uint32_t* e = tmp.data() + start[v + 1];
      std::sort(b, e);
      cnt[v] = static_cast<uint32_t>(dedup ? (std::unique(b, e) - b) : (e - b));
    }
  });
  uint64_t m = 0;
  g->offsets.resize(static_cast<size_t>(n) + 1);
  for (uint32_t v = 0; v < n; ++v) {
    if (m > 0xffffffffULL) return fail("edge count exceeds the uint32 offset range");
    g->offsets[v] = static_cast<uint32_t>(m);
    m += cnt[v];
  }
  if (m > 0xffffffffULL) return fail("edge count exceeds the uint32 offset range");
  g->offsets[n] = static_cast<uint32_t>(m);
  g->m = m;
  g->indices.resize(m);
  par_dynamic(n, 4096, T, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t v = lo; v < hi; ++v)
      std::memcpy(g->indices.data() + g->offsets[v], tmp.data() + start[v],
                  sizeof(uint32_t) * cnt[v]);
  });
  tmp.clear();
  tmp.shrink_to_fit();
  g->outdeg.assign(n, 0);
  par_range(m, T, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t k = lo; k < hi; ++k)
      std::atomic_ref<uint32_t>(g->outdeg[g->indices[k]]).fetch_add(1, std::memory_order_relaxed);
  });
  *out = g.release();
  return 0;
}

int rmat_n_single(uint32_t n, uint64_t draws, uint64_t key, const RmatThresholds& th,
                  nwgr_host_csr** out) {
  uint32_t scale = 0;
  while ((uint64_t{1} << scale) < n) ++scale;
  if (scale == 0) scale = 1;
  std::vector<uint32_t> su, sv;
  su.reserve(draws);
  sv.reserve(draws);
  uint64_t c = 0;
  // Acceptance probability is >= ~1/4 per endpoint pair for n > 2^(S-1), so
  // the loop ends; the bound guards against pathological thresholds.
  const uint64_t cap = draws * 4096 + 1024;
  while (su.size() < draws && c < cap) {
    uint32_t u, v;
    rmat_draw(key, c++, scale, th, u, v);
    if (u < n && v < n) {
      su.push_back(u);
      sv.push_back(v);
    }
  }
  if (su.size() < draws) return fail("R-MAT rejection sampling did not terminate");
  return nwgr_csr_build(n, su.size(), su.data(), sv.data(), NWGR_CSR_DEDUP | NWGR_CSR_DROP_SELF_LOOPS,
                        1, out);
}

}  // namespace

extern "C" {

uint64_t nwgr_splitmix64_at(uint64_t key, uint64_t k) { return sm64_at(key, k); }

const char* nwgr_gen_last_error(void) { return g_err.c_str(); }

int nwgr_csr_build(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, uint32_t flags,
                   int nthreads, nwgr_host_csr** out) {
  if (!out) return fail("out is null");
  *out = nullptr;
  if (m && (!src || !dst)) return fail("src/dst is null");
  if (m > 0xffffffffULL) return fail("edge count exceeds the uint32 offset range");
  for (uint64_t i = 0; i < m; ++i)
    if (src[i] >= n || dst[i] >= n) return fail("vertex id out of range");
  try {
    return build_from_draws(
        n, m, flags, pick_threads(nthreads),
        [&](uint64_t lo, uint64_t hi, auto&& sink) {
          for (uint64_t i = lo; i < hi; ++i) sink(src[i], dst[i]);
        },
        out);
  } catch (const std::bad_alloc&) {
    return fail("out of host memory");
  }
}

int nwgr_gen_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed, double a, double b, double c,
                  int nthreads, nwgr_host_csr** out) {
  if (!out) return fail("out is null");
  *out = nullptr;
  if (scale < 1 || scale > 31) return fail("scale must be in [1, 31]");
  if (edge_factor < 1) return fail("edge_factor must be >= 1");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return fail("invalid quadrant probabilities");
  const uint64_t draws = static_cast<uint64_t>(edge_factor) << scale;
  if (draws > 0xffffffffULL) return fail("edge_factor << scale must be < 2^32");
  const RmatThresholds th = rmat_thresholds(a, b, c);
  try {
    return build_from_draws(
        1u << scale, draws, NWGR_CSR_DEDUP | NWGR_CSR_DROP_SELF_LOOPS, pick_threads(nthreads),
        [&](uint64_t lo, uint64_t hi, auto&& sink) {
          for (uint64_t e = lo; e < hi; ++e) {
            uint32_t u, v;
            rmat_draw(seed, e, scale, th, u, v);
            sink(u, v);
          }
        },
        out);
  } catch (const std::bad_alloc&) {
    return fail("out of host memory");
  }
}

int nwgr_gen_er(uint32_t n, uint64_t m_draws, uint64_t seed, int nthreads, nwgr_host_csr** out) {
  if (!out) return fail("out is null");
  *out = nullptr;
  if (n < 1) return fail("n must be >= 1");
  if (m_draws > 0xffffffffULL) return fail("m_draws must be < 2^32");
  try {
    return build_from_draws(
        n, m_draws, NWGR_CSR_DEDUP | NWGR_CSR_DROP_SELF_LOOPS, pick_threads(nthreads),
        [&](uint64_t lo, uint64_t hi, auto&& sink) {
          for (uint64_t e = lo; e < hi; ++e) {
            uint32_t u, v;
            er_draw(seed, e, n, u, v);
            sink(u, v);
          }
        },
        out);
  } catch (const std::bad_alloc&) {
    return fail("out of host memory");
  }
}

int nwgr_gen_rmat_n(uint32_t n, uint64_t draws, uint64_t seed, double a, double b, double c,
                    nwgr_host_csr** out) {
  if (!out) return fail("out is null");
  *out = nullptr;
  if (n < 2) return fail("n must be >= 2");
  if (draws > 0xffffffffULL) return fail("draws must be < 2^32");
  try {
    return rmat_n_single(n, draws, seed, rmat_thresholds(a, b, c), out);
  } catch (const std::bad_alloc&) {
    return fail("out of host memory");
  }
}

int nwgr_gen_rmat_many(uint32_t count, const uint32_t* ns, uint32_t ef, uint64_t seed, double a,
                       double b, double c, int nthreads, nwgr_host_csr** out) {
  if (!out) return fail("out is null");
  *out = nullptr;
  if (!ns && count) return fail("ns is null");
  uint64_t total_n = 0;
  for (uint32_t g = 0; g < count; ++g) {
    if (ns[g] < 2) return fail("every graph needs >= 2 vertices");
    total_n += ns[g];
  }
  if (total_n > 0xffffffffULL) return fail("total vertex count exceeds uint32");
  const RmatThresholds th = rmat_thresholds(a, b, c);
  std::vector<nwgr_host_csr*> parts(count, nullptr);
  std::atomic<int> bad{0};
  std::string first_err;
  try {
    par_dynamic(count, 1, pick_threads(nthreads), [&](uint64_t lo, uint64_t hi) {
      for (uint64_t g = lo; g < hi; ++g) {
        if (rmat_n_single(ns[g], static_cast<uint64_t>(ef) * ns[g], sm64_at(seed, g), th, &parts[g]))
          bad.store(1);
      }
    });
    if (bad.load()) {
      for (auto* p : parts) delete p;
      return fail("generating one of the graphs failed");
    }
    auto g = std::make_unique<nwgr_host_csr>();
    g->n = static_cast<uint32_t>(total_n);
    uint64_t m = 0;
    for (auto* p : parts) m += p->m;
    if (m > 0xffffffffULL) {
      for (auto* p : parts) delete p;
      return fail("edge count exceeds the uint32 offset range");
    }
    g->m = m;
    g->offsets.resize(total_n + 1);
    g->indices.resize(m);
    g->outdeg.resize(total_n);
    uint64_t vb = 0, eb = 0;
    for (auto* p : parts) {
      for (uint32_t i = 0; i < p->n; ++i) {
        g->offsets[vb + i] = static_cast<uint32_t>(eb + p->offsets[i]);
        g->outdeg[vb + i] = p->outdeg[i];
      }
      for (uint64_t k = 0; k < p->m; ++k) g->indices[eb + k] = static_cast<uint32_t>(vb + p->indices[k]);
      vb += p->n;
      eb += p->m;
      delete p;
    }
    g->offsets[total_n] = static_cast<uint32_t>(eb);
    *out = g.release();
    return 0;
  } catch (const std::bad_alloc&) {
    for (auto* p : parts) delete p;
    return fail("out of host memory");
  }
}

int nwgr_csr_transpose(uint32_t n, uint64_t m, const uint32_t* offsets, const uint32_t* indices,
                       uint32_t* out_offsets, uint32_t* out_indices, uint32_t* perm, int nthreads) {
  if (!offsets || !out_offsets || (m && (!indices || !out_indices || !perm))) return fail("null pointer");
  if (offsets[n] != m) return fail("offsets[n] != m");
  const int T = pick_threads(nthreads);
  try {
    std::vector<uint32_t> cnt(n, 0);
    for (uint64_t k = 0; k < m; ++k) {
      if (indices[k] >= n) return fail("vertex id out of range");
      ++cnt[indices[k]];
    }
    out_offsets[0] = 0;
    for (uint32_t u = 0; u < n; ++u) out_offsets[u + 1] = out_offsets[u] + cnt[u];
    std::vector<uint32_t> cur(out_offsets, out_offsets + n);
    // rows v are visited in ascending order, so each out-row comes out sorted
    (void)T;
    for (uint32_t v = 0; v < n; ++v) {
      for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) {
        uint32_t u = indices[k];
        uint32_t p = cur[u]++;
        out_indices[p] = v;
        perm[p] = k;
      }
    }
    return 0;
  } catch (const std::bad_alloc&) {
    return fail("out of host memory");
  }
}

uint32_t nwgr_host_csr_n(const nwgr_host_csr* g) { return g ? g->n : 0; }
uint64_t nwgr_host_csr_m(const nwgr_host_csr* g) { return g ? g->m : 0; }
const uint32_t* nwgr_host_csr_offsets(const nwgr_host_csr* g) { return g ? g->offsets.data() : nullptr; }
const uint32_t* nwgr_host_csr_indices(const nwgr_host_csr* g) { return g ? g->indices.data() : nullptr; }
const uint32_t* nwgr_host_csr_outdeg(const nwgr_host_csr* g) { return g ? g->outdeg.data() : nullptr; }
void nwgr_host_csr_free(nwgr_host_csr* g) { delete g; }

}  // extern "C"
