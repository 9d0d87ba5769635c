/*
 * nwgr/page_rank.hpp -- C++20 drop-in for nw::graph::page_rank on the MI355X.
 *
 * Include this header instead of <nwgraph/algorithms/page_rank.hpp> and link
 * libnwgr_pagerank.so.  The template keeps the reference signature
 * (include/nwgraph/algorithms/page_rank.hpp:89-91):
 *
 *   template <adjacency_list_graph Graph, typename Real>
 *   void page_rank(const Graph& graph,
 *                  const std::vector<typename Graph::vertex_id_type>& degrees,
 *                  std::vector<Real>& page_rank, Real damping_factor,
 *                  Real threshold, size_t max_iters, size_t num_threads);
 *
 * Graph must expose the CSR the way NWGraph's indexed_struct_of_arrays does
 * (containers/compressed.hpp:67-69): `indices_` (offsets, N+1 entries) and
 * `std::get<0>(to_be_indexed_)` (neighbours), both contiguous uint32, and
 * size().  These are the members the reference's own raw-CSR code reads
 * (experimental/algorithms/page_rank.hpp:118-120, apb/spmv.cpp:48-50).  A
 * plain `nw::graph::csr_view` overload serves callers without NWGraph.
 *
 * Semantics: Jacobi sweeps on the GPU (the reference's v11 sweeps in place,
 * Gauss-Seidel; both reach the same fixed point, the sweep count differs),
 * in the caller's Real: FP32 for float, FP64 for double (long double and
 * other wider types are computed in FP64); dangling mass leaks as in the
 * reference; the result is written into the caller's vector, which must
 * already hold graph.size() entries (page_rank.hpp:102).  An empty graph
 * returns at once, as the reference does for N = 0.  num_threads is ignored
 * (unused in the reference too, page_rank.hpp:91).  Failures (no GPU,
 * malformed CSR) throw std::runtime_error; not converging within max_iters
 * returns normally.
 *
 * The device graph of a call is kept for the next call on the same arrays
 * (nwgr_pr_graph_cached: same addresses, sizes and contents), as the
 * reference's graph outlives its page_rank calls; pr.exe's trials
 * (bench/pr.cpp:94-95) then pay for the solve and the copy of the ranks only.
 * After editing a graph's arrays in place, call page_rank_cache_clear().
 */
#ifndef NWGR_PAGE_RANK_HPP
#define NWGR_PAGE_RANK_HPP

#include <cstddef>
#include <cstdint>
#include <algorithm>
#include <iostream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../nwgr_pagerank.h"

namespace nw::graph {

/* CSR given as raw arrays: offsets[n+1], indices[m] (row i = in-neighbours). */
struct csr_view {
  const uint32_t* offsets;
  const uint32_t* indices;
  uint32_t n;
  uint64_t m;
  using vertex_id_type = uint32_t;
  std::size_t size() const { return n; }
};

template <class G>
concept nwgr_csr_graph = requires(const G& g) {
  { g.indices_.data() } -> std::convertible_to<const uint32_t*>;
  { std::get<0>(g.to_be_indexed_).data() } -> std::convertible_to<const uint32_t*>;
  { g.size() } -> std::convertible_to<std::size_t>;
};

namespace nwgr_detail {

inline void check(int rc) {
  if (rc != NWGR_OK) throw std::runtime_error(std::string("nwgr: ") + nwgr_last_error());
}

/* Device used by the drop-in (default 0). */
inline int& device() {
  static int dev = 0;
  return dev;
}

/* Which solve `run` performs: the Jacobi sweeps of page_rank, or the
 * residual push of page_rank_v13 / page_rank_v3. */
enum class solver { sweeps, residual };

template <typename Real>
void run(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m, const uint32_t* degrees,
         std::size_t degrees_size, std::vector<Real>& page_rank, Real damping_factor, Real threshold,
         std::size_t max_iters, solver kind = solver::sweeps) {
  static_assert(std::is_floating_point_v<Real>, "Real must be a floating-point type");
  if (page_rank.size() < n) throw std::runtime_error("nwgr: page_rank must hold graph.size() entries");
  if (degrees_size < n) throw std::runtime_error("nwgr: degrees must hold graph.size() entries");
  if (n == 0) return;  // page_rank.hpp:92-147 over an empty graph writes nothing
  nwgr_pr_params p;
  nwgr_pr_params_default(&p);
  p.damping = static_cast<double>(damping_factor);
  p.tol = static_cast<double>(threshold);
  p.max_iters = max_iters > 0xffffffffu ? 0xffffffffu : static_cast<uint32_t>(max_iters);
  p.dangling = NWGR_DANGLING_LEAK;
  nwgr_pr_result res;
  // the device graph is the library cache's resident one (nwgr_pr_cached_run)
  auto solve = [&](uint32_t which, void* out) {
    check(nwgr_pr_cached_run(offsets, indices, n, m, degrees, device(), which, &p, out, &res, nullptr));
  };
  if (kind == solver::residual) {
    // the residual push is FP32 on the device
    if constexpr (std::is_same_v<Real, float>) {
      solve(NWGR_SOLVE_DELTA_F32, page_rank.data());
    } else {
      std::vector<float> tmp(n);
      solve(NWGR_SOLVE_DELTA_F32, tmp.data());
      std::copy(tmp.begin(), tmp.end(), page_rank.begin());
    }
  } else if constexpr (std::is_same_v<Real, float>) {
    solve(NWGR_SOLVE_F32, page_rank.data());
  } else if constexpr (std::is_same_v<Real, double>) {
    solve(NWGR_SOLVE_F64, page_rank.data());
  } else {
    std::vector<double> tmp(n);
    solve(NWGR_SOLVE_F64, tmp.data());
    std::copy(tmp.begin(), tmp.end(), page_rank.begin());
  }
}

template <class Graph>
const uint32_t* as_u32(const std::vector<typename Graph::vertex_id_type>& v) {
  static_assert(sizeof(typename Graph::vertex_id_type) == sizeof(uint32_t) &&
                    std::is_integral_v<typename Graph::vertex_id_type>,
                "the device path takes uint32 vertex ids (util/defaults.hpp:23-24)");
  return reinterpret_cast<const uint32_t*>(v.data());
}

}  // namespace nwgr_detail

/* Select the GPU used by page_rank (default 0). */
inline void page_rank_set_device(int device) { nwgr_detail::device() = device; }

/* Release the device graphs kept between calls (needed after a graph's arrays
 * are edited in place). */
inline void page_rank_cache_clear() { nwgr_pr_cache_clear(); }

template <nwgr_csr_graph Graph, typename Real>
void page_rank(const Graph& graph, const std::vector<typename Graph::vertex_id_type>& degrees,
               std::vector<Real>& page_rank, Real damping_factor, Real threshold, std::size_t max_iters,
               std::size_t /*num_threads*/) {
  const auto& offs = graph.indices_;
  const auto& nbrs = std::get<0>(graph.to_be_indexed_);
  const uint32_t n = static_cast<uint32_t>(graph.size());
  nwgr_detail::run<Real>(reinterpret_cast<const uint32_t*>(offs.data()),
                         reinterpret_cast<const uint32_t*>(nbrs.data()), n,
                         static_cast<uint64_t>(offs.empty() ? 0 : offs[n]), nwgr_detail::as_u32<Graph>(degrees),
                         degrees.size(), page_rank, damping_factor, threshold, max_iters);
}

template <typename Real>
void page_rank(const csr_view& graph, const std::vector<uint32_t>& degrees, std::vector<Real>& page_rank,
               Real damping_factor, Real threshold, std::size_t max_iters, std::size_t /*num_threads*/) {
  nwgr_detail::run<Real>(graph.offsets, graph.indices, graph.n, graph.m, degrees.data(), degrees.size(), page_rank,
                         damping_factor, threshold, max_iters);
}

/* Residual ("delta") PageRank, the signature of page_rank_v13
 * (experimental/algorithms/page_rank.hpp:632-634, pr.exe version 13):
 * ranks start at 0, residuals at 1 - d; residuals above `threshold` are
 * pushed into the rank and sent on as residual * d / degree; unnormalised
 * result (nwgr_pr_delta_run_f32, FP32 for every Real). */
template <nwgr_csr_graph Graph, typename Real>
void page_rank_v13(const Graph& graph, const std::vector<typename Graph::vertex_id_type>& degrees,
                   std::vector<Real>& page_rank, Real damping_factor, Real threshold, std::size_t max_iters,
                   std::size_t /*num_threads*/) {
  const auto& offs = graph.indices_;
  const auto& nbrs = std::get<0>(graph.to_be_indexed_);
  const uint32_t n = static_cast<uint32_t>(graph.size());
  nwgr_detail::run<Real>(reinterpret_cast<const uint32_t*>(offs.data()),
                         reinterpret_cast<const uint32_t*>(nbrs.data()), n,
                         static_cast<uint64_t>(offs.empty() ? 0 : offs[n]), nwgr_detail::as_u32<Graph>(degrees),
                         degrees.size(), page_rank, damping_factor, threshold, max_iters,
                         nwgr_detail::solver::residual);
}

template <typename Real>
void page_rank_v13(const csr_view& graph, const std::vector<uint32_t>& degrees, std::vector<Real>& page_rank,
                   Real damping_factor, Real threshold, std::size_t max_iters, std::size_t /*num_threads*/) {
  nwgr_detail::run<Real>(graph.offsets, graph.indices, graph.n, graph.m, degrees.data(), degrees.size(), page_rank,
                         damping_factor, threshold, max_iters, nwgr_detail::solver::residual);
}

/* page_rank_v3 (experimental/algorithms/page_rank.hpp:566-568), the serial
 * residual push: it keeps a residual at or below the threshold until a delta
 * replaces it, and such a residual never pushes, so its ranks are those of
 * page_rank_v13. */
template <typename Graph, typename Real>
void page_rank_v3(const Graph& graph, const std::vector<typename Graph::vertex_id_type>& degrees,
                  std::vector<Real>& page_rank, Real damping_factor, Real threshold, std::size_t max_iters) {
  page_rank_v13(graph, degrees, page_rank, damping_factor, threshold, max_iters, 1);
}

/* The verification print of bench/pr.cpp:47-52: the `n` best-ranked vertices
 * as "id: rank" lines, ranks by non-increasing value and ties by ascending id
 * (ranked on the device, nwgr_top_k_f32). */
template <typename Real>
void print_n_ranks(const std::vector<Real>& rankings, std::size_t n, std::ostream& os = std::cout) {
  std::vector<float> r(rankings.begin(), rankings.end());
  const uint32_t k = static_cast<uint32_t>(std::min<std::size_t>(n, r.size()));
  std::vector<uint32_t> ids(k);
  std::vector<float> vals(k);
  nwgr_detail::check(nwgr_top_k_f32(r.data(), static_cast<uint32_t>(r.size()), k, ids.data(), vals.data(),
                                    nwgr_detail::device()));
  for (uint32_t i = 0; i < k; ++i) os << std::to_string(ids[i]) + ": " << std::to_string(vals[i]) << std::endl;
}

}  // namespace nw::graph

#endif  // NWGR_PAGE_RANK_HPP
