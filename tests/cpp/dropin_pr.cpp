// Calls the C++20 drop-in nw::graph::page_rank the way bench/pr.cpp:137-139
// calls the reference: damping 0.85, tolerance and max_iters from the
// command line.  Input file: "n m" then n+1 offsets, m indices and n degrees
// (whitespace separated).  Prints one rank per line, then with a fourth
// argument k > 0 the "id: rank" lines of the k best-ranked vertices.
//   argv[5]  the solve, as pr.exe's --version (bench/pr.cpp:100-153): 13
//            (page_rank_v13) or 3 (page_rank_v3); default 11 = page_rank
//   argv[6]  the reference's Real: f32 (pr.cpp:89, the default) or f64
//            (test/page_rank_test.cpp:66)
//   argv[7]  trials: the call repeated on the same graph and vectors, as
//            pr.exe's -n (pr.cpp:92-97); each call's wall time goes to
//            stderr ("trial i ms"), and a trial whose ranks differ from the
//            first's fails the run
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "nwgr/page_rank.hpp"

template <typename Real>
int solve_and_print(const nw::graph::csr_view& g, const std::vector<uint32_t>& deg, int version, Real tol,
                    size_t max_iters, long topk, int trials) {
  std::vector<Real> rankings(g.n), first;
  for (int t = 0; t < trials; ++t) {
    const auto t0 = std::chrono::steady_clock::now();
    if (version == 13)
      nw::graph::page_rank_v13(g, deg, rankings, Real(0.85), tol, max_iters, 1);
    else if (version == 3)
      nw::graph::page_rank_v3(g, deg, rankings, Real(0.85), tol, max_iters);
    else
      nw::graph::page_rank(g, deg, rankings, Real(0.85), tol, max_iters, 1);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "trial %d %.3f ms\n", t, ms);
    if (t == 0) {
      first = rankings;
    } else if (std::memcmp(first.data(), rankings.data(), sizeof(Real) * g.n) != 0) {
      std::cerr << "trial " << t << " differs from trial 0\n";
      return 1;
    }
  }
  for (Real r : rankings) std::printf("%.17g\n", static_cast<double>(r));
  if (topk > 0) {  // the pr.exe --verify print (bench/pr.cpp:47-52)
    std::fflush(stdout);
    std::cout << "Verifying\n";
    nw::graph::print_n_ranks(rankings, static_cast<size_t>(topk));
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: dropin_pr <graph.txt> <tol> <max_iters> [k] [version] [f32|f64] [trials]\n";
    return 2;
  }
  std::ifstream in(argv[1]);
  uint64_t n = 0, m = 0;
  in >> n >> m;
  std::vector<uint32_t> offs(n + 1), idx(m), deg(n);
  for (auto& x : offs) in >> x;
  for (auto& x : idx) in >> x;
  for (auto& x : deg) in >> x;
  if (!in) {
    std::cerr << "bad input\n";
    return 2;
  }
  nw::graph::csr_view g{offs.data(), idx.data(), static_cast<uint32_t>(n), m};
  const long topk = argc > 4 ? std::atol(argv[4]) : 0;
  const int version = argc > 5 ? std::atoi(argv[5]) : 11;
  const bool f64 = argc > 6 && std::strcmp(argv[6], "f64") == 0;
  const int trials = argc > 7 ? std::max(1, std::atoi(argv[7])) : 1;
  const size_t max_iters = static_cast<size_t>(std::atol(argv[3]));
  try {
    if (f64) return solve_and_print<double>(g, deg, version, std::strtod(argv[2], nullptr), max_iters, topk, trials);
    return solve_and_print<float>(g, deg, version, std::strtof(argv[2], nullptr), max_iters, topk, trials);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
