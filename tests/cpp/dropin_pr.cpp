// Calls the C++20 drop-in nw::graph::page_rank the way bench/pr.cpp:137-139
// calls the reference: float ranks, damping 0.85f, tolerance and max_iters
// from the command line.  Input file: "n m" then n+1 offsets, m indices and
// n degrees (whitespace separated).  Prints one rank per line, then with a
// fourth argument k the "id: rank" lines of the k best-ranked vertices.  A
// fifth argument picks the solve as pr.exe's --version does (bench/pr.cpp:
// 100-153): 13 (page_rank_v13) or 3 (page_rank_v3); default page_rank.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "nwgr/page_rank.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: dropin_pr <graph.txt> <tol> <max_iters>\n";
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
  std::vector<float> rankings(n);
  try {
    const int version = argc > 5 ? std::atoi(argv[5]) : 11;
    const float tol = std::strtof(argv[2], nullptr);
    const size_t max_iters = static_cast<size_t>(std::atol(argv[3]));
    if (version == 13)
      nw::graph::page_rank_v13(g, deg, rankings, 0.85f, tol, max_iters, 1);
    else if (version == 3)
      nw::graph::page_rank_v3(g, deg, rankings, 0.85f, tol, max_iters);
    else
      nw::graph::page_rank(g, deg, rankings, 0.85f, tol, max_iters, 1);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  for (float r : rankings) std::printf("%.9g\n", r);
  if (argc > 4 && std::atol(argv[4]) > 0) {  // the pr.exe --verify print (bench/pr.cpp:47-52) of the top `argv[4]`
    std::fflush(stdout);
    std::cout << "Verifying\n";
    nw::graph::print_n_ranks(rankings, static_cast<size_t>(std::atol(argv[4])));
  }
  return 0;
}
