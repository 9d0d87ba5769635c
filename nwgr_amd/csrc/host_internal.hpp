// Internal declarations shared by the host-only sources of libnwgr_graphgen.so.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct nwgr_host_csr {
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> offsets;
  std::vector<uint32_t> indices;
  std::vector<uint32_t> outdeg;
};

namespace nwgr_host {
// Records `msg` as this thread's last error (nwgr_gen_last_error); returns 1.
int fail(const std::string& msg);
// nthreads, or the hardware concurrency (at most 64) when nthreads <= 0.
int pick_threads(int nthreads);
}  // namespace nwgr_host
