// Top-k vertices of a rank vector on the device (the verification step of
// bench/pr.cpp:47-52, print_n_ranks over util/proxysort.hpp:36-65).
//
// One stable radix sort of (rank, vertex id) pairs by descending rank: equal
// ranks keep ascending vertex ids, a total order (the reference's proxysort
// uses an unstable std::sort, so its tie order is unspecified).

#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "nwgr_internal.hpp"

namespace nwgr {
namespace {

__global__ void k_iota_u32(uint32_t* v, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}

void top_k(const float* d_ranks, uint32_t n, uint32_t k, uint32_t* d_ids, float* d_vals, hipStream_t s) {
  NWGR_REQUIRE(k <= n, "k must be <= n");
  if (k == 0) return;
  DevBuf ids_in, ids_out, keys_out;
  ids_in.alloc(sizeof(uint32_t) * n);
  ids_out.alloc(sizeof(uint32_t) * n);
  keys_out.alloc(sizeof(float) * n);
  k_iota_u32<<<std::clamp<uint32_t>((n + 255) / 256, 1, 65535), 256, 0, s>>>(ids_in.as<uint32_t>(), n);
  NWGR_HIP(hipGetLastError());
  size_t bytes = 0;
  NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, d_ranks, keys_out.as<float>(),
                                                       ids_in.as<uint32_t>(), ids_out.as<uint32_t>(),
                                                       static_cast<int>(n), 0, 32, s));
  DevBuf tmp;
  tmp.alloc(bytes);
  NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(tmp.p, bytes, d_ranks, keys_out.as<float>(),
                                                       ids_in.as<uint32_t>(), ids_out.as<uint32_t>(),
                                                       static_cast<int>(n), 0, 32, s));
  NWGR_HIP(hipMemcpyAsync(d_ids, ids_out.p, sizeof(uint32_t) * k, hipMemcpyDeviceToDevice, s));
  if (d_vals) NWGR_HIP(hipMemcpyAsync(d_vals, keys_out.p, sizeof(float) * k, hipMemcpyDeviceToDevice, s));
  NWGR_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
}

}  // namespace
}  // namespace nwgr

extern "C" {

int nwgr_top_k_device_f32(const float* d_ranks, uint32_t n, uint32_t k, uint32_t* d_ids, float* d_vals,
                          void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(d_ranks && d_ids, "null argument");
    NWGR_REQUIRE(n < 0x80000000u, "n must be < 2^31");
    nwgr::top_k(d_ranks, n, k, d_ids, d_vals, static_cast<hipStream_t>(stream));
  });
}

int nwgr_top_k_f32(const float* ranks, uint32_t n, uint32_t k, uint32_t* ids, float* vals, int device) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(ranks && ids, "null argument");
    NWGR_REQUIRE(n < 0x80000000u, "n must be < 2^31");
    NWGR_REQUIRE(k <= n, "k must be <= n");
    nwgr::DeviceGuard dg(device);
    nwgr::DevBuf r, i, v;
    r.alloc(sizeof(float) * std::max<uint32_t>(n, 1));
    i.alloc(sizeof(uint32_t) * std::max<uint32_t>(k, 1));
    v.alloc(sizeof(float) * std::max<uint32_t>(k, 1));
    NWGR_HIP(hipMemcpy(r.p, ranks, sizeof(float) * n, hipMemcpyHostToDevice));
    nwgr::top_k(r.as<float>(), n, k, i.as<uint32_t>(), v.as<float>(), nullptr);
    NWGR_HIP(hipMemcpy(ids, i.p, sizeof(uint32_t) * k, hipMemcpyDeviceToHost));
    if (vals) NWGR_HIP(hipMemcpy(vals, v.p, sizeof(float) * k, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
