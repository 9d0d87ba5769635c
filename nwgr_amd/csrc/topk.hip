// Top-k vertices of a rank vector on the device (the verification step of
// bench/pr.cpp:47-52, print_n_ranks over util/proxysort.hpp:36-65).
//
// One stable radix sort of (rank, vertex id) pairs by descending rank: equal
// ranks keep ascending vertex ids, a total order (the reference's proxysort
// uses an unstable std::sort, so its tie order is unspecified).

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>

#include "nwgr_internal.hpp"

namespace nwgr {
namespace {

__global__ void k_iota_u32(uint32_t* v, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free top-k (nwgr_top_k_device_f32_noalloc): two
// key and two id buffers of kNoAllocMaxN entries that the radix sort
// alternates between (its double-buffer form needs no copies of its own),
// then the sort's temporary storage (histograms, lookback state).  Every GPU
// the library runs on holds one copy (zero-initialised program memory,
// 68 MiB) until the process ends; the calls on one GPU share it.
constexpr uint32_t kNoAllocMaxN = 1u << 22;
constexpr size_t kNoAllocTemp = size_t{4} << 20;
constexpr size_t kNoAllocBytes = 4 * sizeof(uint32_t) * size_t{kNoAllocMaxN} + kNoAllocTemp;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each GPU, taken once per device with that
// device current (the host-side name is only a stand-in for it).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

unsigned char* scratch_on(int dev) {
  NWGR_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    NWGR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NWGR_REQUIRE(p != nullptr, "no device address for the top-k scratch");
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch[dev];
}

// The sort of top_k on the scratch of the current device: the ranks are
// copied into the first key buffer, ids 0 .. n-1 into the first id buffer,
// and the stable radix sort alternates each with its twin.  Allocates
// nothing and does not synchronise.
void top_k_noalloc(const float* d_ranks, uint32_t n, uint32_t k, uint32_t* d_ids, float* d_vals, hipStream_t s) {
  NWGR_REQUIRE(n <= kNoAllocMaxN, "n exceeds the 4194304 ranks the allocation-free top-k holds");
  NWGR_REQUIRE(k <= n, "k must be <= n");
  if (k == 0) return;
  int dev = 0;
  NWGR_HIP(hipGetDevice(&dev));
  unsigned char* base = scratch_on(dev);
  const size_t stride = sizeof(uint32_t) * size_t{kNoAllocMaxN};
  float* key0 = reinterpret_cast<float*>(base);
  float* key1 = reinterpret_cast<float*>(base + stride);
  uint32_t* id0 = reinterpret_cast<uint32_t*>(base + 2 * stride);
  uint32_t* id1 = reinterpret_cast<uint32_t*>(base + 3 * stride);
  void* temp = base + 4 * stride;
  hipcub::DoubleBuffer<float> keys(key0, key1);
  hipcub::DoubleBuffer<uint32_t> ids(id0, id1);
  size_t bytes = 0;
  NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, keys, ids, static_cast<int>(n), 0, 32, s));
  if (bytes > kNoAllocTemp)
    throw Error(NWGR_ENOTSUP, "the radix sort needs " + std::to_string(bytes) +
                                  " bytes of temporary storage, more than the allocation-free top-k holds");
  NWGR_HIP(hipMemcpyAsync(key0, d_ranks, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
  k_iota_u32<<<std::clamp<uint32_t>((n + 255) / 256, 1, 65535), 256, 0, s>>>(id0, n);
  NWGR_HIP(hipGetLastError());
  NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(temp, bytes, keys, ids, static_cast<int>(n), 0, 32, s));
  NWGR_HIP(hipMemcpyAsync(d_ids, ids.Current(), sizeof(uint32_t) * k, hipMemcpyDeviceToDevice, s));
  if (d_vals) NWGR_HIP(hipMemcpyAsync(d_vals, keys.Current(), sizeof(float) * k, hipMemcpyDeviceToDevice, s));
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

int nwgr_top_k_device_f32_noalloc(const float* d_ranks, uint32_t n, uint32_t k, uint32_t* d_ids, float* d_vals,
                                  void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(d_ranks && d_ids, "null argument");
    nwgr::top_k_noalloc(d_ranks, n, k, d_ids, d_vals, static_cast<hipStream_t>(stream));
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
