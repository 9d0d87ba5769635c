// In-CSR construction from an edge list on the device.
//
// Replaces the reference's edge_list -> adjacency<1> path on the device:
// index_adjacency<1> (adjacency.hpp:90-97) -> fill_directed<1>
// (build.hpp:209-265), with the optional uniq (build.hpp:425-435) and
// remove_self_loops (build.hpp:437-443), and build_degrees
// (bench/common.hpp:109-120) for the out-degrees.  Row v lists the sources u
// of the edges u -> v in ascending order (one of the orders the reference's
// unstable sort may produce, build.hpp:262).
//
// One 64-bit key (v << 32 | u) per edge, one radix sort, a flagged select
// for self-loops and duplicates, per-row counts and an exclusive scan.

#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "nwgr_internal.hpp"

namespace nwgr {

namespace {

constexpr int kCT = 256;

inline uint32_t cgrid(uint64_t work) {
  return static_cast<uint32_t>(std::clamp<uint64_t>((work + kCT - 1) / kCT, 1, 65535u * 4));
}

#define CSR_STRIDE(i, count)                                                                       \
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < (count);    \
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)

__global__ void k_pack(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t m, uint32_t n,
                       uint64_t* __restrict__ keys, uint32_t* __restrict__ bad) {
  CSR_STRIDE(k, m) {
    const uint32_t u = src[k], v = dst[k];
    if (u >= n || v >= n) atomicOr(bad, 1u);
    keys[k] = (static_cast<uint64_t>(v < n ? v : 0) << 32) | (u < n ? u : 0);
  }
}

__global__ void k_keep(const uint64_t* __restrict__ keys, uint64_t m, uint32_t flags, uint8_t* __restrict__ keep) {
  CSR_STRIDE(k, m) {
    const uint64_t key = keys[k];
    bool ok = true;
    if ((flags & NWGR_BUILD_DROP_SELF_LOOPS) && (key >> 32) == (key & 0xffffffffULL)) ok = false;
    if ((flags & NWGR_BUILD_DEDUP) && k > 0 && keys[k - 1] == key) ok = false;
    keep[k] = ok ? 1 : 0;
  }
}

__global__ void k_split(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ count, uint32_t* __restrict__ idx,
                        uint32_t* __restrict__ rowcnt, uint32_t* __restrict__ outdeg) {
  const uint64_t mm = *count;
  CSR_STRIDE(k, mm) {
    const uint64_t key = keys[k];
    const uint32_t u = static_cast<uint32_t>(key), v = static_cast<uint32_t>(key >> 32);
    idx[k] = u;
    atomicAdd(&rowcnt[v], 1u);
    if (outdeg) atomicAdd(&outdeg[u], 1u);
  }
}

uint32_t nbits(uint64_t v) {
  uint32_t b = 1;
  while (b < 64 && (uint64_t{1} << b) <= v) ++b;
  return b;
}

}  // namespace

// Builds into caller-owned device arrays; returns the number of entries kept.
uint64_t build_csr_device(const uint32_t* d_src, const uint32_t* d_dst, uint64_t m, uint32_t n, uint32_t flags,
                          uint32_t* d_offsets, uint32_t* d_indices, uint32_t* d_outdeg, hipStream_t s) {
  NWGR_REQUIRE(n >= 1, "n must be >= 1");
  NWGR_REQUIRE(m < 0x80000000ULL, "m must be < 2^31");
  NWGR_REQUIRE(d_offsets != nullptr && (m == 0 || (d_src && d_dst && d_indices)), "null argument");
  NWGR_REQUIRE((flags & ~3u) == 0, "unknown flags");
  DevBuf keys, keys2, keep, cnt, bad, rowcnt;
  keys.alloc(sizeof(uint64_t) * std::max<uint64_t>(m, 1));
  keys2.alloc(sizeof(uint64_t) * std::max<uint64_t>(m, 1));
  keep.alloc(std::max<uint64_t>(m, 1));
  cnt.alloc(sizeof(uint32_t));
  bad.alloc(sizeof(uint32_t));
  rowcnt.alloc(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  NWGR_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
  NWGR_HIP(hipMemsetAsync(cnt.p, 0, sizeof(uint32_t), s));
  NWGR_HIP(hipMemsetAsync(rowcnt.p, 0, rowcnt.bytes, s));
  if (d_outdeg) NWGR_HIP(hipMemsetAsync(d_outdeg, 0, sizeof(uint32_t) * n, s));
  if (m) {
    k_pack<<<cgrid(m), kCT, 0, s>>>(d_src, d_dst, m, n, keys.as<uint64_t>(), bad.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    const int end_bit = static_cast<int>(32 + nbits(n - 1));
    size_t tb = 0;
    NWGR_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                               static_cast<int>(m), 0, end_bit, s));
    DevBuf tmp;
    tmp.alloc(tb);
    NWGR_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                               static_cast<int>(m), 0, end_bit, s));
    k_keep<<<cgrid(m), kCT, 0, s>>>(keys2.as<uint64_t>(), m, flags, keep.as<uint8_t>());
    NWGR_HIP(hipGetLastError());
    tb = 0;
    NWGR_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, keys2.as<uint64_t>(), keep.as<uint8_t>(), keys.as<uint64_t>(),
                                           cnt.as<uint32_t>(), static_cast<int>(m), s));
    tmp.alloc(tb);
    NWGR_HIP(hipcub::DeviceSelect::Flagged(tmp.p, tb, keys2.as<uint64_t>(), keep.as<uint8_t>(), keys.as<uint64_t>(),
                                           cnt.as<uint32_t>(), static_cast<int>(m), s));
    k_split<<<cgrid(m), kCT, 0, s>>>(keys.as<uint64_t>(), cnt.as<uint32_t>(), d_indices, rowcnt.as<uint32_t>(),
                                     d_outdeg);
    NWGR_HIP(hipGetLastError());
  }
  size_t tb = 0;
  NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, rowcnt.as<uint32_t>(), d_offsets, static_cast<int>(n) + 1, s));
  DevBuf tmp;
  tmp.alloc(tb);
  NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, rowcnt.as<uint32_t>(), d_offsets, static_cast<int>(n) + 1, s));
  uint32_t h[2] = {0, 0};
  NWGR_HIP(hipMemcpyAsync(&h[0], cnt.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipMemcpyAsync(&h[1], bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  NWGR_REQUIRE(h[1] == 0, "a vertex id is >= n");
  return h[0];
}

}  // namespace nwgr

struct nwgr_pr_graph : nwgr::Graph {};

extern "C" {

int nwgr_csr_build_device(const uint32_t* d_src, const uint32_t* d_dst, uint64_t m, uint32_t n, uint32_t flags,
                          uint32_t* d_offsets, uint32_t* d_indices, uint32_t* d_out_degree, uint64_t* m_out,
                          void* stream) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(m_out != nullptr, "m_out is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    *m_out = nwgr::build_csr_device(d_src, d_dst, m, n, flags, d_offsets, d_indices, d_out_degree, s);
  });
}

int nwgr_pr_graph_create_from_edges_device(const uint32_t* d_src, const uint32_t* d_dst, uint64_t m, uint32_t n,
                                           uint32_t flags, int device, void* stream, nwgr_pr_graph** out) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    nwgr::DeviceGuard dg(device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    nwgr::DevBuf offs, idx, deg;
    offs.alloc(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
    idx.alloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
    deg.alloc(sizeof(uint32_t) * std::max<uint32_t>(n, 1));
    const uint64_t mm = nwgr::build_csr_device(d_src, d_dst, m, n, flags, offs.as<uint32_t>(), idx.as<uint32_t>(),
                                               deg.as<uint32_t>(), s);
    nwgr::Graph* g = nwgr::create_graph(offs.as<uint32_t>(), idx.as<uint32_t>(), n, mm, deg.as<uint32_t>(), device, s,
                                        true);
    *out = static_cast<nwgr_pr_graph*>(g);
  });
}

}  // extern "C"
