// Device sorts and scans over more items than one hipCUB call counts (int).
//
// The reference's CSR holds up to 2^32 - 1 entries (uint32 offsets,
// util/defaults.hpp:23-24); hipCUB's sorts and scans take an int count.
// Counts up to the pass size (sort_batch(), 2^30 by default) go to one call;
// larger ones are cut into passes of at most that many items:
//  * a stable sort by a small key is a stable sort of every pass followed by
//    a placement: item i of a pass goes to bucket_start[key] + (items of its
//    key in earlier passes) + (its rank among the pass's items of that key),
//    so the result is the stable sort of the whole sequence;
//  * an exclusive scan is a scan of every pass plus the total of the passes
//    before it.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>

#include "nwgr_internal.hpp"

namespace nwgr {
namespace {

std::atomic<uint64_t> g_batch{uint64_t{1} << 30};

constexpr int kT = 256;

inline uint32_t grid_for(uint64_t work) {
  return static_cast<uint32_t>(std::clamp<uint64_t>((work + kT - 1) / kT, 1, 65535u * 4));
}

#define GRID_STRIDE(i, count)                                                                       \
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < (count);     \
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)

__global__ void k_iota_from(uint32_t* v, uint64_t count, uint64_t first) {
  GRID_STRIDE(i, count) v[i] = static_cast<uint32_t>(first + i);
}

// run_start[i] = i when a new key starts at i of the sorted pass, else 0 (the
// inclusive max-scan of it is then the start of i's run)
__global__ void k_run_flags(const uint32_t* __restrict__ ks, uint64_t count, uint32_t* __restrict__ rs) {
  GRID_STRIDE(i, count) rs[i] = (i == 0 || ks[i] != ks[i - 1]) ? static_cast<uint32_t>(i) : 0u;
}

__global__ void k_place(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ vs,
                        const uint32_t* __restrict__ rs, uint64_t count, const uint32_t* __restrict__ bucket_start,
                        const uint32_t* __restrict__ done, uint32_t* __restrict__ out_key,
                        uint32_t* __restrict__ out_val) {
  GRID_STRIDE(i, count) {
    const uint32_t key = ks[i];
    const uint64_t pos = static_cast<uint64_t>(bucket_start[key]) + done[key] + (static_cast<uint32_t>(i) - rs[i]);
    out_val[pos] = vs[i];
    if (out_key) out_key[pos] = key;
  }
}

// each key's run in the pass ends once: its last item adds the run's length
__global__ void k_advance(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ rs, uint64_t count,
                          uint32_t* __restrict__ done) {
  GRID_STRIDE(i, count) {
    if (i + 1 == count || ks[i + 1] != ks[i]) done[ks[i]] += static_cast<uint32_t>(i) + 1u - rs[i];
  }
}

// out[i] += in_prev + out_prev, the exclusive total before the pass
__global__ void k_add_carry(uint32_t* __restrict__ out, uint64_t count, const uint32_t* __restrict__ prev_in,
                            const uint32_t* __restrict__ prev_out) {
  const uint32_t carry = *prev_in + *prev_out;
  GRID_STRIDE(i, count) out[i] += carry;
}

struct MaxOp {
  __host__ __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

}  // namespace

uint64_t sort_batch() { return g_batch.load(std::memory_order_relaxed); }

void set_sort_batch(uint64_t items) { g_batch.store(items, std::memory_order_relaxed); }

void stable_sort_place(const uint32_t* keys, uint64_t count, int end_bit, uint32_t nbuckets,
                       const uint32_t* bucket_start, uint32_t* out_key, uint32_t* out_val, hipStream_t s) {
  if (count == 0) return;
  const uint64_t B = sort_batch();
  DevBuf vin, ks, vs, rf, rs, done, temp;
  const uint64_t cap = std::min(count, B);
  vin.alloc(sizeof(uint32_t) * cap);
  if (count <= B) {
    // one pass: the stable sort is the result
    k_iota_from<<<grid_for(count), kT, 0, s>>>(vin.as<uint32_t>(), count, 0);
    NWGR_HIP(hipGetLastError());
    DevBuf kout;
    uint32_t* ko = out_key;
    if (!ko) {
      kout.alloc(sizeof(uint32_t) * count);
      ko = kout.as<uint32_t>();
    }
    size_t bytes = 0;
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, ko, vin.as<uint32_t>(), out_val,
                                                static_cast<int>(count), 0, end_bit, s));
    temp.alloc(bytes);
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, bytes, keys, ko, vin.as<uint32_t>(), out_val,
                                                static_cast<int>(count), 0, end_bit, s));
    NWGR_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
    return;
  }
  ks.alloc(sizeof(uint32_t) * cap);
  vs.alloc(sizeof(uint32_t) * cap);
  rf.alloc(sizeof(uint32_t) * cap);
  rs.alloc(sizeof(uint32_t) * cap);
  done.alloc(sizeof(uint32_t) * std::max<uint32_t>(nbuckets, 1));
  NWGR_HIP(hipMemsetAsync(done.p, 0, done.bytes, s));
  size_t sort_bytes = 0, scan_bytes = 0;
  NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, ks.as<uint32_t>(), vin.as<uint32_t>(),
                                              vs.as<uint32_t>(), static_cast<int>(cap), 0, end_bit, s));
  NWGR_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, scan_bytes, rf.as<uint32_t>(), rs.as<uint32_t>(), MaxOp{},
                                             static_cast<int>(cap), s));
  temp.alloc(std::max(sort_bytes, scan_bytes));
  for (uint64_t b0 = 0; b0 < count; b0 += cap) {
    const uint64_t cnt = std::min(cap, count - b0);
    k_iota_from<<<grid_for(cnt), kT, 0, s>>>(vin.as<uint32_t>(), cnt, b0);
    NWGR_HIP(hipGetLastError());
    size_t bytes = sort_bytes;
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, bytes, keys + b0, ks.as<uint32_t>(), vin.as<uint32_t>(),
                                                vs.as<uint32_t>(), static_cast<int>(cnt), 0, end_bit, s));
    k_run_flags<<<grid_for(cnt), kT, 0, s>>>(ks.as<uint32_t>(), cnt, rf.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    bytes = scan_bytes;
    NWGR_HIP(hipcub::DeviceScan::InclusiveScan(temp.p, bytes, rf.as<uint32_t>(), rs.as<uint32_t>(), MaxOp{},
                                               static_cast<int>(cnt), s));
    k_place<<<grid_for(cnt), kT, 0, s>>>(ks.as<uint32_t>(), vs.as<uint32_t>(), rs.as<uint32_t>(), cnt, bucket_start,
                                         done.as<uint32_t>(), out_key, out_val);
    NWGR_HIP(hipGetLastError());
    k_advance<<<grid_for(cnt), kT, 0, s>>>(ks.as<uint32_t>(), rs.as<uint32_t>(), cnt, done.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  NWGR_HIP(hipStreamSynchronize(s));
}

void exclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t count, hipStream_t s) {
  if (count == 0) return;
  const uint64_t cap = std::min(count, sort_batch());
  size_t bytes = 0;
  NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, static_cast<int>(cap), s));
  DevBuf temp;
  temp.alloc(bytes);
  for (uint64_t b0 = 0; b0 < count; b0 += cap) {
    const uint64_t cnt = std::min(cap, count - b0);
    size_t b = bytes;
    NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, b, in + b0, out + b0, static_cast<int>(cnt), s));
    if (b0 > 0) {
      k_add_carry<<<grid_for(cnt), kT, 0, s>>>(out + b0, cnt, in + b0 - 1, out + b0 - 1);
      NWGR_HIP(hipGetLastError());
    }
  }
  NWGR_HIP(hipStreamSynchronize(s));
}

}  // namespace nwgr

extern "C" {

int nwgr_pr_set_sort_batch(uint64_t items) {
  return nwgr::guarded([&] {
    NWGR_REQUIRE(items >= (uint64_t{1} << 16) && items <= (uint64_t{1} << 30),
                 "the pass size must be in [2^16, 2^30]");
    nwgr::set_sort_batch(items);
  });
}

}  // extern "C"
