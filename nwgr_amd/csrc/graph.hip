// Device-resident graph construction for the PageRank sweeps.
//
// Input is the reference's in-CSR (adjacency<1>, build.hpp:209-265) plus the
// caller's degree array (bench/common.hpp:109-120).  On the device the graph
// is relabelled so that rows appear in non-increasing in-degree order (ties:
// ascending original id).  Degree bins then become contiguous row ranges and
// hub rows come first in the work queue.  Row entries are renamed through
// the same permutation and sorted ascending.  `perm` (new -> original) and
// `iperm` (original -> new) map results back to the caller's order.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <memory>
#include <cstring>

#include "nwgr_internal.hpp"

namespace nwgr {

namespace {

constexpr int kT = 256;

inline uint32_t grid_for(uint64_t work, uint32_t cap = 65535u * 4) {
  uint64_t g = (work + kT - 1) / kT;
  if (g < 1) g = 1;
  return static_cast<uint32_t>(std::min<uint64_t>(g, cap));
}

__global__ void k_validate(const uint32_t* __restrict__ offs, const uint32_t* __restrict__ idx,
                           uint32_t n, uint64_t m, uint32_t* bad) {
  const uint64_t tid = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint32_t flags = 0;
  if (tid == 0 && (offs[0] != 0 || offs[n] != m)) flags |= 1u;
  for (uint64_t i = tid; i < n; i += stride)
    if (offs[i] > offs[i + 1]) flags |= 2u;
  for (uint64_t k = tid; k < m; k += stride)
    if (idx[k] >= n) flags |= 4u;
  if (flags) atomicOr(bad, flags);
}

__global__ void k_out_degree(const uint32_t* __restrict__ idx, uint64_t m, uint32_t* deg) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride)
    atomicAdd(&deg[idx[k]], 1u);
}

__global__ void k_in_degree(const uint32_t* __restrict__ offs, uint32_t n, uint32_t* key, uint32_t* val) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n; v += stride) {
    key[v] = offs[v + 1] - offs[v];
    val[v] = static_cast<uint32_t>(v);
  }
}

__global__ void k_invert(const uint32_t* __restrict__ perm, uint32_t n, uint32_t* iperm) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    iperm[perm[i]] = static_cast<uint32_t>(i);
}

// Edge k of original row v becomes key (iperm[v] << 32 | iperm[idx[k]]),
// value k.  The row of edge k is found by binary search in offs.
__global__ void k_edge_keys(const uint32_t* __restrict__ offs, const uint32_t* __restrict__ idx,
                            const uint32_t* __restrict__ iperm, uint32_t n, uint64_t m,
                            uint64_t* keys, uint32_t* vals) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride) {
    // largest v with offs[v] <= k
    uint32_t lo = 0, hi = n;  // invariant offs[lo] <= k < offs[hi]
    while (hi - lo > 1) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (offs[mid] <= k) lo = mid; else hi = mid;
    }
    keys[k] = (static_cast<uint64_t>(iperm[lo]) << 32) | iperm[idx[k]];
    vals[k] = static_cast<uint32_t>(k);
  }
}

__global__ void k_split_keys(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* idx) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride)
    idx[k] = static_cast<uint32_t>(keys[k]);
}

__global__ void k_degf(const uint32_t* __restrict__ deg, const uint32_t* __restrict__ perm, uint32_t n,
                       float* degf) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    degf[i] = static_cast<float>(deg[perm[i]]);
}

__global__ void k_set_last(uint32_t* offs, uint32_t n, uint32_t m) { offs[n] = m; }

uint32_t bits_for(uint32_t n) {
  uint32_t b = 1;
  while (b < 32 && (uint64_t{1} << b) < n) ++b;
  return b;
}

// Segment plan from the non-increasing in-degree sequence `deg` (new order).
void plan_segments(Graph& g, const std::vector<uint32_t>& deg, const std::vector<uint32_t>& offs_host,
                   std::vector<HubUnit>& hubs) {
  const Tuning& t = g.tune;
  const uint32_t n = g.n;
  auto first_below_or_eq = [&](uint32_t x) -> uint32_t {
    // first row with deg <= x (deg is non-increasing)
    return static_cast<uint32_t>(
        std::partition_point(deg.begin(), deg.end(), [&](uint32_t d) { return d > x; }) - deg.begin());
  };
  SegTable st{};
  uint32_t unit = 0;
  // hub rows: whole workgroups over chunks of hub_chunk edges
  const uint32_t h_end = first_below_or_eq(t.hub_threshold);
  hubs.clear();
  for (uint32_t r = 0; r < h_end; ++r) {
    const uint32_t d = deg[r];
    const uint32_t nch = (d + t.hub_chunk - 1) / t.hub_chunk;
    const uint32_t first = static_cast<uint32_t>(hubs.size());
    for (uint32_t c = 0; c < nch; ++c) {
      HubUnit u;
      u.row = r;
      u.e_lo = offs_host[r] + c * t.hub_chunk;
      u.e_hi = std::min(offs_host[r] + (c + 1) * t.hub_chunk, offs_host[r + 1]);
      u.first = first;
      u.nchunks = nch;
      u.hub = r;
      hubs.push_back(u);
    }
  }
  if (!hubs.empty()) {
    Seg& s = st.s[st.nseg++];
    s.row_lo = 0;
    s.row_hi = h_end;
    s.G = 0;
    s.rows_per_unit = 1;
    s.unit_lo = unit;
    unit += static_cast<uint32_t>(hubs.size());
    s.unit_hi = unit;
  }
  // lane-group segments: G lanes per row for degrees in (G, 2G]; G = 64
  // takes everything in (64, hub_threshold]; G = 1 takes degrees 1 and 2.
  uint32_t row = h_end;
  struct Band { uint32_t G, dlo; };  // rows with deg > dlo (and <= previous band's dlo)
  const Band bands[] = {{64, 64}, {32, 32}, {16, 16}, {8, 8}, {4, 4}, {2, 2}, {1, 0}};
  for (const Band& b : bands) {
    const uint32_t end = first_below_or_eq(b.dlo);
    if (end > row) {
      Seg& s = st.s[st.nseg++];
      s.row_lo = row;
      s.row_hi = end;
      s.G = b.G;
      const uint64_t edges = static_cast<uint64_t>(offs_host[end]) - offs_host[row];
      const double avg = static_cast<double>(edges) / (end - row);
      const uint32_t per_pass = kBlock / b.G;
      double passes = t.unit_edges / std::max(1.0, per_pass * avg);
      uint32_t R = static_cast<uint32_t>(std::clamp(passes, 1.0, 256.0));
      s.rows_per_unit = per_pass * R;
      s.unit_lo = unit;
      unit += (end - row + s.rows_per_unit - 1) / s.rows_per_unit;
      s.unit_hi = unit;
      row = end;
    }
  }
  // rows without in-edges: teleport/dangling terms only
  if (n > row) {
    Seg& s = st.s[st.nseg++];
    s.row_lo = row;
    s.row_hi = n;
    s.G = 1;
    s.rows_per_unit = std::max<uint32_t>(kBlock, t.zero_rows_per_unit / kBlock * kBlock);
    s.unit_lo = unit;
    unit += (n - row + s.rows_per_unit - 1) / s.rows_per_unit;
    s.unit_hi = unit;
  }
  st.total_units = unit;
  g.segs = st;
  g.n_hub_units = static_cast<uint32_t>(hubs.size());
  g.n_hubs = h_end;
}

}  // namespace

Graph::~Graph() {
  drop_graphs();
  if (hflag) (void)hipHostFree(hflag);
  if (own_stream) (void)hipStreamDestroy(own_stream);
}

Graph* create_graph(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                    const uint32_t* degrees, int device, hipStream_t stream, bool on_device) {
  NWGR_REQUIRE(n >= 1, "n must be >= 1");
  NWGR_REQUIRE(m <= 0xffffffffULL, "m must fit the uint32 offsets (m < 2^32)");
  // the device radix sorts count items in int
  NWGR_REQUIRE(n < 0x80000000u && m < 0x80000000ULL, "n and m must be < 2^31");
  NWGR_REQUIRE(offsets != nullptr, "offsets is null");
  NWGR_REQUIRE(m == 0 || indices != nullptr, "indices is null");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    throw Error(NWGR_EDEVICE, "no HIP device available");
  NWGR_REQUIRE(device >= 0 && device < ndev, "device ordinal out of range");
  if (!on_device) {
    NWGR_REQUIRE(offsets[0] == 0, "offsets[0] must be 0");
    NWGR_REQUIRE(offsets[n] == m, "offsets[n] must equal m");
  }
  DeviceGuard dg(device);
  auto t0 = std::chrono::steady_clock::now();
  auto g = std::make_unique<Graph>();
  g->device = device;
  g->n = n;
  g->m = m;
  hipDeviceProp_t prop;
  NWGR_HIP(hipGetDeviceProperties(&prop, device));
  g->num_cu = static_cast<uint32_t>(prop.multiProcessorCount);
  NWGR_HIP(hipStreamCreateWithFlags(&g->own_stream, hipStreamNonBlocking));
  if (!stream) stream = g->own_stream;
  const size_t nb = sizeof(uint32_t) * n;
  const size_t mb = sizeof(uint32_t) * m;

  // input CSR on device
  DevBuf in_offs, in_idx, in_deg;
  const uint32_t* d_offs = offsets;
  const uint32_t* d_idx = m ? indices : nullptr;
  const uint32_t* d_deg = degrees;
  if (!on_device) {
    in_offs.alloc(nb + sizeof(uint32_t));
    NWGR_HIP(hipMemcpyAsync(in_offs.p, offsets, nb + sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    d_offs = in_offs.as<uint32_t>();
    if (m) {
      in_idx.alloc(mb);
      NWGR_HIP(hipMemcpyAsync(in_idx.p, indices, mb, hipMemcpyHostToDevice, stream));
      d_idx = in_idx.as<uint32_t>();
    }
    if (degrees) {
      in_deg.alloc(nb);
      NWGR_HIP(hipMemcpyAsync(in_deg.p, degrees, nb, hipMemcpyHostToDevice, stream));
      d_deg = in_deg.as<uint32_t>();
    }
  }
  // validation
  {
    DevBuf bad;
    bad.alloc(sizeof(uint32_t));
    NWGR_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), stream));
    k_validate<<<grid_for(std::max<uint64_t>(n, m)), kT, 0, stream>>>(d_offs, d_idx ? d_idx : d_offs, n, m,
                                                                        bad.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    uint32_t flags = 0;
    NWGR_HIP(hipMemcpyAsync(&flags, bad.p, sizeof(flags), hipMemcpyDeviceToHost, stream));
    NWGR_HIP(hipStreamSynchronize(stream));
    NWGR_REQUIRE(!(flags & 1u), "offsets[0] must be 0 and offsets[n] must equal m");
    NWGR_REQUIRE(!(flags & 2u), "offsets must be non-decreasing");
    NWGR_REQUIRE(!(flags & 4u), "an index is >= n");
  }
  // degrees: caller's, or out-degree counted from the indices
  if (!d_deg) {
    in_deg.alloc(nb);
    NWGR_HIP(hipMemsetAsync(in_deg.p, 0, nb, stream));
    if (m) k_out_degree<<<grid_for(m), kT, 0, stream>>>(d_idx, m, in_deg.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    d_deg = in_deg.as<uint32_t>();
  }

  // --- relabel rows by non-increasing in-degree (stable => ties by id)
  DevBuf key_in, val_in, key_out;
  key_in.alloc(nb);
  val_in.alloc(nb);
  key_out.alloc(nb);
  g->perm.alloc(nb);
  g->iperm.alloc(nb);
  k_in_degree<<<grid_for(n), kT, 0, stream>>>(d_offs, n, key_in.as<uint32_t>(), val_in.as<uint32_t>());
  NWGR_HIP(hipGetLastError());
  {
    size_t tmp = 0;
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, key_in.as<uint32_t>(),
                                                         key_out.as<uint32_t>(), val_in.as<uint32_t>(),
                                                         g->perm.as<uint32_t>(), static_cast<int>(n), 0,
                                                         32, stream));
    DevBuf t;
    t.alloc(tmp);
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(t.p, tmp, key_in.as<uint32_t>(),
                                                         key_out.as<uint32_t>(), val_in.as<uint32_t>(),
                                                         g->perm.as<uint32_t>(), static_cast<int>(n), 0,
                                                         32, stream));
  }
  k_invert<<<grid_for(n), kT, 0, stream>>>(g->perm.as<uint32_t>(), n, g->iperm.as<uint32_t>());
  NWGR_HIP(hipGetLastError());
  // new offsets = exclusive scan of the sorted in-degrees
  g->offs.alloc(nb + sizeof(uint32_t));
  {
    size_t tmp = 0;
    NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, key_out.as<uint32_t>(), g->offs.as<uint32_t>(),
                                              static_cast<int>(n), stream));
    DevBuf t;
    t.alloc(tmp);
    NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, key_out.as<uint32_t>(), g->offs.as<uint32_t>(),
                                              static_cast<int>(n), stream));
    k_set_last<<<1, 1, 0, stream>>>(g->offs.as<uint32_t>(), n, static_cast<uint32_t>(m));
    NWGR_HIP(hipGetLastError());
  }
  // renamed, row-sorted edges
  g->idx.alloc(mb);
  g->epos.alloc(mb);
  if (m) {
    DevBuf ek_in, ek_out, ev_in;
    ek_in.alloc(sizeof(uint64_t) * m);
    ek_out.alloc(sizeof(uint64_t) * m);
    ev_in.alloc(mb);
    k_edge_keys<<<grid_for(m), kT, 0, stream>>>(d_offs, d_idx, g->iperm.as<uint32_t>(), n, m,
                                                ek_in.as<uint64_t>(), ev_in.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    const int end_bit = 32 + static_cast<int>(bits_for(n));
    size_t tmp = 0;
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ek_in.as<uint64_t>(), ek_out.as<uint64_t>(),
                                                ev_in.as<uint32_t>(), g->epos.as<uint32_t>(),
                                                static_cast<int>(m), 0, end_bit, stream));
    DevBuf t;
    t.alloc(tmp);
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, ek_in.as<uint64_t>(), ek_out.as<uint64_t>(),
                                                ev_in.as<uint32_t>(), g->epos.as<uint32_t>(),
                                                static_cast<int>(m), 0, end_bit, stream));
    k_split_keys<<<grid_for(m), kT, 0, stream>>>(ek_out.as<uint64_t>(), m, g->idx.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  g->degf.alloc(sizeof(float) * n);
  k_degf<<<grid_for(n), kT, 0, stream>>>(d_deg, g->perm.as<uint32_t>(), n, g->degf.as<float>());
  NWGR_HIP(hipGetLastError());

  // --- plan: host copies of the sorted degrees and new offsets
  std::vector<uint32_t> deg_h(n), offs_h(static_cast<size_t>(n) + 1);
  NWGR_HIP(hipMemcpyAsync(deg_h.data(), key_out.p, nb, hipMemcpyDeviceToHost, stream));
  NWGR_HIP(hipMemcpyAsync(offs_h.data(), g->offs.p, nb + sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  NWGR_HIP(hipStreamSynchronize(stream));
  std::vector<HubUnit> hubs;
  plan_segments(*g, deg_h, offs_h, hubs);
  g->hub_units.alloc(sizeof(HubUnit) * std::max<size_t>(hubs.size(), 1));
  if (!hubs.empty())
    NWGR_HIP(hipMemcpyAsync(g->hub_units.p, hubs.data(), sizeof(HubUnit) * hubs.size(),
                            hipMemcpyHostToDevice, stream));
  g->hub_part.alloc(sizeof(uint32_t) * std::max<size_t>(hubs.size(), 1));
  g->hub_ticket.alloc(sizeof(uint32_t) * std::max<uint32_t>(g->n_hubs, 1));
  NWGR_HIP(hipMemsetAsync(g->hub_ticket.p, 0, g->hub_ticket.bytes, stream));

  // --- iteration buffers
  g->grid = std::max<uint32_t>(1, std::min<uint32_t>(g->num_cu * g->tune.blocks_per_cu,
                                                      std::max<uint32_t>(g->segs.total_units, 1)));
  for (int b = 0; b < 2; ++b) {
    g->r[b].alloc(sizeof(float) * n);
    g->c[b].alloc(sizeof(float) * n);
  }
  g->partials.alloc(sizeof(double) * 2 * 2 * g->grid);
  g->state.alloc(sizeof(State));
  NWGR_HIP(hipMemsetAsync(g->state.p, 0, sizeof(State), stream));
  NWGR_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->hflag), sizeof(HostFlag), hipHostMallocMapped));
  NWGR_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&g->hflag_dev), g->hflag, 0));
  std::memset(const_cast<HostFlag*>(g->hflag), 0, sizeof(HostFlag));
  NWGR_HIP(hipStreamSynchronize(stream));
  g->seg_rows.swap(deg_h);
  g->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return g.release();
}

}  // namespace nwgr
