// Device-resident graph construction for the PageRank sweeps.
//
// Input is the reference's in-CSR (adjacency<1>, build.hpp:209-265) plus the
// caller's degree array (bench/common.hpp:109-120).  On the device the graph
// is relabelled so that rows appear in non-increasing in-degree order (ties:
// ascending original id; among the rows without entries, the sources of some
// entry before the isolated vertices).  Degree bins then become contiguous row ranges and
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

// Rows that are a source of some entry (every writer stores the same 1).
__global__ void k_mark_sources(const uint32_t* __restrict__ idx, uint64_t m, uint32_t* has_out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride)
    has_out[idx[k]] = 1u;
}

// Relabelling key: in-degree + 1 for a row with entries; for a row without,
// 1 when it is a source, 0 when the vertex is isolated -- so the
// non-increasing order ends with the isolated vertices, whose contributions
// no sweep reads.
__global__ void k_relabel_key(const uint32_t* __restrict__ offs, const uint32_t* __restrict__ has_out, uint32_t n,
                              uint32_t* key, uint32_t* val) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n; v += stride) {
    const uint32_t d = offs[v + 1] - offs[v];
    key[v] = d ? d + 1 : has_out[v];
    val[v] = static_cast<uint32_t>(v);
  }
}

// Sorted keys back to in-degrees.
__global__ void k_key_to_indeg(uint32_t* key, uint32_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n; v += stride)
    key[v] = key[v] >= 2 ? key[v] - 1 : 0u;
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

// Passes form of the edge order (more entries than one sort pass holds):
// renamed source of every entry, counted per source.
__global__ void k_new_src(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ iperm, uint64_t m,
                          uint32_t* __restrict__ src, uint32_t* __restrict__ cnt) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride) {
    const uint32_t u = iperm[idx[k]];
    src[k] = u;
    atomicAdd(&cnt[u], 1u);
  }
}

// New row of the entry that sits at position j of the source order.
__global__ void k_row_keys(const uint32_t* __restrict__ offs, const uint32_t* __restrict__ iperm,
                           const uint32_t* __restrict__ by_src, uint32_t n, uint64_t m, uint32_t* __restrict__ key) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < m; j += stride) {
    const uint32_t k = by_src[j];
    uint32_t lo = 0, hi = n;  // offs[lo] <= k < offs[hi]
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (offs[mid] <= k) lo = mid; else hi = mid;
    }
    key[j] = iperm[lo];
  }
}

// Position p of the new CSR takes entry by_src[by_row[p]].
__global__ void k_place_entries(const uint32_t* __restrict__ by_src, const uint32_t* __restrict__ by_row,
                                const uint32_t* __restrict__ src, uint64_t m, uint32_t* __restrict__ idx,
                                uint32_t* __restrict__ epos) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < m; p += stride) {
    const uint32_t k = by_src[by_row[p]];
    epos[p] = k;
    idx[p] = src[k];
  }
}

__global__ void k_split_keys(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* idx) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride)
    idx[k] = static_cast<uint32_t>(keys[k]);
}

// Divisors in the new order: as float (what the reference's Real = float
// divides by) and as given (the FP64 solve converts them exactly).
__global__ void k_degf(const uint32_t* __restrict__ deg, const uint32_t* __restrict__ perm, uint32_t n,
                       float* degf, uint32_t* degu) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    const uint32_t d = deg[perm[i]];
    degf[i] = static_cast<float>(d);
    degu[i] = d;
  }
}

__global__ void k_set_last(uint32_t* offs, uint32_t n, uint32_t m) { offs[n] = m; }

uint32_t bits_for(uint32_t n) {
  uint32_t b = 1;
  while (b < 32 && (uint64_t{1} << b) < n) ++b;
  return b;
}

}  // namespace

void build_plan(Graph& g, Plan& plan, const std::vector<uint32_t>& deg, const std::vector<uint32_t>& offs_h,
                const std::vector<uint32_t>& row_of, hipStream_t stream) {
  const Tuning& t = g.tune;
  const uint32_t n = g.n;
  auto row_at = [&](uint32_t pos) { return row_of.empty() ? pos : row_of[pos]; };
  auto first_below_or_eq = [&](uint32_t x) -> uint32_t {
    // first position with deg <= x (deg is non-increasing)
    return static_cast<uint32_t>(
        std::partition_point(deg.begin(), deg.end(), [&](uint32_t d) { return d > x; }) - deg.begin());
  };
  std::vector<uint64_t> pre(static_cast<size_t>(n) + 1, 0);  // edges before a position
  for (uint32_t i = 0; i < n; ++i) pre[i + 1] = pre[i] + deg[i];
  SegTable st{};
  uint32_t unit = 0;
  // hub rows: whole workgroups over chunks of hub_chunk edges
  const uint32_t h_end = first_below_or_eq(t.hub_threshold);
  std::vector<HubUnit> hubs;
  for (uint32_t pos = 0; pos < h_end; ++pos) {
    const uint32_t row = row_at(pos);
    const uint32_t d = deg[pos];
    const uint32_t nch = (d + t.hub_chunk - 1) / t.hub_chunk;
    const uint32_t first = static_cast<uint32_t>(hubs.size());
    for (uint32_t c = 0; c < nch; ++c) {
      HubUnit u;
      u.row = row;
      u.e_lo = offs_h[row] + c * t.hub_chunk;
      u.e_hi = std::min(offs_h[row] + (c + 1) * t.hub_chunk, offs_h[row + 1]);
      u.first = first;
      u.nchunks = nch;
      u.hub = pos;
      hubs.push_back(u);
    }
  }
  if (!hubs.empty()) {
    Seg& sg = st.s[st.nseg++];
    sg.row_lo = 0;
    sg.row_hi = h_end;
    sg.G = 0;
    sg.rows_per_unit = 1;
    sg.unit_lo = unit;
    unit += static_cast<uint32_t>(hubs.size());
    sg.unit_hi = unit;
  }
  // Lane-group segments.  A band takes the rows with dlo < degree <= the
  // previous band's dlo; its rows get G lanes each, every lane loads U
  // entries per batch (one batch covers the band's longest row except in the
  // first band, which loops), and every lane group works on P rows at once so
  // that P independent load chains are in flight.
  uint32_t pos = h_end;
  struct Band { uint32_t G, dlo, U, P; };
  const Band bands[] = {{64, 128, 8, 1}, {64, 64, 2, 2}, {32, 32, 2, 2}, {16, 16, 2, 4}, {8, 8, 2, 4},
                        {4, 4, 2, 4},    {2, 2, 2, 4},   {1, 0, 2, 4}};
  for (const Band& bd : bands) {
    const uint32_t end = first_below_or_eq(bd.dlo);
    if (end > pos) {
      Seg& sg = st.s[st.nseg++];
      sg.row_lo = pos;
      sg.row_hi = end;
      sg.G = bd.G;
      sg.U = bd.U;
      sg.P = bd.P;
      const double avg = static_cast<double>(pre[end] - pre[pos]) / (end - pos);
      const uint32_t per_pass = kWave / bd.G * bd.P;
      const double passes = t.unit_edges / std::max(1.0, per_pass * avg);
      const uint32_t R = static_cast<uint32_t>(std::clamp(passes, 1.0, 256.0));
      sg.rows_per_unit = per_pass * R;
      sg.unit_lo = unit;
      unit += (end - pos + sg.rows_per_unit - 1) / sg.rows_per_unit;
      sg.unit_hi = unit;
      pos = end;
    }
  }
  // rows without entries: teleport/dangling terms only
  if (n > pos) {
    Seg& sg = st.s[st.nseg++];
    sg.row_lo = pos;
    sg.row_hi = n;
    sg.G = 1;
    sg.U = 0;
    sg.P = 4;
    sg.rows_per_unit = std::max<uint32_t>(kWave * 4, t.zero_rows_per_unit / (kWave * 4) * (kWave * 4));
    sg.unit_lo = unit;
    unit += (n - pos + sg.rows_per_unit - 1) / sg.rows_per_unit;
    sg.unit_hi = unit;
  }
  st.total_units = unit;
  plan.segs = st;
  plan.n_hub_units = static_cast<uint32_t>(hubs.size());
  plan.n_hubs = h_end;
  plan.hub_units.alloc(sizeof(HubUnit) * std::max<size_t>(hubs.size(), 1));
  if (!hubs.empty())
    NWGR_HIP(hipMemcpyAsync(plan.hub_units.p, hubs.data(), sizeof(HubUnit) * hubs.size(), hipMemcpyHostToDevice,
                            stream));
  plan.hub_part.alloc(sizeof(uint32_t) * std::max<size_t>(hubs.size(), 1));
  plan.hub_ticket.alloc(sizeof(uint32_t) * std::max<uint32_t>(h_end, 1));
  NWGR_HIP(hipMemsetAsync(plan.hub_ticket.p, 0, plan.hub_ticket.bytes, stream));
  if (!row_of.empty()) {
    plan.rows.alloc(sizeof(uint32_t) * n);
    NWGR_HIP(hipMemcpyAsync(plan.rows.p, row_of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, stream));
  } else {
    plan.rows.reset();
  }
  // the host vectors must outlive the copies
  NWGR_HIP(hipStreamSynchronize(stream));
}

namespace {

// Out-CSR keys of the new in-CSR: edge k (row v, source u) -> (u << 32 | v), value k.
__global__ void k_transpose_keys(const uint32_t* __restrict__ offs, const uint32_t* __restrict__ idx, uint32_t n,
                                 uint64_t m, uint64_t* keys, uint32_t* vals, uint32_t* outdeg) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < m; k += stride) {
    uint32_t lo = 0, hi = n;  // offs[lo] <= k < offs[hi]
    while (hi - lo > 1) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (offs[mid] <= k) lo = mid; else hi = mid;
    }
    const uint32_t u = idx[k];
    keys[k] = (static_cast<uint64_t>(u) << 32) | lo;
    vals[k] = static_cast<uint32_t>(k);
    atomicAdd(&outdeg[u], 1u);
  }
}

__global__ void k_iota(uint32_t n, uint32_t* v) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    v[i] = static_cast<uint32_t>(i);
}

}  // namespace

void ensure_transpose(Graph& g, hipStream_t stream) {
  if (g.has_t) return;
  const uint32_t n = g.n;
  const uint64_t m = g.m;
  // one radix sort of (source, row) keys over every entry, counted in int
  NWGR_REQUIRE(m < 0x80000000ULL, "the out-CSR (weighted solves) needs m < 2^31");
  const size_t nb = sizeof(uint32_t) * n, mb = sizeof(uint32_t) * m;
  DevBuf outdeg, ids, outdeg_sorted, row_list;
  outdeg.alloc(nb);
  NWGR_HIP(hipMemsetAsync(outdeg.p, 0, nb, stream));
  g.t_idx.alloc(mb + sizeof(uint32_t) * kEdgePad);
  NWGR_HIP(hipMemsetAsync(g.t_idx.p, 0, g.t_idx.bytes, stream));
  g.t_pos.alloc(mb);
  if (m) {
    DevBuf k_in, k_out, v_in;
    k_in.alloc(sizeof(uint64_t) * m);
    k_out.alloc(sizeof(uint64_t) * m);
    v_in.alloc(mb);
    k_transpose_keys<<<grid_for(m), kT, 0, stream>>>(g.offs.as<uint32_t>(), g.idx.as<uint32_t>(), n, m,
                                                     k_in.as<uint64_t>(), v_in.as<uint32_t>(), outdeg.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    const int end_bit = 32 + static_cast<int>(bits_for(n));
    size_t tmp = 0;
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k_in.as<uint64_t>(), k_out.as<uint64_t>(),
                                                v_in.as<uint32_t>(), g.t_pos.as<uint32_t>(), static_cast<int>(m), 0,
                                                end_bit, stream));
    DevBuf t;
    t.alloc(tmp);
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, k_in.as<uint64_t>(), k_out.as<uint64_t>(),
                                                v_in.as<uint32_t>(), g.t_pos.as<uint32_t>(), static_cast<int>(m), 0,
                                                end_bit, stream));
    k_split_keys<<<grid_for(m), kT, 0, stream>>>(k_out.as<uint64_t>(), m, g.t_idx.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  g.t_offs.alloc(nb + sizeof(uint32_t));
  {
    size_t tmp = 0;
    NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, outdeg.as<uint32_t>(), g.t_offs.as<uint32_t>(),
                                              static_cast<int>(n), stream));
    DevBuf t;
    t.alloc(tmp);
    NWGR_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, outdeg.as<uint32_t>(), g.t_offs.as<uint32_t>(),
                                              static_cast<int>(n), stream));
    k_set_last<<<1, 1, 0, stream>>>(g.t_offs.as<uint32_t>(), n, static_cast<uint32_t>(m));
    NWGR_HIP(hipGetLastError());
  }
  // rows of the out-CSR in non-increasing out-degree order (stable)
  ids.alloc(nb);
  outdeg_sorted.alloc(nb);
  row_list.alloc(nb);
  k_iota<<<grid_for(n), kT, 0, stream>>>(n, ids.as<uint32_t>());
  {
    size_t tmp = 0;
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, outdeg.as<uint32_t>(),
                                                         outdeg_sorted.as<uint32_t>(), ids.as<uint32_t>(),
                                                         row_list.as<uint32_t>(), static_cast<int>(n), 0, 32, stream));
    DevBuf t;
    t.alloc(tmp);
    NWGR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(t.p, tmp, outdeg.as<uint32_t>(),
                                                         outdeg_sorted.as<uint32_t>(), ids.as<uint32_t>(),
                                                         row_list.as<uint32_t>(), static_cast<int>(n), 0, 32, stream));
  }
  std::vector<uint32_t> deg_h(n), rows_h(n), toffs_h(static_cast<size_t>(n) + 1);
  NWGR_HIP(hipMemcpyAsync(deg_h.data(), outdeg_sorted.p, nb, hipMemcpyDeviceToHost, stream));
  NWGR_HIP(hipMemcpyAsync(rows_h.data(), row_list.p, nb, hipMemcpyDeviceToHost, stream));
  NWGR_HIP(hipMemcpyAsync(toffs_h.data(), g.t_offs.p, nb + sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  NWGR_HIP(hipStreamSynchronize(stream));
  build_plan(g, g.adj, deg_h, toffs_h, rows_h, stream);
  g.t_deg_h.swap(deg_h);
  g.t_rows_h.swap(rows_h);
  g.t_offs_h.swap(toffs_h);
  g.has_t = true;
}

void validate_tuning(const Tuning& t) {
  NWGR_REQUIRE(t.blocks_per_cu >= 1 && t.blocks_per_cu <= 2, "blocks_per_cu must be 1 or 2");
  NWGR_REQUIRE(t.lds_cache <= kLdsCache / t.blocks_per_cu,
               "lds_cache exceeds the LDS share of a sweep workgroup (36864 / blocks_per_cu)");
  NWGR_REQUIRE(t.hub_threshold >= 128, "hub_threshold must be >= 128");
  NWGR_REQUIRE(t.hub_chunk >= kWave, "hub_chunk must be >= 64");
  NWGR_REQUIRE(t.unit_edges >= 1 && t.zero_rows_per_unit >= 1, "unit sizes must be positive");
  NWGR_REQUIRE(t.chunk >= 1 && t.chunk <= Graph::kMaxChunk, "chunk must be in [1, 32]");
  NWGR_REQUIRE(t.engine <= kEngineAuto, "engine must be 0 (degree-binned), 1 (tiled) or 2 (auto)");
  NWGR_REQUIRE(t.tile_bits >= 13 && t.tile_bits <= kMaxTileBits, "tile_bits must be in [13, 15]");
  NWGR_REQUIRE(t.tile_fill_cost <= 65536, "tile_fill_cost must be <= 65536");
  NWGR_REQUIRE(t.tile_combine <= kCombineAuto, "tile_combine must be 0 (gather), 1 (accumulate) or 2 (auto)");
}

void replan(Graph& g, hipStream_t stream) {
  const Tuning& t = g.tune;
  validate_tuning(t);
  build_plan(g, g.fwd, g.in_deg_h, g.offs_h, {}, stream);
  if (g.has_t) build_plan(g, g.adj, g.t_deg_h, g.t_offs_h, g.t_rows_h, stream);
  // tiled layouts follow the tile width; their task ranges follow the fill cost
  for (Tiled* tl : {&g.tfwd, &g.tadj}) {
    if (!tl->built) continue;
    if (tl->bits != t.tile_bits)
      *tl = Tiled{};  // rebuilt on next use
    else if (tl->nblocks != layout_blocks(g, *tl) || tl->fill_cost != t.tile_fill_cost)
      plan_tasks(g, *tl, stream);
  }
  if (!g.gs.empty() && g.gs_bits != t.tile_bits) g.gs.clear();  // the phases' layouts likewise
  for (Tiled& tl : g.gs)
    if (tl.nblocks != layout_blocks(g, tl) || tl.fill_cost != t.tile_fill_cost) plan_tasks(g, tl, stream);
  g.grid = std::max<uint32_t>(
      1, std::min<uint32_t>(g.num_cu * t.blocks_per_cu, std::max<uint32_t>(g.fwd.segs.total_units, 64)));
  // the gather pass runs one wave of resident workgroups taking items from a queue
  g.grid_c = std::max<uint32_t>(1, g.num_cu * combine_blocks_per_cu());
  for (Tiled* tl : {&g.tfwd, &g.tadj})
    if (tl->built && tl->gplan_blocks != g.grid_c) plan_gather(g, *tl, stream);
  for (Tiled& tl : g.gs)
    if (tl.gplan_blocks != g.grid_c) plan_gather(g, tl, stream);
  const uint32_t stride = std::max(g.grid, g.grid_c);
  if (stride != g.part_stride || !g.partials.p) {
    g.part_stride = stride;
    g.partials.alloc(sizeof(double) * 2 * 2 * stride);
  }
  g.drop_graphs();
}

Graph::~Graph() {
  drop_graphs();
  if (hflag) (void)hipHostFree(hflag);
  if (hist_h) (void)hipHostFree(hist_h);
  if (pin_out) (void)hipHostFree(pin_out);
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  if (own_stream) (void)hipStreamDestroy(own_stream);
}

// A copy into pageable memory goes through the runtime's own staging at about
// half the rate of one into pinned memory and a host copy after it (4 M
// floats: 1.0 ms against 0.31 + 0.22 ms), and ran 1.4 ms slower still for the
// first dozen calls after a second large graph was built.
void copy_out(Graph& g, void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (g.pin_bytes < bytes) {
    if (g.pin_out) NWGR_HIP(hipHostFree(g.pin_out));
    g.pin_out = nullptr;
    g.pin_bytes = 0;
    NWGR_HIP(hipHostMalloc(&g.pin_out, bytes, hipHostMallocDefault));
    g.pin_bytes = bytes;
  }
  NWGR_HIP(hipMemcpyAsync(g.pin_out, src, bytes, hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  std::memcpy(dst, g.pin_out, bytes);
}

Graph* create_graph(const uint32_t* offsets, const uint32_t* indices, uint32_t n, uint64_t m,
                    const uint32_t* degrees, int device, hipStream_t stream, bool on_device) {
  if (n == 0) {
    // a graph without vertices (page_rank.hpp:92 with N = 0): no device
    // memory, and every solve on it returns at once
    NWGR_REQUIRE(m == 0, "a graph without vertices has no entries (m must be 0)");
    NWGR_REQUIRE(on_device || offsets == nullptr || offsets[0] == 0, "offsets[0] must be 0");
    auto g = std::make_unique<Graph>();
    g->device = device;
    return g.release();
  }
  NWGR_REQUIRE(m <= kMaxGraphEntries, "m must be <= 2^32 - 2^28 (uint32 offsets with room for the tile padding)");
  // the sorts over vertices count items in int
  NWGR_REQUIRE(n < 0x80000000u, "n must be < 2^31");
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

  // --- relabel rows by non-increasing in-degree (stable => ties by id; among
  // the rows without entries the sources first, then the isolated vertices)
  DevBuf key_in, val_in, key_out;
  key_in.alloc(nb);
  val_in.alloc(nb);
  key_out.alloc(nb);
  g->perm.alloc(nb);
  g->iperm.alloc(nb);
  {
    DevBuf has_out;
    has_out.alloc(nb);
    NWGR_HIP(hipMemsetAsync(has_out.p, 0, nb, stream));
    if (m) k_mark_sources<<<grid_for(m), kT, 0, stream>>>(d_idx, m, has_out.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    k_relabel_key<<<grid_for(n), kT, 0, stream>>>(d_offs, has_out.as<uint32_t>(), n, key_in.as<uint32_t>(),
                                                  val_in.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    NWGR_HIP(hipStreamSynchronize(stream));  // has_out is freed
  }
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
  // the isolated vertices: the zero keys at the end
  {
    std::vector<uint32_t> keys_h(n);
    NWGR_HIP(hipMemcpyAsync(keys_h.data(), key_out.p, nb, hipMemcpyDeviceToHost, stream));
    NWGR_HIP(hipStreamSynchronize(stream));
    uint32_t act = n;
    while (act > 0 && keys_h[act - 1] == 0) --act;
    g->n_act = act;
  }
  k_key_to_indeg<<<grid_for(n), kT, 0, stream>>>(key_out.as<uint32_t>(), n);
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
  g->idx.alloc(mb + sizeof(uint32_t) * kEdgePad);
  NWGR_HIP(hipMemsetAsync(g->idx.p, 0, g->idx.bytes, stream));
  g->epos.alloc(mb);
  if (m > sort_batch()) {
    // More entries than one sort pass: the (row, source) order as two
    // stable passes sorts, by source and then by row, the entries in input
    // order (so equal pairs keep it, as in the one-pass sort below).
    DevBuf src, cnt, first, by_src, key, by_row;
    src.alloc(mb);
    cnt.alloc(nb);
    first.alloc(nb);
    NWGR_HIP(hipMemsetAsync(cnt.p, 0, nb, stream));
    k_new_src<<<grid_for(m), kT, 0, stream>>>(d_idx, g->iperm.as<uint32_t>(), m, src.as<uint32_t>(),
                                              cnt.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    exclusive_sum_u32(cnt.as<uint32_t>(), first.as<uint32_t>(), n, stream);
    const int end_bit = static_cast<int>(bits_for(n));
    by_src.alloc(mb);
    stable_sort_place(src.as<uint32_t>(), m, end_bit, n, first.as<uint32_t>(), nullptr, by_src.as<uint32_t>(),
                      stream);
    cnt.reset();
    first.reset();
    key.alloc(mb);
    k_row_keys<<<grid_for(m), kT, 0, stream>>>(d_offs, g->iperm.as<uint32_t>(), by_src.as<uint32_t>(), n, m,
                                               key.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    by_row.alloc(mb);
    stable_sort_place(key.as<uint32_t>(), m, end_bit, n, g->offs.as<uint32_t>(), nullptr, by_row.as<uint32_t>(),
                      stream);
    key.reset();
    k_place_entries<<<grid_for(m), kT, 0, stream>>>(by_src.as<uint32_t>(), by_row.as<uint32_t>(), src.as<uint32_t>(),
                                                    m, g->idx.as<uint32_t>(), g->epos.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    NWGR_HIP(hipStreamSynchronize(stream));
  } else if (m) {
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
  g->degu.alloc(sizeof(uint32_t) * n);
  k_degf<<<grid_for(n), kT, 0, stream>>>(d_deg, g->perm.as<uint32_t>(), n, g->degf.as<float>(),
                                         g->degu.as<uint32_t>());
  NWGR_HIP(hipGetLastError());

  // --- plan: host copies of the sorted degrees and new offsets
  std::vector<uint32_t> deg_h(n), offs_h(static_cast<size_t>(n) + 1);
  NWGR_HIP(hipMemcpyAsync(deg_h.data(), key_out.p, nb, hipMemcpyDeviceToHost, stream));
  NWGR_HIP(hipMemcpyAsync(offs_h.data(), g->offs.p, nb + sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  NWGR_HIP(hipStreamSynchronize(stream));
  g->in_deg_h.swap(deg_h);
  g->offs_h.swap(offs_h);
  // rows without entries form the tail [nz, n) of the degree order (else no
  // quiet rows are used)
  {
    uint32_t nz = 0;
    while (nz < n && g->offs_h[nz + 1] > g->offs_h[nz]) ++nz;
    for (uint32_t r = nz; r < n; ++r)
      if (g->offs_h[r + 1] != g->offs_h[r]) {
        nz = n;
        break;
      }
    g->nz = nz;
    g->n_act = std::max(g->n_act, nz);
  }
  replan(*g, stream);  // plans and sweep grid for the default tuning

  // --- iteration buffers
  // index n is the zero sentinel of the gathers; the zero tail up to whole
  // maximal tiles lets the tiled engine copy tiles without bounds checks
  for (int b = 0; b < 2; ++b) {
    g->r[b].alloc(sizeof(float) * iter_len(n));
    g->c[b].alloc(sizeof(float) * iter_len(n));
    NWGR_HIP(hipMemsetAsync(g->r[b].p, 0, g->r[b].bytes, stream));
    NWGR_HIP(hipMemsetAsync(g->c[b].p, 0, g->c[b].bytes, stream));
  }
  g->state.alloc(sizeof(State));
  NWGR_HIP(hipMemsetAsync(g->state.p, 0, sizeof(State), stream));
  NWGR_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->hflag), sizeof(HostFlag), hipHostMallocMapped));
  NWGR_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&g->hflag_dev), g->hflag, 0));
  std::memset(const_cast<HostFlag*>(g->hflag), 0, sizeof(HostFlag));
  g->hist_cap = kHistCap;
  NWGR_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->hist_h), sizeof(double) * g->hist_cap, hipHostMallocMapped));
  NWGR_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&g->hist_dev), g->hist_h, 0));
  // the layout the sweeps of the default tuning use is built here, so the
  // first solve costs what every later one does (less its graph captures)
  (void)select_tiled(*g, false, stream);
  prepare_quiet(*g, stream);
  // and its tile pass cut by measured workgroup times (large layouts only)
  calibrate_tile_pass(*g, stream);
  NWGR_HIP(hipStreamSynchronize(stream));
  g->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return g.release();
}

}  // namespace nwgr
