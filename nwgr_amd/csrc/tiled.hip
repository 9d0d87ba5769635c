// Construction of the tiled layout of a CSR (see `Tiled` in nwgr_internal.hpp).
//
// Input: a CSR over rows [0, n) whose entries are sorted within each row (the
// relabelled in-CSR, or its transpose).  Everything is built on the device
// with radix sorts and scans; the host only plans the tile padding, the
// workgroup task ranges and the list of rows the combine pass gives a
// wavefront each.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <numeric>

#include "nwgr_internal.hpp"

namespace nwgr {

namespace {

constexpr int kT = 256;

inline uint32_t grid_for(uint64_t work) {
  uint64_t g = (work + kT - 1) / kT;
  return static_cast<uint32_t>(std::clamp<uint64_t>(g, 1, 65535u * 4));
}

uint32_t bits_for(uint64_t v) {
  uint32_t b = 1;
  while (b < 32 && (uint64_t{1} << b) <= v) ++b;
  return b;
}

#define GRID_STRIDE(i, count)                                                                       \
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < (count);     \
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)

// Row of every CSR entry (binary search in offs), tile of every entry, and
// entries per tile.  With `lds_hist`, per-tile counts are first gathered in an
// LDS histogram of ntiles counters per workgroup (the hottest tiles take most
// entries, so counting straight into global memory serialises on a few
// addresses).
__global__ __launch_bounds__(kT) void k_rows_and_tiles(const uint32_t* __restrict__ offs,
                                                       const uint32_t* __restrict__ idx, uint32_t n, uint64_t m,
                                                       uint32_t bits, uint32_t ntiles, bool lds_hist, uint32_t* erow,
                                                       uint32_t* tkey, uint32_t* tile_cnt) {
  extern __shared__ uint32_t s_hist[];
  if (lds_hist) {
    for (uint32_t i = threadIdx.x; i < ntiles; i += kT) s_hist[i] = 0;
    __syncthreads();
  }
  GRID_STRIDE(k, m) {
    uint32_t lo = 0, hi = n;  // offs[lo] <= k < offs[hi]
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (offs[mid] <= k) lo = mid; else hi = mid;
    }
    erow[k] = lo;
    const uint32_t t = idx[k] >> bits;
    tkey[k] = t;
    if (lds_hist)
      atomicAdd(&s_hist[t], 1u);
    else
      atomicAdd(&tile_cnt[t], 1u);
  }
  if (lds_hist) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ntiles; i += kT)
      if (s_hist[i]) atomicAdd(&tile_cnt[i], s_hist[i]);
  }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t count, uint32_t v) {
  GRID_STRIDE(i, count) p[i] = v;
}

// Place entry k' of the tile-sorted order at its padded position and mark the
// first entry of every fragment.
__global__ void k_tile_place(const uint32_t* __restrict__ tkey, const uint32_t* __restrict__ tperm,
                             const uint32_t* __restrict__ idx, const uint32_t* __restrict__ erow,
                             const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_chunk0,
                             uint64_t m, uint32_t mask, uint16_t* eloc, uint32_t* epos, uint32_t* flagv) {
  GRID_STRIDE(k, m) {
    const uint32_t t = tkey[k];
    const uint32_t i = static_cast<uint32_t>(k) - tile_start[t];
    const uint64_t P = static_cast<uint64_t>(tile_chunk0[t]) * kChunkEdges + i;
    const uint32_t e = tperm[k];
    const uint32_t row = erow[e];
    const bool first = i == 0 || (P % kGroupEdges) == 0 || erow[tperm[k - 1]] != row;
    eloc[eloc_pos(P)] = static_cast<uint16_t>((idx[e] & mask) | (first ? 0x8000u : 0u));
    epos[P] = e;
    flagv[P] = first ? 1u : 0u;
  }
}

// The padding after a tile's entries forms dummy fragments (cut at group
// boundaries like the others).
__global__ void k_pad_flags(const uint32_t* __restrict__ tile_cnt, const uint32_t* __restrict__ tile_chunk0,
                            uint32_t ntiles, uint16_t* eloc, uint32_t* flagv) {
  GRID_STRIDE(t, ntiles) {
    const uint64_t lo = static_cast<uint64_t>(tile_chunk0[t]) * kChunkEdges + tile_cnt[t];
    const uint64_t hi = static_cast<uint64_t>(tile_chunk0[t + 1]) * kChunkEdges;
    for (uint64_t P = lo; P < hi; ++P)
      if (P == lo || P % kGroupEdges == 0) {
        eloc[eloc_pos(P)] = 0x8000u;
        flagv[P] = 1u;
      }
  }
}

// Row of every fragment (n for the dummy ones) and fragments per row.
__global__ void k_subpairs(const uint32_t* __restrict__ flagv, const uint32_t* __restrict__ ord,
                           const uint32_t* __restrict__ epos, const uint32_t* __restrict__ erow, uint64_t mp,
                           uint64_t m, uint32_t n, uint32_t* sprow, uint32_t* rowcnt) {
  GRID_STRIDE(P, mp) {
    if (flagv[P]) {
      const uint32_t e = epos[P];
      const uint32_t row = e < m ? erow[e] : n;
      sprow[ord[P]] = row;
      if (row < n) atomicAdd(&rowcnt[row], 1u);
    }
  }
}

// Ordinal of the fragment holding a chunk's first entry: the exclusive count
// of starts before it, minus one when that entry continues a fragment.
__global__ void k_chunk_sp0(const uint32_t* __restrict__ ord, const uint32_t* __restrict__ flagv, uint64_t nchunks,
                            uint32_t* chunk_sp0) {
  GRID_STRIDE(c, nchunks) chunk_sp0[c] = ord[c * kChunkEdges] + flagv[c * kChunkEdges] - 1u;
}

// (row, tile) runs of a CSR with rows sorted by source: one wavefront per row,
// one RMW atomic per wavefront.
__global__ __launch_bounds__(kT) void k_count_runs(const uint32_t* __restrict__ offs,
                                                   const uint32_t* __restrict__ idx, uint32_t n, uint32_t bits,
                                                   unsigned long long* runs) {
  const uint32_t wl = threadIdx.x & 63;
  const uint32_t nw = gridDim.x * (kT / 64);
  uint32_t cnt = 0;
  for (uint32_t row = blockIdx.x * (kT / 64) + threadIdx.x / 64; row < n; row += nw) {
    const uint32_t b = offs[row], e = offs[row + 1];
    for (uint32_t k = b + wl; k < e; k += 64) cnt += (k == b || (idx[k] >> bits) != (idx[k - 1] >> bits)) ? 1u : 0u;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (wl == 0 && cnt) atomicAdd(runs, static_cast<unsigned long long>(cnt));
}

// The slots of the rows that are not big rows, in row order (lslot).
__global__ void k_light_slots(const uint32_t* __restrict__ pp, const uint32_t* __restrict__ lpp,
                              const uint32_t* __restrict__ heavy, const uint32_t* __restrict__ rslot, uint32_t row_lo,
                              uint32_t row_hi, uint32_t* __restrict__ lslot) {
  GRID_STRIDE(i, static_cast<uint64_t>(row_hi - row_lo)) {
    const uint32_t r = row_lo + static_cast<uint32_t>(i);
    if ((heavy[r >> 5] >> (r & 31)) & 1u) continue;
    const uint32_t b = pp[r], e = pp[r + 1], o = lpp[r];
    for (uint32_t k = b; k < e; ++k) lslot[o + (k - b)] = rslot[k];
  }
}

// Weights in tiled order (stored in the chunk storage order of wgt_pos).
__global__ void k_gather_w(const float* __restrict__ src, const uint32_t* __restrict__ pos, uint64_t count,
                           float* __restrict__ dst) {
  GRID_STRIDE(i, count) dst[wgt_pos(i)] = src[pos[i]];
}

// Accumulate form of the second pass.  Row of every fragment from the
// row-ordered slot list (dummy fragments, which no row lists, keep row n).
__global__ void k_frag_rows(const uint32_t* __restrict__ pair_ptr, const uint32_t* __restrict__ rslot, uint32_t n,
                            uint32_t nslots, uint32_t* __restrict__ sprow) {
  GRID_STRIDE(k, nslots) {
    uint32_t lo = 0, hi = n;  // pair_ptr[lo] <= k < pair_ptr[hi]
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (pair_ptr[mid] <= k) lo = mid; else hi = mid;
    }
    sprow[rslot[k]] = lo;
  }
}

// Row of every fragment relative to the first row of its block.
__global__ void k_frag_local(const uint32_t* __restrict__ sprow, const uint32_t* __restrict__ blo, uint32_t nab,
                             uint32_t nsub, uint32_t n, uint16_t* __restrict__ frow) {
  GRID_STRIDE(k, nsub) {
    const uint32_t r = sprow[k];
    uint32_t v = 0;
    if (r < n) {
      uint32_t lo = 0, hi = nab;  // blo[lo] <= r < blo[hi]
      while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (blo[mid] <= r) lo = mid; else hi = mid;
      }
      v = r - blo[lo];
    }
    frow[k] = static_cast<uint16_t>(v);
  }
}

// First fragment ordinal of every tile (a chunk starts a fragment), nsub
// after the last; an empty tile gets its successor's.
__global__ void k_tile_ord(const uint32_t* __restrict__ chunk_sp0, const uint32_t* __restrict__ chunk0,
                           uint32_t ntiles, uint32_t nchunks, uint32_t nsub, uint32_t* __restrict__ tord) {
  GRID_STRIDE(t, static_cast<uint64_t>(ntiles) + 1) {
    const uint32_t c = t < ntiles ? chunk0[t] : nchunks;
    tord[t] = c < nchunks ? chunk_sp0[c] : nsub;
  }
}

// Segment table: the first ordinal of tile t whose fragment row is >= the
// first row of block b (rows ascend inside a tile; dummies carry row n).
__global__ void k_acc_seg(const uint32_t* __restrict__ sprow, const uint32_t* __restrict__ tord,
                          const uint32_t* __restrict__ blo, uint32_t nab, uint32_t ntiles, uint32_t* __restrict__ seg) {
  GRID_STRIDE(i, static_cast<uint64_t>(nab + 1) * ntiles) {
    const uint32_t b = static_cast<uint32_t>(i / ntiles), t = static_cast<uint32_t>(i % ntiles);
    const uint32_t key = blo[b];
    uint32_t lo = tord[t], hi = tord[t + 1];
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (sprow[mid] < key) lo = mid + 1; else hi = mid;
    }
    seg[i] = lo;
  }
}

template <class Fn>
void cub_call(Fn&& fn, hipStream_t s) {
  size_t bytes = 0;
  NWGR_HIP(fn(nullptr, bytes));
  DevBuf t;
  t.alloc(bytes);
  NWGR_HIP(fn(t.p, bytes));
  (void)s;
}

}  // namespace

// A row-range layout (a phase or a row share) holds a fraction of the entries
// but every workgroup of its tile pass fills a whole tile.  It still gets the
// whole layout's workgroups: with half or a quarter of them (fewer tile
// fills) the slowest of eight R-MAT s22 shares took 43.0 / 57.3 us against
// 37.0 us.
uint32_t layout_blocks(const Graph& g, const Tiled& t) {
  (void)t;
  return tile_blocks(g);
}

uint32_t tile_blocks(const Graph& g) {
  // one 1024-thread workgroup per CU holds a 128 KiB tile; narrower tiles
  // leave room for two
  return g.num_cu * (g.tune.tile_bits >= 15 ? 1u : 2u);
}

// Cut the chunk sequence (tile-major) into one contiguous range per
// workgroup, each split at tile boundaries into tasks.  A chunk costs
// 1 + fragments / kFragsPerChunkCost (each fragment adds a partial to reduce,
// stage and store), plus kSecondWindowCost when its fragments need a second
// pass through the wavefront's LDS window and kThirdWindowCost when they need
// a third; every task costs `fill_cost` more (it copies its tile into LDS and
// refills the load pipeline).  The constants are a least-squares fit of the
// workgroups' measured chunk-loop times at R-MAT s22 (256 workgroups: a chunk
// whose 1024 entries each start a fragment costs 3.3 plain chunks).  The
// greedy cut at a cost target is bisected to the smallest target that needs
// at most G ranges.
namespace {
constexpr double kFragsPerChunkCost = 1500.0;
constexpr double kSecondWindowCost = 0.48;
constexpr double kThirdWindowCost = 1.03;

// Estimated cost (us) of the gather pass's items: windows by partials, rows
// and rows a wavefront adds, quiet-row blocks by rows.  The constants are
// those that, dealing the items measured at R-MAT s22 (one sweep, every
// item's time) by this model, gave the shortest slowest list.
float window_cost(uint32_t partials, uint32_t rows, uint32_t wave_rows) {
  return 1.0f + 3.5e-3f * partials + 6.0e-3f * rows + 0.2f * wave_rows;
}
float big_row_cost(uint32_t partials) { return 2.0f + 3.5e-3f * partials; }
float quiet_block_cost(uint32_t rows) { return 0.5f + 6.0e-4f * rows; }

double chunk_cost(uint32_t frags) {
  return 1.0 + frags / kFragsPerChunkCost + (frags > kTileWindow ? kSecondWindowCost : 0.0) +
         (frags > 2 * kTileWindow ? kThirdWindowCost : 0.0);
}

struct Cut {
  std::vector<TiledTask> tasks;
  std::vector<uint32_t> first;  // first task of each range
};

// pc[c]: cost of chunks [0, c)
Cut greedy_cut(const Tiled& t, const std::vector<double>& pc, double F, double target) {
  Cut r;
  double acc = 0;
  bool open = false;
  for (uint32_t tile = 0; tile < t.ntiles; ++tile) {
    uint32_t c = t.tile_chunk0[tile];
    const uint32_t hi = t.tile_chunk0[tile + 1];
    while (c < hi) {
      if (!open || acc + F + (pc[c + 1] - pc[c]) > target) {
        r.first.push_back(static_cast<uint32_t>(r.tasks.size()));
        acc = 0;
        open = true;
      }
      acc += F;
      // the most chunks from c that fit the remaining budget, at least one
      const double limit = pc[c] + (target - acc);
      uint32_t e = static_cast<uint32_t>(std::upper_bound(pc.begin() + c + 1, pc.begin() + hi + 1, limit) -
                                         pc.begin()) - 1;
      e = std::max(e, c + 1);
      r.tasks.push_back({tile, c, e});
      acc += pc[e] - pc[c];
      c = e;
      if (acc >= target) open = false;
    }
  }
  return r;
}

__global__ void k_chunk_frags(const uint32_t* __restrict__ chunk_sp0, uint32_t nchunks, uint32_t nsub,
                              uint32_t* __restrict__ frags) {
  GRID_STRIDE(c, nchunks) {
    const uint32_t a = chunk_sp0[c];
    const uint32_t b = c + 1 < nchunks ? chunk_sp0[c + 1] : nsub;
    frags[c] = b - a;
  }
}
}  // namespace

void plan_tasks(Graph& g, Tiled& t, hipStream_t s) {
  const uint32_t G = layout_blocks(g, t);
  const double F = g.tune.tile_fill_cost;
  // chunk costs from the fragments of every chunk (a chunk starts a group,
  // so its first lane starts a fragment), unless measured re-cuts scaled them
  std::vector<double> pc(static_cast<size_t>(t.nchunks) + 1, 0.0);
  if (t.ccost.size() != t.nchunks) {
    t.ccost.assign(t.nchunks, 0.0);
    t.calib_rounds = 0;
  }
  if (t.nchunks && t.calib_rounds == 0) {
    std::vector<uint32_t> fr(t.nchunks);
    DevBuf d;
    d.alloc(sizeof(uint32_t) * t.nchunks);
    k_chunk_frags<<<grid_for(t.nchunks), kT, 0, s>>>(t.chunk_sp0.as<uint32_t>(), t.nchunks, t.nsub,
                                                      d.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    NWGR_HIP(hipMemcpyAsync(fr.data(), d.p, sizeof(uint32_t) * t.nchunks, hipMemcpyDeviceToHost, s));
    NWGR_HIP(hipStreamSynchronize(s));
    for (uint32_t c = 0; c < t.nchunks; ++c) t.ccost[c] = chunk_cost(fr[c]);
  }
  for (uint32_t c = 0; c < t.nchunks; ++c) pc[c + 1] = pc[c] + t.ccost[c];
  uint32_t busy = 0;
  for (uint32_t tile = 0; tile < t.ntiles; ++tile) busy += t.tile_chunk0[tile + 1] > t.tile_chunk0[tile];
  // one range holding everything always fits
  double hi = pc.back() + F * busy + F + 2 * chunk_cost(kChunkEdges);
  double lo = std::max(1.0, (pc.back() + F * busy) / G);
  Cut best = greedy_cut(t, pc, F, hi);
  for (int it = 0; it < 48 && hi - lo > 0.25; ++it) {
    const double mid = 0.5 * (lo + hi);
    Cut c = greedy_cut(t, pc, F, mid);
    if (c.first.size() <= G) {
      hi = mid;
      best = std::move(c);
    } else {
      lo = mid;
    }
  }
  std::vector<uint32_t> wg(G + 1, static_cast<uint32_t>(best.tasks.size()));
  for (size_t i = 0; i < best.first.size() && i < G; ++i) wg[i] = best.first[i];
  const std::vector<TiledTask>& tasks = best.tasks;
  t.ntasks = static_cast<uint32_t>(tasks.size());
  t.nblocks = G;
  t.fill_cost = g.tune.tile_fill_cost;
  t.tasks.alloc(sizeof(TiledTask) * std::max<size_t>(tasks.size(), 1));
  t.wg_task.alloc(sizeof(uint32_t) * (G + 1));
  std::vector<uint32_t> tile0(G, kNoTile);
  for (uint32_t b = 0; b < G; ++b)
    if (wg[b] < wg[b + 1]) tile0[b] = tasks[wg[b]].tile;
  t.wg_tile0.alloc(sizeof(uint32_t) * G);
  NWGR_HIP(hipMemcpyAsync(t.wg_tile0.p, tile0.data(), sizeof(uint32_t) * G, hipMemcpyHostToDevice, s));
  if (!tasks.empty())
    NWGR_HIP(hipMemcpyAsync(t.tasks.p, tasks.data(), sizeof(TiledTask) * tasks.size(), hipMemcpyHostToDevice, s));
  NWGR_HIP(hipMemcpyAsync(t.wg_task.p, wg.data(), sizeof(uint32_t) * (G + 1), hipMemcpyHostToDevice, s));
  NWGR_HIP(hipStreamSynchronize(s));
  t.tasks_h = tasks;
  t.wg_h = std::move(wg);
}

// A workgroup's modelled cost is its tasks' fills and chunk costs; its
// measured time over that is its ratio, relative to the ratio of all
// workgroups together.  Every chunk of the workgroup takes that factor
// (clamped to [1/2, 2]), so a re-cut moves chunks out of the slow ranges.
bool recut_tasks(Graph& g, Tiled& t, const std::vector<double>& wg_us, hipStream_t s) {
  const uint32_t G = t.nblocks;
  if (wg_us.size() != G || t.wg_h.size() != G + 1 || t.ccost.size() != t.nchunks) return false;
  const double F = t.fill_cost;
  std::vector<double> model(G, 0.0);
  double sm = 0, su = 0;
  for (uint32_t b = 0; b < G; ++b) {
    for (uint32_t k = t.wg_h[b]; k < t.wg_h[b + 1]; ++k) {
      const TiledTask& task = t.tasks_h[k];
      model[b] += F;
      for (uint32_t c = task.chunk_lo; c < task.chunk_hi; ++c) model[b] += t.ccost[c];
    }
    if (model[b] > 0 && wg_us[b] > 0) {
      sm += model[b];
      su += wg_us[b];
    }
  }
  if (sm <= 0 || su <= 0) return false;
  for (uint32_t b = 0; b < G; ++b) {
    if (model[b] <= 0 || wg_us[b] <= 0) continue;
    const double f = std::clamp((wg_us[b] / model[b]) / (su / sm), 0.5, 2.0);
    for (uint32_t k = t.wg_h[b]; k < t.wg_h[b + 1]; ++k) {
      const TiledTask& task = t.tasks_h[k];
      for (uint32_t c = task.chunk_lo; c < task.chunk_hi; ++c) t.ccost[c] *= f;
    }
  }
  t.calib_rounds += 1;
  plan_tasks(g, t, s);
  g.drop_graphs();
  return true;
}

void build_tiled(Graph& g, Tiled& t, const uint32_t* offs, const uint32_t* idx, uint64_t m, uint32_t row_lo,
                 uint32_t row_hi, hipStream_t s) {
  if (t.built) return;
  const uint32_t n = g.n;
  const uint32_t bits = g.tune.tile_bits;
  t.bits = bits;
  t.row_lo = row_lo;
  t.row_hi = row_hi;
  t.ntiles = static_cast<uint32_t>((static_cast<uint64_t>(n) + (1u << bits) - 1) >> bits);
  const uint32_t ntiles = t.ntiles;
  DevBuf erow, tkey, tile_cnt, tk_sorted, tperm;
  erow.alloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  tkey.alloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  tile_cnt.alloc(sizeof(uint32_t) * ntiles);
  NWGR_HIP(hipMemsetAsync(tile_cnt.p, 0, tile_cnt.bytes, s));
  if (m) {
    // an LDS histogram per workgroup for up to 8192 tiles (32 KiB); fewer,
    // fuller workgroups keep the flush of the histograms small
    const bool lds_hist = ntiles <= 8192;
    const uint32_t grid = lds_hist ? std::min<uint32_t>(grid_for(m), g.num_cu * 4) : grid_for(m);
    k_rows_and_tiles<<<grid, kT, lds_hist ? sizeof(uint32_t) * ntiles : 0, s>>>(
        offs, idx, n, m, bits, ntiles, lds_hist, erow.as<uint32_t>(), tkey.as<uint32_t>(),
        tile_cnt.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  std::vector<uint32_t> cnt(ntiles);
  NWGR_HIP(hipMemcpyAsync(cnt.data(), tile_cnt.p, sizeof(uint32_t) * ntiles, hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> start(ntiles + 1, 0);
  t.tile_chunk0.assign(ntiles + 1, 0);
  for (uint32_t i = 0; i < ntiles; ++i) {
    start[i + 1] = start[i] + cnt[i];
    t.tile_chunk0[i + 1] = t.tile_chunk0[i] + (cnt[i] + kChunkEdges - 1) / kChunkEdges;
  }
  t.nchunks = t.tile_chunk0[ntiles];
  const uint64_t mp = static_cast<uint64_t>(t.nchunks) * kChunkEdges;
  DevBuf d_start, d_chunk0;
  d_start.alloc(sizeof(uint32_t) * (ntiles + 1));
  d_chunk0.alloc(sizeof(uint32_t) * (ntiles + 1));
  NWGR_HIP(hipMemcpyAsync(d_start.p, start.data(), sizeof(uint32_t) * (ntiles + 1), hipMemcpyHostToDevice, s));
  NWGR_HIP(hipMemcpyAsync(d_chunk0.p, t.tile_chunk0.data(), sizeof(uint32_t) * (ntiles + 1), hipMemcpyHostToDevice, s));

  // stable sort of the entries by tile keeps (row, source) order inside a tile
  tk_sorted.alloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  tperm.alloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  stable_sort_place(tkey.as<uint32_t>(), m, static_cast<int>(bits_for(ntiles)), ntiles, d_start.as<uint32_t>(),
                    tk_sorted.as<uint32_t>(), tperm.as<uint32_t>(), s);
  tkey.reset();

  t.eloc.alloc(sizeof(uint16_t) * std::max<uint64_t>(mp, 1));
  t.epos.alloc(sizeof(uint32_t) * std::max<uint64_t>(mp, 1));
  DevBuf flagv, ord;
  flagv.alloc(sizeof(uint32_t) * std::max<uint64_t>(mp, 1));
  ord.alloc(sizeof(uint32_t) * std::max<uint64_t>(mp, 1));
  NWGR_HIP(hipMemsetAsync(t.eloc.p, 0, t.eloc.bytes, s));
  NWGR_HIP(hipMemsetAsync(flagv.p, 0, flagv.bytes, s));
  if (mp) {
    k_fill_u32<<<grid_for(mp), kT, 0, s>>>(t.epos.as<uint32_t>(), mp, static_cast<uint32_t>(m));
    NWGR_HIP(hipGetLastError());
  }
  if (m) {
    k_tile_place<<<grid_for(m), kT, 0, s>>>(tk_sorted.as<uint32_t>(), tperm.as<uint32_t>(), idx, erow.as<uint32_t>(),
                                            d_start.as<uint32_t>(), d_chunk0.as<uint32_t>(), m, (1u << bits) - 1,
                                            t.eloc.as<uint16_t>(), t.epos.as<uint32_t>(), flagv.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  k_pad_flags<<<grid_for(ntiles), kT, 0, s>>>(tile_cnt.as<uint32_t>(), d_chunk0.as<uint32_t>(), ntiles,
                                              t.eloc.as<uint16_t>(), flagv.as<uint32_t>());
  NWGR_HIP(hipGetLastError());
  tk_sorted.reset();
  tperm.reset();

  // fragment ordinals
  exclusive_sum_u32(flagv.as<uint32_t>(), ord.as<uint32_t>(), mp, s);
  uint32_t last[2] = {0, 0};
  if (mp) {
    NWGR_HIP(hipMemcpyAsync(&last[0], ord.as<uint32_t>() + mp - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NWGR_HIP(hipMemcpyAsync(&last[1], flagv.as<uint32_t>() + mp - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NWGR_HIP(hipStreamSynchronize(s));
  }
  t.nsub = last[0] + last[1];
  const uint32_t nsub = t.nsub;
  const uint64_t nch = mp / kChunkEdges;
  t.chunk_sp0.alloc(sizeof(uint32_t) * std::max<uint64_t>(nch, 1));
  if (nch) {
    k_chunk_sp0<<<grid_for(nch), kT, 0, s>>>(ord.as<uint32_t>(), flagv.as<uint32_t>(), nch,
                                             t.chunk_sp0.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  DevBuf sprow, rowcnt;
  sprow.alloc(sizeof(uint32_t) * std::max<uint32_t>(nsub, 1));
  rowcnt.alloc(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  NWGR_HIP(hipMemsetAsync(rowcnt.p, 0, rowcnt.bytes, s));
  if (mp) {
    k_subpairs<<<grid_for(mp), kT, 0, s>>>(flagv.as<uint32_t>(), ord.as<uint32_t>(), t.epos.as<uint32_t>(),
                                           erow.as<uint32_t>(), mp, m, n, sprow.as<uint32_t>(),
                                           rowcnt.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  flagv.reset();
  ord.reset();
  erow.reset();
  // each row's fragment ordinals in the stable row order (tiles ascending)
  t.pair_ptr.alloc(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  cub_call([&](void* tmp, size_t& bytes) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, rowcnt.as<uint32_t>(), t.pair_ptr.as<uint32_t>(),
                                            static_cast<int>(n) + 1, s);
  }, s);
  t.rslot.alloc(sizeof(uint32_t) * std::max<uint32_t>(nsub, 1));
  // (dummy fragments carry row n: they sort last, after pair_ptr[n])
  stable_sort_place(sprow.as<uint32_t>(), nsub, static_cast<int>(bits_for(n)), n + 1, t.pair_ptr.as<uint32_t>(),
                    nullptr, t.rslot.as<uint32_t>(), s);
  t.partial.alloc(sizeof(float) * std::max<uint32_t>(nsub, 1));
  NWGR_HIP(hipMemsetAsync(t.partial.p, 0, t.partial.bytes, s));
  // rows whose partials the whole workgroup adds, each an item of its own
  std::vector<uint32_t> pp(static_cast<size_t>(n) + 1);
  NWGR_HIP(hipMemcpyAsync(pp.data(), t.pair_ptr.p, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> big;
  for (uint32_t r = row_lo; r < row_hi; ++r)
    if (pp[r + 1] - pp[r] > kBigRowMin) big.push_back(r);
  t.nbig = static_cast<uint32_t>(big.size());
  // the windows skip the big rows
  std::vector<uint32_t> hb;
  t.light_split = !big.empty();
  std::vector<uint32_t> lpp;
  if (t.light_split) {
    lpp.resize(static_cast<size_t>(n) + 1);
    hb.assign((static_cast<size_t>(n) + 31) / 32, 0u);
    uint64_t acc = 0;
    for (uint32_t r : big) hb[r >> 5] |= 1u << (r & 31);
    for (uint32_t r = 0; r < n; ++r) {
      lpp[r] = static_cast<uint32_t>(acc);
      if (!((hb[r >> 5] >> (r & 31)) & 1u)) acc += pp[r + 1] - pp[r];
    }
    lpp[n] = static_cast<uint32_t>(acc);
    t.lpp.alloc(sizeof(uint32_t) * lpp.size());
    t.heavy.alloc(sizeof(uint32_t) * hb.size());
    t.lslot.alloc(sizeof(uint32_t) * std::max<uint64_t>(acc, 1));
    NWGR_HIP(hipMemcpyAsync(t.lpp.p, lpp.data(), sizeof(uint32_t) * lpp.size(), hipMemcpyHostToDevice, s));
    NWGR_HIP(hipMemcpyAsync(t.heavy.p, hb.data(), sizeof(uint32_t) * hb.size(), hipMemcpyHostToDevice, s));
    if (row_hi > row_lo) {
      k_light_slots<<<grid_for(row_hi - row_lo), kT, 0, s>>>(t.pair_ptr.as<uint32_t>(), t.lpp.as<uint32_t>(),
                                                             t.heavy.as<uint32_t>(), t.rslot.as<uint32_t>(), row_lo,
                                                             row_hi, t.lslot.as<uint32_t>());
      NWGR_HIP(hipGetLastError());
    }
  }
  const std::vector<uint32_t>& wp = t.light_split ? lpp : pp;
  // windows: greedy runs of consecutive rows, cut before kWindowRows rows or
  // kCombineStage partials would be passed, and at nz
  t.win_h.assign(1, row_lo);
  for (uint32_t r = row_lo; r < row_hi; ++r) {
    const uint32_t w0 = t.win_h.back();
    if (r > w0 && (r - w0 >= kWindowRows || wp[r + 1] - wp[w0] > kCombineStage || r == g.nz)) t.win_h.push_back(r);
  }
  if (row_hi > row_lo) t.win_h.push_back(row_hi);
  t.nwin = static_cast<uint32_t>(t.win_h.size() - 1);
  t.big_cost.resize(big.size());
  t.big_h = big;
  t.big_pp.resize(2 * big.size());
  for (size_t i = 0; i < big.size(); ++i) {
    t.big_cost[i] = big_row_cost(pp[big[i] + 1] - pp[big[i]]);
    t.big_pp[2 * i] = pp[big[i]];
    t.big_pp[2 * i + 1] = pp[big[i] + 1];
  }
  t.win_b.resize(t.win_h.size());
  for (size_t k = 0; k < t.win_h.size(); ++k) t.win_b[k] = wp[t.win_h[k]];
  t.win_cost.resize(t.nwin);
  for (uint32_t k = 0; k < t.nwin; ++k) {
    const uint32_t w0 = t.win_h[k], w1 = t.win_h[k + 1];
    uint32_t wave_rows = 0;
    for (uint32_t r = w0; r < w1; ++r) wave_rows += wp[r + 1] - wp[r] > kWaveRowMin;
    t.win_cost[k] = window_cost(wp[w1] - wp[w0], w1 - w0, wave_rows);
  }
  NWGR_HIP(hipStreamSynchronize(s));  // the host vectors must outlive the copies
  plan_tasks(g, t, s);
  plan_gather(g, t, s);
  t.built = true;
}

// Item lists of the gather pass.  Each item goes, heaviest first, to the
// workgroup whose list is cheapest so far -- but to the cheapest list of its
// home XCD when that list is within kAffinitySlack of the cheapest overall.
// Home XCDs cut the items in row order into contiguous runs of about equal
// cost, so neighbouring windows, which share the partials' cache lines at
// their ends in every tile, mostly run on one XCD (workgroup b runs on XCD
// b % kXcds) while the lists still end together.  Each list is worked through
// in the order it was dealt.
constexpr double kAffinitySlack = 1.0;  // estimated us

void plan_gather(Graph& g, Tiled& t, hipStream_t s) {
  const uint32_t G = g.grid_c;
  const uint32_t X = G % kXcds == 0 ? kXcds : 1u;  // XCDs
  const uint32_t quiet_rows = g.n - std::min(g.nz, g.n);
  for (uint32_t mode = 0; mode < 2; ++mode) {
    Tiled::GatherPlan& gp = t.gplan[mode];
    // mode 1: quiet rows [nz, n) kept as the scalar: windows that start
    // before nz, plus the quiet-row blocks when this layout refreshes them
    const uint32_t stop = mode ? std::min(g.nz, t.row_hi) : t.row_hi;
    gp.nwin = static_cast<uint32_t>(std::lower_bound(t.win_h.begin(), t.win_h.begin() + t.nwin, stop) -
                                    t.win_h.begin());
    // (the stored quiet rows are [nz, n_act): the isolated vertices after them
    // store nothing; one block always holds the common rank)
    const uint32_t stored = std::max(g.n_act, std::min(g.nz, g.n)) - std::min(g.nz, g.n);
    gp.nqb = mode && t.quiet_tail && quiet_rows ? std::max(1u, (stored + kQuietRows - 1) / kQuietRows) : 0;
    const uint32_t nitems = t.nbig + gp.nwin + gp.nqb;
    std::vector<float> cost(nitems);
    double total = 0;
    for (uint32_t i = 0; i < nitems; ++i) {
      if (i < t.nbig)
        cost[i] = t.big_cost[i];
      else if (i - t.nbig < gp.nwin)
        cost[i] = t.win_cost[i - t.nbig];
      else
        cost[i] = quiet_block_cost(std::min(kQuietRows, stored - std::min(stored, (i - t.nbig - gp.nwin) * kQuietRows)));
      total += cost[i];
    }
    std::vector<uint32_t> home(nitems);
    {
      double acc = 0;
      for (uint32_t i = 0; i < nitems; ++i) {
        home[i] = std::min<uint32_t>(X - 1, static_cast<uint32_t>((acc + 0.5 * cost[i]) / total * X));
        acc += cost[i];
      }
    }
    std::vector<std::pair<float, uint32_t>> items(nitems);
    for (uint32_t i = 0; i < nitems; ++i) items[i] = {cost[i], i};
    std::stable_sort(items.begin(), items.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    // one (load, workgroup) min-heap per XCD
    auto cmp = [](const auto& a, const auto& b) { return a.first > b.first || (a.first == b.first && a.second > b.second); };
    std::vector<std::vector<std::pair<double, uint32_t>>> heaps(X);
    for (uint32_t b = 0; b < G; ++b) heaps[b % X].push_back({0.0, b});
    for (auto& h : heaps) std::make_heap(h.begin(), h.end(), cmp);
    std::vector<std::vector<uint32_t>> lists(G);
    for (const auto& it : items) {
      uint32_t best = 0;
      for (uint32_t x = 1; x < X; ++x)
        if (cmp(heaps[best].front(), heaps[x].front())) best = x;
      const uint32_t h = home[it.second];
      const uint32_t x = heaps[h].front().first <= heaps[best].front().first + kAffinitySlack ? h : best;
      auto& hp = heaps[x];
      std::pop_heap(hp.begin(), hp.end(), cmp);
      lists[hp.back().second].push_back(it.second);
      hp.back().first += it.first;
      std::push_heap(hp.begin(), hp.end(), cmp);
    }
    std::vector<uint32_t> ptr(static_cast<size_t>(G) + 1, 0), flat;
    flat.reserve(nitems);
    for (uint32_t b = 0; b < G; ++b) {
      ptr[b] = static_cast<uint32_t>(flat.size());
      flat.insert(flat.end(), lists[b].begin(), lists[b].end());
    }
    ptr[G] = static_cast<uint32_t>(flat.size());
    std::vector<GatherDesc> desc(flat.size());
    for (size_t i = 0; i < flat.size(); ++i) {
      const uint32_t it = flat[i];
      if (it < t.nbig) {
        desc[i] = {t.big_h[it], t.big_pp[2 * it], t.big_pp[2 * it + 1], 0u};
      } else if (it - t.nbig < gp.nwin) {
        const uint32_t k = it - t.nbig;
        desc[i] = {t.win_h[k], t.win_h[k + 1], t.win_b[k], t.win_b[k + 1]};
      } else {
        desc[i] = {it - t.nbig - gp.nwin, 0u, 0u, 0u};
      }
    }
    gp.ptr.alloc(sizeof(uint32_t) * ptr.size());
    gp.items.alloc(sizeof(uint32_t) * std::max<size_t>(flat.size(), 1));
    gp.desc.alloc(sizeof(GatherDesc) * std::max<size_t>(flat.size(), 1));
    NWGR_HIP(hipMemcpyAsync(gp.ptr.p, ptr.data(), sizeof(uint32_t) * ptr.size(), hipMemcpyHostToDevice, s));
    if (!flat.empty()) {
      NWGR_HIP(hipMemcpyAsync(gp.items.p, flat.data(), sizeof(uint32_t) * flat.size(), hipMemcpyHostToDevice, s));
      NWGR_HIP(hipMemcpyAsync(gp.desc.p, desc.data(), sizeof(GatherDesc) * desc.size(), hipMemcpyHostToDevice, s));
    }
    NWGR_HIP(hipStreamSynchronize(s));  // the host vectors must outlive the copies
  }
  t.gplan_blocks = G;
}

void build_accum(Graph& g, Tiled& t, hipStream_t s) {
  if (t.acc_built) return;
  const uint32_t n = g.n, nt = t.ntiles, nsub = t.nsub;
  std::vector<uint32_t> pp(static_cast<size_t>(n) + 1);
  NWGR_HIP(hipMemcpyAsync(pp.data(), t.pair_ptr.p, sizeof(uint32_t) * (static_cast<size_t>(n) + 1),
                          hipMemcpyDeviceToHost, s));
  NWGR_HIP(hipStreamSynchronize(s));
  // blocks of consecutive rows: at most kAccRows rows, cut before the
  // fragments would pass kAccFrags (a row alone may exceed it)
  t.acc_rows = kAccRows;
  std::vector<uint32_t> blo{0};
  uint64_t frags = 0;
  for (uint32_t r = 0; r < n; ++r) {
    const uint32_t f = pp[r + 1] - pp[r];
    if (r > blo.back() && (r - blo.back() >= t.acc_rows || frags + f > kAccFrags)) {
      blo.push_back(r);
      frags = 0;
    }
    frags += f;
  }
  blo.push_back(n);
  t.nab = static_cast<uint32_t>(blo.size() - 1);
  t.ablo.alloc(sizeof(uint32_t) * blo.size());
  NWGR_HIP(hipMemcpyAsync(t.ablo.p, blo.data(), sizeof(uint32_t) * blo.size(), hipMemcpyHostToDevice, s));
  DevBuf sprow, tord, chunk0;
  sprow.alloc(sizeof(uint32_t) * std::max<uint32_t>(nsub, 1));
  if (nsub) {
    k_fill_u32<<<grid_for(nsub), kT, 0, s>>>(sprow.as<uint32_t>(), nsub, n);
    NWGR_HIP(hipGetLastError());
  }
  if (pp[n]) {
    k_frag_rows<<<grid_for(pp[n]), kT, 0, s>>>(t.pair_ptr.as<uint32_t>(), t.rslot.as<uint32_t>(), n, pp[n],
                                                sprow.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  t.afrow.alloc(sizeof(uint16_t) * std::max<uint32_t>(nsub, 1));
  if (nsub) {
    k_frag_local<<<grid_for(nsub), kT, 0, s>>>(sprow.as<uint32_t>(), t.ablo.as<uint32_t>(), t.nab, nsub, n,
                                                t.afrow.as<uint16_t>());
    NWGR_HIP(hipGetLastError());
  }
  tord.alloc(sizeof(uint32_t) * (static_cast<size_t>(nt) + 1));
  chunk0.alloc(sizeof(uint32_t) * (static_cast<size_t>(nt) + 1));
  NWGR_HIP(hipMemcpyAsync(chunk0.p, t.tile_chunk0.data(), sizeof(uint32_t) * (static_cast<size_t>(nt) + 1),
                          hipMemcpyHostToDevice, s));
  k_tile_ord<<<grid_for(static_cast<uint64_t>(nt) + 1), kT, 0, s>>>(t.chunk_sp0.as<uint32_t>(), chunk0.as<uint32_t>(),
                                                                     nt, t.nchunks, nsub, tord.as<uint32_t>());
  NWGR_HIP(hipGetLastError());
  const uint64_t nseg = static_cast<uint64_t>(t.nab + 1) * nt;
  t.aseg.alloc(sizeof(uint32_t) * std::max<uint64_t>(nseg, 1));
  if (nseg) {
    k_acc_seg<<<grid_for(nseg), kT, 0, s>>>(sprow.as<uint32_t>(), tord.as<uint32_t>(), t.ablo.as<uint32_t>(), t.nab,
                                             nt, t.aseg.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
  }
  NWGR_HIP(hipStreamSynchronize(s));
  t.acc_built = true;
}

void tiled_weights(Graph& g, Tiled& t, const float* weights, hipStream_t s) {
  const uint64_t mp = static_cast<uint64_t>(t.nchunks) * kChunkEdges;
  if (!t.w.p || t.w.bytes < sizeof(float) * std::max<uint64_t>(mp, 1)) t.w.alloc(sizeof(float) * std::max<uint64_t>(mp, 1));
  if (mp) {
    k_gather_w<<<grid_for(mp), kT, 0, s>>>(weights, t.epos.as<uint32_t>(), mp, t.w.as<float>());
    NWGR_HIP(hipGetLastError());
  }
  (void)g;
}

namespace {
// Offsets of the rows [lo, hi) of a CSR as a CSR of its own over all n rows
// (the other rows empty): offs[i] clamped to [offs[lo], offs[hi]], shifted.
__global__ void k_sub_offsets(const uint32_t* __restrict__ offs, uint32_t n, uint32_t e_lo, uint32_t e_hi,
                              uint32_t* __restrict__ out) {
  GRID_STRIDE(i, static_cast<uint64_t>(n) + 1) out[i] = min(max(offs[i], e_lo), e_hi) - e_lo;
}
}  // namespace

std::vector<uint32_t> gs_cut_rows(const Graph& g, uint32_t phases, const std::vector<uint32_t>* old_cuts,
                                  const std::vector<double>* scale) {
  // Ranges of about equal sweep cost.  A row with e > 0 entries costs
  // kCutEntry * e + kCutFrag * (its fragments, about min(e, tiles) + e / 256)
  // + kCutRow, a row without entries kCutQuiet (the quiet scalar leaves it
  // little work).  The weights are a least-squares fit of the sweep times of
  // the eight row shares of R-MAT s22 under three different cuts (17.2 us
  // fixed per share, then 0.94 per million entries, 0.34 per million
  // fragments, 0.93 per 100 K rows).  Phase q starts at the first row at which
  // the cost of the rows before it reaches q / P of the total.
  constexpr double kCutEntry = 8.0, kCutFrag = 2.9, kCutRow = 80.0, kCutQuiet = 1.0;
  std::vector<uint32_t> cuts(phases + 1, 0);
  const auto& oh = g.offs_h;
  // the tiles that hold sources (the isolated vertices, last, are none)
  const double tiles = static_cast<double>((static_cast<uint64_t>(std::max(g.n_act, 1u)) +
                                            (uint64_t{1} << g.tune.tile_bits) - 1) >> g.tune.tile_bits);
  auto cost = [&](uint32_t r) -> double {
    const double e = static_cast<double>(oh[r + 1] - oh[r]);
    if (e > 0) return kCutEntry * e + kCutFrag * (std::min(e, tiles) + std::floor(e / 256)) + kCutRow;
    return r < g.n_act ? kCutQuiet : 0.0;  // an isolated vertex is never stored
  };
  // (a measured re-cut: every row of old range k costs scale[k] times more)
  const bool scaled = old_cuts && scale && scale->size() + 1 == old_cuts->size();
  std::vector<double> rc(g.n);
  uint32_t k = 0;
  for (uint32_t r = 0; r < g.n; ++r) {
    if (scaled)
      while (k + 1 < scale->size() && r >= (*old_cuts)[k + 1]) ++k;
    rc[r] = cost(r) * (scaled ? (*scale)[k] : 1.0);
  }
  double total = 0;
  for (uint32_t r = 0; r < g.n; ++r) total += rc[r];
  double acc = 0;
  uint32_t q = 1;
  for (uint32_t r = 0; r < g.n && q < phases; ++r) {
    while (q < phases && acc >= total * q / phases) cuts[q++] = r;
    acc += rc[r];
  }
  for (; q < phases; ++q) cuts[q] = g.n;
  cuts[phases] = g.n;
  return cuts;
}

void ensure_gs(Graph& g, uint32_t phases, hipStream_t s) {
  if (g.gs.size() == phases && g.gs_bits == g.tune.tile_bits) {
    bool replanned = false;
    for (Tiled& t : g.gs)
      if (t.nblocks != layout_blocks(g, t) || t.fill_cost != g.tune.tile_fill_cost) {
        plan_tasks(g, t, s);
        replanned = true;
      }
    // captured sweeps hold the old task lists
    if (replanned) g.drop_graphs(static_cast<int>(kExecGs));
    return;
  }
  build_gs(g, gs_cut_rows(g, phases), s);
}

void build_gs(Graph& g, const std::vector<uint32_t>& cuts, hipStream_t s) {
  const uint32_t phases = static_cast<uint32_t>(cuts.size() - 1);
  g.gs.clear();
  g.drop_graphs(static_cast<int>(kExecGs));
  DevBuf sub;
  sub.alloc(sizeof(uint32_t) * (static_cast<size_t>(g.n) + 1));
  std::vector<Tiled> layouts(phases);
  for (uint32_t q = 0; q < phases; ++q) {
    const uint32_t lo = cuts[q], hi = cuts[q + 1];
    const uint32_t e_lo = g.offs_h[lo], e_hi = g.offs_h[hi];
    k_sub_offsets<<<grid_for(static_cast<uint64_t>(g.n) + 1), kT, 0, s>>>(g.offs.as<uint32_t>(), g.n, e_lo, e_hi,
                                                                          sub.as<uint32_t>());
    NWGR_HIP(hipGetLastError());
    layouts[q].quiet_tail = q + 1 == phases;
    build_tiled(g, layouts[q], sub.as<uint32_t>(), g.idx.as<uint32_t>() + e_lo, e_hi - e_lo, lo, hi, s);
    NWGR_HIP(hipStreamSynchronize(s));  // `sub` is rewritten for the next phase
  }
  g.gs = std::move(layouts);
  g.gs_cuts = cuts;
  g.gs_bits = g.tune.tile_bits;
}

Tiled* select_tiled(Graph& g, bool adj, hipStream_t s) {
  if (g.tune.engine == kEngineBinned) return nullptr;
  const uint32_t* offs = adj ? g.t_offs.as<uint32_t>() : g.offs.as<uint32_t>();
  const uint32_t* idx = adj ? g.t_idx.as<uint32_t>() : g.idx.as<uint32_t>();
  Tiled& t = adj ? g.tadj : g.tfwd;
  if (g.tune.engine == kEngineAuto) {
    if (g.auto_bits != g.tune.tile_bits) {
      g.auto_tiled[0] = g.auto_tiled[1] = -1;
      g.auto_bits = g.tune.tile_bits;
    }
    int& choice = g.auto_tiled[adj ? 1 : 0];
    if (choice < 0) {
      unsigned long long runs = 0;
      if (g.m) {
        DevBuf d;
        d.alloc(sizeof(unsigned long long));
        NWGR_HIP(hipMemsetAsync(d.p, 0, sizeof(unsigned long long), s));
        const uint32_t blocks = std::clamp<uint32_t>((g.n + kT / 64 - 1) / (kT / 64), 1, 8192);
        k_count_runs<<<blocks, kT, 0, s>>>(offs, idx, g.n, g.tune.tile_bits, d.as<unsigned long long>());
        NWGR_HIP(hipGetLastError());
        NWGR_HIP(hipMemcpyAsync(&runs, d.p, sizeof(runs), hipMemcpyDeviceToHost, s));
        NWGR_HIP(hipStreamSynchronize(s));
      }
      choice = (runs == 0 || static_cast<double>(g.m) >= kTiledMinRun * static_cast<double>(runs) ||
                g.n > kTiledMinN) ? 1 : 0;
    }
    if (!choice) return nullptr;
  }
  build_tiled(g, t, offs, idx, g.m, 0, g.n, s);
  if (t.nblocks != layout_blocks(g, t) || t.fill_cost != g.tune.tile_fill_cost) plan_tasks(g, t, s);
  // second pass: accumulate when rows have about one entry per tile run
  const uint32_t mode = g.tune.tile_combine;
  const bool acc = mode == kCombineAccumulate ||
                   (mode == kCombineAuto && 2 * static_cast<uint64_t>(g.m) < 3 * static_cast<uint64_t>(t.nsub));
  if (acc) build_accum(g, t, s);
  if (acc != t.acc_on) {
    t.acc_on = acc;
    g.drop_graphs();
  }
  return &t;
}

}  // namespace nwgr
