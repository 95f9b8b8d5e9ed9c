// Changed-ghost-label delta pack: the compute half of the reference's
// exchangeplan_exchange_delta (src/exchange_plan.c:325-357).
//
// For every entry idx of send_to_flat (the owned vertices each neighbour
// rank asked for, concatenated per destination in dests order):
//   lbl = comp_label[comp_of[v - v_start]] if v is owned, else v;
//   if lbl < prev_sent[idx]: prev_sent[idx] = lbl and emit
//     ((uint64_t)(idx - segment_start) << 32) | lbl
// in ascending idx order, i.e. exactly the reference's send buffer, plus
// per-destination counts and displacements.  comp_of is fixed through the
// merge loop, so the reps comp_of[v - v_start] can be gathered once
// (cc_hip_delta_prepare_device) and each pack streams them instead.
//
// The pack is an order-preserving stream compaction in three passes over
// tiles of 4096 entries (64 rows of 64): a count pass (every row's keep
// ballot stored as a 64-bit mask, prev_sent lowered for the kept entries),
// a one-workgroup scan of the tile counts, and an emit pass in which one
// wave per tile ranks its rows from their masks and writes the kept
// entries' pairs.  No workgroup waits on another.  (A single pass with a
// decoupled look-back, knob delta_passes = 1, measured 2.7x slower.)  The
// pairs are staged in HBM and moved to the caller's buffer (the pinned
// buffer MPI sends from) by one DMA copy of exactly the emitted pairs.
// Kernel stores straight into pinned host memory measured 8 GB/s on a
// pack that emits every entry (46.4M pairs: 45 ms); the copy engine moves
// them far faster.

#include <hip/hip_runtime.h>

#include <cstring>

#include "ccx_internal.h"
#include "lookback.h"
#include "../../../include/cc_hip.h"

using ccx::kWave;

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
// Rows per wave (tiles of 4096 entries) in both pack designs.  The single
// pass measured 543 us per pack of 46.4M entries at 16 rows against 638 us
// at 32, whose extra registers leave fewer waves to cover the gathers; the
// three passes measured 183 us per late pack at 16 rows, 210 at 32 and
// 220 at 64 (the count pass runs out of waves the same way).
constexpr int kRows = 16;
// 3: the three-pass pack (count, scan, emit; default); 1: one pass with a
// decoupled look-back, 440-500 us per late pack at the 4-rank share against
// 183 us for the three passes: its 11.4K tiles queue on one tile counter
// (same-address atomics are served ~11 ns apart) and on each other's
// look-backs
int g_delta_passes = 3;

// rep_of_send[idx] for an entry that is not owned: its label is the id
// itself, read from send_to_flat.
constexpr uint32_t kNotOwned = 0xFFFFFFFFu;

struct DeltaOp {
  const uint32_t *send_to_flat;
  const uint32_t *comp_label;
  const uint32_t *comp_of;
  // comp_of[send_to_flat[idx] - v_start] gathered once (or kNotOwned), or
  // null: then each pack gathers it from comp_of
  const uint32_t *rep_of_send;
  uint32_t v_start, v_end;
  uint32_t *prev_sent;
  uint64_t *out;
  const int *seg_start;  // nseg + 1 entries, seg_start[0] = 0
  int nseg;
  unsigned long long *seg_rank;  // pairs emitted before each segment start
  // per-thread cursor over segment starts
  uint64_t cur_base;
  int next;

  __device__ void begin(uint64_t first) {
    // lower_bound of `first` in seg_start[0..nseg]
    int lo = 0, hi = nseg + 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((uint64_t)seg_start[mid] < first) lo = mid + 1;
      else hi = mid;
    }
    next = lo;
    cur_base = lo > 0 ? (uint64_t)seg_start[lo - 1] : 0;
  }
  __device__ bool eval(uint64_t j, uint64_t &v) const {
    uint32_t lbl;
    if (rep_of_send) {
      const uint32_t r = rep_of_send[j];
      lbl = r != kNotOwned ? comp_label[r] : send_to_flat[j];
    } else {
      const uint32_t x = send_to_flat[j];
      lbl = x;
      if (x >= v_start && x < v_end) lbl = comp_label[comp_of[x - v_start]];
    }
    v = lbl;
    return lbl < prev_sent[j];
  }
  // Start of the last segment starting at or before j (seg_start[0] = 0).
  __device__ uint64_t seg_base_of(uint64_t j) const {
    int lo = 0, hi = nseg + 1;  // first k with seg_start[k] > j
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((uint64_t)seg_start[mid] <= j) lo = mid + 1;
      else hi = mid;
    }
    return (uint64_t)seg_start[lo - 1];
  }
  // A kept entry of the three-pass pack (its label already in prev_sent):
  // write its pair at `rank`, relative to its segment's start.
  __device__ void emit_kept(uint64_t j, uint64_t rank, uint64_t sbase) const {
    const uint32_t v = prev_sent[j];
    out[rank] = ((j - sbase) << 32) | v;
  }
  __device__ void apply(uint64_t j, bool keep, uint64_t rank, uint64_t v) {
    while (next <= nseg && (uint64_t)seg_start[next] == j) {
      seg_rank[next] = rank;
      cur_base = j;
      ++next;
    }
    if (keep) {
      prev_sent[j] = (uint32_t)v;
      out[rank] = ((j - cur_base) << 32) | (uint32_t)v;
    }
  }
  // apply() without the cursor, for an entry reached out of order: the
  // segments starting at j (several when some are empty) record the rank,
  // and the entry belongs to the last segment starting at or before j
  // (seg_start[0] = 0, so there is one).
  __device__ void apply_at(uint64_t j, bool keep, uint64_t rank, uint32_t v) const {
    int lo = 0, hi = nseg + 1;  // first k with seg_start[k] >= j
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((uint64_t)seg_start[mid] < j) lo = mid + 1;
      else hi = mid;
    }
    int k = lo;
    while (k <= nseg && (uint64_t)seg_start[k] == j) seg_rank[k++] = rank;
    if (keep) {
      prev_sent[j] = v;
      out[rank] = ((j - (uint64_t)seg_start[k - 1]) << 32) | v;
    }
  }
};

__global__ __launch_bounds__(kThreads) void k_delta(DeltaOp op, uint64_t n, uint32_t *__restrict__ tile_ctr,
                                                    uint64_t *__restrict__ status, uint32_t epoch,
                                                    unsigned long long *__restrict__ total) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wsum[kWaves];
  __shared__ uint64_t s_base;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  // Tile ids come from a counter zeroed before the launch and run 0..tiles-1;
  // an id past the input (never expected) must not touch memory.
  if (tile * kTile >= n) return;
  const uint64_t first = tile * kTile + (uint64_t)t * kItems;
  uint32_t keep = 0;
  uint32_t val[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint64_t j = first + k;
    uint64_t v = 0;
    if (j < n && op.eval(j, v)) keep |= 1u << k;
    val[k] = (uint32_t)v;
  }
  const uint32_t cnt = (uint32_t)__popc(keep);
  const uint32_t incl = ccx::wave_incl_scan<uint32_t>(cnt, lane);
  if (lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  uint32_t wprefix = 0, tile_total = 0;
  for (int ww = 0; ww < kWaves; ++ww) {
    if (ww < w) wprefix += s_wsum[ww];
    tile_total += s_wsum[ww];
  }
  if (w == 0) {  // the whole first wave looks back
    uint64_t *my = status + tile;
    uint64_t excl = 0;
    if (tile == 0) {
      if (lane == 0) ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, tile_total));
    } else {
      if (lane == 0) ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagAggregate, tile_total));
      excl = ccx::lookback_wave(status, tile, 1, 0, epoch, lane);
      if (lane == 0)
        ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, excl + tile_total));
    }
    if (lane == 0) {
      s_base = excl;
      if ((tile + 1) * kTile >= n) *total = excl + tile_total;
    }
  }
  __syncthreads();
  uint64_t rank = s_base + wprefix + incl - cnt;
  op.begin(first);
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint64_t j = first + k;
    if (j < n) {
      const bool kf = (keep >> k) & 1u;
      op.apply(j, kf, rank, val[k]);
      rank += kf;
    }
  }
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
  return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// The same pack over row-striped tiles: row i of wave w is the 64
// consecutive entries wbase + 64 i + lane, so the reads of send_to_flat and
// prev_sent and the stores of a row's emitted pairs coalesce; ranks come
// from the rows' ballots.
template <int Rows>
__global__ __launch_bounds__(kThreads) void k_delta_rows(DeltaOp op, uint64_t n,
                                                         uint32_t *__restrict__ tile_ctr,
                                                         uint64_t *__restrict__ status, uint32_t epoch,
                                                         unsigned long long *__restrict__ total) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wsum[kWaves];
  __shared__ uint64_t s_base;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  // Tile ids come from a counter zeroed before the launch and run 0..tiles-1;
  // an id past the input (never expected) must not touch memory.
  constexpr uint64_t kTileEntries = (uint64_t)kThreads * Rows;
  if (tile * kTileEntries >= n) return;
  const uint64_t wbase = tile * kTileEntries + (uint64_t)w * (kWave * Rows);
  uint32_t val[Rows];
  uint64_t bal[Rows];
#pragma unroll
  for (int i = 0; i < Rows; ++i) {
    const uint64_t j = wbase + (uint64_t)i * kWave + lane;
    uint64_t v = 0;
    const bool keep = j < n && op.eval(j, v);
    val[i] = (uint32_t)v;
    bal[i] = __ballot(keep);
  }
  uint32_t wcnt = 0;
#pragma unroll
  for (int i = 0; i < Rows; ++i) wcnt += (uint32_t)__popcll(bal[i]);
  if (lane == 0) s_wsum[w] = wcnt;
  __syncthreads();
  uint32_t wprefix = 0, tile_total = 0;
  for (int ww = 0; ww < kWaves; ++ww) {
    if (ww < w) wprefix += s_wsum[ww];
    tile_total += s_wsum[ww];
  }
  if (w == 0) {  // the whole first wave looks back
    uint64_t *my = status + tile;
    uint64_t excl = 0;
    if (tile == 0) {
      if (lane == 0) ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, tile_total));
    } else {
      if (lane == 0) ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagAggregate, tile_total));
      excl = ccx::lookback_wave(status, tile, 1, 0, epoch, lane);
      if (lane == 0)
        ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, excl + tile_total));
    }
    if (lane == 0) {
      s_base = excl;
      if ((tile + 1) * kTileEntries >= n) *total = excl + tile_total;
    }
  }
  __syncthreads();
  const uint64_t lt = lanemask_lt(lane);
  uint64_t run = s_base + wprefix;
#pragma unroll
  for (int i = 0; i < Rows; ++i) {
    const uint64_t j = wbase + (uint64_t)i * kWave + lane;
    if (j < n) op.apply_at(j, (bal[i] >> lane) & 1ull, run + (uint64_t)__popcll(bal[i] & lt), val[i]);
    run += (uint64_t)__popcll(bal[i]);
  }
}

// ---- the pack in three passes, with no workgroup waiting on another ----
//
// 1. k_delta_count: every tile of kThreads * Rows entries evaluates its
//    entries row by row (64 consecutive entries a row), stores each row's
//    keep ballot as one 64-bit mask and the tile's count.
// 2. k_delta_scan: one workgroup scans the tile counts into offsets, and
//    ranks every segment start from the offsets and the masks.
// 3. k_delta_emit: every tile reads its rows' masks; the kept entries
//    (few, after the first exchange) recompute their label, lower
//    prev_sent and write their pair at offset + rank within the tile.
// The tiles of passes 1 and 3 are independent: no counter, no look-back.

template <int Rows>
__global__ __launch_bounds__(kThreads) void k_delta_count(DeltaOp op, uint64_t n,
                                                          uint64_t *__restrict__ mask,
                                                          uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_cnt[kWaves];
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  constexpr uint64_t kTileEntries = (uint64_t)kThreads * Rows;
  const uint64_t wbase = (uint64_t)blockIdx.x * kTileEntries + (uint64_t)w * (kWave * Rows);
  uint32_t cnt = 0;
#pragma unroll
  for (int i = 0; i < Rows; ++i) {
    const uint64_t row0 = wbase + (uint64_t)i * kWave;
    const uint64_t j = row0 + lane;
    uint64_t v = 0;
    const bool keep = j < n && op.eval(j, v);
    // the kept label is final for this exchange: prev_sent takes it now,
    // and the emit pass reads it back from there
    if (keep) op.prev_sent[j] = (uint32_t)v;
    const uint64_t b = __ballot(keep);
    if (lane == 0 && row0 < n) mask[row0 / kWave] = b;
    cnt += (uint32_t)__popcll(b);
  }
  if (lane == 0) s_cnt[w] = cnt;
  __syncthreads();
  if (t == 0) {
    uint32_t total = 0;
    for (int ww = 0; ww < kWaves; ++ww) total += s_cnt[ww];
    counts[blockIdx.x] = total;
  }
}

constexpr int kScanThreads = 1024;

// offsets[i] = sum of counts[0..i) for i <= tiles; seg_rank[k] = pairs
// emitted before entry seg_start[k] (total for a start at n).  One
// workgroup: each thread sums a run of consecutive counts, one block scan
// of the runs, then each thread writes its run's offsets.
__global__ __launch_bounds__(kScanThreads) void k_delta_scan(const uint32_t *__restrict__ counts,
                                                              uint64_t tiles, uint64_t tile_entries,
                                                              const uint64_t *__restrict__ mask,
                                                              uint64_t n, const int *__restrict__ seg,
                                                              int nseg,
                                                              uint64_t *__restrict__ offsets,
                                                              unsigned long long *__restrict__ seg_rank,
                                                              unsigned long long *__restrict__ total) {
  __shared__ uint64_t s_wsum[kScanThreads / kWave];
  __shared__ uint64_t s_all;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  const uint64_t per = (tiles + kScanThreads - 1) / kScanThreads;
  const uint64_t lo = (uint64_t)t * per;
  const uint64_t hi = lo + per < tiles ? lo + per : tiles;
  uint64_t mine = 0;
  for (uint64_t i = lo; i < hi; ++i) mine += counts[i];
  const uint64_t incl = ccx::wave_incl_scan<uint64_t>(mine, lane);
  if (lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  uint64_t run = incl - mine, sum = 0;
  for (int ww = 0; ww < kScanThreads / kWave; ++ww) {
    if (ww < w) run += s_wsum[ww];
    sum += s_wsum[ww];
  }
  for (uint64_t i = lo; i < hi; ++i) {
    offsets[i] = run;
    run += counts[i];
  }
  if (t == 0) {
    offsets[tiles] = sum;
    *total = sum;
    s_all = sum;
  }
  __syncthreads();  // offsets[] above are read by other threads of this workgroup
  const uint64_t all = s_all;
  for (int k = t; k <= nseg; k += kScanThreads) {
    const uint64_t pos = (uint64_t)seg[k];
    if (pos >= n) {
      seg_rank[k] = all;
      continue;
    }
    const uint64_t tile = pos / tile_entries;
    uint64_t r = offsets[tile];
    for (uint64_t row = tile * tile_entries / kWave; row < pos / kWave; ++row)
      r += (uint64_t)__popcll(mask[row]);
    const int bit = (int)(pos % kWave);
    if (bit) r += (uint64_t)__popcll(mask[pos / kWave] & (~0ull >> (kWave - bit)));
    seg_rank[k] = r;
  }
}

// One wave per tile, grid-stride over tiles: a tile of kThreads * Rows
// entries is 4 * Rows rows of 64, so lane i holds the masks of rows i,
// i + 64, ... (one coalesced load each), a wave scan of their popcounts
// gives every row its rank, and the wave then visits only the rows with
// kept entries.  No LDS, no barrier.
template <int Rows>
__global__ __launch_bounds__(kThreads) void k_delta_emit(DeltaOp op, uint64_t n, uint64_t tiles,
                                                         const uint64_t *__restrict__ mask,
                                                         const uint64_t *__restrict__ offsets) {
  constexpr int kChunks = (4 * Rows) / kWave;  // rows per lane
  static_assert(kChunks * kWave == 4 * Rows, "a tile holds whole chunks of 64 rows");
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t rows_total = (n + kWave - 1) / kWave;
  const uint64_t waves = (uint64_t)gridDim.x * (kThreads / kWave);
  const uint64_t lt = lanemask_lt(lane);
  for (uint64_t tile = (uint64_t)blockIdx.x * (kThreads / kWave) + threadIdx.x / kWave;
       tile < tiles; tile += waves) {
    uint64_t base = offsets[tile];
#pragma unroll
    for (int q = 0; q < kChunks; ++q) {
      const uint64_t row = tile * (uint64_t)(4 * Rows) + (uint64_t)q * kWave + lane;
      const uint64_t m = row < rows_total ? mask[row] : 0;
      const uint64_t c = (uint64_t)__popcll(m);
      const uint64_t incl = ccx::wave_incl_scan<uint64_t>(c, lane);
      const uint64_t excl = incl - c;
      uint64_t live = __ballot(m != 0);
      while (live) {
        const int r = __builtin_ctzll(live);
        live &= live - 1;
        const uint64_t mr = __shfl(m, r, kWave);
        const uint64_t rbase = base + __shfl(excl, r, kWave);
        const uint64_t row0 = (tile * (uint64_t)(4 * Rows) + (uint64_t)q * kWave + r) * kWave;
        // the segment holding the row's first entry, and where the next starts
        int lo = 0, hi = op.nseg + 1;  // first k with seg_start[k] > row0
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((uint64_t)op.seg_start[mid] <= row0) lo = mid + 1;
          else hi = mid;
        }
        const uint64_t sbase = (uint64_t)op.seg_start[lo - 1];
        const uint64_t snext = lo <= op.nseg ? (uint64_t)op.seg_start[lo] : ~0ull;
        if ((mr >> lane) & 1ull) {
          const uint64_t j = row0 + lane;
          op.emit_kept(j, rbase + (uint64_t)__popcll(mr & lt), j < snext ? sbase : op.seg_base_of(j));
        }
      }
      base += __shfl(incl, kWave - 1, kWave);
    }
  }
}

// rep_of_send[i] = comp_of[send_to_flat[i] - v_start] for owned entries,
// kNotOwned for the others.
__global__ __launch_bounds__(kThreads) void k_delta_prepare(const uint32_t *__restrict__ stf,
                                                            uint64_t n,
                                                            const uint32_t *__restrict__ comp_of,
                                                            uint32_t v_start, uint32_t v_end,
                                                            uint32_t *__restrict__ rep) {
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads) {
    const uint32_t x = stf[i];
    rep[i] = (x >= v_start && x < v_end) ? comp_of[x - v_start] : kNotOwned;
  }
}

int pack(const uint32_t *d_send_to_flat, const uint32_t *d_rep_of_send, uint64_t total_send_to,
         const int *seg_start, int nseg, const uint32_t *d_comp_label, const uint32_t *d_comp_of,
         uint32_t v_start, uint32_t v_end, uint32_t *d_prev_sent, uint64_t *h_sendbuf,
         int *sendcounts, int *sdispls, uint64_t *total_pairs, void *stream, const char *who);

}  // namespace

namespace ccx {
// Passes of the delta pack (cc_hip_sort_set key "delta_passes"): 3 or 1.
int delta_passes_set(int v) {
  if (v != 1 && v != 3) return 4;
  g_delta_passes = v;
  return 0;
}
int delta_passes_get() { return g_delta_passes; }
}  // namespace ccx

extern "C" int cc_hip_delta_prepare_device(const uint32_t *d_send_to_flat, uint64_t total_send_to,
                                           const uint32_t *d_comp_of, uint32_t v_start,
                                           uint32_t v_end, uint32_t *d_rep_of_send, void *stream) {
  const ccx::Context *c = ccx::ctx();
  if (!c) return 2;
  if (total_send_to == 0) return 0;
  if (!d_send_to_flat || !d_comp_of || !d_rep_of_send)
    return ccx::set_error(4, "cc_hip_delta_prepare_device: null pointer");
  hipStream_t s = ccx::pick_stream(stream);
  const uint64_t need = (total_send_to + kThreads - 1) / kThreads;
  const uint64_t cap = (uint64_t)(c->cu_count > 0 ? c->cu_count : 256) * 8;
  const unsigned g = (unsigned)(need < cap ? need : cap);
  CCX_LAUNCH("delta_prepare", s,
             k_delta_prepare<<<g, kThreads, 0, s>>>(d_send_to_flat, total_send_to, d_comp_of,
                                                    v_start, v_end, d_rep_of_send));
  return 0;
}

extern "C" int cc_hip_delta_pack_reps_device(const uint32_t *d_send_to_flat,
                                             const uint32_t *d_rep_of_send,
                                             uint64_t total_send_to, const int *seg_start,
                                             int nseg, const uint32_t *d_comp_label,
                                             uint32_t *d_prev_sent, uint64_t *h_sendbuf,
                                             int *sendcounts, int *sdispls, uint64_t *total_pairs,
                                             void *stream) {
  if (total_send_to > 0 && !d_rep_of_send)
    return ccx::set_error(4, "cc_hip_delta_pack_reps_device: null rep list");
  return pack(d_send_to_flat, d_rep_of_send, total_send_to, seg_start, nseg, d_comp_label,
              nullptr, 0, 0, d_prev_sent, h_sendbuf, sendcounts, sdispls, total_pairs, stream,
              "cc_hip_delta_pack_reps_device");
}

extern "C" int cc_hip_delta_pack_device(const uint32_t *d_send_to_flat, uint64_t total_send_to,
                                        const int *seg_start, int nseg,
                                        const uint32_t *d_comp_label, const uint32_t *d_comp_of,
                                        uint32_t v_start, uint32_t v_end, uint32_t *d_prev_sent,
                                        uint64_t *h_sendbuf, int *sendcounts, int *sdispls,
                                        uint64_t *total_pairs, void *stream) {
  if (total_send_to > 0 && !d_comp_of)
    return ccx::set_error(4, "cc_hip_delta_pack_device: null pointer");
  return pack(d_send_to_flat, nullptr, total_send_to, seg_start, nseg, d_comp_label, d_comp_of,
              v_start, v_end, d_prev_sent, h_sendbuf, sendcounts, sdispls, total_pairs, stream,
              "cc_hip_delta_pack_device");
}

namespace {

int pack(const uint32_t *d_send_to_flat, const uint32_t *d_rep_of_send, uint64_t total_send_to,
         const int *seg_start, int nseg, const uint32_t *d_comp_label, const uint32_t *d_comp_of,
         uint32_t v_start, uint32_t v_end, uint32_t *d_prev_sent, uint64_t *h_sendbuf,
         int *sendcounts, int *sdispls, uint64_t *total_pairs, void *stream, const char *who) {
  if (!ccx::ctx()) return 2;
  if (total_pairs) *total_pairs = 0;
  if (nseg < 0 || (nseg > 0 && (!seg_start || !sendcounts || !sdispls)))
    return ccx::set_error(4, "%s: bad segment arguments", who);
  for (int k = 0; k < nseg; ++k) {
    sendcounts[k] = 0;
    sdispls[k] = 0;
  }
  if (nseg == 0 || total_send_to == 0) return 0;
  if (seg_start[0] != 0 || (uint64_t)seg_start[nseg] != total_send_to)
    return ccx::set_error(4, "%s: seg_start must run from 0 to total_send_to", who);
  for (int k = 0; k < nseg; ++k)
    if (seg_start[k + 1] < seg_start[k])
      return ccx::set_error(4, "%s: seg_start not ascending", who);
  if (!d_send_to_flat || !d_comp_label || !d_prev_sent || !h_sendbuf)
    return ccx::set_error(4, "%s: null pointer", who);
  if ((size_t)(nseg + 1) * (sizeof(int) + sizeof(uint64_t)) > 65536)
    return ccx::set_error(4, "%s: too many segments (%d)", who, nseg);

  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  // three passes by default; the single look-back pass (row-striped or
  // runs of items) stays selectable
  const bool three = g_delta_passes == 3;
  const int rows = (three || ccx::g_compact_rows) ? kRows : 0;
  const uint64_t tile_items = rows ? (uint64_t)kThreads * rows : (uint64_t)kTile;
  const uint64_t tiles = (total_send_to + tile_items - 1) / tile_items;
  const uint64_t launches = tiles;
  if (int rc = ccx::ensure_status(st, tiles)) return rc;
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t seg_bytes = (size_t)(nseg + 1) * sizeof(int);
  const size_t rank_off = up(seg_bytes);
  const size_t stage_off = up(rank_off + (size_t)(nseg + 1) * sizeof(uint64_t));
  // three-pass scratch: one keep mask per row of 64 entries, tile counts
  // and tile offsets
  const size_t mask_off = up(stage_off + total_send_to * sizeof(uint64_t));
  const size_t counts_off = up(mask_off + ((total_send_to + kWave - 1) / kWave) * sizeof(uint64_t));
  const size_t offsets_off = up(counts_off + tiles * sizeof(uint32_t));
  const size_t tmp_bytes = three ? offsets_off + (tiles + 1) * sizeof(uint64_t)
                                 : stage_off + total_send_to * sizeof(uint64_t);
  char *tmp = (char *)ccx::stream_tmp(st, tmp_bytes);
  if (!tmp) return 3;
  int *d_seg = (int *)tmp;
  unsigned long long *d_rank = (unsigned long long *)(tmp + rank_off);
  uint64_t *d_stage = (uint64_t *)(tmp + stage_off);  // room for every entry
  uint64_t *d_mask = (uint64_t *)(tmp + mask_off);
  uint32_t *d_counts = (uint32_t *)(tmp + counts_off);
  uint64_t *d_offsets = (uint64_t *)(tmp + offsets_off);
  memcpy(st->h_stage, seg_start, seg_bytes);
  CCX_TRY(hipMemcpyAsync(d_seg, st->h_stage, seg_bytes, hipMemcpyHostToDevice, s));
  uint32_t epoch;
  if (int rc = ccx::next_epoch(st, &epoch)) return rc;
  uint32_t *ctr = st->d_words + 8;
  unsigned long long *tot = (unsigned long long *)(st->d_words + 16);
  CCX_TRY(hipMemsetAsync(ctr, 0, sizeof(uint32_t), s));
  DeltaOp op{d_send_to_flat, d_comp_label, d_comp_of, d_rep_of_send, v_start, v_end,
             d_prev_sent, d_stage, d_seg, nseg, d_rank, 0, 0};
#define CCX_DELTA_ROWS(R)                                                                          \
  CCX_LAUNCH("delta_pack", s,                                                                      \
             k_delta_rows<R><<<(unsigned)launches, kThreads, 0, s>>>(op, total_send_to, ctr,       \
                                                                      st->d_status, epoch, tot))
  if (three) {
    // one wave per tile
    const uint64_t cap = (uint64_t)(ccx::ctx()->cu_count > 0 ? ccx::ctx()->cu_count : 256) * 8;
    const uint64_t wgs = (tiles + kWaves - 1) / kWaves;
    const unsigned ge = (unsigned)(wgs < cap ? wgs : cap);
#define CCX_DELTA3(R)                                                                              \
  do {                                                                                             \
    CCX_LAUNCH("delta_pack", s,                                                                    \
               k_delta_count<R><<<(unsigned)tiles, kThreads, 0, s>>>(op, total_send_to, d_mask,    \
                                                                     d_counts));                   \
    CCX_LAUNCH("delta_scan", s,                                                                    \
               k_delta_scan<<<1, kScanThreads, 0, s>>>(d_counts, tiles, tile_items, d_mask,        \
                                                       total_send_to, d_seg, nseg, d_offsets,      \
                                                       d_rank, tot));                              \
    CCX_LAUNCH("delta_emit", s,                                                                    \
               k_delta_emit<R><<<ge, kThreads, 0, s>>>(op, total_send_to, tiles, d_mask,           \
                                                       d_offsets));                                \
  } while (0)
    CCX_DELTA3(kRows);
#undef CCX_DELTA3
  } else if (rows) CCX_DELTA_ROWS(kRows);
#undef CCX_DELTA_ROWS
  else
    CCX_LAUNCH("delta_pack", s,
               k_delta<<<(unsigned)tiles, kThreads, 0, s>>>(op, total_send_to, ctr, st->d_status,
                                                            epoch, tot));
  uint64_t *h = st->h_stage;
  CCX_TRY(hipMemcpyAsync(h, tot, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipMemcpyAsync(h + 1, d_rank, (size_t)(nseg + 1) * sizeof(uint64_t),
                         hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  const uint64_t total = h[0];
  if (total > total_send_to)
    return ccx::set_error(1, "%s: %llu pairs from %llu entries", who,
                          (unsigned long long)total, (unsigned long long)total_send_to);
  if (total > 0) {
    ccx::prof_begin("delta_copy", s);
    CCX_TRY(hipMemcpyAsync(h_sendbuf, d_stage, total * sizeof(uint64_t), hipMemcpyDefault, s));
    ccx::prof_end(s);
    CCX_TRY(hipStreamSynchronize(s));
  }
  uint64_t *rank = h + 1;
  for (int k = 0; k <= nseg; ++k)
    if ((uint64_t)seg_start[k] >= total_send_to) rank[k] = total;
  for (int k = 0; k < nseg; ++k) {
    sendcounts[k] = (int)(rank[k + 1] - rank[k]);
    sdispls[k] = (int)rank[k];
  }
  if (total_pairs) *total_pairs = total;
  return 0;
}

}  // namespace
