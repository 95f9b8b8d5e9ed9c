// Changed-ghost-label delta pack: the compute half of the reference's
// exchangeplan_exchange_delta (src/exchange_plan.c:325-357).
//
// For every entry idx of send_to_flat (the owned vertices each neighbour
// rank asked for, concatenated per destination in dests order):
//   lbl = comp_label[comp_of[v - v_start]] if v is owned, else v;
//   if lbl < prev_sent[idx]: prev_sent[idx] = lbl and emit
//     ((uint64_t)(idx - segment_start) << 32) | lbl
// in ascending idx order, i.e. exactly the reference's send buffer, plus
// per-destination counts and displacements.  On the device this is one
// order-preserving stream compaction (decoupled look-back over row-striped
// tiles, so every row of 64 entries is read in one coalesced access); the pairs are
// staged in HBM and moved to the caller's buffer (the pinned buffer MPI
// sends from) by one DMA copy of exactly the emitted pairs.  Kernel stores
// straight into pinned host memory measured 8 GB/s on a pack that emits
// every entry (46.4M pairs: 45 ms); the copy engine moves them far faster.

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
// rows per wave in k_delta_rows: 16 (4096-entry tiles) measured 543 us per
// pack of 46.4M entries against 638 us with 32 rows, whose extra registers
// leave fewer waves to cover the label gathers
constexpr int kRows = 16;
constexpr int kTileRows = kThreads * kRows;

struct DeltaOp {
  const uint32_t *send_to_flat;
  const uint32_t *comp_label;
  const uint32_t *comp_of;
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
    const uint32_t x = send_to_flat[j];
    uint32_t lbl = x;
    if (x >= v_start && x < v_end) lbl = comp_label[comp_of[x - v_start]];
    v = lbl;
    return lbl < prev_sent[j];
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
  if (t == 0) {
    uint64_t *my = status + tile;
    uint64_t excl = 0;
    if (tile == 0) {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, tile_total));
    } else {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagAggregate, tile_total));
      excl = ccx::lookback(status, tile, 1, 0, epoch);
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, excl + tile_total));
    }
    s_base = excl;
    if ((tile + 1) * kTile >= n) *total = excl + tile_total;
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
__global__ __launch_bounds__(kThreads) void k_delta_rows(DeltaOp op, uint64_t n,
                                                         uint32_t *__restrict__ tile_ctr,
                                                         uint64_t *__restrict__ status, uint32_t epoch,
                                                         unsigned long long *__restrict__ total) {
  constexpr int Rows = kRows;
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
  if (tile * kTileRows >= n) return;
  const uint64_t wbase = tile * kTileRows + (uint64_t)w * (kWave * Rows);
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
  if (t == 0) {
    uint64_t *my = status + tile;
    uint64_t excl = 0;
    if (tile == 0) {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, tile_total));
    } else {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagAggregate, tile_total));
      excl = ccx::lookback(status, tile, 1, 0, epoch);
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, excl + tile_total));
    }
    s_base = excl;
    if ((tile + 1) * kTileRows >= n) *total = excl + tile_total;
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

}  // namespace

extern "C" int cc_hip_delta_pack_device(const uint32_t *d_send_to_flat, uint64_t total_send_to,
                                        const int *seg_start, int nseg,
                                        const uint32_t *d_comp_label, const uint32_t *d_comp_of,
                                        uint32_t v_start, uint32_t v_end, uint32_t *d_prev_sent,
                                        uint64_t *h_sendbuf, int *sendcounts, int *sdispls,
                                        uint64_t *total_pairs, void *stream) {
  if (!ccx::ctx()) return 2;
  if (total_pairs) *total_pairs = 0;
  if (nseg < 0 || (nseg > 0 && (!seg_start || !sendcounts || !sdispls)))
    return ccx::set_error(4, "cc_hip_delta_pack_device: bad segment arguments");
  for (int k = 0; k < nseg; ++k) {
    sendcounts[k] = 0;
    sdispls[k] = 0;
  }
  if (nseg == 0 || total_send_to == 0) return 0;
  if (seg_start[0] != 0 || (uint64_t)seg_start[nseg] != total_send_to)
    return ccx::set_error(4, "cc_hip_delta_pack_device: seg_start must run from 0 to total_send_to");
  for (int k = 0; k < nseg; ++k)
    if (seg_start[k + 1] < seg_start[k])
      return ccx::set_error(4, "cc_hip_delta_pack_device: seg_start not ascending");
  if (!d_send_to_flat || !d_comp_label || !d_comp_of || !d_prev_sent || !h_sendbuf)
    return ccx::set_error(4, "cc_hip_delta_pack_device: null pointer");
  if ((size_t)(nseg + 1) * (sizeof(int) + sizeof(uint64_t)) > 65536)
    return ccx::set_error(4, "cc_hip_delta_pack_device: too many segments (%d)", nseg);

  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  const uint64_t tile_items = ccx::g_compact_rows ? kTileRows : kTile;
  const uint64_t tiles = (total_send_to + tile_items - 1) / tile_items;
  if (int rc = ccx::ensure_status(st, tiles)) return rc;
  const size_t seg_bytes = (size_t)(nseg + 1) * sizeof(int);
  const size_t rank_off = (seg_bytes + 255) & ~(size_t)255;
  const size_t stage_off = (rank_off + (size_t)(nseg + 1) * sizeof(uint64_t) + 255) & ~(size_t)255;
  char *tmp = (char *)ccx::stream_tmp(st, stage_off + total_send_to * sizeof(uint64_t));
  if (!tmp) return 3;
  int *d_seg = (int *)tmp;
  unsigned long long *d_rank = (unsigned long long *)(tmp + rank_off);
  uint64_t *d_stage = (uint64_t *)(tmp + stage_off);  // room for every entry
  memcpy(st->h_stage, seg_start, seg_bytes);
  CCX_TRY(hipMemcpyAsync(d_seg, st->h_stage, seg_bytes, hipMemcpyHostToDevice, s));
  uint32_t epoch;
  if (int rc = ccx::next_epoch(st, &epoch)) return rc;
  uint32_t *ctr = st->d_words + 8;
  unsigned long long *tot = (unsigned long long *)(st->d_words + 16);
  CCX_TRY(hipMemsetAsync(ctr, 0, sizeof(uint32_t), s));
  DeltaOp op{d_send_to_flat, d_comp_label, d_comp_of, v_start, v_end, d_prev_sent, d_stage,
             d_seg, nseg, d_rank, 0, 0};
  if (ccx::g_compact_rows)
    CCX_LAUNCH("delta_pack", s,
               k_delta_rows<<<(unsigned)tiles, kThreads, 0, s>>>(op, total_send_to, ctr,
                                                                 st->d_status, epoch, tot));
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
    return ccx::set_error(1, "cc_hip_delta_pack_device: %llu pairs from %llu entries",
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
