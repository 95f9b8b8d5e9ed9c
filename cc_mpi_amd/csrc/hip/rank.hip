// Per-rank preprocessing and merge kernels of the distributed CC driver
// (reference src/cc_mpi.c), so the driver keeps the owned CSR, the local
// labels and the boundary records resident in HBM:
//
//   * split (src/cc_mpi.c:252-304 and 364-394): one pass counts the owned
//     entries whose column is owned (local-local) and those that are not but
//     are < n_global (boundary); a second pass, an ordered dual compaction
//     over row-striped tiles, emits the local CSR (offsets, columns rebased
//     to local ids) and the boundary records (remote << 32) | li in row
//     order;
//   * rep fill: li -> comp_of[li] in each boundary record (:379, :391);
//   * merge step (:547-622): comp_min = comp_label over the reps, hook every
//     boundary record (comp_min[rep] = min(comp_min[rep], ghost_labels[gidx])),
//     then for each rep follow up to four label hops (owned vertex:
//     comp_label[comp_of[x - v_start]]; ghost: binary search of the sorted
//     ghost list; anything else stops the walk) and lower comp_label;
//   * labels + root count (:730-742): label of owned li is
//     comp_label[comp_of[li]]; li is a root when that equals v_start + li;
//   * delta unpack (src/exchange_plan.c:390-413).
//
// Every step computes a minimum or a pure function of its inputs, so the
// atomics' order never changes a result.

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
// rows per wave of the row-striped split: 32 (8192-entry tiles) measured
// 3.94 ms on 854M entries against 4.76 ms with 16
constexpr int kSplitRows = 32;
constexpr int kSplitTile = kThreads * kSplitRows;

#define GRID_STRIDE(i, n) \
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * kThreads)

unsigned grid_for(uint64_t n) {
  const uint64_t need = (n + kThreads - 1) / kThreads;
  ccx::Context *c = ccx::ctx();
  const uint64_t cap = (uint64_t)((c && c->cu_count > 0) ? c->cu_count : 256) * 8;
  return (unsigned)(need == 0 ? 1 : (need < cap ? need : cap));
}

// Exclusive scan of n uint32 counts into n+1 uint64 offsets
// (out[0] = 0, out[n] = total), decoupled look-back over 4096-item tiles.
__global__ __launch_bounds__(kThreads) void k_scan_u32(const uint32_t *__restrict__ in, uint64_t n,
                                                       uint64_t *__restrict__ out,
                                                       uint32_t *__restrict__ tile_ctr,
                                                       uint64_t *__restrict__ status, uint32_t epoch) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wsum[kWaves];
  __shared__ uint64_t s_base;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  if (tile * kTile >= n) return;
  const uint64_t first = tile * kTile + (uint64_t)t * kItems;
  uint32_t v[kItems];
  uint64_t sum = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = (first + k < n) ? in[first + k] : 0u;
    sum += v[k];
  }
  const uint64_t incl = ccx::wave_incl_scan<uint64_t>(sum, lane);
  if (lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  uint64_t wprefix = 0, tile_total = 0;
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
    if ((tile + 1) * kTile >= n) out[n] = excl + tile_total;
  }
  __syncthreads();
  uint64_t run = s_base + wprefix + incl - sum;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (first + k < n) out[first + k] = run;
    run += v[k];
  }
}

// Entries j in [e0, e1) of the owned rows (row_ptr may carry global
// offsets: e0 = row_ptr[0]).  Totals of local-local and boundary entries.
constexpr int kSplitCountItems = 16;

__global__ __launch_bounds__(kThreads) void k_split_count(const uint32_t *__restrict__ col,
                                                          uint64_t e0, uint64_t e1,
                                                          uint32_t v_start, uint32_t v_end,
                                                          uint32_t n_global,
                                                          unsigned long long *__restrict__ totals) {
  uint64_t l = 0, b = 0;
  // Each workgroup walks chunks of kSplitCountItems x kThreads consecutive
  // entries, a thread's loads kThreads apart (coalesced across the wave) and
  // all in flight together: faster than loads a whole grid apart.
  constexpr int U = kSplitCountItems;
  const uint64_t step = (uint64_t)gridDim.x * kThreads * U;
  uint64_t j = e0 + (uint64_t)blockIdx.x * kThreads * U + threadIdx.x;
  for (; j + (uint64_t)(U - 1) * kThreads < e1; j += step) {
    uint32_t v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = col[j + (uint64_t)k * kThreads];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool loc = v[k] >= v_start && v[k] < v_end;
      l += loc;
      b += (!loc && v[k] < n_global);
    }
  }
  // at most one partial chunk per thread: the next chunk starts past e1
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint64_t pos = j + (uint64_t)k * kThreads;
    if (pos < e1) {
      const uint32_t v = col[pos];
      const bool loc = v >= v_start && v < v_end;
      l += loc;
      b += (!loc && v < n_global);
    }
  }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    l += __shfl_xor(l, d, kWave);
    b += __shfl_xor(b, d, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    if (l) atomicAdd(&totals[0], (unsigned long long)l);
    if (b) atomicAdd(&totals[1], (unsigned long long)b);
  }
}

// First index r in [lo, hi) with row_ptr[r] >= x, or hi if none
// (row_ptr ascending).
__device__ __forceinline__ uint32_t rows_lower_bound(const uint64_t *row_ptr, uint32_t lo,
                                                     uint32_t hi, uint64_t x) {
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (row_ptr[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The split as one order-preserving dual stream compaction over the
// entries: each thread owns 16 consecutive entries, a 4096-entry tile
// publishes both of its counts through decoupled look-back (status words
// 2*tile and 2*tile+1), and a per-thread cursor over the row starts gives
// each row's local/boundary offsets and each boundary record's row li.
// col_idx is read once, coalesced, whatever the degree distribution.  The
// tile's local entries and records are staged in LDS at their tile-local
// ranks and written out as two contiguous runs, so the stores coalesce
// (Staged; 48 KiB of LDS, which costs occupancy), or stored straight from
// each thread's cursor.  The rows overlapping the tile are found once per
// workgroup, and each thread's search for its first row stays inside them.
template <bool Staged>
__global__ __launch_bounds__(kThreads) void k_split_emit(const uint64_t *__restrict__ row_ptr,
                                                         const uint32_t *__restrict__ col,
                                                         uint32_t n_local, uint32_t v_start,
                                                         uint32_t v_end, uint32_t n_global,
                                                         uint64_t *__restrict__ lrow,
                                                         uint64_t *__restrict__ boff,
                                                         uint32_t *__restrict__ lcol,
                                                         uint64_t *__restrict__ brec,
                                                         uint32_t *__restrict__ tile_ctr,
                                                         uint64_t *__restrict__ status, uint32_t epoch) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wsum[kWaves];
  __shared__ uint64_t s_base[2];
  __shared__ uint32_t s_rows[2];
  __shared__ uint32_t s_l[Staged ? kTile : 1];
  __shared__ uint64_t s_b[Staged ? kTile : 1];
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  const uint64_t e0 = row_ptr[0], e1 = row_ptr[n_local];
  const uint64_t n = e1 - e0;
  if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  if (tile * kTile >= n) return;
  const uint64_t tile_first = e0 + tile * kTile;
  const uint64_t tile_end = (tile_first + kTile < e1) ? tile_first + kTile : e1;
  if (t == 0) {
    // rows [r0, r1]: r0 = last row starting at or before tile_first (its
    // entries may begin the tile), r1 = first row starting at or after
    // tile_end (or n_local)
    const uint32_t a = rows_lower_bound(row_ptr, 0, n_local + 1, tile_first);
    s_rows[0] = (a <= n_local && row_ptr[a] == tile_first) ? a : a - 1;
    s_rows[1] = rows_lower_bound(row_ptr, a, n_local + 1, tile_end);
  }
  const uint64_t first = tile_first + (uint64_t)t * kItems;
  uint32_t v[kItems];
  uint32_t lmask = 0, bmask = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = (first + k < e1) ? col[first + k] : 0xFFFFFFFFu;
    const bool in = first + k < e1;
    const bool loc = in && v[k] >= v_start && v[k] < v_end;
    if (loc) lmask |= 1u << k;
    else if (in && v[k] < n_global) bmask |= 1u << k;
  }
  // both counts in one word: local low 32 bits, boundary high 32 bits
  const uint64_t mine = (uint64_t)__popc(lmask) | ((uint64_t)__popc(bmask) << 32);
  const uint64_t incl = ccx::wave_incl_scan<uint64_t>(mine, lane);
  if (lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  uint64_t wprefix = 0, tile_total = 0;
  for (int ww = 0; ww < kWaves; ++ww) {
    if (ww < w) wprefix += s_wsum[ww];
    tile_total += s_wsum[ww];
  }
  const uint32_t tile_l = (uint32_t)(tile_total & 0xFFFFFFFFull);
  const uint32_t tile_b = (uint32_t)(tile_total >> 32);
  (void)tile_l;
  (void)tile_b;
  if (t < 2) {  // thread 0: local count, thread 1: boundary count
    const uint64_t cnt = (t == 0) ? tile_l : tile_b;
    uint64_t *my = status + 2 * tile + t;
    uint64_t excl = 0;
    if (tile == 0) {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, cnt));
    } else {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagAggregate, cnt));
      excl = ccx::lookback(status, tile, 2, t, epoch);
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, excl + cnt));
    }
    s_base[t] = excl;
  }
  // tile-local ranks: stage the outputs while the look-back runs
  const uint64_t ex = wprefix + incl - mine;
  uint32_t ll = (uint32_t)(ex & 0xFFFFFFFFull);
  uint32_t bl = (uint32_t)(ex >> 32);
  const uint32_t r_lo = s_rows[0], r_hi = s_rows[1];
  uint32_t r = 0, cur = 0;
  const uint64_t stop = (first + kItems < e1) ? first + kItems : e1;
  const bool has = first < e1;
  if (has) {
    // rows starting inside [first, first + 16) get their offsets below
    r = rows_lower_bound(row_ptr, r_lo, r_hi + 1, first);
    // row of entry `first`: the last row starting at or before it
    cur = (r <= n_local && row_ptr[r] == first) ? r : r - 1;
  }
  if constexpr (Staged) {
    if (has) {
      uint32_t rr = r, cc = cur;
#pragma unroll
      for (int k = 0; k < kItems; ++k) {
        const uint64_t j = first + k;
        if (j >= stop) break;
        while (rr <= n_local && row_ptr[rr] == j) cc = rr++;
        if ((lmask >> k) & 1u) s_l[ll++] = v[k] - v_start;
        else if ((bmask >> k) & 1u) s_b[bl++] = ((uint64_t)v[k] << 32) | cc;
      }
    }
    __syncthreads();
    for (uint32_t i = t; i < tile_l; i += kThreads) lcol[s_base[0] + i] = s_l[i];
    for (uint32_t i = t; i < tile_b; i += kThreads) brec[s_base[1] + i] = s_b[i];
  } else {
    __syncthreads();
  }
  if (!has) return;
  // row offsets: global ranks of each row's first local / boundary entry
  uint64_t lr = s_base[0] + (uint32_t)(ex & 0xFFFFFFFFull);
  uint64_t br = s_base[1] + (uint32_t)(ex >> 32);
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint64_t j = first + k;
    if (j >= stop) break;
    while (r <= n_local && row_ptr[r] == j) {  // rows (possibly empty) starting at j
      lrow[r] = lr;
      boff[r] = br;
      cur = r;
      ++r;
    }
    if constexpr (!Staged) {
      if ((lmask >> k) & 1u) lcol[lr] = v[k] - v_start;
      else if ((bmask >> k) & 1u) brec[br] = ((uint64_t)v[k] << 32) | cur;
    }
    lr += (lmask >> k) & 1u;
    br += (bmask >> k) & 1u;
  }
  if (stop == e1) {  // trailing rows start at e1 (at least row n_local)
    while (r <= n_local) {
      lrow[r] = lr;
      boff[r] = br;
      ++r;
    }
  }
}

// ---- the split over row-striped tiles -------------------------------------
// Same outputs as k_split_emit, but wave w of a tile reads its entries as
// kSplitRows rows of 64 consecutive entries (coalesced), and ranks come from the
// rows' ballots, so a row's local columns and boundary records are stored
// contiguously.  An entry's row comes from marks in LDS: every row starting
// strictly inside the tile writes (row + 1) at its start position (the
// largest row wins where empty rows share a start), and a running max over
// the positions gives each entry's row.  The rows around each tile boundary
// are found by k_split_tile_rows beforehand, all tiles in parallel.

// For tile boundary b (position e0 + min(b * tile_items, n)): first[b] = the
// first row starting at or after it, own[b] = the last row starting at or
// before it (the row of the entry there).
__global__ __launch_bounds__(kThreads) void k_split_tile_rows(const uint64_t *__restrict__ row_ptr,
                                                              uint32_t n_local, uint64_t e0, uint64_t n,
                                                              uint64_t tiles, uint64_t tile_items,
                                                              uint32_t *__restrict__ first,
                                                              uint32_t *__restrict__ own) {
  GRID_STRIDE(b, tiles + 1) {
    const uint64_t pos = e0 + (b * tile_items < n ? b * tile_items : n);
    uint32_t lo = 0, hi = n_local + 1;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (row_ptr[mid] < pos) lo = mid + 1;
      else hi = mid;
    }
    first[b] = lo;
    hi = n_local + 1;
    while (lo < hi) {  // row_ptr[0] = e0 <= pos, so the upper bound is at least 1
      const uint32_t mid = lo + (hi - lo) / 2;
      if (row_ptr[mid] <= pos) lo = mid + 1;
      else hi = mid;
    }
    own[b] = lo - 1;
  }
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
  return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

__global__ __launch_bounds__(kThreads) void k_split_rows(
    const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col, uint32_t n_local,
    uint32_t v_start, uint32_t v_end, uint32_t n_global, const uint32_t *__restrict__ tfirst,
    const uint32_t *__restrict__ town, uint64_t *__restrict__ lrow, uint64_t *__restrict__ boff,
    uint32_t *__restrict__ lcol, uint64_t *__restrict__ brec, uint32_t *__restrict__ tile_ctr,
    uint64_t *__restrict__ status, uint32_t epoch) {
  constexpr int R = kSplitRows;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_mark[kSplitTile];
  __shared__ uint32_t s_wmax[kWaves];
  __shared__ uint64_t s_wsum[kWaves];
  __shared__ uint64_t s_base[2];
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  const uint64_t e0 = row_ptr[0], e1 = row_ptr[n_local];
  const uint64_t n = e1 - e0;
  if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
  for (int i = t; i < kSplitTile; i += kThreads) s_mark[i] = 0;
  __syncthreads();
  const uint64_t tile = s_tile;
  // Tile ids come from a counter zeroed before the launch and run 0..tiles-1;
  // an id past the input (never expected) must not touch memory.
  if (tile * kSplitTile >= n) return;
  const uint64_t tile_first = e0 + tile * kSplitTile;
  const uint64_t tile_end = (tile_first + kSplitTile < e1) ? tile_first + kSplitTile : e1;
  const uint32_t r_own = town[tile];    // row of the tile's first entry
  const uint32_t r_first = tfirst[tile];  // first row starting at the tile's first entry
  const uint32_t r_next = tfirst[tile + 1];  // first row starting at or after tile_end
  for (uint32_t r = r_own + 1 + t; r < r_next; r += kThreads)
    atomicMax(&s_mark[(uint32_t)(row_ptr[r] - tile_first)], r + 1);
  const uint32_t wpos = (uint32_t)w * (kWave * R);
  uint32_t v[R], mk[R];
  uint32_t lcount = 0, bcount = 0, wm = 0;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint64_t j = tile_first + wpos + (uint32_t)i * kWave + lane;
    v[i] = j < tile_end ? col[j] : 0xFFFFFFFFu;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < R; ++i) {
    mk[i] = s_mark[wpos + (uint32_t)i * kWave + lane];
    wm = mk[i] > wm ? mk[i] : wm;
    const bool loc = v[i] >= v_start && v[i] < v_end;
    const bool bnd = !loc && v[i] < n_global;  // false past the tile's end (0xFFFFFFFF)
    lcount += (uint32_t)__popcll(__ballot(loc));
    bcount += (uint32_t)__popcll(__ballot(bnd));
  }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const uint32_t o = __shfl_xor(wm, d, kWave);
    wm = o > wm ? o : wm;
  }
  if (lane == 0) {
    s_wmax[w] = wm;
    s_wsum[w] = (uint64_t)lcount | ((uint64_t)bcount << 32);
  }
  __syncthreads();
  uint32_t carry = r_own + 1;  // (row + 1) of the position before this wave's first
  uint64_t wpre = 0, ttot = 0;
  for (int ww = 0; ww < kWaves; ++ww) {
    if (ww < w) {
      carry = s_wmax[ww] > carry ? s_wmax[ww] : carry;
      wpre += s_wsum[ww];
    }
    ttot += s_wsum[ww];
  }
  if (t < 2) {  // thread 0: local count, thread 1: boundary count
    const uint64_t cnt = (t == 0) ? (uint32_t)ttot : (uint32_t)(ttot >> 32);
    uint64_t *my = status + 2 * tile + t;
    uint64_t excl = 0;
    if (tile == 0) {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, cnt));
    } else {
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagAggregate, cnt));
      excl = ccx::lookback(status, tile, 2, t, epoch);
      ccx::status_store(my, ccx::tag_word(epoch, ccx::kFlagInclusive, excl + cnt));
    }
    s_base[t] = excl;
  }
  __syncthreads();
  const uint64_t lt = lanemask_lt(lane);
  uint64_t lr = s_base[0] + (uint32_t)wpre, br = s_base[1] + (uint32_t)(wpre >> 32);
  // row of the position before this wave's first (the tile's first position
  // instead owns the rows starting there: r_first .. r_own)
  uint32_t prev_last = (w == 0) ? r_first - 1 : carry - 1;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    // The running max of the marks up to this lane.  Marks grow along the
    // entries (row + 1 at each row start), so it is the mark of the nearest
    // marked lane at or below this one: one ballot and one lane permute
    // instead of a six-step shuffle scan (4.19 -> 3.82 ms for 910M entries).
    const uint64_t upto = __ballot(mk[i] != 0) & ((2ull << lane) - 1);
    const int src = upto ? 63 - __builtin_clzll(upto) : lane;
    const uint32_t near = __shfl(mk[i], src, kWave);
    uint32_t m = upto ? near : 0u;
    m = carry > m ? carry : m;
    const uint32_t row = m - 1;
    uint32_t prow = __shfl_up(row, 1, kWave);
    if (lane == 0) prow = prev_last;
    const bool loc = v[i] >= v_start && v[i] < v_end;
    const bool bnd = !loc && v[i] < n_global;
    const uint64_t bl = __ballot(loc), bb = __ballot(bnd);
    const uint64_t lex = lr + (uint64_t)__popcll(bl & lt), bex = br + (uint64_t)__popcll(bb & lt);
    const uint64_t j = tile_first + wpos + (uint32_t)i * kWave + lane;
    if (j < tile_end) {
      for (uint32_t r = prow + 1; r <= row; ++r) {  // rows starting at j (several if empty)
        lrow[r] = lex;
        boff[r] = bex;
      }
      if (loc) lcol[lex] = v[i] - v_start;
      else if (bnd) brec[bex] = ((uint64_t)v[i] << 32) | row;
    }
    carry = __shfl(m, kWave - 1, kWave);
    prev_last = carry - 1;
    lr += (uint64_t)__popcll(bl);
    br += (uint64_t)__popcll(bb);
  }
  if (tile_end == e1) {  // trailing rows start at e1 (at least row n_local)
    const uint64_t tl = s_base[0] + (uint32_t)ttot, tb = s_base[1] + (uint32_t)(ttot >> 32);
    for (uint32_t r = r_next + t; r <= n_local; r += kThreads) {
      lrow[r] = tl;
      boff[r] = tb;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_set_rep(uint64_t *__restrict__ rec, uint64_t nb,
                                                      const uint32_t *__restrict__ comp_of) {
  GRID_STRIDE(i, nb) {
    const uint64_t r = rec[i];
    rec[i] = (r & 0xFFFFFFFF00000000ull) | comp_of[(uint32_t)r];
  }
}

__global__ __launch_bounds__(kThreads) void k_fill_u64(uint64_t *__restrict__ a, uint64_t *__restrict__ b,
                                                       uint64_t n) {
  GRID_STRIDE(i, n) {
    a[i] = 0;
    b[i] = 0;
  }
}

__global__ __launch_bounds__(kThreads) void k_iota(uint32_t *__restrict__ a, uint32_t n, uint32_t base) {
  GRID_STRIDE(i, n) a[i] = base + (uint32_t)i;
}

// comp_min[r] = comp_label[r] for every rep r: the rep_count listed in
// reps, or, with reps null, every r with comp_of[r] == r.
__global__ __launch_bounds__(kThreads) void k_merge_init(const uint32_t *__restrict__ comp_of,
                                                         const uint32_t *__restrict__ comp_label,
                                                         uint32_t *__restrict__ comp_min,
                                                         uint32_t n_local,
                                                         const uint32_t *__restrict__ reps,
                                                         uint32_t rep_count) {
  if (reps) {
    GRID_STRIDE(k, rep_count) {
      const uint32_t r = reps[k];
      comp_min[r] = comp_label[r];
    }
    return;
  }
  GRID_STRIDE(i, n_local) {
    if (comp_of[i] == (uint32_t)i) comp_min[i] = comp_label[i];
  }
}

// comp_min[rep] = min(comp_min[rep], ghost_labels[gidx]) over the records
// (gidx << 32) | rep.
__global__ __launch_bounds__(kThreads) void k_merge_hook(const uint64_t *__restrict__ rec, uint64_t nb,
                                                         const uint32_t *__restrict__ ghost_labels,
                                                         uint32_t *__restrict__ comp_min) {
  GRID_STRIDE(i, nb) {
    const uint64_t r = rec[i];
    const uint32_t rep = (uint32_t)r;
    const uint32_t lbl = ghost_labels[r >> 32];
    if (lbl < comp_min[rep]) atomicMin(&comp_min[rep], lbl);
  }
}

// ghost_min[k] = comp_label[reps[k]] for every rep, before the hooks lower
// it by the ghost labels (ghost_min is indexed by position in the rep
// list).  Starting from the label rather than from ~0 lets the hooks'
// plain read skip the atomic for nearly every record, and leaves the
// round's result unchanged: comp_label only decreases, so
// min(comp_label now, ghost_min) is the same either way.
__global__ __launch_bounds__(kThreads) void k_ghost_min_init(uint32_t *__restrict__ ghost_min,
                                                             const uint32_t *__restrict__ comp_label,
                                                             const uint32_t *__restrict__ reps,
                                                             uint32_t rep_count) {
  GRID_STRIDE(k, rep_count) ghost_min[k] = comp_label[reps[k]];
}

// ghost_min[dominant] lowered to the minimum ghost label over the set bits
// of the dominant rep's bitmap over the ghosts: one streaming pass over the
// bitmap and ghost_labels, where the record list costs 8 bytes per record
// plus the label.  Reduced per workgroup, one atomic each: all of them
// target one address, which serves atomics about 11 ns apart.
__global__ __launch_bounds__(kThreads) void k_merge_hook_dom(const uint32_t *__restrict__ bits,
                                                             uint32_t ghost_count,
                                                             const uint32_t *__restrict__ ghost_labels,
                                                             uint32_t dominant,
                                                             uint32_t *__restrict__ ghost_min) {
  constexpr int kIlp = 4;  // loads in flight per thread
  __shared__ uint32_t s_min[kThreads / kWave];
  uint32_t m = 0xFFFFFFFFu;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  uint64_t g = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  for (; g + (kIlp - 1) * stride < ghost_count; g += kIlp * stride) {
    uint32_t w[kIlp], l[kIlp];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      const uint64_t x = g + u * stride;
      w[u] = bits[x >> 5];
      l[u] = ghost_labels[x];
    }
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      const uint64_t x = g + u * stride;
      const uint32_t v = ((w[u] >> (x & 31)) & 1u) ? l[u] : 0xFFFFFFFFu;
      m = v < m ? v : m;
    }
  }
  for (; g < ghost_count; g += stride) {
    if ((bits[g >> 5] >> (g & 31)) & 1u) {
      const uint32_t l = ghost_labels[g];
      m = l < m ? l : m;
    }
  }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const uint32_t o = __shfl_xor(m, d, kWave);
    m = o < m ? o : m;
  }
  if ((threadIdx.x & (kWave - 1)) == 0) s_min[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / kWave; ++w) m = s_min[w] < m ? s_min[w] : m;
    if (m < ghost_min[dominant]) atomicMin(&ghost_min[dominant], m);
  }
}

// The side records' hook of a prepared round: ghost_min[k] lowered by
// every record (gidx << 32) | k, k the rep's index in the rep list (a
// compact array, where indexing by local id spread the reps' entries over
// a 4n-byte array: 125 us for 5.7M records at one rank's share).
// record -> label -> ghost_min is a chain of three dependent loads, so each
// thread takes four records per step with every load of the step issued
// before the first use.  Atomics are rare (ghost_min only moves when a
// label drops below it), so they go lane by lane.
__global__ __launch_bounds__(kThreads) void k_merge_hook_side(const uint64_t *__restrict__ rec,
                                                              uint64_t nb,
                                                              const uint32_t *__restrict__ ghost_labels,
                                                              uint32_t *__restrict__ ghost_min) {
  constexpr int kIlp = 4;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i0 = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i0 < nb; i0 += kIlp * stride) {
    uint64_t r[kIlp];
    uint32_t l[kIlp], g[kIlp];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      const uint64_t i = i0 + u * stride;
      r[u] = i < nb ? rec[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kIlp; ++u) l[u] = i0 + u * stride < nb ? ghost_labels[r[u] >> 32] : 0xFFFFFFFFu;
#pragma unroll
    for (int u = 0; u < kIlp; ++u) g[u] = i0 + u * stride < nb ? ghost_min[(uint32_t)r[u]] : 0u;
#pragma unroll
    for (int u = 0; u < kIlp; ++u)
      if (l[u] < g[u]) atomicMin(&ghost_min[(uint32_t)r[u]], l[u]);
  }
}

// Half-width of the window an interpolation probe brackets.
constexpr uint32_t kFindWindow = 4096;

// Position of x in the sorted unique list a[0..count), or count if absent.
// The ghosts of a permuted graph are spread evenly over their id range, so
// one interpolation probe brackets x in a window of 2 kFindWindow entries
// (its ends checked, both loads in flight together) and a binary search
// runs there: about 15 dependent loads instead of 26 for 47M ghosts.  When
// the window misses x (uneven ids), the search covers the whole list, so
// the result is exact for any sorted list.
__device__ __forceinline__ uint32_t find_sorted(const uint32_t *a, uint32_t count, uint32_t x) {
  uint32_t lo = 0, hi = count;
  if (count > 4 * kFindWindow) {
    const uint32_t first = a[0], last = a[count - 1];
    if (x < first || x > last) return count;
    // last > first: the list is unique and longer than one entry
    const uint64_t est = (uint64_t)(x - first) * (count - 1) / (uint64_t)(last - first);
    const uint32_t l = est > kFindWindow ? (uint32_t)est - kFindWindow : 0u;
    const uint32_t h = est + kFindWindow < (uint64_t)(count - 1) ? (uint32_t)est + kFindWindow
                                                                  : count - 1;
    if (a[l] <= x && x <= a[h]) {
      lo = l;
      hi = h + 1;
    }
  }
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (lo < count && a[lo] == x) ? lo : count;
}

__device__ __forceinline__ bool label_of(uint32_t x, const uint32_t *comp_label,
                                         const uint32_t *comp_of, uint32_t v_start, uint32_t v_end,
                                         const uint32_t *ghosts, uint32_t ghost_count,
                                         const uint32_t *ghost_labels, uint32_t &y) {
  if (x >= v_start && x < v_end) {
    y = comp_label[comp_of[x - v_start]];
    return true;
  }
  const uint32_t g = find_sorted(ghosts, ghost_count, x);
  if (g < ghost_count) {
    y = ghost_labels[g];
    return true;
  }
  return false;
}

// Label hops and the comp_label update of rep r (src/cc_mpi.c:588-621);
// true if comp_label[r] dropped.
__device__ __forceinline__ bool merge_update_rep(uint32_t r, const uint32_t *comp_of,
                                                 uint32_t *comp_label, uint32_t *comp_min,
                                                 uint32_t v_start, uint32_t v_end,
                                                 const uint32_t *ghosts, uint32_t ghost_count,
                                                 const uint32_t *ghost_labels,
                                                 const uint32_t *ghost_min) {
  if (ghost_min) {  // the init of this round, fused: comp_min = min(comp_label, this rep's ghost_min)
    const uint32_t l = comp_label[r], g = *ghost_min;
    comp_min[r] = g < l ? g : l;
  }
  uint32_t x = comp_min[r];
  for (int t = 0; t < 4; ++t) {
    uint32_t y;
    if (!label_of(x, comp_label, comp_of, v_start, v_end, ghosts, ghost_count, ghost_labels, y))
      break;
    if (y >= x) break;
    x = y;
  }
  if (x < comp_min[r]) comp_min[r] = x;
  if (comp_min[r] < comp_label[r]) {
    comp_label[r] = comp_min[r];
    return true;
  }
  return false;
}

// The update for every rep: the rep_count listed in reps, or, with reps
// null, every r with comp_of[r] == r.
__global__ __launch_bounds__(kThreads) void k_merge_update(const uint32_t *__restrict__ comp_of,
                                                           uint32_t *__restrict__ comp_label,
                                                           uint32_t *__restrict__ comp_min,
                                                           uint32_t n_local,
                                                           const uint32_t *__restrict__ reps,
                                                           uint32_t rep_count, uint32_t v_start,
                                                           uint32_t v_end,
                                                           const uint32_t *__restrict__ ghosts,
                                                           uint32_t ghost_count,
                                                           const uint32_t *__restrict__ ghost_labels,
                                                           uint32_t *__restrict__ changed,
                                                           const uint32_t *__restrict__ ghost_min) {
  bool any = false;
  if (reps) {
    GRID_STRIDE(k, rep_count) {
      any |= merge_update_rep(reps[k], comp_of, comp_label, comp_min, v_start, v_end, ghosts,
                              ghost_count, ghost_labels, ghost_min ? ghost_min + k : nullptr);
    }
  } else {
    GRID_STRIDE(i, n_local) {
      if (comp_of[i] != (uint32_t)i) continue;
      any |= merge_update_rep((uint32_t)i, comp_of, comp_label, comp_min, v_start, v_end, ghosts,
                              ghost_count, ghost_labels, nullptr);
    }
  }
  if (__ballot(any) != 0 && (threadIdx.x & (kWave - 1)) == 0) atomicOr(changed, 1u);
}

__global__ __launch_bounds__(kThreads) void k_rank_labels(const uint32_t *__restrict__ comp_label,
                                                          const uint32_t *__restrict__ comp_of,
                                                          uint32_t n_local, uint32_t v_start,
                                                          uint32_t *__restrict__ labels,
                                                          unsigned long long *__restrict__ roots) {
  uint32_t mine = 0;
  GRID_STRIDE(i, n_local) {
    const uint32_t l = comp_label[comp_of[i]];
    if (labels) labels[i] = l;
    mine += (l == v_start + (uint32_t)i);
  }
  for (int d = kWave / 2; d > 0; d >>= 1) mine += __shfl_xor(mine, d, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0 && mine) atomicAdd(roots, (unsigned long long)mine);
}

// Received pairs (pos << 32) | lbl, grouped per source k at rdispls[k];
// ghost_labels[need_from_gidx_flat[need_base[k] + pos]] = lbl.
__global__ __launch_bounds__(kThreads) void k_delta_unpack(const uint64_t *__restrict__ pairs,
                                                           uint64_t total,
                                                           const int *__restrict__ rdispls,
                                                           const int *__restrict__ need_base,
                                                           int nsrc,
                                                           const uint32_t *__restrict__ gidx_flat,
                                                           uint32_t *__restrict__ ghost_labels) {
  GRID_STRIDE(i, total) {
    // last source whose displacement is <= i
    int lo = 0, hi = nsrc - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((uint64_t)rdispls[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const uint64_t pr = pairs[i];
    ghost_labels[gidx_flat[need_base[lo] + (int)(pr >> 32)]] = (uint32_t)pr;
  }
}

// key = (owner(remote) << 32) | li for raw records (remote << 32) | li;
// owner by binary search of bounds[0..P].
__global__ __launch_bounds__(kThreads) void k_owner_keys(const uint64_t *__restrict__ rec, uint64_t nb,
                                                         const uint32_t *__restrict__ bounds, int P,
                                                         uint64_t *__restrict__ keys) {
  GRID_STRIDE(i, nb) {
    const uint64_t r = rec[i];
    const uint32_t v = (uint32_t)(r >> 32);
    int lo = 0, hi = P - 1;
    while (lo < hi) {  // last q with bounds[q] <= v
      const int mid = (lo + hi + 1) >> 1;
      if (bounds[mid] <= v) lo = mid;
      else hi = mid - 1;
    }
    keys[i] = ((uint64_t)lo << 32) | (uint32_t)r;
  }
}

// Segments of sorted unique keys (q << 32) | x: starts[q] = first index
// with high word >= q (q in [0, P]); vals[i] = low word + base.
__global__ __launch_bounds__(kThreads) void k_segments(const uint64_t *__restrict__ keys, uint64_t m,
                                                       int P, uint32_t base,
                                                       int *__restrict__ starts,
                                                       uint32_t *__restrict__ vals) {
  GRID_STRIDE(i, m) {
    const uint64_t q = keys[i] >> 32;
    const uint64_t prev = (i == 0) ? 0 : (keys[i - 1] >> 32) + 1;
    for (uint64_t r = prev; r <= q && r <= (uint64_t)P; ++r) starts[r] = (int)i;
    vals[i] = (uint32_t)keys[i] + base;
    if (i == m - 1)
      for (uint64_t r = q + 1; r <= (uint64_t)P; ++r) starts[r] = (int)m;
  }
}

// out[i] = table[idx[i]].
__global__ __launch_bounds__(kThreads) void k_gather_u32(const uint32_t *__restrict__ table,
                                                         const uint32_t *__restrict__ idx, uint64_t n,
                                                         uint32_t *__restrict__ out) {
  GRID_STRIDE(i, n) out[i] = table[idx[i]];
}

}  // namespace

extern "C" int cc_hip_owner_keys_device(const uint64_t *d_records, uint64_t count,
                                        const uint32_t *d_bounds, int comm_size,
                                        uint64_t *d_keys, void *stream) {
  if (!ccx::ctx()) return 2;
  if (count == 0) return 0;
  if (comm_size < 1) return ccx::set_error(4, "cc_hip_owner_keys_device: comm_size < 1");
  hipStream_t s = ccx::pick_stream(stream);
  CCX_LAUNCH("owner_keys", s,
             k_owner_keys<<<grid_for(count), kThreads, 0, s>>>(d_records, count, d_bounds, comm_size,
                                                               d_keys));
  return 0;
}

extern "C" int cc_hip_segments_device(const uint64_t *d_keys, uint64_t count, int nseg,
                                      uint32_t base, int *d_starts, uint32_t *d_vals,
                                      void *stream) {
  if (!ccx::ctx()) return 2;
  hipStream_t s = ccx::pick_stream(stream);
  if (count == 0) {
    CCX_TRY(hipMemsetAsync(d_starts, 0, (size_t)(nseg + 1) * sizeof(int), s));
    return 0;
  }
  if (count > 0x7FFFFFFFull) return ccx::set_error(4, "cc_hip_segments_device: count exceeds int");
  CCX_LAUNCH("segments", s,
             k_segments<<<grid_for(count), kThreads, 0, s>>>(d_keys, count, nseg, base, d_starts,
                                                             d_vals));
  return 0;
}

extern "C" int cc_hip_gather_u32_device(const uint32_t *d_table, const uint32_t *d_idx,
                                        uint64_t count, uint32_t *d_out, void *stream) {
  if (!ccx::ctx()) return 2;
  if (count == 0) return 0;
  hipStream_t s = ccx::pick_stream(stream);
  CCX_LAUNCH("gather", s, k_gather_u32<<<grid_for(count), kThreads, 0, s>>>(d_table, d_idx, count, d_out));
  return 0;
}

extern "C" int cc_hip_exclusive_scan_u32_device(const uint32_t *d_in, uint64_t n, uint64_t *d_out,
                                                void *stream) {
  if (!ccx::ctx()) return 2;
  hipStream_t s = ccx::pick_stream(stream);
  if (n == 0) {
    CCX_TRY(hipMemsetAsync(d_out, 0, sizeof(uint64_t), s));
    return 0;
  }
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  const uint64_t tiles = (n + kTile - 1) / kTile;
  if (int rc = ccx::ensure_status(st, tiles)) return rc;
  uint32_t epoch;
  if (int rc = ccx::next_epoch(st, &epoch)) return rc;
  uint32_t *ctr = st->d_words + 9;
  CCX_TRY(hipMemsetAsync(ctr, 0, sizeof(uint32_t), s));
  CCX_LAUNCH("scan_u32", s,
             k_scan_u32<<<(unsigned)tiles, kThreads, 0, s>>>(d_in, n, d_out, ctr, st->d_status, epoch));
  return 0;
}

extern "C" int cc_hip_split_count_device(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                                         uint32_t n_local, uint32_t v_start, uint32_t v_end,
                                         uint32_t n_global, uint64_t *m_local_local,
                                         uint64_t *n_boundary, void *stream) {
  if (!ccx::ctx()) return 2;
  *m_local_local = 0;
  *n_boundary = 0;
  if (n_local == 0) return 0;
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  uint64_t *h = st->h_stage;
  CCX_TRY(hipMemcpyAsync(h, d_row_ptr, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipMemcpyAsync(h + 1, d_row_ptr + n_local, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  const uint64_t e0 = h[0], e1 = h[1];
  if (e1 <= e0) return 0;
  unsigned long long *tot = (unsigned long long *)(st->d_words + 28);
  CCX_TRY(hipMemsetAsync(tot, 0, 2 * sizeof(uint64_t), s));
  CCX_LAUNCH("split_count", s,
             k_split_count<<<grid_for((e1 - e0 + kSplitCountItems - 1) / kSplitCountItems),
                             kThreads, 0, s>>>(d_col_idx, e0, e1, v_start, v_end, n_global, tot));
  CCX_TRY(hipMemcpyAsync(h, tot, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  *m_local_local = h[0];
  *n_boundary = h[1];
  return 0;
}

extern "C" int cc_hip_split_emit_device(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                                        uint32_t n_local, uint32_t v_start, uint32_t v_end,
                                        uint32_t n_global, uint64_t *d_lrow, uint64_t *d_boff,
                                        uint32_t *d_lcol, uint64_t *d_brec, void *stream) {
  if (!ccx::ctx()) return 2;
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  uint64_t *h = st->h_stage;
  CCX_TRY(hipMemcpyAsync(h, d_row_ptr, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipMemcpyAsync(h + 1, d_row_ptr + n_local, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  const uint64_t e0 = h[0], n = h[1] - h[0];
  if (n_local == 0 || n == 0) {
    // no entries: every offset is 0
    CCX_LAUNCH("split_zero", s, k_fill_u64<<<grid_for((uint64_t)n_local + 1), kThreads, 0, s>>>(d_lrow, d_boff, (uint64_t)n_local + 1));
    return 0;
  }
  const bool rows = ccx::g_compact_rows && !ccx::g_split_staged;
  const uint64_t tile_items = rows ? kSplitTile : kTile;
  const uint64_t tiles = (n + tile_items - 1) / tile_items;
  if (tiles > 0xFFFFFFFFull) return ccx::set_error(4, "cc_hip_split_emit_device: input too large");
  if (int rc = ccx::ensure_status(st, 2 * tiles)) return rc;
  uint32_t epoch;
  if (int rc = ccx::next_epoch(st, &epoch)) return rc;
  uint32_t *ctr = st->d_words + 10;
  CCX_TRY(hipMemsetAsync(ctr, 0, sizeof(uint32_t), s));
  if (rows) {
    // the rows around every tile boundary, then the row-striped split
    uint32_t *tb = (uint32_t *)ccx::stream_tmp(st, 2 * (tiles + 1) * sizeof(uint32_t));
    if (!tb) return 3;
    CCX_LAUNCH("split_tiles", s,
               k_split_tile_rows<<<grid_for(tiles + 1), kThreads, 0, s>>>(d_row_ptr, n_local, e0, n,
                                                                         tiles, kSplitTile, tb,
                                                                         tb + tiles + 1));
    CCX_LAUNCH("split_emit", s,
               k_split_rows<<<(unsigned)tiles, kThreads, 0, s>>>(
                   d_row_ptr, d_col_idx, n_local, v_start, v_end, n_global, tb, tb + tiles + 1,
                   d_lrow, d_boff, d_lcol, d_brec, ctr, st->d_status, epoch));
    return 0;
  }
  if (ccx::g_split_staged)
    CCX_LAUNCH("split_emit", s,
               k_split_emit<true><<<(unsigned)tiles, kThreads, 0, s>>>(
                   d_row_ptr, d_col_idx, n_local, v_start, v_end, n_global, d_lrow, d_boff, d_lcol,
                   d_brec, ctr, st->d_status, epoch));
  else
    CCX_LAUNCH("split_emit", s,
               k_split_emit<false><<<(unsigned)tiles, kThreads, 0, s>>>(
                   d_row_ptr, d_col_idx, n_local, v_start, v_end, n_global, d_lrow, d_boff, d_lcol,
                   d_brec, ctr, st->d_status, epoch));
  return 0;
}

extern "C" int cc_hip_boundary_set_rep_device(uint64_t *d_records, uint64_t count,
                                              const uint32_t *d_comp_of, void *stream) {
  if (!ccx::ctx()) return 2;
  if (count == 0) return 0;
  hipStream_t s = ccx::pick_stream(stream);
  CCX_LAUNCH("boundary_set_rep", s,
             k_set_rep<<<grid_for(count), kThreads, 0, s>>>(d_records, count, d_comp_of));
  return 0;
}

extern "C" int cc_hip_iota_u32_device(uint32_t *d_out, uint32_t n, uint32_t base, void *stream) {
  if (!ccx::ctx()) return 2;
  if (n == 0) return 0;
  hipStream_t s = ccx::pick_stream(stream);
  CCX_LAUNCH("iota", s, k_iota<<<grid_for(n), kThreads, 0, s>>>(d_out, n, base));
  return 0;
}

namespace {
// The dominant rep's part of a prepared merge round (null bits: none).
struct DomPart {
  const uint32_t *bits;
  uint32_t dominant;
  uint32_t *ghost_min;
  int ghosts_changed;
};

// One merge round; reps null: the reps are found by scanning comp_of.  With
// dom.bits the records are the side records of cc_hip_merge_prepare_device
// and the per-rep ghost minimum lives in dom.ghost_min, recomputed only when
// the ghost labels changed; the update then starts each rep from
// min(comp_label, ghost_min), which is what the init and hook give.
int merge_step(const uint64_t *d_records, uint64_t count, const uint32_t *d_ghost_vertices,
               const uint32_t *d_ghost_labels, uint32_t ghost_count, const uint32_t *d_comp_of,
               uint32_t *d_comp_label, uint32_t *d_comp_min, uint32_t n_local,
               const uint32_t *d_reps, uint32_t rep_count, uint32_t v_start, uint32_t v_end,
               int *changed, void *stream, const DomPart &dom = DomPart{nullptr, 0, nullptr, 0}) {
  if (!ccx::ctx()) return 2;
  if (changed) *changed = 0;
  if (n_local == 0) return 0;
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  uint32_t *d_changed = st->d_words + 24;
  CCX_TRY(hipMemsetAsync(d_changed, 0, sizeof(uint32_t), s));
  const unsigned g = grid_for(d_reps ? rep_count : n_local);
  if (dom.bits) {
    // 2: the first round (ghost_min seeded from comp_label); 1: ghost labels
    // lowered since the last round, so the hooks lower ghost_min in place:
    // min over the records of the lowered labels is at most the old one,
    // and only records whose label dropped below it pass the hooks' filter
    if (dom.ghosts_changed >= 2)
      CCX_LAUNCH("merge_ghost_min_init", s,
                 k_ghost_min_init<<<g, kThreads, 0, s>>>(dom.ghost_min, d_comp_label, d_reps,
                                                         rep_count));
    if (dom.ghosts_changed) {
      if (ghost_count > 0)
        CCX_LAUNCH("merge_hook_dom", s,
                   k_merge_hook_dom<<<grid_for(ghost_count), kThreads, 0, s>>>(
                       dom.bits, ghost_count, d_ghost_labels, dom.dominant, dom.ghost_min));
      if (count > 0)
        CCX_LAUNCH("merge_hook_side", s,
                   k_merge_hook_side<<<grid_for(count), kThreads, 0, s>>>(d_records, count,
                                                                          d_ghost_labels,
                                                                          dom.ghost_min));
    }
    // the round's init is fused into the update (comp_min from ghost_min)
  } else {
    CCX_LAUNCH("merge_init", s,
               k_merge_init<<<g, kThreads, 0, s>>>(d_comp_of, d_comp_label, d_comp_min, n_local,
                                                   d_reps, rep_count));
    if (count > 0)
      CCX_LAUNCH("merge_hook", s,
                 k_merge_hook<<<grid_for(count), kThreads, 0, s>>>(d_records, count, d_ghost_labels,
                                                                   d_comp_min));
  }
  CCX_LAUNCH("merge_update", s,
             k_merge_update<<<g, kThreads, 0, s>>>(d_comp_of, d_comp_label, d_comp_min, n_local,
                                                   d_reps, rep_count, v_start, v_end,
                                                   d_ghost_vertices, ghost_count, d_ghost_labels,
                                                   d_changed, dom.bits ? dom.ghost_min : nullptr));
  uint32_t *h = (uint32_t *)st->h_stage;
  CCX_TRY(hipMemcpyAsync(h, d_changed, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  if (changed) *changed = (int)h[0];
  return 0;
}
}  // namespace

extern "C" int cc_hip_merge_step_device(const uint64_t *d_records, uint64_t count,
                                        const uint32_t *d_ghost_vertices,
                                        const uint32_t *d_ghost_labels, uint32_t ghost_count,
                                        const uint32_t *d_comp_of, uint32_t *d_comp_label,
                                        uint32_t *d_comp_min, uint32_t n_local, uint32_t v_start,
                                        uint32_t v_end, int *changed, void *stream) {
  return merge_step(d_records, count, d_ghost_vertices, d_ghost_labels, ghost_count, d_comp_of,
                    d_comp_label, d_comp_min, n_local, nullptr, 0, v_start, v_end, changed, stream);
}

extern "C" int cc_hip_merge_step_reps_device(const uint64_t *d_records, uint64_t count,
                                             const uint32_t *d_ghost_vertices,
                                             const uint32_t *d_ghost_labels, uint32_t ghost_count,
                                             const uint32_t *d_comp_of, uint32_t *d_comp_label,
                                             uint32_t *d_comp_min, uint32_t n_local,
                                             const uint32_t *d_reps, uint32_t rep_count,
                                             uint32_t v_start, uint32_t v_end, int *changed,
                                             void *stream) {
  if (n_local > 0 && (!d_reps || rep_count == 0 || rep_count > n_local))
    return ccx::set_error(4, "cc_hip_merge_step_reps_device: rep list missing or larger than n_local");
  return merge_step(d_records, count, d_ghost_vertices, d_ghost_labels, ghost_count, d_comp_of,
                    d_comp_label, d_comp_min, n_local, d_reps, rep_count, v_start, v_end, changed,
                    stream);
}

extern "C" int cc_hip_merge_step_prepared_device(
    const uint32_t *d_dom_bits, uint32_t dominant, const uint64_t *d_side, uint64_t side_count,
    const uint32_t *d_ghost_vertices, const uint32_t *d_ghost_labels, uint32_t ghost_count,
    const uint32_t *d_comp_of, uint32_t *d_comp_label, uint32_t *d_comp_min,
    uint32_t *d_ghost_min, uint32_t n_local, const uint32_t *d_reps, uint32_t rep_count,
    uint32_t v_start, uint32_t v_end, int ghosts_changed, int *changed, void *stream) {
  if (n_local > 0 && (!d_reps || rep_count == 0 || rep_count > n_local))
    return ccx::set_error(4, "cc_hip_merge_step_prepared_device: rep list missing or larger than n_local");
  if (n_local > 0 && (!d_ghost_min || (ghost_count > 0 && !d_dom_bits) ||
                      (side_count > 0 && !d_side)))
    return ccx::set_error(4, "cc_hip_merge_step_prepared_device: null pointer");
  if (n_local > 0 && dominant >= rep_count)
    return ccx::set_error(4, "cc_hip_merge_step_prepared_device: dominant index %u not below rep_count %u",
                          dominant, rep_count);
  // a valid address for the dominant part even without ghosts (no bit is read then)
  const uint32_t *bits = d_dom_bits ? d_dom_bits : d_ghost_min;
  return merge_step(d_side, side_count, d_ghost_vertices, d_ghost_labels, ghost_count, d_comp_of,
                    d_comp_label, d_comp_min, n_local, d_reps, rep_count, v_start, v_end, changed,
                    stream, DomPart{bits, dominant, d_ghost_min, ghosts_changed});
}

extern "C" int cc_hip_rank_labels_device(const uint32_t *d_comp_label, const uint32_t *d_comp_of,
                                         uint32_t n_local, uint32_t v_start, uint32_t *d_labels,
                                         uint64_t *roots, void *stream) {
  if (!ccx::ctx()) return 2;
  if (roots) *roots = 0;
  if (n_local == 0) return 0;
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  unsigned long long *d_roots = (unsigned long long *)(st->d_words + 26);
  CCX_TRY(hipMemsetAsync(d_roots, 0, sizeof(uint64_t), s));
  CCX_LAUNCH("rank_labels", s,
             k_rank_labels<<<grid_for(n_local), kThreads, 0, s>>>(d_comp_label, d_comp_of, n_local,
                                                                  v_start, d_labels, d_roots));
  CCX_TRY(hipMemcpyAsync(st->h_stage, d_roots, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  if (roots) *roots = st->h_stage[0];
  return 0;
}

extern "C" int cc_hip_delta_unpack_device(const uint64_t *h_pairs, uint64_t total,
                                          const int *rdispls, const int *need_base, int nsrc,
                                          const uint32_t *d_need_from_gidx_flat,
                                          uint32_t *d_ghost_labels, void *stream) {
  if (!ccx::ctx()) return 2;
  if (total == 0 || nsrc <= 0) return 0;
  if ((size_t)nsrc * 2 * sizeof(int) > 65536)
    return ccx::set_error(4, "cc_hip_delta_unpack_device: too many sources (%d)", nsrc);
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  int *d_meta = (int *)ccx::stream_tmp(st, (size_t)nsrc * 2 * sizeof(int));
  if (!d_meta) return 3;
  int *h = (int *)st->h_stage;
  memcpy(h, rdispls, (size_t)nsrc * sizeof(int));
  memcpy(h + nsrc, need_base, (size_t)nsrc * sizeof(int));
  CCX_TRY(hipMemcpyAsync(d_meta, h, (size_t)nsrc * 2 * sizeof(int), hipMemcpyHostToDevice, s));
  CCX_LAUNCH("delta_unpack", s,
             k_delta_unpack<<<grid_for(total), kThreads, 0, s>>>(h_pairs, total, d_meta,
                                                                 d_meta + nsrc, nsrc,
                                                                 d_need_from_gidx_flat,
                                                                 d_ghost_labels));
  CCX_TRY(hipStreamSynchronize(s));
  return 0;
}
