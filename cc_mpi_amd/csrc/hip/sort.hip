// Boundary-record sort and the ordered-compaction kernels behind
// include/boundary_sort.h (reference src/boundary_sort.c:5-119) and the
// ghost extraction of src/cc_mpi.c:431-457.
//
// Sort: the reference sorts BoundaryEdge records by the 64-bit key
// (remote << 32) | rep with an in-place MSD American-flag radix sort.  The
// order is total on the key and the key is the whole record, so any correct
// sort gives the same bytes.  Here it is an LSD radix sort on the raw
// little-endian records (8-bit digits), one "onesweep" pass per digit:
//   * one pass over the keys builds all eight 256-bin histograms (wave64
//     ballot matching, so skewed digits do not serialise LDS atomics) and
//     the AND/OR of the keys; digits that are constant over the input get
//     no scatter pass;
//   * each digit pass: a workgroup takes the next 8192-key tile (atomic tile
//     counter; 512 threads, 16 keys each), ranks its keys stably with wave64 ballot matching into
//     per-wave LDS counters, publishes its 256 digit counts, resolves the
//     global offsets by decoupled look-back over the preceding tiles, and
//     scatters through LDS so that each digit's run of the tile is written
//     to HBM contiguously.
//
// Compaction: an order-preserving stream compaction (decoupled look-back
// over 8192-item tiles whose waves read 64 consecutive items per row, ranks
// from the rows' ballots) drives dedup (adjacent-unique), ghost extraction
// and the send lists.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "ccx_internal.h"
#include "lookback.h"
#include "../../../include/cc_hip.h"
#include "../../../include/boundary_sort.h"

using ccx::kWave;

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // 4096 keys
constexpr int kBins = 256;

// Lanes (among `active`) whose 8-bit digit equals this lane's.  Only the
// digit bits in `vary` can differ between keys; the rest are skipped.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active, uint32_t vary) {
  uint64_t peers = active;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    if (!((vary >> b) & 1u)) continue;
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  return peers;
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
  return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// A value that is equal in every lane of the wave, as a scalar.
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// AND and OR of all keys: a bit that is equal in both is constant.  Each
// thread keeps 16 coalesced loads in flight per chunk (one load at a time
// held the pass to 4 TB/s).
constexpr int kAndOrItems = 16;

__global__ __launch_bounds__(kThreads) void k_and_or(const uint64_t *__restrict__ keys, uint64_t n,
                                                     unsigned long long *__restrict__ out) {
  uint64_t a = ~0ull, o = 0;
  const uint64_t step = (uint64_t)gridDim.x * kThreads * kAndOrItems;
  uint64_t i = (uint64_t)blockIdx.x * kThreads * kAndOrItems + threadIdx.x;
  for (; i + (uint64_t)(kAndOrItems - 1) * kThreads < n; i += step) {
    uint64_t k[kAndOrItems];
#pragma unroll
    for (int j = 0; j < kAndOrItems; ++j) k[j] = keys[i + (uint64_t)j * kThreads];
#pragma unroll
    for (int j = 0; j < kAndOrItems; ++j) {
      a &= k[j];
      o |= k[j];
    }
  }
  // at most one partial chunk per thread: the next chunk starts past n
#pragma unroll
  for (int j = 0; j < kAndOrItems; ++j) {
    const uint64_t pos = i + (uint64_t)j * kThreads;
    if (pos < n) {
      a &= keys[pos];
      o |= keys[pos];
    }
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    a &= __shfl_xor(a, d, kWave);
    o |= __shfl_xor(o, d, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicAnd(&out[0], (unsigned long long)a);
    atomicOr(&out[1], (unsigned long long)o);
  }
}

struct DigitPlan {
  uint32_t vary[8];  // varying bits of each digit (0 = constant digit)
};

// 256-bin histograms of all eight digits and the AND/OR of all keys, one
// pass over the keys.  Each workgroup walks 4096-key tiles (16 keys per
// thread in flight, loaded like the scatter passes load them) and counts
// into per-wave LDS histograms.  A wave first reduces the AND and OR of its
// 1024 keys: a digit constant over them is counted with one add; a digit
// with at most 3 varying bits has at most 8 values, so the 64 lanes would
// collide heavily on its bins: it is counted by ballot matching (one add
// per group of equal digits); the others add per lane, except that the
// lanes sharing lane 0's digit add once.
constexpr int kHistWaveBins = 8 * kBins;

__global__ __launch_bounds__(kThreads) void k_histograms(const uint64_t *__restrict__ keys, uint64_t n,
                                                         unsigned long long *__restrict__ hist,
                                                         unsigned long long *__restrict__ and_or) {
  __shared__ uint32_t h[kWaves][kHistWaveBins];
  for (int i = threadIdx.x; i < kWaves * kHistWaveBins; i += kThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt(lane);
  uint32_t *hw = h[w];
  uint64_t all_and = ~0ull, all_or = 0;  // wave-uniform
  const uint64_t tiles = (n + kTile - 1) / kTile;
  for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const uint64_t wbase = tile * kTile + (uint64_t)w * (kWave * kItems);
    if (wbase >= n) continue;  // wave-uniform: this wave's keys are all past n
    const uint64_t base = wbase + lane;
    uint64_t k[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint64_t pos = base + (uint64_t)i * kWave;
      k[i] = pos < n ? keys[pos] : 0;
    }
    uint64_t a = ~0ull, o = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if (base + (uint64_t)i * kWave < n) {
        a &= k[i];
        o |= k[i];
      }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
      a &= __shfl_xor(a, d, kWave);
      o |= __shfl_xor(o, d, kWave);
    }
    // equal in every lane after the butterfly: keep them in scalar registers
    // so the per-digit branches below are uniform
    a = uniform64(a);
    o = uniform64(o);
    all_and &= a;
    all_or |= o;
    const uint64_t left = n - wbase;
    const uint32_t nvalid = left < (uint64_t)(kWave * kItems) ? (uint32_t)left : kWave * kItems;
    const uint64_t diff = a ^ o;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const uint32_t vary = (uint32_t)(diff >> (8 * d)) & 0xFFu;
      if (vary == 0) {
        if (lane == 0) atomicAdd(&hw[d * kBins + ((uint32_t)(a >> (8 * d)) & 0xFFu)], nvalid);
      } else if (__popc(vary) <= 3) {
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
          const bool valid = base + (uint64_t)i * kWave < n;
          const uint32_t dig = (uint32_t)(k[i] >> (8 * d)) & 0xFFu;
          const uint64_t peers = match_digit(dig, __ballot(valid), vary);
          if (valid && (peers & lt) == 0) atomicAdd(&hw[d * kBins + dig], (uint32_t)__popcll(peers));
        }
      } else {
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
          // Wide digits can still be skewed (one value for most keys, e.g.
          // the giant tree's rep): lanes equal to lane 0's digit are counted
          // by lane 0 with one add, the rest add one each.  Valid lanes are
          // a prefix of the wave, so lane 0 is valid whenever any lane is.
          // (one atomic instruction: lane 0's digit is `first`)
          const bool valid = base + (uint64_t)i * kWave < n;
          const uint32_t dig = (uint32_t)(k[i] >> (8 * d)) & 0xFFu;
          const uint32_t first = __shfl(dig, 0, kWave);
          const bool same = valid && dig == first;
          const uint64_t bal = __ballot(same);
          if (valid && (lane == 0 || !same))
            atomicAdd(&hw[d * kBins + dig], lane == 0 ? (uint32_t)__popcll(bal) : 1u);
        }
      }
    }
  }
  if (lane == 0 && all_or != 0) atomicOr(&and_or[1], (unsigned long long)all_or);
  if (lane == 0 && all_and != ~0ull) atomicAnd(&and_or[0], (unsigned long long)all_and);
  __syncthreads();
  for (int i = threadIdx.x; i < kHistWaveBins; i += kThreads) {
    uint32_t c = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) c += h[ww][i];
    if (c) atomicAdd(&hist[i], (unsigned long long)c);
  }
}

// goff[d][b] = exclusive scan over b of hist[d][b]: workgroup d scans
// digit d's 256 bins (one per thread).
__global__ __launch_bounds__(kBins) void k_digit_offsets(const unsigned long long *__restrict__ hist,
                                                         uint64_t *__restrict__ goff) {
  __shared__ uint64_t s_wave[kBins / kWave];
  const int b = threadIdx.x, lane = b & (kWave - 1), w = b / kWave;
  const uint64_t v = hist[blockIdx.x * kBins + b];
  uint64_t inc = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t u = __shfl_up(inc, d, kWave);
    if (lane >= d) inc += u;
  }
  if (lane == kWave - 1) s_wave[w] = inc;
  __syncthreads();
  uint64_t before = 0;
  for (int ww = 0; ww < w; ++ww) before += s_wave[ww];
  goff[blockIdx.x * kBins + b] = before + inc - v;
}

// One LSD pass on digit `shift/8`: stable scatter of in[] into out[].
// goff[d] = first output position of digit d (exclusive scan of the global
// histogram).  tile_ctr hands out tiles in order; status holds per-tile,
// per-digit look-back words (kBins per tile).
template <int Items, int Threads = kThreads>
__global__ __launch_bounds__(Threads) void k_onesweep(const uint64_t *__restrict__ in,
                                                      uint64_t *__restrict__ out, uint64_t n,
                                                      int shift, uint32_t vary,
                                                      const uint64_t *__restrict__ goff,
                                                      uint32_t *__restrict__ tile_ctr,
                                                      uint64_t *__restrict__ status, uint32_t epoch) {
  static_assert(Threads >= kBins && Threads % kWave == 0, "a thread per digit");
  constexpr int TileK = Threads * Items;
  constexpr int Waves = Threads / kWave;
  __shared__ uint64_t s_keys[TileK];
  __shared__ uint32_t s_wcnt[Waves][kBins];  // per-wave digit counts, then exclusive offsets
  __shared__ uint32_t s_dstart[kBins];       // tile-local start of each digit
  __shared__ uint64_t s_gbase[kBins];        // global position of tile-local index 0 of digit d
  __shared__ uint32_t s_wsum[kBins / kWave];
  __shared__ uint32_t s_tile;

  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  const bool owner = t < kBins;  // threads 0..255 each own one digit
  if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
  for (int i = t; i < Waves * kBins; i += Threads) (&s_wcnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t tile = s_tile;
  // Tile ids come from a counter zeroed before the launch and run 0..tiles-1;
  // an id past the input (never expected) must not touch memory.
  if (tile * TileK >= n) return;
  const uint64_t tile_base = tile * TileK;
  const uint64_t lt = lanemask_lt(lane);

  uint64_t key[Items];
  uint32_t rank[Items];
#pragma unroll
  for (int i = 0; i < Items; ++i) {
    const uint64_t pos = tile_base + (uint64_t)w * (kWave * Items) + (uint64_t)i * kWave + lane;
    // keys are streamed once per pass: non-temporal loads and stores leave
    // the caches to the look-back status words (268M keys: 11.05 -> 10.90 ms)
    key[i] = (pos < n) ? __builtin_nontemporal_load(&in[pos]) : ~0ull;
  }
#pragma unroll
  for (int i = 0; i < Items; ++i) {
    const uint64_t pos = tile_base + (uint64_t)w * (kWave * Items) + (uint64_t)i * kWave + lane;
    const bool valid = pos < n;
    const uint32_t d = (uint32_t)(key[i] >> shift) & 0xFFu;
    const uint64_t active = __ballot(valid);
    const uint64_t peers = match_digit(d, active, vary);
    uint32_t old = 0;
    if (valid) old = s_wcnt[w][d];
    __builtin_amdgcn_wave_barrier();
    // the highest peer bumps the wave's counter after every peer has read it
    if (valid && (peers >> lane) == 1ull) s_wcnt[w][d] = old + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rank[i] = old + (uint32_t)__popcll(peers & lt);
  }
  __syncthreads();

  // Thread t < 256 owns digit t: per-wave exclusive offsets and the tile count.
  uint32_t cnt = 0;
  uint64_t *my_status = status + tile * kBins + (owner ? t : 0);
  if (owner) {
#pragma unroll
    for (int ww = 0; ww < Waves; ++ww) {
      const uint32_t c = s_wcnt[ww][t];
      s_wcnt[ww][t] = cnt;
      cnt += c;
    }
    // publish this tile's count for digit t as early as possible
    if (tile == 0)
      ccx::status_store(my_status, ccx::tag_word(epoch, ccx::kFlagInclusive, cnt));
    else
      ccx::status_store(my_status, ccx::tag_word(epoch, ccx::kFlagAggregate, cnt));
  }

  // tile-local exclusive scan over digits (the owner waves)
  const uint32_t incl = ccx::wave_incl_scan<uint32_t>(cnt, lane);
  if (owner && lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  if (owner) {
    uint32_t wprefix = 0;
    for (int ww = 0; ww < w; ++ww) wprefix += s_wsum[ww];
    const uint32_t dstart = wprefix + incl - cnt;
    s_dstart[t] = dstart;
    uint64_t excl = 0;
    if (tile > 0) {
      excl = ccx::lookback(status, tile, kBins, t, epoch);
      ccx::status_store(my_status, ccx::tag_word(epoch, ccx::kFlagInclusive, excl + cnt));
    }
    s_gbase[t] = goff[t] + excl - dstart;
  }
  __syncthreads();

  // Stable scatter into LDS in digit order.
#pragma unroll
  for (int i = 0; i < Items; ++i) {
    const uint64_t pos = tile_base + (uint64_t)w * (kWave * Items) + (uint64_t)i * kWave + lane;
    if (pos < n) {
      const uint32_t d = (uint32_t)(key[i] >> shift) & 0xFFu;
      s_keys[s_dstart[d] + s_wcnt[w][d] + rank[i]] = key[i];
    }
  }
  __syncthreads();
  const uint64_t valid_in_tile = (n - tile_base) < (uint64_t)TileK ? (n - tile_base) : TileK;
  for (uint32_t j = t; j < valid_in_tile; j += Threads) {
    const uint64_t k = s_keys[j];
    const uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
    __builtin_nontemporal_store(k, &out[s_gbase[d] + j]);
  }
}

// ---- ordered compaction --------------------------------------------------

// Op interface: bool eval(uint64_t j, uint64_t &val) decides whether item j
// is kept (val carries data to apply).  A staged op (kStaged) writes only
// kept values: store(rank, val) is called for each, from a contiguous run
// staged in LDS.  Otherwise void apply(uint64_t j, bool keep, uint64_t
// rank, uint64_t val) is called for every item in order, rank = number of
// kept items before j, after begin(first_item) once per thread.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_compact(Op op, uint64_t n, uint32_t *__restrict__ tile_ctr,
                                                      uint64_t *__restrict__ status, uint32_t epoch,
                                                      unsigned long long *__restrict__ total) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wsum[kWaves];
  __shared__ uint64_t s_base;
  constexpr int Items = Op::kItemsPerThread;
  constexpr int TileK = kThreads * Items;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  // Tile ids come from a counter zeroed before the launch and run 0..tiles-1;
  // an id past the input (never expected) must not touch memory.
  if (tile * TileK >= n) return;
  const uint64_t first = tile * TileK + (uint64_t)t * Items;
  uint32_t keep = 0;
  uint64_t val[Items];
#pragma unroll
  for (int k = 0; k < Items; ++k) {
    const uint64_t j = first + k;
    val[k] = 0;
    if (j < n && op.eval(j, val[k])) keep |= 1u << k;
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
    if ((tile + 1) * TileK >= n) *total = excl + tile_total;
  }
  if constexpr (Op::kStaged) {
    // kept values go to LDS at their tile-local rank, then out as one
    // contiguous run (coalesced stores)
    __shared__ uint64_t s_out[TileK];
    uint32_t lr = wprefix + incl - cnt;
#pragma unroll
    for (int k = 0; k < Items; ++k)
      if ((keep >> k) & 1u) s_out[lr++] = val[k];
    __syncthreads();
    const uint64_t base = s_base;
    for (uint32_t i = t; i < tile_total; i += kThreads) op.store(base + i, s_out[i]);
  } else {
    __syncthreads();
    uint64_t rank = s_base + wprefix + incl - cnt;
    op.begin(first);
#pragma unroll
    for (int k = 0; k < Items; ++k) {
      const uint64_t j = first + k;
      if (j < n) {
        const bool kf = (keep >> k) & 1u;
        op.apply(j, kf, rank, val[k]);
        rank += kf;
      }
    }
  }
}

// Row-striped ordered compaction.  A tile is kThreads x Rows items (Rows = 32:
// 8192-item tiles measured 1.46x faster than 4096 on a 268M-key dedup); row
// i of wave w holds the 64 consecutive items wbase + 64 i + lane, so each row is
// one coalesced load, and the kept items of a row get consecutive ranks, so
// their stores coalesce too (in k_compact each thread owns a run of
// consecutive items, which makes every load instruction touch 64 separate
// cache lines).  An item's rank is the tile's base, the wave's prefix, the
// counts of the rows before it and the kept lanes below it in its row, all
// from the rows' ballots.  Op provides Val, kRows,
//   rows(wbase, lane, n, val[], bal[]): load the rows, ballot the kept items,
//   apply(j, keep, rank, val): called for every item j < n.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_compact_rows(Op op, uint64_t n,
                                                           uint32_t *__restrict__ tile_ctr,
                                                           uint64_t *__restrict__ status, uint32_t epoch,
                                                           unsigned long long *__restrict__ total) {
  constexpr int Rows = Op::kRows;
  constexpr int TileK = kThreads * Rows;
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
  if (tile * TileK >= n) return;
  const uint64_t wbase = tile * TileK + (uint64_t)w * (kWave * Rows);
  typename Op::Val val[Rows];
  uint64_t bal[Rows];
  op.template rows<Rows>(wbase, lane, n, val, bal);
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
    if ((tile + 1) * TileK >= n) *total = excl + tile_total;
  }
  __syncthreads();
  const uint64_t lt = lanemask_lt(lane);
  uint64_t run = s_base + wprefix;
#pragma unroll
  for (int i = 0; i < Rows; ++i) {
    const uint64_t j = wbase + (uint64_t)i * kWave + lane;
    if (j < n) op.apply(j, (bal[i] >> lane) & 1ull, run + (uint64_t)__popcll(bal[i] & lt), val[i]);
    run += (uint64_t)__popcll(bal[i]);
  }
}

// Rows of 64-bit items with the keep flag "key differs from the item
// before": the item before is the lane below in the same row, or the last
// lane of the row before; the one before the wave's first row is read once.
template <int Rows, class KeyOf>
__device__ __forceinline__ void adjacent_rows(const uint64_t *__restrict__ in, uint64_t wbase, int lane,
                                              uint64_t n, uint64_t (&v)[Rows], uint64_t (&bal)[Rows],
                                              KeyOf key) {
#pragma unroll
  for (int i = 0; i < Rows; ++i) {
    const uint64_t j = wbase + (uint64_t)i * kWave + lane;
    v[i] = j < n ? in[j] : 0;
  }
  const uint64_t before = (wbase > 0 && wbase - 1 < n) ? in[wbase - 1] : 0;
#pragma unroll
  for (int i = 0; i < Rows; ++i) {
    const uint64_t j = wbase + (uint64_t)i * kWave + lane;
    uint64_t prev = __shfl_up(v[i], 1, kWave);
    const uint64_t last = (i == 0) ? before : __shfl(v[i > 0 ? i - 1 : 0], kWave - 1, kWave);
    if (lane == 0) prev = last;
    bal[i] = __ballot(j < n && (j == 0 || key(v[i]) != key(prev)));
  }
}

struct KeyAll {
  __device__ uint64_t operator()(uint64_t v) const { return v; }
};
struct KeyHigh {
  __device__ uint64_t operator()(uint64_t v) const { return v >> 32; }
};

// Adjacent-unique, row-striped.
struct DedupRows {
  using Val = uint64_t;
  static constexpr int kRows = 32;
  const uint64_t *in;
  uint64_t *out;
  template <int Rows>
  __device__ void rows(uint64_t wbase, int lane, uint64_t n, uint64_t (&v)[Rows],
                       uint64_t (&bal)[Rows]) const {
    adjacent_rows<Rows>(in, wbase, lane, n, v, bal, KeyAll{});
  }
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) out[rank] = v;
  }
};

// Adjacent-unique (reference src/boundary_sort.c:100-119).
template <bool Staged, int Items = kItems>
struct DedupOp {
  static constexpr bool kStaged = Staged;
  static constexpr int kItemsPerThread = Items;
  const uint64_t *in;
  uint64_t *out;
  __device__ bool eval(uint64_t j, uint64_t &v) const {
    v = in[j];
    return j == 0 || v != in[j - 1];
  }
  __device__ void store(uint64_t rank, uint64_t v) const { out[rank] = v; }
  __device__ void begin(uint64_t) {}
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) out[rank] = v;
  }
};

int run_dedup(const uint64_t *in, uint64_t *out, uint64_t n, hipStream_t s, ccx::StreamState *st,
              uint64_t *count_out);

// Ghost extraction over sorted unique records (reference src/cc_mpi.c:431-457):
// ghosts[g] = g-th distinct remote; each record's remote becomes its ghost
// index g.  Records are (remote << 32) | rep.
struct GhostOp {
  static constexpr bool kStaged = false;
  static constexpr int kItemsPerThread = kItems;
  const uint64_t *in;
  uint64_t *out;
  uint32_t *ghosts;
  __device__ void begin(uint64_t) {}
  __device__ bool eval(uint64_t j, uint64_t &v) const {
    v = in[j];
    return j == 0 || (uint32_t)(v >> 32) != (uint32_t)(in[j - 1] >> 32);
  }
  __device__ void apply(uint64_t j, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) ghosts[rank] = (uint32_t)(v >> 32);
    const uint64_t g = rank + (keep ? 1 : 0) - 1;
    out[j] = (g << 32) | (uint32_t)v;
  }
};

// Ghost extraction, row-striped (same output as GhostOp).
struct GhostRows {
  using Val = uint64_t;
  static constexpr int kRows = 32;
  const uint64_t *in;
  uint64_t *out;
  uint32_t *ghosts;
  template <int Rows>
  __device__ void rows(uint64_t wbase, int lane, uint64_t n, uint64_t (&v)[Rows],
                       uint64_t (&bal)[Rows]) const {
    adjacent_rows<Rows>(in, wbase, lane, n, v, bal, KeyHigh{});
  }
  __device__ void apply(uint64_t j, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) ghosts[rank] = (uint32_t)(v >> 32);
    const uint64_t g = rank + (keep ? 1 : 0) - 1;
    out[j] = (g << 32) | (uint32_t)v;
  }
};

// ---- send lists without sorting ---------------------------------------------
// For the exchange plan, rank q must receive every owned vertex li that has
// a boundary record whose remote end q owns.  mask[li] collects those q as
// bits: records are read coalesced, each window of 64 records OR-reduces
// the bits of each run of equal li (records of one row are contiguous, so a
// row costs one atomic per window it spans), and the lane ending each run
// ORs the result into mask[li].
constexpr int kMaskWindows = 8;

__global__ __launch_bounds__(kThreads) void k_owner_mask(const uint64_t *__restrict__ rec, uint64_t nb,
                                                         const uint32_t *__restrict__ bounds, int P,
                                                         uint32_t n_local,
                                                         unsigned long long *__restrict__ mask) {
  __shared__ uint32_t s_bounds[64];
  if (threadIdx.x < P) s_bounds[threadIdx.x] = bounds[threadIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwaves = (uint64_t)gridDim.x * kWaves;
  // A wave takes kMaskWindows consecutive windows of 64 records per step
  // and issues all their loads before using any (one window at a time kept
  // too few bytes in flight: 2 TB/s).  The loop bound is uniform across the
  // wave, so every lane takes part in the shuffles.
  for (uint64_t c = ((uint64_t)blockIdx.x * kThreads + threadIdx.x) / kWave;
       c * (kWave * kMaskWindows) < nb; c += nwaves) {
    uint64_t r[kMaskWindows];
#pragma unroll
    for (int u = 0; u < kMaskWindows; ++u) {
      const uint64_t i = (c * kMaskWindows + u) * kWave + lane;
      r[u] = i < nb ? rec[i] : ~0ull;
    }
#pragma unroll
    for (int u = 0; u < kMaskWindows; ++u) {
      const uint64_t i = (c * kMaskWindows + u) * kWave + lane;
      const bool valid = i < nb;
      uint32_t li = 0xFFFFFFFFu;
      int q = 0;
      if (valid) {
        const uint32_t v = (uint32_t)(r[u] >> 32);
        li = (uint32_t)r[u];
        int hi = P - 1;
        while (q < hi) {  // last q with bounds[q] <= v
          const int mid = (q + hi + 1) >> 1;
          if (s_bounds[mid] <= v) q = mid;
          else hi = mid - 1;
        }
      }
      // Runs of equal li: a run starts where li differs from the lane below.
      // This lane's run is [first start at or below it, lane]; its OR of
      // owner bits takes one ballot per owner present in the window (valid
      // lanes are a prefix, so lane 0 starts a run whenever any lane is valid).
      const uint32_t prev = __shfl_up(li, 1, kWave);
      const uint64_t starts = __ballot(valid && (lane == 0 || prev != li));
      const uint64_t upto = lane == kWave - 1 ? ~0ull : (2ull << lane) - 1;
      const uint64_t below = starts & upto;
      const int first = below ? 63 - __clzll((long long)below) : 0;
      const uint64_t run = upto & (~0ull << first);
      unsigned long long m = 0;
      uint64_t pending = __ballot(valid);
      while (pending) {
        const int src = __ffsll((long long)pending) - 1;
        const int owner = __builtin_amdgcn_readlane(q, src);
        const uint64_t same = __ballot(valid && q == owner);
        pending &= ~same;
        if (same & run) m |= 1ull << owner;
      }
      const bool next_starts = lane == kWave - 1 || ((starts >> (lane + 1)) & 1ull) || i + 1 >= nb;
      if (valid && next_starts && li < n_local) atomicOr(&mask[li], m);
    }
  }
}

// Ordered compaction of {j : bit q of mask[j]} into vals[] as j + base.
struct MaskOp {
  static constexpr bool kStaged = true;
  static constexpr int kItemsPerThread = kItems;
  const unsigned long long *mask;
  uint32_t *vals;
  uint32_t base;
  int q;
  __device__ bool eval(uint64_t j, uint64_t &v) const {
    v = (uint32_t)j + base;
    return (mask[j] >> q) & 1ull;
  }
  __device__ void store(uint64_t rank, uint64_t v) const { vals[rank] = (uint32_t)v; }
};

// The same compaction, row-striped.
struct MaskRows {
  using Val = uint32_t;
  static constexpr int kRows = 32;
  const unsigned long long *mask;
  uint32_t *vals;
  uint32_t base;
  int q;
  template <int Rows>
  __device__ void rows(uint64_t wbase, int lane, uint64_t n, uint32_t (&v)[Rows],
                       uint64_t (&bal)[Rows]) const {
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      v[i] = (uint32_t)j + base;
      bal[i] = __ballot(j < n && ((mask[j < n ? j : 0] >> q) & 1ull));
    }
  }
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint32_t v) const {
    if (keep) vals[rank] = v;
  }
};

// The local reps {r : comp_of[r] == r}, ascending (the rep list of
// src/cc_mpi.c:328-362), row-striped.
struct RepRows {
  using Val = uint32_t;
  static constexpr int kRows = 32;
  const uint32_t *comp_of;
  uint32_t *reps;
  template <int Rows>
  __device__ void rows(uint64_t wbase, int lane, uint64_t n, uint32_t (&v)[Rows],
                       uint64_t (&bal)[Rows]) const {
    uint32_t c[Rows];
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      c[i] = j < n ? comp_of[j] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      v[i] = (uint32_t)j;
      bal[i] = __ballot(j < n && c[i] == (uint32_t)j);
    }
  }
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint32_t v) const {
    if (keep) reps[rank] = v;
  }
};

inline unsigned tiles_for(uint64_t n) { return (unsigned)((n + kTile - 1) / kTile); }

template <class Op>
int run_compact(const char *name, Op op, uint64_t n, hipStream_t s, ccx::StreamState *st,
                uint64_t *count_out) {
  if (n == 0) {
    *count_out = 0;
    return 0;
  }
  const uint64_t tile_keys = (uint64_t)kThreads * Op::kItemsPerThread;
  const uint64_t tiles = (n + tile_keys - 1) / tile_keys;
  if (tiles > 0xFFFFFFFFull) return ccx::set_error(4, "%s: input too large", name);
  if (int rc = ccx::ensure_status(st, tiles)) return rc;
  uint32_t epoch;
  if (int rc = ccx::next_epoch(st, &epoch)) return rc;
  uint32_t *ctr = st->d_words + 8;
  unsigned long long *tot = (unsigned long long *)(st->d_words + 16);
  CCX_TRY(hipMemsetAsync(ctr, 0, sizeof(uint32_t), s));
  CCX_LAUNCH(name, s,
             k_compact<Op><<<(unsigned)tiles, kThreads, 0, s>>>(op, n, ctr, st->d_status, epoch, tot));
  CCX_TRY(hipMemcpyAsync(st->h_stage, tot, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  *count_out = st->h_stage[0];
  return 0;
}

// ---- boundary build with compressed reps ----------------------------------
// The reps of a rank's boundary records are roots of its local forest, of
// which there are usually few (one per local component touching the
// boundary, often a few thousand against 2^24 ids).  Replacing each rep by
// its rank among the roots keeps the key order and equality (the map is
// increasing) while the key shrinks from 32 + 32 bits to 32 + ceil(log2
// roots), which drops whole digit passes from the sort; the dedup writes
// the records back with the original reps.
__global__ __launch_bounds__(kThreads) void k_root_flags(const uint32_t *__restrict__ comp_of,
                                                         uint32_t n, uint32_t *__restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    flags[i] = comp_of[i] == (uint32_t)i;
}

__global__ __launch_bounds__(kThreads) void k_rep_tables(const uint32_t *__restrict__ comp_of,
                                                         uint32_t n, const uint64_t *__restrict__ scan,
                                                         uint32_t *__restrict__ packed,
                                                         uint32_t *__restrict__ rep_of_rank) {
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads) {
    const uint32_t c = comp_of[i];
    packed[i] = (uint32_t)scan[c < n ? c : (uint32_t)i];
    if (c == (uint32_t)i) rep_of_rank[scan[i]] = (uint32_t)i;
  }
}

__global__ __launch_bounds__(kThreads) void k_set_rep_packed(uint64_t *__restrict__ rec, uint64_t nb,
                                                             const uint32_t *__restrict__ packed,
                                                             uint32_t n, int bits) {
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < nb;
       i += (uint64_t)gridDim.x * kThreads) {
    const uint64_t r = rec[i];
    const uint32_t li = (uint32_t)r;
    rec[i] = ((r >> 32) << bits) | (li < n ? packed[li] : 0u);
  }
}

// ---- boundary build, dominant rep split off through a bitmap ---------------
// Most boundary records of a rank carry one rep, the root of its giant local
// component (at the 4-rank share of the 65.6M-vertex graph, 46.5M of the
// 52.8M distinct pairs, from 638M raw records).  Their distinct pairs are
// just the distinct remotes, so those records only set one bit per remote;
// the rest go through pack -> sort -> dedup as before, and a merge of the
// two sorted lists gives the bytes of the full sort + dedup.
constexpr int kModeSamples = 1024;

// Most frequent packed rep (rank among the roots) among kModeSamples records
// spread evenly over the input; ties go to the smaller rank.  One workgroup.
// With packed null the records' reps are taken as they are.
__global__ __launch_bounds__(kModeSamples) void k_mode_rank(const uint64_t *__restrict__ rec,
                                                            uint64_t count,
                                                            const uint32_t *__restrict__ packed,
                                                            uint32_t n, uint32_t *__restrict__ out) {
  constexpr uint32_t kSlots = 2 * kModeSamples;
  constexpr uint32_t kEmpty = 0xFFFFFFFFu;
  __shared__ uint32_t keys[kSlots];
  __shared__ uint32_t cnts[kSlots];
  __shared__ uint64_t best[kModeSamples / kWave];
  const int t = threadIdx.x;
  for (uint32_t i = t; i < kSlots; i += kModeSamples) {
    keys[i] = kEmpty;
    cnts[i] = 0;
  }
  __syncthreads();
  {
    const uint32_t li = (uint32_t)rec[(uint64_t)t * count / kModeSamples];
    const uint32_t x = packed ? (li < n ? packed[li] : 0u) : li;
    uint32_t h = (x * 0x9E3779B1u) >> 21;  // 11-bit slot
    for (;;) {
      const uint32_t prev = atomicCAS(&keys[h], kEmpty, x);
      if (prev == kEmpty || prev == x) {
        atomicAdd(&cnts[h], 1u);
        break;
      }
      h = (h + 1) & (kSlots - 1);
    }
  }
  __syncthreads();
  uint64_t mine = 0;  // (count, ~rank): the max is the most frequent, then the smallest
  for (uint32_t i = t; i < kSlots; i += kModeSamples)
    if (cnts[i]) {
      const uint64_t cand = ((uint64_t)cnts[i] << 32) | (uint32_t)~keys[i];
      if (cand > mine) mine = cand;
    }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const uint64_t o = __shfl_xor(mine, d, kWave);
    if (o > mine) mine = o;
  }
  if ((t & (kWave - 1)) == 0) best[t / kWave] = mine;
  __syncthreads();
  if (t == 0) {
    uint64_t b = best[0];
    for (int w = 1; w < kModeSamples / kWave; ++w)
      if (best[w] > b) b = best[w];
    out[0] = ~(uint32_t)b;
  }
}

// The split as an ordered compaction (k_compact_rows): records whose
// packed rep is *g_rank set bit `remote` of the bitmap; the others are kept,
// packed as (remote << bits) | rank, in input order.
template <int R>
struct SplitRows {
  using Val = uint64_t;
  static constexpr int kRows = R;
  const uint64_t *rec;
  const uint32_t *packed;
  uint32_t n;
  int bits;
  const uint32_t *g_rank;
  uint32_t *bitmap;
  uint64_t *others;
  template <int Rows>
  __device__ void rows(uint64_t wbase, int lane, uint64_t count, uint64_t (&v)[Rows],
                       uint64_t (&bal)[Rows]) const {
    const uint32_t g = *g_rank;
    uint32_t rank[Rows];
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      v[i] = j < count ? __builtin_nontemporal_load(&rec[j]) : 0;
    }
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      const uint32_t li = (uint32_t)v[i];
      rank[i] = (j < count && li < n) ? packed[li] : 0u;
    }
    // Read first: most bits are already set, and a plain read is far
    // cheaper than an atomic (an atomic on every record measured 3x slower
    // on 638M records).  All rows' words are read before any atomic, so the
    // reads are in flight together (a read after an atomic on the same array
    // waits for it); a stale read only costs a redundant atomic.
    uint32_t word[Rows];
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      const uint32_t remote = (uint32_t)(v[i] >> 32);
      word[i] = (j < count && rank[i] == g) ? bitmap[remote >> 5] : ~0u;
    }
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      const uint32_t remote = (uint32_t)(v[i] >> 32);
      const bool dominant = j < count && rank[i] == g;
      const uint32_t bit = 1u << (remote & 31);
      if (dominant && !(word[i] & bit)) atomicOr(&bitmap[remote >> 5], bit);
      bal[i] = __ballot(j < count && !dominant);
      v[i] = ((uint64_t)remote << bits) | rank[i];
    }
  }
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) others[rank] = v;
  }
};

// The same split by partitioning instead of probing (boundary_split 2).
// The probing split above reads one random bitmap word per dominant record
// (638M reads of a few-MB bitmap), which the L2s serve at a fixed request
// rate.  Here the dominant records' remotes are first scattered into
// buckets of 2^shift consecutive remotes, so that one workgroup per bucket
// sets that bucket's bits in LDS and stores the window's words once:
//   count:   per tile of kPartTile records, how many dominant records fall
//            in each bucket and how many others there are (LDS counters);
//   (the counts, stored tile-major, are transposed to bucket-major; two
//    exclusive scans give every tile its place in each bucket region and
//    in the others' list; the bucket offsets are transposed back);
//   scatter: the same tiles again; the tile's dominant remotes are grouped
//            by bucket in LDS and each bucket's run is stored in one piece
//            in its region; each other record, packed as in SplitRows,
//            goes to the others' list;
//   bitmap:  one workgroup per bucket.
// Order inside a tile is not kept (positions come from LDS atomics): the
// bitmap is a set and the others are sorted afterwards, so the result is
// the same bytes as the probing split's.
constexpr uint64_t kPartTile = 16384;
constexpr int kPartUnroll = 8;
constexpr uint32_t kPartMaxBuckets = 1024;
constexpr int kPartMaxShift = 17;  // 2^17 remotes = 4096 words = 16 KiB of LDS per bucket
constexpr int kPartBitmapThreads = 1024;

__global__ __launch_bounds__(kThreads) void k_part_count(
    const uint64_t *__restrict__ rec, uint64_t count, const uint32_t *__restrict__ packed,
    uint32_t n, const uint32_t *__restrict__ g_rank, int shift, uint32_t nb, uint64_t tiles,
    uint32_t *__restrict__ cnt_dom, uint32_t *__restrict__ cnt_other) {
  __shared__ uint32_t c[kPartMaxBuckets];
  __shared__ uint32_t c_other;
  const uint32_t g = *g_rank;
  const int lane = threadIdx.x & (kWave - 1);
  for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    for (uint32_t b = threadIdx.x; b < nb; b += kThreads) c[b] = 0;
    if (threadIdx.x == 0) c_other = 0;
    __syncthreads();
    const uint64_t base = tile * kPartTile;
    for (uint64_t r = 0; r < kPartTile; r += (uint64_t)kThreads * kPartUnroll) {
      uint64_t v[kPartUnroll];
      uint32_t rank[kPartUnroll];
#pragma unroll
      for (int u = 0; u < kPartUnroll; ++u) {
        const uint64_t j = base + r + (uint64_t)u * kThreads + threadIdx.x;
        v[u] = j < count ? __builtin_nontemporal_load(&rec[j]) : 0;
      }
#pragma unroll
      for (int u = 0; u < kPartUnroll; ++u) {
        const uint64_t j = base + r + (uint64_t)u * kThreads + threadIdx.x;
        const uint32_t li = (uint32_t)v[u];
        rank[u] = (j < count && li < n) ? packed[li] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kPartUnroll; ++u) {
        const uint64_t j = base + r + (uint64_t)u * kThreads + threadIdx.x;
        const bool dom = j < count && rank[u] == g;
        if (dom) atomicAdd(&c[(uint32_t)(v[u] >> 32) >> shift], 1u);
        const uint64_t other = __ballot(j < count && !dom);
        if (lane == 0 && other) atomicAdd(&c_other, (uint32_t)__popcll(other));
      }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += kThreads) cnt_dom[tile * nb + b] = c[b];
    if (threadIdx.x == 0) cnt_other[tile] = c_other;
    __syncthreads();
  }
}

// Exclusive prefix sums over a 512-thread workgroup (x from every thread);
// returns this thread's prefix and sets *total.  Uses `wsum` (8 words of
// LDS); ends with a barrier.
constexpr int kPartScatterThreads = 512;
__device__ __forceinline__ uint32_t block_excl_scan_512(uint32_t x, uint32_t *wsum,
                                                       uint32_t *total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t up = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += up;
  }
  if (lane == kWave - 1) wsum[w] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kPartScatterThreads / kWave; ++k) {
    const uint32_t t = wsum[k];
    before += (k < w) ? t : 0u;
    all += t;
  }
  *total = all;
  __syncthreads();
  return before + incl - x;
}

// Scatter by tiles of kPartTile records: the tile's dominant remotes are
// first grouped by bucket in LDS (bucket b's run starts at the prefix sum
// of the tile's counts, which the count pass stored), then each run is
// written to the bucket's region in one contiguous piece, so the stores
// coalesce; the others go straight to their list.
__global__ __launch_bounds__(kPartScatterThreads) void k_part_scatter(
    const uint64_t *__restrict__ rec, uint64_t count, const uint32_t *__restrict__ packed,
    uint32_t n, const uint32_t *__restrict__ g_rank, int shift, uint32_t nb, uint64_t tiles,
    int bits, const uint32_t *__restrict__ cnt_dom, const uint64_t *__restrict__ off_tile,
    const uint64_t *__restrict__ off_other, uint32_t *__restrict__ remotes,
    uint64_t *__restrict__ others) {
  __shared__ uint32_t stage[kPartTile];
  __shared__ uint64_t base_b[kPartMaxBuckets];
  __shared__ uint32_t loc[kPartMaxBuckets];
  __shared__ uint32_t cur[kPartMaxBuckets];
  __shared__ uint32_t wsum[kPartScatterThreads / kWave];
  __shared__ uint32_t c_other;
  constexpr uint32_t kPer = kPartMaxBuckets / kPartScatterThreads;  // buckets per thread in the scan
  const uint32_t g = *g_rank;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t below = (lane == 0) ? 0 : (~0ull >> (kWave - lane));  // lanes under this one
  for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    // this tile's runs: bucket b at loc[b] in stage
    uint32_t mine[kPer];
    uint32_t part = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t b = threadIdx.x * kPer + k;
      mine[k] = b < nb ? cnt_dom[tile * nb + b] : 0u;
      part += mine[k];
    }
    uint32_t tile_dom = 0;
    uint32_t at = block_excl_scan_512(part, wsum, &tile_dom);
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t b = threadIdx.x * kPer + k;
      if (b < nb) {
        loc[b] = at;
        cur[b] = at;
        base_b[b] = off_tile[tile * nb + b];
      }
      at += mine[k];
    }
    if (threadIdx.x == 0) c_other = 0;
    __syncthreads();
    const uint64_t base_o = off_other[tile];
    const uint64_t base = tile * kPartTile;
    for (uint64_t r = 0; r < kPartTile; r += (uint64_t)kPartScatterThreads * kPartUnroll) {
      uint64_t v[kPartUnroll];
      uint32_t rank[kPartUnroll];
#pragma unroll
      for (int u = 0; u < kPartUnroll; ++u) {
        const uint64_t j = base + r + (uint64_t)u * kPartScatterThreads + threadIdx.x;
        v[u] = j < count ? __builtin_nontemporal_load(&rec[j]) : 0;
      }
#pragma unroll
      for (int u = 0; u < kPartUnroll; ++u) {
        const uint64_t j = base + r + (uint64_t)u * kPartScatterThreads + threadIdx.x;
        const uint32_t li = (uint32_t)v[u];
        rank[u] = (j < count && li < n) ? packed[li] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kPartUnroll; ++u) {
        const uint64_t j = base + r + (uint64_t)u * kPartScatterThreads + threadIdx.x;
        const bool dom = j < count && rank[u] == g;
        const uint32_t remote = (uint32_t)(v[u] >> 32);
        if (dom) stage[atomicAdd(&cur[remote >> shift], 1u)] = remote;
        const bool keep = j < count && !dom;
        const uint64_t other = __ballot(keep);
        uint32_t first = 0;
        if (lane == 0 && other) first = atomicAdd(&c_other, (uint32_t)__popcll(other));
        first = __shfl(first, 0, kWave);
        if (keep)
          others[base_o + first + (uint32_t)__popcll(other & below)] =
              ((uint64_t)remote << bits) | rank[u];
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tile_dom; i += kPartScatterThreads) {
      const uint32_t x = stage[i];
      const uint32_t b = x >> shift;
      remotes[base_b[b] + (i - loc[b])] = x;
    }
    __syncthreads();
  }
}

// out[c * rows + r] = in[r * cols + c] for an rows x cols matrix, through
// 32 x 32 LDS tiles (the count pass stores each tile's bucket counts
// contiguously; the scan needs them bucket-major, and the scatter pass the
// resulting offsets tile-major again).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_transpose(const T *__restrict__ in, uint64_t rows,
                                                         uint64_t cols, T *__restrict__ out) {
  __shared__ T t[32][33];
  const uint64_t tiles_c = (cols + 31) / 32;
  const uint64_t tiles_r = (rows + 31) / 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x / 32;  // 32 x 8 threads
  for (uint64_t blk = blockIdx.x; blk < tiles_r * tiles_c; blk += gridDim.x) {
    const uint64_t r0 = (blk / tiles_c) * 32, c0 = (blk % tiles_c) * 32;
    for (int k = ty; k < 32; k += kThreads / 32)
      if (r0 + k < rows && c0 + tx < cols) t[k][tx] = in[(r0 + k) * cols + c0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += kThreads / 32)
      if (c0 + k < cols && r0 + tx < rows) out[(c0 + k) * rows + r0 + tx] = t[tx][k];
    __syncthreads();
  }
}

// One workgroup per bucket: the bucket's remotes set bits of a window of
// `ww` words in LDS, which is then stored over the bitmap's words
// [b * ww, (b + 1) * ww) (clipped to `words`); every bitmap word belongs to
// exactly one bucket, so nothing else writes it.
__global__ __launch_bounds__(kPartBitmapThreads) void k_part_bitmap(
    const uint32_t *__restrict__ remotes, const uint64_t *__restrict__ off_dom, uint64_t tiles,
    uint32_t ww, uint64_t words, uint32_t *__restrict__ bitmap) {
  __shared__ uint32_t win[1u << (kPartMaxShift - 5)];
  const uint64_t b = blockIdx.x;
  for (uint32_t i = threadIdx.x; i < ww; i += kPartBitmapThreads) win[i] = 0;
  __syncthreads();
  const uint64_t lo = off_dom[b * tiles];
  const uint64_t hi = off_dom[(b + 1) * tiles];
  const uint32_t mask = ww - 1;
  uint64_t k = lo + threadIdx.x;
  for (; k + 3 * kPartBitmapThreads < hi; k += 4 * kPartBitmapThreads) {
    uint32_t r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) r[u] = __builtin_nontemporal_load(&remotes[k + u * kPartBitmapThreads]);
#pragma unroll
    for (int u = 0; u < 4; ++u) atomicOr(&win[(r[u] >> 5) & mask], 1u << (r[u] & 31));
  }
  for (; k < hi; k += kPartBitmapThreads) {
    const uint32_t r = remotes[k];
    atomicOr(&win[(r >> 5) & mask], 1u << (r & 31));
  }
  __syncthreads();
  const uint64_t w0 = b * ww;
  for (uint32_t i = threadIdx.x; i < ww && w0 + i < words; i += kPartBitmapThreads)
    bitmap[w0 + i] = win[i];
}

// Index of rep r in the ascending rep list (rep_count if absent).
__device__ __forceinline__ uint32_t rep_index(const uint32_t *reps, uint32_t rep_count, uint32_t r) {
  uint32_t lo = 0, hi = rep_count;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (reps[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  return (lo < rep_count && reps[lo] == r) ? lo : rep_count;
}

// The merge loop's records split by rep (cc_hip_merge_prepare_device):
// records (gidx << 32) | rep whose rep is *g set bit gidx of a bitmap over
// the ghosts; the others are kept in input order.  The records are sorted,
// so the lanes' bitmap words are nondecreasing along a row: each run of
// lanes on one word ORs its bits together (a segmented reduction across
// the wave) and the run's first lane sets them with one atomic.  The
// records are unique, so every bit is new (no read-before-atomic filter).
// The kept records' reps become indices in the rep list afterwards
// (k_side_rep_index), so that the rounds keep their per-rep state in an
// array of rep_count entries.
template <int R>
struct DomRows {
  using Val = uint64_t;
  static constexpr int kRows = R;
  const uint64_t *rec;
  const uint32_t *g;
  uint32_t nbits;  // ghost count: a record with gidx >= nbits is kept as a side record
  uint32_t *bitmap;
  uint64_t *side;

  template <int Rows>
  __device__ void rows(uint64_t wbase, int lane, uint64_t count, uint64_t (&v)[Rows],
                       uint64_t (&bal)[Rows]) const {
    const uint32_t d = *g;
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      v[i] = j < count ? __builtin_nontemporal_load(&rec[j]) : 0;
    }
#pragma unroll
    for (int i = 0; i < Rows; ++i) {
      const uint64_t j = wbase + (uint64_t)i * kWave + lane;
      const bool live = j < count;
      const uint32_t gidx = (uint32_t)(v[i] >> 32);
      const bool dom = live && (uint32_t)v[i] == d && gidx < nbits;
      const uint32_t word = live ? (gidx >> 5) : 0xFFFFFFFFu;
      uint32_t bits = dom ? (1u << (gidx & 31)) : 0u;
#pragma unroll
      for (int dd = 1; dd < kWave; dd <<= 1) {
        const uint32_t ob = __shfl_down(bits, dd, kWave);
        const uint32_t ow = __shfl_down(word, dd, kWave);
        if (lane + dd < kWave && ow == word) bits |= ob;
      }
      const uint32_t up = __shfl_up(word, 1, kWave);
      if (bits && (lane == 0 || up != word)) atomicOr(&bitmap[word], bits);
      bal[i] = __ballot(live && !dom);
    }
  }
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) side[rank] = v;
  }
};

// side[i] = (gidx << 32) | index of its rep in the rep list; *bad = 1 if
// some rep is not in it.  A kernel of its own over the side records: the
// same binary searches inside the split pass, row after row in a wave,
// made that pass 4-7x slower.
// Up to kLdsReps reps the list is staged in LDS first, so the 11 or so
// dependent probes of each search are LDS reads rather than L1 ones.
constexpr uint32_t kLdsReps = 8192;

__global__ __launch_bounds__(kThreads) void k_side_rep_index(uint64_t *__restrict__ side, uint64_t n,
                                                             const uint32_t *__restrict__ reps,
                                                             uint32_t rep_count,
                                                             uint32_t *__restrict__ bad) {
  __shared__ uint32_t s_reps[kLdsReps];
  const bool staged = rep_count <= kLdsReps;
  if (staged) {
    for (uint32_t k = threadIdx.x; k < rep_count; k += kThreads) s_reps[k] = reps[k];
    __syncthreads();
  }
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads) {
    const uint64_t v = side[i];
    const uint32_t k = staged ? rep_index(s_reps, rep_count, (uint32_t)v)
                              : rep_index(reps, rep_count, (uint32_t)v);
    if (k == rep_count) *bad = 1u;
    side[i] = (v & 0xFFFFFFFF00000000ull) | k;
  }
}

// out[0] = index of *rep in the rep list (rep_count if absent).  One thread.
__global__ void k_rep_index(const uint32_t *__restrict__ reps, uint32_t rep_count,
                            const uint32_t *__restrict__ rep, uint32_t *__restrict__ out) {
  out[0] = rep_index(reps, rep_count, rep[0]);
}

__global__ __launch_bounds__(kThreads) void k_fill_u32(uint32_t *__restrict__ a, uint64_t n,
                                                       uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    a[i] = v;
}

__global__ __launch_bounds__(kThreads) void k_word_counts(const uint32_t *__restrict__ bitmap,
                                                          uint64_t words, uint32_t *__restrict__ cnt) {
  for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < words;
       w += (uint64_t)gridDim.x * kThreads)
    cnt[w] = (uint32_t)__popc(bitmap[w]);
}

// Set bits in ascending order as records (remote << 32) | rep of *g_rank,
// word w's bits starting at off[w].
__global__ __launch_bounds__(kThreads) void k_emit_bits(const uint32_t *__restrict__ bitmap,
                                                        uint64_t words,
                                                        const uint64_t *__restrict__ off,
                                                        const uint32_t *__restrict__ rep_of_rank,
                                                        const uint32_t *__restrict__ g_rank,
                                                        uint64_t *__restrict__ out) {
  const uint64_t rep = rep_of_rank[*g_rank];
  for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < words;
       w += (uint64_t)gridDim.x * kThreads) {
    uint32_t b = bitmap[w];
    uint64_t o = off[w];
    while (b) {
      const uint32_t k = (uint32_t)__ffs(b) - 1;
      b &= b - 1;
      out[o++] = ((w * 32 + k) << 32) | rep;
    }
  }
}

// Merge of two sorted arrays with no key in common (merge path), one tile
// of kMergeTile outputs per workgroup: the tile's first and last diagonals
// are found by binary search in HBM, the two input runs they bound are
// staged in LDS (coalesced), each thread merges its 16 outputs after a
// binary search of its own diagonal in LDS, and the tile is written out
// from LDS in one coalesced pass.
constexpr int kMergePer = 16;
constexpr int kMergeTile = kThreads * kMergePer;

// How many of the first d outputs come from a: the smallest i with
// a[i] > b[d - i - 1] (the arrays share no key).
template <class A, class B>
__device__ __forceinline__ uint64_t merge_split(const A &a, uint64_t na, const B &b, uint64_t nb,
                                                uint64_t d) {
  uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < b[d - mid - 1]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kThreads) void k_merge_tiles(const uint64_t *__restrict__ a, uint64_t na,
                                                          const uint64_t *__restrict__ b, uint64_t nb,
                                                          uint64_t *__restrict__ out) {
  __shared__ uint64_t s_in[kMergeTile];
  __shared__ uint64_t s_out[kMergeTile];
  __shared__ uint64_t s_split[2];
  const uint64_t total = na + nb;
  const int t = threadIdx.x;
  for (uint64_t tile = blockIdx.x; tile * kMergeTile < total; tile += gridDim.x) {
    const uint64_t d0 = tile * kMergeTile;
    const uint64_t d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
    if (t < 2) s_split[t] = merge_split(a, na, b, nb, t == 0 ? d0 : d1);
    __syncthreads();
    const uint64_t i0 = s_split[0], i1 = s_split[1];
    const uint32_t ta = (uint32_t)(i1 - i0), tn = (uint32_t)(d1 - d0), tb = tn - ta;
    const uint64_t j0 = d0 - i0;
    for (uint32_t x = t; x < tn; x += kThreads) s_in[x] = x < ta ? a[i0 + x] : b[j0 + (x - ta)];
    __syncthreads();
    const uint64_t *la = s_in, *lb = s_in + ta;
    const uint32_t dk = (uint32_t)t * kMergePer;
    if (dk < tn) {
      uint32_t i = (uint32_t)merge_split(la, ta, lb, tb, dk), j = dk - i;
      const uint32_t end = dk + kMergePer < tn ? dk + kMergePer : tn;
      for (uint32_t o = dk; o < end; ++o) s_out[o] = (j >= tb || (i < ta && la[i] < lb[j])) ? la[i++] : lb[j++];
    }
    __syncthreads();
    for (uint32_t x = t; x < tn; x += kThreads) out[d0 + x] = s_out[x];
    __syncthreads();  // s_split and the LDS runs are rewritten by the next tile
  }
}

// Adjacent-unique over packed keys, writing (remote << 32) | rep.
template <int Items = kItems>
struct DedupUnpackOp {
  static constexpr bool kStaged = false;
  static constexpr int kItemsPerThread = Items;
  const uint64_t *in;
  uint64_t *out;
  const uint32_t *rep_of_rank;
  int bits;
  __device__ bool eval(uint64_t j, uint64_t &v) const {
    v = in[j];
    return j == 0 || v != in[j - 1];
  }
  __device__ void store(uint64_t, uint64_t) const {}
  __device__ void begin(uint64_t) {}
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) out[rank] = ((v >> bits) << 32) | rep_of_rank[(uint32_t)(v & ((1ull << bits) - 1))];
  }
};

// The same, row-striped.
struct DedupUnpackRows {
  using Val = uint64_t;
  static constexpr int kRows = 32;
  const uint64_t *in;
  uint64_t *out;
  const uint32_t *rep_of_rank;
  int bits;
  template <int Rows>
  __device__ void rows(uint64_t wbase, int lane, uint64_t n, uint64_t (&v)[Rows],
                       uint64_t (&bal)[Rows]) const {
    adjacent_rows<Rows>(in, wbase, lane, n, v, bal, KeyAll{});
  }
  __device__ void apply(uint64_t, bool keep, uint64_t rank, uint64_t v) const {
    if (keep) out[rank] = ((v >> bits) << 32) | rep_of_rank[(uint32_t)(v & ((1ull << bits) - 1))];
  }
};

template <class Op>
int run_compact_rows(const char *name, Op op, uint64_t n, hipStream_t s, ccx::StreamState *st,
                     uint64_t *count_out) {
  if (n == 0) {
    *count_out = 0;
    return 0;
  }
  const uint64_t tile_items = (uint64_t)kThreads * Op::kRows;
  const uint64_t tiles = (n + tile_items - 1) / tile_items;
  if (tiles > 0xFFFFFFFFull) return ccx::set_error(4, "%s: input too large", name);
  if (int rc = ccx::ensure_status(st, tiles)) return rc;
  uint32_t epoch;
  if (int rc = ccx::next_epoch(st, &epoch)) return rc;
  uint32_t *ctr = st->d_words + 8;
  unsigned long long *tot = (unsigned long long *)(st->d_words + 16);
  CCX_TRY(hipMemsetAsync(ctr, 0, sizeof(uint32_t), s));
  CCX_LAUNCH(name, s,
             k_compact_rows<Op><<<(unsigned)tiles, kThreads, 0, s>>>(op, n, ctr, st->d_status, epoch,
                                                                     tot));
  CCX_TRY(hipMemcpyAsync(st->h_stage, tot, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  *count_out = st->h_stage[0];
  return 0;
}

int run_dedup(const uint64_t *in, uint64_t *out, uint64_t n, hipStream_t s, ccx::StreamState *st,
              uint64_t *count_out) {
  if (ccx::g_compact_rows)
    return run_compact_rows("dedup_compact", DedupRows{in, out}, n, s, st, count_out);
  if (ccx::g_dedup_staged)
    return run_compact("dedup_compact", DedupOp<true>{in, out}, n, s, st, count_out);
  if (ccx::g_dedup_items == 32)
    return run_compact("dedup_compact", DedupOp<false, 32>{in, out}, n, s, st, count_out);
  return run_compact("dedup_compact", DedupOp<false>{in, out}, n, s, st, count_out);
}

}  // namespace

// ---- sort ------------------------------------------------------------------

namespace ccx {
int g_split_staged = 0;
int g_compact_rows = 1;
int g_dedup_staged = 0;
int g_dedup_items = 32;
}  // namespace ccx

namespace {
// Scatter-pass tile: threads x items keys.  512 x 16 (8192 keys, 16 waves
// per CU at 2 workgroups) measured 1.44 ms per pass over 268M keys, against
// 1.65 ms for 256 x 32 and 1.89 ms for 1024 x 8.
int g_sort_items = 16;     // keys per thread (8, 16, 24, 32)
int g_sort_threads = 512;  // threads per workgroup (256; 512 with 8 or 16 items; 1024 with 8)
// boundary build: dominant rep through a bitmap, set by partitioning the
// remotes into LDS windows (2; at the 4-rank share of the 65.6M-vertex
// graph the split takes 5.3 ms against 7.5 ms) or by probing (1); 0 sorts
// every record
int g_boundary_split = 2;
int g_split_dominant_rows = 32;  // rows of 64 records per wave step in the bitmap split (16, 32)
// workgroups for `threads` threads, at most per_cu per CU
unsigned grid_cap(uint64_t threads, int per_cu);
}  // namespace

extern "C" int cc_hip_sort_set(const char *key, int value) {
  if (key && strcmp(key, "delta_passes") == 0) {
    if (ccx::delta_passes_set(value) == 0) return 0;
    return ccx::set_error(4, "cc_hip_sort_set: bad key or value (%s=%d)", key, value);
  }

  if (key && strcmp(key, "items_per_thread") == 0 &&
      (value == 8 || value == 16 || value == 24 || value == 32)) {
    g_sort_items = value;
    return 0;
  }
  if (key && strcmp(key, "split_staged") == 0 && (value == 0 || value == 1)) {
    ccx::g_split_staged = value;
    return 0;
  }
  if (key && strcmp(key, "sort_threads") == 0 && (value == 256 || value == 512 || value == 1024)) {
    g_sort_threads = value;
    return 0;
  }
  if (key && strcmp(key, "compact_rows") == 0 && (value == 0 || value == 1)) {
    ccx::g_compact_rows = value;
    return 0;
  }
  if (key && strcmp(key, "boundary_split") == 0 && (value >= 0 && value <= 2)) {
    g_boundary_split = value;
    return 0;
  }
  if (key && strcmp(key, "split_dominant_rows") == 0 && (value == 16 || value == 32)) {
    g_split_dominant_rows = value;
    return 0;
  }
  if (key && strcmp(key, "dedup_staged") == 0 && (value == 0 || value == 1)) {
    ccx::g_dedup_staged = value;
    return 0;
  }
  if (key && strcmp(key, "dedup_items") == 0 && (value == 16 || value == 32)) {
    ccx::g_dedup_items = value;
    return 0;
  }
  return ccx::set_error(4, "cc_hip_sort_set: bad key or value (%s=%d)", key ? key : "", value);
}

extern "C" int cc_hip_sort_get(const char *key) {
  if (key && strcmp(key, "items_per_thread") == 0) return g_sort_items;
  if (key && strcmp(key, "delta_passes") == 0) return ccx::delta_passes_get();

  if (key && strcmp(key, "sort_threads") == 0) return g_sort_threads;
  if (key && strcmp(key, "split_staged") == 0) return ccx::g_split_staged;
  if (key && strcmp(key, "compact_rows") == 0) return ccx::g_compact_rows;
  if (key && strcmp(key, "boundary_split") == 0) return g_boundary_split;
  if (key && strcmp(key, "split_dominant_rows") == 0) return g_split_dominant_rows;
  if (key && strcmp(key, "dedup_staged") == 0) return ccx::g_dedup_staged;
  if (key && strcmp(key, "dedup_items") == 0) return ccx::g_dedup_items;
  return -1;
}

namespace ccx {
// Sorts keys[0..n) ascending; alt is a buffer of n keys.  On return
// *in_alt says which buffer holds the result.  Blocks until the digit plan
// is known (one small device-to-host read), then enqueues the passes.
int sort_u64(uint64_t *keys, uint64_t *alt, uint64_t n, int *in_alt, hipStream_t s) {
  *in_alt = 0;
  if (n <= 1) return 0;
  StreamState *st = stream_state(s);
  if (!st) return 3;
  unsigned long long *d_ao = (unsigned long long *)(st->d_words + 32);    // 2 x u64
  unsigned long long *d_hist = nullptr;
  uint64_t *d_goff = nullptr;
  // histograms (8 x 256 u64) and offsets (8 x 256 u64) live in the temp buffer
  char *tmp = (char *)stream_tmp(st, 2 * 8 * kBins * sizeof(uint64_t));
  if (!tmp) return 3;
  d_hist = (unsigned long long *)tmp;
  d_goff = (uint64_t *)(tmp + 8 * kBins * sizeof(uint64_t));

  uint64_t init_ao[2] = {~0ull, 0ull};
  memcpy(st->h_stage, init_ao, sizeof(init_ao));
  CCX_TRY(hipMemcpyAsync(d_ao, st->h_stage, sizeof(init_ao), hipMemcpyHostToDevice, s));
  CCX_TRY(hipMemsetAsync(d_hist, 0, 8 * kBins * sizeof(uint64_t), s));
  {
    // 32 KiB of LDS per workgroup: 4 per CU resident
    const uint64_t tiles = tiles_for(n);
    const uint64_t cap = (uint64_t)(ctx()->cu_count > 0 ? ctx()->cu_count : 256) * 4;
    const unsigned hb = (unsigned)(tiles < cap ? tiles : cap);
    CCX_LAUNCH("sort_histograms", s, k_histograms<<<hb, kThreads, 0, s>>>(keys, n, d_hist, d_ao));
  }
  CCX_TRY(hipMemcpyAsync(st->h_stage, d_ao, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  const uint64_t diff = st->h_stage[0] ^ st->h_stage[1];
  DigitPlan plan;
  int passes = 0;
  for (int d = 0; d < 8; ++d) {
    plan.vary[d] = (uint32_t)(diff >> (8 * d)) & 0xFFu;
    passes += plan.vary[d] != 0;
  }
  if (passes == 0) return 0;  // all keys equal
  CCX_LAUNCH("sort_digit_offsets", s, k_digit_offsets<<<8, kBins, 0, s>>>(d_hist, d_goff));
  const int items = g_sort_items;
  // larger workgroups only with tiles that fit their LDS (512 threads: 8 or
  // 16 keys each; 1024 threads: 8)
  const int threads = (g_sort_threads == 1024 && items == 8)  ? 1024
                      : (g_sort_threads >= 512 && items <= 16) ? 512
                                                               : kThreads;
  const uint64_t tile_keys = (uint64_t)threads * items;
  const uint64_t tiles = (n + tile_keys - 1) / tile_keys;
  if (tiles > 0xFFFFFFFFull) return set_error(4, "sort: input too large");
  if (int rc = ensure_status(st, tiles * kBins)) return rc;
  uint64_t *src = keys, *dst = alt;
  int pass = 0;
  // one tile counter per pass (at most 8), zeroed together
  CCX_TRY(hipMemsetAsync(st->d_words + 8, 0, 8 * sizeof(uint32_t), s));
  for (int d = 0; d < 8; ++d) {
    if (!plan.vary[d]) continue;
    uint32_t epoch;
    if (int rc = next_epoch(st, &epoch)) return rc;
    uint32_t *ctr = st->d_words + 8 + pass;
#define CCX_ONESWEEP(I, T)                                                                   \
  CCX_LAUNCH("sort_onesweep", s,                                                             \
             (k_onesweep<I, T><<<(unsigned)tiles, T, 0, s>>>(src, dst, n, 8 * d, plan.vary[d], \
                                                             d_goff + d * kBins, ctr,          \
                                                             st->d_status, epoch)))
    if (threads == 1024) CCX_ONESWEEP(8, 1024);
    else if (threads == 512 && items == 8) CCX_ONESWEEP(8, 512);
    else if (threads == 512) CCX_ONESWEEP(16, 512);
    else if (items == 8) CCX_ONESWEEP(8, 256);
    else if (items == 24) CCX_ONESWEEP(24, 256);
    else if (items == 32) CCX_ONESWEEP(32, 256);
    else CCX_ONESWEEP(16, 256);
#undef CCX_ONESWEEP
    uint64_t *tt = src;
    src = dst;
    dst = tt;
    ++pass;
  }
  *in_alt = (src == alt);
  return 0;
}
}  // namespace ccx

extern "C" int cc_hip_sort_u64_device(uint64_t *d_keys, uint64_t *d_alt, uint64_t count,
                                      int *result_in_alt, void *stream) {
  if (!ccx::ctx()) return 2;
  int in_alt = 0;
  if (result_in_alt) *result_in_alt = 0;
  if (count <= 1) return 0;
  if (!d_keys || !d_alt) return ccx::set_error(4, "cc_hip_sort_u64_device: null pointer");
  int rc = ccx::sort_u64(d_keys, d_alt, count, &in_alt, ccx::pick_stream(stream));
  if (rc) return rc;
  if (result_in_alt) {
    *result_in_alt = in_alt;
  } else if (in_alt) {
    CCX_TRY(hipMemcpyAsync(d_keys, d_alt, count * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                           ccx::pick_stream(stream)));
  }
  return 0;
}

extern "C" int cc_hip_dedup_u64_device(const uint64_t *d_in, uint64_t count, uint64_t *d_out,
                                       uint64_t *out_count, void *stream) {
  if (!ccx::ctx()) return 2;
  if (!out_count) return ccx::set_error(4, "cc_hip_dedup_u64_device: null count");
  if (count > 0 && (!d_in || !d_out)) return ccx::set_error(4, "cc_hip_dedup_u64_device: null pointer");
  if (d_in == d_out && count > 1)
    return ccx::set_error(4, "cc_hip_dedup_u64_device: input and output must differ");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  return run_dedup(d_in, d_out, count, s, st, out_count);
}

extern "C" int cc_hip_ghost_extract_device(const uint64_t *d_records, uint64_t count,
                                           uint64_t *d_out_records, uint32_t *d_ghost_vertices,
                                           uint64_t *ghost_count, void *stream) {
  if (!ccx::ctx()) return 2;
  if (!ghost_count) return ccx::set_error(4, "cc_hip_ghost_extract_device: null count");
  if (count > 0 && (!d_records || !d_out_records || !d_ghost_vertices))
    return ccx::set_error(4, "cc_hip_ghost_extract_device: null pointer");
  if (d_records == d_out_records && count > 1)
    return ccx::set_error(4, "cc_hip_ghost_extract_device: input and output must differ");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  if (ccx::g_compact_rows)
    return run_compact_rows("ghost_compact", GhostRows{d_records, d_out_records, d_ghost_vertices},
                            count, s, st, ghost_count);
  return run_compact("ghost_compact", GhostOp{d_records, d_out_records, d_ghost_vertices}, count,
                     s, st, ghost_count);
}

extern "C" int cc_hip_merge_prepare_device(const uint64_t *d_records, uint64_t count,
                                           uint32_t ghost_count, const uint32_t *d_reps,
                                           uint32_t rep_count, uint32_t *d_dom_bits,
                                           uint64_t *d_side, uint64_t *side_count,
                                           uint32_t *dominant, void *stream) {
  if (!ccx::ctx()) return 2;
  if (!side_count || !dominant) return ccx::set_error(4, "cc_hip_merge_prepare_device: null output");
  *side_count = 0;
  *dominant = 0;
  const uint64_t words = ((uint64_t)ghost_count + 31) / 32;
  if ((words > 0 && !d_dom_bits) || (count > 0 && (!d_records || !d_side || !d_reps)))
    return ccx::set_error(4, "cc_hip_merge_prepare_device: null pointer");
  if (count > 0 && rep_count == 0)
    return ccx::set_error(4, "cc_hip_merge_prepare_device: records but no reps");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  if (words > 0)
    CCX_LAUNCH("merge_prepare_clear", s,
               k_fill_u32<<<grid_cap(words, 8), kThreads, 0, s>>>(d_dom_bits, words, 0u));
  if (count == 0) return 0;
  uint32_t *d_g = st->d_words + 40;  // the dominant rep, its index, the bad-rep flag
  CCX_TRY(hipMemsetAsync(d_g + 2, 0, sizeof(uint32_t), s));
  CCX_LAUNCH("merge_prepare_mode", s,
             k_mode_rank<<<1, kModeSamples, 0, s>>>(d_records, count, nullptr, 0, d_g));
  uint64_t n_side = 0;
  if (int rc = run_compact_rows(
          "merge_prepare",
          DomRows<32>{d_records, d_g, ghost_count, d_dom_bits, d_side}, count,
          s, st, &n_side))
    return rc;
  if (n_side > 0)
    CCX_LAUNCH("merge_prepare_side", s,
               k_side_rep_index<<<grid_cap(n_side, 8), kThreads, 0, s>>>(d_side, n_side, d_reps,
                                                                         rep_count, d_g + 2));
  CCX_LAUNCH("merge_prepare_index", s, k_rep_index<<<1, 1, 0, s>>>(d_reps, rep_count, d_g, d_g + 1));
  uint32_t *h = (uint32_t *)st->h_stage;
  CCX_TRY(hipMemcpyAsync(h, d_g + 1, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  if (h[0] >= rep_count || h[1] != 0)
    return ccx::set_error(4, "cc_hip_merge_prepare_device: a record's rep is not in the rep list");
  *dominant = h[0];
  *side_count = n_side;
  return 0;
}

extern "C" int cc_hip_local_reps_device(const uint32_t *d_comp_of, uint32_t n_local,
                                        uint32_t *d_reps, uint32_t *rep_count, void *stream) {
  if (!ccx::ctx()) return 2;
  if (!rep_count) return ccx::set_error(4, "cc_hip_local_reps_device: null count");
  *rep_count = 0;
  if (n_local == 0) return 0;
  if (!d_comp_of || !d_reps) return ccx::set_error(4, "cc_hip_local_reps_device: null pointer");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  uint64_t cnt = 0;
  if (int rc = run_compact_rows("local_reps", RepRows{d_comp_of, d_reps}, n_local, s, st, &cnt))
    return rc;
  *rep_count = (uint32_t)cnt;
  return 0;
}

extern "C" int cc_hip_send_lists_device(const uint64_t *d_records, uint64_t count,
                                        uint32_t n_local, const uint32_t *d_bounds, int comm_size,
                                        uint32_t base, int *h_starts, uint32_t *d_vals,
                                        void *stream) {
  if (!ccx::ctx()) return 2;
  if (comm_size < 1 || comm_size > 64)
    return ccx::set_error(4, "cc_hip_send_lists_device: comm_size %d outside 1..64", comm_size);
  if (!h_starts || !d_bounds || (count > 0 && (!d_records || !d_vals)))
    return ccx::set_error(4, "cc_hip_send_lists_device: null pointer");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  h_starts[0] = 0;
  if (count == 0 || n_local == 0) {
    for (int q = 1; q <= comm_size; ++q) h_starts[q] = 0;
    return 0;
  }
  unsigned long long *mask =
      (unsigned long long *)ccx::stream_tmp(st, (size_t)n_local * sizeof(unsigned long long));
  if (!mask) return 3;
  CCX_TRY(hipMemsetAsync(mask, 0, (size_t)n_local * sizeof(unsigned long long), s));
  const uint64_t waves = (count + kWave * kMaskWindows - 1) / (kWave * kMaskWindows);
  const uint64_t cap = (uint64_t)(ccx::ctx()->cu_count > 0 ? ccx::ctx()->cu_count : 256) * 8 * kWaves;
  const unsigned grid = (unsigned)(((waves < cap ? waves : cap) + kWaves - 1) / kWaves);
  CCX_LAUNCH("owner_mask", s,
             k_owner_mask<<<grid, kThreads, 0, s>>>(d_records, count, d_bounds, comm_size, n_local,
                                                   mask));
  uint64_t total = 0;
  for (int q = 0; q < comm_size; ++q) {
    uint64_t cnt = 0;
    const int rc = ccx::g_compact_rows
                       ? run_compact_rows("send_compact", MaskRows{mask, d_vals + total, base, q},
                                          n_local, s, st, &cnt)
                       : run_compact("send_compact", MaskOp{mask, d_vals + total, base, q}, n_local,
                                     s, st, &cnt);
    if (rc) return rc;
    total += cnt;
    if (total > 0x7FFFFFFFull)
      return ccx::set_error(4, "cc_hip_send_lists_device: lists exceed int offsets");
    h_starts[q + 1] = (int)total;
  }
  return 0;
}

namespace {
// Adjacent-unique over packed keys (remote << bits) | rank into
// (remote << 32) | rep, in the layout the knobs select.
int dedup_unpack(const uint64_t *in, uint64_t *out, uint64_t count, const uint32_t *rep_of_rank,
                 int bits, hipStream_t s, ccx::StreamState *st, uint64_t *kept) {
  if (ccx::g_compact_rows)
    return run_compact_rows("dedup_compact", DedupUnpackRows{in, out, rep_of_rank, bits}, count, s,
                            st, kept);
  if (ccx::g_dedup_items == 32)
    return run_compact("dedup_compact", DedupUnpackOp<32>{in, out, rep_of_rank, bits}, count, s, st,
                       kept);
  return run_compact("dedup_compact", DedupUnpackOp<>{in, out, rep_of_rank, bits}, count, s, st,
                     kept);
}

unsigned grid_cap(uint64_t threads, int per_cu) {
  const uint64_t need = (threads + kThreads - 1) / kThreads;
  const ccx::Context *c = ccx::ctx();
  const uint64_t cap = (uint64_t)(c && c->cu_count > 0 ? c->cu_count : 256) * per_cu;
  return (unsigned)(need == 0 ? 1 : (need < cap ? need : cap));
}

// Boundary build with the dominant rep split off (see SplitRows):
// records of the most frequent rep set bits of a bitmap over the remotes,
// the others are packed, sorted and deduplicated, and the set bits, read
// in order, are merged with them.  On return the result is in d_records or
// d_alt (*result_in_alt), *out_count records.  `packed` holds each local
// vertex's rep as its rank among the roots, rep_of_rank the inverse.
int build_split_dominant(uint64_t *d_records, uint64_t *d_alt, uint64_t count,
                         const uint32_t *packed, const uint32_t *rep_of_rank, uint32_t n_local,
                         int bits, hipStream_t s, ccx::StreamState *st, uint64_t *out_count,
                         int *result_in_alt) {
  unsigned long long *d_ao = (unsigned long long *)(st->d_words + 32);  // AND, OR
  uint32_t *d_g = st->d_words + 38;                                      // dominant rank
  const uint64_t init[2] = {~0ull, 0ull};
  memcpy(st->h_stage, init, sizeof(init));
  CCX_TRY(hipMemcpyAsync(d_ao, st->h_stage, sizeof(init), hipMemcpyHostToDevice, s));
  CCX_LAUNCH("sort_and_or", s, k_and_or<<<grid_cap((count + kAndOrItems - 1) / kAndOrItems, 8), kThreads, 0, s>>>(d_records, count, d_ao));
  CCX_LAUNCH("mode_rep", s,
             k_mode_rank<<<1, kModeSamples, 0, s>>>(d_records, count, packed, n_local, d_g));
  CCX_TRY(hipMemcpyAsync(st->h_stage, d_ao, sizeof(init), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  // the OR's high half bounds every remote
  const uint64_t words = (st->h_stage[1] >> 37) + 1;
  const size_t map_bytes = (words * 4 + 7) & ~(size_t)7;
  char *bm = nullptr;
  const size_t bm_bytes = 2 * map_bytes + (words + 1) * 8;
  if (hipMalloc(&bm, bm_bytes) != hipSuccess)
    return ccx::set_error(3, "cc_hip_boundary_build_device: hipMalloc of %zu bytes failed", bm_bytes);
  uint32_t *bitmap = (uint32_t *)bm;
  uint32_t *wcnt = (uint32_t *)(bm + map_bytes);
  uint64_t *woff = (uint64_t *)(bm + 2 * map_bytes);
  int rc = 0;
  // partitioned split: buckets of 2^shift remotes, at most kPartMaxBuckets
  // of them (else the probing split runs)
  const uint64_t max_remote = st->h_stage[1] >> 32;
  int rbits = 1;
  while (rbits < 64 && (max_remote >> rbits) != 0) ++rbits;
  int shift = rbits - 8;
  if (shift < 5) shift = 5;
  if (shift > kPartMaxShift) shift = kPartMaxShift;
  const uint64_t nb = (words * 32 + (1ull << shift) - 1) >> shift;
  const bool part = g_boundary_split == 2 && nb <= kPartMaxBuckets;
  const uint64_t tiles = (count + kPartTile - 1) / kPartTile;
  char *pt = nullptr;
  uint32_t *cnt_dom = nullptr, *cnt_bucket = nullptr, *cnt_other = nullptr;
  uint64_t *off_dom = nullptr, *off_tile = nullptr, *off_other = nullptr;
  if (part) {
    // counts tile-major and bucket-major (u32), offsets bucket-major (with
    // the total) and tile-major (u64), the others' counts and offsets
    const size_t a = (nb * tiles * 4 + 7) & ~(size_t)7;
    const size_t b2 = (tiles * 4 + 7) & ~(size_t)7;
    const size_t pt_bytes = 2 * a + b2 + (nb * tiles + 1) * 8 + nb * tiles * 8 + (tiles + 1) * 8;
    if (hipMalloc(&pt, pt_bytes) != hipSuccess) {
      (void)hipFree(bm);
      return ccx::set_error(3, "cc_hip_boundary_build_device: hipMalloc of %zu bytes failed", pt_bytes);
    }
    cnt_dom = (uint32_t *)pt;
    cnt_bucket = (uint32_t *)(pt + a);
    cnt_other = (uint32_t *)(pt + 2 * a);
    off_dom = (uint64_t *)(pt + 2 * a + b2);
    off_tile = off_dom + nb * tiles + 1;
    off_other = off_tile + nb * tiles;
  }
  do {
    uint64_t n_other = 0;
    if (part) {
      const uint64_t gcap = grid_cap(tiles * kThreads, 8);
      const unsigned gt = (unsigned)(tiles < gcap ? tiles : gcap);
      CCX_LAUNCH_RC(rc, "split_part_count", s,
                    k_part_count<<<gt, kThreads, 0, s>>>(d_records, count, packed, n_local, d_g, shift,
                                                         (uint32_t)nb, tiles, cnt_dom, cnt_other));
      if (rc) break;
      const unsigned gtr = (unsigned)grid_cap(((tiles + 31) / 32) * ((nb + 31) / 32) * kThreads, 8);
      CCX_LAUNCH_RC(rc, "split_part_transpose", s,
                    k_transpose<uint32_t><<<gtr, kThreads, 0, s>>>(cnt_dom, tiles, nb, cnt_bucket));
      if (rc) break;
      if ((rc = cc_hip_exclusive_scan_u32_device(cnt_bucket, nb * tiles, off_dom, s))) break;
      CCX_LAUNCH_RC(rc, "split_part_transpose", s,
                    k_transpose<uint64_t><<<gtr, kThreads, 0, s>>>(off_dom, nb, tiles, off_tile));
      if (rc) break;
      if ((rc = cc_hip_exclusive_scan_u32_device(cnt_other, tiles, off_other, s))) break;
      uint64_t *h = st->h_stage;
      if (hipMemcpyAsync(&h[0], off_dom + nb * tiles, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(&h[1], off_other + tiles, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = ccx::set_error(1, "cc_hip_boundary_build_device: reading the split counts failed");
        break;
      }
      const uint64_t n_dom = h[0];
      n_other = h[1];
      if (n_dom + n_other != count) {
        rc = ccx::set_error(1, "cc_hip_boundary_build_device: split counted %llu + %llu of %llu records",
                            (unsigned long long)n_dom, (unsigned long long)n_other,
                            (unsigned long long)count);
        break;
      }
      // the others fill d_alt[0, n_other); the dominant remotes (4 bytes
      // each) go right after them, which fits: 8 n_other + 4 n_dom <= 8 count
      uint32_t *remotes = (uint32_t *)(d_alt + n_other);
      CCX_LAUNCH_RC(rc, "split_part_scatter", s,
                    k_part_scatter<<<gt, kPartScatterThreads, 0, s>>>(
                        d_records, count, packed, n_local, d_g, shift, (uint32_t)nb, tiles, bits,
                        cnt_dom, off_tile, off_other, remotes, d_alt));
      if (rc) break;
      CCX_LAUNCH_RC(rc, "split_part_bitmap", s,
                    k_part_bitmap<<<(unsigned)nb, kPartBitmapThreads, 0, s>>>(
                        remotes, off_dom, tiles, (uint32_t)(1u << (shift - 5)), words, bitmap));
      if (rc) break;
    } else if (hipMemsetAsync(bitmap, 0, words * 4, s) != hipSuccess) {
      rc = ccx::set_error(1, "cc_hip_boundary_build_device: bitmap clear failed");
      break;
    } else if (g_split_dominant_rows == 16)
      rc = run_compact_rows("split_dominant",
                            SplitRows<16>{d_records, packed, n_local, bits, d_g, bitmap, d_alt},
                            count, s, st, &n_other);
    else
      rc = run_compact_rows("split_dominant",
                            SplitRows<32>{d_records, packed, n_local, bits, d_g, bitmap, d_alt},
                            count, s, st, &n_other);
    if (rc) break;
    if (n_other > count) {
      rc = ccx::set_error(1, "cc_hip_boundary_build_device: %llu others from %llu records",
                          (unsigned long long)n_other, (unsigned long long)count);
      break;
    }
    // the others: sort (d_alt <-> d_records; every raw record has been read)
    // and dedup from S into T
    uint64_t *S = d_alt, *T = d_records;
    if (n_other > 1) {
      int in_alt = 0;
      if ((rc = ccx::sort_u64(d_alt, d_records, n_other, &in_alt, s))) break;
      if (in_alt) {
        S = d_records;
        T = d_alt;
      }
    }
    uint64_t n_o = 0;
    if ((rc = dedup_unpack(S, T, n_other, rep_of_rank, bits, s, st, &n_o))) break;
    // the dominant rep's pairs, in remote order, after the others in T (they
    // fit: n_o <= n_other and the set bits are at most count - n_other)
    CCX_LAUNCH_RC(rc, "bitmap_counts", s,
                  k_word_counts<<<grid_cap(words, 8), kThreads, 0, s>>>(bitmap, words, wcnt));
    if (rc) break;
    if ((rc = cc_hip_exclusive_scan_u32_device(wcnt, words, woff, s))) break;
    if (hipMemcpyAsync(st->h_stage, woff + words, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = ccx::set_error(1, "cc_hip_boundary_build_device: reading the bitmap count failed");
      break;
    }
    const uint64_t n_g = st->h_stage[0];
    if (n_o + n_g > count) {
      rc = ccx::set_error(1, "cc_hip_boundary_build_device: %llu + %llu pairs from %llu records",
                          (unsigned long long)n_o, (unsigned long long)n_g,
                          (unsigned long long)count);
      break;
    }
    CCX_LAUNCH_RC(rc, "bitmap_emit", s,
                  k_emit_bits<<<grid_cap(words, 8), kThreads, 0, s>>>(bitmap, words, woff, rep_of_rank,
                                                                   d_g, T + n_o));
    if (rc) break;
    uint64_t *result = T;
    if (n_o > 0 && n_g > 0) {
      const uint64_t tiles = (n_o + n_g + kMergeTile - 1) / kMergeTile;
      const uint64_t cap = (uint64_t)(ccx::ctx()->cu_count > 0 ? ccx::ctx()->cu_count : 256) * 2;
      CCX_LAUNCH_RC(rc, "bitmap_merge", s,
                    k_merge_tiles<<<(unsigned)(tiles < cap ? tiles : cap), kThreads, 0, s>>>(
                        T + n_o, n_g, T, n_o, S));
      if (rc) break;
      result = S;
    }
    *out_count = n_o + n_g;
    *result_in_alt = (result == d_alt);
  } while (0);
  (void)hipStreamSynchronize(s);
  (void)hipFree(bm);
  if (pt) (void)hipFree(pt);
  return rc;
}
}  // namespace

extern "C" int cc_hip_boundary_build_device(uint64_t *d_records, uint64_t *d_alt, uint64_t count,
                                            const uint32_t *d_comp_of, uint32_t n_local,
                                            uint64_t *out_count, int *result_in_alt,
                                            void *stream) {
  ccx::Context *c = ccx::ctx();
  if (!c) return 2;
  if (!out_count || !result_in_alt)
    return ccx::set_error(4, "cc_hip_boundary_build_device: null output");
  *out_count = 0;
  *result_in_alt = 0;
  if (count == 0) return 0;
  if (!d_records || !d_alt || !d_comp_of || n_local == 0)
    return ccx::set_error(4, "cc_hip_boundary_build_device: null pointer or no local vertices");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  // tables: flags/packed u32[n], scan u64[n + 1], rep_of_rank u32[n]
  const size_t n = n_local;
  char *tab = nullptr;
  const size_t bytes = n * 4 + (n + 1) * 8 + n * 4 + 256;
  if (hipMalloc(&tab, bytes) != hipSuccess)
    return ccx::set_error(3, "cc_hip_boundary_build_device: hipMalloc of %zu bytes failed", bytes);
  uint32_t *flags = (uint32_t *)tab;
  uint64_t *scan = (uint64_t *)(tab + ((n * 4 + 7) & ~(size_t)7));
  uint32_t *rep_of_rank = (uint32_t *)(scan + n + 1);
  int rc = 0;
  do {
    const unsigned g = (unsigned)((n + kThreads - 1) / kThreads < 8192 ? (n + kThreads - 1) / kThreads : 8192);
    CCX_LAUNCH_RC(rc, "rep_flags", s, k_root_flags<<<g, kThreads, 0, s>>>(d_comp_of, n_local, flags));
    if (rc) break;
    if ((rc = cc_hip_exclusive_scan_u32_device(flags, n, scan, s))) break;
    // flags is rewritten in place as the packed rep of every local vertex
    CCX_LAUNCH_RC(rc, "rep_tables", s,
                  k_rep_tables<<<g, kThreads, 0, s>>>(d_comp_of, n_local, scan, flags, rep_of_rank));
    if (rc) break;
    if (hipMemcpyAsync(st->h_stage, scan + n, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = ccx::set_error(1, "cc_hip_boundary_build_device: reading the root count failed");
      break;
    }
    const uint64_t roots = st->h_stage[0];
    int bits = 1;
    while (bits < 32 && (1ull << bits) < roots) ++bits;
    if (g_boundary_split) {
      rc = build_split_dominant(d_records, d_alt, count, flags, rep_of_rank, n_local, bits, s, st,
                                out_count, result_in_alt);
      break;
    }
    const uint64_t gb = (count + kThreads - 1) / kThreads;
    CCX_LAUNCH_RC(rc, "boundary_set_rep", s,
                  k_set_rep_packed<<<(unsigned)(gb < 65536 ? gb : 65536), kThreads, 0, s>>>(
                      d_records, count, flags, n_local, bits));
    if (rc) break;
    int in_alt = 0;
    if ((rc = ccx::sort_u64(d_records, d_alt, count, &in_alt, s))) break;
    uint64_t *sorted = in_alt ? d_alt : d_records, *other = in_alt ? d_records : d_alt;
    uint64_t kept = 0;
    if ((rc = dedup_unpack(sorted, other, count, rep_of_rank, bits, s, st, &kept))) break;
    *out_count = kept;
    *result_in_alt = (other == d_alt);
  } while (0);
  (void)hipStreamSynchronize(s);
  (void)hipFree(tab);
  return rc;
}

// ---- host drop-ins (include/boundary_sort.h) --------------------------------

// Both run on a host slot of their own (stream and workspace), so calls
// from several host threads at once share no device memory.
static int host_sort(BoundaryEdge *edges, uint64_t count) {
  if (!ccx::ctx()) return 2;
  ccx::SlotGuard g;
  if (!g.slot) return 3;
  hipStream_t s = g.slot->stream;
  const size_t bytes = count * sizeof(uint64_t);
  char *ws = (char *)ccx::slot_workspace(g.slot, 2 * bytes + 256);
  if (!ws) return 3;
  uint64_t *a = (uint64_t *)ws;
  uint64_t *b = (uint64_t *)(ws + ((bytes + 255) & ~(size_t)255));
  CCX_TRY(hipMemcpyAsync(a, edges, bytes, hipMemcpyHostToDevice, s));
  int in_alt = 0;
  if (int rc = ccx::sort_u64(a, b, count, &in_alt, s)) return rc;
  CCX_TRY(hipMemcpyAsync(edges, in_alt ? b : a, bytes, hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  return 0;
}

static int host_dedup(BoundaryEdge *edges, uint64_t count, uint64_t *out) {
  if (!ccx::ctx()) return 2;
  ccx::SlotGuard g;
  if (!g.slot) return 3;
  hipStream_t s = g.slot->stream;
  const size_t bytes = count * sizeof(uint64_t);
  char *ws = (char *)ccx::slot_workspace(g.slot, 2 * bytes + 256);
  if (!ws) return 3;
  uint64_t *a = (uint64_t *)ws;
  uint64_t *b = (uint64_t *)(ws + ((bytes + 255) & ~(size_t)255));
  CCX_TRY(hipMemcpyAsync(a, edges, bytes, hipMemcpyHostToDevice, s));
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  if (int rc = run_dedup(a, b, count, s, st, out)) return rc;
  CCX_TRY(hipMemcpyAsync(edges, b, *out * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  return 0;
}

// Drop-in for include/boundary_sort.h:8 (reference src/boundary_sort.c:93-98).
extern "C" void boundary_edges_sort(BoundaryEdge *edges, uint64_t count) {
  if (!edges || count <= 1) return;
  if (int rc = host_sort(edges, count)) {
    fprintf(stderr, "boundary_edges_sort: %s (rc=%d)\n", cc_hip_last_error(), rc);
    exit(EXIT_FAILURE);
  }
}

// Drop-in for include/boundary_sort.h:9 (reference src/boundary_sort.c:100-119).
extern "C" uint64_t boundary_edges_dedup(BoundaryEdge *edges, uint64_t count) {
  if (!edges || count <= 1) return count;
  uint64_t out = count;
  if (int rc = host_dedup(edges, count, &out)) {
    fprintf(stderr, "boundary_edges_dedup: %s (rc=%d)\n", cc_hip_last_error(), rc);
    exit(EXIT_FAILURE);
  }
  return out;
}
