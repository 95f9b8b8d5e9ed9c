// Device helpers shared by the look-back kernels (radix sort, ordered
// compaction): wave64 scans and the epoch-tagged status words.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ccx_internal.h"

namespace ccx {

constexpr int kWave = 64;

__device__ __forceinline__ uint64_t tag_word(uint32_t epoch, uint64_t flag, uint64_t value) {
  return ((uint64_t)epoch << kEpochShift) | flag | (value & kValueMask);
}

__device__ __forceinline__ void status_store(uint64_t *p, uint64_t w) {
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t status_load(uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back over the status words of tiles tile-1, tile-2, ...
// (stride `stride` words apart, starting at `col`): sums aggregates until an
// inclusive prefix is found.  Tiles are handed out in order by an atomic
// counter, so every predecessor is already running and publishes its
// aggregate without waiting on anyone: the spin always ends.
__device__ __forceinline__ uint64_t lookback(uint64_t *status, uint64_t tile, size_t stride,
                                             size_t col, uint32_t epoch) {
  uint64_t excl = 0;
  int64_t t = (int64_t)tile - 1;
  while (t >= 0) {
    const uint64_t w = status_load(&status[(size_t)t * stride + col]);
    if ((uint32_t)(w >> kEpochShift) != epoch) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += w & kValueMask;
    if (w & kFlagInclusive) break;
    --t;
  }
  return excl;
}

// The same look-back done by a whole wave (call it with all 64 lanes active;
// every lane gets the result): lane i reads the status of tile hi - i, so
// each step inspects 64 predecessors with one coalesced load instead of
// one.  Nearest first: the run of published words before the first
// unpublished one is summed, up to and including the nearest inclusive
// prefix if it lies in that run.  When thousands of tiles reach their
// look-back together most predecessors hold only aggregates, and a
// one-word-per-step walk is a chain of dependent L2 loads as long as the
// distance to the nearest inclusive prefix.  Tile 0 always publishes an
// inclusive prefix, so a lane with no tile (hi - i < 0) is only ever seen
// past it.
__device__ __forceinline__ uint64_t lookback_wave(uint64_t *status, uint64_t tile, size_t stride,
                                                  size_t col, uint32_t epoch, int lane) {
  uint64_t excl = 0;
  int64_t hi = (int64_t)tile - 1;  // nearest predecessor not yet summed
  for (;;) {
    const int64_t t = hi - lane;
    uint64_t w = 0;
    if (t >= 0) w = status_load(&status[(size_t)t * stride + col]);
    const bool ready = t >= 0 && (uint32_t)(w >> kEpochShift) == epoch;
    const uint64_t incl = __ballot(ready && (w & kFlagInclusive) != 0);
    const uint64_t unready = __ballot(!ready);
    const int first_in = incl ? __builtin_ctzll(incl) : kWave;
    const int first_un = unready ? __builtin_ctzll(unready) : kWave;
    const int take = first_in < first_un ? first_in + 1 : first_un;
    uint64_t v = lane < take ? (w & kValueMask) : 0;
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    excl += v;
    if (first_in < first_un) return excl;
    if (take == 0) __builtin_amdgcn_s_sleep(1);
    hi -= take;
  }
}

// Inclusive prefix sum across the 64 lanes of a wave.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const T up = __shfl_up(x, d, kWave);
    if (lane >= d) x += up;
  }
  return x;
}

}  // namespace ccx
