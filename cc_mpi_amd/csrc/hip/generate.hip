// Device-side construction of the seeded benchmark graphs.
//
// The pair generators compute, for every pair index, exactly what the host
// generators in host/graphgen.c compute (same SplitMix64 counters, same
// integer thresholds, same double-precision inverse-CDF search), so a graph
// built here is byte-identical to the host-built one; only the speed
// differs.  The CSR builder turns undirected pairs into the symmetric
// simple CSR: both directions as 64-bit keys (u << 32) | v (self-loops
// become an all-ones sentinel), onesweep radix sort, adjacent-unique, then
// row offsets from the run boundaries of u.

#include <hip/hip_runtime.h>

#include <cstring>

#include "ccx_internal.h"
#include "../../../include/cc_hip.h"

namespace {

constexpr int kThreads = 256;

#define GRID_STRIDE(i, n) \
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * kThreads)

unsigned grid_for(uint64_t n) {
  const uint64_t need = (n + kThreads - 1) / kThreads;
  ccx::Context *c = ccx::ctx();
  const uint64_t cap = (uint64_t)((c && c->cu_count > 0) ? c->cu_count : 256) * 16;
  return (unsigned)(need == 0 ? 1 : (need < cap ? need : cap));
}

__device__ __forceinline__ uint64_t rand_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t below(uint64_t r, uint32_t n) {
  return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

__global__ __launch_bounds__(kThreads) void k_gen_er(uint32_t n, uint64_t pairs, uint64_t seed,
                                                     uint32_t *__restrict__ src,
                                                     uint32_t *__restrict__ dst) {
  GRID_STRIDE(i, pairs) {
    src[i] = below(rand_at(seed, 2 * i), n);
    dst[i] = below(rand_at(seed, 2 * i + 1), n);
  }
}

__global__ __launch_bounds__(kThreads) void k_gen_rmat(int scale, uint64_t edges, uint32_t ta,
                                                       uint32_t tab, uint32_t tabc, uint64_t seed,
                                                       uint32_t *__restrict__ src,
                                                       uint32_t *__restrict__ dst) {
  const uint64_t draws = (uint64_t)(scale + 3) / 4;
  GRID_STRIDE(e, edges) {
    uint32_t u = 0, v = 0;
    uint64_t r = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
      if ((lvl & 3) == 0) r = rand_at(seed, e * draws + (uint64_t)(lvl >> 2));
      const uint32_t x = (uint32_t)(r & 0xFFFF);
      r >>= 16;
      const uint32_t q = (x < ta) ? 0 : (x < tab) ? 1 : (x < tabc) ? 2 : 3;
      u = (u << 1) | (q >> 1);
      v = (v << 1) | (q & 1);
    }
    src[e] = u;
    dst[e] = v;
  }
}

__global__ __launch_bounds__(kThreads) void k_gen_grid(uint32_t rows, uint32_t cols, uint32_t thr,
                                                       uint64_t seed, uint32_t *__restrict__ src,
                                                       uint32_t *__restrict__ dst) {
  const uint64_t n = (uint64_t)rows * cols;
  GRID_STRIDE(id, n) {
    const uint32_t r = (uint32_t)(id / cols), c = (uint32_t)(id % cols);
    const uint64_t rnd = rand_at(seed, id);
    const bool right = (c + 1 < cols) && ((uint32_t)(rnd & 0xFFFF) < thr);
    const bool down = (r + 1 < rows) && ((uint32_t)((rnd >> 16) & 0xFFFF) < thr);
    src[2 * id] = (uint32_t)id;
    dst[2 * id] = right ? (uint32_t)(id + 1) : (uint32_t)id;
    src[2 * id + 1] = (uint32_t)id;
    dst[2 * id + 1] = down ? (uint32_t)(id + cols) : (uint32_t)id;
  }
}

__global__ __launch_bounds__(kThreads) void k_gen_chung_lu(uint32_t n, uint64_t pairs,
                                                           const double *__restrict__ cdf,
                                                           double total, uint64_t seed,
                                                           uint32_t *__restrict__ src,
                                                           uint32_t *__restrict__ dst) {
  GRID_STRIDE(e, 2 * pairs) {
    const double x = (double)(rand_at(seed, e) >> 11) * (1.0 / 9007199254740992.0) * total;
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (cdf[mid] <= x) lo = mid + 1;
      else hi = mid;
    }
    if (e & 1) dst[e >> 1] = lo;
    else src[e >> 1] = lo;
  }
}

// keys[2i], keys[2i+1] = both directions of pair i (renamed by perm);
// pairs out of range or self-loops become the sentinel ~0.
__global__ __launch_bounds__(kThreads) void k_expand(const uint32_t *__restrict__ src,
                                                     const uint32_t *__restrict__ dst,
                                                     uint64_t count, uint32_t n,
                                                     const uint32_t *__restrict__ perm,
                                                     uint64_t *__restrict__ keys) {
  GRID_STRIDE(i, count) {
    uint32_t u = src[i], v = dst[i];
    uint64_t a = ~0ull, b = ~0ull;
    if (u < n && v < n && u != v) {
      if (perm) {
        u = perm[u];
        v = perm[v];
      }
      a = ((uint64_t)u << 32) | v;
      b = ((uint64_t)v << 32) | u;
    }
    keys[2 * i] = a;
    keys[2 * i + 1] = b;
  }
}

// row_ptr from sorted unique keys: every row index in (u[i-1], u[i]] starts
// at i; rows after the last key's row end at m.
__global__ __launch_bounds__(kThreads) void k_rows(const uint64_t *__restrict__ keys, uint64_t m,
                                                   uint32_t n, uint64_t *__restrict__ row_ptr,
                                                   uint32_t *__restrict__ col) {
  GRID_STRIDE(i, m) {
    const uint64_t k = keys[i];
    const uint64_t u = k >> 32;
    const uint64_t prev = (i == 0) ? 0 : (keys[i - 1] >> 32) + 1;
    for (uint64_t r = prev; r <= u; ++r) row_ptr[r] = i;
    col[i] = (uint32_t)k;
    if (i == m - 1)
      for (uint64_t r = u + 1; r <= n; ++r) row_ptr[r] = m;
  }
}

__global__ void k_fill_rows(uint64_t *__restrict__ row_ptr, uint32_t n) {
  GRID_STRIDE(i, (uint64_t)n + 1) row_ptr[i] = 0;
}

}  // namespace

extern "C" int cc_hip_gen_pairs_device(int kind, const uint64_t *params, const double *d_cdf,
                                       uint64_t seed, uint32_t *d_src, uint32_t *d_dst,
                                       uint64_t *count, void *stream) {
  if (!ccx::ctx()) return 2;
  hipStream_t s = ccx::pick_stream(stream);
  switch (kind) {
    case 0: {  // Erdos-Renyi: params = {n, pairs}
      *count = params[1];
      CCX_LAUNCH("gen_er", s,
                 k_gen_er<<<grid_for(params[1]), kThreads, 0, s>>>((uint32_t)params[0], params[1],
                                                                    seed, d_src, d_dst));
      return 0;
    }
    case 1: {  // R-MAT: params = {scale, edge_factor, ta, tab, tabc}
      const int scale = (int)params[0];
      const uint64_t edges = params[1] << scale;
      *count = edges;
      CCX_LAUNCH("gen_rmat", s,
                 k_gen_rmat<<<grid_for(edges), kThreads, 0, s>>>(scale, edges, (uint32_t)params[2],
                                                                  (uint32_t)params[3],
                                                                  (uint32_t)params[4], seed, d_src,
                                                                  d_dst));
      return 0;
    }
    case 2: {  // grid: params = {rows, cols, thr}
      const uint64_t n = params[0] * params[1];
      *count = 2 * n;
      CCX_LAUNCH("gen_grid", s,
                 k_gen_grid<<<grid_for(n), kThreads, 0, s>>>((uint32_t)params[0], (uint32_t)params[1],
                                                             (uint32_t)params[2], seed, d_src,
                                                             d_dst));
      return 0;
    }
    case 3: {  // Chung-Lu: params = {n, pairs, bits of the cdf total}
      double total;
      memcpy(&total, &params[2], sizeof(double));
      *count = params[1];
      CCX_LAUNCH("gen_chung_lu", s,
                 k_gen_chung_lu<<<grid_for(2 * params[1]), kThreads, 0, s>>>(
                     (uint32_t)params[0], params[1], d_cdf, total, seed, d_src, d_dst));
      return 0;
    }
    default:
      return ccx::set_error(4, "cc_hip_gen_pairs_device: unknown kind %d", kind);
  }
}

extern "C" int cc_hip_csr_from_pairs_device(uint32_t n, const uint32_t *d_src,
                                            const uint32_t *d_dst, uint64_t count,
                                            const uint32_t *d_perm, uint64_t *d_keys,
                                            uint64_t *d_alt, uint64_t *d_row_ptr,
                                            uint32_t *d_col, uint64_t *m_out, void *stream) {
  if (!ccx::ctx()) return 2;
  hipStream_t s = ccx::pick_stream(stream);
  *m_out = 0;
  const uint64_t nk = 2 * count;
  if (nk == 0) {
    CCX_LAUNCH("csr_rows", s, k_fill_rows<<<grid_for((uint64_t)n + 1), kThreads, 0, s>>>(d_row_ptr, n));
    return 0;
  }
  CCX_LAUNCH("csr_expand", s,
             k_expand<<<grid_for(count), kThreads, 0, s>>>(d_src, d_dst, count, n, d_perm, d_keys));
  int in_alt = 0;
  if (int rc = cc_hip_sort_u64_device(d_keys, d_alt, nk, &in_alt, s)) return rc;
  uint64_t *sorted = in_alt ? d_alt : d_keys, *uniq = in_alt ? d_keys : d_alt;
  uint64_t m = 0;
  if (int rc = cc_hip_dedup_u64_device(sorted, nk, uniq, &m, s)) return rc;
  uint64_t last = 0;
  CCX_TRY(hipMemcpyAsync(&last, uniq + (m - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  if (last == ~0ull) --m;  // the self-loop sentinel
  if (m == 0) {
    CCX_LAUNCH("csr_rows", s, k_fill_rows<<<grid_for((uint64_t)n + 1), kThreads, 0, s>>>(d_row_ptr, n));
  } else {
    CCX_LAUNCH("csr_rows", s, k_rows<<<grid_for(m), kThreads, 0, s>>>(uniq, m, n, d_row_ptr, d_col));
  }
  CCX_TRY(hipStreamSynchronize(s));
  *m_out = m;
  return 0;
}
