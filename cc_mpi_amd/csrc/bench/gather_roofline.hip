// Random-gather ceiling of the memory system, the roofline that bounds the
// solve's hooks: sum p[idx[i]] over K indices into an array of N 32-bit
// words, for arrays from 16 MB (inside the MALL) to 1 GiB, with the index
// stream read non-temporally.  Each gather of a random word costs one
// 128-byte line, so gathers/s x 128 B is the fabric traffic it drives.
// Also: dependent chains (q = p[q], as a root walk), strided gathers (one
// word every `stride` words, like the first entries of CSR rows) and a
// plain streaming read for comparison.  Every index is below N (checked on
// the host before the first launch).  Not part of the product build:
// `make -C cc_mpi_amd/csrc bench`, then run ../lib/gather_roofline.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                            \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);     \
      exit(1);                                                                           \
    }                                                                                    \
  } while (0)

namespace {

constexpr int kBlock = 256;
constexpr int kIlp = 4;

template <bool Nt>
__device__ __forceinline__ uint32_t load(const uint32_t *p) {
  if constexpr (Nt) return __builtin_nontemporal_load(p);
  else return *p;
}

// acc += p[idx[i]]; kIlp gathers in flight per thread
template <bool Nt>
__global__ __launch_bounds__(kBlock) void k_gather(const uint32_t *__restrict__ p,
                                                   const uint32_t *__restrict__ idx, uint64_t k,
                                                   uint32_t *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t acc = 0;
  for (; i + (kIlp - 1) * stride < k; i += kIlp * stride) {
    uint32_t j[kIlp];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) j[u] = __builtin_nontemporal_load(&idx[i + u * stride]);
#pragma unroll
    for (int u = 0; u < kIlp; ++u) acc += load<Nt>(&p[j[u]]);
  }
  for (; i < k; i += stride) acc += load<Nt>(&p[__builtin_nontemporal_load(&idx[i])]);
  out[(uint64_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}

// `hops` dependent steps q = p[q] from each start idx[i] (p holds indices < N)
__global__ __launch_bounds__(kBlock) void k_chase(const uint32_t *__restrict__ p,
                                                  const uint32_t *__restrict__ idx, uint64_t k,
                                                  int hops, uint32_t *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint32_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < k; i += stride) {
    uint32_t q = __builtin_nontemporal_load(&idx[i]);
    for (int h = 0; h < hops; ++h) q = p[q];
    acc += q;
  }
  out[(uint64_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}

// acc += p[i * step] for i < k (k * step <= N): one word per `step` words
__global__ __launch_bounds__(kBlock) void k_strided(const uint32_t *__restrict__ p, uint64_t k,
                                                    uint64_t step, uint32_t *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t acc = 0;
  for (; i + (kIlp - 1) * stride < k; i += kIlp * stride) {
    uint32_t v[kIlp];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) v[u] = p[(i + u * stride) * step];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) acc += v[u];
  }
  for (; i < k; i += stride) acc += p[i * step];
  out[(uint64_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}

uint64_t g_rng = 88172645463325252ull;
uint64_t next_rand() {
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return g_rng;
}

}  // namespace

int main() {
  const uint64_t kMaxWords = 1ull << 28;  // 1 GiB
  const uint64_t kGathers = 1ull << 26;
  int dev = 0, cus = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = (unsigned)cus * 8;
  std::vector<uint32_t> h_p(kMaxWords), h_idx(kGathers);
  uint32_t *d_p, *d_idx, *d_out;
  CK(hipMalloc(&d_p, kMaxWords * 4));
  CK(hipMalloc(&d_idx, kGathers * 4));
  CK(hipMalloc(&d_out, (size_t)grid * kBlock * 4));
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  auto best_ms = [&](auto launch) {
    launch();
    CK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
      CK(hipEventRecord(a, 0));
      launch();
      CK(hipEventRecord(b, 0));
      CK(hipEventSynchronize(b));
      float ms = 0;
      CK(hipEventElapsedTime(&ms, a, b));
      if (ms < best) best = ms;
    }
    return best;
  };
  printf("test,array_bytes,count,us,rate_G_per_s,GB_per_s\n");
  for (uint64_t n : {1ull << 22, 1ull << 24, 1ull << 26, 1ull << 28}) {
    for (uint64_t v = 0; v < n; ++v) h_p[v] = (uint32_t)(next_rand() % n);
    for (uint64_t i = 0; i < kGathers; ++i) h_idx[i] = (uint32_t)(next_rand() % n);
    for (uint64_t v = 0; v < n; ++v)
      if (h_p[v] >= n) return 2;
    for (uint64_t i = 0; i < kGathers; ++i)
      if (h_idx[i] >= n) return 2;
    CK(hipMemcpy(d_p, h_p.data(), n * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_idx, h_idx.data(), kGathers * 4, hipMemcpyHostToDevice));
    auto line = [&](const char *name, uint64_t count, float ms, double bytes_per) {
      const double rate = count / (ms * 1e-3) / 1e9;
      printf("%s,%llu,%llu,%.1f,%.2f,%.0f\n", name, (unsigned long long)(n * 4),
             (unsigned long long)count, ms * 1e3, rate, rate * bytes_per);
      fflush(stdout);
    };
    line("random_gather", kGathers,
         best_ms([&] { k_gather<false><<<grid, kBlock>>>(d_p, d_idx, kGathers, d_out); }), 128.0);
    line("random_gather_nt", kGathers,
         best_ms([&] { k_gather<true><<<grid, kBlock>>>(d_p, d_idx, kGathers, d_out); }), 128.0);
    line("chase_4_hops", kGathers,
         best_ms([&] { k_chase<<<grid, kBlock>>>(d_p, d_idx, kGathers / 4, 4, d_out); }), 128.0);
    for (uint64_t step : {16ull, 32ull, 64ull}) {
      const uint64_t k = n / step;
      char name[32];
      snprintf(name, sizeof(name), "strided_%llu", (unsigned long long)step);
      line(name, k, best_ms([&] { k_strided<<<grid, kBlock>>>(d_p, k, step, d_out); }), 128.0);
    }
    line("stream_read", n, best_ms([&] { k_strided<<<grid, kBlock>>>(d_p, n, 1, d_out); }), 4.0);
  }
  return 0;
}
