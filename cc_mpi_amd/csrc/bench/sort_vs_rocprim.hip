// The boundary sort (cc_hip_sort_u64_device) against rocPRIM's
// radix_sort_keys on the same device keys: uniform 50-bit keys from a
// counter hash, 7 non-constant 8-bit digits, the shape of one rank's packed
// boundary records.  Both sorts run 4 times; the best time of each is
// printed (HIP events on the null stream around the whole call, so ours
// includes its host wait for the digit plan), and the two outputs are
// compared key for key on the device.
//
//   cc_mpi_amd/lib/sort_vs_rocprim [count]      (default 638,000,000)
//
// Exit status 0 when the outputs agree.

#include <cstring>  // rocPRIM's texture_cache_iterator uses memset

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../../../include/cc_hip.h"

namespace {

constexpr int kBits = 50;
constexpr int kRuns = 4;

__global__ void k_keys(uint64_t *k, uint64_t n, uint64_t salt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = (i + salt) * 0x9E3779B97F4A7C15ull;  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    k[i] = (z ^ (z >> 31)) & ((1ull << kBits) - 1);
  }
}

__global__ void k_count_diff(const uint64_t *a, const uint64_t *b, uint64_t n,
                             unsigned long long *bad) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (a[i] != b[i]) atomicAdd(bad, 1ull);
}

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main(int argc, char **argv) {
  const uint64_t n = argc > 1 ? strtoull(argv[1], nullptr, 10) : 638000000ull;
  if (n < 2) {
    fprintf(stderr, "count must be at least 2\n");
    return 2;
  }
  if (cc_hip_init(0) != 0) {
    fprintf(stderr, "cc_hip_init: %s\n", cc_hip_last_error());
    return 1;
  }
  uint64_t *src = nullptr, *a = nullptr, *b = nullptr, *ref = nullptr;
  void *tmp = nullptr;
  unsigned long long *bad = nullptr;
  if (!ok(hipMalloc(&src, n * 8), "hipMalloc") || !ok(hipMalloc(&a, n * 8), "hipMalloc") ||
      !ok(hipMalloc(&b, n * 8), "hipMalloc") || !ok(hipMalloc(&ref, n * 8), "hipMalloc") ||
      !ok(hipMalloc(&bad, 8), "hipMalloc"))
    return 1;
  k_keys<<<4096, 256>>>(src, n, 12345);
  size_t tmp_bytes = 0;
  if (!ok(rocprim::radix_sort_keys(nullptr, tmp_bytes, src, ref, n, 0, kBits), "rocprim size") ||
      !ok(hipMalloc(&tmp, tmp_bytes), "hipMalloc") || !ok(hipDeviceSynchronize(), "keys"))
    return 1;

  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return 1;
  float best_rp = 1e30f, best_ours = 1e30f, ms = 0.f;
  for (int r = 0; r < kRuns; ++r) {
    if (!ok(hipEventRecord(e0, 0), "record") ||
        !ok(rocprim::radix_sort_keys(tmp, tmp_bytes, src, ref, n, 0, kBits), "rocprim sort") ||
        !ok(hipEventRecord(e1, 0), "record") || !ok(hipEventSynchronize(e1), "sync") ||
        !ok(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
      return 1;
    if (ms < best_rp) best_rp = ms;
  }
  int in_alt = 0;
  for (int r = 0; r < kRuns; ++r) {
    if (!ok(hipMemcpy(a, src, n * 8, hipMemcpyDeviceToDevice), "copy") ||
        !ok(hipDeviceSynchronize(), "sync") || !ok(hipEventRecord(e0, 0), "record"))
      return 1;
    if (cc_hip_sort_u64_device(a, b, n, &in_alt, nullptr) != 0) {
      fprintf(stderr, "cc_hip_sort_u64_device: %s\n", cc_hip_last_error());
      return 1;
    }
    if (!ok(hipDeviceSynchronize(), "sync") || !ok(hipEventRecord(e1, 0), "record") ||
        !ok(hipEventSynchronize(e1), "sync") || !ok(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
      return 1;
    if (ms < best_ours) best_ours = ms;
  }
  unsigned long long h_bad = 0;
  if (!ok(hipMemset(bad, 0, 8), "memset")) return 1;
  k_count_diff<<<4096, 256>>>(in_alt ? b : a, ref, n, bad);
  if (!ok(hipMemcpy(&h_bad, bad, 8, hipMemcpyDeviceToHost), "copy")) return 1;
  printf("%llu keys of %d bits: rocPRIM radix_sort_keys %.2f ms, cc_hip_sort_u64_device %.2f ms "
         "(best of %d), %llu keys differ\n",
         (unsigned long long)n, kBits, best_rp, best_ours, kRuns, h_bad);
  for (void *p : {tmp, (void *)bad, (void *)ref, (void *)b, (void *)a, (void *)src})
    (void)hipFree(p);
  return h_bad == 0 ? 0 : 1;
}
