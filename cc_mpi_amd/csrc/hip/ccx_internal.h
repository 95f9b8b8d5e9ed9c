// Internal helpers shared by the HIP translation units of libccmi355x.so.
// Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ccx {

// Record a failure message for cc_hip_last_error() and return its code.
int set_error(int code, const char *fmt, ...);
void clear_error();

// Lazily initialised per-process state for the current device.
struct Context {
  int device = -1;
  hipStream_t stream = nullptr;  // library stream, default priority
  // Grow-only device workspace reused across calls (never shrinks).
  void *ws = nullptr;
  size_t ws_bytes = 0;
  hipStream_t ws_stream = nullptr;  // stream of the workspace's last user
  // Small device scratch: counters, sampled giant root, flags.
  uint32_t *d_small = nullptr;
  // Host-pinned mirror of a few counters read back after a launch.
  uint32_t *h_small = nullptr;
  // Radix-sort look-back status words, epoch-tagged so they never need
  // clearing between passes.
  uint64_t *d_status = nullptr;
  size_t status_words = 0;
  uint32_t epoch = 0;
  int cu_count = 0;
};

// Returns the context, initialising device 0 on first use.  nullptr on
// failure (error recorded).
Context *ctx();
// Grow the workspace to at least `bytes` for use on stream `s`; if its last
// user was another stream, waits for that stream first.  Returns nullptr on
// failure.
void *workspace(size_t bytes, hipStream_t s);
hipStream_t pick_stream(void *stream);

// Kernel timing: call before/after a launch; no-ops unless enabled.
void prof_begin(const char *name, hipStream_t s);
void prof_end(hipStream_t s);

}  // namespace ccx

#define CCX_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return ccx::set_error(1, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                            __FILE__, __LINE__);                                   \
  } while (0)

#define CCX_LAUNCH(name, stream, ...)                                              \
  do {                                                                             \
    ccx::prof_begin(name, stream);                                                 \
    __VA_ARGS__;                                                                   \
    ccx::prof_end(stream);                                                         \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      return ccx::set_error(1, "launch of %s failed: %s", name, hipGetErrorString(_e)); \
  } while (0)
