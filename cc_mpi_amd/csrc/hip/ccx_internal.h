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

// State private to one stream, so work enqueued on different streams never
// shares scratch memory.
struct StreamState {
  hipStream_t stream = nullptr;
  // 256 device words: [0] sampled root, [1] frontier link flag, [8..] tile
  // counters and flags of the other passes
  uint32_t *d_words = nullptr;
  uint64_t *d_status = nullptr;  // look-back status words (epoch-tagged)
  size_t status_words = 0;
  uint32_t epoch = 0;           // last epoch handed out; 0 is never valid
  uint64_t *h_stage = nullptr;  // 64 KiB pinned host staging
  void *tmp = nullptr;          // grow-only temp buffer (sort ping-pong etc.)
  size_t tmp_bytes = 0;
};

// Largest vertex count a solve accepts: every id stays below the
// 0xFFFFFFFF "no neighbour" sentinel, and a 32-bit grid-stride loop (stride
// below 2^24 threads) cannot wrap past n.
constexpr uint32_t kMaxVertices = 0xFF000000u;

// Look-back status word layout: [63:42] epoch, [41:40] flag, [39:0] value.
constexpr int kEpochShift = 42;
constexpr uint32_t kEpochMax = (1u << 22) - 1;
constexpr uint64_t kFlagAggregate = 1ull << 40;
constexpr uint64_t kFlagInclusive = 2ull << 40;
constexpr uint64_t kValueMask = (1ull << 40) - 1;

StreamState *stream_state(hipStream_t s);
// Wait for stream s and free its state (call before destroying s).
void release_stream_state(hipStream_t s);
// Grow the status array to `words` (new memory zeroed, epochs restart).
int ensure_status(StreamState *st, size_t words);
// Fresh epoch for one look-back pass; recycles the array on wrap-around.
int next_epoch(StreamState *st, uint32_t *epoch);
// Grow the per-stream temp buffer to at least `bytes`.
void *stream_tmp(StreamState *st, size_t bytes);

// Store strategy of the split and dedup compactions (0: each thread stores
// its own items; 1: staged in LDS, stored as contiguous runs).  Set through
// cc_hip_sort_set.
extern int g_split_staged;
// Ordered compactions (dedup, ghost extraction, send lists, delta pack) in
// row-striped tiles (1) or with a run of consecutive items per thread (0).
extern int g_compact_rows;
extern int g_dedup_staged;
extern int g_dedup_items;  // items per thread of the dedup compaction (16, 32)
// Passes of the delta pack: 3 (count, scan, emit) or 1 (decoupled look-back).
int delta_passes_set(int v);
int delta_passes_get();

// Returns the context, initialising device 0 on first use.  nullptr on
// failure (error recorded).
Context *ctx();
hipStream_t pick_stream(void *stream);

// A host slot: a stream of its own plus a grow-only device workspace, held
// by one call of a host-buffer entry point (the reference-prototype
// drop-ins) at a time.  Concurrent callers get different slots, so they
// share no device memory and no stream state; slots are created on demand
// and kept until the process ends.
struct HostSlot {
  hipStream_t stream = nullptr;
  void *ws = nullptr;
  size_t ws_bytes = 0;
  bool busy = false;
};
// A free slot (created if every slot is busy); nullptr on failure.
HostSlot *acquire_host_slot();
void release_host_slot(HostSlot *slot);
// Grow the slot's workspace to at least `bytes` (waits for the slot's
// stream before freeing the old buffer); nullptr on failure.
void *slot_workspace(HostSlot *slot, size_t bytes);
struct SlotGuard {
  HostSlot *slot;
  SlotGuard() : slot(acquire_host_slot()) {}
  ~SlotGuard() {
    if (slot) release_host_slot(slot);
  }
  SlotGuard(const SlotGuard &) = delete;
  SlotGuard &operator=(const SlotGuard &) = delete;
};

// Kernel timing: call before/after a launch; no-ops unless enabled.
void prof_begin(const char *name, hipStream_t s);
void prof_end(hipStream_t s);
// While one is alive, this thread's launches record no timing events
// (which may create events), whatever cc_hip_prof_enable says.
void prof_mute(bool on);
struct ProfMute {
  ProfMute() { prof_mute(true); }
  ~ProfMute() { prof_mute(false); }
};

// The context if the library is initialised, else nullptr (never
// initialises it, so never allocates).
const Context *ctx_if_ready();
// Sets up the scratch array of the allocation-free solve on `device`
// (current at the call).
int components_noalloc_setup(int device);

}  // namespace ccx

#define CCX_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return ccx::set_error(1, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                            __FILE__, __LINE__);                                   \
  } while (0)

// CCX_LAUNCH for code that must clean up: sets `rc` instead of returning.
#define CCX_LAUNCH_RC(rc, name, stream, ...)                                       \
  do {                                                                             \
    ccx::prof_begin(name, stream);                                                 \
    __VA_ARGS__;                                                                   \
    ccx::prof_end(stream);                                                         \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      rc = ccx::set_error(1, "launch of %s failed: %s", name, hipGetErrorString(_e)); \
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
