// Runtime layer of libccmi355x.so: device context, errors, memory helpers
// and per-kernel HIP-event timing.  One context per process (device chosen
// by cc_hip_init or device 0), one library stream at default priority.

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ccx_internal.h"
#include "../../../include/cc_hip.h"

namespace ccx {

static thread_local char g_err[512];

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = '\0'; }

static std::mutex g_ctx_mu;
static Context g_ctx;
static std::atomic<bool> g_ctx_ok{false};  // set once, after g_ctx is complete

static int init_locked(int device) {
  if (g_ctx_ok) {
    if (device >= 0 && device != g_ctx.device)
      return set_error(2, "cc_hip_init: already initialised on device %d", g_ctx.device);
    return 0;
  }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    return set_error(2, "no HIP device available (%s)",
                     e == hipSuccess ? "count=0" : hipGetErrorString(e));
  if (device < 0) device = 0;
  if (device >= count) return set_error(2, "device %d out of range (%d devices)", device, count);
  CCX_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  CCX_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(2, "device %d is %s; this library is built for gfx950 only", device,
                     prop.gcnArchName);
  g_ctx.device = device;
  g_ctx.cu_count = prop.multiProcessorCount;
  CCX_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  CCX_TRY(hipMalloc(&g_ctx.d_small, 4096));
  CCX_TRY(hipMemset(g_ctx.d_small, 0, 4096));
  CCX_TRY(hipHostMalloc(&g_ctx.h_small, 4096, hipHostMallocDefault));
  if (int rc = components_noalloc_setup(device)) return rc;
  g_ctx_ok.store(true, std::memory_order_release);
  return 0;
}

Context *ctx() {
  if (g_ctx_ok.load(std::memory_order_acquire)) {
    // Every call runs on the context's device, whatever the caller set.
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != g_ctx.device) (void)hipSetDevice(g_ctx.device);
    return &g_ctx;
  }
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (init_locked(-1) != 0) return nullptr;
  return &g_ctx;
}

const Context *ctx_if_ready() {
  if (!g_ctx_ok.load(std::memory_order_acquire)) return nullptr;
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur != g_ctx.device) (void)hipSetDevice(g_ctx.device);
  return &g_ctx;
}

static std::mutex g_slot_mu;
static std::vector<HostSlot *> g_slots;

HostSlot *acquire_host_slot() {
  if (!ctx()) return nullptr;
  std::lock_guard<std::mutex> lk(g_slot_mu);
  for (HostSlot *s : g_slots)
    if (!s->busy) {
      s->busy = true;
      return s;
    }
  HostSlot *s = new HostSlot();
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    set_error(1, "host slot: stream creation failed");
    return nullptr;
  }
  s->busy = true;
  g_slots.push_back(s);
  return s;
}

void release_host_slot(HostSlot *slot) {
  std::lock_guard<std::mutex> lk(g_slot_mu);
  slot->busy = false;
}

void *slot_workspace(HostSlot *slot, size_t bytes) {
  if (bytes <= slot->ws_bytes) return slot->ws;
  if (slot->ws) {
    // work still queued on the slot's stream may use the old buffer
    if (hipStreamSynchronize(slot->stream) != hipSuccess) {
      set_error(1, "slot workspace: synchronize failed");
      return nullptr;
    }
    (void)hipFree(slot->ws);
    slot->ws = nullptr;
    slot->ws_bytes = 0;
  }
  const size_t want = bytes + bytes / 8 + 4096;
  if (hipMalloc(&slot->ws, want) != hipSuccess) {
    slot->ws = nullptr;
    set_error(3, "slot workspace: hipMalloc of %zu bytes failed", want);
    return nullptr;
  }
  slot->ws_bytes = want;
  return slot->ws;
}

static std::mutex g_ss_mu;
static std::map<hipStream_t, StreamState *> g_ss;

StreamState *stream_state(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_ss_mu);
  auto it = g_ss.find(s);
  if (it != g_ss.end()) return it->second;
  StreamState *st = new StreamState();
  st->stream = s;
  if (hipMalloc(&st->d_words, 256 * sizeof(uint32_t)) != hipSuccess ||
      hipMemset(st->d_words, 0, 256 * sizeof(uint32_t)) != hipSuccess ||
      hipHostMalloc(&st->h_stage, 65536, hipHostMallocDefault) != hipSuccess) {
    set_error(3, "stream_state: allocation failed");
    delete st;
    return nullptr;
  }
  g_ss[s] = st;
  return st;
}

void release_stream_state(hipStream_t s) {
  StreamState *st = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_ss_mu);
    auto it = g_ss.find(s);
    if (it != g_ss.end()) {
      st = it->second;
      g_ss.erase(it);
    }
  }
  (void)hipStreamSynchronize(s);
  if (!st) return;
  (void)hipFree(st->d_words);
  (void)hipFree(st->d_status);
  (void)hipFree(st->tmp);
  (void)hipHostFree(st->h_stage);
  delete st;
}

int ensure_status(StreamState *st, size_t words) {
  if (words <= st->status_words) return 0;
  CCX_TRY(hipStreamSynchronize(st->stream));
  if (st->d_status) CCX_TRY(hipFree(st->d_status));
  st->d_status = nullptr;
  st->status_words = 0;
  size_t want = words + words / 4 + 4096;
  CCX_TRY(hipMalloc(&st->d_status, want * sizeof(uint64_t)));
  CCX_TRY(hipMemsetAsync(st->d_status, 0, want * sizeof(uint64_t), st->stream));
  st->status_words = want;
  st->epoch = 0;
  return 0;
}

int next_epoch(StreamState *st, uint32_t *epoch) {
  if (st->epoch >= kEpochMax) {
    // Every tag value has been used: clear so no stale word can match.
    CCX_TRY(hipMemsetAsync(st->d_status, 0, st->status_words * sizeof(uint64_t), st->stream));
    st->epoch = 0;
  }
  *epoch = ++st->epoch;
  return 0;
}

void *stream_tmp(StreamState *st, size_t bytes) {
  if (bytes <= st->tmp_bytes) return st->tmp;
  if (hipStreamSynchronize(st->stream) != hipSuccess) {
    set_error(1, "stream_tmp: synchronize failed");
    return nullptr;
  }
  if (st->tmp) (void)hipFree(st->tmp);
  st->tmp = nullptr;
  st->tmp_bytes = 0;
  size_t want = bytes + bytes / 8 + 4096;
  if (hipMalloc(&st->tmp, want) != hipSuccess) {
    set_error(3, "stream_tmp: hipMalloc of %zu bytes failed", want);
    return nullptr;
  }
  st->tmp_bytes = want;
  return st->tmp;
}

hipStream_t pick_stream(void *stream) {
  if (stream) return (hipStream_t)stream;
  Context *c = ctx();
  return c ? c->stream : nullptr;
}

// ---- kernel timing -------------------------------------------------------

struct PendingPair {
  std::string name;
  hipEvent_t a, b;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<PendingPair> g_pending;
static std::vector<hipEvent_t> g_event_pool;
static std::map<std::string, std::pair<uint64_t, double>> g_prof_acc;
static thread_local PendingPair *g_open = nullptr;
static thread_local bool g_prof_muted = false;
static thread_local PendingPair g_open_slot;
static std::string g_names;

static hipEvent_t take_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void prof_mute(bool on) { g_prof_muted = on; }

void prof_begin(const char *name, hipStream_t s) {
  if (!g_prof_on || g_prof_muted) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_open_slot.name = name;
  g_open_slot.a = take_event();
  g_open_slot.b = take_event();
  if (!g_open_slot.a || !g_open_slot.b) return;
  (void)hipEventRecord(g_open_slot.a, s);
  g_open = &g_open_slot;
}

void prof_end(hipStream_t s) {
  if (!g_prof_on || !g_open) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  (void)hipEventRecord(g_open->b, s);
  g_pending.push_back(*g_open);
  g_open = nullptr;
}

static void prof_collect_locked() {
  for (auto &p : g_pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto &acc = g_prof_acc[p.name];
      acc.first += 1;
      acc.second += ms;
    }
    g_event_pool.push_back(p.a);
    g_event_pool.push_back(p.b);
  }
  g_pending.clear();
}

}  // namespace ccx

using namespace ccx;

extern "C" {

int cc_hip_init(int device) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  clear_error();
  return init_locked(device);
}

int cc_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char *cc_hip_last_error(void) { return g_err; }

const char *cc_hip_build_info(void) {
  return "libccmi355x: gfx950, wave64, HIP " __DATE__;
}

int cc_hip_malloc(void **d_ptr, size_t bytes) {
  if (!ctx()) return 2;
  if (!d_ptr) return set_error(4, "cc_hip_malloc: null out pointer");
  *d_ptr = nullptr;
  if (bytes == 0) bytes = 1;
  CCX_TRY(hipMalloc(d_ptr, bytes));
  return 0;
}

int cc_hip_free(void *d_ptr) {
  if (!d_ptr) return 0;
  if (!ctx()) return 2;
  CCX_TRY(hipFree(d_ptr));
  return 0;
}

int cc_hip_host_alloc(void **h_ptr, size_t bytes) {
  if (!ctx()) return 2;
  if (!h_ptr) return set_error(4, "cc_hip_host_alloc: null out pointer");
  *h_ptr = nullptr;
  if (bytes == 0) bytes = 1;
  CCX_TRY(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
  return 0;
}

int cc_hip_host_free(void *h_ptr) {
  if (!h_ptr) return 0;
  if (!ctx()) return 2;
  CCX_TRY(hipHostFree(h_ptr));
  return 0;
}

int cc_hip_memcpy_h2d(void *d_dst, const void *src, size_t bytes, void *stream) {
  if (!ctx()) return 2;
  if (bytes == 0) return 0;
  hipStream_t s = pick_stream(stream);
  CCX_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, s));
  CCX_TRY(hipStreamSynchronize(s));
  return 0;
}

int cc_hip_memcpy_d2h(void *dst, const void *d_src, size_t bytes, void *stream) {
  if (!ctx()) return 2;
  if (bytes == 0) return 0;
  hipStream_t s = pick_stream(stream);
  CCX_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  return 0;
}

int cc_hip_memcpy_d2h_async(void *h_dst, const void *d_src, size_t bytes, void *stream) {
  if (!ctx()) return 2;
  if (bytes == 0) return 0;
  CCX_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, pick_stream(stream)));
  return 0;
}

int cc_hip_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream) {
  if (!ctx()) return 2;
  if (bytes == 0) return 0;
  CCX_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, pick_stream(stream)));
  return 0;
}

int cc_hip_stream_create(void **stream) {
  if (!ctx()) return 2;
  if (!stream) return set_error(4, "cc_hip_stream_create: null out pointer");
  hipStream_t s = nullptr;
  CCX_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return 0;
}

int cc_hip_stream_destroy(void *stream) {
  if (!stream) return 0;
  if (!ctx()) return 2;
  release_stream_state((hipStream_t)stream);
  CCX_TRY(hipStreamDestroy((hipStream_t)stream));
  return 0;
}

int cc_hip_stream_sync(void *stream) {
  if (!ctx()) return 2;
  CCX_TRY(hipStreamSynchronize(pick_stream(stream)));
  return 0;
}

int cc_hip_device_sync(void) {
  if (!ctx()) return 2;
  CCX_TRY(hipDeviceSynchronize());
  return 0;
}

int cc_hip_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = (on != 0);
  return 0;
}

int cc_hip_prof_reset(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_collect_locked();
  g_prof_acc.clear();
  return 0;
}

int cc_hip_prof_query(const char *name, uint64_t *launches, double *total_ms) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_collect_locked();
  auto it = g_prof_acc.find(name ? name : "");
  if (launches) *launches = (it == g_prof_acc.end()) ? 0 : it->second.first;
  if (total_ms) *total_ms = (it == g_prof_acc.end()) ? 0.0 : it->second.second;
  return 0;
}

const char *cc_hip_prof_names(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_collect_locked();
  g_names.clear();
  for (auto &kv : g_prof_acc) {
    if (!g_names.empty()) g_names += ",";
    g_names += kv.first;
  }
  return g_names.c_str();
}

}  // extern "C"
