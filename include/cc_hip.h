#ifndef CC_HIP_H
#define CC_HIP_H

/*
 * Device-resident C ABI of the MI355X connected-components library
 * (libccmi355x.so).  Everything here takes plain pointers and sizes; device
 * pointers are named d_*, pinned host pointers h_*.  A `stream` argument is a
 * hipStream_t passed as void*; NULL means the library's own stream.
 *
 * These are the building blocks the reference-compatible entry points
 * (cc.h, boundary_sort.h, exchange_plan.h) are made of, exposed so a caller
 * can keep the CSR, labels and boundary records resident in HBM between
 * calls instead of paying a host round trip each time.
 *
 * Return value convention: 0 on success, nonzero on failure with a message
 * available from cc_hip_last_error().  Nothing here falls back to the CPU.
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ---------------------------------------------------------- */

/* Select and initialise the device (idempotent; the other calls initialise
 * device 0 lazily).  Returns nonzero if no usable gfx950 device exists. */
int cc_hip_init(int device);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int cc_hip_device_count(void);
/* Text of the last error on this thread ("" if none). */
const char *cc_hip_last_error(void);
/* Library build identification string (arch and options). */
const char *cc_hip_build_info(void);

int cc_hip_malloc(void **d_ptr, size_t bytes);
int cc_hip_free(void *d_ptr);
int cc_hip_host_alloc(void **h_ptr, size_t bytes); /* pinned, device-visible */
int cc_hip_host_free(void *h_ptr);
int cc_hip_memcpy_h2d(void *d_dst, const void *src, size_t bytes, void *stream);
int cc_hip_memcpy_d2h(void *dst, const void *d_src, size_t bytes, void *stream);
int cc_hip_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
int cc_hip_stream_create(void **stream);
int cc_hip_stream_destroy(void *stream);
int cc_hip_stream_sync(void *stream);
int cc_hip_device_sync(void);

/* ---- kernel timing ---------------------------------------------------- */

/* When enabled, every kernel launch of the library is bracketed by HIP
 * events on the stream it runs on, and its duration is accumulated per
 * kernel name.  Disabled by default. */
int cc_hip_prof_enable(int on);
int cc_hip_prof_reset(void);
/* Collect pending event pairs (synchronises on them) and return, for kernel
 * name `name`, the number of launches and their total milliseconds. */
int cc_hip_prof_query(const char *name, uint64_t *launches, double *total_ms);
/* Comma-separated list of kernel names seen since the last reset. */
const char *cc_hip_prof_names(void);

/* ---- local connected components --------------------------------------- */

/*
 * d_labels[v] = smallest vertex id in v's component of the undirected graph
 * (d_row_ptr[n+1], d_col_idx[m]) resident in device memory.  The graph must
 * be symmetric (both directions stored), as the reference's local CSR is
 * (src/cc_mpi.c:252-304).  Self-loops and duplicate entries are allowed.
 * Algorithm: sampled hooking + full hooking of the non-giant frontier with
 * atomicMin on parent pointers, then pointer jumping; deterministic result.
 * Asynchronous on `stream`.
 */
int cc_hip_components_device(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                             uint32_t n, uint64_t m, uint32_t *d_labels, void *stream);

/* Same result through host buffers (H2D, solve, D2H; blocking). */
int cc_hip_components_host(const uint64_t *row_ptr, const uint32_t *col_idx,
                           uint32_t n, uint64_t m, uint32_t *labels);

#ifdef __cplusplus
}
#endif

#endif /* CC_HIP_H */
