#ifndef CC_H
#define CC_H

/*
 * Local connected components: the drop-in for the reference's
 * include/cc.h:27-28.  Backed by HIP kernels on the MI355X (gfx950); the
 * reference's pthreads label propagation is not linked in.
 */

#include <stdint.h>
#include "graph.h"

/* include/cc.h:7 */
#define DEFAULT_CHUNK_SIZE 4096

#ifdef __cplusplus
#define CC_RESTRICT __restrict__
extern "C" {
#else
#define CC_RESTRICT restrict
#endif

/*
 * labels[v] = smallest vertex id in v's connected component, for every
 * v in [0, G->n).  Same contract as the reference (src/cc_pthreads.c:152-229:
 * labels start at the vertex id and are only ever lowered to a neighbour's
 * label, so the fixed point is the per-component minimum).  G must be
 * symmetric (both directions of every edge stored), as the reference's
 * inputs and its local-local CSR are; duplicates, self-loops and unsorted
 * rows are fine, and column ids >= n are ignored.
 *
 * G is read-only host memory; labels is caller-owned host memory of G->n
 * entries and is fully written.  Blocking: returns when the labels are back
 * in host memory.  num_threads is accepted for ABI compatibility and does not
 * select anything (the device does the work); chunk_size keeps the
 * reference's meaning only as a scheduling hint (1 = static, >0 = dynamic
 * chunk, <=0 = DEFAULT_CHUNK_SIZE) and never changes the result.
 * A HIP failure prints to stderr and exits with EXIT_FAILURE, as the
 * reference does on allocation failure (src/cc_pthreads.c:163-167).
 */
void compute_connected_components_pthreads(const CSRGraph *CC_RESTRICT G, uint32_t *CC_RESTRICT labels,
                                           int num_threads, int chunk_size);

#ifdef __cplusplus
}
#endif

#endif /* CC_H */
