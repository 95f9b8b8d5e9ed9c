#ifndef BOUNDARY_SORT_H
#define BOUNDARY_SORT_H

/*
 * Boundary-edge sort and deduplication: the drop-in for the reference's
 * include/boundary_sort.h:8-9.  Backed by a device radix sort.
 */

#include <stdint.h>
#include "vec_helpers.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Sort edges[0..count) ascending by the 64-bit key (remote << 32) | rep
 * (reference src/boundary_sort.c:93-98).  In place on caller-owned host
 * memory; NULL or count <= 1 is a no-op.  The order is total on the key and
 * the key is the whole record, so the output is uniquely determined.
 */
void boundary_edges_sort(BoundaryEdge *edges, uint64_t count);

/*
 * Remove adjacent exact duplicates in place and return the new count
 * (reference src/boundary_sort.c:100-119).  NULL or count <= 1 returns count.
 */
uint64_t boundary_edges_dedup(BoundaryEdge *edges, uint64_t count);

#ifdef __cplusplus
}
#endif

#endif /* BOUNDARY_SORT_H */
