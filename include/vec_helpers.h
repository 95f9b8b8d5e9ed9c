#ifndef VEC_HELPERS_H
#define VEC_HELPERS_H

/*
 * Boundary-edge record types, ABI-identical to the reference's
 * include/vec_helpers.h:34-38 (BoundaryEdge) and :43-47 (BoundaryPair).
 *
 * On a little-endian machine (x86-64 host, AMD GPU) the 8 bytes of a
 * BoundaryEdge read as one uint64_t are ((uint64_t)remote << 32) | rep,
 * which is exactly the reference's sort key (src/boundary_sort.c:5-8), so
 * the device sorts the raw records as 64-bit integers.
 *
 * The reference header also pulls in <mpi.h> for its growable vectors; this
 * one does not, so the GPU library has no MPI dependency.
 */

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/** Local representative and remote vertex describing a boundary edge. */
typedef struct
{
  uint32_t rep;
  uint32_t remote;
} BoundaryEdge;

/** Local representative and ghost index pair. */
typedef struct
{
  uint32_t rep;
  uint32_t gidx;
} BoundaryPair;

#ifdef __cplusplus
}
#endif

#endif /* VEC_HELPERS_H */
