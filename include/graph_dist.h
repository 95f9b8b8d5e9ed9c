#ifndef GRAPH_DIST_H
#define GRAPH_DIST_H

/*
 * Row-partitioned CSR, ABI-identical to the reference's DistCSRGraph
 * (include/graph_dist.h:15-28): each rank owns the contiguous vertex block
 * [v_start, v_end); row_ptr is local (n_local+1 entries, starting at 0) and
 * col_idx holds GLOBAL vertex ids.  part_kind 0 = vertex blocks, 1 = the
 * edge-balanced bounds in part_bounds[part_size+1].
 */

#include <stdint.h>
#include <mpi.h>
#include "graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  uint32_t n_global;
  uint64_t m_global;
  uint32_t n_local;
  uint64_t m_local;
  uint32_t v_start;
  uint32_t v_end;
  uint64_t *row_ptr;
  uint32_t *col_idx;
  uint32_t *part_bounds;
  int part_size;
  int part_kind;
} DistCSRGraph;

/*
 * Distribute a graph held by rank 0 (`full`; ignored on other ranks) the
 * way the reference's loader does after reading its file
 * (src/graph_dist.c:142-342): broadcast n and m, compute the partition
 * (use_edge = 1: edge-balanced bounds with vertex weight 1, else vertex
 * blocks), scatter each rank's row_ptr window (rebased to 0) and its
 * col_idx slice.  Collective on comm; returns 0 on success.
 */
int dist_csr_from_global(const CSRGraph *full, int use_edge, DistCSRGraph *out, MPI_Comm comm);

/* Replaces include/graph_dist.h:41-46 (src/graph_dist.c:344-357): a ".mat"
 * path is read on rank 0 by load_csr_from_file, the return code is
 * broadcast, and on success the graph is distributed as by
 * dist_csr_from_global; any other extension returns -1.  symmetrize and
 * drop_self_loops are ignored, as there.  Collective on comm. */
int load_dist_csr_from_file(const char *path, int symmetrize, int drop_self_loops, int use_edge,
                            DistCSRGraph *out, MPI_Comm comm);

/* Free the local arrays (include/graph_dist.h:49, src/graph_dist.c:360-378). */
void free_dist_csr(DistCSRGraph *graph);

/* Vertex-block partition: the first n % P ranks get one extra vertex
 * (src/graph_dist.c:22-51). */
void cc_partition_block_range(uint32_t n, int comm_size, int rank, uint32_t *v_start,
                              uint32_t *v_end);

/* Edge-balanced bounds[comm_size+1], weight deg+vertex_weight per vertex,
 * every rank at least one vertex (src/graph_dist.c:53-92). */
void cc_partition_weighted_bounds(const uint64_t *row_ptr, uint32_t n, int comm_size,
                                  uint32_t vertex_weight, uint32_t *bounds);

/* Owner rank of v under a partition (src/cc_mpi.c:41-76): binary search of
 * bounds when part_kind != 0, else the block formula. */
int cc_partition_owner(uint32_t v, uint32_t n_global, int comm_size, const uint32_t *bounds,
                       int part_kind);

#ifdef __cplusplus
}
#endif

#endif /* GRAPH_DIST_H */
