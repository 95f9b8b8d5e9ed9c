#ifndef GRAPH_H
#define GRAPH_H

/*
 * Compressed sparse row graph, ABI-identical to the reference's CSRGraph
 * (include/graph.h:8-14): n vertices, m directed entries, row_ptr[n+1]
 * 64-bit offsets, col_idx[m] 32-bit neighbour ids.  An undirected graph
 * stores both directions of every edge.
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  uint32_t n;
  uint64_t m;
  uint64_t *row_ptr;
  uint32_t *col_idx;
} CSRGraph;

/* Free the arrays of a CSRGraph built by this library (malloc'd) and zero it.
 * Replaces include/graph.h:33 (free_csr). */
void free_csr(CSRGraph *g);

#ifdef __cplusplus
}
#endif

#endif /* GRAPH_H */
