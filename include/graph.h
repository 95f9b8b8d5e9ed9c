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

/* Read a graph from a level 5 MAT file (plain or zlib-compressed
 * variables, either byte order).  Replaces include/graph.h:24
 * (src/graph.c:38-128): the first struct named "Problem" supplies its field
 * "A", otherwise the first sparse variable is used; rows become vertices
 * (n = number of rows) and row r lists, in increasing order, the columns
 * holding an entry in row r; values are ignored.  Returns 0 on success,
 * 1 if the file cannot be opened or is not a level 5 MAT file, 2 if it
 * holds no usable sparse matrix, 3 if memory runs out. */
int load_csr_from_mat(const char *path, CSRGraph *out);

/* Dispatch on the file extension, as include/graph.h:30 (src/graph.c:14-35):
 * ".mat" -> load_csr_from_mat; ".mtx"/".txt" -> -1 (not supported, as in
 * the reference); anything else -> 99.  symmetrize and drop_self_loops are
 * accepted and ignored, as there.  The reference also declares
 * load_csr_from_mtx (include/graph.h:20) but defines it nowhere, so it is
 * not declared here. */
int load_csr_from_file(const char *path, int symmetrize, int drop_self_loops, CSRGraph *out);

/* Write `g` as an n x n sparse logical matrix to a level 5 MAT file, so that
 * load_csr_from_mat returns `g` again (rows sorted, no duplicates, every
 * column id < n).  No counterpart in the reference, whose inputs come from
 * MATLAB; here it produces inputs for the reference's file path.  flags:
 * CC_MAT_COMPRESS wraps the variable in a zlib stream, CC_MAT_PROBLEM
 * stores it as field "A" of a struct "Problem" (field "name" holds `name`),
 * CC_MAT_BIG_ENDIAN writes big-endian.  The format limits the matrix to
 * fewer than 2^31 entries.  Returns 0, 1 (cannot write the file), 3 (out of
 * memory) or 4 (invalid graph or too large). */
#define CC_MAT_COMPRESS 1
#define CC_MAT_PROBLEM 2
#define CC_MAT_BIG_ENDIAN 4
int cc_mat_write_sparse(const char *path, const char *name, const CSRGraph *g, int flags);

#ifdef __cplusplus
}
#endif

#endif /* GRAPH_H */
