#ifndef CC_MPI_H
#define CC_MPI_H

/*
 * Distributed connected components: the drop-in for the reference's
 * include/cc_mpi.h:29-47.  Host C and MPI as in the reference; the local
 * CC, the local-local split, the boundary sort/dedup/ghost extraction, the
 * merge hook/update and the delta pack/unpack run on the GPU through
 * libccmi355x.so.
 */

#include <stdint.h>
#include <mpi.h>
#include "graph_dist.h"
#include "cc.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Collective on comm (reference src/cc_mpi.c:221-763).  On return
 * labels_global[v] (n_global entries, if labels_global is non-NULL) is the
 * smallest vertex id of v's component, identical on all ranks — the
 * reference header's contract (include/cc_mpi.h:26-27), which its body
 * leaves unimplemented.  chunk_size keeps its meaning as a scheduling hint;
 * exchange_interval (< 1 means 1) is the number of merge rounds between
 * halo exchanges.  Prints the component count on rank 0 like the reference.
 * Every rank takes part in every collective, including ranks that own no
 * vertices or have no ghosts.
 */
void compute_connected_components_mpi_advanced(const DistCSRGraph *Gd, uint32_t *labels_global,
                                               int chunk_size, int exchange_interval,
                                               MPI_Comm comm);

/* Number of distinct labels < n_global (src/cc_mpi.c:765-792). */
uint32_t count_connected_components(const uint32_t *labels_global, uint32_t n_global);

/* Thread count hint (src/cc_mpi.c:30-35); accepted for interface parity. */
void cc_mpi_set_num_threads(int nthreads);

/* Component count of the last compute_connected_components_mpi_advanced
 * call (all ranks), and the number of merge rounds it took. */
uint64_t cc_mpi_last_component_count(void);
int cc_mpi_last_merge_rounds(void);

#ifdef __cplusplus
}
#endif

#endif /* CC_MPI_H */
