/*
 * Distributed connected components driver (host C + MPI): the drop-in for
 * the reference's include/cc_mpi.h (src/cc_mpi.c:221-792).
 *
 * Same phases and the same merge schedule as the reference; the per-rank
 * data stays in HBM and every O(m) or O(n) loop runs on the GPU through
 * libccmi355x.so:
 *   1. local-local split + raw boundary records   (src/cc_mpi.c:252-304, 376-394)
 *   2. local CC                                    (:315)
 *   3. rep fill, sort, dedup, ghost extraction    (:379, 396-461)
 *   4. exchange plan                               (:463-487)   host + MPI
 *   5. merge loop: init/hook/update on the GPU, two LOR allreduces, delta
 *      exchange every exchange_interval rounds or when no rank changed
 *                                                  (:541-710)
 *   6. root count + MPI_Reduce, labels_global gather (:730-742; the gather
 *      honours include/cc_mpi.h:26-27, which the reference leaves undone)
 *
 * Deviation on purpose: every collective is entered by every rank (the
 * reference returns early on an empty rank and builds the plan and runs the
 * merge loop only where ghosts exist, which hangs mixed jobs).
 */
#define _POSIX_C_SOURCE 200112L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mpi.h>

#include "../../../include/cc_hip.h"
#include "../../../include/cc_mpi.h"
#include "../../../include/exchange_plan.h"

#ifndef MPI_UINT64_T
#define MPI_UINT64_T MPI_UNSIGNED_LONG_LONG
#endif
#ifndef MPI_UINT32_T
#define MPI_UINT32_T MPI_UNSIGNED
#endif

static int g_threads = 0;
static const uint32_t *g_part_bounds = NULL;
static int g_part_size = 0;
static int g_part_kind = 0;
static uint64_t g_last_components = 0;
static int g_last_rounds = 0;

void cc_mpi_set_num_threads(int nthreads) { g_threads = nthreads > 0 ? nthreads : 0; }
uint64_t cc_mpi_last_component_count(void) { return g_last_components; }
int cc_mpi_last_merge_rounds(void) { return g_last_rounds; }

static int owner_of_vertex(uint32_t v, uint32_t n_global, int comm_size)
{
  const uint32_t *b = (g_part_size == comm_size) ? g_part_bounds : NULL;
  return cc_partition_owner(v, n_global, comm_size, b, g_part_kind);
}

static void check(int rc, MPI_Comm comm, const char *what)
{
  if (rc != 0)
  {
    fprintf(stderr, "compute_connected_components_mpi_advanced: %s: %s\n", what,
            cc_hip_last_error());
    MPI_Abort(comm, EXIT_FAILURE);
  }
}

static void *dalloc(size_t bytes, MPI_Comm comm, const char *what)
{
  void *p = NULL;
  check(cc_hip_malloc(&p, bytes ? bytes : 4), comm, what);
  return p;
}

void compute_connected_components_mpi_advanced(const DistCSRGraph *Gd, uint32_t *labels_global,
                                               int chunk_size, int exchange_interval,
                                               MPI_Comm comm)
{
  (void)chunk_size;
  if (exchange_interval < 1)
    exchange_interval = 1;
  int rank = 0, size = 1;
  MPI_Comm_rank(comm, &rank);
  MPI_Comm_size(comm, &size);
  g_part_kind = Gd->part_kind;
  g_part_size = Gd->part_size;
  g_part_bounds = Gd->part_bounds;
  const uint32_t n_global = Gd->n_global, v_start = Gd->v_start, v_end = Gd->v_end;
  const uint32_t n_local = Gd->n_local;
  const uint64_t m_local = Gd->m_local;
  g_last_components = 0;
  g_last_rounds = 0;
  if (n_global == 0)
    return; /* same on every rank */
  const double t0 = MPI_Wtime();

  /* 1. owned CSR -> device; split into the local CSR and boundary records */
  uint64_t *d_row = (uint64_t *)dalloc(((size_t)n_local + 1) * 8, comm, "alloc row_ptr");
  uint32_t *d_col = (uint32_t *)dalloc((size_t)m_local * 4, comm, "alloc col_idx");
  check(cc_hip_memcpy_h2d(d_row, Gd->row_ptr, ((size_t)n_local + 1) * 8, NULL), comm, "upload row_ptr");
  check(cc_hip_memcpy_h2d(d_col, Gd->col_idx, (size_t)m_local * 4, NULL), comm, "upload col_idx");
  uint64_t *d_lrow = (uint64_t *)dalloc(((size_t)n_local + 1) * 8, comm, "alloc lrow");
  uint64_t *d_boff = (uint64_t *)dalloc(((size_t)n_local + 1) * 8, comm, "alloc boff");
  uint64_t m_ll = 0, nb = 0;
  check(cc_hip_split_count_device(d_row, d_col, n_local, v_start, v_end, n_global, &m_ll, &nb,
                                  NULL),
        comm, "split count");
  uint32_t *d_lcol = (uint32_t *)dalloc((size_t)m_ll * 4, comm, "alloc lcol");
  uint64_t *d_rec = (uint64_t *)dalloc((size_t)nb * 8, comm, "alloc boundary");
  check(cc_hip_split_emit_device(d_row, d_col, n_local, v_start, v_end, n_global, d_lrow, d_boff,
                                 d_lcol, d_rec, NULL),
        comm, "split emit");
  check(cc_hip_device_sync(), comm, "split");
  cc_hip_free(d_row);
  cc_hip_free(d_col);
  cc_hip_free(d_boff);

  /* 2. local CC */
  uint32_t *d_comp_of = (uint32_t *)dalloc((size_t)n_local * 4, comm, "alloc comp_of");
  const double t_cc0 = MPI_Wtime();
  check(cc_hip_components_device(d_lrow, d_lcol, n_local, m_ll, d_comp_of, NULL), comm, "local CC");
  check(cc_hip_device_sync(), comm, "local CC");
  const double t_cc1 = MPI_Wtime();
  printf("[rank %d] Local CC done: n_local=%u m_local=%llu device time=%.3fs\n", rank, n_local,
         (unsigned long long)m_ll, t_cc1 - t_cc0);
  cc_hip_free(d_lrow);
  cc_hip_free(d_lcol);

  /* 3. boundary: rep fill, sort, dedup, ghost extraction */
  uint32_t ghost_count = 0;
  uint32_t *ghost_vertices = NULL;
  uint32_t *d_ghosts = NULL, *d_ghost_labels = NULL;
  uint64_t *d_bnd = NULL; /* (gidx << 32) | rep */
  uint64_t nbd = 0;
  if (nb > 0)
  {
    /* rep fill + sort + dedup in one call (src/cc_mpi.c:364-394,
     * src/boundary_sort.c:93-119) */
    uint64_t *d_alt = (uint64_t *)dalloc((size_t)nb * 8, comm, "alloc sort buffer");
    int in_alt = 0;
    check(cc_hip_boundary_build_device(d_rec, d_alt, nb, d_comp_of, n_local, &nbd, &in_alt, NULL),
          comm, "boundary build");
    uint64_t *deduped = in_alt ? d_alt : d_rec, *other = in_alt ? d_rec : d_alt;
    d_ghosts = (uint32_t *)dalloc((size_t)(nbd ? nbd : 1) * 4, comm, "alloc ghosts");
    uint64_t gc = 0;
    check(cc_hip_ghost_extract_device(deduped, nbd, other, d_ghosts, &gc, NULL), comm, "ghosts");
    ghost_count = (uint32_t)gc;
    d_bnd = other;
    cc_hip_free(deduped);
    ghost_vertices = (uint32_t *)malloc((size_t)(ghost_count ? ghost_count : 1) * 4);
    if (!ghost_vertices)
      MPI_Abort(comm, EXIT_FAILURE);
    check(cc_hip_memcpy_d2h(ghost_vertices, d_ghosts, (size_t)ghost_count * 4, NULL), comm,
          "download ghosts");
    d_ghost_labels = (uint32_t *)dalloc((size_t)ghost_count * 4, comm, "alloc ghost labels");
    check(cc_hip_memcpy_d2d(d_ghost_labels, d_ghosts, (size_t)ghost_count * 4, NULL), comm,
          "ghost labels init");
  }
  else
  {
    cc_hip_free(d_rec);
    ghost_vertices = (uint32_t *)malloc(4);
  }
  printf("[rank %d] boundary_edges raw=%llu dedup=%llu ghost_count=%u\n", rank,
         (unsigned long long)nb, (unsigned long long)nbd, ghost_count);

  /* 4. exchange plan (collective on every rank) */
  ExchangePlan plan;
  memset(&plan, 0, sizeof(plan));
  exchangeplan_build(&plan, ghost_vertices, ghost_count, n_global, comm, owner_of_vertex);
  ExchangeDevice xd;
  check(exchangeplan_device_init(&plan, &xd), comm, "exchange device init");
  printf("[rank %d] plan indegree=%d outdegree=%d send=%d recv=%d\n", rank, plan.indegree,
         plan.outdegree, plan.total_send_to, plan.total_need_from);
  uint32_t *d_prev = (uint32_t *)dalloc((size_t)plan.total_send_to * 4, comm, "alloc prev_sent");
  {
    uint32_t *ones = (uint32_t *)malloc((size_t)(plan.total_send_to ? plan.total_send_to : 1) * 4);
    if (!ones)
      MPI_Abort(comm, EXIT_FAILURE);
    memset(ones, 0xFF, (size_t)plan.total_send_to * 4);
    check(cc_hip_memcpy_h2d(d_prev, ones, (size_t)plan.total_send_to * 4, NULL), comm, "prev_sent");
    free(ones);
  }
  uint32_t *d_comp_label = (uint32_t *)dalloc((size_t)n_local * 4, comm, "alloc comp_label");
  uint32_t *d_comp_min = (uint32_t *)dalloc((size_t)n_local * 4, comm, "alloc comp_min");
  check(cc_hip_iota_u32_device(d_comp_label, n_local, v_start, NULL), comm, "comp_label init");
  /* the rep list (cc_mpi.c:328-362 of the reference): the merge round's
     per-rep passes walk it instead of scanning comp_of */
  uint32_t *d_reps = (uint32_t *)dalloc((size_t)n_local * 4, comm, "alloc reps");
  uint32_t rep_count = 0;
  check(cc_hip_local_reps_device(d_comp_of, n_local, d_reps, &rep_count, NULL), comm, "local reps");
  /* comp_of is fixed from here on: the packs gather the send entries' reps once */
  check(exchangeplan_device_prepare(&plan, &xd, d_comp_of, v_start, v_end), comm,
        "exchange device prepare");
  /* the merge rounds' records, split once: the most frequent rep's as a
     bitmap over the ghosts, the others as a side list; each rep's minimum
     ghost label is kept until the ghost labels change */
  uint32_t *d_dom_bits = (uint32_t *)dalloc((size_t)((ghost_count + 31) / 32) * 4, comm,
                                            "alloc dominant bitmap");
  uint64_t *d_side = (uint64_t *)dalloc((size_t)nbd * 8, comm, "alloc side records");
  uint32_t *d_ghost_min = (uint32_t *)dalloc((size_t)rep_count * 4, comm, "alloc ghost minimum");
  uint64_t n_side = 0;
  uint32_t dominant = 0;
  check(cc_hip_merge_prepare_device(d_bnd, nbd, ghost_count, d_reps, rep_count, d_dom_bits, d_side,
                                    &n_side, &dominant, NULL),
        comm, "merge prepare");
  int ghosts_changed = 2; /* first round: seed the per-rep ghost minimum */

  /* 5. merge loop */
  MPI_Barrier(comm);
  int local_pending = (ghost_count > 0 && nbd > 0 && n_local > 0) ? 1 : 0;
  int global_pending = 0;
  MPI_Allreduce(&local_pending, &global_pending, 1, MPI_INT, MPI_LOR, comm);
  int pending_since_exchange = local_pending;
  int steps_since_exchange = 0, iter = 0;
  while (global_pending)
  {
    int changed = 0;
    check(cc_hip_merge_step_prepared_device(d_dom_bits, dominant, d_side, n_side, d_ghosts,
                                            d_ghost_labels, ghost_count, d_comp_of, d_comp_label,
                                            d_comp_min, d_ghost_min, n_local, d_reps, rep_count,
                                            v_start, v_end, ghosts_changed, &changed, NULL),
          comm, "merge step");
    ghosts_changed = 0;
    if (changed)
      pending_since_exchange = 1;
    int global_changed = 0;
    MPI_Allreduce(&pending_since_exchange, &global_pending, 1, MPI_INT, MPI_LOR, comm);
    MPI_Allreduce(&changed, &global_changed, 1, MPI_INT, MPI_LOR, comm);
    ++iter;
    if (!global_pending)
      break;
    ++steps_since_exchange;
    if (steps_since_exchange >= exchange_interval || !global_changed)
    {
      exchangeplan_exchange_delta_device(&plan, &xd, d_comp_label, d_comp_of, v_start, v_end,
                                         d_prev, d_ghost_labels);
      ghosts_changed = xd.last_received > 0;
      pending_since_exchange = 0;
      steps_since_exchange = 0;
    }
  }
  if (rank == 0)
    printf("CC merge converged in %d rounds\n", iter);
  g_last_rounds = iter;

  /* 6. labels, root count, gather */
  uint32_t *d_labels = (uint32_t *)dalloc((size_t)n_local * 4, comm, "alloc labels");
  uint64_t local_roots = 0, global_roots = 0;
  check(cc_hip_rank_labels_device(d_comp_label, d_comp_of, n_local, v_start, d_labels,
                                  &local_roots, NULL),
        comm, "labels");
  MPI_Allreduce(&local_roots, &global_roots, 1, MPI_UINT64_T, MPI_SUM, comm);
  g_last_components = global_roots;
  if (rank == 0)
    printf("Number of connected components: %llu\n", (unsigned long long)global_roots);
  if (labels_global)
  {
    uint32_t *mine = (uint32_t *)malloc((size_t)(n_local ? n_local : 1) * 4);
    int *counts = (int *)malloc((size_t)size * sizeof(int));
    int *displs = (int *)malloc((size_t)size * sizeof(int));
    if (!mine || !counts || !displs)
      MPI_Abort(comm, EXIT_FAILURE);
    check(cc_hip_memcpy_d2h(mine, d_labels, (size_t)n_local * 4, NULL), comm, "download labels");
    int nl = (int)n_local;
    MPI_Allgather(&nl, 1, MPI_INT, counts, 1, MPI_INT, comm);
    uint32_t vs = v_start;
    uint32_t *starts = (uint32_t *)malloc((size_t)size * sizeof(uint32_t));
    if (!starts)
      MPI_Abort(comm, EXIT_FAILURE);
    MPI_Allgather(&vs, 1, MPI_UINT32_T, starts, 1, MPI_UINT32_T, comm);
    for (int r = 0; r < size; ++r)
      displs[r] = (int)starts[r];
    MPI_Allgatherv(mine, nl, MPI_UINT32_T, labels_global, counts, displs, MPI_UINT32_T, comm);
    free(starts);
    free(mine);
    free(counts);
    free(displs);
  }
  if (rank == 0)
    printf("[rank 0] total_cc_time=%.3fs\n", MPI_Wtime() - t0);

  cc_hip_free(d_labels);
  cc_hip_free(d_comp_label);
  cc_hip_free(d_comp_min);
  cc_hip_free(d_reps);
  cc_hip_free(d_prev);
  cc_hip_free(d_dom_bits);
  cc_hip_free(d_side);
  cc_hip_free(d_ghost_min);
  exchangeplan_device_free(&xd);
  exchangeplan_free(&plan);
  cc_hip_free(d_bnd);
  cc_hip_free(d_ghosts);
  cc_hip_free(d_ghost_labels);
  cc_hip_free(d_comp_of);
  free(ghost_vertices);
}

uint32_t count_connected_components(const uint32_t *labels_global, uint32_t n_global)
{
  if (n_global == 0 || labels_global == NULL)
    return 0;
  unsigned char *seen = (unsigned char *)calloc((size_t)n_global, 1);
  if (!seen)
  {
    fprintf(stderr, "count_connected_components: Failed to allocate bitmap\n");
    return 0;
  }
  uint32_t count = 0;
  for (uint32_t v = 0; v < n_global; ++v)
  {
    const uint32_t r = labels_global[v];
    if (r < n_global && !seen[r])
    {
      seen[r] = 1;
      ++count;
    }
  }
  free(seen);
  return count;
}
