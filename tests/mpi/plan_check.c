/*
 * MPI check of exchangeplan_build and the legacy label exchange (host-only
 * parts of include/exchange_plan.h).  Each rank makes a seeded sorted ghost
 * list of vertices owned elsewhere (rank 1 has none, to check that a rank
 * without ghosts does not hang the others), builds the plan, and verifies it
 * against an independent reconstruction from an allgather of every ghost
 * list.  Then labels f(v) = v ^ 0x5a5a are exchanged with
 * exchangeplan_exchange and checked.  Prints "plan_check OK" on rank 0.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mpi.h>

#include "exchange_plan.h"
#include "graph_dist.h"

static uint32_t g_n;

static int owner(uint32_t v, uint32_t n, int P) { return cc_partition_owner(v, n, P, NULL, 0); }

static uint64_t mix(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

#define FAIL(...)                                  \
  do                                               \
  {                                                \
    fprintf(stderr, "rank %d: ", rank);            \
    fprintf(stderr, __VA_ARGS__);                  \
    fprintf(stderr, "\n");                         \
    MPI_Abort(MPI_COMM_WORLD, 3);                  \
  } while (0)

int main(int argc, char **argv)
{
  MPI_Init(&argc, &argv);
  int rank, P;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &P);
  g_n = (argc > 1) ? (uint32_t)atoi(argv[1]) : 1000;
  uint32_t vs, ve;
  cc_partition_block_range(g_n, P, rank, &vs, &ve);

  /* sorted unique ghosts: a seeded subset of the non-owned vertices */
  uint32_t *ghosts = malloc(g_n * sizeof(uint32_t));
  uint32_t gc = 0;
  for (uint32_t v = 0; v < g_n; ++v)
    if ((v < vs || v >= ve) && rank != 1 && (mix((uint64_t)rank * 1000003u + v) % 5 == 0))
      ghosts[gc++] = v;

  ExchangePlan plan;
  memset(&plan, 0, sizeof(plan));
  exchangeplan_build(&plan, ghosts, gc, g_n, MPI_COMM_WORLD, owner);

  /* independent reconstruction */
  int *counts = malloc(P * sizeof(int)), *displs = malloc(P * sizeof(int));
  int gci = (int)gc;
  MPI_Allgather(&gci, 1, MPI_INT, counts, 1, MPI_INT, MPI_COMM_WORLD);
  int tot = 0;
  for (int r = 0; r < P; ++r)
  {
    displs[r] = tot;
    tot += counts[r];
  }
  uint32_t *all = malloc((tot ? tot : 1) * sizeof(uint32_t));
  MPI_Allgatherv(ghosts, gci, MPI_UNSIGNED, all, counts, displs, MPI_UNSIGNED, MPI_COMM_WORLD);

  int ts = 0;
  for (int q = 0; q < P; ++q)
  {
    /* what q needs from me, in q's ghost order */
    int want = 0;
    for (int i = 0; i < counts[q]; ++i)
      if (q != rank && owner(all[displs[q] + i], g_n, P) == rank)
        ++want;
    if (plan.send_to_counts[q] != want)
      FAIL("send_to_counts[%d]=%d want %d", q, plan.send_to_counts[q], want);
    if (plan.send_to_displs[q] != ts)
      FAIL("send_to_displs[%d]=%d want %d", q, plan.send_to_displs[q], ts);
    int k = 0;
    for (int i = 0; i < counts[q]; ++i)
    {
      uint32_t v = all[displs[q] + i];
      if (q != rank && owner(v, g_n, P) == rank)
      {
        if (plan.send_to_flat[ts + k] != v)
          FAIL("send_to_flat mismatch for dest %d", q);
        ++k;
      }
    }
    ts += want;
    /* what I need from q, in my ghost order, with ghost indices */
    int need = 0;
    for (uint32_t gi = 0; gi < gc; ++gi)
      if (owner(ghosts[gi], g_n, P) == q && q != rank)
      {
        int at = plan.need_from_displs[q] + need;
        if (plan.need_from_flat[at] != ghosts[gi] || plan.need_from_gidx_flat[at] != gi)
          FAIL("need_from mismatch for source %d", q);
        ++need;
      }
    if (plan.need_from_counts[q] != need)
      FAIL("need_from_counts[%d]=%d want %d", q, plan.need_from_counts[q], need);
  }
  if (plan.total_send_to != ts)
    FAIL("total_send_to %d want %d", plan.total_send_to, ts);
  for (int k = 1; k < plan.outdegree; ++k)
    if (plan.dests[k] <= plan.dests[k - 1])
      FAIL("dests not ascending");
  for (int k = 0; k < plan.outdegree; ++k)
    if (plan.sendcounts[k] != plan.send_to_counts[plan.dests[k]])
      FAIL("sendcounts layout");
  for (int k = 0; k < plan.indegree; ++k)
    if (plan.recvcounts[k] != plan.need_from_counts[plan.sources[k]])
      FAIL("recvcounts layout");
  if (plan.comm_graph == MPI_COMM_NULL)
    FAIL("no neighbourhood communicator");

  /* legacy exchange: fill send labels, exchange, check ghost labels */
  for (int i = 0; i < plan.total_send_to; ++i)
    plan.send_labels_flat[i] = plan.send_to_flat[i] ^ 0x5a5au;
  uint32_t *glab = malloc((gc ? gc : 1) * sizeof(uint32_t));
  for (uint32_t gi = 0; gi < gc; ++gi)
    glab[gi] = 0xFFFFFFFFu;
  exchangeplan_exchange(&plan, glab, MPI_COMM_WORLD);
  for (uint32_t gi = 0; gi < gc; ++gi)
    if (glab[gi] != (ghosts[gi] ^ 0x5a5au))
      FAIL("ghost label %u wrong", gi);

  exchangeplan_free(&plan);
  MPI_Barrier(MPI_COMM_WORLD);
  if (rank == 0)
    printf("plan_check OK ranks=%d n=%u\n", P, g_n);
  free(ghosts);
  free(counts);
  free(displs);
  free(all);
  free(glab);
  MPI_Finalize();
  return 0;
}
