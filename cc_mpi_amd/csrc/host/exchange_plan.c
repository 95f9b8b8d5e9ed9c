/*
 * Halo-exchange plan and delta exchange (host C + MPI), the drop-in for
 * include/exchange_plan.h of the reference (src/exchange_plan.c).  The plan
 * build is host work done once; the per-exchange pack and unpack run on the
 * GPU (cc_hip_delta_pack_device / cc_hip_delta_unpack_device) and MPI moves
 * the pinned buffers.
 */
#define _POSIX_C_SOURCE 200112L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mpi.h>
#include <pthread.h>

#include "../../../include/cc_hip.h"
#include "../../../include/exchange_plan.h"

#ifndef MPI_UINT32_T
#define MPI_UINT32_T MPI_UNSIGNED
#endif
#ifndef MPI_UINT64_T
#define MPI_UINT64_T MPI_UNSIGNED_LONG_LONG
#endif

static void die(MPI_Comm comm, const char *msg)
{
  fprintf(stderr, "%s\n", msg);
  MPI_Abort(comm, EXIT_FAILURE);
}

static void *xmalloc(size_t bytes, MPI_Comm comm, const char *what)
{
  void *p = malloc(bytes ? bytes : 1);
  if (!p)
    die(comm, what);
  return p;
}

static void hip_or_die(int rc, MPI_Comm comm, const char *what)
{
  if (rc != 0)
  {
    fprintf(stderr, "%s: %s\n", what, cc_hip_last_error());
    MPI_Abort(comm, EXIT_FAILURE);
  }
}

static void delta_cache_drop(const ExchangePlan *P);

void exchangeplan_free(ExchangePlan *P)
{
  if (!P)
    return;
  delta_cache_drop(P);
  free(P->sources);
  free(P->dests);
  free(P->need_from_counts);
  free(P->send_to_counts);
  free(P->need_from_displs);
  free(P->send_to_displs);
  free(P->need_from_flat);
  free(P->send_to_flat);
  free(P->need_from_gidx_flat);
  free(P->sendcounts);
  free(P->sdispls);
  free(P->recvcounts);
  free(P->rdispls);
  free(P->send_labels_flat);
  free(P->recv_labels_flat);
  memset(P, 0, sizeof(*P));
  P->comm_graph = MPI_COMM_NULL;
}

int exchangeplan_build(ExchangePlan *P, const uint32_t *ghost_vertices, uint32_t ghost_count,
                       uint32_t n_global, MPI_Comm comm, OwnerFn owner_fn)
{
  if (!P)
    return 1;
  exchangeplan_free(P);
  MPI_Comm_rank(comm, &P->comm_rank);
  MPI_Comm_size(comm, &P->comm_size);
  const int size = P->comm_size, rank = P->comm_rank;

  P->need_from_counts = (int *)calloc((size_t)size, sizeof(int));
  P->send_to_counts = (int *)calloc((size_t)size, sizeof(int));
  P->need_from_displs = (int *)calloc((size_t)size, sizeof(int));
  P->send_to_displs = (int *)calloc((size_t)size, sizeof(int));
  if (!P->need_from_counts || !P->send_to_counts || !P->need_from_displs || !P->send_to_displs)
    die(comm, "exchangeplan: OOM counts/displs");

  /* Ghosts not owned by this rank, bucketed by owner, each bucket in ghost
   * order (src/exchange_plan.c:74-83). */
  int *owner = (int *)xmalloc((size_t)ghost_count * sizeof(int), comm, "exchangeplan: OOM owner");
  for (uint32_t gi = 0; gi < ghost_count; ++gi)
  {
    owner[gi] = owner_fn(ghost_vertices[gi], n_global, size);
    if (owner[gi] != rank)
      P->need_from_counts[owner[gi]]++;
  }
  MPI_Alltoall(P->need_from_counts, 1, MPI_INT, P->send_to_counts, 1, MPI_INT, comm);
  int tn = 0, ts = 0;
  for (int p = 0; p < size; ++p)
  {
    P->need_from_displs[p] = tn;
    tn += P->need_from_counts[p];
    P->send_to_displs[p] = ts;
    ts += P->send_to_counts[p];
  }
  P->total_need_from = tn;
  P->total_send_to = ts;
  P->need_from_flat = (uint32_t *)xmalloc((size_t)tn * sizeof(uint32_t), comm, "exchangeplan: OOM need_from");
  P->need_from_gidx_flat = (uint32_t *)xmalloc((size_t)tn * sizeof(uint32_t), comm, "exchangeplan: OOM need_gidx");
  P->send_to_flat = (uint32_t *)xmalloc((size_t)ts * sizeof(uint32_t), comm, "exchangeplan: OOM send_to");
  int *fill = (int *)xmalloc((size_t)size * sizeof(int), comm, "exchangeplan: OOM fill");
  memcpy(fill, P->need_from_displs, (size_t)size * sizeof(int));
  for (uint32_t gi = 0; gi < ghost_count; ++gi)
  {
    if (owner[gi] == rank)
      continue;
    int at = fill[owner[gi]]++;
    P->need_from_flat[at] = ghost_vertices[gi];
    P->need_from_gidx_flat[at] = gi;
  }
  free(fill);
  free(owner);
  MPI_Alltoallv(P->need_from_flat, P->need_from_counts, P->need_from_displs, MPI_UINT32_T,
                P->send_to_flat, P->send_to_counts, P->send_to_displs, MPI_UINT32_T, comm);

  for (int p = 0; p < size; ++p)
  {
    P->indegree += P->need_from_counts[p] > 0;
    P->outdegree += P->send_to_counts[p] > 0;
  }
  P->sources = (int *)xmalloc((size_t)P->indegree * sizeof(int), comm, "exchangeplan: OOM sources");
  P->dests = (int *)xmalloc((size_t)P->outdegree * sizeof(int), comm, "exchangeplan: OOM dests");
  for (int p = 0, i = 0, o = 0; p < size; ++p)
  {
    if (P->need_from_counts[p] > 0)
      P->sources[i++] = p;
    if (P->send_to_counts[p] > 0)
      P->dests[o++] = p;
  }
  /* Every rank creates the neighbourhood communicator: the call is
   * collective on comm. */
  MPI_Dist_graph_create_adjacent(comm, P->indegree, P->sources, MPI_UNWEIGHTED, P->outdegree,
                                 P->dests, MPI_UNWEIGHTED, MPI_INFO_NULL, 0, &P->comm_graph);

  P->sendcounts = (int *)xmalloc((size_t)P->outdegree * sizeof(int), comm, "exchangeplan: OOM layouts");
  P->sdispls = (int *)xmalloc((size_t)P->outdegree * sizeof(int), comm, "exchangeplan: OOM layouts");
  P->recvcounts = (int *)xmalloc((size_t)P->indegree * sizeof(int), comm, "exchangeplan: OOM layouts");
  P->rdispls = (int *)xmalloc((size_t)P->indegree * sizeof(int), comm, "exchangeplan: OOM layouts");
  for (int k = 0; k < P->outdegree; ++k)
  {
    P->sendcounts[k] = P->send_to_counts[P->dests[k]];
    P->sdispls[k] = P->send_to_displs[P->dests[k]];
  }
  for (int k = 0; k < P->indegree; ++k)
  {
    P->recvcounts[k] = P->need_from_counts[P->sources[k]];
    P->rdispls[k] = P->need_from_displs[P->sources[k]];
  }
  P->send_labels_flat = (uint32_t *)xmalloc((size_t)ts * sizeof(uint32_t), comm, "exchangeplan: OOM labels");
  P->recv_labels_flat = (uint32_t *)xmalloc((size_t)tn * sizeof(uint32_t), comm, "exchangeplan: OOM labels");
  return 0;
}

/* ---- device-resident exchange -------------------------------------------- */

int exchangeplan_device_init(const ExchangePlan *P, ExchangeDevice *D)
{
  memset(D, 0, sizeof(*D));
  const int outd = P->outdegree, ind = P->indegree;
  const size_t ts = (size_t)P->total_send_to, tn = (size_t)P->total_need_from;
  if (cc_hip_malloc((void **)&D->d_send_to_flat, ts * sizeof(uint32_t)) ||
      cc_hip_malloc((void **)&D->d_need_from_gidx_flat, tn * sizeof(uint32_t)) ||
      cc_hip_host_alloc((void **)&D->h_sendbuf, ts * sizeof(uint64_t)) ||
      cc_hip_host_alloc((void **)&D->h_recvbuf, tn * sizeof(uint64_t)))
    return 1;
  if (cc_hip_memcpy_h2d(D->d_send_to_flat, P->send_to_flat, ts * sizeof(uint32_t), NULL) ||
      cc_hip_memcpy_h2d(D->d_need_from_gidx_flat, P->need_from_gidx_flat, tn * sizeof(uint32_t), NULL))
    return 1;
  D->seg_start = (int *)malloc((size_t)(outd + 1) * sizeof(int));
  D->need_base = (int *)malloc((size_t)(ind ? ind : 1) * sizeof(int));
  D->sendcounts = (int *)calloc((size_t)(outd ? outd : 1), sizeof(int));
  D->sdispls = (int *)calloc((size_t)(outd ? outd : 1), sizeof(int));
  D->recvcounts = (int *)calloc((size_t)(ind ? ind : 1), sizeof(int));
  D->rdispls = (int *)calloc((size_t)(ind ? ind : 1), sizeof(int));
  if (!D->seg_start || !D->need_base || !D->sendcounts || !D->sdispls || !D->recvcounts || !D->rdispls)
    return 1;
  /* dests are ascending ranks and send_to_displs ascend with rank, so the
   * destination segments tile send_to_flat in dests order */
  for (int k = 0; k < outd; ++k)
    D->seg_start[k] = P->send_to_displs[P->dests[k]];
  D->seg_start[outd] = P->total_send_to;
  for (int k = 0; k < ind; ++k)
    D->need_base[k] = P->need_from_displs[P->sources[k]];
  return 0;
}

int exchangeplan_device_prepare(const ExchangePlan *P, ExchangeDevice *D,
                                const uint32_t *d_comp_of, uint32_t v_start, uint32_t v_end)
{
  const size_t ts = (size_t)P->total_send_to;
  if (!D->d_rep_of_send && cc_hip_malloc((void **)&D->d_rep_of_send, ts * sizeof(uint32_t)))
    return 1;
  return cc_hip_delta_prepare_device(D->d_send_to_flat, ts, d_comp_of, v_start, v_end,
                                     D->d_rep_of_send, NULL);
}

void exchangeplan_device_free(ExchangeDevice *D)
{
  if (!D)
    return;
  cc_hip_free(D->d_rep_of_send);
  cc_hip_free(D->d_send_to_flat);
  cc_hip_free(D->d_need_from_gidx_flat);
  cc_hip_host_free(D->h_sendbuf);
  cc_hip_host_free(D->h_recvbuf);
  free(D->seg_start);
  free(D->need_base);
  free(D->sendcounts);
  free(D->sdispls);
  free(D->recvcounts);
  free(D->rdispls);
  memset(D, 0, sizeof(*D));
}

uint64_t exchangeplan_exchange_delta_device(const ExchangePlan *P, ExchangeDevice *D,
                                            const uint32_t *d_comp_label,
                                            const uint32_t *d_comp_of, uint32_t v_start,
                                            uint32_t v_end, uint32_t *d_prev_sent,
                                            uint32_t *d_ghost_labels)
{
  if (!P || P->comm_graph == MPI_COMM_NULL)
    return 0;
  const int outd = P->outdegree, ind = P->indegree;
  uint64_t sent = 0;
  if (outd > 0 && D->d_rep_of_send)
    hip_or_die(cc_hip_delta_pack_reps_device(D->d_send_to_flat, D->d_rep_of_send,
                                             (uint64_t)P->total_send_to, D->seg_start, outd,
                                             d_comp_label, d_prev_sent, D->h_sendbuf,
                                             D->sendcounts, D->sdispls, &sent, NULL),
               P->comm_graph, "exchange delta pack");
  else if (outd > 0)
    hip_or_die(cc_hip_delta_pack_device(D->d_send_to_flat, (uint64_t)P->total_send_to,
                                        D->seg_start, outd, d_comp_label, d_comp_of, v_start,
                                        v_end, d_prev_sent, D->h_sendbuf, D->sendcounts,
                                        D->sdispls, &sent, NULL),
               P->comm_graph, "exchange delta pack");
  MPI_Neighbor_alltoall(D->sendcounts, 1, MPI_INT, D->recvcounts, 1, MPI_INT, P->comm_graph);
  int total_recv = 0;
  for (int k = 0; k < ind; ++k)
  {
    D->rdispls[k] = total_recv;
    total_recv += D->recvcounts[k];
  }
  MPI_Neighbor_alltoallv(D->h_sendbuf, D->sendcounts, D->sdispls, MPI_UINT64_T, D->h_recvbuf,
                         D->recvcounts, D->rdispls, MPI_UINT64_T, P->comm_graph);
  if (total_recv > 0)
    hip_or_die(cc_hip_delta_unpack_device(D->h_recvbuf, (uint64_t)total_recv, D->rdispls,
                                          D->need_base, ind, D->d_need_from_gidx_flat,
                                          d_ghost_labels, NULL),
               P->comm_graph, "exchange delta unpack");
  D->last_received = (uint64_t)total_recv;
  return sent;
}

/* ---- host-array drop-in (include/exchange_plan.h:75-83) ---------------- */

/*
 * Device state of the host-array exchange, kept per plan so that only the
 * first call on a plan allocates: the plan's index lists on the device and
 * the pinned MPI buffers (ExchangeDevice), plus device copies of the
 * caller's comp_label, comp_of and prev_sent, which each call uploads.  An
 * entry is dropped by exchangeplan_free (which exchangeplan_build also
 * calls first), and replaced if a plan at the same address no longer
 * matches it.  The list is guarded by a mutex; calls on one plan from
 * several threads at once are not supported (as in the reference).
 */
typedef struct DeltaCache
{
  const ExchangePlan *plan;
  const uint32_t *send_to_flat; /* identity of the plan it was built for */
  MPI_Comm comm_graph;
  int total_send_to;
  ExchangeDevice D;
  uint32_t *d_cl, *d_co, *d_prev;
  size_t nl_cap; /* entries of d_cl and d_co */
  struct DeltaCache *next;
} DeltaCache;

static DeltaCache *g_delta_cache;
static pthread_mutex_t g_delta_mu = PTHREAD_MUTEX_INITIALIZER;

static void delta_cache_release(DeltaCache *c)
{
  cc_hip_free(c->d_cl);
  cc_hip_free(c->d_co);
  cc_hip_free(c->d_prev);
  exchangeplan_device_free(&c->D);
  free(c);
}

static void delta_cache_drop(const ExchangePlan *P)
{
  pthread_mutex_lock(&g_delta_mu);
  for (DeltaCache **pp = &g_delta_cache; *pp;)
  {
    DeltaCache *c = *pp;
    if (c->plan == P)
    {
      *pp = c->next;
      delta_cache_release(c);
    }
    else
      pp = &c->next;
  }
  pthread_mutex_unlock(&g_delta_mu);
}

/* The cache entry of P with d_cl/d_co holding at least n_local entries. */
static DeltaCache *delta_cache_get(const ExchangePlan *P, size_t n_local, MPI_Comm comm)
{
  pthread_mutex_lock(&g_delta_mu);
  DeltaCache *c = g_delta_cache;
  while (c && c->plan != P)
    c = c->next;
  pthread_mutex_unlock(&g_delta_mu);
  if (c && (c->send_to_flat != P->send_to_flat || c->comm_graph != P->comm_graph ||
            c->total_send_to != P->total_send_to))
  {
    delta_cache_drop(P);
    c = NULL;
  }
  if (!c)
  {
    c = (DeltaCache *)calloc(1, sizeof(DeltaCache));
    if (!c)
      die(comm, "exchangeplan_exchange_delta: OOM cache");
    c->plan = P;
    c->send_to_flat = P->send_to_flat;
    c->comm_graph = P->comm_graph;
    c->total_send_to = P->total_send_to;
    hip_or_die(exchangeplan_device_init(P, &c->D), comm, "exchange delta: device init");
    hip_or_die(cc_hip_malloc((void **)&c->d_prev, (size_t)P->total_send_to * sizeof(uint32_t)),
               comm, "exchange delta: device alloc");
    pthread_mutex_lock(&g_delta_mu);
    c->next = g_delta_cache;
    g_delta_cache = c;
    pthread_mutex_unlock(&g_delta_mu);
  }
  if (c->nl_cap < n_local)
  {
    cc_hip_free(c->d_cl);
    cc_hip_free(c->d_co);
    c->d_cl = c->d_co = NULL;
    hip_or_die(cc_hip_malloc((void **)&c->d_cl, n_local * sizeof(uint32_t)) ||
                   cc_hip_malloc((void **)&c->d_co, n_local * sizeof(uint32_t)),
               comm, "exchange delta: device alloc");
    c->nl_cap = n_local;
  }
  return c;
}

void exchangeplan_exchange_delta(const ExchangePlan *P, const uint32_t *comp_label,
                                 const uint32_t *comp_of, uint32_t v_start, uint32_t v_end,
                                 uint32_t *prev_sent, uint32_t *ghost_labels, MPI_Comm comm,
                                 uint64_t **sendbuf_io, int *sendcap_io, uint64_t **recvbuf_io,
                                 int *recvcap_io)
{
  if (!P || P->comm_graph == MPI_COMM_NULL)
    return;
  const int outd = P->outdegree, ind = P->indegree;
  const uint32_t n_local = v_end - v_start;
  /* reference buffer contract: at least one element, grown with realloc */
  if (*sendbuf_io == NULL)
  {
    *sendbuf_io = (uint64_t *)xmalloc(sizeof(uint64_t), comm, "OOM: delta sendbuf init");
    *sendcap_io = 1;
  }
  if (*recvbuf_io == NULL)
  {
    *recvbuf_io = (uint64_t *)xmalloc(sizeof(uint64_t), comm, "OOM: delta recvbuf init");
    *recvcap_io = 1;
  }
  DeltaCache *c = delta_cache_get(P, n_local, comm);
  ExchangeDevice *D = &c->D;
  for (int k = 0; k < outd; ++k)
    D->sendcounts[k] = D->sdispls[k] = 0;
  if (outd > 0)
  {
    if (*sendcap_io < P->total_send_to)
    {
      uint64_t *nb = (uint64_t *)realloc(*sendbuf_io, (size_t)P->total_send_to * sizeof(uint64_t));
      if (!nb)
        die(comm, "OOM: delta sendbuf grow");
      *sendbuf_io = nb;
      *sendcap_io = P->total_send_to;
    }
    /* Stage the caller's labels on the device and pack there. */
    const size_t nl = (size_t)n_local * sizeof(uint32_t);
    const size_t ps = (size_t)P->total_send_to * sizeof(uint32_t);
    hip_or_die(cc_hip_memcpy_h2d(c->d_cl, comp_label, nl, NULL) ||
                   cc_hip_memcpy_h2d(c->d_co, comp_of, nl, NULL) ||
                   cc_hip_memcpy_h2d(c->d_prev, prev_sent, ps, NULL),
               comm, "exchange delta: upload");
    uint64_t sent = 0;
    hip_or_die(cc_hip_delta_pack_device(D->d_send_to_flat, (uint64_t)P->total_send_to,
                                        D->seg_start, outd, c->d_cl, c->d_co, v_start, v_end,
                                        c->d_prev, D->h_sendbuf, D->sendcounts, D->sdispls, &sent,
                                        NULL),
               comm, "exchange delta pack");
    memcpy(*sendbuf_io, D->h_sendbuf, (size_t)sent * sizeof(uint64_t));
    hip_or_die(cc_hip_memcpy_d2h(prev_sent, c->d_prev, ps, NULL), comm, "exchange delta: download");
  }
  MPI_Neighbor_alltoall(D->sendcounts, 1, MPI_INT, D->recvcounts, 1, MPI_INT, P->comm_graph);
  int total_recv = 0;
  for (int k = 0; k < ind; ++k)
  {
    D->rdispls[k] = total_recv;
    total_recv += D->recvcounts[k];
  }
  if (total_recv > 0 && *recvcap_io < total_recv)
  {
    uint64_t *nb = (uint64_t *)realloc(*recvbuf_io, (size_t)total_recv * sizeof(uint64_t));
    if (!nb)
      die(comm, "OOM: delta recvbuf grow");
    *recvbuf_io = nb;
    *recvcap_io = total_recv;
  }
  MPI_Neighbor_alltoallv(*sendbuf_io, D->sendcounts, D->sdispls, MPI_UINT64_T, *recvbuf_io,
                         D->recvcounts, D->rdispls, MPI_UINT64_T, P->comm_graph);
  /* Unpack: a scatter of at most total_need_from labels (src/exchange_plan.c:390-413). */
  for (int k = 0; k < ind; ++k)
  {
    const int base = D->need_base[k];
    for (int i = 0; i < D->recvcounts[k]; ++i)
    {
      const uint64_t pr = (*recvbuf_io)[D->rdispls[k] + i];
      ghost_labels[P->need_from_gidx_flat[base + (int)(pr >> 32)]] = (uint32_t)pr;
    }
  }
}

/* ---- legacy full-label exchange (include/exchange_plan.h:90-107) --------- */

void exchangeplan_exchange_start(ExchangePlan *P, MPI_Request *req)
{
  if (!P || !req)
    return;
  *req = MPI_REQUEST_NULL;
  if (P->comm_graph != MPI_COMM_NULL && (P->total_send_to > 0 || P->total_need_from > 0))
    MPI_Ineighbor_alltoallv(P->send_labels_flat, P->sendcounts, P->sdispls, MPI_UINT32_T,
                            P->recv_labels_flat, P->recvcounts, P->rdispls, MPI_UINT32_T,
                            P->comm_graph, req);
}

void exchangeplan_exchange_finish(ExchangePlan *P, uint32_t *ghost_labels, MPI_Request *req)
{
  if (!P || !ghost_labels || !req)
    return;
  if (*req != MPI_REQUEST_NULL)
    MPI_Wait(req, MPI_STATUS_IGNORE);
  for (int i = 0; i < P->total_need_from; ++i)
    ghost_labels[P->need_from_gidx_flat[i]] = P->recv_labels_flat[i];
}

void exchangeplan_exchange(ExchangePlan *P, uint32_t *ghost_labels, MPI_Comm comm)
{
  (void)comm;
  MPI_Request req;
  exchangeplan_exchange_start(P, &req);
  exchangeplan_exchange_finish(P, ghost_labels, &req);
}
