#ifndef EXCHANGE_PLAN_H
#define EXCHANGE_PLAN_H

/*
 * Halo-exchange plan and delta exchange: the drop-in for the reference's
 * include/exchange_plan.h:7-107 (same types and prototypes), plus a
 * device-resident variant whose pack and unpack run on the GPU.  MPI stays
 * on the host thread.
 */

#include <stdint.h>
#include <mpi.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/exchange_plan.h:7 */
typedef int (*OwnerFn)(uint32_t v, uint32_t n_global, int comm_size);

/* include/exchange_plan.h:10-40 (field for field). */
typedef struct
{
  int comm_size;
  int comm_rank;

  int indegree;
  int outdegree;
  int *sources;
  int *dests;
  MPI_Comm comm_graph;

  int *need_from_counts;
  int *send_to_counts;
  int *need_from_displs;
  int *send_to_displs;

  uint32_t *need_from_flat;
  uint32_t *send_to_flat;
  uint32_t *need_from_gidx_flat;

  int total_need_from;
  int total_send_to;

  int *sendcounts;
  int *sdispls;
  int *recvcounts;
  int *rdispls;

  uint32_t *send_labels_flat;
  uint32_t *recv_labels_flat;
} ExchangePlan;

/*
 * include/exchange_plan.h:51-56 (src/exchange_plan.c:48-225).  ghost_vertices
 * sorted unique.  Collective on comm.  Unlike the reference, every rank
 * creates the neighbourhood communicator (a rank with no neighbours gets an
 * empty one), so ranks without ghosts may call it too without hanging the
 * others.  Returns 0; aborts via MPI_Abort on allocation failure.
 */
int exchangeplan_build(ExchangePlan *P, const uint32_t *ghost_vertices, uint32_t ghost_count,
                       uint32_t n_global, MPI_Comm comm, OwnerFn owner_fn);

/* include/exchange_plan.h:59.  Safe on a zeroed or partially built plan. */
void exchangeplan_free(ExchangePlan *P);

/*
 * include/exchange_plan.h:75-83 (src/exchange_plan.c:242-425) with host
 * arrays: the labels and prev_sent are staged to the device, the changed
 * labels are packed there into a pinned send buffer, MPI moves them, and
 * the received pairs are applied to ghost_labels.  Buffers in *_io keep the
 * reference's meaning (grown as needed; owned by the caller, freed with
 * free()).
 */
void exchangeplan_exchange_delta(const ExchangePlan *P, const uint32_t *comp_label,
                                 const uint32_t *comp_of, uint32_t v_start, uint32_t v_end,
                                 uint32_t *prev_sent, uint32_t *ghost_labels, MPI_Comm comm,
                                 uint64_t **sendbuf_io, int *sendcap_io, uint64_t **recvbuf_io,
                                 int *recvcap_io);

/* Legacy full-label exchange (include/exchange_plan.h:90-107; unused by the
 * reference's driver, kept for interface parity). */
void exchangeplan_exchange_start(ExchangePlan *P, MPI_Request *req);
void exchangeplan_exchange_finish(ExchangePlan *P, uint32_t *ghost_labels, MPI_Request *req);
void exchangeplan_exchange(ExchangePlan *P, uint32_t *ghost_labels, MPI_Comm comm);

/* ---- device-resident variant (no reference counterpart) ---------------- */

/* Device copies of a plan's index lists plus pinned MPI buffers. */
typedef struct
{
  uint32_t *d_send_to_flat;
  uint32_t *d_need_from_gidx_flat;
  uint64_t *h_sendbuf; /* pinned, total_send_to pairs */
  uint64_t *h_recvbuf; /* pinned, total_need_from pairs */
  int *seg_start;      /* outdegree+1 segment starts of send_to_flat in dests order */
  int *need_base;      /* need_from_displs of each source, in sources order */
  int *sendcounts, *sdispls, *recvcounts, *rdispls;
  uint32_t *d_rep_of_send; /* send entries' reps (exchangeplan_device_prepare), or NULL */
  uint64_t last_received;  /* pairs the last exchange delivered into d_ghost_labels */
} ExchangeDevice;

int exchangeplan_device_init(const ExchangePlan *P, ExchangeDevice *D);
void exchangeplan_device_free(ExchangeDevice *D);

/*
 * Gather the send entries' reps once (cc_hip_delta_prepare_device) for a
 * comp_of that stays fixed from here on, as it does through the merge loop:
 * the packs of exchangeplan_exchange_delta_device then gather only
 * comp_label at the reps.  Returns 0, or nonzero on a device failure.
 */
int exchangeplan_device_prepare(const ExchangePlan *P, ExchangeDevice *D,
                                const uint32_t *d_comp_of, uint32_t v_start, uint32_t v_end);

/*
 * exchangeplan_exchange_delta with comp_label, comp_of, prev_sent and
 * ghost_labels resident on the device: the pack compacts the changed pairs
 * in HBM and copies them into the pinned send buffer, MPI exchanges them, and the unpack
 * kernel scatters the received pairs into d_ghost_labels.  With a prepared
 * rep list (exchangeplan_device_prepare) d_comp_of is not read.  Returns the
 * number of pairs this rank sent.
 */
uint64_t exchangeplan_exchange_delta_device(const ExchangePlan *P, ExchangeDevice *D,
                                            const uint32_t *d_comp_label,
                                            const uint32_t *d_comp_of, uint32_t v_start,
                                            uint32_t v_end, uint32_t *d_prev_sent,
                                            uint32_t *d_ghost_labels);

#ifdef __cplusplus
}
#endif

#endif /* EXCHANGE_PLAN_H */
