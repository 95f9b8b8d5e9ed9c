/*
 * MPI check of the delta exchange of include/exchange_plan.h
 * (exchangeplan_build + exchangeplan_exchange_delta, reference
 * src/exchange_plan.c:48-225 and :242-425) against the oracle.
 *
 * The same source is linked either with the reference's exchange_plan.c
 * compiled from its own sources (oracle/_ref/libccref_mpi.so) or with this
 * project's libccmpi.so, whose pack runs on the GPU.  Each rank owns a block
 * of n vertices, makes a seeded sorted list of ghosts owned by other ranks,
 * builds the plan, and then for `rounds` exchanges lowers a seeded subset of
 * its component labels (monotone, as in the merge loop) and calls
 * exchangeplan_exchange_delta.  After every call it checks:
 *   - prev_sent and the send buffer's emitted pairs equal oracle_delta_pack
 *     run on the same plan and labels;
 *   - ghost_labels equal oracle_delta_unpack of the pairs every other rank's
 *     oracle pack produced, moved by a plain MPI_Alltoallv of this program,
 *     and the receive buffer holds exactly those pairs.
 * Every rank prints one digest line per plan and per round (FNV-1a of the
 * observable state), so two builds can be compared line for line, and rank
 * 0 prints "delta_check OK" when every rank passed.
 *
 * usage: delta_check N ROUNDS SEED
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mpi.h>

#include "exchange_plan.h"

/* oracle/cc_oracle.c */
uint64_t oracle_delta_pack(const uint32_t *send_to_flat, const int *seg_start, int nseg,
                           const uint32_t *comp_label, const uint32_t *comp_of, uint32_t v_start,
                           uint32_t v_end, uint32_t *prev_sent, uint64_t *sendbuf, int *sendcounts,
                           int *sdispls);
void oracle_delta_unpack(const uint64_t *recvbuf, const int *recvcounts, const int *rdispls,
                         const int *need_base, int nsrc, const uint32_t *need_from_gidx_flat,
                         uint32_t *ghost_labels);

static int g_rank;

#define FAIL(...)                                  \
  do                                               \
  {                                                \
    fprintf(stderr, "rank %d: ", g_rank);          \
    fprintf(stderr, __VA_ARGS__);                  \
    fprintf(stderr, "\n");                         \
    MPI_Abort(MPI_COMM_WORLD, 3);                  \
  } while (0)

static uint64_t mix(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t fnv(uint64_t h, const void *data, size_t bytes)
{
  const unsigned char *p = (const unsigned char *)data;
  for (size_t i = 0; i < bytes; ++i)
  {
    h ^= p[i];
    h *= 0x100000001B3ull;
  }
  return h;
}

/* Block partition: the first n % P ranks own one vertex more. */
static void block_range(uint32_t n, int P, int r, uint32_t *vs, uint32_t *ve)
{
  const uint32_t base = n / (uint32_t)P, rem = n % (uint32_t)P, ur = (uint32_t)r;
  *vs = ur * base + (ur < rem ? ur : rem);
  *ve = *vs + base + (ur < rem ? 1 : 0);
}

static int owner_block(uint32_t v, uint32_t n, int P)
{
  int lo = 0, hi = P - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) / 2;
    uint32_t s, e;
    block_range(n, P, mid, &s, &e);
    if (v >= s)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

static int cmp_u32(const void *a, const void *b)
{
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

int main(int argc, char **argv)
{
  MPI_Init(&argc, &argv);
  int P;
  MPI_Comm_rank(MPI_COMM_WORLD, &g_rank);
  MPI_Comm_size(MPI_COMM_WORLD, &P);
  const int rank = g_rank;
  if (argc < 4)
    FAIL("usage: delta_check N ROUNDS SEED");
  const uint32_t n = (uint32_t)strtoul(argv[1], NULL, 10);
  const int rounds = atoi(argv[2]);
  const uint64_t seed = strtoull(argv[3], NULL, 10);
  if (P < 2 || n < (uint32_t)(4 * P))
    FAIL("needs at least 2 ranks and 4 vertices per rank");
  uint32_t vs, ve;
  block_range(n, P, rank, &vs, &ve);
  const uint32_t nl = ve - vs;

  /* ghosts: sorted unique remote vertices (about 3 per owned vertex drawn,
   * plus one vertex of every other rank, so every pair of ranks talks) */
  const uint32_t draws = 3 * nl + (uint32_t)P;
  uint32_t *ghosts = (uint32_t *)malloc((size_t)draws * sizeof(uint32_t));
  uint32_t ng = 0;
  for (int q = 0; q < P; ++q)
    if (q != rank)
    {
      uint32_t s, e;
      block_range(n, P, q, &s, &e);
      ghosts[ng++] = s + (uint32_t)(mix(seed ^ (uint64_t)(rank * 131 + q)) % (e - s));
    }
  for (uint32_t i = 0; ng < draws && i < 3 * nl; ++i)
  {
    const uint32_t v = (uint32_t)(mix(seed * 7919u + (uint64_t)rank * 1000003u + i) % n);
    if (v < vs || v >= ve)
      ghosts[ng++] = v;
  }
  qsort(ghosts, ng, sizeof(uint32_t), cmp_u32);
  uint32_t u = 0;
  for (uint32_t i = 0; i < ng; ++i)
    if (u == 0 || ghosts[u - 1] != ghosts[i])
      ghosts[u++] = ghosts[i];
  ng = u;

  ExchangePlan plan;
  memset(&plan, 0, sizeof(plan));
  plan.comm_graph = MPI_COMM_NULL;
  if (exchangeplan_build(&plan, ghosts, ng, n, MPI_COMM_WORLD, owner_block) != 0)
    FAIL("exchangeplan_build failed");
  {
    uint64_t h = 0xCBF29CE484222325ull;
    h = fnv(h, &plan.indegree, sizeof(int));
    h = fnv(h, &plan.outdegree, sizeof(int));
    h = fnv(h, plan.sources, (size_t)plan.indegree * sizeof(int));
    h = fnv(h, plan.dests, (size_t)plan.outdegree * sizeof(int));
    h = fnv(h, plan.need_from_counts, (size_t)P * sizeof(int));
    h = fnv(h, plan.send_to_counts, (size_t)P * sizeof(int));
    h = fnv(h, plan.need_from_flat, (size_t)plan.total_need_from * sizeof(uint32_t));
    h = fnv(h, plan.need_from_gidx_flat, (size_t)plan.total_need_from * sizeof(uint32_t));
    h = fnv(h, plan.send_to_flat, (size_t)plan.total_send_to * sizeof(uint32_t));
    printf("plan rank=%d ghosts=%u need=%d send=%d in=%d out=%d hash=%016llx\n", rank, ng,
           plan.total_need_from, plan.total_send_to, plan.indegree, plan.outdegree,
           (unsigned long long)h);
  }

  /* a local forest: every vertex under one of about nl/8 reps */
  uint32_t *comp_of = (uint32_t *)malloc((size_t)nl * sizeof(uint32_t));
  uint32_t *comp_label = (uint32_t *)malloc((size_t)nl * sizeof(uint32_t));
  const uint32_t nreps = nl / 8 + 1;
  for (uint32_t li = 0; li < nl; ++li)
  {
    comp_of[li] = (uint32_t)(mix(seed + 17u * li) % nreps) * 8u;
    if (comp_of[li] >= nl)
      comp_of[li] = 0;
    comp_label[li] = vs + li;
  }
  for (uint32_t r = 0; r < nl; r += 8)
    comp_of[r] = r;

  const int ts = plan.total_send_to, tn = plan.total_need_from;
  const int outd = plan.outdegree, ind = plan.indegree;
  uint32_t *prev_sent = (uint32_t *)malloc((size_t)(ts ? ts : 1) * sizeof(uint32_t));
  uint32_t *prev_o = (uint32_t *)malloc((size_t)(ts ? ts : 1) * sizeof(uint32_t));
  for (int i = 0; i < ts; ++i)
    prev_sent[i] = UINT32_MAX;
  uint32_t *gl = (uint32_t *)malloc((size_t)(ng ? ng : 1) * sizeof(uint32_t));
  uint32_t *gl_o = (uint32_t *)malloc((size_t)(ng ? ng : 1) * sizeof(uint32_t));
  for (uint32_t g = 0; g < ng; ++g)
    gl[g] = ghosts[g];
  int *seg = (int *)malloc((size_t)(outd + 1) * sizeof(int));
  for (int k = 0; k < outd; ++k)
    seg[k] = plan.send_to_displs[plan.dests[k]];
  seg[outd] = ts;
  uint64_t *pairs_o = (uint64_t *)malloc((size_t)(ts ? ts : 1) * sizeof(uint64_t));
  int *cnt_o = (int *)calloc((size_t)(outd ? outd : 1), sizeof(int));
  int *dis_o = (int *)calloc((size_t)(outd ? outd : 1), sizeof(int));
  int *scount = (int *)calloc((size_t)P, sizeof(int)), *sdisp = (int *)calloc((size_t)P, sizeof(int));
  int *rcount = (int *)calloc((size_t)P, sizeof(int)), *rdisp = (int *)calloc((size_t)P, sizeof(int));
  uint64_t *recv_o = (uint64_t *)malloc((size_t)(tn ? tn : 1) * sizeof(uint64_t));
  int *rc_src = (int *)calloc((size_t)(ind ? ind : 1), sizeof(int));
  int *rd_src = (int *)calloc((size_t)(ind ? ind : 1), sizeof(int));
  int *need_base = (int *)calloc((size_t)(ind ? ind : 1), sizeof(int));
  for (int k = 0; k < ind; ++k)
    need_base[k] = plan.need_from_displs[plan.sources[k]];

  uint64_t *sendbuf = NULL, *recvbuf = NULL;
  int sendcap = 0, recvcap = 0;
  for (int round = 0; round < rounds; ++round)
  {
    /* lower a seeded subset of labels (round 0: none, so the first exchange
     * sends every vertex's initial label; later rounds lower 1/(round+1) of
     * the reps, some to ids of other ranks' blocks) */
    if (round > 0)
      for (uint32_t r = 0; r < nl; r += 8)
      {
        const uint64_t z = mix(seed * 31u + (uint64_t)round * 1000003u + r + (uint64_t)rank * 7777u);
        if (z % (uint64_t)(round + 1) == 0)
        {
          const uint32_t cand = (uint32_t)((z >> 20) % ((uint64_t)comp_label[r] + 1));
          if (cand < comp_label[r])
            comp_label[r] = cand;
        }
      }
    memcpy(prev_o, prev_sent, (size_t)ts * sizeof(uint32_t));
    memcpy(gl_o, gl, (size_t)ng * sizeof(uint32_t));
    const uint64_t total_o = oracle_delta_pack(plan.send_to_flat, seg, outd, comp_label, comp_of,
                                               vs, ve, prev_o, pairs_o, cnt_o, dis_o);

    exchangeplan_exchange_delta(&plan, comp_label, comp_of, vs, ve, prev_sent, gl, MPI_COMM_WORLD,
                                &sendbuf, &sendcap, &recvbuf, &recvcap);

    if (memcmp(prev_sent, prev_o, (size_t)ts * sizeof(uint32_t)) != 0)
      FAIL("round %d: prev_sent differs from the oracle's pack", round);
    if (total_o > 0 && memcmp(sendbuf, pairs_o, (size_t)total_o * sizeof(uint64_t)) != 0)
      FAIL("round %d: send buffer differs from the oracle's %llu pairs", round,
           (unsigned long long)total_o);

    /* the oracle's pairs moved by a plain all-to-all, then unpacked */
    memset(scount, 0, (size_t)P * sizeof(int));
    for (int k = 0; k < outd; ++k)
      scount[plan.dests[k]] = cnt_o[k];
    MPI_Alltoall(scount, 1, MPI_INT, rcount, 1, MPI_INT, MPI_COMM_WORLD);
    int sd = 0, rd = 0;
    for (int q = 0; q < P; ++q)
    {
      sdisp[q] = sd;
      sd += scount[q];
      rdisp[q] = rd;
      rd += rcount[q];
    }
    if (rd > tn)
      FAIL("round %d: %d pairs received, more than the %d ghosts needed", round, rd, tn);
    MPI_Alltoallv(pairs_o, scount, sdisp, MPI_UINT64_T, recv_o, rcount, rdisp, MPI_UINT64_T,
                  MPI_COMM_WORLD);
    int off = 0;
    for (int k = 0; k < ind; ++k)
    {
      rc_src[k] = rcount[plan.sources[k]];
      rd_src[k] = off;
      off += rc_src[k];
    }
    if (off != rd)
      FAIL("round %d: pairs from a rank that is not a source", round);
    oracle_delta_unpack(recv_o, rc_src, rd_src, need_base, ind, plan.need_from_gidx_flat, gl_o);
    if (memcmp(gl, gl_o, (size_t)ng * sizeof(uint32_t)) != 0)
      FAIL("round %d: ghost_labels differ from the oracle's unpack", round);
    if (rd > 0 && memcmp(recvbuf, recv_o, (size_t)rd * sizeof(uint64_t)) != 0)
      FAIL("round %d: receive buffer differs from the pairs sent", round);

    uint64_t h = 0xCBF29CE484222325ull;
    h = fnv(h, prev_sent, (size_t)ts * sizeof(uint32_t));
    h = fnv(h, gl, (size_t)ng * sizeof(uint32_t));
    h = fnv(h, sendbuf, (size_t)total_o * sizeof(uint64_t));
    h = fnv(h, recvbuf, (size_t)rd * sizeof(uint64_t));
    printf("round rank=%d round=%d sent=%llu received=%d hash=%016llx\n", rank, round,
           (unsigned long long)total_o, rd, (unsigned long long)h);
  }
  fflush(stdout);
  MPI_Barrier(MPI_COMM_WORLD);
  if (rank == 0)
  {
    printf("delta_check OK ranks=%d n=%u rounds=%d\n", P, n, rounds);
    fflush(stdout);
  }
  exchangeplan_free(&plan);
  free(sendbuf);
  free(recvbuf);
  free(ghosts);
  free(comp_of);
  free(comp_label);
  free(prev_sent);
  free(prev_o);
  free(gl);
  free(gl_o);
  free(seg);
  free(pairs_o);
  free(cnt_o);
  free(dis_o);
  free(scount);
  free(sdisp);
  free(rcount);
  free(rdisp);
  free(recv_o);
  free(rc_src);
  free(rd_src);
  free(need_base);
  MPI_Finalize();
  return 0;
}
