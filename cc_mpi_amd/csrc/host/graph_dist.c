/*
 * Row partitioning and distribution of a CSR held by rank 0 (host C + MPI).
 * Partition math follows the reference's src/graph_dist.c:22-92 and
 * src/cc_mpi.c:41-76; the scatter follows src/graph_dist.c:142-342 (row_ptr
 * windows rebased to 0, col_idx slices sent in INT_MAX-sized pieces).
 */
#define _POSIX_C_SOURCE 200112L
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <mpi.h>

#include "../../../include/graph_dist.h"

#ifndef MPI_UINT32_T
#define MPI_UINT32_T MPI_UNSIGNED
#endif
#ifndef MPI_UINT64_T
#define MPI_UINT64_T MPI_UNSIGNED_LONG_LONG
#endif

void cc_partition_block_range(uint32_t n, int comm_size, int rank, uint32_t *v_start,
                              uint32_t *v_end)
{
  if (comm_size <= 0)
  {
    *v_start = *v_end = 0;
    return;
  }
  const uint32_t P = (uint32_t)comm_size, r = (uint32_t)rank;
  const uint32_t base = n / P, rem = n % P;
  if (r < rem)
  {
    *v_start = r * (base + 1);
    *v_end = *v_start + base + 1;
  }
  else
  {
    *v_start = rem * (base + 1) + (r - rem) * base;
    *v_end = *v_start + base;
  }
}

void cc_partition_weighted_bounds(const uint64_t *row_ptr, uint32_t n, int comm_size,
                                  uint32_t vertex_weight, uint32_t *bounds)
{
  if (comm_size <= 0)
    return;
  const uint64_t total = (row_ptr ? row_ptr[n] : 0) + (uint64_t)vertex_weight * n;
  bounds[0] = 0;
  uint32_t v = 0;
  uint64_t cum = 0;
  for (int r = 1; r < comm_size; ++r)
  {
    const uint64_t target = total * (uint64_t)r / (uint64_t)comm_size;
    while (v < n && cum < target)
    {
      cum += row_ptr[v + 1] - row_ptr[v] + vertex_weight;
      ++v;
    }
    const uint32_t lo = bounds[r - 1] + 1u, hi = n - (uint32_t)(comm_size - r);
    if (v < lo)
      v = lo;
    if (v > hi)
      v = hi;
    bounds[r] = v;
  }
  bounds[comm_size] = n;
}

int cc_partition_owner(uint32_t v, uint32_t n_global, int comm_size, const uint32_t *bounds,
                       int part_kind)
{
  if (part_kind != 0 && bounds && comm_size > 1)
  {
    int lo = 0, hi = comm_size - 1;
    while (lo <= hi)
    {
      const int mid = lo + ((hi - lo) >> 1);
      if (v < bounds[mid])
        hi = mid - 1;
      else if (v >= bounds[mid + 1])
        lo = mid + 1;
      else
        return mid;
    }
    return v >= bounds[comm_size] ? comm_size - 1 : 0;
  }
  if (comm_size <= 1)
    return 0;
  const uint32_t P = (uint32_t)comm_size, base = n_global / P, rem = n_global % P;
  const uint64_t split = (uint64_t)(base + 1) * rem;
  if ((uint64_t)v < split)
    return (int)(v / (base + 1));
  if (base == 0)
    return (int)rem;
  return (int)(rem + (uint32_t)(((uint64_t)v - split) / base));
}

void free_dist_csr(DistCSRGraph *g)
{
  if (!g)
    return;
  free(g->row_ptr);
  free(g->col_idx);
  free(g->part_bounds);
  memset(g, 0, sizeof(*g));
}

static void range_of(const DistCSRGraph *g, uint32_t n, int size, int r, uint32_t *a, uint32_t *b)
{
  if (g->part_kind == 1)
  {
    *a = g->part_bounds[r];
    *b = g->part_bounds[r + 1];
  }
  else
  {
    cc_partition_block_range(n, size, r, a, b);
  }
}

int dist_csr_from_global(const CSRGraph *full, int use_edge, DistCSRGraph *out, MPI_Comm comm)
{
  int rank, size;
  MPI_Comm_rank(comm, &rank);
  MPI_Comm_size(comm, &size);
  memset(out, 0, sizeof(*out));
  uint32_t n = (rank == 0) ? full->n : 0;
  uint64_t m = (rank == 0) ? full->m : 0;
  MPI_Bcast(&n, 1, MPI_UINT32_T, 0, comm);
  MPI_Bcast(&m, 1, MPI_UINT64_T, 0, comm);
  out->n_global = n;
  out->m_global = m;
  if (use_edge && size > 1)
  {
    out->part_bounds = (uint32_t *)malloc((size_t)(size + 1) * sizeof(uint32_t));
    if (!out->part_bounds)
      return 1;
    if (rank == 0)
      cc_partition_weighted_bounds(full->row_ptr, n, size, 1, out->part_bounds);
    MPI_Bcast(out->part_bounds, size + 1, MPI_UINT32_T, 0, comm);
    out->part_size = size;
    out->part_kind = 1;
  }
  range_of(out, n, size, rank, &out->v_start, &out->v_end);
  out->n_local = out->v_end - out->v_start;

  /* row_ptr window (n_local + 1 entries), rebased to 0 */
  out->row_ptr = (uint64_t *)malloc(((size_t)out->n_local + 1) * sizeof(uint64_t));
  if (!out->row_ptr)
    return 1;
  if (rank == 0)
  {
    for (int r = 0; r < size; ++r)
    {
      uint32_t a, b;
      range_of(out, n, size, r, &a, &b);
      if (r == 0)
        memcpy(out->row_ptr, full->row_ptr + a, ((size_t)(b - a) + 1) * sizeof(uint64_t));
      else
        MPI_Send(full->row_ptr + a, (int)(b - a + 1), MPI_UINT64_T, r, 1, comm);
    }
  }
  else
  {
    MPI_Recv(out->row_ptr, (int)(out->n_local + 1), MPI_UINT64_T, 0, 1, comm, MPI_STATUS_IGNORE);
  }
  const uint64_t base = out->row_ptr[0];
  for (uint32_t i = 0; i <= out->n_local; ++i)
    out->row_ptr[i] -= base;
  out->m_local = out->row_ptr[out->n_local];

  /* col_idx slice, in pieces of at most INT_MAX entries */
  out->col_idx = (uint32_t *)malloc((size_t)(out->m_local ? out->m_local : 1) * sizeof(uint32_t));
  if (!out->col_idx)
    return 1;
  if (rank == 0)
  {
    for (int r = 0; r < size; ++r)
    {
      uint32_t a, b;
      range_of(out, n, size, r, &a, &b);
      const uint64_t off = full->row_ptr[a], cnt = full->row_ptr[b] - off;
      if (r == 0)
      {
        memcpy(out->col_idx, full->col_idx + off, (size_t)cnt * sizeof(uint32_t));
        continue;
      }
      for (uint64_t s = 0; s < cnt;)
      {
        const int piece = (int)((cnt - s) > (uint64_t)INT_MAX ? INT_MAX : (cnt - s));
        MPI_Send(full->col_idx + off + s, piece, MPI_UINT32_T, r, 2, comm);
        s += (uint64_t)piece;
      }
    }
  }
  else
  {
    for (uint64_t s = 0; s < out->m_local;)
    {
      const int piece = (int)((out->m_local - s) > (uint64_t)INT_MAX ? INT_MAX : (out->m_local - s));
      MPI_Recv(out->col_idx + s, piece, MPI_UINT32_T, 0, 2, comm, MPI_STATUS_IGNORE);
      s += (uint64_t)piece;
    }
  }
  return 0;
}

int load_dist_csr_from_file(const char *path, int symmetrize, int drop_self_loops, int use_edge,
                            DistCSRGraph *out, MPI_Comm comm)
{
  const char *ext = strrchr(path, '.');
  if (!ext || strcasecmp(ext, ".mat") != 0)
    return -1;
  int rank;
  MPI_Comm_rank(comm, &rank);
  CSRGraph full;
  memset(&full, 0, sizeof(full));
  int rc = 0;
  if (rank == 0)
    rc = load_csr_from_file(path, symmetrize, drop_self_loops, &full);
  MPI_Bcast(&rc, 1, MPI_INT, 0, comm);
  if (rc != 0)
  {
    if (rank == 0)
      free_csr(&full);
    return rc;
  }
  rc = dist_csr_from_global(&full, use_edge, out, comm);
  if (rank == 0)
    free_csr(&full);
  return rc;
}
