/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatements of the reference's per-rank connected-components path,
 * used as the checker for the HIP library and as the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; the product library (libccmi355x.so) never links,
 * calls or falls back to it.
 *
 * Each function names the reference file:line it restates.  All of it is
 * this project's own code.  How each function is pinned (golden vectors,
 * the reference compiled from its own sources into oracle/_ref, closed
 * forms, brute force) is recorded in the test files that check it; every
 * function below is pinned by at least one such check.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Connected components                                               */
/* ------------------------------------------------------------------ */

static uint32_t uf_root(uint32_t *parent, uint32_t x)
{
  uint32_t r = x;
  while (parent[r] != r)
    r = parent[r];
  while (parent[x] != r)
  {
    uint32_t nx = parent[x];
    parent[x] = r;
    x = nx;
  }
  return r;
}

/*
 * labels[v] = minimum vertex id of v's component: the closed form of the
 * fixed point of src/cc_pthreads.c:32-73 (labels start at the id,
 * :169-170, and are only ever lowered to a neighbour's label, :43-67).
 * Sequential union-find that always hangs the larger root under the
 * smaller, so every root is its tree's minimum.  Entries >= n are ignored.
 */
void oracle_cc_min_label(uint32_t n, const uint64_t *row_ptr, const uint32_t *col_idx,
                         uint32_t *labels)
{
  for (uint32_t v = 0; v < n; ++v)
    labels[v] = v;
  for (uint32_t u = 0; u < n; ++u)
  {
    for (uint64_t j = row_ptr[u]; j < row_ptr[u + 1]; ++j)
    {
      uint32_t w = col_idx[j];
      if (w >= n)
        continue;
      uint32_t a = uf_root(labels, u), b = uf_root(labels, w);
      if (a == b)
        continue;
      if (a < b)
        labels[b] = a;
      else
        labels[a] = b;
    }
  }
  for (uint32_t v = 0; v < n; ++v)
    labels[v] = uf_root(labels, v);
}

/* ---- pthreads label propagation (src/cc_pthreads.c:14-229) ---------- */

typedef struct
{
  const uint64_t *row_ptr;
  const uint32_t *col_idx;
  _Atomic uint32_t *lab;
  atomic_int *flag;            /* 1 = changed this round, 0 = not, -1 = done */
  _Atomic uint64_t *cursor;    /* dynamic chunk cursor */
  pthread_barrier_t *bar;
  uint32_t n;
  uint32_t chunk;              /* 0 = static blocks */
  uint32_t lo, hi;             /* static block */
  int tid;
  int rounds;                  /* rounds completed (thread 0 only) */
} LpWorker;

static int g_lp_verbose = 0;

/* Print one line to stderr per completed round of oracle_cc_lp_pthreads
 * (progress for long runs). */
void oracle_set_verbose(int on) { g_lp_verbose = on; }

/* One vertex: pull-min over the row, lower own label, push the new label to
 * every neighbour (src/cc_pthreads.c:32-73).  Returns 1 if u improved. */
static int lp_relax(uint32_t u, const uint64_t *row_ptr, const uint32_t *col_idx,
                    _Atomic uint32_t *lab)
{
  uint32_t mine = atomic_load_explicit(&lab[u], memory_order_relaxed);
  uint32_t best = mine;
  const uint64_t b = row_ptr[u], e = row_ptr[u + 1];
  for (uint64_t j = b; j < e; ++j)
  {
    uint32_t l = atomic_load_explicit(&lab[col_idx[j]], memory_order_relaxed);
    if (l < best)
      best = l;
  }
  if (best >= mine)
    return 0;
  uint32_t cur = mine;
  while (cur > best &&
         !atomic_compare_exchange_weak_explicit(&lab[u], &cur, best, memory_order_relaxed,
                                                memory_order_relaxed))
    ;
  for (uint64_t j = b; j < e; ++j)
  {
    _Atomic uint32_t *t = &lab[col_idx[j]];
    uint32_t c = atomic_load_explicit(t, memory_order_relaxed);
    while (c > best &&
           !atomic_compare_exchange_weak_explicit(t, &c, best, memory_order_relaxed,
                                                  memory_order_relaxed))
      ;
  }
  return 1;
}

/* Round loop of src/cc_pthreads.c:76-150: cursor reset, barrier, relax a
 * dynamic chunk at a time (or the static block), publish "changed",
 * barrier, thread 0 turns 0 into -1 (done) or clears it, barrier. */
static void *lp_thread(void *arg)
{
  LpWorker *w = (LpWorker *)arg;
  for (;;)
  {
    int any = 0;
    if (w->chunk && w->tid == 0)
      atomic_store_explicit(w->cursor, 0, memory_order_relaxed);
    pthread_barrier_wait(w->bar);
    if (w->chunk)
    {
      for (;;)
      {
        uint64_t s = atomic_fetch_add_explicit(w->cursor, w->chunk, memory_order_relaxed);
        if (s >= w->n)
          break;
        uint64_t e = s + w->chunk;
        if (e > w->n)
          e = w->n;
        for (uint64_t u = s; u < e; ++u)
          any |= lp_relax((uint32_t)u, w->row_ptr, w->col_idx, w->lab);
      }
    }
    else
    {
      for (uint32_t u = w->lo; u < w->hi; ++u)
        any |= lp_relax(u, w->row_ptr, w->col_idx, w->lab);
    }
    if (any)
      atomic_store_explicit(w->flag, 1, memory_order_relaxed);
    pthread_barrier_wait(w->bar);
    if (w->tid == 0)
    {
      int f = atomic_load_explicit(w->flag, memory_order_acquire);
      atomic_store_explicit(w->flag, f == 0 ? -1 : 0, memory_order_release);
      if (g_lp_verbose)
        fprintf(stderr, "[oracle lp] round %d done (%s)\n", ++w->rounds,
                f == 0 ? "converged" : "changed");
    }
    pthread_barrier_wait(w->bar);
    if (atomic_load_explicit(w->flag, memory_order_acquire) == -1)
      break;
  }
  return NULL;
}

/*
 * Restatement of compute_connected_components_pthreads
 * (src/cc_pthreads.c:152-229): chunk_size 1 = static blocks of
 * ceil(n/T), > 0 = dynamic chunks of that size, <= 0 = 4096
 * (include/cc.h:7).  Returns 0, or -1 on allocation failure.
 */
int oracle_cc_lp_pthreads(uint32_t n, const uint64_t *row_ptr, const uint32_t *col_idx,
                          uint32_t *labels, int num_threads, int chunk_size)
{
  if (num_threads < 1)
    num_threads = 1;
  const int dynamic = (chunk_size != 1);
  const uint32_t chunk = dynamic ? (uint32_t)(chunk_size > 0 ? chunk_size : 4096) : 0;
  const uint32_t block = (uint32_t)(((uint64_t)n + (uint64_t)num_threads - 1) / (uint64_t)num_threads);

  _Atomic uint32_t *lab = NULL;
  if (posix_memalign((void **)&lab, 64, ((size_t)n + 1) * sizeof(*lab)) != 0)
    return -1;
  for (uint32_t i = 0; i < n; ++i)
    atomic_init(&lab[i], i);
  atomic_int flag;
  atomic_init(&flag, 1);
  _Atomic uint64_t cursor;
  atomic_init(&cursor, 0);
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, NULL, (unsigned)num_threads);
  pthread_t *th = (pthread_t *)malloc((size_t)num_threads * sizeof(pthread_t));
  LpWorker *w = (LpWorker *)malloc((size_t)num_threads * sizeof(LpWorker));
  if (!th || !w)
  {
    free(th);
    free(w);
    free(lab);
    pthread_barrier_destroy(&bar);
    return -1;
  }
  for (int t = 0; t < num_threads; ++t)
  {
    uint64_t lo = (uint64_t)t * block, hi = lo + block;
    if (lo > n)
      lo = n;
    if (hi > n)
      hi = n;
    w[t] = (LpWorker){row_ptr, col_idx, lab, &flag, &cursor, &bar, n, chunk,
                      (uint32_t)lo, (uint32_t)hi, t, 0};
  }
  for (int t = 1; t < num_threads; ++t)
    pthread_create(&th[t], NULL, lp_thread, &w[t]);
  lp_thread(&w[0]);
  for (int t = 1; t < num_threads; ++t)
    pthread_join(th[t], NULL);
  for (uint32_t i = 0; i < n; ++i)
    labels[i] = atomic_load_explicit(&lab[i], memory_order_relaxed);
  pthread_barrier_destroy(&bar);
  free(th);
  free(w);
  free(lab);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Boundary records: sort, dedup, ghost extraction                     */
/* ------------------------------------------------------------------ */

/*
 * Ascending sort of 64-bit keys: the total order src/boundary_sort.c:93-98
 * puts on BoundaryEdge records read as (remote << 32) | rep.  LSD radix,
 * 8-bit digits, skipping digits that are constant over the input.
 * Returns -1 on allocation failure.
 */
int oracle_sort_u64(uint64_t *keys, uint64_t count)
{
  if (!keys || count <= 1)
    return 0;
  uint64_t *tmp = (uint64_t *)malloc((size_t)count * sizeof(uint64_t));
  if (!tmp)
    return -1;
  uint64_t hist[8][256];
  memset(hist, 0, sizeof(hist));
  for (uint64_t i = 0; i < count; ++i)
    for (int d = 0; d < 8; ++d)
      hist[d][(keys[i] >> (8 * d)) & 0xFF]++;
  uint64_t *src = keys, *dst = tmp;
  for (int d = 0; d < 8; ++d)
  {
    int constant = 0;
    for (int b = 0; b < 256; ++b)
      if (hist[d][b] == count)
        constant = 1;
    if (constant)
      continue;
    uint64_t off[256], s = 0;
    for (int b = 0; b < 256; ++b)
    {
      off[b] = s;
      s += hist[d][b];
    }
    for (uint64_t i = 0; i < count; ++i)
      dst[off[(src[i] >> (8 * d)) & 0xFF]++] = src[i];
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  if (src != keys)
    memcpy(keys, src, (size_t)count * sizeof(uint64_t));
  free(tmp);
  return 0;
}

/* Adjacent-unique in place; returns the new count
 * (src/boundary_sort.c:100-119). */
uint64_t oracle_dedup_u64(uint64_t *keys, uint64_t count)
{
  if (!keys || count <= 1)
    return count;
  uint64_t w = 1;
  for (uint64_t i = 1; i < count; ++i)
    if (keys[i] != keys[w - 1])
      keys[w++] = keys[i];
  return w;
}

/*
 * Ghost extraction over sorted, deduplicated records (src/cc_mpi.c:431-457):
 * ghosts = the distinct `remote` values in order, and each record's remote
 * is replaced by its ghost index.  Records are (remote << 32) | rep.
 * Returns the ghost count.
 */
uint32_t oracle_ghost_extract(uint64_t *records, uint64_t count, uint32_t *ghost_vertices)
{
  uint32_t g = 0;
  uint32_t prev = 0xFFFFFFFFu;
  for (uint64_t i = 0; i < count; ++i)
  {
    uint32_t remote = (uint32_t)(records[i] >> 32);
    uint32_t rep = (uint32_t)records[i];
    if (i == 0 || remote != prev)
    {
      ghost_vertices[g++] = remote;
      prev = remote;
    }
    records[i] = ((uint64_t)(g - 1) << 32) | rep;
  }
  return g;
}

/* ------------------------------------------------------------------ */
/* Delta exchange pack (src/exchange_plan.c:242-357)                   */
/* ------------------------------------------------------------------ */

/*
 * For each destination segment k (entries [seg_start[k], seg_start[k+1])
 * of send_to_flat), in order: for each i in the segment,
 *   lbl = comp_label[comp_of[v - v_start]] if v in [v_start, v_end) else v,
 *   and if lbl < prev_sent[idx] set prev_sent[idx] = lbl and emit
 *   ((uint64_t)i << 32) | lbl   (i relative to the segment start).
 * sendcounts[k] = pairs emitted for segment k; sdispls[k] = exclusive sum.
 * Returns the total number of pairs written to sendbuf.
 */
uint64_t oracle_delta_pack(const uint32_t *send_to_flat, const int *seg_start, int nseg,
                           const uint32_t *comp_label, const uint32_t *comp_of,
                           uint32_t v_start, uint32_t v_end, uint32_t *prev_sent,
                           uint64_t *sendbuf, int *sendcounts, int *sdispls)
{
  uint64_t total = 0;
  for (int k = 0; k < nseg; ++k)
  {
    const int base = seg_start[k];
    const int cnt = seg_start[k + 1] - base;
    sdispls[k] = (int)total;
    int wrote = 0;
    for (int i = 0; i < cnt; ++i)
    {
      const int idx = base + i;
      const uint32_t v = send_to_flat[idx];
      uint32_t lbl = v;
      if (v >= v_start && v < v_end)
        lbl = comp_label[comp_of[v - v_start]];
      if (lbl < prev_sent[idx])
      {
        prev_sent[idx] = lbl;
        sendbuf[total + (uint64_t)wrote] = ((uint64_t)(uint32_t)i << 32) | lbl;
        ++wrote;
      }
    }
    sendcounts[k] = wrote;
    total += (uint64_t)wrote;
  }
  return total;
}

/* Unpack of src/exchange_plan.c:390-413: for each source segment k of the
 * received pairs, ghost_labels[need_from_gidx_flat[need_base[k] + pos]] = lbl. */
void oracle_delta_unpack(const uint64_t *recvbuf, const int *recvcounts, const int *rdispls,
                         const int *need_base, int nsrc, const uint32_t *need_from_gidx_flat,
                         uint32_t *ghost_labels)
{
  for (int k = 0; k < nsrc; ++k)
    for (int i = 0; i < recvcounts[k]; ++i)
    {
      uint64_t pr = recvbuf[rdispls[k] + i];
      uint32_t pos = (uint32_t)(pr >> 32), lbl = (uint32_t)pr;
      ghost_labels[need_from_gidx_flat[need_base[k] + (int)pos]] = lbl;
    }
}

/* ------------------------------------------------------------------ */
/* Merge round (src/cc_mpi.c:547-622, single-thread path)              */
/* ------------------------------------------------------------------ */

/* get_vertex_label_global (src/cc_mpi.c:78-115): the label of vertex x as
 * this rank sees it; *missing = 1 if x is neither owned nor a ghost. */
static uint32_t merge_label_of(uint32_t x, const uint32_t *comp_label, const uint32_t *comp_of,
                               uint32_t v_start, uint32_t v_end, const uint32_t *ghost_vertices,
                               uint32_t ghost_count, const uint32_t *ghost_labels, int *missing)
{
  if (x >= v_start && x < v_end)
  {
    *missing = 0;
    return comp_label[comp_of[x - v_start]];
  }
  uint32_t lo = 0, hi = ghost_count;
  while (lo < hi)
  {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (ghost_vertices[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < ghost_count && ghost_vertices[lo] == x)
  {
    *missing = 0;
    return ghost_labels[lo];
  }
  *missing = 1;
  return x;
}

/*
 * One merge iteration's init, hook and update phases (src/cc_mpi.c:547-622,
 * the nthreads == 1 branches) over records (gidx << 32) | rep (BoundaryEdge
 * {rep, remote} after ghost extraction) and the rep list: comp_min[r] =
 * comp_label[r]; comp_min[rep] lowered by every record's ghost label; then,
 * rep by rep in list order, up to 4 label hops from comp_min[r], and
 * comp_label[r] lowered to the result.  Returns 1 if any comp_label dropped.
 */
int oracle_merge_step(const uint64_t *records, uint64_t count, const uint32_t *ghost_vertices,
                      const uint32_t *ghost_labels, uint32_t ghost_count, const uint32_t *comp_of,
                      uint32_t *comp_label, uint32_t *comp_min, const uint32_t *rep_list,
                      uint32_t rep_count, uint32_t v_start, uint32_t v_end)
{
  for (uint32_t i = 0; i < rep_count; ++i)
    comp_min[rep_list[i]] = comp_label[rep_list[i]];
  for (uint64_t e = 0; e < count; ++e)
  {
    const uint32_t rep = (uint32_t)records[e];
    const uint32_t lbl = ghost_labels[(uint32_t)(records[e] >> 32)];
    if (lbl < comp_min[rep])
      comp_min[rep] = lbl;
  }
  int any = 0;
  for (uint32_t i = 0; i < rep_count; ++i)
  {
    const uint32_t r = rep_list[i];
    uint32_t x = comp_min[r];
    for (int t = 0; t < 4; ++t)
    {
      int missing = 0;
      const uint32_t y = merge_label_of(x, comp_label, comp_of, v_start, v_end, ghost_vertices,
                                        ghost_count, ghost_labels, &missing);
      if (missing || y >= x)
        break;
      x = y;
    }
    if (x < comp_min[r])
      comp_min[r] = x;
    if (comp_min[r] < comp_label[r])
    {
      comp_label[r] = comp_min[r];
      any = 1;
    }
  }
  return any;
}

/* ------------------------------------------------------------------ */
/* Partition math (src/graph_dist.c:22-92, src/cc_mpi.c:41-76)          */
/* ------------------------------------------------------------------ */

/* Block partition: the first n % P ranks get one extra vertex
 * (src/graph_dist.c:22-51). */
void oracle_vertex_range(uint32_t n, int P, int rank, uint32_t *v_start, uint32_t *v_end)
{
  if (P <= 0)
  {
    *v_start = *v_end = 0;
    return;
  }
  uint32_t base = n / (uint32_t)P, rem = n % (uint32_t)P, r = (uint32_t)rank;
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

/* Edge-balanced bounds with vertex weight w (src/graph_dist.c:53-92):
 * bounds[r] is the first vertex at which the cumulative weight deg+w
 * reaches r/P of the total, clamped so every rank keeps >= 1 vertex. */
void oracle_weighted_bounds(const uint64_t *row_ptr, uint32_t n, int P, uint32_t w,
                            uint32_t *bounds)
{
  const uint64_t total = row_ptr[n] + (uint64_t)w * n;
  bounds[0] = 0;
  uint32_t v = 0;
  uint64_t cum = 0;
  for (int r = 1; r < P; ++r)
  {
    uint64_t target = total * (uint64_t)r / (uint64_t)P;
    while (v < n && cum < target)
    {
      cum += row_ptr[v + 1] - row_ptr[v] + w;
      ++v;
    }
    uint32_t lo = bounds[r - 1] + 1, hi = n - (uint32_t)(P - r);
    if (v < lo)
      v = lo;
    if (v > hi)
      v = hi;
    bounds[r] = v;
  }
  bounds[P] = n;
}

/* Owner rank of v (src/cc_mpi.c:41-76): binary search of bounds when an
 * edge-balanced partition is in use, else the block formula. */
int oracle_owner_of_vertex(uint32_t v, uint32_t n, int P, const uint32_t *bounds, int part_kind)
{
  if (part_kind != 0 && bounds && P > 1)
  {
    int lo = 0, hi = P - 1;
    while (lo <= hi)
    {
      int mid = lo + ((hi - lo) >> 1);
      if (v < bounds[mid])
        hi = mid - 1;
      else if (v >= bounds[mid + 1])
        lo = mid + 1;
      else
        return mid;
    }
    return v >= bounds[P] ? P - 1 : 0;
  }
  if (P <= 1)
    return 0;
  uint32_t base = n / (uint32_t)P, rem = n % (uint32_t)P;
  uint64_t split = (uint64_t)(base + 1) * rem;
  if ((uint64_t)v < split)
    return (int)(v / (base + 1));
  if (base == 0)
    return (int)rem;
  return (int)(rem + (uint32_t)(((uint64_t)v - split) / base));
}

/* ------------------------------------------------------------------ */
/* Per-rank preprocessing (src/cc_mpi.c:252-304, 364-394)              */
/* ------------------------------------------------------------------ */

/*
 * Local-local split: keep the entries of the owned rows whose column lies in
 * [v_start, v_end), rebased to local ids.  lrow has n_local+1 entries; lcol
 * must hold row_ptr[n_local] entries.  Returns m_local_local and stores the
 * raw boundary-entry count (columns outside the range but < n_global).
 */
uint64_t oracle_local_split(uint32_t n_local, const uint64_t *row_ptr, const uint32_t *col_idx,
                            uint32_t v_start, uint32_t v_end, uint32_t n_global, uint64_t *lrow,
                            uint32_t *lcol, uint64_t *boundary_raw)
{
  uint64_t w = 0, braw = 0;
  lrow[0] = 0;
  for (uint32_t li = 0; li < n_local; ++li)
  {
    for (uint64_t j = row_ptr[li]; j < row_ptr[li + 1]; ++j)
    {
      uint32_t v = col_idx[j];
      if (v >= v_start && v < v_end)
        lcol[w++] = v - v_start;
      else if (v < n_global)
        ++braw;
    }
    lrow[li + 1] = w;
  }
  *boundary_raw = braw;
  return w;
}

/* Boundary records {rep = comp_of[li], remote = v} for every owned entry
 * whose column is outside [v_start, v_end) and < n_global, in row order,
 * written as (remote << 32) | rep.  Returns the count. */
uint64_t oracle_boundary_build(uint32_t n_local, const uint64_t *row_ptr, const uint32_t *col_idx,
                               uint32_t v_start, uint32_t v_end, uint32_t n_global,
                               const uint32_t *comp_of, uint64_t *records)
{
  uint64_t w = 0;
  for (uint32_t li = 0; li < n_local; ++li)
    for (uint64_t j = row_ptr[li]; j < row_ptr[li + 1]; ++j)
    {
      uint32_t v = col_idx[j];
      if ((v >= v_start && v < v_end) || v >= n_global)
        continue;
      records[w++] = ((uint64_t)v << 32) | comp_of[li];
    }
  return w;
}
