/*
 * Seeded synthetic graph generators and the CSR builder.
 *
 * Every random draw is a pure function of (seed, counter): the i-th output
 * of a SplitMix64 stream (Steele, Lea & Flood, "Fast splittable
 * pseudorandom number generators", OOPSLA 2014), so a graph depends only on
 * its parameters and seed, never on the thread count.
 *
 * Output is an undirected simple graph stored as a symmetric CSR with
 * sorted rows: self-loops dropped, both directions present, duplicates
 * removed (the input contract of the reference's loader, src/graph.c:38-128,
 * and of its local CSR, src/cc_mpi.c:252-304).  Vertex ids are uint32,
 * offsets uint64.
 */
#define _GNU_SOURCE
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/graph.h"
#include "../../../include/graphgen.h"

static inline uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* i-th output of the SplitMix64 stream seeded with `seed`. */
uint64_t ccgen_rand_at(uint64_t seed, uint64_t i)
{
  return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

static inline uint32_t below(uint64_t r, uint32_t n)
{
  return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

void free_csr(CSRGraph *g)
{
  if (!g)
    return;
  free(g->row_ptr);
  free(g->col_idx);
  memset(g, 0, sizeof(*g));
}

/* Random permutation of [0, n) by Fisher-Yates driven by the seed. */
static uint32_t *make_permutation(uint32_t n, uint64_t seed)
{
  uint32_t *perm = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
  if (!perm)
    return NULL;
  for (uint32_t i = 0; i < n; ++i)
    perm[i] = i;
  for (uint32_t i = n; i > 1; --i)
  {
    uint32_t j = below(ccgen_rand_at(seed, i), i);
    uint32_t t = perm[i - 1];
    perm[i - 1] = perm[j];
    perm[j] = t;
  }
  return perm;
}

static int cmp_u32(const void *a, const void *b)
{
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

static void sort_small_u32(uint32_t *a, uint64_t n)
{
  if (n <= 32)
  {
    for (uint64_t i = 1; i < n; ++i)
    {
      uint32_t x = a[i];
      uint64_t j = i;
      while (j > 0 && a[j - 1] > x)
      {
        a[j] = a[j - 1];
        --j;
      }
      a[j] = x;
    }
  }
  else
  {
    qsort(a, (size_t)n, sizeof(uint32_t), cmp_u32);
  }
}

/*
 * Build a symmetric simple CSR over n vertices from `count` undirected edge
 * pairs (src[i], dst[i]).  Pairs with an endpoint >= n and self-loops are
 * dropped; both directions are stored; duplicates are removed; rows are
 * sorted.  If perm is non-NULL, vertex x is renamed perm[x] first.
 * Returns 0, or nonzero on allocation failure.
 */
int ccgen_csr_from_pairs(uint32_t n, const uint32_t *src, const uint32_t *dst, uint64_t count,
                         const uint32_t *perm, CSRGraph *out)
{
  memset(out, 0, sizeof(*out));
  uint64_t *deg = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  if (!deg)
    return 1;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < count; ++i)
  {
    uint32_t u = src[i], v = dst[i];
    if (u >= n || v >= n || u == v)
      continue;
    if (perm)
    {
      u = perm[u];
      v = perm[v];
    }
    __atomic_fetch_add(&deg[u + 1], 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&deg[v + 1], 1, __ATOMIC_RELAXED);
  }
  for (uint32_t i = 0; i < n; ++i)
    deg[i + 1] += deg[i];
  const uint64_t m_raw = deg[n];
  uint32_t *col = (uint32_t *)malloc((size_t)(m_raw ? m_raw : 1) * sizeof(uint32_t));
  uint64_t *cursor = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  if (!col || !cursor)
  {
    free(deg);
    free(col);
    free(cursor);
    return 1;
  }
  memcpy(cursor, deg, ((size_t)n + 1) * sizeof(uint64_t));
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < count; ++i)
  {
    uint32_t u = src[i], v = dst[i];
    if (u >= n || v >= n || u == v)
      continue;
    if (perm)
    {
      u = perm[u];
      v = perm[v];
    }
    col[__atomic_fetch_add(&cursor[u], 1, __ATOMIC_RELAXED)] = v;
    col[__atomic_fetch_add(&cursor[v], 1, __ATOMIC_RELAXED)] = u;
  }
  free(cursor);
  /* Sort + unique each row; the unique count goes to ucnt. */
  uint64_t *ucnt = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  if (!ucnt)
  {
    free(deg);
    free(col);
    return 1;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint32_t u = 0; u < n; ++u)
  {
    uint32_t *row = col + deg[u];
    uint64_t len = deg[u + 1] - deg[u];
    sort_small_u32(row, len);
    uint64_t w = 0;
    for (uint64_t j = 0; j < len; ++j)
      if (w == 0 || row[j] != row[w - 1])
        row[w++] = row[j];
    ucnt[u + 1] = w;
  }
  ucnt[0] = 0;
  for (uint32_t i = 0; i < n; ++i)
    ucnt[i + 1] += ucnt[i];
  const uint64_t m = ucnt[n];
  uint32_t *col2 = (uint32_t *)malloc((size_t)(m ? m : 1) * sizeof(uint32_t));
  if (!col2)
  {
    free(deg);
    free(col);
    free(ucnt);
    return 1;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint32_t u = 0; u < n; ++u)
    memcpy(col2 + ucnt[u], col + deg[u], (size_t)(ucnt[u + 1] - ucnt[u]) * sizeof(uint32_t));
  free(col);
  free(deg);
  out->n = n;
  out->m = m;
  out->row_ptr = ucnt;
  out->col_idx = col2;
  return 0;
}

/* The vertex renaming the generators apply when `permute` is set. */
int ccgen_permutation(uint32_t n, uint64_t seed, uint32_t *out)
{
  uint32_t *perm = make_permutation(n, seed ^ 0x5EED5EED5EED5EEDull);
  if (!perm)
    return 1;
  memcpy(out, perm, (size_t)n * sizeof(uint32_t));
  free(perm);
  return 0;
}

/* Chung-Lu weight prefix sums (the inverse-CDF table ccgen_chung_lu
 * searches); returns the total. */
double ccgen_chung_lu_cdf(uint32_t n, double gamma, double *cdf)
{
  const double expo = -1.0 / (gamma - 1.0);
  const double i0 = (double)(n >> 20) + 1.0;
  double acc = 0.0;
  for (uint32_t i = 0; i < n; ++i)
  {
    acc += pow((double)i + i0, expo);
    cdf[i] = acc;
  }
  return acc;
}

static int finish_pairs(uint32_t n, uint32_t *src, uint32_t *dst, uint64_t count, int permute,
                        uint64_t seed, CSRGraph *out)
{
  uint32_t *perm = NULL;
  if (permute)
  {
    perm = make_permutation(n, seed ^ 0x5EED5EED5EED5EEDull);
    if (!perm)
    {
      free(src);
      free(dst);
      return 1;
    }
  }
  int rc = ccgen_csr_from_pairs(n, src, dst, count, perm, out);
  free(perm);
  free(src);
  free(dst);
  return rc;
}

/*
 * Erdos-Renyi G(n, M): M uniformly random vertex pairs (self-loops and
 * repeats dropped by the builder).  Average degree ~ 2M/n.
 */
int ccgen_erdos_renyi(uint32_t n, uint64_t pairs, uint64_t seed, CSRGraph *out)
{
  memset(out, 0, sizeof(*out));
  uint32_t *src = (uint32_t *)malloc((size_t)(pairs ? pairs : 1) * sizeof(uint32_t));
  uint32_t *dst = (uint32_t *)malloc((size_t)(pairs ? pairs : 1) * sizeof(uint32_t));
  if (!src || !dst)
  {
    free(src);
    free(dst);
    return 1;
  }
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < pairs; ++i)
  {
    src[i] = below(ccgen_rand_at(seed, 2 * i), n);
    dst[i] = below(ccgen_rand_at(seed, 2 * i + 1), n);
  }
  return finish_pairs(n, src, dst, pairs, 0, seed, out);
}

/*
 * R-MAT (Chakrabarti, Zhan & Faloutsos, SDM 2004): 2^scale vertices,
 * edge_factor * 2^scale edges, each placed by `scale` quadrant choices with
 * probabilities (a, b, c, 1-a-b-c), 16-bit resolution per choice.  Vertex
 * ids are then permuted by a seeded random permutation if `permute`.
 */
int ccgen_rmat(int scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
               int permute, CSRGraph *out)
{
  memset(out, 0, sizeof(*out));
  if (scale < 1 || scale > 31)
    return 2;
  const uint32_t n = (uint32_t)1 << scale;
  const uint64_t edges = (uint64_t)edge_factor << scale;
  uint32_t *src = (uint32_t *)malloc((size_t)(edges ? edges : 1) * sizeof(uint32_t));
  uint32_t *dst = (uint32_t *)malloc((size_t)(edges ? edges : 1) * sizeof(uint32_t));
  if (!src || !dst)
  {
    free(src);
    free(dst);
    return 1;
  }
  const uint32_t ta = (uint32_t)(a * 65536.0);
  const uint32_t tab = (uint32_t)((a + b) * 65536.0);
  const uint32_t tabc = (uint32_t)((a + b + c) * 65536.0);
  const uint64_t draws = (uint64_t)(scale + 3) / 4; /* four 16-bit choices per draw */
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < edges; ++e)
  {
    uint32_t u = 0, v = 0;
    uint64_t r = 0;
    for (int lvl = 0; lvl < scale; ++lvl)
    {
      if ((lvl & 3) == 0)
        r = ccgen_rand_at(seed, e * draws + (uint64_t)(lvl >> 2));
      uint32_t x = (uint32_t)(r & 0xFFFF);
      r >>= 16;
      uint32_t q = (x < ta) ? 0 : (x < tab) ? 1 : (x < tabc) ? 2 : 3;
      u = (u << 1) | (q >> 1);
      v = (v << 1) | (q & 1);
    }
    src[e] = u;
    dst[e] = v;
  }
  return finish_pairs(n, src, dst, edges, permute, seed, out);
}

/*
 * 2-D grid of rows x cols vertices (id = r*cols + c) with 4-neighbour
 * edges, each kept independently with probability keep (16-bit
 * resolution).  Vertex ids are permuted if `permute`.
 */
int ccgen_grid(uint32_t rows, uint32_t cols, double keep, uint64_t seed, int permute,
               CSRGraph *out)
{
  memset(out, 0, sizeof(*out));
  const uint64_t n64 = (uint64_t)rows * cols;
  if (n64 == 0 || n64 > 0xFFFFFFFFull)
    return 2;
  const uint32_t n = (uint32_t)n64;
  const uint32_t thr = (uint32_t)(keep * 65536.0);
  const uint64_t slots = 2 * n64;
  uint32_t *src = (uint32_t *)malloc((size_t)slots * sizeof(uint32_t));
  uint32_t *dst = (uint32_t *)malloc((size_t)slots * sizeof(uint32_t));
  if (!src || !dst)
  {
    free(src);
    free(dst);
    return 1;
  }
#pragma omp parallel for schedule(static)
  for (uint64_t id = 0; id < n64; ++id)
  {
    uint32_t r = (uint32_t)(id / cols), c = (uint32_t)(id % cols);
    uint64_t rnd = ccgen_rand_at(seed, id);
    /* right edge: low 16 bits; down edge: next 16 bits.  A dropped or
     * absent edge is written as a self-loop, which the builder drops. */
    int right = (c + 1 < cols) && ((uint32_t)(rnd & 0xFFFF) < thr);
    int down = (r + 1 < rows) && ((uint32_t)((rnd >> 16) & 0xFFFF) < thr);
    src[2 * id] = (uint32_t)id;
    dst[2 * id] = right ? (uint32_t)(id + 1) : (uint32_t)id;
    src[2 * id + 1] = (uint32_t)id;
    dst[2 * id + 1] = down ? (uint32_t)(id + cols) : (uint32_t)id;
  }
  return finish_pairs(n, src, dst, slots, permute, seed, out);
}

/*
 * Chung-Lu power-law graph: expected degree of vertex i proportional to
 * (i + i0)^(-1/(gamma-1)); `pairs` endpoint pairs are drawn independently
 * with probability proportional to the weights (inverse-CDF search), then
 * ids are permuted if `permute`.  i0 = n / 2^20 + 1 softens the head.
 */
int ccgen_chung_lu(uint32_t n, uint64_t pairs, double gamma, uint64_t seed, int permute,
                   CSRGraph *out)
{
  memset(out, 0, sizeof(*out));
  if (n == 0 || gamma <= 1.0)
    return 2;
  double *cdf = (double *)malloc((size_t)n * sizeof(double));
  uint32_t *src = (uint32_t *)malloc((size_t)(pairs ? pairs : 1) * sizeof(uint32_t));
  uint32_t *dst = (uint32_t *)malloc((size_t)(pairs ? pairs : 1) * sizeof(uint32_t));
  if (!cdf || !src || !dst)
  {
    free(cdf);
    free(src);
    free(dst);
    return 1;
  }
  const double acc = ccgen_chung_lu_cdf(n, gamma, cdf);
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < 2 * pairs; ++e)
  {
    double x = (double)(ccgen_rand_at(seed, e) >> 11) * (1.0 / 9007199254740992.0) * acc;
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi)
    {
      uint32_t mid = lo + (hi - lo) / 2;
      if (cdf[mid] <= x)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (e & 1)
      dst[e >> 1] = lo;
    else
      src[e >> 1] = lo;
  }
  free(cdf);
  return finish_pairs(n, src, dst, pairs, permute, seed, out);
}
