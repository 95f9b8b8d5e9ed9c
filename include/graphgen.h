#ifndef GRAPHGEN_H
#define GRAPHGEN_H

/*
 * Seeded synthetic inputs (libcchost.so).  The reference reads its graphs
 * from .mat files (src/graph.c:38-128); these generators produce the same
 * kind of graph — undirected, symmetric CSR, no self-loops, no duplicates,
 * sorted rows — from a seed, so CPU and GPU runs see identical inputs.
 * Each returns 0 on success; the CSRGraph arrays are malloc'd and released
 * with free_csr().
 */

#include <stdint.h>
#include "graph.h"

#ifdef __cplusplus
extern "C" {
#endif

uint64_t ccgen_rand_at(uint64_t seed, uint64_t i);

int ccgen_csr_from_pairs(uint32_t n, const uint32_t *src, const uint32_t *dst, uint64_t count,
                         const uint32_t *perm, CSRGraph *out);
int ccgen_erdos_renyi(uint32_t n, uint64_t pairs, uint64_t seed, CSRGraph *out);
int ccgen_rmat(int scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
               int permute, CSRGraph *out);
int ccgen_grid(uint32_t rows, uint32_t cols, double keep, uint64_t seed, int permute,
               CSRGraph *out);
int ccgen_chung_lu(uint32_t n, uint64_t pairs, double gamma, uint64_t seed, int permute,
                   CSRGraph *out);

/* The vertex renaming the generators apply when `permute` is set. */
int ccgen_permutation(uint32_t n, uint64_t seed, uint32_t *out);
/* Chung-Lu inverse-CDF table (n doubles); returns the weight total. */
double ccgen_chung_lu_cdf(uint32_t n, double gamma, double *cdf);

#ifdef __cplusplus
}
#endif

#endif /* GRAPHGEN_H */
