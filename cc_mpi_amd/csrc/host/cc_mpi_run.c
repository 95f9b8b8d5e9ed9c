/*
 * cc_mpi_run: MPI driver program for the distributed CC path.
 *
 * Rank 0 generates a seeded graph (libcchost) or reads a level 5 MAT file,
 * distributes it with the reference's partitioning (dist_csr_from_global /
 * load_dist_csr_from_file, as src/mpi_cc_benchmark.c:178), every rank calls
 * compute_connected_components_mpi_advanced, and rank 0 writes the gathered
 * labels (uint32 per vertex, little-endian) to --out.  Plays the role of the
 * reference's bench entry (src/mpi_cc_benchmark.c:31-251) for generated
 * inputs; -e/--exchange and -c/--chunk-size keep their meaning.
 *
 *   cc_mpi_run --gen rmat  --scale S [--ef F] [--seed X] [options]
 *   cc_mpi_run --gen er    --n N --deg D [--seed X] [options]
 *   cc_mpi_run --gen grid  --rows R --cols C [--keep P] [--seed X] [options]
 *   cc_mpi_run --input GRAPH.mat [options]
 *   options: -e E (exchange interval, default 1), -c C (chunk, default 2048),
 *            --vertex-blocks (default: edge-balanced, as the reference bench),
 *            --out FILE, -r RUNS, --write-mat FILE (save the generated graph)
 */
#define _POSIX_C_SOURCE 200112L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mpi.h>

#include "../../../include/cc_mpi.h"
#include "../../../include/graphgen.h"

static const char *arg_str(int argc, char **argv, const char *name, const char *def)
{
  for (int i = 1; i + 1 < argc; ++i)
    if (strcmp(argv[i], name) == 0)
      return argv[i + 1];
  return def;
}

static int has_flag(int argc, char **argv, const char *name)
{
  for (int i = 1; i < argc; ++i)
    if (strcmp(argv[i], name) == 0)
      return 1;
  return 0;
}

/* Rank 0 generates the graph and distributes it; 0 on success on every rank. */
static int generate(int argc, char **argv, const char *gen, uint64_t seed, int use_edge,
                    const char *write_mat, DistCSRGraph *Gd, MPI_Comm comm)
{
  int rank, size;
  MPI_Comm_rank(comm, &rank);
  MPI_Comm_size(comm, &size);
  CSRGraph full;
  memset(&full, 0, sizeof(full));
  int rc = 0;
  if (rank == 0)
  {
    if (strcmp(gen, "rmat") == 0)
      rc = ccgen_rmat(atoi(arg_str(argc, argv, "--scale", "16")),
                      (uint32_t)atoi(arg_str(argc, argv, "--ef", "16")), 0.57, 0.19, 0.19, seed, 1,
                      &full);
    else if (strcmp(gen, "er") == 0)
      rc = ccgen_erdos_renyi((uint32_t)atol(arg_str(argc, argv, "--n", "65536")),
                             (uint64_t)(atof(arg_str(argc, argv, "--deg", "8")) *
                                        atof(arg_str(argc, argv, "--n", "65536")) / 2.0),
                             seed, &full);
    else if (strcmp(gen, "grid") == 0)
      rc = ccgen_grid((uint32_t)atol(arg_str(argc, argv, "--rows", "256")),
                      (uint32_t)atol(arg_str(argc, argv, "--cols", "256")),
                      atof(arg_str(argc, argv, "--keep", "0.7")), seed, 1, &full);
    else
      rc = 2;
    if (rc)
      fprintf(stderr, "cc_mpi_run: graph generation failed (%d)\n", rc);
    else
      printf("graph %s: n=%u m=%llu ranks=%d\n", gen, full.n, (unsigned long long)full.m, size);
  }
  if (rank == 0 && rc == 0 && write_mat)
  {
    rc = cc_mat_write_sparse(write_mat, gen, &full, 0);
    if (rc)
      fprintf(stderr, "cc_mpi_run: cannot write %s (%d)\n", write_mat, rc);
  }
  MPI_Bcast(&rc, 1, MPI_INT, 0, comm);
  if (rc)
  {
    if (rank == 0)
      free_csr(&full);
    return 1;
  }
  if (dist_csr_from_global(&full, use_edge, Gd, comm) != 0)
  {
    fprintf(stderr, "cc_mpi_run: distribution failed\n");
    MPI_Abort(comm, EXIT_FAILURE);
  }
  if (rank == 0)
    free_csr(&full);
  return 0;
}

int main(int argc, char **argv)
{
  MPI_Init(&argc, &argv);
  int rank, size;
  MPI_Comm comm = MPI_COMM_WORLD;
  MPI_Comm_rank(comm, &rank);
  MPI_Comm_size(comm, &size);
  const char *gen = arg_str(argc, argv, "--gen", "rmat");
  const uint64_t seed = strtoull(arg_str(argc, argv, "--seed", "1002"), NULL, 10);
  const int exch = atoi(arg_str(argc, argv, "-e", "1"));
  const int chunk = atoi(arg_str(argc, argv, "-c", "2048"));
  const int runs = atoi(arg_str(argc, argv, "-r", "1"));
  const char *out = arg_str(argc, argv, "--out", NULL);
  const int use_edge = !has_flag(argc, argv, "--vertex-blocks");

  const char *input = arg_str(argc, argv, "--input", NULL);
  const char *write_mat = arg_str(argc, argv, "--write-mat", NULL);

  DistCSRGraph Gd;
  if (input)
  {
    const int lrc = load_dist_csr_from_file(input, 1, 1, use_edge, &Gd, comm);
    if (lrc != 0)
    {
      if (rank == 0)
        fprintf(stderr, "cc_mpi_run: cannot load %s (%d)\n", input, lrc);
      MPI_Finalize();
      return 1;
    }
    if (rank == 0)
      printf("graph %s: n=%u m=%llu ranks=%d\n", input, Gd.n_global,
             (unsigned long long)Gd.m_global, size);
  }
  else if (generate(argc, argv, gen, seed, use_edge, write_mat, &Gd, comm) != 0)
  {
    MPI_Finalize();
    return 1;
  }
  uint32_t *labels = (uint32_t *)malloc((size_t)(Gd.n_global ? Gd.n_global : 1) * 4);
  if (!labels)
    MPI_Abort(comm, EXIT_FAILURE);
  for (int r = 0; r < (runs > 0 ? runs : 1); ++r)
  {
    MPI_Barrier(comm);
    const double t0 = MPI_Wtime();
    compute_connected_components_mpi_advanced(&Gd, labels, chunk, exch, comm);
    const double t1 = MPI_Wtime();
    if (rank == 0)
      printf("run %d: %.6f s, components %llu (labels: %u distinct), merge rounds %d\n", r, t1 - t0,
             (unsigned long long)cc_mpi_last_component_count(),
             count_connected_components(labels, Gd.n_global), cc_mpi_last_merge_rounds());
  }
  if (rank == 0 && out)
  {
    FILE *f = fopen(out, "wb");
    if (!f || fwrite(labels, 4, Gd.n_global, f) != Gd.n_global)
    {
      fprintf(stderr, "cc_mpi_run: cannot write %s\n", out);
      MPI_Abort(comm, EXIT_FAILURE);
    }
    fclose(f);
  }
  free(labels);
  free_dist_csr(&Gd);
  MPI_Finalize();
  return 0;
}
