/*
 * ref_knn_driver.c - runs the reference's own candidate-list builder
 * aco_mpi_shared_init (src/openmp-mpi/aco_parallel.c:218-250) on a .vrp file
 * and prints the candidate indices, one row per node.  TEST INFRASTRUCTURE:
 * built only by oracle/ref.mk from the reference sources where they lie
 * (the translation unit below includes aco_parallel.c so its static functions
 * are callable); its output pins the oracle's reference-key kNN.
 *
 * usage: ref_knn <instance.vrp> <k>
 */
#include "aco_parallel.c"

#include "instance_parser.h"

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s <instance.vrp> <k>\n", argv[0]);
    return 1;
  }
  int n = 0;
  double **c = NULL;
  VrpInstanceMeta meta = {0};
  if (vrp_load_tsplib_euc2d_matrix_ex(argv[1], &n, &c, &meta) != 0) return 1;
  int k = atoi(argv[2]);
  /* the float matrix exactly as aco_vrp_run builds it (aco_parallel.c:511-514) */
  MatrixFloat *cm = matrix_create_float(n);
  for (int i = 0; i <= n; i++)
    for (int j = 0; j <= n; j++) cm->rows[i][j] = (float)c[i][j];
  AcoMpiRankShared s;
  if (!aco_mpi_shared_init(&s, n, k, cm, 2.0)) return 1;
  printf("%d %d\n", n, k);
  for (int i = 0; i <= n; i++) {
    for (int t = 0; t < k; t++) printf("%s%d", t ? " " : "", s.cand_idx[(size_t)i * s.stride + t]);
    printf("\n");
  }
  aco_mpi_shared_free(&s);
  matrix_free_float(cm);
  matrix_free(c);
  return 0;
}
