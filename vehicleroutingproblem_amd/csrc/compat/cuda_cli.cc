/*
 * aco_vrp_cuda_compat.out - the reference CUDA CLI's flow
 * (src/cuda/main_vrp.cu:43-124) over the drop-in aco_vrp_cuda_with_capacity:
 *   aco_vrp_cuda_compat.out <instance.vrp> <K> <m> [seed]
 * It declares the entry point itself, exactly as main_vrp.cu:29-35 does (C++
 * linkage, float coordinates, customer capacity, Solution and double cost),
 * so building it proves that a caller written against the reference's
 * prototype links against libaco_hip.so unchanged.  The instance is read with
 * unit demand required, like the reference parser (instance_parser.c:357-366).
 */
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "aco_host.h"

typedef enum {
  ACO_OK = 0,
  ACO_ERR_INVALID_INPUT = 1,
  ACO_ERR_ALLOCATION = 2,
  ACO_ERR_NO_SOLUTION = 3,
  ACO_ERR_BACKEND = 4
} AcoStatus;

AcoStatus aco_vrp_cuda_with_capacity(int n, int K, int vehicle_capacity_customers, int m,
                                     float *coords_x, float *coords_y, double alpha, double beta,
                                     double rho, double tau0, double Q, unsigned int seed,
                                     Solution *best_solution, double *best_cost);

int main(int argc, char **argv) {
  int K = 0, m = 0, ok = 1;
  unsigned int seed = 1234u;
  if (argc < 4 || argc > 5) {
    fprintf(stderr, "usage: %s <instance.vrp> <K> <m> [seed]\n", argv[0]);
    return 1;
  }
  ok = ok && cli_parse_int_arg(argv[2], &K);
  ok = ok && cli_parse_int_arg(argv[3], &m);
  if (argc == 5) seed = cli_parse_uint_arg(argv[4], &ok);
  if (!ok || K <= 0 || m < 0) {
    fprintf(stderr, "usage: %s <instance.vrp> <K> <m> [seed]\n", argv[0]);
    return 1;
  }
  VrpInstance inst;
  if (vrp_read(argv[1], &inst, 1) != 0) return 1;
  if (inst.K > 0 && inst.K != K) {
    fprintf(stderr, "instance VEHICLES mismatch: CLI K=%d, file VEHICLES=%d\n", K, inst.K);
    vrp_free(&inst);
    return 1;
  }
  const int n = inst.n;
  std::vector<float> x((size_t)n + 1), y((size_t)n + 1);
  for (int i = 0; i <= n; ++i) {
    x[(size_t)i] = (float)inst.x[i];
    y[(size_t)i] = (float)inst.y[i];
  }
  Solution *best = solution_create(K, n);
  if (!best) {
    fprintf(stderr, "failed to allocate solution\n");
    vrp_free(&inst);
    return 1;
  }
  double best_cost = 0.0;
  const AcoStatus st = aco_vrp_cuda_with_capacity(n, K, inst.capacity, m, x.data(), y.data(), 1.0,
                                                  2.0, 0.5, 1.0, 1.0, seed, best, &best_cost);
  int rc = 0;
  if (st != ACO_OK) {
    fprintf(stderr, "solver failed: %s\n", aco_status_string(st));
    rc = 1;
  } else if (!cli_validate_solution_or_report(best, n, K, best_cost)) {
    rc = 1;
  } else {
    cli_print_solution_routes(best, K);
    cli_print_solution_cost(best_cost);
  }
  solution_free(best);
  vrp_free(&inst);
  return rc;
}
