/*
 * aco_vrp_hip.out - command line of the MI355X solver.
 *
 * Same surface as the reference CUDA CLI (src/cuda/main_vrp.cu:43-124):
 *   aco_vrp_hip.out <instance.vrp> <K> <m> [seed]
 * with the ACO_SOLVER_* environment knobs, the VEHICLES mismatch check, the
 * final validation, and the "Route i: ...", "Cost: %.3f", "best cost: %.3f"
 * output.  alpha=1, beta=2, rho=0.5, tau0=1, Q=1 as in the reference
 * (main_vrp.cu:53-57).  Like the reference it optimises and prints the
 * unrounded route length (instance_parser.c:17-19); ACO_HIP_DIST=nint
 * optimises and prints the integer TSPLIB (nint) length instead.  Weighted
 * DEMAND_SECTION is accepted and capacity-checked.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "aco_hip.h"
#include "aco_host.h"

static void usage(const char *argv0) {
  fprintf(stderr, "usage: %s <instance.vrp> <K> <m> [seed]\n", argv0);
  fprintf(stderr, "example: %s instance.vrp 8 0 1234\n", argv0);
}

int main(int argc, char **argv) {
  int K = 0, m = 0, ok = 1;
  unsigned int seed = 1234u;
  if (argc < 4 || argc > 5) {
    usage(argv[0]);
    return 1;
  }
  ok = ok && cli_parse_int_arg(argv[2], &K);
  ok = ok && cli_parse_int_arg(argv[3], &m);
  if (argc == 5) seed = cli_parse_uint_arg(argv[4], &ok);
  if (!ok || K <= 0 || m < 0) {
    usage(argv[0]);
    return 1;
  }
  VrpInstance inst;
  if (vrp_read(argv[1], &inst, 0) != 0) return 1;
  if (inst.K > 0 && inst.K != K) {
    fprintf(stderr, "instance VEHICLES mismatch: CLI K=%d, file VEHICLES=%d\n", K, inst.K);
    vrp_free(&inst);
    return 1;
  }
  const int n = inst.n;
  int32_t *tour = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + K + 1 + 64));
  Solution *best = solution_create(K, n);
  if (!tour || !best) {
    fprintf(stderr, "failed to allocate solution\n");
    free(tour);
    solution_free(best);
    vrp_free(&inst);
    return 1;
  }
  const char *dist = getenv("ACO_HIP_DIST");
  const int real = !(dist && strcmp(dist, "nint") == 0);
  int32_t len = 0;
  int64_t cost = 0;
  int status = aco_vrp_hip_with_capacity(n, K, inst.capacity, m, inst.x, inst.y, inst.demand, 1.0,
                                         2.0, 0.5, 1.0, 1.0, seed,
                                         real ? ACO_HIP_DIST_REAL : ACO_HIP_DIST_NINT, tour, &len,
                                         &cost);
  int rc = 0;
  if (status != ACO_HIP_OK) {
    fprintf(stderr, "HIP solver failed: %s\n", aco_status_string(status));
    rc = 1;
  } else if (!solution_from_tour(best, tour, len)) {
    fprintf(stderr, "invalid solution: tour does not hold %d routes\n", K);
    rc = 1;
  } else {
    /* the cost of the metric optimised: the exact sum of the float edge
     * lengths (real) or the integer length (nint) */
    const double printed = real ? (double)cost * ACO_HIP_REAL_COST_UNIT : (double)cost;
    const int64_t check = real ? solution_cost_real_units(best, &inst)
                               : solution_cost_nint(best, &inst);
    char err[160];
    if (!cli_validate_solution_or_report(best, n, K, printed)) {
      rc = 1;
    } else if (!solution_validate_capacity(best, &inst, err, sizeof(err))) {
      fprintf(stderr, "invalid solution: %s\n", err);
      rc = 1;
    } else if (check != cost) {
      fprintf(stderr, "invalid solution: cost mismatch\n");
      rc = 1;
    } else {
      cli_print_solution_routes(best, K);
      cli_print_solution_cost(printed);
    }
  }
  solution_free(best);
  free(tour);
  vrp_free(&inst);
  return rc;
}
