/*
 * aco_vrp_cuda.cc - aco_vrp_cuda_with_capacity with the reference's exact
 * prototype and C++ linkage (src/cuda/main_vrp.cu:29-35), over the one-call
 * solver of the C ABI in real-distance mode.  Host code only (g++).
 */
#include "aco_cuda_compat.h"

#include <float.h>
#include <math.h>
#include <stdio.h>

#include <vector>

#include "aco_hip.h"

AcoStatus aco_vrp_cuda_with_capacity(int n, int K, int vehicle_capacity_customers, int m,
                                     float *coords_x, float *coords_y, double alpha, double beta,
                                     double rho, double tau0, double Q, unsigned int seed,
                                     Solution *best_solution, double *best_cost) {
  /* the argument checks of aco_cuda.cu:193-196 */
  if (n <= 0 || K <= 0 || !coords_x || !coords_y || !best_solution || !best_cost)
    return ACO_ERR_INVALID_INPUT;
  if (m < 0 || vehicle_capacity_customers <= 0 || best_solution->K != K ||
      best_solution->route_cap < n + 2) {
    fprintf(stderr, "aco_vrp_cuda_with_capacity: need m >= 0, a capacity >= 1 and a solution "
                    "made by solution_create(K, n)\n");
    return ACO_ERR_INVALID_INPUT;
  }
  std::vector<int32_t> x((size_t)n + 1), y((size_t)n + 1), demand((size_t)n + 1, 1);
  demand[0] = 0;
  for (int i = 0; i <= n; ++i) {
    const float fx = coords_x[i], fy = coords_y[i];
    if (!(fx >= 0.0f && fx <= 32767.0f && fy >= 0.0f && fy <= 32767.0f) || fx != floorf(fx) ||
        fy != floorf(fy)) {
      fprintf(stderr, "aco_vrp_cuda_with_capacity: node %d coordinates (%g, %g) are not "
                      "integers in [0, 32767]\n", i, (double)fx, (double)fy);
      return ACO_ERR_INVALID_INPUT;
    }
    x[i] = (int32_t)fx;
    y[i] = (int32_t)fy;
  }
  /* unit demand: a capacity above n never binds */
  const int cap = vehicle_capacity_customers > n ? n : vehicle_capacity_customers;
  if (cap > 32767) {
    fprintf(stderr, "aco_vrp_cuda_with_capacity: capacity %d above 32767 customers\n", cap);
    return ACO_ERR_INVALID_INPUT;
  }
  std::vector<int32_t> tour((size_t)n + (size_t)K + 1 + 64);
  int32_t len = 0;
  int64_t cost = 0;
  /* m == 0 selects 256 ants (aco_cuda.cu:206) inside the one-call solver */
  const int rc = aco_vrp_hip_with_capacity(n, K, cap, m, x.data(), y.data(), demand.data(), alpha,
                                           beta, rho, tau0, Q, seed, ACO_HIP_DIST_REAL,
                                           tour.data(), &len, &cost);
  if (rc != ACO_HIP_OK) {
    if (rc == ACO_HIP_ERR_NO_SOLUTION) *best_cost = DBL_MAX;
    return (AcoStatus)rc;
  }
  if (!solution_from_tour(best_solution, tour.data(), len)) return ACO_ERR_BACKEND;
  *best_cost = (double)cost * ACO_HIP_REAL_COST_UNIT;
  return ACO_OK;
}
