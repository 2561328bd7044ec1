/*
 * aco_cuda_compat.h - the reference CUDA back-end's entry point with exactly
 * its prototype and linkage, so that the reference CLI links against
 * libaco_hip.so unchanged.
 *
 * The reference declares it in src/cuda/main_vrp.cu:29-35 (outside its
 * extern "C" block, so with C++ linkage) and defines it in
 * src/cuda/aco_cuda.cu:188-192.  libaco_hip.so exports the same C++ symbol.
 * Semantics kept: unit demand with the capacity counted in customers, the
 * reference's unrounded distances (ACO_HIP_DIST_REAL), candidate lists of 32
 * (n when n <= 32), m = 0 meaning 256 ants, stop rules from ACO_SOLVER_*,
 * the best tour copied into the caller's Solution (made by solution_create(K,
 * n)) route by route with both depots, and its length in *best_cost.
 * Coordinates must be integers in [0, 32767] (the reference's instance
 * generator writes integer grid coordinates); others are refused with
 * ACO_ERR_INVALID_INPUT.  C callers use aco_vrp_hip_with_capacity
 * (aco_hip.h).
 */
#ifndef ACO_CUDA_COMPAT_H
#define ACO_CUDA_COMPAT_H

#ifndef __cplusplus
#error "aco_vrp_cuda_with_capacity has C++ linkage, as in the reference"
#endif

#include "aco_host.h"

#ifndef ACO_H /* the reference's include/aco.h:82-88 defines the same enum */
typedef enum {
  ACO_OK = 0,
  ACO_ERR_INVALID_INPUT = 1,
  ACO_ERR_ALLOCATION = 2,
  ACO_ERR_NO_SOLUTION = 3,
  ACO_ERR_BACKEND = 4
} AcoStatus;
#endif

AcoStatus aco_vrp_cuda_with_capacity(int n, int K, int vehicle_capacity_customers, int m,
                                     float *coords_x, float *coords_y, double alpha, double beta,
                                     double rho, double tau0, double Q, unsigned int seed,
                                     Solution *best_solution, double *best_cost);

#endif /* ACO_CUDA_COMPAT_H */
