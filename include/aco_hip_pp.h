/*
 * aco_hip_pp.h - plain-pointer entry points: the reference's per-kernel
 * launchers on caller-owned device buffers, with no context.
 *
 * The reference's host loop (src/cuda/aco_cuda.cu:267-399) cudaMallocs its
 * buffers and calls the launchers of include/aco_cuda_kernels.h:151-181 on
 * them.  These calls take the same buffers in the same layouts:
 *   - pheromone: (n+1)*(n+1) uint8 levels, row stride n+1, no padding
 *     (aco_cuda.cu:267-268), the log-quantized format of
 *     src/cuda/aco_cuda_kernels.cu:25-45 (log range 1e-4 .. 100, 255 steps);
 *   - colony routes: int32 routes[steps][m], step-major, route v of ant a
 *     at steps off_v+1 .. off_v+len-1 with off_v = sum over u < v of
 *     (lens[a*K+u] - 1), and lens[m][K] (aco_cuda_kernels.cu:592-627);
 *   - flat routes: int32 routes[K][n+1], route v's nodes 0..len_v-1 in row
 *     v, and lens[K] (aco_cuda.cu:373-379, aco_cuda_kernels.cu:637-668);
 *   - tours: the giant-tour sequence of aco_hip.h (0, c.., 0, c.., 0).
 * Coordinates are int32 x[n+1], y[n+1] in [0, 32767] (aco_hip.h), distances
 * ACO_HIP_DIST_NINT or ACO_HIP_DIST_REAL.
 *
 * Every pointer named d_* is device memory of the current HIP device.  The
 * calls only enqueue work on `stream` (a hipStream_t, NULL = the null
 * stream) and return: synchronize the stream (aco_hip_pp_sync) before
 * reading results on the host.  They allocate nothing.  Sizes and host
 * arguments are checked before anything is enqueued (ACO_HIP_ERR_INVALID_INPUT);
 * data in device buffers cannot be checked from the host, so the kernels
 * treat it defensively: a node outside 0..n or a route position outside
 * the buffer deposits nothing, and a tour with such a node or a length
 * outside 0..stride gets cost -1.  ACO_HIP_ERR_BACKEND reports a failed
 * launch.  aco_hip_pp_last_error() gives the calling thread's last message.
 *
 * Deposits equal the reference's sequential per-route edge order whenever no
 * customer appears twice in the deposited routes (the reference's own
 * one-thread-per-route kernels race otherwise, and so do these).
 */
#ifndef ACO_HIP_PP_H
#define ACO_HIP_PP_H

#include <stdint.h>

#include "aco_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* launch_init_tau (include/aco_cuda_kernels.h:151, kernel_init_tau
 * aco_cuda_kernels.cu:131-136): every byte of d_tau[(n+1)^2] = q_tau0. */
int aco_hip_pp_init_tau(uint8_t *d_tau, int n, uint8_t q_tau0, void *stream);

/* launch_evaporate_tau (include/aco_cuda_kernels.h:171, kernel_evaporate_tau
 * aco_cuda_kernels.cu:145-155): q > q_min + delta ? q - delta : q_min on
 * every byte, the diagonal included. */
int aco_hip_pp_evaporate_tau(uint8_t *d_tau, int n, uint8_t delta, uint8_t q_min,
                             void *stream);

/* launch_build_candidate_lists (include/aco_cuda_kernels.h:156,
 * kernel_build_candidate_lists aco_cuda_kernels.cu:164-214): for each node
 * i in 0..n the k (1..512) nearest customers j in 1..n, j != i, nearest
 * first, ties to the lower index (knn_key ACO_HIP_KNN_EXACT: exact integer
 * key; ACO_HIP_KNN_REF_F32: the float key of aco_parallel.c:218-250), into
 * d_cand[(n+1)*k]; slots past the n-1 (n for the depot row) customers hold
 * 0.  In place of the reference's eta_beta floats, d_code[(n+1)*k] holds
 * each slot's heuristic-table index (the nint distance, or in real mode the
 * distance bin; 0 for padding): eta_beta = table[code] with the table of
 * aco_hip_pp_eta_table. */
int aco_hip_pp_build_candidate_lists(const int32_t *d_x, const int32_t *d_y, int n, int k,
                                     int knn_key, int dist_mode, int32_t *d_cand,
                                     int32_t *d_code, void *stream);

/* The heuristic table of a distance mode, on the host: eta[code] =
 * (1/d)^beta for each table index code < *len (aco_cuda_kernels.cu:207-212).
 * With eta == NULL only *len is set.  The table covers every distance of the
 * coordinate range (nint distances 0..46340, or the real-mode bins of
 * distances up to 32767*sqrt(2)), beta any real value. */
int aco_hip_pp_eta_table(int dist_mode, float beta, float *eta, int *len);

/* launch_deposit_solution (include/aco_cuda_kernels.h:176,
 * kernel_deposit_solution aco_cuda_kernels.cu:592-627): deposit `deposit`
 * on every edge of ant best_ant's K routes in d_routes[steps][m] /
 * d_route_lengths[m][K] (layout above): per edge (u, v),
 * new = quantize(dequantize(tau[u][v]) + deposit), tau[u][v] and tau[v][u]
 * raised to new. */
int aco_hip_pp_deposit_solution(uint8_t *d_tau, int n, int K, int m, int steps,
                                const int32_t *d_routes, const int32_t *d_route_lengths,
                                float deposit, int best_ant, void *stream);

/* launch_deposit_flat_solution (include/aco_cuda_kernels.h:181,
 * kernel_deposit_flat_solution aco_cuda_kernels.cu:637-668): the same
 * deposit on the K routes of d_flat_routes[K][n+1] / d_flat_lengths[K]. */
int aco_hip_pp_deposit_flat_solution(uint8_t *d_tau, int n, int K, const int32_t *d_flat_routes,
                                     const int32_t *d_flat_lengths, float deposit, void *stream);

/* The same deposit on one giant tour d_tour[len] (aco_hip_deposit_tour on a
 * caller-owned matrix). */
int aco_hip_pp_deposit_tour(uint8_t *d_tau, int n, const int32_t *d_tour, int len,
                            float deposit, void *stream);

/* Route costs (the in-kernel sums of aco_cuda_kernels.cu:540,550): d_costs[t]
 * = cost units of tour t (d_tours[t*stride ..], d_lens[t] nodes), summed over
 * its edges in order; -1 for a tour with a length outside 0..stride or a node
 * outside 0..n. */
int aco_hip_pp_route_costs(const int32_t *d_x, const int32_t *d_y, int n, const int32_t *d_tours,
                           const int32_t *d_lens, int count, int stride, int dist_mode,
                           int64_t *d_costs, void *stream);

/* hipStreamSynchronize(stream), for C callers without HIP headers. */
int aco_hip_pp_sync(void *stream);

/* Message of the calling thread's last failed plain-pointer call. */
const char *aco_hip_pp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ACO_HIP_PP_H */
