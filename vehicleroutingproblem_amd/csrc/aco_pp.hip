/*
 * aco_pp.hip - the plain-pointer entry points: the reference's per-kernel
 * launchers (include/aco_cuda_kernels.h:151-181) on caller-owned device
 * buffers, no context, no allocation, asynchronous on the caller's stream.
 * Host arguments are checked before anything is enqueued; the kernels
 * (aco_kernels.hip, k_pp_*) check what lives in device memory themselves.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "aco_hip_pp.h"
#include "aco_host.h"
#include "aco_kernels.h"

static thread_local char g_msg[256];

static int pp_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  return code;
}

/* Clear a stale error of an earlier call, launch, then report this launch's. */
#define PP_LAUNCH(name, ...)                                                                  \
  do {                                                                                        \
    (void)hipGetLastError();                                                                  \
    hipLaunchKernelGGL(__VA_ARGS__);                                                          \
    const hipError_t e__ = hipGetLastError();                                                 \
    if (e__ != hipSuccess)                                                                    \
      return pp_err(ACO_HIP_ERR_BACKEND, "launch of %s failed: %s", name,                     \
                    hipGetErrorString(e__));                                                  \
  } while (0)

/* deq / thr of the reference's quantization (they depend on the log range
 * only, not on alpha, rho or tau0), built once */
static const PpTables &pp_tables() {
  static PpTables t;
  static std::once_flag once;
  std::call_once(once, [] {
    AcoTauTables h;
    aco_tau_tables_build(&h, 1.0f, 0.5f, 1.0f);
    std::copy(h.deq, h.deq + 256, t.deq);
    std::copy(h.thr, h.thr + 255, t.thr);
  });
  return t;
}

static int check_n(int n) {
  if (n < 1 || n > 65534) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "n must be in [1, 65534]");
  return ACO_HIP_OK;
}

static unsigned grid_for(size_t items, size_t per_block, unsigned cap) {
  const size_t b = (items + per_block - 1) / per_block;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(b, cap));
}

extern "C" int aco_hip_pp_init_tau(uint8_t *d_tau, int n, uint8_t q_tau0, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (!d_tau) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null pheromone buffer");
  const size_t bytes = (size_t)(n + 1) * (size_t)(n + 1);
  PP_LAUNCH("k_pp_fill", k_pp_fill, dim3(grid_for(bytes / 16, 256, 4096)), dim3(256), 0,
            (hipStream_t)stream, d_tau, bytes, (int)q_tau0);
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_evaporate_tau(uint8_t *d_tau, int n, uint8_t delta, uint8_t q_min,
                                        void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (!d_tau) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null pheromone buffer");
  const size_t bytes = (size_t)(n + 1) * (size_t)(n + 1);
  PP_LAUNCH("k_pp_evaporate", k_pp_evaporate, dim3(grid_for(bytes / 16, 256, 4096)), dim3(256),
            0, (hipStream_t)stream, d_tau, bytes, (int)delta, (int)q_min);
  return ACO_HIP_OK;
}

static int check_dist(int dist_mode) {
  if (dist_mode != ACO_HIP_DIST_NINT && dist_mode != ACO_HIP_DIST_REAL)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "unknown dist_mode");
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_build_candidate_lists(const int32_t *d_x, const int32_t *d_y, int n,
                                                int k, int knn_key, int dist_mode,
                                                int32_t *d_cand, int32_t *d_code, void *stream) {
  int rc;
  if ((rc = check_n(n)) || (rc = check_dist(dist_mode))) return rc;
  if (k < 1 || k > 512) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "k must be in [1, 512]");
  if (knn_key != ACO_HIP_KNN_EXACT && knn_key != ACO_HIP_KNN_REF_F32)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "unknown knn_key");
  if (!d_x || !d_y || !d_cand || !d_code)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null candidate-list argument");
  /* one wavefront per row, four rows per workgroup */
  PP_LAUNCH("k_pp_knn", k_pp_knn, dim3((unsigned)((n + 1 + 3) / 4)), dim3(256), 0,
            (hipStream_t)stream, d_x, d_y, n, k, knn_key, dist_mode == ACO_HIP_DIST_REAL ? 1 : 0,
            d_cand, d_code);
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_eta_table(int dist_mode, float beta, float *eta, int *len) {
  int rc;
  if ((rc = check_dist(dist_mode))) return rc;
  if (!len) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null length");
  int l;
  if (dist_mode == ACO_HIP_DIST_REAL)
    l = (int)aco_real_eta_bin(aco_real_dist_s(2u * 32767u * 32767u)) + 1;
  else
    l = aco_nint_dist(0, 0, 32767, 32767) + 1;
  *len = l;
  if (eta) {
    if (dist_mode == ACO_HIP_DIST_REAL)
      aco_eta_table_build_real(eta, l, beta);
    else
      aco_eta_table_build(eta, l - 1, beta);
  }
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_deposit_solution(uint8_t *d_tau, int n, int K, int m, int steps,
                                           const int32_t *d_routes,
                                           const int32_t *d_route_lengths, float deposit,
                                           int best_ant, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (K < 1 || m < 1 || steps < 1)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "K, m and steps must be >= 1");
  if (best_ant < 0 || best_ant >= m)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "best_ant must be in [0, m)");
  if (!d_tau || !d_routes || !d_route_lengths)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null deposit argument");
  PP_LAUNCH("k_pp_deposit_routes<true>", k_pp_deposit_routes<true>, dim3(1), dim3(1024), 0,
            (hipStream_t)stream, d_tau, n, K, m, (long long)steps, d_routes, d_route_lengths,
            best_ant, deposit, pp_tables());
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_deposit_flat_solution(uint8_t *d_tau, int n, int K,
                                                const int32_t *d_flat_routes,
                                                const int32_t *d_flat_lengths, float deposit,
                                                void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (K < 1) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "K must be >= 1");
  if (!d_tau || !d_flat_routes || !d_flat_lengths)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null deposit argument");
  PP_LAUNCH("k_pp_deposit_routes<false>", k_pp_deposit_routes<false>, dim3(1), dim3(1024), 0,
            (hipStream_t)stream, d_tau, n, K, 1, 0LL, d_flat_routes, d_flat_lengths, 0, deposit,
            pp_tables());
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_deposit_tour(uint8_t *d_tau, int n, const int32_t *d_tour, int len,
                                       float deposit, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (len < 0) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "len must be >= 0");
  if (len < 2) return ACO_HIP_OK;
  if (!d_tau || !d_tour) return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null deposit argument");
  PP_LAUNCH("k_pp_deposit_tour", k_pp_deposit_tour, dim3(1), dim3(1024), 0, (hipStream_t)stream,
            d_tau, n, d_tour, len, deposit, pp_tables());
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_route_costs(const int32_t *d_x, const int32_t *d_y, int n,
                                      const int32_t *d_tours, const int32_t *d_lens, int count,
                                      int stride, int dist_mode, int64_t *d_costs,
                                      void *stream) {
  int rc;
  if ((rc = check_n(n)) || (rc = check_dist(dist_mode))) return rc;
  if (count < 0 || stride < 1)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "count must be >= 0 and stride >= 1");
  if (count == 0) return ACO_HIP_OK;
  if (!d_x || !d_y || !d_tours || !d_lens || !d_costs)
    return pp_err(ACO_HIP_ERR_INVALID_INPUT, "null route-cost argument");
  /* one wavefront per tour */
  PP_LAUNCH("k_pp_route_costs", k_pp_route_costs, dim3((unsigned)(((size_t)count + 3) / 4)),
            dim3(256), 0, (hipStream_t)stream, d_x, d_y, n, d_tours, d_lens, count, stride,
            dist_mode == ACO_HIP_DIST_REAL ? 1 : 0, reinterpret_cast<long long *>(d_costs));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_pp_sync(void *stream) {
  const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess)
    return pp_err(ACO_HIP_ERR_BACKEND, "stream synchronize: %s", hipGetErrorString(e));
  return ACO_HIP_OK;
}

extern "C" const char *aco_hip_pp_last_error(void) { return g_msg; }
