/*
 * Quantization and heuristic tables.  The kernels never evaluate logf/expf/
 * powf: dequantization, tau^alpha, eta^beta and quantization all go through
 * these host-built tables, which the CPU oracle rebuilds from the same
 * formulas, so GPU and CPU agree bit for bit.
 */
#include "aco_host.h"
#include "aco_inline.h"

#include <math.h>

/* Quantization parameters of src/cuda/aco_cuda.cu:209-215 and :242, evaluated
 * with host float math exactly as there.  deq follows dequantize_tau
 * (src/cuda/aco_cuda_kernels.cu:25-27), whose q*step+log_min nvcc contracts to
 * one fused multiply-add; exp is evaluated in double and rounded once.
 * thr[b] is the smallest float whose ideal level
 * (log(tau)-log_min)/step is >= b+1/2, i.e. the rounding boundary of
 * quantize_tau (aco_cuda_kernels.cu:38-45) computed in double. */
void aco_tau_tables_build(AcoTauTables *t, float alpha, float rho, float tau0) {
  const float log_min = logf(0.0001f);
  const float log_max = logf(100.0f);
  const float step = (log_max - log_min) / (float)(255 - 0);
  const float log_rho = logf(1.0f - rho);
  t->log_tau_min = log_min;
  t->log_tau_step = step;
  t->q_evap = (uint8_t)fmaxf(1.0f, roundf(-log_rho / step));
  t->q_tau0 = (uint8_t)fmaxf(0.0f, fminf(255.0f, roundf((logf(tau0) - log_min) / step)));
  for (int q = 0; q < 256; ++q) {
    float e = fmaf((float)q, step, log_min);
    t->deq[q] = (float)exp((double)e);
    t->taupow[q] = (alpha == 1.0f) ? t->deq[q] : (float)pow((double)t->deq[q], (double)alpha);
  }
  for (int b = 0; b < 255; ++b) {
    const double target = (double)log_min + ((double)b + 0.5) * (double)step;
    float f = (float)exp(target);
    while (log((double)f) < target) f = nextafterf(f, INFINITY);
    for (;;) {
      float p = nextafterf(f, 0.0f);
      if (log((double)p) >= target) f = p;
      else break;
    }
    t->thr[b] = f;
  }
}

/* eta^beta for integer distance d: (1/(d + 1e-9f))^beta as in
 * kernel_build_candidate_lists (aco_cuda_kernels.cu:207-212) and the depot
 * score (aco_cuda_kernels.cu:504-506).  For d >= 1 the 1e-9f is below half an
 * ulp of d, so the value is d^-beta, evaluated in double and rounded once. */
void aco_eta_table_build(float *eta, int dmax, float beta) {
  for (int d = 0; d <= dmax; ++d) {
    double base = (d == 0) ? (double)1e-9f : (double)d;
    eta[d] = (float)pow(base, -(double)beta);
  }
}

/* eta^beta in real mode, by heuristic-table index (aco_real_eta_bin):
 * index 0 is d = 0, which uses 1e-9f like 1/(0 + CUDA_EPS); index b >= 1 is
 * the middle float c of its bin, eta = c^-beta in double, rounded once.
 * The reference evaluates powf(1/(d + 1e-9f), beta) of the float distance
 * (aco_cuda_kernels.cu:207-212) and rsqrtf(d^2 + 1e-9f)^beta in the fallback
 * (:400-406); within a bin of relative width 2^-10 the table value is within
 * beta * 2^-11 (relative) of d^-beta. */
void aco_eta_table_build_real(float *eta, int len, float beta) {
  for (int b = 0; b < len; ++b) {
    const double base = b == 0 ? (double)1e-9f : (double)aco_real_bin_centre((uint32_t)b);
    eta[b] = (float)pow(base, -(double)beta);
  }
}

/* Real distances of squared lengths (the arithmetic of the kernels). */
void aco_real_dist_batch(const uint32_t *s, int64_t count, float *out) {
  for (int64_t i = 0; i < count; ++i) out[i] = aco_real_dist_s(s[i]);
}

void aco_philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  aco_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

uint8_t aco_tau_quantize(const AcoTauTables *t, float tau) {
  return (uint8_t)aco_quantize_thr(t->thr, tau);
}

int vrp_nint_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2) {
  return aco_nint_dist(x1, y1, x2, y2);
}

double vrp_real_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2) {
  double dx = (double)x1 - (double)x2, dy = (double)y1 - (double)y2;
  return sqrt(dx * dx + dy * dy);
}
