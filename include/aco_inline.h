/*
 * aco_inline.h - inline arithmetic shared by the host library and the gfx950
 * kernels, so that both evaluate bit-identical values:
 *
 *  - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 *    "Parallel random numbers: as easy as 1, 2, 3", SC'11), written from the
 *    published round function.  It replaces the per-ant LCG of
 *    src/cuda/aco_cuda_kernels.cu:13-16,253.  Draw layout of the solver:
 *      key = {instance seed, ACO_PHILOX_KEY1}
 *      ctr = {decision index within the ant's tour, ant, epoch, colony}
 *      out[0] -> move roulette, out[1] -> depot-close roulette;
 *  - TSPLIB EUC_2D nearest-integer distance (exact integer square root);
 *  - pheromone quantization through a threshold table.
 */
#ifndef ACO_INLINE_H
#define ACO_INLINE_H

#include <stdint.h>

#if defined(__HIPCC__)
#define ACO_HD __host__ __device__
#else
#define ACO_HD
#endif

#define ACO_PHILOX_M0 0xD2511F53u
#define ACO_PHILOX_M1 0xCD9E8D57u
#define ACO_PHILOX_W0 0x9E3779B9u
#define ACO_PHILOX_W1 0xBB67AE85u
#define ACO_PHILOX_KEY1 0x5CA1AB1Eu

static inline ACO_HD void aco_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1, uint32_t out[4]) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)ACO_PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)ACO_PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += ACO_PHILOX_W0;
    k1 += ACO_PHILOX_W1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Uniform float in [0,1) from the top 24 bits (exact conversion). */
static inline ACO_HD float aco_u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

/* nint(sqrt(s)) for s <= 2*32767^2, exact.  A float estimate rounded to the
 * nearest integer is within one of the answer; one integer test fixes it:
 * nint(sqrt(s)) = r  <=>  r^2 - r < s <= r^2 + r. */
static inline ACO_HD float aco_sqrt_estimate(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x); /* one v_sqrt_f32 (1 ulp): the fix-up below is exact */
#else
  return __builtin_sqrtf(x);
#endif
}

/* Products of operands below 2^23 whose result fits in 32 bits: one
 * full-rate 24-bit multiply on the GPU (the 32-bit integer multiply runs at
 * a quarter of the rate), a plain multiply elsewhere; exact either way. */
static inline ACO_HD uint32_t aco_umul24(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul24(a, b);
#else
  return a * b;
#endif
}

static inline ACO_HD uint32_t aco_sq15(int32_t d) { /* d^2 for |d| <= 32767 */
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint32_t)__mul24(d, d);
#else
  return (uint32_t)(d * d);
#endif
}

static inline ACO_HD int32_t aco_nint_sqrt_u32(uint32_t s) {
  uint32_t r = (uint32_t)(aco_sqrt_estimate((float)s) + 0.5f);
  const uint32_t sq = aco_umul24(r, r); /* r <= 46341: r^2 < 2^32 */
  if (s > sq + r) ++r;
  else if (r > 0u && s <= sq - r) --r;
  return (int32_t)r;
}

/* TSPLIB EUC_2D distance for coordinates in [0, 32767]. */
static inline ACO_HD int32_t aco_nint_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2) {
  const uint32_t s = aco_sq15(x1 - x2) + aco_sq15(y1 - y2);
  return aco_nint_sqrt_u32(s);
}

/* Quantize tau through the 255 increasing thresholds thr[b] (the smallest
 * float whose level is >= b+1): level = number of thresholds <= tau. */
static inline ACO_HD uint32_t aco_quantize_thr(const float *thr, float tau) {
  if (!(tau > 0.0f)) return 0u;
  uint32_t lo = 0u, hi = 255u;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (tau >= thr[mid]) lo = mid + 1u;
    else hi = mid;
  }
  return lo;
}

#endif /* ACO_INLINE_H */
