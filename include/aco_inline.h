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

/* ---- real (unrounded) distances ----------------------------------------------
 * The real distance of an edge is d = the float nearest to sqrt(s), s =
 * dx^2 + dy^2 <= 2*32767^2 < 2^31 (exactly what (float)sqrt((double)s)
 * gives with a correctly rounded double square root).  A hardware estimate
 * from (float)s is within two floats of d.  The midpoint of two neighbouring
 * floats has at most 25 significant bits, so its square is exact in double
 * and comparing s with it decides exactly which float is nearest; sqrt(s)
 * is never such a midpoint (it is an integer below 2^24, hence a float, or
 * irrational). */
static inline ACO_HD float aco_bits_float(uint32_t b) {
  union { uint32_t u; float f; } v;
  v.u = b;
  return v.f;
}
static inline ACO_HD uint32_t aco_float_bits(float f) {
  union { uint32_t u; float f; } v;
  v.f = f;
  return v.u;
}

static inline ACO_HD float aco_real_dist_s(uint32_t s) {
  if (s == 0u) return 0.0f;
  float d = aco_sqrt_estimate((float)s);
  const double S = (double)s;
  for (int it = 0; it < 4; ++it) {
    const float up = aco_bits_float(aco_float_bits(d) + 1u);
    const double mu = 0.5 * ((double)d + (double)up);
    if (S > mu * mu) {
      d = up;
      continue;
    }
    const float dn = aco_bits_float(aco_float_bits(d) - 1u);
    const double md = 0.5 * ((double)dn + (double)d);
    if (S < md * md) {
      d = dn;
      continue;
    }
    break;
  }
  return d;
}

/* Route-cost units of a real distance: d * 2^23, an exact integer (every
 * float >= 1 is a multiple of 2^-23), so tour costs are exact int64 sums
 * whatever the order of the additions.  Cost = units / ACO_REAL_UNITS. */
#define ACO_REAL_UNITS 8388608.0
static inline ACO_HD int64_t aco_real_units(float d) {
  return (int64_t)((double)d * ACO_REAL_UNITS);
}

/* Heuristic-table index of a real distance: 0 for d = 0, else 1 + the bin of
 * d among 1024 equal bins of its binade (bins of relative width 2^-10; at
 * most 15787 for d <= 46341).  The table holds eta at each bin's middle float. */
#define ACO_REAL_BIN_SHIFT 13
static inline ACO_HD uint32_t aco_real_eta_bin(float d) {
  if (!(d >= 1.0f)) return 0u;
  return 1u + ((aco_float_bits(d) - 0x3F800000u) >> ACO_REAL_BIN_SHIFT);
}
static inline ACO_HD float aco_real_bin_centre(uint32_t b) {
  return aco_bits_float(0x3F800000u + ((b - 1u) << ACO_REAL_BIN_SHIFT) +
                        (1u << (ACO_REAL_BIN_SHIFT - 1)));
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
