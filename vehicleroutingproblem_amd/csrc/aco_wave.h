/*
 * aco_wave.h - wavefront primitives of the gfx950 kernels: row-shift DPP
 * scans, lane reads, ballots and wave-wide reductions.
 */
#ifndef ACO_WAVE_H
#define ACO_WAVE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ float f_dpp_row_shr1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, true));
}
__device__ __forceinline__ float f_dpp_row_shr2(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xF, 0xF, true));
}
__device__ __forceinline__ float f_dpp_row_shr4(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xF, 0xF, true));
}
__device__ __forceinline__ float f_dpp_row_shr8(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xF, 0xF, true));
}

__device__ __forceinline__ float f_readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

/* Inclusive scan: Kogge-Stone by row_shr 1,2,4,8 inside each 16-lane row
 * (lanes shifted in from outside the row read 0), then row r adds the last
 * prefix of row r-1 for r = 1..rows-1.  The CPU oracle restates this order. */
__device__ __forceinline__ float wave_scan_rows(float v, int lane, int rows) {
  v = v + f_dpp_row_shr1(v);
  v = v + f_dpp_row_shr2(v);
  v = v + f_dpp_row_shr4(v);
  v = v + f_dpp_row_shr8(v);
  if (rows > 1) {
    const float c = f_readlane(v, 15);
    if ((lane >> 4) == 1) v = v + c;
  }
  if (rows > 2) {
    const float c = f_readlane(v, 31);
    if ((lane >> 4) == 2) v = v + c;
  }
  if (rows > 3) {
    const float c = f_readlane(v, 47);
    if ((lane >> 4) == 3) v = v + c;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ int first_lane(uint64_t m) { return __builtin_ctzll(m); }
/* index of the lowest set bit of a wave-uniform word, -1 when it is 0
 * (s_ff1_i32_b32 gives -1 for 0 itself; ffs() would add a compare and a
 * select).  The readfirstlane costs nothing when x is already scalar and
 * keeps the operand an SGPR when the compiler holds it in a VGPR. */
__device__ __forceinline__ int sff1(uint32_t x) {
  int r;
  asm("s_ff1_i32_b32 %0, %1" : "=s"(r) : "s"(__builtin_amdgcn_readfirstlane((int)x)));
  return r;
}
__device__ __forceinline__ int last_lane(uint64_t m) { return 63 - __builtin_clzll(m); }

#endif /* ACO_WAVE_H */
