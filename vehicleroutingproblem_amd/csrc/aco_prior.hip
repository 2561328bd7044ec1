/*
 * aco_prior.hip - the learned heuristic prior: an edge-scoring network on the
 * kNN candidate graph whose output replaces (1/d)^beta at the seam where the
 * reference multiplies eta_beta into a candidate's weight
 * (aco_cuda_kernels.cu:295).  The network has no reference counterpart.
 *
 * Network: 8 edge features -> 64 -> 64 -> 1 logit, ReLU, BF16 weights and
 * activations into the matrix cores, FP32 accumulation, FP32 master weights
 * with Adam.  The products over the 64 hidden units (layer 2 and dH1 =
 * dZ2 W2) use v_mfma_f32_16x16x32_bf16, two K=16 slices per instruction;
 * layer 1 (K = 8 features, padded to 16) and the products over a tile's 16
 * edges (dW2, dW1) use v_mfma_f32_16x16x16bf16_1k.  Trained on the GPU by the kernels below with a
 * mean binary cross-entropy against "the edge joins consecutive nodes of an
 * elite tour".  Installed as eta = (1/d)^beta * gain * sigmoid(logit), so the
 * candidate, depot-close and fallback weights stay on one scale.
 *
 * Work split: one wavefront per 16-edge tile.  MFMA lane layout (16x16x16):
 * A: lane l holds A[l%16][4(l/16)+i], B: B[4(l/16)+i][l%16], D: D[4(l/16)+r][l%16].
 * The backward products dW2 = dZ2^T H1 and dW1 = dZ1^T X take both operands
 * straight from D-layout registers; H1 (layer 2) and dZ2 (dH1 = dZ2 W2) are
 * transposed through LDS (rows padded against bank conflicts, as are the LDS
 * copies of W2 and W2^T).  Gradients are summed per wavefront in registers
 * over its tiles, the wavefronts of a workgroup are added in LDS, one row is
 * written per workgroup, and the rows are reduced in a fixed order in two
 * stages (deterministic) before the update.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "aco_ctx.h"
#include "aco_inline.h"

/* parameter layout (FP32 master copy) */
#define MLP_F 8
#define MLP_H 64
#define PW1 0      /* W1[64][8] */
#define PB1 512    /* b1[64] */
#define PW2 576    /* W2[64][64] (row = output neuron) */
#define PB2 4672   /* b2[64] */
#define PW3 4736   /* w3[64] */
#define PB3 4800   /* b3 */
#define NPAR 4801
#define PROW (NPAR + 1) /* gradient row: parameters, then the loss sum */
/* BF16 copies of the matrices */
#define BW1 0
#define BW2 512
#define NBW 4608

#define MLP_GRID 1024 /* wavefronts of the gradient kernel (rows of partial sums) */
#define RED_SPLIT 32  /* first-stage splits of the row reduction */
#define GRAD_WAVES 2  /* wavefronts per gradient workgroup */
/* LDS row stride (BF16 elements) of W2, W2^T and the transpose buffers: 68
 * (34 dwords) puts the 16 rows a 16-lane group reads with one ds_read_b64 on
 * 16 distinct bank pairs; a 64-element stride put them all on one bank. */
#define MLP_HS 68

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float fl4 __attribute__((ext_vector_type(4)));

/* FP32 -> BF16, round to nearest even (finite inputs). */
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t u = __float_as_uint(f);
#else
  uint32_t u;
  memcpy(&u, &f, 4);
#endif
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

/* four FP32 -> BF16 (round to nearest even, as f2bf) with the hardware's
 * packed conversion, two values per instruction */
__device__ __forceinline__ bf4 pack4(float a, float b, float c, float d) {
  const bf16x2 lo = __builtin_convertvector(fl2{a, b}, bf16x2);
  const bf16x2 hi = __builtin_convertvector(fl2{c, d}, bf16x2);
  return __builtin_bit_cast(bf4, u32x2{__builtin_bit_cast(uint32_t, lo),
                                       __builtin_bit_cast(uint32_t, hi)});
}

__device__ __forceinline__ fl4 mfma16(bf4 a, bf4 b, fl4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short sh8 __attribute__((ext_vector_type(8)));

/* v_mfma_f32_16x16x32_bf16 (gfx950, twice the flops per cycle of the
 * 16x16x16 form) on two K=16 slices at once: lane l's A values are
 * (a0, a1) and its B values (b0, b1), so k-slot i of a0 meets k-slot i of
 * b0 and so on; any such pairing of two 16x16x16 products gives their sum. */
__device__ __forceinline__ fl4 mfma32(bf4 a0, bf4 a1, bf4 b0, bf4 b1, fl4 c) {
  const sh8 a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
  const sh8 b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                 __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL> __device__ __forceinline__ float dpp_add(float v) {
  return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                 0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

/* sum over the 16 lanes of a row (lanes with equal l/16) with DPP: quad swaps
 * (quad_perm [1,0,3,2], [2,3,0,1]), then row rotations by 4 and 8 */
__device__ __forceinline__ float row_sum16(float v) {
  v = dpp_add<0xB1>(v);
  v = dpp_add<0x4E>(v);
  v = dpp_add<0x124>(v);
  return dpp_add<0x128>(v);
}

/* ---- edge features -------------------------------------------------------- */

/* Features of candidate slot (i, t), j = cand[i][t], depot 0, L = the
 * instance's coordinate scale:
 *   0,1: (x_i - x_0)/L, (y_i - y_0)/L      2,3: (x_j - x_0)/L, (y_j - y_0)/L
 *   4: demand_j / Q                         5: nint d(i,j) / L
 *   6: cosine of the angle at i between i->j and i->depot (0 if degenerate,
 *      computed in double and rounded once)
 *   7: t / k (rank in the candidate list)
 * each rounded to BF16.  One thread per slot. */
__global__ void __launch_bounds__(256) k_edge_features(DevCtx P, int inst, uint16_t *X) {
  const InstDesc I = P.inst[inst];
  const int k = I.k;
  const long slots = (long)(I.n + 1) * k;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots) return;
  const int i = (int)(e / k), t = (int)(e - (long)i * k);
  const uint32_t r = P.rec[I.cand_off + e];
  const int j = (int)(r & 0xFFFFu);
  const int2 *xy = P.xy + I.node_off;
  const int2 p0 = xy[0], pi = xy[i], pj = xy[j];
  /* nint distance in either distance mode; 0 on padding slots (index 0,
   * which no candidate list holds otherwise) */
  const int d = j ? aco_nint_dist(pi.x, pi.y, pj.x, pj.y) : 0;
  const float L = I.scale;
  float f[MLP_F];
  f[0] = (float)(pi.x - p0.x) / L;
  f[1] = (float)(pi.y - p0.y) / L;
  f[2] = (float)(pj.x - p0.x) / L;
  f[3] = (float)(pj.y - p0.y) / L;
  f[4] = (float)P.dem[I.node_off + j] / (float)I.Q;
  f[5] = (float)d / L;
  const long ux = pj.x - pi.x, uy = pj.y - pi.y, vx = p0.x - pi.x, vy = p0.y - pi.y;
  const long nu = ux * ux + uy * uy, nv = vx * vx + vy * vy;
  f[6] = (nu > 0 && nv > 0)
             ? (float)((double)(ux * vx + uy * vy) / (sqrt((double)nu) * sqrt((double)nv)))
             : 0.0f;
  f[7] = (float)t / (float)k;
  uint16_t *o = X + e * MLP_F;
#pragma unroll
  for (int q = 0; q < MLP_F; ++q) o[q] = f2bf(f[q]);
}

/* ---- network kernels -------------------------------------------------------- */

struct __align__(16) MlpLds {
  uint16_t w2[MLP_H * MLP_HS];  /* W2 row-major [out][in], rows padded to MLP_HS */
  uint16_t w2t[MLP_H * MLP_HS]; /* W2 transposed [in][out] */
  uint16_t tr[16 * MLP_HS];     /* transpose buffer of one 16-edge tile */
};

/* Thread l < 64 stages row l of W2 (64 BF16, 128 B) with eight independent
 * 16-byte loads, so a wavefront waits for one memory latency, not 64; the
 * caller synchronises (wave_sync for one wavefront, __syncthreads for more). */
__device__ __forceinline__ void stage_w2(uint16_t *w2, uint16_t *w2t, const uint16_t *wb,
                                         int l) {
  if (l >= MLP_H) return;
  const uint4 *src = reinterpret_cast<const uint4 *>(wb + BW2 + l * MLP_H);
  uint4 v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = src[q];
  uint2 *dst = reinterpret_cast<uint2 *>(w2 + l * MLP_HS); /* 8-byte aligned rows */
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    dst[2 * q] = make_uint2(v[q].x, v[q].y);
    dst[2 * q + 1] = make_uint2(v[q].z, v[q].w);
    const uint32_t u[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
    for (int h = 0; h < 8; ++h)
      w2t[(8 * q + h) * MLP_HS + l] = (uint16_t)(u[h >> 1] >> (16 * (h & 1)));
  }
}

/* per-lane weights that stay in registers over all tiles of a wavefront */
struct LaneW {
  bf4 w1[4];                   /* B operand of layer 1: W1[16nt+c][4g..4g+3] (g < 2) */
  float b1[4], b2[4], w3[4];   /* neuron 16nt+c */
  float b3;
};

__device__ __forceinline__ void load_lane_w(LaneW &W, const float *__restrict__ w,
                                            const uint16_t *__restrict__ wb, int l) {
  const int g = l >> 4, c = l & 15;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    W.w1[nt] = bf4{0, 0, 0, 0};
    if (g < 2) W.w1[nt] = *reinterpret_cast<const bf4 *>(wb + BW1 + (16 * nt + c) * MLP_F + 4 * g);
    W.b1[nt] = w[PB1 + 16 * nt + c];
    W.b2[nt] = w[PB2 + 16 * nt + c];
    W.w3[nt] = w[PW3 + 16 * nt + c];
  }
  W.b3 = w[PB3];
}

/* Tile inputs are loaded from clamped indices (always inside [0, E)) with no
 * branch, and masked where they are used: a load issued for the next tile then
 * stays in flight while the current tile computes. */
struct TileIn {
  bf4 xa;      /* lane (g < 2, c): features 4g..4g+3 of edge e0+c */
  bf4 xb;      /* lane (g, c < 8): feature c of edges e0+4g..e0+4g+3 */
  uint32_t y[4]; /* lane (g, .): labels of edges e0+4g..e0+4g+3 */
};

__device__ __forceinline__ long clamp_e(long e, long E) { return e < E ? e : E - 1; }

__device__ __forceinline__ TileIn load_tile(const uint16_t *__restrict__ X,
                                            const uint8_t *__restrict__ Y, long E, long e0,
                                            int l) {
  const int g = l >> 4, c = l & 15;
  TileIn t;
  t.xa = *reinterpret_cast<const bf4 *>(X + clamp_e(e0 + c, E) * MLP_F + 4 * (g & 1));
#pragma unroll
  for (int i = 0; i < 4; ++i) t.xb[i] = (short)X[clamp_e(e0 + 4 * g + i, E) * MLP_F + (c & 7)];
#pragma unroll
  for (int r = 0; r < 4; ++r) t.y[r] = Y ? (uint32_t)Y[clamp_e(e0 + 4 * g + r, E)] : 0u;
  return t;
}

/* A operand of layer 1 (zero for g >= 2 and past E) */
__device__ __forceinline__ bf4 tile_xa(const TileIn &t, long E, long e0, int l) {
  const int g = l >> 4, c = l & 15;
  const bool ok = g < 2 && e0 + c < E;
  bf4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = ok ? t.xa[i] : (short)0;
  return r;
}

/* B operand of dW1 = dZ1^T X (zero for c >= 8 and past E) */
__device__ __forceinline__ bf4 tile_xb(const TileIn &t, long E, long e0, int l) {
  const int g = l >> 4, c = l & 15;
  bf4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = (c < MLP_F && e0 + 4 * g + i < E) ? t.xb[i] : (short)0;
  return r;
}

/* Forward pass of one 16-edge tile (A operand xa; rows past E are zero).
 * Returns the pre-activations z1, z2 in D layout (edge 4g+r, neuron 16nt+c)
 * and the logits of edges 4g+r (identical in the 16 lanes of row g). */
__device__ __forceinline__ void tile_forward(const LaneW &W, bf4 xa, const uint16_t *w2,
                                             uint16_t *tr, int l, fl4 (&z1)[4], bf4 (&h1b)[4],
                                             fl4 (&z2)[4], fl4 &z3) {
  const int g = l >> 4, c = l & 15;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float b = W.b1[nt];
    fl4 acc = {b, b, b, b};
    z1[nt] = mfma16(xa, W.w1[nt], acc);
    h1b[nt] = pack4(fmaxf(z1[nt][0], 0.0f), fmaxf(z1[nt][1], 0.0f), fmaxf(z1[nt][2], 0.0f),
                    fmaxf(z1[nt][3], 0.0f));
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int r = 0; r < 4; ++r) tr[(4 * g + r) * MLP_HS + 16 * nt + c] = (uint16_t)h1b[nt][r];
  wave_sync();
  bf4 ha[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
    ha[ks] = *reinterpret_cast<const bf4 *>(&tr[c * MLP_HS + 16 * ks + 4 * g]);
  wave_sync();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float b = W.b2[nt];
    fl4 acc = {b, b, b, b};
#pragma unroll
    for (int ks = 0; ks < 4; ks += 2) {
      const uint16_t *wr = &w2[(16 * nt + c) * MLP_HS + 16 * ks + 4 * g];
      acc = mfma32(ha[ks], ha[ks + 1], *reinterpret_cast<const bf4 *>(wr),
                   *reinterpret_cast<const bf4 *>(wr + 16), acc);
    }
    z2[nt] = acc;
  }
  fl4 s = {0, 0, 0, 0};
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int r = 0; r < 4; ++r) s[r] = fmaf(fmaxf(z2[nt][r], 0.0f), W.w3[nt], s[r]);
#pragma unroll
  for (int r = 0; r < 4; ++r) z3[r] = row_sum16(s[r]) + W.b3;
}

/* logits of E edges; grid-stride over 16-edge tiles, one wavefront each */
__global__ void __launch_bounds__(64) k_mlp_forward(const uint16_t *__restrict__ X, long E,
                                                    const float *__restrict__ w,
                                                    const uint16_t *__restrict__ wb,
                                                    float *__restrict__ logits) {
  __shared__ MlpLds L;
  const int l = threadIdx.x, g = l >> 4, c = l & 15;
  LaneW W;
  load_lane_w(W, w, wb, l);
  stage_w2(L.w2, L.w2t, wb, l);
  wave_sync();
  long tile = blockIdx.x;
  TileIn in = load_tile(X, nullptr, E, tile * 16, l);
  for (; tile * 16 < E; tile += gridDim.x) {
    const long e0 = tile * 16;
    const TileIn next = load_tile(X, nullptr, E, (tile + gridDim.x) * 16, l);
    fl4 z1[4], z2[4], z3;
    bf4 h1b[4];
    tile_forward(W, tile_xa(in, E, e0, l), L.w2, L.tr, l, z1, h1b, z2, z3);
    if (c == 0)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (e0 + 4 * g + r < E) logits[e0 + 4 * g + r] = z3[r];
    in = next;
  }
}

/* LDS of the gradient kernel: W1 (read per tile: in registers it would
 * spill at two wavefronts per SIMD), W2 and W2^T shared by the workgroup, a
 * transpose buffer per wavefront; after the tiles the W2/transpose bytes carry
 * the partial sums of wavefronts 1.. to wavefront 0 (GRAD_XCHG floats per
 * lane). */
#define GRAD_XCHG 94
struct __align__(16) MlpGradLds {
  uint16_t w1[MLP_H * MLP_F];
  union {
    struct {
      uint16_t w2[MLP_H * MLP_HS];
      uint16_t w2t[MLP_H * MLP_HS];
      uint16_t tr[GRAD_WAVES][16 * MLP_HS];
    } m;
    float xchg[(GRAD_WAVES - 1) * GRAD_XCHG * 64];
  };
};

/* Gradient of sum_e BCE(sigmoid(z_e), y_e) * inv_e and, when want_loss, the
 * BCE sum (zero otherwise), one row of PROW partial sums per workgroup
 * (row = blockIdx.x).  GRAD_WAVES wavefronts per workgroup (two per SIMD fit
 * the register budget) take tiles gw, gw + total, ... (gw = the global
 * wavefront index); their sums meet in LDS, added in wavefront order, before
 * wavefront 0 writes the row.  The next tile's features and labels are loaded
 * while the current tile computes. */
__global__ void __launch_bounds__(64 * GRAD_WAVES, 2)
    k_mlp_grad(const uint16_t *__restrict__ X, const uint8_t *__restrict__ Y, long E,
               float inv_e, const float *__restrict__ w, const uint16_t *__restrict__ wb,
               float *__restrict__ part, int want_loss) {
  __shared__ MlpGradLds L;
  const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, g = l >> 4, c = l & 15;
  uint16_t *tr = L.m.tr[wave];
  LaneW W;
  load_lane_w(W, w, wb, l);
  stage_w2(L.m.w2, L.m.w2t, wb, threadIdx.x);
  for (int i = threadIdx.x; i < MLP_H * MLP_F; i += blockDim.x) L.w1[i] = wb[BW1 + i];
  __syncthreads();
  fl4 acc1[4], acc2[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    acc1[a] = fl4{0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < 4; ++b) acc2[a][b] = fl4{0, 0, 0, 0};
  }
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0};
  float gb3 = 0.0f, loss = 0.0f;
  const long stride = (long)gridDim.x * GRAD_WAVES;
  long tile = (long)blockIdx.x * GRAD_WAVES + wave;
  TileIn in = load_tile(X, Y, E, tile * 16, l);
  for (; tile * 16 < E; tile += stride) {
    const long e0 = tile * 16;
    const TileIn next = load_tile(X, Y, E, (tile + stride) * 16, l);
    fl4 z1[4], z2[4], z3;
    bf4 h1b[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const bf4 v = *reinterpret_cast<const bf4 *>(&L.w1[(16 * nt + c) * MLP_F + 4 * (g & 1)]);
      W.w1[nt] = g < 2 ? v : bf4{0, 0, 0, 0};
    }
    tile_forward(W, tile_xa(in, E, e0, l), L.m.w2, tr, l, z1, h1b, z2, z3);
    fl4 dz3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      dz3[r] = 0.0f;
      if (e0 + 4 * g + r < E) {
        const float z = z3[r], y = (float)in.y[r];
        const float p = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
        dz3[r] = (p - y) * inv_e;
        if (want_loss && c == 0) loss += fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
      }
    }
    if (c == 0) gb3 += (dz3[0] + dz3[1]) + (dz3[2] + dz3[3]);
    fl4 dz2[4];
    bf4 dz2b[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const float w3n = W.w3[nt];
      float sw = 0.0f, sb = 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sw = fmaf(dz3[r], fmaxf(z2[nt][r], 0.0f), sw);
        dz2[nt][r] = z2[nt][r] > 0.0f ? dz3[r] * w3n : 0.0f;
        sb += dz2[nt][r];
      }
      gw3[nt] += sw;
      gb2[nt] += sb;
      dz2b[nt] = pack4(dz2[nt][0], dz2[nt][1], dz2[nt][2], dz2[nt][3]);
    }
    /* dW2 += dZ2^T H1 (K = the 16 edges) */
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc2[mt][nt] = mfma16(dz2b[mt], h1b[nt], acc2[mt][nt]);
    /* dH1 = dZ2 W2: dZ2 transposed through LDS */
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int r = 0; r < 4; ++r) tr[(4 * g + r) * MLP_HS + 16 * nt + c] = (uint16_t)dz2b[nt][r];
    wave_sync();
    const bf4 xb = tile_xb(in, E, e0, l);
    bf4 da[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      da[ks] = *reinterpret_cast<const bf4 *>(&tr[c * MLP_HS + 16 * ks + 4 * g]);
    wave_sync();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      fl4 dh = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < 4; ks += 2) {
        const uint16_t *wr = &L.m.w2t[(16 * nt + c) * MLP_HS + 16 * ks + 4 * g];
        dh = mfma32(da[ks], da[ks + 1], *reinterpret_cast<const bf4 *>(wr),
                    *reinterpret_cast<const bf4 *>(wr + 16), dh);
      }
      fl4 dz1;
      float sb = 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        dz1[r] = z1[nt][r] > 0.0f ? dh[r] : 0.0f;
        sb += dz1[r];
      }
      gb1[nt] += sb;
      /* dW1 += dZ1^T X (K = the 16 edges; feature columns 8..15 are zero) */
      acc1[nt] = mfma16(pack4(dz1[0], dz1[1], dz1[2], dz1[3]), xb, acc1[nt]);
    }
    in = next;
  }
  /* wavefronts 1.. hand their sums to wavefront 0 through LDS (the weight
   * and transpose buffers are dead from here) */
  __syncthreads();
  if (wave > 0) {
    float *x = L.xchg + (size_t)(wave - 1) * GRAD_XCHG * 64 + l;
    int v = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) x[64 * v++] = acc2[a][b][r];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) x[64 * v++] = acc1[a][r];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      x[64 * v++] = gb1[a];
      x[64 * v++] = gb2[a];
      x[64 * v++] = gw3[a];
    }
    x[64 * v++] = gb3;
    x[64 * v++] = loss;
  }
  __syncthreads();
  if (wave > 0) return;
#pragma unroll
  for (int q = 0; q < GRAD_WAVES - 1; ++q) {
    const float *x = L.xchg + (size_t)q * GRAD_XCHG * 64 + l;
    int v = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc2[a][b][r] += x[64 * v++];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc1[a][r] += x[64 * v++];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      gb1[a] += x[64 * v++];
      gb2[a] += x[64 * v++];
      gw3[a] += x[64 * v++];
    }
    gb3 += x[64 * v++];
    loss += x[64 * v++];
  }
  float *row = part + (size_t)blockIdx.x * PROW;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 16 * mt + 4 * g + r;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) row[PW2 + n * MLP_H + 16 * nt + c] = acc2[mt][nt][r];
      if (c < MLP_F) row[PW1 + n * MLP_F + c] = acc1[mt][r];
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float a = gb1[nt], b = gb2[nt], d = gw3[nt];
    a += __shfl_xor(a, 16, 64);
    a += __shfl_xor(a, 32, 64);
    b += __shfl_xor(b, 16, 64);
    b += __shfl_xor(b, 32, 64);
    d += __shfl_xor(d, 16, 64);
    d += __shfl_xor(d, 32, 64);
    if (g == 0) {
      row[PB1 + 16 * nt + c] = a;
      row[PB2 + 16 * nt + c] = b;
      row[PW3 + 16 * nt + c] = d;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    gb3 += __shfl_xor(gb3, off, 64);
    loss += __shfl_xor(loss, off, 64);
  }
  if (l == 0) {
    row[PB3] = gb3;
    row[NPAR] = loss;
  }
}

/* First stage of the row reduction: split y (blockIdx.y) of RED_SPLIT sums
 * rows [y*rows/RED_SPLIT, (y+1)*rows/RED_SPLIT) in row order into row y of
 * part2.  The loads of a split are independent, so a thread waits for about
 * one memory latency. */
__global__ void __launch_bounds__(256) k_mlp_reduce_rows(const float *__restrict__ part, int rows,
                                                         float *__restrict__ part2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= PROW) return;
  const int y = blockIdx.y;
  const int r0 = (int)((long)y * rows / RED_SPLIT), r1 = (int)((long)(y + 1) * rows / RED_SPLIT);
  float s = 0.0f;
#pragma unroll 8
  for (int r = r0; r < r1; ++r) s += part[(size_t)r * PROW + p];
  part2[(size_t)y * PROW + p] = s;
}

/* Sum the RED_SPLIT first-stage rows in order (deterministic); optionally
 * write the sums, and with lr > 0 take one Adam step (bias-corrected rate
 * lr_t) and refresh the BF16 matrix copies. */
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ part2,
                                                    float *__restrict__ gsum, float *__restrict__ w,
                                                    float *__restrict__ m, float *__restrict__ v,
                                                    uint16_t *__restrict__ wb, float lr_t,
                                                    float beta1, float beta2, float eps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= PROW) return;
  float s = 0.0f;
#pragma unroll
  for (int r = 0; r < RED_SPLIT; ++r) s += part2[(size_t)r * PROW + p];
  if (gsum) gsum[p] = s;
  if (p >= NPAR || !(lr_t > 0.0f)) return;
  const float mm = beta1 * m[p] + (1.0f - beta1) * s;
  const float vv = beta2 * v[p] + (1.0f - beta2) * s * s;
  m[p] = mm;
  v[p] = vv;
  const float nw = w[p] - lr_t * mm / (sqrtf(vv) + eps);
  w[p] = nw;
  if (p < PB1) wb[BW1 + p - PW1] = f2bf(nw);
  else if (p >= PW2 && p < PB2) wb[BW2 + p - PW2] = f2bf(nw);
}

/* eta of every candidate slot of an instance = (1/d)^beta * gain *
 * sigmoid(logit), into the learned-heuristic array. */
__global__ void __launch_bounds__(256) k_prior_eta(DevCtx P, int inst, const float *logits,
                                                   float gain, float *ceta) {
  const InstDesc I = P.inst[inst];
  const long slots = (long)(I.n + 1) * I.k;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots) return;
  const float s = 1.0f / (1.0f + expf(-logits[e]));
  ceta[I.cand_off + e] = P.eta[P.rec[I.cand_off + e] >> 16] * (gain * s);
}

/* ---- host side -------------------------------------------------------------- */

struct AcoHipPrior {
  int device = 0;
  hipStream_t stream = nullptr;
  char err[512] = {0};
  float *d_w = nullptr, *d_m = nullptr, *d_v = nullptr, *d_part = nullptr, *d_gsum = nullptr;
  float *d_part2 = nullptr; /* first-stage sums of the gradient rows */
  uint16_t *d_wb = nullptr;
  uint16_t *d_X = nullptr; /* training set features, E x 8 BF16 */
  uint8_t *d_y = nullptr;  /* training set labels */
  int64_t E = 0, cap = 0;
  uint16_t *d_tx = nullptr; /* scratch: features / logits of one call */
  float *d_tl = nullptr;
  uint8_t *d_ty = nullptr;
  int64_t tcap = 0;
  int step = 0;
};

static int perr(AcoHipPrior *pr, int code, const char *fmt, ...) {
  if (pr) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(pr->err, sizeof(pr->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define PHIP(pr, call)                                                                          \
  do {                                                                                          \
    hipError_t e__ = (call);                                                                    \
    if (e__ != hipSuccess)                                                                      \
      return perr(pr, ACO_HIP_ERR_BACKEND, "%s failed: %s", #call, hipGetErrorString(e__));     \
  } while (0)

#define PLAUNCHED(pr, name)                                                                     \
  do {                                                                                          \
    hipError_t e__ = hipGetLastError();                                                         \
    if (e__ != hipSuccess)                                                                      \
      return perr(pr, ACO_HIP_ERR_BACKEND, "launch of %s failed: %s", name,                     \
                  hipGetErrorString(e__));                                                      \
  } while (0)

static int prior_upload_w(AcoHipPrior *pr, const float *w) {
  std::vector<uint16_t> wb(NBW);
  for (int i = 0; i < 512; ++i) wb[BW1 + i] = f2bf(w[PW1 + i]);
  for (int i = 0; i < 4096; ++i) wb[BW2 + i] = f2bf(w[PW2 + i]);
  PHIP(pr, hipMemcpy(pr->d_w, w, sizeof(float) * NPAR, hipMemcpyHostToDevice));
  PHIP(pr, hipMemcpy(pr->d_wb, wb.data(), sizeof(uint16_t) * NBW, hipMemcpyHostToDevice));
  PHIP(pr, hipMemset(pr->d_m, 0, sizeof(float) * NPAR));
  PHIP(pr, hipMemset(pr->d_v, 0, sizeof(float) * NPAR));
  pr->step = 0;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_prior_create(AcoHipPrior **out, int device, uint32_t seed) {
  if (!out) return ACO_HIP_ERR_INVALID_INPUT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return ACO_HIP_ERR_BACKEND;
  if (device < 0 || device >= count) return ACO_HIP_ERR_INVALID_INPUT;
  AcoHipPrior *pr = new AcoHipPrior();
  pr->device = device;
  int rc = ACO_HIP_OK;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&pr->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&pr->d_w, sizeof(float) * NPAR) != hipSuccess ||
      hipMalloc((void **)&pr->d_m, sizeof(float) * NPAR) != hipSuccess ||
      hipMalloc((void **)&pr->d_v, sizeof(float) * NPAR) != hipSuccess ||
      hipMalloc((void **)&pr->d_gsum, sizeof(float) * PROW) != hipSuccess ||
      hipMalloc((void **)&pr->d_part, sizeof(float) * PROW * MLP_GRID) != hipSuccess ||
      hipMalloc((void **)&pr->d_part2, sizeof(float) * PROW * RED_SPLIT) != hipSuccess ||
      hipMalloc((void **)&pr->d_wb, sizeof(uint16_t) * NBW) != hipSuccess)
    rc = ACO_HIP_ERR_ALLOCATION;
  if (rc == ACO_HIP_OK) {
    /* Glorot-uniform weights from Philox (stream 0x9E3779B9, counter = index),
     * zero biases */
    std::vector<float> w(NPAR, 0.0f);
    auto uni = [&](uint32_t idx) {
      uint32_t r[4];
      aco_philox4x32_10(idx, 0x9E3779B9u, 0u, 0u, seed, ACO_PHILOX_KEY1, r);
      return 2.0f * aco_u01(r[0]) - 1.0f;
    };
    const float a1 = sqrtf(6.0f / (MLP_F + MLP_H)), a2 = sqrtf(6.0f / (2 * MLP_H)),
                a3 = sqrtf(6.0f / (MLP_H + 1));
    for (int i = 0; i < 512; ++i) w[PW1 + i] = a1 * uni((uint32_t)(PW1 + i));
    for (int i = 0; i < 4096; ++i) w[PW2 + i] = a2 * uni((uint32_t)(PW2 + i));
    for (int i = 0; i < 64; ++i) w[PW3 + i] = a3 * uni((uint32_t)(PW3 + i));
    rc = prior_upload_w(pr, w.data());
  }
  if (rc) {
    aco_hip_prior_destroy(pr);
    return rc;
  }
  *out = pr;
  return ACO_HIP_OK;
}

extern "C" void aco_hip_prior_destroy(AcoHipPrior *pr) {
  if (!pr) return;
  (void)hipSetDevice(pr->device);
  if (pr->stream) (void)hipStreamSynchronize(pr->stream);
  void *ptrs[] = {pr->d_w, pr->d_m, pr->d_v, pr->d_part, pr->d_part2, pr->d_gsum, pr->d_wb,
                  pr->d_X, pr->d_y, pr->d_tx, pr->d_tl, pr->d_ty};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (pr->stream) (void)hipStreamDestroy(pr->stream);
  delete pr;
}

extern "C" const char *aco_hip_prior_last_error(const AcoHipPrior *pr) {
  return pr ? pr->err : "null prior";
}

extern "C" int aco_hip_prior_get_params(AcoHipPrior *pr, float *params) {
  if (!pr || !params) return ACO_HIP_ERR_INVALID_INPUT;
  PHIP(pr, hipSetDevice(pr->device));
  PHIP(pr, hipStreamSynchronize(pr->stream));
  PHIP(pr, hipMemcpy(params, pr->d_w, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_prior_set_params(AcoHipPrior *pr, const float *params) {
  if (!pr || !params) return ACO_HIP_ERR_INVALID_INPUT;
  for (int i = 0; i < NPAR; ++i)
    if (!(fabsf(params[i]) < 1e30f))
      return perr(pr, ACO_HIP_ERR_INVALID_INPUT, "parameter %d is not finite", i);
  PHIP(pr, hipSetDevice(pr->device));
  PHIP(pr, hipStreamSynchronize(pr->stream));
  return prior_upload_w(pr, params);
}

/* scratch for `count` edges of one call */
static int prior_scratch(AcoHipPrior *pr, int64_t count) {
  if (count <= pr->tcap) return ACO_HIP_OK;
  if (pr->d_tx) (void)hipFree(pr->d_tx);
  if (pr->d_tl) (void)hipFree(pr->d_tl);
  if (pr->d_ty) (void)hipFree(pr->d_ty);
  pr->d_tx = nullptr;
  pr->d_tl = nullptr;
  pr->d_ty = nullptr;
  pr->tcap = 0;
  const int64_t cap = std::max<int64_t>(count, 1 << 16);
  PHIP(pr, hipMalloc((void **)&pr->d_tx, sizeof(uint16_t) * MLP_F * cap));
  PHIP(pr, hipMalloc((void **)&pr->d_tl, sizeof(float) * cap));
  PHIP(pr, hipMalloc((void **)&pr->d_ty, cap));
  pr->tcap = cap;
  return ACO_HIP_OK;
}

static int prior_launch_forward(AcoHipPrior *pr, hipStream_t st, const uint16_t *X, int64_t E,
                                float *logits) {
  const long tiles = (long)((E + 15) / 16);
  const int grid = (int)std::max<long>(1, std::min<long>(tiles, 4096));
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(64), 0, st, X, (long)E, pr->d_w, pr->d_wb,
                     logits);
  PLAUNCHED(pr, "k_mlp_forward");
  return ACO_HIP_OK;
}

/* gradient rows for (X, Y, E), reduced into d_gsum; lr_t > 0 also steps */
static int prior_launch_grad(AcoHipPrior *pr, const uint16_t *X, const uint8_t *Y, int64_t E,
                             float lr_t, int want_loss) {
  const long tiles = (long)((E + 15) / 16);
  const int rows = (int)std::max<long>(1, std::min<long>((tiles + GRAD_WAVES - 1) / GRAD_WAVES, MLP_GRID));
  hipLaunchKernelGGL(k_mlp_grad, dim3(rows), dim3(64 * GRAD_WAVES), 0, pr->stream, X, Y, (long)E,
                     1.0f / (float)E, pr->d_w, pr->d_wb, pr->d_part, want_loss);
  PLAUNCHED(pr, "k_mlp_grad");
  hipLaunchKernelGGL(k_mlp_reduce_rows, dim3((PROW + 255) / 256, RED_SPLIT), dim3(256), 0,
                     pr->stream, pr->d_part, rows, pr->d_part2);
  PLAUNCHED(pr, "k_mlp_reduce_rows");
  hipLaunchKernelGGL(k_mlp_reduce, dim3((PROW + 255) / 256), dim3(256), 0, pr->stream,
                     pr->d_part2, pr->d_gsum, pr->d_w, pr->d_m, pr->d_v, pr->d_wb, lr_t, 0.9f,
                     0.999f, 1e-8f);
  PLAUNCHED(pr, "k_mlp_reduce");
  return ACO_HIP_OK;
}

extern "C" int aco_hip_prior_forward(AcoHipPrior *pr, const uint16_t *features, int64_t count,
                                     float *logits) {
  if (!pr || !features || !logits || count <= 0) return ACO_HIP_ERR_INVALID_INPUT;
  PHIP(pr, hipSetDevice(pr->device));
  int rc;
  if ((rc = prior_scratch(pr, count))) return rc;
  PHIP(pr, hipMemcpyAsync(pr->d_tx, features, sizeof(uint16_t) * MLP_F * count,
                          hipMemcpyHostToDevice, pr->stream));
  if ((rc = prior_launch_forward(pr, pr->stream, pr->d_tx, count, pr->d_tl))) return rc;
  PHIP(pr, hipMemcpyAsync(logits, pr->d_tl, sizeof(float) * count, hipMemcpyDeviceToHost,
                          pr->stream));
  PHIP(pr, hipStreamSynchronize(pr->stream));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_prior_gradient(AcoHipPrior *pr, const uint16_t *features,
                                      const uint8_t *labels, int64_t count, float *grad,
                                      float *loss) {
  if (!pr || !features || !labels || count <= 0) return ACO_HIP_ERR_INVALID_INPUT;
  PHIP(pr, hipSetDevice(pr->device));
  int rc;
  if ((rc = prior_scratch(pr, count))) return rc;
  PHIP(pr, hipMemcpyAsync(pr->d_tx, features, sizeof(uint16_t) * MLP_F * count,
                          hipMemcpyHostToDevice, pr->stream));
  PHIP(pr, hipMemcpyAsync(pr->d_ty, labels, count, hipMemcpyHostToDevice, pr->stream));
  if ((rc = prior_launch_grad(pr, pr->d_tx, pr->d_ty, count, 0.0f, 1))) return rc;
  std::vector<float> g(PROW);
  PHIP(pr, hipMemcpyAsync(g.data(), pr->d_gsum, sizeof(float) * PROW, hipMemcpyDeviceToHost,
                          pr->stream));
  PHIP(pr, hipStreamSynchronize(pr->stream));
  if (grad) memcpy(grad, g.data(), sizeof(float) * NPAR);
  if (loss) *loss = g[NPAR] / (float)count;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_edge_features(AcoHipContext *c, int instance, uint16_t *features) {
  NEED_LOADED(c);
  if (instance < 0 || instance >= c->num_inst || !features)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "bad instance or null features");
  const InstDesc &d = c->inst[instance];
  const size_t slots = ((size_t)d.n + 1) * d.k;
  uint16_t *dx = nullptr;
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipMalloc((void **)&dx, sizeof(uint16_t) * MLP_F * slots));
  hipLaunchKernelGGL(k_edge_features, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0,
                     c->stream, c->P, instance, dx);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpy(features, dx, sizeof(uint16_t) * MLP_F * slots, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  if (e != hipSuccess)
    return set_err(c, ACO_HIP_ERR_BACKEND, "edge features: %s", hipGetErrorString(e));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_prior_add_instance(AcoHipPrior *pr, AcoHipContext *c, int instance,
                                          const int32_t *tour, int len) {
  if (!pr || !c || !c->loaded || !tour || len < 2) return ACO_HIP_ERR_INVALID_INPUT;
  if (instance < 0 || instance >= c->num_inst)
    return perr(pr, ACO_HIP_ERR_INVALID_INPUT, "instance %d out of range", instance);
  if (c->device != pr->device)
    return perr(pr, ACO_HIP_ERR_INVALID_INPUT, "prior and context are on different devices");
  const InstDesc &d = c->inst[instance];
  const int n1 = d.n + 1, k = d.k;
  const size_t slots = (size_t)n1 * k;
  /* labels: candidate j of node i is positive iff i and j are consecutive in
   * the tour (the depot's neighbours are every route's first and last node) */
  std::vector<int> prv(n1, -1), nxt(n1, -1);
  std::vector<uint8_t> depot_nb(n1, 0);
  for (int p = 0; p + 1 < len; ++p) {
    const int a = tour[p], b = tour[p + 1];
    if (a < 0 || a >= n1 || b < 0 || b >= n1)
      return perr(pr, ACO_HIP_ERR_INVALID_INPUT, "tour entry out of range at %d", p);
    if (a == 0) depot_nb[b] = 1;
    else nxt[a] = b;
    if (b == 0) depot_nb[a] = 1;
    else prv[b] = a;
  }
  std::vector<uint32_t> rec(slots);
  PHIP(pr, hipSetDevice(pr->device));
  PHIP(pr, hipStreamSynchronize(c->stream));
  PHIP(pr, hipMemcpy(rec.data(), c->d_rec + d.cand_off, sizeof(uint32_t) * slots,
                     hipMemcpyDeviceToHost));
  std::vector<uint8_t> y(slots);
  for (int i = 0; i < n1; ++i)
    for (int t = 0; t < k; ++t) {
      const int j = (int)(rec[(size_t)i * k + t] & 0xFFFFu);
      y[(size_t)i * k + t] = (i == 0) ? depot_nb[j] : (uint8_t)(j == prv[i] || j == nxt[i]);
    }
  PHIP(pr, hipStreamSynchronize(pr->stream));
  if (pr->E + (int64_t)slots > pr->cap) {
    const int64_t cap = std::max<int64_t>(2 * pr->cap, pr->E + (int64_t)slots + (1 << 16));
    uint16_t *nx = nullptr;
    uint8_t *ny = nullptr;
    PHIP(pr, hipMalloc((void **)&nx, sizeof(uint16_t) * MLP_F * cap));
    PHIP(pr, hipMalloc((void **)&ny, cap));
    if (pr->E) {
      PHIP(pr, hipMemcpy(nx, pr->d_X, sizeof(uint16_t) * MLP_F * pr->E, hipMemcpyDeviceToDevice));
      PHIP(pr, hipMemcpy(ny, pr->d_y, pr->E, hipMemcpyDeviceToDevice));
    }
    if (pr->d_X) (void)hipFree(pr->d_X);
    if (pr->d_y) (void)hipFree(pr->d_y);
    pr->d_X = nx;
    pr->d_y = ny;
    pr->cap = cap;
  }
  hipLaunchKernelGGL(k_edge_features, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0,
                     c->stream, c->P, instance, pr->d_X + (size_t)pr->E * MLP_F);
  PLAUNCHED(pr, "k_edge_features");
  PHIP(pr, hipStreamSynchronize(c->stream));
  PHIP(pr, hipMemcpy(pr->d_y + pr->E, y.data(), slots, hipMemcpyHostToDevice));
  pr->E += (int64_t)slots;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_prior_clear(AcoHipPrior *pr) {
  if (!pr) return ACO_HIP_ERR_INVALID_INPUT;
  pr->E = 0;
  return ACO_HIP_OK;
}

extern "C" int64_t aco_hip_prior_size(const AcoHipPrior *pr) { return pr ? pr->E : 0; }

extern "C" int aco_hip_prior_train(AcoHipPrior *pr, int steps, float lr, float *loss_out) {
  if (!pr || steps < 0 || !(lr > 0.0f)) return ACO_HIP_ERR_INVALID_INPUT;
  if (pr->E <= 0) return perr(pr, ACO_HIP_ERR_INVALID_INPUT, "empty training set");
  PHIP(pr, hipSetDevice(pr->device));
  int rc;
  for (int s = 0; s < steps; ++s) {
    ++pr->step;
    const double b1t = 1.0 - pow(0.9, pr->step), b2t = 1.0 - pow(0.999, pr->step);
    const float lr_t = (float)(lr * sqrt(b2t) / b1t);
    if ((rc = prior_launch_grad(pr, pr->d_X, pr->d_y, pr->E, lr_t, s == steps - 1))) return rc;
  }
  if (loss_out) {
    float l = 0.0f;
    PHIP(pr, hipMemcpyAsync(&l, pr->d_gsum + NPAR, sizeof(float), hipMemcpyDeviceToHost,
                            pr->stream));
    PHIP(pr, hipStreamSynchronize(pr->stream));
    *loss_out = l / (float)pr->E;
  }
  PHIP(pr, hipStreamSynchronize(pr->stream));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_prior_apply(AcoHipPrior *pr, AcoHipContext *c, int instance, float gain) {
  if (!pr || !c || !c->loaded || !(gain > 0.0f)) return ACO_HIP_ERR_INVALID_INPUT;
  if (instance < 0 || instance >= c->num_inst)
    return perr(pr, ACO_HIP_ERR_INVALID_INPUT, "instance %d out of range", instance);
  if (c->device != pr->device)
    return perr(pr, ACO_HIP_ERR_INVALID_INPUT, "prior and context are on different devices");
  const InstDesc &d = c->inst[instance];
  const int64_t slots = ((int64_t)d.n + 1) * d.k;
  PHIP(pr, hipSetDevice(pr->device));
  PHIP(pr, hipStreamSynchronize(pr->stream));
  int rc;
  if ((rc = prior_scratch(pr, slots))) return rc;
  hipLaunchKernelGGL(k_edge_features, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0,
                     c->stream, c->P, instance, pr->d_tx);
  PLAUNCHED(pr, "k_edge_features");
  if ((rc = prior_launch_forward(pr, c->stream, pr->d_tx, slots, pr->d_tl))) return rc;
  hipLaunchKernelGGL(k_prior_eta, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, c->stream,
                     c->P, instance, pr->d_tl, gain, c->d_ceta);
  PLAUNCHED(pr, "k_prior_eta");
  PHIP(pr, hipStreamSynchronize(c->stream));
  rc = ctx_set_learned(c, instance, 1);
  if (rc) return perr(pr, rc, "%s", aco_hip_last_error(c));
  return ACO_HIP_OK;
}
