/*
 * aco_construct.h - the gfx950 tour-construction kernels, as templates.
 *
 * k_construct: one wavefront (64 lanes) builds one ant's tour.  Per decision
 * the lanes hold one candidate each: a wavefront scan of the candidate
 * weights, one 64-bit ballot and one scalar bit-scan pick the next customer.
 * k_construct_pair: two ants per wavefront, one per 32-lane half, for
 * candidate lists of at most 32.  The per-ant visited/demand state and tour
 * buffer live in LDS, the candidate rows in LDS or global memory.
 *
 * FSM selects the fallback roulette's weight source: the all-pairs weight
 * matrix (FSM) or the pheromone row and the taupow / eta tables.  Each
 * choice is instantiated in its own translation unit, so a kernel carries
 * only the registers of its own fallback.
 *
 * Compiled with -ffp-contract=off: every float expression is evaluated as
 * written, in the order the oracle restates.
 */
#ifndef ACO_CONSTRUCT_H
#define ACO_CONSTRUCT_H

#include <hip/hip_runtime.h>

#include <type_traits>

#include "aco_inline.h"
#include "aco_kernels.h"
#include "aco_wave.h"

#define VISITED 0xFFFFu
#define FB_V 16
#define FB_CHUNK 1024

/* ---- tour construction --------------------------------------------------- */

/* Copies `bytes` (multiple of 16, 16-byte aligned on both sides) with the
 * whole workgroup, eight loads in flight per thread. */
__device__ __forceinline__ void wg_copy16(void *dst, const void *src, size_t bytes) {
  const int nvec = (int)(bytes / 16);
  const int nt = (int)blockDim.x;
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  int e = (int)threadIdx.x;
  for (; e + 7 * nt < nvec; e += 8 * nt) {
    uint4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = s[e + u * nt];
#pragma unroll
    for (int u = 0; u < 8; ++u) d[e + u * nt] = v[u];
  }
  for (; e < nvec; e += nt) d[e] = s[e];
}

struct FbOut {
  int move;
  int mdem;
  float T;
};

/* The fallback roulette partitions the nodes 0..n into chunks of 1024:
 * node j = c*1024 + l*16 + i belongs to chunk c, lane l, slot i (the depot is
 * visited from the start, so it never carries weight).  A lane's 16 nodes are
 * contiguous, and so are its fallback weights and states. */

/* Prefixes of the 16 weights of nodes b..b+15 of one lane from the all-pairs
 * weight row: p[i] = w[b] + ... + w[b+i], summed in slot order (a weight is 0
 * when the node is visited, too heavy, or past the last node).  Returns the
 * lane's sum p[15].  Lanes past the last chunk read node 0's group and keep
 * nothing. */
template <class SE>
__device__ __forceinline__ float fb_lane_fs(float (&p)[FB_V], int b0, int n1, int room,
                                            const float *__restrict__ frow, const SE *state) {
  const bool live = b0 < n1;
  const int b = live ? b0 : 0;
  const float4 *f4 = reinterpret_cast<const float4 *>(frow + b);
  const uint4 *s4 = reinterpret_cast<const uint4 *>(state + b);
  const float4 fa = f4[0], fb = f4[1], fc = f4[2], fd = f4[3];
  /* the 16 states: two 16-byte loads of u16, or one of bytes */
  constexpr int SB = 8 * (int)sizeof(SE);    /* bits per state */
  constexpr int SPW = 32 / SB;               /* states per 32-bit word */
  constexpr uint32_t SM = (1u << SB) - 1u;
  uint32_t sw[8];
  {
    const uint4 sa = s4[0];
    sw[0] = sa.x; sw[1] = sa.y; sw[2] = sa.z; sw[3] = sa.w;
    if constexpr (sizeof(SE) == 2) {
      const uint4 sb = s4[1];
      sw[4] = sb.x; sw[5] = sb.y; sw[6] = sb.z; sw[7] = sb.w;
    }
  }
  const float fv[FB_V] = {fa.x, fa.y, fa.z, fa.w, fb.x, fb.y, fb.z, fb.w,
                          fc.x, fc.y, fc.z, fc.w, fd.x, fd.y, fd.z, fd.w};
  int lim = live ? room : -1;
  asm("" : "+v"(lim)); /* one compare per node against the lane's limit */
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < FB_V; ++i) {
    const int st = (int)((sw[i / SPW] >> (SB * (i % SPW))) & SM);
    s = s + (st <= lim ? fv[i] : 0.0f);
    p[i] = s;
  }
  return s;
}

/* The weight of one node j from the all-pairs weight row, as fb_lane_fs
 * computes it, and the node's state. */
template <class SE>
__device__ __forceinline__ float fb_node_fs(int j, int n1, int room, const float *__restrict__ frow,
                                            const SE *state, int &st) {
  const bool live = j < n1;
  const int jc = live ? j : 0;
  st = (int)state[jc];
  const float f = frow[jc];
  return (live && st <= room) ? f : 0.0f;
}

/* Same prefixes with the weights computed from the pheromone row, the
 * coordinates and the taupow / eta tables (used when the all-pairs weight
 * matrix is off). */
template <class SE>
__device__ __forceinline__ float fb_lane_tables(float (&p)[FB_V], int b0, int n1, int room,
                                                int2 pc, const int2 *xy, const float *taupow,
                                                const float *eta, const SE *state,
                                                const uint8_t *__restrict__ trow, bool real) {
  uint32_t st[FB_V], tq[FB_V];
  int2 pj[FB_V];
#pragma unroll
  for (int i = 0; i < FB_V; ++i) {
    const int j = b0 + i;
    const int jc = j < n1 ? j : 0;
    st[i] = state[jc];
    tq[i] = trow[jc];
    pj[i] = xy[jc];
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < FB_V; ++i) {
    const uint32_t d = real ? edge_code_t<true>(pc, pj[i]) : edge_code_t<false>(pc, pj[i]);
    const float v = taupow[tq[i]] * eta[d];
    s = s + ((b0 + i < n1 && (int)st[i] <= room) ? v : 0.0f);
    p[i] = s;
  }
  return s;
}

template <class SE>
__device__ __forceinline__ float fb_node_tables(int j, int n1, int room, int2 pc, const int2 *xy,
                                                const float *taupow, const float *eta,
                                                const SE *state,
                                                const uint8_t *__restrict__ trow, bool real,
                                                int &st) {
  const bool live = j < n1;
  const int jc = live ? j : 0;
  st = (int)state[jc];
  const uint32_t d = real ? edge_code_t<true>(pc, xy[jc]) : edge_code_t<false>(pc, xy[jc]);
  const float v = taupow[trow[jc]] * eta[d];
  return (live && st <= room) ? v : 0.0f;
}

/* Fallback roulette over every unvisited customer that fits (the exact
 * two-pass restatement of select_global_fallback, aco_cuda_kernels.cu:336-430):
 * chunk totals, then the chunk, the lane and the slot holding the threshold.
 * FSM selects the all-pairs weight row frow; otherwise the weights are
 * computed from trow and the tables.  States are u16 or bytes (SE), the
 * visited mark above any room passed.  xs: 64 bytes of LDS private to the
 * wavefront, 16-byte aligned.  Returns the node (or -1), its demand and the
 * total weight. */
template <bool FSM, class SE>
__device__ __forceinline__ FbOut fallback_select(const float *__restrict__ frow, const int2 *xy,
                                              const float *taupow, const float *eta,
                                              const SE *state,
                                              const uint8_t *__restrict__ trow, int n1, int cur,
                                              int room, float u, int lane, bool real, float *xs) {
  const int nch = (n1 + FB_CHUNK - 1) / FB_CHUNK;
  const int2 pc = xy[cur];
  float p[FB_V];
  float Rv = 0.0f, Cv = 0.0f, R = 0.0f, s = 0.0f, Pl = 0.0f;
  for (int c = 0; c < nch; ++c) {
    const int b0 = c * FB_CHUNK + lane * FB_V;
    if constexpr (FSM)
      s = fb_lane_fs(p, b0, n1, room, frow, state);
    else
      s = fb_lane_tables(p, b0, n1, room, pc, xy, taupow, eta, state, trow, real);
    Pl = wave_scan_rows(s, lane, 4);
    const float Cc = f_readlane(Pl, 63);
    R = R + Cc;
    if (lane == c) {
      Rv = R;
      Cv = Cc;
    }
  }
  FbOut out;
  out.T = R;
  out.move = -1;
  out.mdem = 0;
  if (!(R > 0.0f)) return out;
  const float thr = u * R;
  const uint64_t m = ballot64(lane < nch && Rv > thr);
  const int cs = m ? first_lane(m) : last_lane(ballot64(lane < nch && Cv > 0.0f));
  const float base = cs > 0 ? f_readlane(Rv, cs - 1) : 0.0f;
  const float thr_c = thr - base;
  if (cs != nch - 1) {
    const int b0 = cs * FB_CHUNK + lane * FB_V;
    if constexpr (FSM)
      s = fb_lane_fs(p, b0, n1, room, frow, state);
    else
      s = fb_lane_tables(p, b0, n1, room, pc, xy, taupow, eta, state, trow, real);
    Pl = wave_scan_rows(s, lane, 4);
  }
  const uint64_t lm = ballot64(s > 0.0f && Pl > thr_c);
  const int ls = lm ? first_lane(lm) : last_lane(ballot64(s > 0.0f));
  const float pprev = ls > 0 ? f_readlane(Pl, ls - 1) : 0.0f;
  const float thr_l = thr_c - pprev;
  /* The slot of lane ls: the first slot i with w[i] > 0 and p[i] > thr_l,
   * else the last slot with w[i] > 0 (lane ls has a positive sum, so one
   * exists).  Lane ls publishes its 16 prefixes, lane q < 16 recomputes the
   * weight of slot q, and two ballots answer both questions at once. */
  if (lane == ls) {
    float4 *x4 = reinterpret_cast<float4 *>(xs);
    x4[0] = make_float4(p[0], p[1], p[2], p[3]);
    x4[1] = make_float4(p[4], p[5], p[6], p[7]);
    x4[2] = make_float4(p[8], p[9], p[10], p[11]);
    x4[3] = make_float4(p[12], p[13], p[14], p[15]);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int q = lane & (FB_V - 1);
  const int node0 = cs * FB_CHUNK + ls * FB_V;
  int stq;
  float wq;
  if constexpr (FSM)
    wq = fb_node_fs(node0 + q, n1, room, frow, state, stq);
  else
    wq = fb_node_tables(node0 + q, n1, room, pc, xy, taupow, eta, state, trow, real, stq);
  const float pq = xs[q];
  const uint32_t wm = (uint32_t)ballot64(lane < FB_V && wq > 0.0f);
  const uint32_t sm = wm & (uint32_t)ballot64(pq > thr_l);
  const int isel = sm ? (int)__builtin_ctz(sm) : 31 - (int)__builtin_clz(wm | 1u);
  out.move = node0 + isel;
  out.mdem = __builtin_amdgcn_readlane(stq, isel);
  return out;
}

/* 32-bit LDS byte address of an LDS pointer; a u16 store to one. */
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) char *)p;
}
__device__ __forceinline__ void lds_store_u16(uint32_t addr, uint16_t v) {
  *(__attribute__((address_space(3))) uint16_t *)(uintptr_t)addr = v;
}
__device__ __forceinline__ void lds_store_u8(uint32_t addr, uint8_t v) {
  *(__attribute__((address_space(3))) uint8_t *)(uintptr_t)addr = v;
}
/* a * b + c on the low 24 bits of a and b (both operands < 2^24): one
 * v_mad_u32_u24, which the compiler does not pick when it cannot prove a's
 * range itself */
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

/* *(base + byte_off): a 32-bit byte offset from a wave-uniform base, so a
 * global load takes the scalar-base form and an LDS load one address add. */
template <class V>
__device__ __forceinline__ V ld_at(const V *base, uint32_t byte_off) {
  return *reinterpret_cast<const V *>(reinterpret_cast<const char *>(base) + byte_off);
}

/* Inclusive scan of the candidate weights with the row order of
 * wave_scan_rows, specialised for ROWS rows: row 1 takes lane 15 through a
 * row_bcast:15 DPP add (the same single addition as a readlane carry). */
template <int ROWS>
__device__ __forceinline__ float cand_scan(float v, int lane) {
  v = v + f_dpp_row_shr1(v);
  v = v + f_dpp_row_shr2(v);
  v = v + f_dpp_row_shr4(v);
  v = v + f_dpp_row_shr8(v);
  if constexpr (ROWS >= 2) /* row 1: v = v[lane 15] + v, one DPP add (see pair_scan) */
    asm("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0x2 bank_mask:0xf"
        : "+v"(v));
  if constexpr (ROWS >= 3) {
    const float c = f_readlane(v, 31);
    if ((lane >> 4) == 2) v = v + c;
  }
  if constexpr (ROWS >= 4) {
    const float c = f_readlane(v, 47);
    if ((lane >> 4) == 3) v = v + c;
  }
  return v;
}

/* One wavefront per ant.  Per decision every lane reads entry `lane` of the
 * current node's candidate row (lanes >= k read past the row, in bounds by
 * the 64-entry padding, and are masked out of the ballots), then that node's
 * state (its demand, or VISITED); a DPP scan, two ballots and one bit-scan
 * select the move.  All control flow is uniform (scalar).  ROWS = 16-lane rows
 * the widest candidate list spans. */
template <bool STAGE, bool STATS, int ROWS, bool XCD, bool FSM>
__global__ void __launch_bounds__(512) k_construct(DevCtx P, int groups_per_colony, int W,
                                                   const int *cols) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const unsigned long long t_kernel = STATS ? clock64() : 0ull;
  /* cols: the colonies of this launch (one launch per size bucket) */
  /* Workgroups are dispatched to the 8 XCDs round-robin (workgroup i to
   * XCD i mod 8), and each XCD has its own L2.  With XCD set (launches of
   * many colonies) they are renumbered so that XCD x runs the x-th
   * contiguous run of logical workgroups: the ants of a colony (consecutive
   * logical workgroups) stay on one XCD, whose L2 then holds the candidate
   * rows of an eighth of the colonies instead of all of them. */
  unsigned lb = blockIdx.x;
  if constexpr (XCD) {
    const unsigned nxcd = 8;
    const unsigned xq = gridDim.x / nxcd, xr = gridDim.x % nxcd;
    const unsigned xid = blockIdx.x % nxcd, xi = blockIdx.x / nxcd;
    lb = xid * xq + (xid < xr ? xid : xr) + xi;
  }
  const int slot = lb / groups_per_colony;
  const int group = lb - slot * groups_per_colony;
  const int colony = cols ? cols[slot] : slot;
  const ColonyDesc D = P.col[colony];
  const AcoHipColonyState *S = P.state + colony;
  if (S->halted) return; /* uniform over the workgroup */
  const int epoch = S->epoch;
  const InstDesc I = P.inst[D.inst];
  const int n = I.n, k = I.k, K = I.K, Q = I.Q, n1 = n + 1, Sn = I.fs_stride;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); /* wave-uniform */
  constexpr bool fs_on = FSM; /* fallback weights from the all-pairs weight matrix */

  /* ---- LDS carve-up and staging ---- */
  const size_t ent = (size_t)n1 * k + 64; /* candidate entries incl. padding */
  size_t off = 0;
  const uint16_t *c_idx;
  const float *c_score;
  const float *c_dsc;
  const int2 *c_xy;
  if (STAGE) {
    uint16_t *s_idx = reinterpret_cast<uint16_t *>(smem + off);
    off += round16(ent * 2);
    float *s_score = reinterpret_cast<float *>(smem + off);
    off += round16(ent * 4);
    float *s_dsc = reinterpret_cast<float *>(smem + off);
    off += round16((size_t)n1 * 4);
    int2 *s_xy = reinterpret_cast<int2 *>(smem + off);
    off += round16((size_t)n1 * 8);
    wg_copy16(s_idx, P.cidx + I.cand_off, round16(ent * 2));
    wg_copy16(s_score, P.score + D.score_off, round16(ent * 4));
    wg_copy16(s_dsc, P.dscore + I.node_off, round16((size_t)n1 * 4));
    wg_copy16(s_xy, P.xy + I.node_off, round16((size_t)n1 * 8));
    c_idx = s_idx;
    c_score = s_score;
    c_dsc = s_dsc;
    c_xy = s_xy;
  } else {
    c_idx = P.cidx + I.cand_off;
    c_score = P.score + D.score_off;
    c_dsc = P.dscore + I.node_off;
    c_xy = P.xy + I.node_off;
  }
  /* the tables of the computed fallback (only without the weight matrix);
   * large coordinate ranges keep the eta table in global memory */
  const float *c_taupow = P.taupow;
  const float *c_eta = P.eta;
  if constexpr (!fs_on) {
    float *s_taupow = reinterpret_cast<float *>(smem + off);
    off += 256 * 4;
    wg_copy16(s_taupow, P.taupow, 256 * 4);
    c_taupow = s_taupow;
    if (P.eta_lds) {
      float *s_eta = reinterpret_cast<float *>(smem + off);
      off += round16((size_t)P.eta_len * 4);
      wg_copy16(s_eta, P.eta, round16((size_t)P.eta_len * 4));
      c_eta = s_eta;
    }
  }
  uint16_t *s_state = reinterpret_cast<uint16_t *>(smem + off + (size_t)wave * Sn * 2);
  off += (size_t)W * Sn * 2;
  /* tour buffer of this wave: entry p of the tour (u16 node ids) */
  uint16_t *s_tour = reinterpret_cast<uint16_t *>(smem + off + (size_t)wave * round16((size_t)D.stride * 2));
  off += (size_t)W * round16((size_t)D.stride * 2);
  /* 64 u16 per wave that lanes 1..63 write instead of the state and tour
   * entries lane 0 writes, so those stores need no exec mask */
  uint16_t *s_sink = reinterpret_cast<uint16_t *>(smem + off + (size_t)wave * 128);
  __syncthreads();
  if (STATS && threadIdx.x == 0)
    atomicAdd(P.stats + (size_t)colony * ACO_STATS + 5, clock64() - t_kernel);

  const int ant = group * W + wave;
  if (ant >= D.m) return;

  /* ---- reset (kernel_reset_ant_state, aco_cuda_kernels.cu:227-257): the
   * state of node j is its demand while unvisited, VISITED after; the depot
   * and the padding up to S start visited ---- */
  const uint16_t *g_dem = P.dem + I.node_off;
  for (int j = lane; j < Sn; j += 64)
    s_state[j] = (j == 0 || j >= n1) ? (uint16_t)VISITED : g_dem[j];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const uint8_t *__restrict__ tau = P.pher + D.pher_off;
  const float *__restrict__ fsm = fs_on ? P.fscore + D.fs_off : nullptr;
  int32_t *tour_out = P.tours + D.tour_off + (size_t)ant * D.stride;
  const uint32_t key0 = I.seed;
  const uint32_t cid = (uint32_t)D.colony_in_inst;
  const uint64_t kmask = k >= 64 ? ~0ull : ((1ull << k) - 1ull);

  const unsigned long long t_start = STATS ? clock64() : 0ull;
  unsigned long long fallbacks = 0, fb_cycles = 0;
  int rem_cnt = n;
  int rem_dem = (int)I.total_demand;
  uint32_t s = 0;
  float rf0 = 0.0f, rf1 = 0.0f; /* lane l: uniform draws of decision (s & ~63) + l */
  int pos = 0;
  if (lane == 0) s_tour[0] = 0;
  /* lane 0 stores to state[v] and tour[p + 1], the others to their sink */
  const int wscale = lane == 0 ? 1 : 0;
  uint16_t *const st_w = lane == 0 ? s_state : s_sink + lane;
  uint16_t *const tr_w = lane == 0 ? s_tour + 1 : s_sink + lane;
  /* byte offsets of this lane's candidate entry in a row (index, score) */
  uint32_t l2 = 2u * (uint32_t)lane, l4 = 4u * (uint32_t)lane;
  asm("" : "+v"(l2), "+v"(l4)); /* keeps row offset + lane offset one add each */

  for (int veh = 0; veh < K; ++veh) {
    int cur = 0;
    int room = Q; /* Q - load */
    const int close_cap = (K - veh - 1) * Q; /* close allowed iff rem_dem <= close_cap */
    /* the current node's row: lane t holds candidate t (index, score, state)
     * and every lane the node's depot score.  After a move the next row is
     * loaded at once, ahead of the close roulette and the bookkeeping: none
     * of its states can change before its decision, since the only state
     * written meanwhile is the move's own, which is not in its candidate
     * list (lanes past k read other rows and are masked out). */
    uint32_t j = c_idx[lane];
    float sc = c_score[lane];
    float dsc_cur = c_dsc[0];
    uint32_t st = s_state[j];
    /* aco_cuda_kernels.cu:492: the route goes on while rem_cnt > 0 and
     * load < Q; cur = -1 marks a route closed by a roulette (one exit test) */
    while (((rem_cnt - 1) | (room - 1) | cur) >= 0) {
      const int sl = (int)(s & 63u);
      if (sl == 0) {
        uint32_t r[4];
        aco_philox4x32_10(s + (uint32_t)lane, (uint32_t)ant, (uint32_t)epoch, cid, key0,
                          ACO_PHILOX_KEY1, r);
        rf0 = aco_u01(r[0]);
        rf1 = aco_u01(r[1]);
      }
      ++s;
      const float u0 = f_readlane(rf0, sl);

      /* candidate roulette (select_from_candidates, aco_cuda_kernels.cu:271-322).
       * Lanes past the list weigh 0, so the total T is the prefix at the
       * last lane of the last row, which equals that of the instance's own
       * last row when the launch spans more rows than its list */
      const float w = (((int)st <= room) & __builtin_amdgcn_inverse_ballot_w64(kmask)) ? sc : 0.0f;
      const uint64_t emask = ballot64(w > 0.0f);
      const float Pf = cand_scan<ROWS>(w, lane);
      int move, mdem;
      float T;
      if (__builtin_expect(emask != 0, 1)) {
        /* T > 0 exactly when some weight is > 0 (sums of non-negatives) */
        T = f_readlane(Pf, 16 * ROWS - 1);
        const float thr = u0 * T;
        const uint64_t m = ballot64(Pf > thr) & emask;
        const int sel = m ? first_lane(m) : last_lane(emask);
        move = __builtin_amdgcn_readlane((int)j, sel);
        mdem = __builtin_amdgcn_readlane((int)st, sel);
      } else {
        const unsigned long long f0 = STATS ? clock64() : 0ull;
        const FbOut fb = fallback_select<FSM, uint16_t>(fsm ? fsm + (size_t)cur * Sn : nullptr, c_xy, c_taupow,
                                         c_eta, s_state, tau + (size_t)cur * Sn, n1, cur, room,
                                         u0, lane, P.dist_real != 0,
                                         reinterpret_cast<float *>(s_sink));
        move = fb.move;
        mdem = fb.mdem;
        T = fb.T;
        if (STATS) {
          fb_cycles += clock64() - f0;
          ++fallbacks;
        }
      }
      /* prefetch of the next row (row 0 when nothing was found), from the
       * row's wave-uniform base */
      const int mv = move > 0 ? move : 0;
      const uint32_t e2 = (uint32_t)(mv * k);
      const uint32_t j2 = ld_at(c_idx, 2u * e2 + l2);
      const float sc2 = ld_at(c_score, 4u * e2 + l4);
      const float dsc2 = c_dsc[mv];
      /* depot-close roulette (aco_cuda_kernels.cu:501-520) on remaining demand,
       * without branches: dsc_cur is the same in every lane; depot scores
       * are finite or 0 (never NaN), so dsc > 0 is a test of the
       * sign-and-magnitude bits.  A fallback that found nothing (move -1)
       * closes the route whatever this roulette draws. */
      {
        const float thr2 = f_readlane(rf1, sl) * (T + dsc_cur);
        const uint64_t cm = ballot64(__float_as_int(dsc_cur) > 0) & ballot64(thr2 >= T);
        if (rem_dem <= close_cap && cm != 0) move = 0;
      }
      /* the move's bookkeeping, also without branches: with no move (a close,
       * or a fallback that found nothing) lane 0 marks the depot visited
       * again and writes tour[pos + 1] = 0, which the route close writes
       * anyway; the loop then ends on cur = -1 */
      const bool moved = move > 0;
      const int mvw = moved ? move : 0;
      const uint32_t st2 = s_state[j2]; /* read before state[move] changes: j2 != move */
      st_w[mvw * wscale] = (uint16_t)VISITED;
      tr_w[pos * wscale] = (uint16_t)mvw;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int md = moved ? mdem : 0;
      pos += moved ? 1 : 0;
      room -= md;
      rem_dem -= md;
      rem_cnt -= moved ? 1 : 0;
      cur = moved ? move : -1;
      j = j2;
      sc = sc2;
      dsc_cur = dsc2;
      st = st2;
    }
    /* route close (aco_cuda_kernels.cu:547-552) */
    ++pos;
    if (lane == 0) s_tour[pos] = 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  /* write the tour out and sum its edge costs (int64 cost units, exact) */
  long long cost = 0;
  for (int p = lane; p <= pos; p += 64) {
    const int v = s_tour[p];
    tour_out[p] = v;
    if (p >= 1) cost += edge_units(P, c_xy[s_tour[p - 1]], c_xy[v]);
  }
  cost = wave_sum_i64(cost);
  if (lane == 0) {
    const int a = D.ant_off + ant;
    P.len[a] = pos + 1;
    P.cost[a] = cost;
    P.unvis[a] = rem_cnt;
    if (rem_cnt == 0) P.feas[colony] = epoch + 1; /* some ant of this epoch is feasible */
    if (STATS) {
      unsigned long long *sc = P.stats + (size_t)colony * ACO_STATS;
      atomicAdd(sc + 0, (unsigned long long)s);
      atomicAdd(sc + 1, fallbacks);
      atomicAdd(sc + 2, fb_cycles);
      atomicAdd(sc + 3, clock64() - t_start);
      atomicAdd(sc + 4, 1ull);
    }
  }
}


/* ---- tour construction, two ants per wavefront ----------------------------- */

/* The scan of cand_scan<2> in both 32-lane halves at once: row_shr 1, 2, 4,
 * 8 inside each 16-lane row, then rows 1 and 3 add lane 15 of rows 0 and 2
 * (row_bcast:15 with row mask 0xA), so every half holds exactly the prefix
 * sums the one-ant scan gives for candidate lists of up to 32. */
__device__ __forceinline__ float pair_scan(float v) {
  v = v + f_dpp_row_shr1(v);
  v = v + f_dpp_row_shr2(v);
  v = v + f_dpp_row_shr4(v);
  v = v + f_dpp_row_shr8(v);
  /* rows 1 and 3: v = v[lane 15 of the row below] + v; rows 0 and 2 are not
   * written.  One DPP add (the s_nop gives the two wait states a DPP read
   * needs after the VALU write of v). */
  asm("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf"
      : "+v"(v));
  return v;
}

/* Each half's total in all of its 32 lanes: lane 15 of every row to the
 * whole row (row_newbcast:15), then rows 1 and 3 over rows 0 and 2
 * (v_permlane16_swap: its second result holds the odd rows of the first
 * operand in both rows of each pair).  Two VALU operations, no LDS. */
__device__ __forceinline__ float pair_total(float v) {
  const int b = __builtin_amdgcn_mov_dpp(__float_as_int(v), 0x15F, 0xF, 0xF, false);
  return __int_as_float(__builtin_amdgcn_permlane16_swap(b, b, false, false)[1]);
}

/* Two ants per wavefront, for candidate lists of at most 32: lanes 32h+t
 * (h = 0, 1) hold slot t of ant h's current row.  Each ant has its own LDS
 * state array and tour buffer, and its control state (current node, room,
 * customers left, close slack, tour position, route) lives in the lanes of
 * its half, the same value in all 32 of them, so one vector instruction
 * updates both ants and no per-ant branch is taken on the hot path.  Every
 * ant makes the decisions, draws and writes of k_construct exactly (same
 * Philox counters (decision, ant, epoch, colony), same scan order, selection
 * rule, fallback and close rule), so the tours are bit-identical: both ants
 * of a wavefront decide in lockstep, one decision each per iteration, and an
 * ant whose tour is complete idles until its partner's is too.  Ant
 * 2(gW + w) + h of the colony is half h of wavefront w of workgroup g.
 * B8: byte states (demand, 0xFF once visited) for instances whose demands
 * are all at most 254, half the LDS of u16 states; the room a state is
 * compared with is then capped at 254, which keeps every test exact.  With
 * B8 the staged layout leaves the coordinates in global memory (they are
 * read only for the tour cost), so a workgroup of up to 16 wavefronts can
 * share one staged copy of an n=1000 colony's candidate rows. */
template <bool STAGE, bool STATS, bool XCD, bool FSM, bool RING, bool B8>
__global__ void __launch_bounds__(B8 ? 1024 : 512) k_construct_pair(DevCtx P, int groups_per_colony,
                                                                   int W, const int *cols) {
  typedef typename std::conditional<B8, uint8_t, uint16_t>::type SE;
  constexpr uint32_t VIS = B8 ? 0xFFu : (uint32_t)VISITED;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const unsigned long long t_kernel = STATS ? clock64() : 0ull;
  unsigned lb = blockIdx.x;
  if constexpr (XCD) {
    const unsigned nxcd = 8;
    const unsigned xq = gridDim.x / nxcd, xr = gridDim.x % nxcd;
    const unsigned xid = blockIdx.x % nxcd, xi = blockIdx.x / nxcd;
    lb = xid * xq + (xid < xr ? xid : xr) + xi;
  }
  const int slot = lb / groups_per_colony;
  const int group = lb - slot * groups_per_colony;
  const int colony = cols ? cols[slot] : slot;
  const ColonyDesc D = P.col[colony];
  const AcoHipColonyState *S = P.state + colony;
  if (S->halted) return; /* uniform over the workgroup */
  const int epoch = S->epoch;
  const InstDesc I = P.inst[D.inst];
  const int n = I.n, k = I.k, K = I.K, Q = I.Q, n1 = n + 1, Sn = I.fs_stride;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); /* wave-uniform */
  const int h = lane >> 5, t = lane & 31;
  constexpr bool fs_on = FSM; /* fallback weights from the all-pairs weight matrix */
  const bool real = P.dist_real != 0;

  /* ---- LDS: staged rows as k_construct, then per wave two state arrays
   * and two tour buffers ---- */
  const size_t ent = (size_t)n1 * k + 64;
  size_t off = 0;
  const uint16_t *c_idx;
  const float *c_score;
  const float *c_dsc;
  const int2 *c_xy = P.xy + I.node_off;
  if (STAGE) {
    uint16_t *s_idx = reinterpret_cast<uint16_t *>(smem + off);
    off += round16(ent * 2);
    float *s_score = reinterpret_cast<float *>(smem + off);
    off += round16(ent * 4);
    float *s_dsc = reinterpret_cast<float *>(smem + off);
    off += round16((size_t)n1 * 4);
    wg_copy16(s_idx, P.cidx + I.cand_off, round16(ent * 2));
    wg_copy16(s_score, P.score + D.score_off, round16(ent * 4));
    wg_copy16(s_dsc, P.dscore + I.node_off, round16((size_t)n1 * 4));
    if constexpr (!B8) {
      int2 *s_xy = reinterpret_cast<int2 *>(smem + off);
      off += round16((size_t)n1 * 8);
      wg_copy16(s_xy, P.xy + I.node_off, round16((size_t)n1 * 8));
      c_xy = s_xy;
    }
    c_idx = s_idx;
    c_score = s_score;
    c_dsc = s_dsc;
  } else {
    c_idx = P.cidx + I.cand_off;
    c_score = P.score + D.score_off;
    c_dsc = P.dscore + I.node_off;
  }
  const float *c_taupow = P.taupow;
  const float *c_eta = P.eta;
  if constexpr (!fs_on) {
    float *s_taupow = reinterpret_cast<float *>(smem + off);
    off += 256 * 4;
    wg_copy16(s_taupow, P.taupow, 256 * 4);
    c_taupow = s_taupow;
    if (P.eta_lds) {
      float *s_eta = reinterpret_cast<float *>(smem + off);
      off += round16((size_t)P.eta_len * 4);
      wg_copy16(s_eta, P.eta, round16((size_t)P.eta_len * 4));
      c_eta = s_eta;
    }
  }
  SE *const s_state0 = reinterpret_cast<SE *>(smem + off) + (size_t)wave * 2 * Sn;
  off += (size_t)W * 2 * Sn * sizeof(SE);
  /* without RING, two LDS tour buffers per wave */
  const size_t tour_elems = round16((size_t)D.stride * 2) / 2;
  uint16_t *const s_tour0 =
      RING ? nullptr : reinterpret_cast<uint16_t *>(smem + off) + (size_t)wave * 2 * tour_elems;
  if (!RING) off += (size_t)W * 2 * tour_elems * 2;
  /* 64 u16 per wave that the lanes with no entry to write store to, so the
   * per-decision stores need no exec mask */
  uint16_t *const s_sink = reinterpret_cast<uint16_t *>(smem + off) + (size_t)wave * 64;
  __syncthreads();
  if (STATS && threadIdx.x == 0)
    atomicAdd(P.stats + (size_t)colony * ACO_STATS + 5, clock64() - t_kernel);

  const int ant0 = (group * W + wave) * 2;
  if (ant0 >= D.m) return;
  const bool has1 = ant0 + 1 < D.m;
  const int ant_h = ant0 + h;
  /* this lane's ant's state array and tour.  Without RING the tour is
   * built in an LDS buffer and written out at the end.  With RING it goes
   * to its output row in whole 128-byte chunks (rows are 256-byte aligned):
   * lane t of a half holds entry 32c + t of the chunk c being filled (entry
   * 0, the start depot, is 0), and a chunk is stored by its 32 lanes once
   * complete, so no line is written back half-filled and LDS holds only the
   * two state arrays. */
  SE *const st_h = s_state0 + (size_t)h * Sn;
  int32_t *const tour_h = P.tours + D.tour_off + (size_t)ant_h * D.stride;
  uint16_t *const tr_h = RING ? nullptr : s_tour0 + (size_t)h * tour_elems;
  int ring = 0;
  const uint16_t *g_dem = P.dem + I.node_off;
  for (int j = t; j < Sn; j += 32)
    st_h[j] = (j == 0 || j >= n1) ? (SE)VIS : (SE)g_dem[j];
  if (!RING && t == 0) tr_h[0] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const uint8_t *__restrict__ tau = P.pher + D.pher_off;
  const float *__restrict__ fsm = fs_on ? P.fscore + D.fs_off : nullptr;
  const uint32_t key0 = I.seed;
  const uint32_t cid = (uint32_t)D.colony_in_inst;
  /* slots t < k of a half */
  const uint32_t km32 = k >= 32 ? 0xFFFFFFFFu : ((1u << k) - 1u);
  const uint64_t km64 = (uint64_t)km32 | ((uint64_t)km32 << 32);
  /* lanes of the ants still constructing, and their candidate slots per half */
  uint64_t alive_m = has1 ? ~0ull : 0xFFFFFFFFull;
  uint64_t ka = alive_m & km64;
  /* byte address for ds_bpermute: lane 32h */
  int ub = (lane & 32) << 2;
  asm("" : "+v"(ub)); /* keeps ub + 4 sl one add */
  const uint32_t hs = (uint32_t)h << 4;
  const uint32_t k2 = 2u * (uint32_t)k, k4 = 4u * (uint32_t)k;
  const uint32_t t2 = 2u * (uint32_t)t, t4 = 4u * (uint32_t)t;

  /* the lane's ant's state.  pos = moves + routes closed, so the customers
   * left are lim - pos with lim = n + routes closed.  The depot close is
   * allowed iff slack = close_cap - rem_dem >= 0 with close_cap =
   * (K - route - 1) Q (aco_cuda_kernels.cu:501-520).  A move of demand d
   * raises slack by d and lowers room by d, so slack + room is constant
   * within a route: the test is room <= cl with cl = slack + room, which
   * changes only when a route closes (cl -= room, then room = Q). */
  int cur = 0, room = Q, pos = 0, veh = 0, lim = n;
  int cl = K * Q - (int)I.total_demand;
  /* lanes 0 and 32 store their constructing ant's state and tour entries
   * (byte scale sizeof(SE) or 2 per index), the others (and those of a
   * finished ant) to their own sink entry (scale 0): one multiply-add per
   * store address */
  const bool writer = t == 0 && (h == 0 || has1);
  uint32_t wsc = writer ? (uint32_t)sizeof(SE) : 0u;
  uint32_t wsc2 = writer ? 2u : 0u;
  uint32_t st_w = writer ? lds_addr(st_h) : lds_addr(s_sink + lane);
  uint32_t tr_w = lds_addr(writer && !RING ? tr_h + 1 : s_sink + lane);

  const unsigned long long t_start = STATS ? clock64() : 0ull;
  unsigned long long fb_n = 0, fb_cyc = 0, t_done = 0;
  uint32_t dec = 0;
  uint32_t s = 0;
  float rf0 = 0.0f, rf1 = 0.0f; /* half h, lane t: draws of decision (s & ~31) + t of ant h */

  /* the depot rows of both ants.  States are u16 loads; the empty asm
   * makes each an opaque 32-bit value, so the loop-carried state needs no
   * 16-bit mask at its uses */
  uint32_t j = c_idx[t];
  float sc = c_score[t];
  float dsc = c_dsc[0];
  uint32_t st = st_h[j];
  asm("" : "+v"(st));

  while (alive_m != 0ull) {
    const uint32_t sl = s & 31u;
    if (sl == 0) {
      uint32_t r[4];
      aco_philox4x32_10(s + (uint32_t)t, (uint32_t)ant_h, (uint32_t)epoch, cid, key0,
                        ACO_PHILOX_KEY1, r);
      rf0 = aco_u01(r[0]);
      rf1 = aco_u01(r[1]);
    }
    ++s;
    const int ua = ub + (int)(sl << 2);
    const float u0 = __int_as_float(__builtin_amdgcn_ds_bpermute(ua, __float_as_int(rf0)));
    const float u1 = __int_as_float(__builtin_amdgcn_ds_bpermute(ua, __float_as_int(rf1)));

    /* candidate roulettes of both ants (select_from_candidates,
     * aco_cuda_kernels.cu:271-322) */
    /* slots past k weigh 0, so lane 31 of a half holds its total T: the
     * prefix at the last lane of the half's second row */
    const int rl = B8 ? min(room, 254) : room;
    const float w = (((int)st <= rl) & __builtin_amdgcn_inverse_ballot_w64(km64)) ? sc : 0.0f;
    const uint64_t wm = ballot64(w > 0.0f);
    const uint64_t e = wm & ka;
    const float Pf = pair_scan(w);
    float T = pair_total(Pf);
    const uint64_t mm = ballot64(Pf > u0 * T);
    /* each half's first slot above its threshold; a half with none takes
     * its last positive slot (a constructing half whose threshold rounded up
     * to its total), or slot 0 (a half with no weight: the fallback's, or a
     * finished ant's) */
    const uint64_t m = mm & e;
    int sel0 = sff1((uint32_t)m), sel1 = sff1((uint32_t)(m >> 32));
    if (__builtin_expect((sel0 | sel1) < 0, 0)) {
      if (sel0 < 0) sel0 = 31 - __builtin_clz((uint32_t)e | 1u);
      if (sel1 < 0) sel1 = 31 - __builtin_clz((uint32_t)(e >> 32) | 1u);
    }
    /* each half fetches the node and the demand of its own selected lane;
     * node 0 = no move */
    const uint32_t sela = ((uint32_t)sel0 << 2) | ((128u + ((uint32_t)sel1 << 2)) << 16);
    const int pa = (int)__builtin_amdgcn_ubfe(sela, hs, 16u);
    uint32_t mvn = (uint32_t)__builtin_amdgcn_ds_bpermute(pa, (int)j);
    int md = __builtin_amdgcn_ds_bpermute(pa, (int)st);
    /* an ant with no feasible candidate takes the whole-wave fallback.  The
     * weights are >= 0, so a half's T (the sum of its candidates' weights) is
     * > 0 exactly when one of them is > 0: one compare over the wave finds
     * every constructing half with none (or a NaN sum, which the per-half
     * test below settles as before) */
    if (__builtin_expect((ballot64(!(T > 0.0f)) & alive_m) != 0ull, 0)) {
#pragma unroll 1
      for (int q = 0; q < 2; ++q) {
        if ((uint32_t)(e >> (32 * q)) != 0u || (uint32_t)(ka >> (32 * q)) == 0u) continue;
        const unsigned long long f0 = STATS ? clock64() : 0ull;
        const int cq = __builtin_amdgcn_readlane(cur, 32 * q);
        const int rq = __builtin_amdgcn_readlane(rl, 32 * q);
        const float uq = f_readlane(u0, 32 * q);
        const FbOut fb = fallback_select<FSM, SE>(fsm ? fsm + (size_t)cq * Sn : nullptr, c_xy, c_taupow,
                                         c_eta, s_state0 + (size_t)q * Sn,
                                         tau + (size_t)cq * Sn, n1, cq, rq, uq, lane, real,
                                         reinterpret_cast<float *>(s_sink));
        if (h == q) {
          mvn = fb.move > 0 ? (uint32_t)fb.move : 0u;
          md = fb.move > 0 ? fb.mdem : 0;
          T = fb.T;
          if (STATS) {
            fb_cyc += clock64() - f0;
            ++fb_n;
          }
        }
      }
    }
    /* prefetch of both next rows (row 0 when nothing was found) */
    uint32_t j2 = ld_at(c_idx, __umul24(mvn, k2) + t2);
    asm("" : "+v"(j2));
    const float sc2 = ld_at(c_score, __umul24(mvn, k4) + t4);
    const float dsc2 = ld_at(c_dsc, mvn << 2);
    /* depot-close roulettes (aco_cuda_kernels.cu:501-520) on the half's values */
    const float thr2 = u1 * (T + dsc);
    const bool closes = (__float_as_int(dsc) > 0) & (thr2 >= T) & (room <= cl);
    /* the moves' bookkeeping: with no move lane 0 marks the depot visited
     * again and writes tour[pos + 1] = 0, which the route close writes anyway */
    const uint32_t mvw = closes ? 0u : mvn;
    const int mdw = closes ? 0 : md;
    const uint64_t mvm = ballot64(mvw != 0u) & alive_m;
    uint32_t st2 = st_h[j2]; /* before state[move] changes: j2 != move */
    asm("" : "+v"(st2));
    if constexpr (B8)
      lds_store_u8(mad24(mvw, wsc, st_w), (uint8_t)VIS);
    else
      lds_store_u16(mad24(mvw, wsc, st_w), (uint16_t)VIS);
    if constexpr (!RING) lds_store_u16(mad24((uint32_t)pos, wsc2, tr_w), (uint16_t)mvw);
    const int wl = (pos + 1) & 31; /* chunk slot of this decision's tour entry */
    if constexpr (RING)
      if (__builtin_amdgcn_inverse_ballot_w64(alive_m) && t == wl) ring = (int)mvw;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    j = j2;
    sc = sc2;
    dsc = dsc2;
    st = st2;
    room -= mdw;
    /* pos += 1 in the lanes of mvm: one add with the mask as its carry-in */
    asm("v_addc_co_u32_e64 %0, vcc, 0, %0, %1" : "+v"(pos) : "s"(mvm) : "vcc");
    cur = (int)mvw;
    if constexpr (RING) { /* a move that completes a chunk stores it */
      const uint64_t fl = mvm & ballot64(wl == 31);
      if (__builtin_expect(fl != 0ull, 0))
        if (__builtin_amdgcn_inverse_ballot_w64(fl)) tour_h[(pos & ~31) + t] = ring;
    }
    /* aco_cuda_kernels.cu:492: the route goes on while customers remain,
     * load < Q and no roulette closed it */
    const uint64_t cm = alive_m & (~mvm | ballot64(pos == lim) | ballot64(room == 0));
    if (__builtin_expect(cm != 0ull, 0)) {
      /* route close (aco_cuda_kernels.cu:547-552), the next vehicle starts at
       * the depot; with no customer left the remaining routes are empty */
      if (__builtin_amdgcn_inverse_ballot_w64(cm)) {
        const int zeros = pos == lim ? K - veh : 1;
        if constexpr (RING) {
          /* entries pos+1 .. pos+zeros are 0; every chunk they complete is stored */
          const int p0 = pos + 1, p1 = pos + zeros;
          for (int c0 = p0 & ~31;; c0 += 32) {
            const int e = c0 + t;
            if (e >= p0 && e <= p1) ring = 0;
            if (c0 + 31 > p1) break;
            tour_h[e] = ring;
          }
        } else {
          for (int z = t; z < zeros; z += 32) tr_h[pos + 1 + z] = 0;
        }
        pos += zeros;
        veh += zeros;
        lim += zeros;
        if (veh >= K) {
          if constexpr (RING) /* the last, partial chunk */
            if (t <= (pos & 31)) tour_h[(pos & ~31) + t] = ring;
          dec = s;
          if (STATS) t_done = clock64();
        } else {
          cur = 0;
          cl -= room;
          room = Q;
          j = c_idx[t];
          sc = c_score[t];
          dsc = c_dsc[0];
        }
      }
      alive_m &= ~ballot64(veh >= K);
      if (veh >= K) { /* a finished ant's writer lane stores to its sink */
        wsc = 0u;
        wsc2 = 0u;
        st_w = lds_addr(s_sink + lane);
        tr_w = st_w;
      }
      ka = alive_m & km64;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      st = st_h[j];
      asm("" : "+v"(st));
    }
  }
  const int rem_cnt = lim - pos;
  const bool mine = h == 0 || has1;
  long long cost = 0;
  if constexpr (RING) {
    /* the tour stores are complete and visible to the wave's loads (the rows
     * were never read before, so no stale L1 line holds them) */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    /* sum both tours' edge costs (int64 cost units) */
    if (mine)
      for (int p = t + 1; p <= pos; p += 32)
        cost += edge_units(P, c_xy[tour_h[p - 1]], c_xy[tour_h[p]]);
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    /* write both tours out and sum their edge costs (int64 cost units) */
    if (mine)
      for (int p = t; p <= pos; p += 32) {
        const int v = tr_h[p];
        tour_h[p] = v;
        if (p >= 1) cost += edge_units(P, c_xy[tr_h[p - 1]], c_xy[v]);
      }
  }
  for (int o = 16; o > 0; o >>= 1) cost += __shfl_xor(cost, o, 64);
  if (t == 0 && mine) {
    const int a = D.ant_off + ant_h;
    P.len[a] = pos + 1;
    P.cost[a] = cost;
    P.unvis[a] = rem_cnt;
    if (rem_cnt == 0) P.feas[colony] = epoch + 1; /* some ant of this epoch is feasible */
    if (STATS) {
      unsigned long long *sc_ = P.stats + (size_t)colony * ACO_STATS;
      atomicAdd(sc_ + 0, (unsigned long long)dec);
      atomicAdd(sc_ + 1, fb_n);
      atomicAdd(sc_ + 2, fb_cyc);
      atomicAdd(sc_ + 3, t_done - t_start);
      atomicAdd(sc_ + 4, 1ull);
    }
  }
}

#endif /* ACO_CONSTRUCT_H */
