/*
 * aco_kernels.hip - gfx950 kernels of the ACO-CVRP hot path.
 *
 * One wavefront (64 lanes) builds one ant's tour.  Per decision the lanes
 * hold one candidate each: a wavefront scan of the candidate weights, one
 * 64-bit ballot and one scalar bit-scan pick the next customer.  Candidate
 * records, candidate pheromone, heuristic tables and the per-ant
 * visited/demand state live in LDS.  All draws come from Philox4x32-10, all
 * transcendental math from host-built tables, so tours replay bit for bit
 * from (seed, epoch, ant) and equal the CPU oracle's.
 *
 * Compiled with -ffp-contract=off: every float expression is evaluated as
 * written, in the order the oracle restates.
 */
#include <hip/hip_runtime.h>

#include "aco_kernels.h"
#include "aco_inline.h"

#define VISITED 0xFFFFu
#define FB_V 16
#define FB_CHUNK 1024

/* ---- wavefront primitives ------------------------------------------------ */

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
__device__ __forceinline__ int last_lane(uint64_t m) { return 63 - __builtin_clzll(m); }

/* ---- kNN candidate lists ------------------------------------------------- */

/* One wavefront per row i: k rounds, each taking the wave-wide minimum of the
 * unique keys above the previous round's key.  Key = (primary << 24) | j with
 * primary dx^2+dy^2 (exact) or the float bits of (float)sqrt((double)s)
 * (reference OpenMP order).  Follows kernel_build_candidate_lists
 * (aco_cuda_kernels.cu:164-214): customers only, self excluded, ties to the
 * lower index, slots past n-1 padded with index 0.  Each record carries the
 * slot's distance code (the eta-table index) in its upper 16 bits. */
__global__ void __launch_bounds__(256) k_knn(DevCtx P, int total_rows, int knn_key,
                                              float depot_w) {
  const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= total_rows) return;
  const int ii = P.row_inst[row];
  const InstDesc I = P.inst[ii];
  const int i = row - I.row_base;
  const int n = I.n, k = I.k;
  const int2 *xy = P.xy + I.node_off;
  const int2 pi = xy[i];
  uint64_t last = 0;
  uint32_t *rec = P.rec + I.cand_off + (size_t)i * k;
  for (int t = 0; t < k; ++t) {
    uint64_t best = ~0ull;
    for (int j = 1 + lane; j <= n; j += 64) {
      if (j == i) continue;
      const int2 pj = xy[j];
      const int dx = pi.x - pj.x, dy = pi.y - pj.y;
      const uint32_t s = (uint32_t)(dx * dx) + (uint32_t)(dy * dy);
      uint64_t hi;
      if (knn_key == 1) {
        const float fd = (float)__builtin_sqrt((double)s);
        hi = (uint64_t)(uint32_t)__float_as_int(fd);
      } else {
        hi = s;
      }
      const uint64_t key = (hi << 24) | (uint64_t)j;
      if (key > last && key < best) best = key;
    }
    best = wave_min_u64(best);
    if (lane == 0) {
      uint32_t r = 0;
      if (best != ~0ull) {
        const int j = (int)(best & 0xFFFFFFu);
        const uint32_t d = edge_code(P, pi, xy[j]);
        r = (uint32_t)j | (d << 16);
      }
      rec[t] = r;
      P.cidx[I.cand_off + (size_t)i * k + t] = (uint16_t)(r & 0xFFFFu);
    }
    last = best;
  }
  if (lane == 0) {
    const float ds = depot_w * P.eta[edge_code(P, pi, xy[0])];
    P.dscore[I.node_off + i] = ds == ds ? ds : 0.0f; /* NaN scores never close a route */
  }
}

/* slot_of row i: zero, then 1 + t at column cidx[i][t] (first slot wins for a
 * repeated index, which only the padding index 0 can be).  One wavefront per
 * row. */
__global__ void __launch_bounds__(256) k_build_slots(DevCtx P, int total_rows) {
  const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= total_rows) return;
  const InstDesc I = P.inst[P.row_inst[row]];
  const int i = row - I.row_base;
  const int n1 = I.n + 1, k = I.k;
  uint8_t *so = P.slot_of + I.slot_off + (size_t)i * n1;
  for (int j = lane; j < n1; j += 64) so[j] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane == 0) {
    const uint16_t *ci = P.cidx + I.cand_off + (size_t)i * k;
    for (int t = k - 1; t >= 0; --t) so[ci[t]] = (uint8_t)(t + 1);
  }
}

/* ---- pheromone fill ------------------------------------------------------ */

__global__ void __launch_bounds__(256) k_fill_tau(uint8_t *pher, size_t bytes, uint32_t word) {
  const size_t nvec = bytes / 16;
  uint4 v = make_uint4(word, word, word, word);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nvec;
       e += (size_t)gridDim.x * blockDim.x)
    reinterpret_cast<uint4 *>(pher)[e] = v;
}

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

/* Weights of the 16 nodes b..b+15 of one lane from the all-pairs weight row
 * (0 when visited, too heavy, or past the last node).  Lanes past the last
 * chunk read node 0's group and keep nothing. */
__device__ __forceinline__ float fb_lane_fs(float (&w)[FB_V], int b0, int n1, int room,
                                            const float *__restrict__ frow,
                                            const uint16_t *state) {
  const bool live = b0 < n1;
  const int b = live ? b0 : 0;
  const float4 *f4 = reinterpret_cast<const float4 *>(frow + b);
  const uint4 *s4 = reinterpret_cast<const uint4 *>(state + b);
  const float4 fa = f4[0], fb = f4[1], fc = f4[2], fd = f4[3];
  const uint4 sa = s4[0], sb = s4[1];
  const float fv[FB_V] = {fa.x, fa.y, fa.z, fa.w, fb.x, fb.y, fb.z, fb.w,
                          fc.x, fc.y, fc.z, fc.w, fd.x, fd.y, fd.z, fd.w};
  const uint32_t sw[8] = {sa.x, sa.y, sa.z, sa.w, sb.x, sb.y, sb.z, sb.w};
  const int lim = live ? room : -1;
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < FB_V; ++i) {
    const int st = (int)((sw[i >> 1] >> (16 * (i & 1))) & 0xFFFFu);
    w[i] = st <= lim ? fv[i] : 0.0f;
    s = s + w[i];
  }
  return s;
}

/* Same weights computed from the pheromone row, the coordinates and the
 * taupow / eta tables (used when the all-pairs weight matrix is off). */
__device__ __forceinline__ float fb_lane_tables(float (&w)[FB_V], int b0, int n1, int room,
                                                int2 pc, const int2 *xy, const float *taupow,
                                                const float *eta, const uint16_t *state,
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
    w[i] = (b0 + i < n1 && (int)st[i] <= room) ? v : 0.0f;
    s = s + w[i];
  }
  return s;
}

/* Fallback roulette over every unvisited customer that fits (the exact
 * two-pass restatement of select_global_fallback, aco_cuda_kernels.cu:336-430):
 * chunk totals, then the chunk, the lane and the slot holding the threshold.
 * frow != NULL selects the all-pairs weight row; otherwise the weights are
 * computed from trow and the tables.  Returns the node (or -1), its demand and
 * the total weight. */
__device__ __forceinline__ FbOut fallback_select(const float *__restrict__ frow, const int2 *xy,
                                              const float *taupow, const float *eta,
                                              const uint16_t *state,
                                              const uint8_t *__restrict__ trow, int n1, int cur,
                                              int room, float u, int lane, bool real) {
  const int nch = (n1 + FB_CHUNK - 1) / FB_CHUNK;
  const int2 pc = xy[cur];
  float w[FB_V];
  float Rv = 0.0f, Cv = 0.0f, R = 0.0f, s = 0.0f, Pl = 0.0f;
  for (int c = 0; c < nch; ++c) {
    const int b0 = c * FB_CHUNK + lane * FB_V;
    s = frow ? fb_lane_fs(w, b0, n1, room, frow, state)
             : fb_lane_tables(w, b0, n1, room, pc, xy, taupow, eta, state, trow, real);
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
    s = frow ? fb_lane_fs(w, b0, n1, room, frow, state)
             : fb_lane_tables(w, b0, n1, room, pc, xy, taupow, eta, state, trow, real);
    Pl = wave_scan_rows(s, lane, 4);
  }
  const uint64_t lm = ballot64(s > 0.0f && Pl > thr_c);
  const int ls = lm ? first_lane(lm) : last_lane(ballot64(s > 0.0f));
  const float pprev = ls > 0 ? f_readlane(Pl, ls - 1) : 0.0f;
  const float thr_l = thr_c - pprev;
  int is = -1, il = -1;
  float p = 0.0f;
#pragma unroll
  for (int i = 0; i < FB_V; ++i) {
    p = p + w[i];
    if (w[i] > 0.0f) {
      il = i;
      if (is < 0 && p > thr_l) is = i;
    }
  }
  if (is < 0) is = il;
  const int isel = __builtin_amdgcn_readlane(is, ls);
  out.move = cs * FB_CHUNK + ls * FB_V + isel;
  out.mdem = __builtin_amdgcn_readfirstlane((int)state[out.move]);
  return out;
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
  if constexpr (ROWS >= 2)
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0x2, 0xF,
                                                       false));
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
template <bool STAGE, bool STATS, int ROWS, bool XCD>
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
  const int wave = threadIdx.x >> 6;
  const bool fs_on = P.fscore != nullptr;

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
  if (!fs_on) {
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
    int j = (int)c_idx[lane];
    float sc = c_score[lane];
    float dsc_cur = c_dsc[0];
    uint16_t st = s_state[j]; /* 16-bit loop variable: widened at its use */
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

      /* candidate roulette (select_from_candidates, aco_cuda_kernels.cu:271-322) */
      const float w = ((int)st <= room) ? sc : 0.0f;
      const uint64_t emask = ballot64(w > 0.0f) & kmask;
      const float Pf = cand_scan<ROWS>(w, lane);
      int move, mdem;
      float T;
      if (__builtin_expect(emask != 0, 1)) {
        /* T > 0 exactly when some weight is > 0 (sums of non-negatives) */
        T = f_readlane(Pf, k - 1);
        const float thr = u0 * T;
        const uint64_t m = ballot64(Pf > thr) & emask;
        const int sel = m ? first_lane(m) : last_lane(emask);
        move = __builtin_amdgcn_readlane(j, sel);
        mdem = __builtin_amdgcn_readlane((int)st, sel);
      } else {
        const unsigned long long f0 = STATS ? clock64() : 0ull;
        const FbOut fb = fallback_select(fsm ? fsm + (size_t)cur * Sn : nullptr, c_xy, c_taupow,
                                         c_eta, s_state, tau + (size_t)cur * Sn, n1, cur, room,
                                         u0, lane, P.dist_real != 0);
        move = fb.move;
        mdem = fb.mdem;
        T = fb.T;
        if (STATS) {
          fb_cycles += clock64() - f0;
          ++fallbacks;
        }
      }
      /* prefetch of the next row (row 0 when nothing was found) */
      const int mv = move > 0 ? move : 0;
      const int e2 = mv * k + lane;
      const int j2 = (int)c_idx[e2];
      const float sc2 = c_score[e2];
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
      const uint16_t st2 = s_state[j2]; /* read before state[move] changes: j2 != move */
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

#define ACO_CONSTRUCT_VARIANTS(R)                                                          \
  template __global__ void k_construct<true, false, R, false>(DevCtx, int, int, const int *);  \
  template __global__ void k_construct<false, false, R, false>(DevCtx, int, int, const int *); \
  template __global__ void k_construct<true, true, R, false>(DevCtx, int, int, const int *);   \
  template __global__ void k_construct<false, true, R, false>(DevCtx, int, int, const int *);  \
  template __global__ void k_construct<true, false, R, true>(DevCtx, int, int, const int *);   \
  template __global__ void k_construct<false, false, R, true>(DevCtx, int, int, const int *);
ACO_CONSTRUCT_VARIANTS(1)
ACO_CONSTRUCT_VARIANTS(2)
ACO_CONSTRUCT_VARIANTS(3)
ACO_CONSTRUCT_VARIANTS(4)
#undef ACO_CONSTRUCT_VARIANTS

/* ---- evaporation ---------------------------------------------------------- */

__device__ __forceinline__ uint32_t evap_word(uint32_t x, uint32_t d) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    uint32_t q = (x >> (8 * b)) & 0xFFu;
    q = (q > d) ? q - d : 0u; /* q > qmin + delta ? q - delta : qmin, qmin = 0 */
    r |= q << (8 * b);
  }
  return r;
}

/* Items of the all-pairs weight matrix: item it covers row u = it / G and
 * the 16 columns 16g..16g+15 (G = S/16 items per row).  fscore[u][v] =
 * taupow[tau[u][v]] * eta[nint d(u,v)], 0 past the last node.  Pheromone and
 * weight rows share the 16-aligned stride S, so an item is one 16-byte
 * pheromone load (and, with evap, one evaporated 16-byte store: q > delta ?
 * q - delta : 0 per byte) and four 16-byte weight stores.  `taupow` may be an
 * LDS copy of the table.  Coordinates of a group are read as 16-byte pairs;
 * the coordinate pool is padded by 16 entries past the last instance. */
__device__ __forceinline__ void fs_item(const DevCtx &P, const ColonyDesc &D, const InstDesc &I,
                                        const float *taupow, const float *eta, uint32_t it,
                                        bool evap, uint32_t dq) {
  const int n1 = I.n + 1, S = I.fs_stride;
  const uint32_t G = (uint32_t)S >> 4;
  uint8_t *__restrict__ tau = P.pher + D.pher_off;
  float *__restrict__ fs = P.fscore + D.fs_off;
  const int2 *__restrict__ xy = P.xy + I.node_off;
  const uint32_t u = it / G;
  const int v0 = (int)(it - u * G) * 16;
  const size_t row = (size_t)u * S + v0;
  uint4 t = *reinterpret_cast<const uint4 *>(tau + row);
  if (evap) {
    /* levels at 0 stay at 0: an all-zero group keeps its weights */
    if ((t.x | t.y | t.z | t.w) == 0u) return;
    t.x = evap_word(t.x, dq);
    t.y = evap_word(t.y, dq);
    t.z = evap_word(t.z, dq);
    t.w = evap_word(t.w, dq);
    *reinterpret_cast<uint4 *>(tau + row) = t;
  }
  const int2 pu = xy[u];
  const int4 *x4 = reinterpret_cast<const int4 *>(xy + v0);
  int4 q[8];
#pragma unroll
  for (int h = 0; h < 8; ++h) q[h] = x4[h];
  const uint32_t tw[4] = {t.x, t.y, t.z, t.w};
  float o[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int px = (i & 1) ? q[i >> 1].z : q[i >> 1].x;
    const int py = (i & 1) ? q[i >> 1].w : q[i >> 1].y;
    const uint32_t b = (tw[i >> 2] >> (8 * (i & 3))) & 0xFFu;
    const uint32_t d = edge_code(P, pu, make_int2(px, py));
    o[i] = v0 + i < n1 ? taupow[b] * eta[v0 + i < n1 ? d : 0] : 0.0f;
  }
  float4 *f4 = reinterpret_cast<float4 *>(fs + row);
#pragma unroll
  for (int h = 0; h < 4; ++h)
    f4[h] = make_float4(o[4 * h], o[4 * h + 1], o[4 * h + 2], o[4 * h + 3]);
}

/* The evaporating form of fs_item for the flat evaporation loop, on one
 * quad: the 4 columns v0..v0+3 (v0 = 16g + 4r) of row u of item `it`.
 * Consecutive lanes take consecutive quads, so a wavefront's pheromone
 * loads and stores and weight stores are contiguous (256 B and 1 KB per
 * instruction).  The quad's pheromone word t is already loaded; once the
 * coordinate loads are issued, the word at `next` is loaded into t_next for
 * the thread's next quad, so its latency overlaps this quad's arithmetic. */
template <bool REAL>
__device__ __forceinline__ void fs_quad_evap(const DevCtx &P, const ColonyDesc &D,
                                             const InstDesc &I, const float *taupow,
                                             const float *eta, uint32_t it, uint32_t r,
                                             uint32_t t, uint32_t dq, const uint32_t *next,
                                             uint32_t &t_next) {
  const int n1 = I.n + 1, S = I.fs_stride;
  const uint32_t G = (uint32_t)S >> 4;
  uint8_t *__restrict__ tau = P.pher + D.pher_off;
  float *__restrict__ fs = P.fscore + D.fs_off;
  const int2 *__restrict__ xy = P.xy + I.node_off;
  const uint32_t u = it / G;
  const int v0 = (int)(it - u * G) * 16 + 4 * (int)r;
  const size_t row = (size_t)u * S + v0;
  /* levels at 0 stay at 0: an all-zero quad keeps its weights */
  if (t == 0u) {
    t_next = *next;
    return;
  }
  t = evap_word(t, dq);
  *reinterpret_cast<uint32_t *>(tau + row) = t;
  const int2 pu = xy[u];
  const int4 *x4 = reinterpret_cast<const int4 *>(xy + v0);
  const int4 qa = x4[0], qb = x4[1];
  t_next = *next;
  const int px[4] = {qa.x, qa.z, qb.x, qb.z};
  const int py[4] = {qa.y, qa.w, qb.y, qb.w};
  float o[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    /* past the last node the coordinates are the pool's padding (in
     * bounds); the table reads are unconditional, at index 0 there */
    const bool live = v0 + i < n1;
    const uint32_t d = edge_code_t<REAL>(pu, make_int2(px[i], py[i]));
    const uint32_t b = (t >> (8 * i)) & 0xFFu;
    const float w = taupow[b] * eta[live ? d : 0];
    o[i] = live ? w : 0.0f;
  }
  *reinterpret_cast<float4 *>(fs + row) = make_float4(o[0], o[1], o[2], o[3]);
}

/* Items of the all-pairs weight matrix: item it covers row u = it / G and
 * the 16 columns 16g..16g+15 (G = S/16 items per row).  fscore[u][v] =
 * taupow[tau[u][v]] * eta[nint d(u,v)], 0 past the last node.  Pheromone and
 * weight rows share the 16-aligned stride S, so an item is one 16-byte
 * pheromone load (and, with evap, one evaporated 16-byte store: q > delta ?
 * q - delta : 0 per byte) and four 16-byte weight stores.  `taupow` may be an
 * LDS copy of the table.  Coordinates of a group are read as 16-byte pairs;
 * the coordinate pool is padded by 16 entries past the last instance. */
__device__ __forceinline__ void fs_items(const DevCtx &P, const ColonyDesc &D, const InstDesc &I,
                                         const float *taupow, uint32_t first, uint32_t step,
                                         bool evap, uint32_t dq) {
  const uint32_t items = (uint32_t)(I.n + 1) * ((uint32_t)I.fs_stride >> 4);
  for (uint32_t it = first; it < items; it += step) fs_item(P, D, I, taupow, P.eta, it, evap, dq);
}

/* Saturating subtract (kernel_evaporate_tau, aco_cuda_kernels.cu:145-155).
 * Blocks [0, mat_blocks) stream the (n+1)^2 matrix 16 bytes per lane; the
 * remaining blocks evaporate the candidate-aligned copy slot by slot and
 * refresh each slot's score taupow[q] * eta[d] (the per-epoch score refresh
 * of the reference OpenMP path, aco_parallel.c:535-540).
 * from_ants = 0: evaporate iff the last best selection found a feasible ant
 * (the reference order, aco_cuda.cu:355-385).  from_ants = 1 (fused epoch,
 * before the best selection): iff the colony is not halted and some ant of
 * this epoch is feasible - the same condition, read from the ants. */
__global__ void __launch_bounds__(256) k_evaporate(DevCtx P, int mat_blocks, int from_ants) {
  const int colony = blockIdx.y;
  const AcoHipColonyState *S = P.state + colony;
  const ColonyDesc D = P.col[colony];
  if (from_ants) {
    if (S->halted) return;
    int any = 0;
    for (int a = threadIdx.x; a < D.m; a += blockDim.x) any |= P.unvis[D.ant_off + a] == 0;
    if (!__syncthreads_or(any)) return;
  } else if (!S->found) {
    return;
  }
  const uint32_t d = P.q_evap;
  if ((int)blockIdx.x < mat_blocks) {
    if (P.fscore) {
      fs_items(P, D, P.inst[D.inst], P.taupow, blockIdx.x * blockDim.x + threadIdx.x,
               (uint32_t)mat_blocks * blockDim.x, true, d);
      return;
    }
    uint4 *base = reinterpret_cast<uint4 *>(P.pher + D.pher_off);
    const size_t nvec = (size_t)D.mat_bytes / 16;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nvec;
         e += (size_t)mat_blocks * blockDim.x) {
      uint4 v = base[e];
      v.x = evap_word(v.x, d);
      v.y = evap_word(v.y, d);
      v.z = evap_word(v.z, d);
      v.w = evap_word(v.w, d);
      base[e] = v;
    }
    return;
  }
  const InstDesc I = P.inst[D.inst];
  const size_t slots = (size_t)(I.n + 1) * I.k;
  uint8_t *tc = P.pher + D.tcand_off;
  float *score = P.score + D.score_off;
  const uint32_t *rec = P.rec + I.cand_off;
  const size_t nb = gridDim.x - mat_blocks;
  for (size_t e = (blockIdx.x - mat_blocks) * (size_t)blockDim.x + threadIdx.x; e < slots;
       e += nb * blockDim.x) {
    uint32_t q = tc[e];
    q = q > d ? q - d : 0u;
    tc[e] = (uint8_t)q;
    score[e] = P.taupow[q] * slot_eta(P, I, e, rec[e]);
  }
}

/* Largest c in [0, n) with off[c] <= g (off ascending, off[0] = 0). */
__device__ __forceinline__ int colony_of(const long long *off, int n, long long g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

/* Evaporation of every colony of the context in one flat grid: items
 * [0, item_off[ncol]) are the colonies' matrix items (all-pairs weight items
 * of fs_item, or 16-byte pheromone vectors without the weight matrix),
 * items after them the candidate slots.  With cmap, offset entry i is colony
 * cmap[i] (the colonies of one size bucket).  Each workgroup takes a contiguous
 * range of items and its threads stride through it by the workgroup size, so
 * a thread's consecutive items nearly always belong to one colony: the
 * colony (a binary search of the prefix offsets, staged in LDS when there are
 * at most EVAP_LDS_COLONIES) and its descriptors are looked up again only
 * when an item leaves the cached range.  A colony evaporates iff it is not
 * halted and some ant of this epoch was feasible (P.feas, tagged with the
 * epoch by k_construct) - the condition of k_evaporate with from_ants. */
__global__ void __launch_bounds__(512) k_evaporate_flat(DevCtx P, const long long *item_off,
                                                        const long long *slot_off, int ncol,
                                                        int eta_lds, const int *cmap) {
  extern __shared__ __attribute__((aligned(16))) unsigned char dyn_smem[];
  long long *s_ioff = reinterpret_cast<long long *>(dyn_smem);
  long long *s_soff = s_ioff + (ncol + 1);
  const bool off_lds = ncol <= EVAP_LDS_COLONIES;
  float *s_tp = reinterpret_cast<float *>(off_lds ? (unsigned char *)(s_soff + (ncol + 1))
                                                  : dyn_smem);
  float *s_eta = s_tp + 256;
  if (off_lds)
    for (int i = threadIdx.x; i <= ncol; i += blockDim.x) {
      s_ioff[i] = item_off[i];
      s_soff[i] = slot_off[i];
    }
  for (int i = threadIdx.x; i < 256; i += blockDim.x) s_tp[i] = P.taupow[i];
  if (eta_lds)
    for (int i = threadIdx.x; i < P.eta_len; i += blockDim.x) s_eta[i] = P.eta[i];
  __syncthreads();
  const long long *ioff = off_lds ? (const long long *)s_ioff : item_off;
  const long long *soff = off_lds ? (const long long *)s_soff : slot_off;
  const float *eta = eta_lds ? (const float *)s_eta : P.eta;
  const long long ti = ioff[ncol], total = ti + soff[ncol];
  const long long per = ((total + gridDim.x - 1) / gridDim.x + 255) & ~255LL;
  const long long g0 = (long long)blockIdx.x * per;
  const long long g1 = g0 + per < total ? g0 + per : total;
  const uint32_t d = P.q_evap;
  long long lo = 0, hi = 0; /* flat range of the cached colony */
  bool mat = false, ok = false;
  ColonyDesc D;
  InstDesc I;
  /* cache the colony of flat index g: its range [lo, hi), whether it is in
   * the matrix part, whether it evaporates this epoch, its descriptors */
  auto locate = [&](long long g) __attribute__((always_inline)) {
    if (g >= lo && g < hi) return;
    mat = g < ti;
    const long long *off = mat ? ioff : soff;
    const long long base = mat ? 0 : ti;
    const int ci = colony_of(off, ncol, g - base);
    lo = base + off[ci];
    hi = base + off[ci + 1];
    const int c = cmap ? cmap[ci] : ci;
    const AcoHipColonyState *S = P.state + c;
    ok = !S->halted && P.feas[c] == S->epoch + 1;
    D = P.col[c];
    I = P.inst[D.inst];
  };
  long long gs = g0; /* first flat index of the generic loop below */
  if (P.fscore) {
    /* the matrix part of this block's range, quad by quad (4 quads per item) */
    const long long q0 = 4 * (g0 < ti ? g0 : ti), q1 = 4 * (g1 < ti ? g1 : ti);
    uint32_t tn = 0u; /* pheromone word of quad qn, loaded ahead */
    long long qn = -1;
    for (long long qq = q0 + threadIdx.x; qq < q1; qq += blockDim.x) {
      const long long g = qq >> 2;
      locate(g);
      if (!ok) continue;
      const uint32_t *tw = reinterpret_cast<const uint32_t *>(P.pher + D.pher_off);
      const long long qe = qq - 4 * lo; /* quad within the colony */
      uint32_t t;
      if (qq == qn) t = tn;
      else t = tw[qe];
      const long long q2 = qq + blockDim.x;
      const bool pf = q2 < q1 && (q2 >> 2) < hi;
      /* always a load (of this quad again when there is no next one), so
       * that the wait counts are the same on every path */
      const uint32_t *nx = tw + (pf ? qe + blockDim.x : qe);
      const uint32_t it = (uint32_t)(g - lo), r = (uint32_t)(qq & 3);
      /* copies per table location (so that the LDS table is read as LDS)
       * and per distance mode */
      if (P.dist_real) {
        if (eta_lds) fs_quad_evap<true>(P, D, I, s_tp, s_eta, it, r, t, d, nx, tn);
        else fs_quad_evap<true>(P, D, I, s_tp, P.eta, it, r, t, d, nx, tn);
      } else {
        if (eta_lds) fs_quad_evap<false>(P, D, I, s_tp, s_eta, it, r, t, d, nx, tn);
        else fs_quad_evap<false>(P, D, I, s_tp, P.eta, it, r, t, d, nx, tn);
      }
      qn = pf ? q2 : -1;
    }
    gs = g0 > ti ? g0 : ti;
  }
  for (long long g = gs + threadIdx.x; g < g1; g += blockDim.x) {
    locate(g);
    if (!ok) continue;
    const long long e = g - lo;
    if (mat) {
      uint4 *base = reinterpret_cast<uint4 *>(P.pher + D.pher_off);
      uint4 v = base[e];
      v.x = evap_word(v.x, d);
      v.y = evap_word(v.y, d);
      v.z = evap_word(v.z, d);
      v.w = evap_word(v.w, d);
      base[e] = v;
    } else {
      uint8_t *tc = P.pher + D.tcand_off;
      uint32_t q = tc[e];
      q = q > d ? q - d : 0u;
      tc[e] = (uint8_t)q;
      P.score[D.score_off + e] = s_tp[q] * slot_eta(P, I, e, P.rec[I.cand_off + e]);
    }
  }
}

/* score[slot] = taupow[tau_cand[slot]] * eta[d(slot)] for every slot, and
 * the all-pairs weights when enabled (after pheromone init, candidate build
 * or a pheromone upload). */
__global__ void __launch_bounds__(256) k_refresh_scores(DevCtx P) {
  const int colony = blockIdx.y;
  const ColonyDesc D = P.col[colony];
  const InstDesc I = P.inst[D.inst];
  const size_t slots = (size_t)(I.n + 1) * I.k;
  const uint8_t *tc = P.pher + D.tcand_off;
  float *score = P.score + D.score_off;
  const uint32_t *rec = P.rec + I.cand_off;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < slots;
       e += (size_t)gridDim.x * blockDim.x)
    score[e] = P.taupow[tc[e]] * slot_eta(P, I, e, rec[e]);
  if (P.fscore)
    fs_items(P, D, I, P.taupow, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x,
             false, 0u);
}

/* ---- best selection, stop rules and deposits (one workgroup per colony) --- */

/* Tables of the update workgroup, staged in LDS. */
struct UpdTables {
  const float *deq;    /* 256 */
  const float *thr;    /* 255 */
  const float *taupow; /* 256 */
  const float *eta;    /* eta table (global memory, L2-resident) */
};

struct BestAnt {
  long long cost; /* INT64_MAX when no ant is feasible */
  int ant;        /* -1 when no ant is feasible */
  int feasible;   /* feasible ants */
};

/* Iteration best of a colony (select_iter_best_host, aco_cuda.cu:104-126):
 * the lowest cost among feasible ants, ties to the lower ant, as two
 * block-wide minimum reductions (the cost, then the lowest ant at that
 * cost), so costs of any size compare exactly.  `skip` (uniform) reports no
 * feasible ant.  Every thread of the block gets the result. */
__device__ BestAnt block_best_ant(const DevCtx &P, const ColonyDesc &D, bool skip) {
  __shared__ unsigned long long s_red[16];
  __shared__ int s_cnt[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  unsigned long long c = ~0ull;
  int cnt = 0;
  if (!skip)
    for (int a = threadIdx.x; a < D.m; a += blockDim.x)
      if (P.unvis[D.ant_off + a] == 0) {
        const unsigned long long ca = (unsigned long long)P.cost[D.ant_off + a];
        c = ca < c ? ca : c;
        ++cnt;
      }
  c = wave_min_u64(c);
  cnt = (int)wave_sum_i64(cnt);
  if (lane == 0) {
    s_red[wave] = c;
    s_cnt[wave] = cnt;
  }
  __syncthreads();
  unsigned long long best = s_red[0];
  int feas = s_cnt[0];
  for (int w = 1; w < nw; ++w) {
    best = s_red[w] < best ? s_red[w] : best;
    feas += s_cnt[w];
  }
  __syncthreads();
  unsigned long long am = ~0ull;
  if (best != ~0ull)
    for (int a = threadIdx.x; a < D.m; a += blockDim.x)
      if (P.unvis[D.ant_off + a] == 0 && (unsigned long long)P.cost[D.ant_off + a] == best) {
        am = (unsigned long long)a;
        break; /* a thread's ants ascend: its first match is its lowest */
      }
  am = wave_min_u64(am);
  if (lane == 0) s_red[wave] = am;
  __syncthreads();
  unsigned long long ant = s_red[0];
  for (int w = 1; w < nw; ++w) ant = s_red[w] < ant ? s_red[w] : ant;
  __syncthreads();
  BestAnt r;
  r.cost = best == ~0ull ? INT64_MAX : (long long)best;
  r.ant = best == ~0ull ? -1 : (int)ant;
  r.feasible = feas;
  return r;
}

/* Iteration best over feasible ants (lowest cost, ties to the lower ant) and
 * the global-best / stagnation bookkeeping of aco_cuda.cu:104-126,331-399.
 * Returns (to every thread) whether a feasible ant was found. */
__device__ bool select_best_block(const DevCtx &P, int colony) {
  __shared__ int s_copy, s_found;
  AcoHipColonyState *S = P.state + colony;
  const ColonyDesc D = P.col[colony];
  const BestAnt bb = block_best_ant(P, D, false);
  if (threadIdx.x == 0) {
    S->feasible_ants = bb.feasible;
    S->improved = 0;
    s_copy = -1;
    if (bb.ant >= 0) {
      const int ant = bb.ant;
      const long long c = bb.cost;
      S->found = 1;
      S->best_ant = ant;
      S->iter_best_cost = c;
      if (c < S->global_best_cost) {
        const double prev =
            S->global_best_cost == INT64_MAX ? 1.7976931348623157e308 : (double)S->global_best_cost;
        S->improved = 1;
        if (aco_significant(prev, (double)c, P.min_rel)) S->iter_since_best = 0;
        else S->iter_since_best++;
        S->global_best_cost = c;
        S->best_len = P.len[D.ant_off + ant];
        s_copy = ant;
      } else {
        S->iter_since_best++;
      }
      const float qeff = P.qdep * 100.0f;
      S->dep_iter = (0.3f * qeff) / cost_len(P.dist_real, c);
      S->dep_best = (0.7f * qeff) / cost_len(P.dist_real, S->global_best_cost);
    } else {
      S->found = 0;
      S->best_ant = -1;
      S->iter_best_cost = INT64_MAX;
      S->iter_since_best++;
    }
    S->epoch++;
    if (P.fixed_epochs > 0) S->halted = S->epoch >= P.fixed_epochs;
    else if (P.stagnation > 0 && S->iter_since_best >= P.stagnation) S->halted = 1;
    if (P.max_epochs > 0 && S->epoch >= P.max_epochs) S->halted = 1;
    HistEntry *h = reinterpret_cast<HistEntry *>(P.hist) + (size_t)colony * ACO_HIST +
                   (S->epoch - 1) % ACO_HIST;
    h->epoch = S->epoch;
    h->iter_since_best = S->iter_since_best;
    h->improved = S->improved;
    h->halted = S->halted;
    h->global_best = S->global_best_cost;
    s_found = S->found;
  }
  __syncthreads();
  const int ant = s_copy;
  if (ant >= 0) {
    const int32_t *src = P.tours + D.tour_off + (size_t)ant * D.stride;
    int32_t *dst = P.best + D.best_off;
    for (int p = threadIdx.x; p < D.stride; p += blockDim.x) dst[p] = src[p];
  }
  __syncthreads();
  return s_found != 0;
}

__device__ __forceinline__ uint32_t dep_new(const UpdTables &T, uint32_t q, float delta) {
  return aco_quantize_thr(T.thr, T.deq[q] + delta);
}

/* Patch tau_cand and score of row u for candidate v after tau[u][v] became
 * `val`.  slot_of gives the slot in the same memory round trip as the
 * pheromone bytes; the distance code d of (u,v) comes from the coordinates
 * (the kNN stored the same code).  Rows with padding slots (k >= n, index 0 repeated) patch
 * every matching slot by scanning the row. */
__device__ __forceinline__ void patch_slot(const DevCtx &P, const UpdTables &T,
                                           const InstDesc &I, uint8_t *__restrict__ tc,
                                           float *__restrict__ score, int u, int v, int slot1,
                                           int d, uint8_t val) {
  const int k = I.k;
  if (I.k >= I.n) {
    const uint32_t *__restrict__ row = P.rec + I.cand_off + (size_t)u * k;
    for (int t = 0; t < k; ++t) {
      const uint32_t r = row[t];
      if ((int)(r & 0xFFFFu) == v) {
        tc[(size_t)u * k + t] = val;
        score[(size_t)u * k + t] =
            T.taupow[val] * (I.learned ? P.ceta[I.cand_off + (size_t)u * k + t] : T.eta[r >> 16]);
      }
    }
    return;
  }
  if (slot1) {
    const size_t e = (size_t)u * k + slot1 - 1;
    tc[e] = val;
    score[e] = T.taupow[val] * (I.learned ? P.ceta[I.cand_off + e] : T.eta[d]);
  }
}

/* Deposit one tour: per edge new = quantize(dequantize(tau[u][v]) + delta),
 * max into tau[u][v] and tau[v][u] (kernel_deposit_solution /
 * kernel_deposit_flat_solution, aco_cuda_kernels.cu:592-668).  One thread per
 * edge; the only entries two edges share are those of a single-customer
 * route 0-c-0 (one thread does both edges) and the (0,0) entry of empty
 * routes (one thread applies them all), so the result equals the sequential
 * edge order.  The candidate-aligned copy, its scores and the all-pairs
 * weights are patched alongside.  Two dependent memory round trips per edge.  Ends with a
 * workgroup barrier. */
__device__ void deposit_tour(const DevCtx &P, const UpdTables &T, const ColonyDesc &D,
                             const InstDesc &I, const int32_t *tour, int len, float delta) {
  __shared__ int s_zz_count, s_zz_first;
  const int n1 = I.n + 1;
  uint8_t *__restrict__ tau = P.pher + D.pher_off;
  uint8_t *__restrict__ tc = P.pher + D.tcand_off;
  float *__restrict__ score = P.score + D.score_off;
  const uint8_t *__restrict__ so = P.slot_of + I.slot_off;
  const int2 *__restrict__ xy = P.xy + I.node_off;
  float *__restrict__ fs = P.fscore ? P.fscore + D.fs_off : nullptr;
  const int S = I.fs_stride;
  if (threadIdx.x == 0) {
    s_zz_count = 0;
    s_zz_first = 0x7FFFFFFF;
  }
  __syncthreads();
  for (int p = threadIdx.x; p + 1 < len; p += blockDim.x) {
    if (tour[p] == 0 && tour[p + 1] == 0) {
      atomicAdd(&s_zz_count, 1);
      atomicMin(&s_zz_first, p);
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p + 1 < len; p += blockDim.x) {
    const int u = tour[p], v = tour[p + 1];
    const int um1 = p >= 1 ? tour[p - 1] : -1;
    const int vp1 = p + 2 < len ? tour[p + 2] : -1;
    if (u == 0 && v == 0) {
      if (p == s_zz_first) {
        uint32_t x = tau[0];
        for (int r = 0; r < s_zz_count; ++r) {
          const uint32_t nq = dep_new(T, x, delta);
          x = nq > x ? nq : x;
        }
        tau[0] = (uint8_t)x;
        patch_slot(P, T, I, tc, score, 0, 0, so[0], 0, (uint8_t)x);
        if (fs) fs[0] = T.taupow[x] * T.eta[0];
      }
      continue;
    }
    if (u != 0 && v == 0 && um1 == 0) continue; /* done by thread p-1 */
    const uint32_t a = tau[(size_t)u * S + v];
    const uint32_t b = tau[(size_t)v * S + u];
    const int su = so[(size_t)u * n1 + v];
    const int sv = so[(size_t)v * n1 + u];
    const int d = (int)edge_code(P, xy[u], xy[v]);
    uint32_t tuv, tvu;
    if (u == 0 && vp1 == 0) {
      /* single-customer route 0 -> v -> 0: both edges in order */
      uint32_t nq = dep_new(T, a, delta);
      tuv = nq > a ? nq : a;
      tvu = nq > b ? nq : b;
      nq = dep_new(T, tvu, delta);
      tvu = nq > tvu ? nq : tvu;
      tuv = nq > tuv ? nq : tuv;
    } else {
      const uint32_t nq = dep_new(T, a, delta);
      tuv = nq > a ? nq : a;
      tvu = nq > b ? nq : b;
    }
    tau[(size_t)u * S + v] = (uint8_t)tuv;
    tau[(size_t)v * S + u] = (uint8_t)tvu;
    patch_slot(P, T, I, tc, score, u, v, su, d, (uint8_t)tuv);
    patch_slot(P, T, I, tc, score, v, u, sv, d, (uint8_t)tvu);
    if (fs) {
      fs[(size_t)u * S + v] = T.taupow[tuv] * T.eta[d];
      fs[(size_t)v * S + u] = T.taupow[tvu] * T.eta[d];
    }
  }
  __syncthreads();
}

__device__ __forceinline__ UpdTables stage_tables(const DevCtx &P) {
  __shared__ float s_deq[256], s_thr[256], s_tp[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    s_deq[i] = P.deq[i];
    s_thr[i] = i < 255 ? P.thr[i] : 0.0f;
    s_tp[i] = P.taupow[i];
  }
  __syncthreads();
  UpdTables T;
  T.deq = s_deq;
  T.thr = s_thr;
  T.taupow = s_tp;
  T.eta = P.eta;
  return T;
}

/* ---- best selection and both deposits, spread over many workgroups -------- */

/* Events of tour t on the unordered pheromone pair {x, y} (not both 0):
 * returns their count (0..2) and each event's forward node (the event
 * "from -> to" is an edge of the tour, in tour order).  Every customer
 * appears once in a complete tour, so the pair is an edge iff x and y are
 * adjacent; a depot-customer pair can be both edges of a route 0-c-0. */
__device__ __forceinline__ int pair_events(const int32_t *t, int len, const uint16_t *pos, int x,
                                           int y, int (&from)[2]) {
  int n = 0;
  if (x != 0 && y != 0) {
    const int p = pos[x];
    if (p + 1 < len && t[p + 1] == y) from[n++] = x;
    else if (p >= 1 && t[p - 1] == y) from[n++] = y;
    return n;
  }
  const int c = x ? x : y;
  const int p = pos[c];
  if (p >= 1 && t[p - 1] == 0) from[n++] = 0;
  if (p + 1 < len && t[p + 1] == 0) from[n++] = c;
  return n;
}

/* Applies one deposit event to the pair (A = tau[x][y], B = tau[y][x]):
 * the forward entry's new level max-es into both. */
__device__ __forceinline__ void pair_apply(const UpdTables &T, bool fwd_x, float delta, uint32_t &A,
                                           uint32_t &B) {
  const uint32_t nq = dep_new(T, fwd_x ? A : B, delta);
  A = nq > A ? nq : A;
  B = nq > B ? nq : B;
}

/* One epoch's best selection and deposits after k_evaporate(from_ants = 1).
 * Every workgroup repeats the (cheap, deterministic) iteration-best
 * selection, stages both tours and the position of every customer in each
 * (LDS), and owns a share of the unordered pheromone pairs {u, v} the two
 * tours touch: for its pair it applies tour 1's events (iteration best,
 * delta1) then tour 2's (global best, delta2) - per pair exactly the result
 * of the sequential deposits of kernel_deposit_solution, since pairs share
 * no entries.  The empty-route (0,0) entry is one thread's.  The last
 * workgroup to finish (device counter) commits the colony state, the
 * global-best tour and the stop rules (aco_cuda.cu:104-126,331-399). */
__global__ void __launch_bounds__(256) k_select_deposit(DevCtx P, int stage_eta,
                                                        const int *cmap) {
  extern __shared__ __attribute__((aligned(16))) unsigned char dyn_smem[];
  __shared__ int s_last, s_zz[2];
  __shared__ float s_deq[256], s_thr[256], s_tp[256];
  /* grid row y: colony y, or colony cmap[y] of a size bucket's list */
  const int colony = cmap ? cmap[blockIdx.y] : (int)blockIdx.y;
  AcoHipColonyState *S = P.state + colony;
  const ColonyDesc D = P.col[colony];
  const InstDesc I = P.inst[D.inst];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nt = (int)blockDim.x;
  const bool halted = S->halted != 0;
  const long long gbest = S->global_best_cost;
  const int best_len0 = S->best_len;

  /* iteration best: feasible ants only, lowest cost, ties to the lower ant */
  for (int i = threadIdx.x; i < 256; i += nt) {
    s_deq[i] = P.deq[i];
    s_thr[i] = i < 255 ? P.thr[i] : 0.0f;
    s_tp[i] = P.taupow[i];
  }
  if (threadIdx.x < 2) s_zz[threadIdx.x] = 0;
  const BestAnt bb = block_best_ant(P, D, halted); /* its barriers order the stores above */
  const int feasible = bb.feasible;
  const bool found = bb.ant >= 0;
  const int ant = bb.ant;
  const long long c = bb.cost;
  const bool improved = found && c < gbest;
  const long long gnew = improved ? c : gbest;
  const float qeff = P.qdep * 100.0f;
  const float delta1 = found ? (0.3f * qeff) / cost_len(P.dist_real, c) : 0.0f;
  const float delta2 = found ? (0.7f * qeff) / cost_len(P.dist_real, gnew) : 0.0f;
  const int32_t *g1 = found ? P.tours + D.tour_off + (size_t)ant * D.stride : nullptr;
  const int len1 = found ? P.len[D.ant_off + ant] : 0;
  const int32_t *g2 = improved ? g1 : P.best + D.best_off;
  const int len2 = improved ? len1 : best_len0;

  if (found) {
    const int n1 = I.n + 1, Sn = I.fs_stride, stride = D.stride;
    int32_t *t1 = reinterpret_cast<int32_t *>(dyn_smem);
    int32_t *t2 = t1 + stride;
    uint16_t *pos1 = reinterpret_cast<uint16_t *>(t2 + stride);
    uint16_t *pos2 = pos1 + ((n1 + 7) & ~7);
    UpdTables T;
    T.deq = s_deq;
    T.thr = s_thr;
    T.taupow = s_tp;
    T.eta = P.eta;
    if (stage_eta) {
      float *s_eta = reinterpret_cast<float *>(pos2 + ((n1 + 7) & ~7));
      for (int i = threadIdx.x; i < P.eta_len; i += nt) s_eta[i] = P.eta[i];
      T.eta = s_eta;
    }
    for (int p = threadIdx.x; p < len1; p += nt) {
      const int v = g1[p];
      t1[p] = v;
      if (v != 0) pos1[v] = (uint16_t)p;
    }
    for (int p = threadIdx.x; p < len2; p += nt) {
      const int v = g2[p];
      t2[p] = v;
      if (v != 0) pos2[v] = (uint16_t)p;
    }
    __syncthreads();
    uint8_t *__restrict__ tau = P.pher + D.pher_off;
    uint8_t *__restrict__ tc = P.pher + D.tcand_off;
    float *__restrict__ score = P.score + D.score_off;
    float *__restrict__ fs = P.fscore ? P.fscore + D.fs_off : nullptr;
    const uint8_t *__restrict__ so = P.slot_of + I.slot_off;
    const int2 *__restrict__ xy = P.xy + I.node_off;
    const int e1 = len1 - 1, e2 = len2 - 1;
    /* the (0,0) pair: empty routes of either tour, one thread */
    if (blockIdx.x == 0) {
      for (int p = threadIdx.x; p < e1; p += nt)
        if (t1[p] == 0 && t1[p + 1] == 0) atomicAdd(&s_zz[0], 1);
      for (int p = threadIdx.x; p < e2; p += nt)
        if (t2[p] == 0 && t2[p + 1] == 0) atomicAdd(&s_zz[1], 1);
      __syncthreads();
      if (threadIdx.x == 0 && (s_zz[0] | s_zz[1])) {
        uint32_t x = tau[0];
        for (int t = 0; t < 2; ++t)
          for (int r = 0; r < s_zz[t]; ++r) {
            const uint32_t nq = dep_new(T, x, t ? delta2 : delta1);
            x = nq > x ? nq : x;
          }
        tau[0] = (uint8_t)x;
        patch_slot(P, T, I, tc, score, 0, 0, so[0], 0, (uint8_t)x);
        if (fs) fs[0] = T.taupow[x] * T.eta[0];
      }
    }
    for (int it = blockIdx.x * nt + threadIdx.x; it < e1 + e2; it += gridDim.x * nt) {
      const bool second = it >= e1;
      const int p = second ? it - e1 : it;
      const int32_t *tt = second ? t2 : t1;
      const int x = tt[p], y = tt[p + 1];
      if (x == 0 && y == 0) continue;
      if (y == 0 && p >= 1 && tt[p - 1] == 0) continue; /* pair {0,x} done at p-1 */
      int f1[2], f2[2];
      const int n1e = pair_events(t1, len1, pos1, x, y, f1);
      if (second && n1e) continue; /* the tour-1 item of this pair applies both */
      const int n2e = pair_events(t2, len2, pos2, x, y, f2);
      const uint32_t xy_ = (uint32_t)x * (uint32_t)Sn + (uint32_t)y;
      const uint32_t yx_ = (uint32_t)y * (uint32_t)Sn + (uint32_t)x;
      uint32_t A = tau[xy_], B = tau[yx_];
      const int sxy = so[(uint32_t)x * (uint32_t)n1 + (uint32_t)y];
      const int syx = so[(uint32_t)y * (uint32_t)n1 + (uint32_t)x];
      const int2 px = xy[x], py = xy[y];
      for (int e = 0; e < n1e; ++e) pair_apply(T, f1[e] == x, delta1, A, B);
      for (int e = 0; e < n2e; ++e) pair_apply(T, f2[e] == x, delta2, A, B);
      const int d = (int)edge_code(P, px, py);
      tau[xy_] = (uint8_t)A;
      tau[yx_] = (uint8_t)B;
      patch_slot(P, T, I, tc, score, x, y, sxy, d, (uint8_t)A);
      patch_slot(P, T, I, tc, score, y, x, syx, d, (uint8_t)B);
      if (fs) {
        fs[xy_] = T.taupow[A] * T.eta[d];
        fs[yx_] = T.taupow[B] * T.eta[d];
      }
    }
  }

  /* the last workgroup commits the colony state */
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int prev = atomicAdd(P.done + colony, 1);
    s_last = prev == (int)gridDim.x - 1;
    if (s_last) {
      __threadfence();
      P.done[colony] = 0;
    }
  }
  __syncthreads();
  if (!s_last) return;
  if (halted) {
    if (threadIdx.x == 0) S->found = 0;
    return;
  }
  if (threadIdx.x == 0) {
    S->feasible_ants = feasible;
    S->improved = improved ? 1 : 0;
    if (found) {
      S->found = 1;
      S->best_ant = ant;
      S->iter_best_cost = c;
      if (improved) {
        const double prev =
            gbest == INT64_MAX ? 1.7976931348623157e308 : (double)gbest;
        if (aco_significant(prev, (double)c, P.min_rel)) S->iter_since_best = 0;
        else S->iter_since_best++;
        S->global_best_cost = c;
        S->best_len = len1;
      } else {
        S->iter_since_best++;
      }
      S->dep_iter = delta1;
      S->dep_best = delta2;
    } else {
      S->found = 0;
      S->best_ant = -1;
      S->iter_best_cost = INT64_MAX;
      S->iter_since_best++;
    }
    S->epoch++;
    if (P.fixed_epochs > 0) S->halted = S->epoch >= P.fixed_epochs;
    else if (P.stagnation > 0 && S->iter_since_best >= P.stagnation) S->halted = 1;
    if (P.max_epochs > 0 && S->epoch >= P.max_epochs) S->halted = 1;
    HistEntry *h = reinterpret_cast<HistEntry *>(P.hist) + (size_t)colony * ACO_HIST +
                   (S->epoch - 1) % ACO_HIST;
    h->epoch = S->epoch;
    h->iter_since_best = S->iter_since_best;
    h->improved = S->improved;
    h->halted = S->halted;
    h->global_best = S->global_best_cost;
  }
  if (improved) {
    int32_t *dst = P.best + D.best_off;
    for (int p = threadIdx.x; p < D.stride; p += nt) dst[p] = g1[p];
  }
}

/* mode bit 1: best selection and stop rules; bit 2: iteration-best deposit;
 * bit 4: global-best deposit.  Deposits run only when the last best
 * selection found a feasible ant (aco_cuda.cu:355-396).  Global writes of one
 * phase are visible to the next through the workgroup barriers. */
__global__ void __launch_bounds__(1024) k_update(DevCtx P, int mode) {
  const int colony = blockIdx.x;
  AcoHipColonyState *S = P.state + colony;
  if (mode & 1) {
    if (S->halted) {
      if (threadIdx.x == 0) S->found = 0;
      return;
    }
    if (!select_best_block(P, colony)) return;
  } else if (!S->found) {
    return;
  }
  if (!(mode & 6)) return;
  const ColonyDesc D = P.col[colony];
  const InstDesc I = P.inst[D.inst];
  const UpdTables T = stage_tables(P);
  if (mode & 2) {
    const int ant = S->best_ant;
    deposit_tour(P, T, D, I, P.tours + D.tour_off + (size_t)ant * D.stride,
                 P.len[D.ant_off + ant], S->dep_iter);
  }
  if (mode & 4) deposit_tour(P, T, D, I, P.best + D.best_off, S->best_len, S->dep_best);
}

/* One epoch's pheromone update in one launch: every block evaporates its
 * share (deciding feasibility from the ants, as k_evaporate with from_ants),
 * then the last block to finish - found with a device-wide counter, no
 * waiting - runs the best selection and both deposits (k_update mode 7).
 * __threadfence() before the counter increment and after reading it makes
 * every block's evaporation visible to that last block. */
__global__ void __launch_bounds__(1024) k_evap_update(DevCtx P, int mat_blocks) {
  __shared__ int s_last;
  __shared__ float s_tp[256];
  const int colony = blockIdx.y;
  AcoHipColonyState *S = P.state + colony;
  const ColonyDesc D = P.col[colony];
  const bool halted = S->halted != 0;
  unsigned long long *stc = P.stats ? P.stats + (size_t)colony * ACO_STATS : nullptr;
  if (stc && threadIdx.x == 0) atomicMax(stc + 10, ~wall_clock64());
  int any = 0;
  if (!halted)
    for (int a = threadIdx.x; a < D.m; a += blockDim.x) any |= P.unvis[D.ant_off + a] == 0;
  const bool run = !halted && __syncthreads_or(any);
  if (run) {
    const uint32_t d = P.q_evap;
    if ((int)blockIdx.x < mat_blocks && P.fscore) {
      for (int i = threadIdx.x; i < 256; i += blockDim.x) s_tp[i] = P.taupow[i];
      __syncthreads();
      fs_items(P, D, P.inst[D.inst], s_tp, blockIdx.x * blockDim.x + threadIdx.x,
               (uint32_t)mat_blocks * blockDim.x, true, d);
    } else if ((int)blockIdx.x < mat_blocks) {
      uint4 *base = reinterpret_cast<uint4 *>(P.pher + D.pher_off);
      const size_t nvec = (size_t)D.mat_bytes / 16;
      for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nvec;
           e += (size_t)mat_blocks * blockDim.x) {
        uint4 v = base[e];
        v.x = evap_word(v.x, d);
        v.y = evap_word(v.y, d);
        v.z = evap_word(v.z, d);
        v.w = evap_word(v.w, d);
        base[e] = v;
      }
    } else {
      const InstDesc I = P.inst[D.inst];
      const size_t slots = (size_t)(I.n + 1) * I.k;
      uint8_t *tc = P.pher + D.tcand_off;
      float *score = P.score + D.score_off;
      const uint32_t *rec = P.rec + I.cand_off;
      const size_t nb = gridDim.x - mat_blocks;
      for (size_t e = (blockIdx.x - mat_blocks) * (size_t)blockDim.x + threadIdx.x; e < slots;
           e += nb * blockDim.x) {
        uint32_t q = tc[e];
        q = q > d ? q - d : 0u;
        tc[e] = (uint8_t)q;
        score[e] = P.taupow[q] * slot_eta(P, I, e, rec[e]);
      }
    }
  }
  /* the barrier orders every wave's stores before thread 0's device-scope
   * release; one fence per block (each fence writes back the L2) */
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int prev = atomicAdd(P.done + colony, 1);
    s_last = prev == (int)gridDim.x - 1;
    if (s_last) {
      __threadfence(); /* acquire: the other blocks' evaporation */
      P.done[colony] = 0; /* ready for the next epoch */
    }
  }
  __syncthreads();
  if (!s_last) return;
  const unsigned long long t0 = stc ? wall_clock64() : 0ull;
  if (stc && threadIdx.x == 0) {
    const unsigned long long first = ~stc[10];
    stc[10] = 0ull;
    stc[7] += t0 - first;
  }
  if (halted) {
    if (threadIdx.x == 0) S->found = 0;
    return;
  }
  const bool found = select_best_block(P, colony);
  const unsigned long long t1 = stc ? wall_clock64() : 0ull;
  if (stc && threadIdx.x == 0) {
    stc[6] += 1;
    stc[8] += t1 - t0;
  }
  if (!found) return;
  const InstDesc I = P.inst[D.inst];
  const UpdTables T = stage_tables(P);
  const int ant = S->best_ant;
  deposit_tour(P, T, D, I, P.tours + D.tour_off + (size_t)ant * D.stride, P.len[D.ant_off + ant],
               S->dep_iter);
  deposit_tour(P, T, D, I, P.best + D.best_off, S->best_len, S->dep_best);
  if (stc && threadIdx.x == 0) stc[9] += wall_clock64() - t1;
}

/* Deposit of a host-provided tour (staged in the scratch tour buffer). */
__global__ void __launch_bounds__(1024) k_deposit_tour(DevCtx P, int colony, const int32_t *tour,
                                                        int len, float delta) {
  const ColonyDesc D = P.col[colony];
  const InstDesc I = P.inst[D.inst];
  const UpdTables T = stage_tables(P);
  deposit_tour(P, T, D, I, tour, len, delta);
}

/* ---- route-cost reduction ------------------------------------------------- */

/* One wavefront per tour: int64 sum of its edge costs in cost units (the
 * reference adds calc_dist per move in-kernel, aco_cuda_kernels.cu:540,550). */
__global__ void __launch_bounds__(256) k_route_costs(DevCtx P, int instance, const int32_t *tours,
                                                      const int32_t *lens, int count, int stride,
                                                      long long *out) {
  const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= count) return;
  const InstDesc I = P.inst[instance];
  const int2 *xy = P.xy + I.node_off;
  const int32_t *tour = tours + (size_t)t * stride;
  const int len = lens[t];
  long long acc = 0;
  for (int p = lane; p + 1 < len; p += 64) {
    acc += edge_units(P, xy[tour[p]], xy[tour[p + 1]]);
  }
  acc = wave_sum_i64(acc);
  if (lane == 0) out[t] = acc;
}
