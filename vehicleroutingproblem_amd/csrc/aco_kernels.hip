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
 * lower index, slots past n-1 padded with index 0. */
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
        const int2 pj = xy[j];
        const uint32_t d = (uint32_t)aco_nint_dist(pi.x, pi.y, pj.x, pj.y);
        r = (uint32_t)j | (d << 16);
      }
      rec[t] = r;
    }
    last = best;
  }
  if (lane == 0) {
    const int2 p0 = xy[0];
    const int d0 = aco_nint_dist(pi.x, pi.y, p0.x, p0.y);
    P.dscore[I.node_off + i] = depot_w * P.eta[d0];
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
 * whole workgroup. */
__device__ __forceinline__ void wg_copy16(void *dst, const void *src, size_t bytes) {
  const int nvec = (int)(bytes / 16);
  const int nt = (int)blockDim.x;
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  /* eight loads in flight per thread before the LDS stores */
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

/* Weights of the 16 customers j0..j0+15 of one lane (0 when not eligible);
 * every load is issued before any is used. */
__device__ __forceinline__ float fb_lane_weights(float (&w)[FB_V], int j0, int n, int room,
                                                 int2 pc, const int2 *xy, const float *taupow,
                                                 const float *eta, const uint16_t *state,
                                                 const uint8_t *trow) {
  uint32_t st[FB_V], tq[FB_V];
  int2 pj[FB_V];
#pragma unroll
  for (int i = 0; i < FB_V; ++i) {
    const int j = j0 + i;
    const int jc = j <= n ? j : n;
    st[i] = state[jc];
    tq[i] = trow[jc];
    pj[i] = xy[jc];
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < FB_V; ++i) {
    const int d = aco_nint_dist(pc.x, pc.y, pj[i].x, pj[i].y);
    const float v = taupow[tq[i]] * eta[d];
    w[i] = (j0 + i <= n && (int)st[i] <= room) ? v : 0.0f;
    s = s + w[i];
  }
  return s;
}

/* Fallback roulette over every unvisited customer that fits (the exact
 * two-pass restatement of select_global_fallback, aco_cuda_kernels.cu:336-430).
 * Customer j = 1 + c*1024 + lane*16 + i.  Returns the node (or -1), its demand
 * and the total weight. */
__device__ __forceinline__ FbOut fallback_select(const int2 *xy, const float *taupow,
                                                 const float *eta, const uint16_t *state,
                                                 const uint8_t *__restrict__ trow, int n, int cur,
                                                 int room, float u, int lane) {
  const int nch = (n + FB_CHUNK - 1) / FB_CHUNK;
  const int2 pc = xy[cur];
  float w[FB_V];
  float Rv = 0.0f, Cv = 0.0f, R = 0.0f, s = 0.0f, Pl = 0.0f;
  for (int c = 0; c < nch; ++c) {
    s = fb_lane_weights(w, 1 + c * FB_CHUNK + lane * FB_V, n, room, pc, xy, taupow, eta, state,
                        trow);
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
    s = fb_lane_weights(w, 1 + cs * FB_CHUNK + lane * FB_V, n, room, pc, xy, taupow, eta, state,
                        trow);
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
  out.move = 1 + cs * FB_CHUNK + ls * FB_V + isel;
  out.mdem = __builtin_amdgcn_readfirstlane((int)state[out.move]);
  return out;
}

template <bool STAGE>
__global__ void __launch_bounds__(512) k_construct(DevCtx P, int groups_per_colony, int W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int colony = blockIdx.x / groups_per_colony;
  const int group = blockIdx.x - colony * groups_per_colony;
  const ColonyDesc D = P.col[colony];
  const AcoHipColonyState *S = P.state + colony;
  if (S->halted) return; /* uniform over the workgroup */
  const int epoch = S->epoch;
  const InstDesc I = P.inst[D.inst];
  const int n = I.n, k = I.k, K = I.K, Q = I.Q, n1 = n + 1;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;

  /* ---- LDS carve-up and staging ---- */
  const unsigned long long t_kernel = P.stats ? clock64() : 0ull;
  size_t off = 0;
  const uint32_t *c_rec;
  const uint8_t *c_tc;
  const float *c_dsc;
  const int2 *c_xy;
  if (STAGE) {
    uint32_t *s_rec = reinterpret_cast<uint32_t *>(smem + off);
    off += round16((size_t)n1 * k * 4);
    uint8_t *s_tc = smem + off;
    off += round16((size_t)n1 * k);
    float *s_dsc = reinterpret_cast<float *>(smem + off);
    off += round16((size_t)n1 * 4);
    int2 *s_xy = reinterpret_cast<int2 *>(smem + off);
    off += round16((size_t)n1 * 8);
    wg_copy16(s_rec, P.rec + I.cand_off, round16((size_t)n1 * k * 4));
    wg_copy16(s_tc, P.pher + D.tcand_off, round16((size_t)n1 * k));
    wg_copy16(s_dsc, P.dscore + I.node_off, round16((size_t)n1 * 4));
    wg_copy16(s_xy, P.xy + I.node_off, round16((size_t)n1 * 8));
    c_rec = s_rec;
    c_tc = s_tc;
    c_dsc = s_dsc;
    c_xy = s_xy;
  } else {
    c_rec = P.rec + I.cand_off;
    c_tc = P.pher + D.tcand_off;
    c_dsc = P.dscore + I.node_off;
    c_xy = P.xy + I.node_off;
  }
  float *s_taupow = reinterpret_cast<float *>(smem + off);
  off += 256 * 4;
  float *s_eta = reinterpret_cast<float *>(smem + off);
  off += round16((size_t)P.eta_len * 4);
  wg_copy16(s_taupow, P.taupow, 256 * 4);
  wg_copy16(s_eta, P.eta, round16((size_t)P.eta_len * 4));
  const size_t state_stride = round16((size_t)n1 * 2);
  uint16_t *s_state = reinterpret_cast<uint16_t *>(smem + off + (size_t)wave * state_stride);
  __syncthreads();
  if (P.stats && threadIdx.x == 0)
    atomicAdd(P.stats + (size_t)colony * ACO_STATS + 5, clock64() - t_kernel);

  const int ant = group * W + wave;
  if (ant >= D.m) return;

  /* ---- reset (kernel_reset_ant_state, aco_cuda_kernels.cu:227-257): the
   * state of node j is its demand while unvisited, VISITED after; the depot
   * starts visited ---- */
  const uint16_t *g_dem = P.dem + I.node_off;
  for (int j = lane; j < n1; j += 64) s_state[j] = (j == 0) ? (uint16_t)VISITED : g_dem[j];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const uint8_t *__restrict__ tau = P.pher + D.pher_off;
  int32_t *tour_out = P.tours + D.tour_off + (size_t)ant * D.stride;
  const uint32_t key0 = I.seed;
  const uint32_t cid = (uint32_t)D.colony_in_inst;
  const int rows = (k + 15) >> 4;
  const int klane = lane < k ? lane : 0;

  const unsigned long long t_start = P.stats ? clock64() : 0ull;
  unsigned long long fallbacks = 0, fb_cycles = 0;
  int rem_cnt = n;
  long long rem_dem = I.total_demand;
  uint32_t s = 0;
  uint32_t rng0 = 0, rng1 = 0;
  int pos = 0;
  int tb = 0;      /* tour buffer: lane (p & 63) holds entry p of the current block */
  int carry = 0;   /* last entry of the previous block */
  long long cost = 0;

  auto flush = [&](int base, int cnt) {
    const int e = tb;
    int prev = __shfl_up(e, 1, 64);
    if (lane == 0) prev = carry;
    long long dd = 0;
    if (lane < cnt) {
      if (base + lane >= 1) {
        const int2 a = c_xy[prev], b = c_xy[e];
        dd = aco_nint_dist(a.x, a.y, b.x, b.y);
      }
      tour_out[base + lane] = e;
    }
    cost += wave_sum_i64(dd);
    carry = __builtin_amdgcn_readlane(e, cnt - 1);
  };
  auto append = [&](int node) {
    ++pos;
    const int slot = pos & 63;
    if (lane == slot) tb = node;
    if (slot == 63) flush(pos - 63, 64);
  };

  for (int veh = 0; veh < K; ++veh) {
    int cur = 0;
    int load = 0;
    for (;;) {
      if (rem_cnt <= 0 || load >= Q) break; /* aco_cuda_kernels.cu:492 */
      const int room = Q - load;
      if ((s & 63u) == 0u) {
        uint32_t r[4];
        aco_philox4x32_10(s + (uint32_t)lane, (uint32_t)ant, (uint32_t)epoch, cid, key0,
                          ACO_PHILOX_KEY1, r);
        rng0 = r[0];
        rng1 = r[1];
      }
      const uint32_t u0 = (uint32_t)__builtin_amdgcn_readlane((int)rng0, (int)(s & 63u));
      const uint32_t u1 = (uint32_t)__builtin_amdgcn_readlane((int)rng1, (int)(s & 63u));
      ++s;

      /* candidate roulette (select_from_candidates, aco_cuda_kernels.cu:271-322):
       * lanes >= k read slot 0 and are masked out */
      const size_t e = (size_t)cur * k + klane;
      const uint32_t rc = c_rec[e];
      const uint32_t tq = c_tc[e];
      const int j = lane < k ? (int)(rc & 0xFFFFu) : 0;
      const uint32_t st = s_state[j];
      const float v = s_taupow[tq] * s_eta[rc >> 16];
      const float w = ((int)st <= room) ? v : 0.0f;
      const float Pf = wave_scan_rows(w, lane, rows);
      float T = f_readlane(Pf, k - 1);
      int move, mdem = 0;
      if (T > 0.0f) {
        const float thr = aco_u01(u0) * T;
        const uint64_t m = ballot64(w > 0.0f && Pf > thr);
        const int sel = m ? first_lane(m) : last_lane(ballot64(w > 0.0f));
        move = __builtin_amdgcn_readlane(j, sel);
        mdem = __builtin_amdgcn_readlane((int)st, sel);
      } else {
        const unsigned long long f0 = P.stats ? clock64() : 0ull;
        const FbOut fb = fallback_select(c_xy, s_taupow, s_eta, s_state, tau + (size_t)cur * n1,
                                         n, cur, room, aco_u01(u0), lane);
        move = fb.move;
        mdem = fb.mdem;
        T = fb.T;
        if (P.stats) {
          fb_cycles += clock64() - f0;
          ++fallbacks;
        }
      }
      /* depot-close roulette (aco_cuda_kernels.cu:501-520) on remaining demand */
      const bool close_allowed = rem_dem <= (long long)(K - veh - 1) * Q;
      const float dsc =
          close_allowed ? __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(c_dsc[cur])))
                        : 0.0f;
      if (move > 0 && dsc > 0.0f) {
        const float thr2 = aco_u01(u1) * (T + dsc);
        if (thr2 >= T) move = 0;
      }
      if (move <= 0) break;
      append(move);
      if (lane == 0) s_state[move] = (uint16_t)VISITED;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      load += mdem;
      rem_dem -= mdem;
      --rem_cnt;
      cur = move;
    }
    append(0); /* route close (aco_cuda_kernels.cu:547-552) */
  }
  if ((pos & 63) != 63) flush(pos & ~63, (pos & 63) + 1);
  if (lane == 0) {
    const int a = D.ant_off + ant;
    P.len[a] = pos + 1;
    P.cost[a] = cost;
    P.unvis[a] = rem_cnt;
    if (P.stats) {
      unsigned long long *sc = P.stats + (size_t)colony * ACO_STATS;
      atomicAdd(sc + 0, (unsigned long long)s);
      atomicAdd(sc + 1, fallbacks);
      atomicAdd(sc + 2, fb_cycles);
      atomicAdd(sc + 3, clock64() - t_start);
      atomicAdd(sc + 4, 1ull);
    }
  }
}

template __global__ void k_construct<true>(DevCtx, int, int);
template __global__ void k_construct<false>(DevCtx, int, int);

/* ---- best selection and stop rules --------------------------------------- */

/* Iteration best over feasible ants (lowest cost, ties to the lower ant) and
 * the global-best / stagnation bookkeeping of aco_cuda.cu:104-126,331-399,
 * all on device. */
__global__ void __launch_bounds__(256) k_select_best(DevCtx P) {
  __shared__ unsigned long long s_min[4];
  __shared__ int s_cnt[4];
  __shared__ int s_copy;
  const int colony = blockIdx.x;
  AcoHipColonyState *S = P.state + colony;
  const ColonyDesc D = P.col[colony];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (S->halted) {
    if (threadIdx.x == 0) S->found = 0;
    return;
  }
  unsigned long long key = ~0ull;
  int cnt = 0;
  for (int a = threadIdx.x; a < D.m; a += blockDim.x) {
    if (P.unvis[D.ant_off + a] == 0) {
      const unsigned long long kk =
          ((unsigned long long)P.cost[D.ant_off + a] << 20) | (unsigned long long)a;
      key = kk < key ? kk : key;
      ++cnt;
    }
  }
  key = wave_min_u64(key);
  cnt = (int)wave_sum_i64(cnt);
  if (lane == 0) {
    s_min[wave] = key;
    s_cnt[wave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      if (s_min[w] < s_min[0]) s_min[0] = s_min[w];
      s_cnt[0] += s_cnt[w];
    }
    const unsigned long long best = s_min[0];
    S->feasible_ants = s_cnt[0];
    S->improved = 0;
    s_copy = -1;
    if (best != ~0ull) {
      const int ant = (int)(best & 0xFFFFFu);
      const long long c = (long long)(best >> 20);
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
      S->dep_iter = (0.3f * qeff) / (float)c;
      S->dep_best = (0.7f * qeff) / (float)S->global_best_cost;
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
  __syncthreads();
  const int ant = s_copy;
  if (ant >= 0) {
    const int32_t *src = P.tours + D.tour_off + (size_t)ant * D.stride;
    int32_t *dst = P.best + D.best_off;
    for (int p = threadIdx.x; p < D.stride; p += blockDim.x) dst[p] = src[p];
  }
}

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

/* Saturating subtract over the colony's whole pheromone block (matrix and
 * candidate copy): kernel_evaporate_tau (aco_cuda_kernels.cu:145-155). */
__global__ void __launch_bounds__(256) k_evaporate(DevCtx P) {
  const int colony = blockIdx.y;
  if (!P.state[colony].found) return;
  const ColonyDesc D = P.col[colony];
  uint4 *base = reinterpret_cast<uint4 *>(P.pher + D.pher_off);
  const size_t nvec = (size_t)D.pher_bytes / 16;
  const uint32_t d = P.q_evap;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nvec;
       e += (size_t)gridDim.x * blockDim.x) {
    uint4 v = base[e];
    v.x = evap_word(v.x, d);
    v.y = evap_word(v.y, d);
    v.z = evap_word(v.z, d);
    v.w = evap_word(v.w, d);
    base[e] = v;
  }
}

/* ---- deposits -------------------------------------------------------------- */

__device__ __forceinline__ uint32_t dep_f(const DevCtx &P, uint32_t q, float delta) {
  const uint32_t nq = aco_quantize_thr(P.thr, P.deq[q] + delta);
  return nq > q ? nq : q;
}

__device__ __forceinline__ void patch_cand(const uint32_t *rec, uint8_t *tc, int k, int u, int v,
                                           uint8_t val) {
  for (int t = 0; t < k; ++t)
    if ((int)(rec[(size_t)u * k + t] & 0xFFFFu) == v) tc[(size_t)u * k + t] = val;
}

/* Deposit one tour: per edge new = quantize(dequantize(tau[u][v]) + delta),
 * max into tau[u][v] and tau[v][u] (kernel_deposit_solution /
 * kernel_deposit_flat_solution, aco_cuda_kernels.cu:592-668).  One thread per
 * edge; the only entries two edges share are those of a single-customer
 * route 0-c-0 (one thread does both edges) and the (0,0) entry of empty
 * routes (one thread applies them all), so the result equals the sequential
 * edge order.  The candidate-aligned copy is patched alongside. */
__device__ void deposit_tour(const DevCtx &P, const ColonyDesc &D, const InstDesc &I,
                             const int32_t *tour, int len, float delta) {
  __shared__ int s_zz_count, s_zz_first;
  const int n1 = I.n + 1, k = I.k;
  uint8_t *tau = P.pher + D.pher_off;
  uint8_t *tc = P.pher + D.tcand_off;
  const uint32_t *rec = P.rec + I.cand_off;
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
    if (u == 0 && v == 0) {
      if (p == s_zz_first) {
        uint32_t x = tau[0];
        for (int r = 0; r < s_zz_count; ++r) x = dep_f(P, x, delta);
        tau[0] = (uint8_t)x;
        patch_cand(rec, tc, k, 0, 0, (uint8_t)x);
      }
      continue;
    }
    if (u != 0 && v == 0 && p >= 1 && tour[p - 1] == 0) continue; /* done by thread p-1 */
    if (u == 0 && p + 2 < len && tour[p + 2] == 0) {
      /* single-customer route 0 -> v -> 0: both edges in order */
      uint32_t t0v = tau[v], tv0 = tau[(size_t)v * n1];
      uint32_t nq = aco_quantize_thr(P.thr, P.deq[t0v] + delta);
      t0v = nq > t0v ? nq : t0v;
      tv0 = nq > tv0 ? nq : tv0;
      nq = aco_quantize_thr(P.thr, P.deq[tv0] + delta);
      tv0 = nq > tv0 ? nq : tv0;
      t0v = nq > t0v ? nq : t0v;
      tau[v] = (uint8_t)t0v;
      tau[(size_t)v * n1] = (uint8_t)tv0;
      patch_cand(rec, tc, k, 0, v, (uint8_t)t0v);
      patch_cand(rec, tc, k, v, 0, (uint8_t)tv0);
      continue;
    }
    const uint32_t a = tau[(size_t)u * n1 + v];
    const uint32_t nq = aco_quantize_thr(P.thr, P.deq[a] + delta);
    const uint32_t tuv = nq > a ? nq : a;
    const uint32_t b = tau[(size_t)v * n1 + u];
    const uint32_t tvu = nq > b ? nq : b;
    tau[(size_t)u * n1 + v] = (uint8_t)tuv;
    tau[(size_t)v * n1 + u] = (uint8_t)tvu;
    patch_cand(rec, tc, k, u, v, (uint8_t)tuv);
    patch_cand(rec, tc, k, v, u, (uint8_t)tvu);
  }
}

__global__ void __launch_bounds__(1024) k_deposit(DevCtx P, int which) {
  const int colony = blockIdx.x;
  const AcoHipColonyState *S = P.state + colony;
  if (!S->found) return;
  const ColonyDesc D = P.col[colony];
  const InstDesc I = P.inst[D.inst];
  if (which == 0) {
    const int ant = S->best_ant;
    deposit_tour(P, D, I, P.tours + D.tour_off + (size_t)ant * D.stride, P.len[D.ant_off + ant],
                 S->dep_iter);
  } else {
    deposit_tour(P, D, I, P.best + D.best_off, S->best_len, S->dep_best);
  }
}

/* Deposit of a host-provided tour (staged in the scratch tour buffer). */
__global__ void __launch_bounds__(1024) k_deposit_tour(DevCtx P, int colony, const int32_t *tour,
                                                        int len, float delta) {
  const ColonyDesc D = P.col[colony];
  const InstDesc I = P.inst[D.inst];
  deposit_tour(P, D, I, tour, len, delta);
}

/* ---- route-cost reduction ------------------------------------------------- */

/* One wavefront per tour: int64 sum of nint edge lengths (the reference adds
 * calc_dist per move in-kernel, aco_cuda_kernels.cu:540,550). */
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
    const int2 a = xy[tour[p]], b = xy[tour[p + 1]];
    acc += aco_nint_dist(a.x, a.y, b.x, b.y);
  }
  acc = wave_sum_i64(acc);
  if (lane == 0) out[t] = acc;
}
