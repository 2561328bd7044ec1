/*
 * aco_kernels.hip - gfx950 kernels of the ACO-CVRP hot path around tour
 * construction (aco_construct.h): kNN candidate lists, pheromone fill,
 * evaporation, best selection with the stop rules, deposits and route costs.
 *
 * All draws come from Philox4x32-10, all transcendental math from host-built
 * tables, so tours replay bit for bit from (seed, epoch, ant) and equal the
 * CPU oracle's.  Compiled with -ffp-contract=off: every float expression is
 * evaluated as written, in the order the oracle restates.
 */
#include <hip/hip_runtime.h>

#include "aco_inline.h"
#include "aco_kernels.h"
#include "aco_wave.h"

/* ---- kNN candidate lists ------------------------------------------------- */

/* One wavefront per row i: k rounds, each taking the wave-wide minimum of the
 * unique keys above the previous round's key.  Key = (primary << 24) | j with
 * primary dx^2+dy^2 (exact) or the float bits of (float)sqrt((double)s)
 * (reference OpenMP order).  Follows kernel_build_candidate_lists
 * (aco_cuda_kernels.cu:164-214): customers only, self excluded, ties to the
 * lower index, slots past n-1 padded with index 0.  Each record carries the
 * slot's distance code (the eta-table index) in its upper 16 bits. */
template <class XY, class Emit>
__device__ __forceinline__ void knn_row(XY &&xy, int n, int i, int k, int knn_key, int real,
                                        int lane, Emit &&emit) {
  const int2 pi = xy(i);
  uint64_t last = 0;
  for (int t = 0; t < k; ++t) {
    uint64_t best = ~0ull;
    for (int j = 1 + lane; j <= n; j += 64) {
      if (j == i) continue;
      const int2 pj = xy(j);
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
      if (best != ~0ull) {
        const int j = (int)(best & 0xFFFFFFu);
        emit(t, j, edge_code_r(real, pi, xy(j)));
      } else {
        emit(t, 0, 0u);
      }
    }
    last = best;
  }
}

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
  uint32_t *rec = P.rec + I.cand_off + (size_t)i * k;
  uint16_t *cidx = P.cidx + I.cand_off + (size_t)i * k;
  knn_row([&](int j) { return xy[j]; }, n, i, k, knn_key, P.dist_real, lane,
          [&](int t, int j, uint32_t d) {
            const uint32_t r = (uint32_t)j | (d << 16);
            rec[t] = r;
            cidx[t] = (uint16_t)j;
          });
  if (lane == 0) {
    const float ds = depot_w * P.eta[edge_code(P, xy[i], xy[0])];
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
        if (aco_significant(prev, (double)c, P.rules->min_rel)) S->iter_since_best = 0;
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
    const RunRules R = *P.rules;
    if (R.fixed_epochs > 0) S->halted = S->epoch >= R.fixed_epochs;
    else if (R.stagnation > 0 && S->iter_since_best >= R.stagnation) S->halted = 1;
    if (R.max_epochs > 0 && S->epoch >= R.max_epochs) S->halted = 1;
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

/* Deposit one tour on the pheromone matrix tau (row stride S): per edge
 * new = quantize(dequantize(tau[u][v]) + delta), max into tau[u][v] and
 * tau[v][u] (kernel_deposit_solution / kernel_deposit_flat_solution,
 * aco_cuda_kernels.cu:592-668).  One thread per edge; the only entries two
 * edges share are those of a single-customer route 0-c-0 (one thread does
 * both edges) and the (0,0) entry of empty routes (one thread applies them
 * all), so the result equals the sequential edge order when every customer
 * appears at most once.  Edges with a node outside 0..n1-1 deposit nothing.
 * commit(u, v, tuv, tvu) stores an edge's two new levels (u, v != both 0),
 * commit00(x) the (0,0) level.  Ends with a workgroup barrier. */
__device__ __forceinline__ void dep_edge(const uint8_t *__restrict__ tau, int S, int u, int v,
                                         bool both, float delta, const float *deq,
                                         const float *thr, uint32_t &tuv, uint32_t &tvu) {
  const uint32_t a = tau[(size_t)u * S + v];
  const uint32_t b = tau[(size_t)v * S + u];
  uint32_t nq = aco_quantize_thr(thr, deq[a] + delta);
  tuv = nq > a ? nq : a;
  tvu = nq > b ? nq : b;
  if (both) {
    /* single-customer route 0 -> v -> 0: the second edge (v, 0) in order */
    nq = aco_quantize_thr(thr, deq[tvu] + delta);
    tvu = nq > tvu ? nq : tvu;
    tuv = nq > tuv ? nq : tuv;
  }
}

template <class Commit, class Commit00>
__device__ __forceinline__ void deposit_core(const uint8_t *__restrict__ tau, int S, int n1,
                                             const int32_t *tour, int len, float delta,
                                             const float *deq, const float *thr,
                                             Commit &&commit, Commit00 &&commit00) {
  __shared__ int s_zz_count, s_zz_first;
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
          const uint32_t nq = aco_quantize_thr(thr, deq[x] + delta);
          x = nq > x ? nq : x;
        }
        commit00(x);
      }
      continue;
    }
    if ((unsigned)u >= (unsigned)n1 || (unsigned)v >= (unsigned)n1) continue;
    if (u != 0 && v == 0 && um1 == 0) continue; /* done by thread p-1 */
    uint32_t tuv, tvu;
    dep_edge(tau, S, u, v, u == 0 && vp1 == 0, delta, deq, thr, tuv, tvu);
    commit(u, v, tuv, tvu);
  }
  __syncthreads();
}

/* deposit_core on a colony of the context: the candidate-aligned copy, its
 * scores and the all-pairs weights are patched alongside.  Two dependent
 * memory round trips per edge. */
__device__ void deposit_tour(const DevCtx &P, const UpdTables &T, const ColonyDesc &D,
                             const InstDesc &I, const int32_t *tour, int len, float delta) {
  const int n1 = I.n + 1;
  uint8_t *__restrict__ tau = P.pher + D.pher_off;
  uint8_t *__restrict__ tc = P.pher + D.tcand_off;
  float *__restrict__ score = P.score + D.score_off;
  const uint8_t *__restrict__ so = P.slot_of + I.slot_off;
  const int2 *__restrict__ xy = P.xy + I.node_off;
  float *__restrict__ fs = P.fscore ? P.fscore + D.fs_off : nullptr;
  const int S = I.fs_stride;
  deposit_core(
      tau, S, n1, tour, len, delta, T.deq, T.thr,
      [&](int u, int v, uint32_t tuv, uint32_t tvu) {
        const int su = so[(size_t)u * n1 + v];
        const int sv = so[(size_t)v * n1 + u];
        const int d = (int)edge_code(P, xy[u], xy[v]);
        tau[(size_t)u * S + v] = (uint8_t)tuv;
        tau[(size_t)v * S + u] = (uint8_t)tvu;
        patch_slot(P, T, I, tc, score, u, v, su, d, (uint8_t)tuv);
        patch_slot(P, T, I, tc, score, v, u, sv, d, (uint8_t)tvu);
        if (fs) {
          fs[(size_t)u * S + v] = T.taupow[tuv] * T.eta[d];
          fs[(size_t)v * S + u] = T.taupow[tvu] * T.eta[d];
        }
      },
      [&](uint32_t x) {
        tau[0] = (uint8_t)x;
        patch_slot(P, T, I, tc, score, 0, 0, so[0], 0, (uint8_t)x);
        if (fs) fs[0] = T.taupow[x] * T.eta[0];
      });
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
        if (aco_significant(prev, (double)c, P.rules->min_rel)) S->iter_since_best = 0;
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
    const RunRules R = *P.rules;
    if (R.fixed_epochs > 0) S->halted = S->epoch >= R.fixed_epochs;
    else if (R.stagnation > 0 && S->iter_since_best >= R.stagnation) S->halted = 1;
    if (R.max_epochs > 0 && S->epoch >= R.max_epochs) S->halted = 1;
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

/* ---- plain-pointer kernels ------------------------------------------------ */

/* f applied to every byte of p[0, bytes): the unaligned head and tail byte by
 * byte, the rest as 16-byte vectors (one grid-stride pass). */
template <class F>
__device__ __forceinline__ void bytes_map(uint8_t *p, size_t bytes, F &&f) {
  const size_t head = ((16 - ((uintptr_t)p & 15)) & 15) < bytes ? ((16 - ((uintptr_t)p & 15)) & 15)
                                                               : bytes;
  const size_t nvec = (bytes - head) / 16;
  const size_t tail0 = head + nvec * 16;
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t nth = (size_t)gridDim.x * blockDim.x;
  if (tid < head) p[tid] = (uint8_t)f((uint32_t)p[tid]);
  if (tid < bytes - tail0) p[tail0 + tid] = (uint8_t)f((uint32_t)p[tail0 + tid]);
  uint4 *body = reinterpret_cast<uint4 *>(p + head);
  for (size_t e = tid; e < nvec; e += nth) {
    uint4 v = body[e];
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      uint32_t r = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) r |= f((w[i] >> (8 * b)) & 0xFFu) << (8 * b);
      w[i] = r;
    }
    body[e] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

/* kernel_init_tau (aco_cuda_kernels.cu:131-136): every byte = q */
__global__ void __launch_bounds__(256) k_pp_fill(uint8_t *tau, size_t bytes, int q) {
  bytes_map(tau, bytes, [=](uint32_t) { return (uint32_t)q; });
}

/* kernel_evaporate_tau (aco_cuda_kernels.cu:145-155): q > q_min + delta ?
 * q - delta : q_min, with q_min + delta in int (no uint8 wrap). */
__global__ void __launch_bounds__(256) k_pp_evaporate(uint8_t *tau, size_t bytes, int delta,
                                                      int q_min) {
  bytes_map(tau, bytes, [=](uint32_t q) {
    return (int)q > q_min + delta ? q - (uint32_t)delta : (uint32_t)q_min;
  });
}

/* Candidate lists on caller buffers: row i of cand / code (stride k) holds
 * the k nearest customers of node i and their heuristic-table index, in the
 * order of k_knn (one wavefront per row). */
__global__ void __launch_bounds__(256) k_pp_knn(const int32_t *x, const int32_t *y, int n, int k,
                                                int knn_key, int real, int32_t *cand,
                                                int32_t *code) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i > n) return;
  int32_t *cr = cand + (size_t)i * k;
  int32_t *dr = code + (size_t)i * k;
  knn_row([&](int j) { return make_int2(x[j], y[j]); }, n, i, k, knn_key, real, lane,
          [&](int t, int j, uint32_t d) {
            cr[t] = j;
            dr[t] = (int32_t)d;
          });
}

__device__ __forceinline__ void pp_stage_tables(const PpTables &tab, float *s_deq, float *s_thr) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    s_deq[i] = tab.deq[i];
    s_thr[i] = i < 255 ? tab.thr[i] : 0.0f;
  }
  __syncthreads();
}

/* deposit_core on a caller-owned matrix (row stride n+1), one workgroup. */
__global__ void __launch_bounds__(1024) k_pp_deposit_tour(uint8_t *tau, int n,
                                                          const int32_t *tour, int len,
                                                          float delta, PpTables tab) {
  __shared__ float s_deq[256], s_thr[256];
  pp_stage_tables(tab, s_deq, s_thr);
  const int S = n + 1;
  deposit_core(
      tau, S, S, tour, len, delta, s_deq, s_thr,
      [&](int u, int v, uint32_t tuv, uint32_t tvu) {
        tau[(size_t)u * S + v] = (uint8_t)tuv;
        tau[(size_t)v * S + u] = (uint8_t)tvu;
      },
      [&](uint32_t xq) { tau[0] = (uint8_t)xq; });
}

/* The reference's per-route deposits on caller buffers, one workgroup:
 * kernel_deposit_solution (aco_cuda_kernels.cu:592-627, STEP: node pos of
 * route v is routes[(off_v + pos) * m + ant] with off_v the sum of
 * lens[ant*K + u] - 1 over u < v) and kernel_deposit_flat_solution (:637-668,
 * !STEP: node pos of route v is routes[v * (n+1) + pos]).  Each route starts
 * from the depot whatever its entry 0 holds, and deposits its edges
 * pos-1 -> pos for pos in 1..len-1.  Entries outside the buffer (steps rows of
 * m, or K rows of n+1) and nodes outside 0..n deposit nothing.  Conflicts as
 * deposit_core: a route's 0 -> c -> 0 is one lane's, and the (0,0) entry of
 * every 0 -> 0 edge is applied by one thread at the end. */
template <bool STEP>
__global__ void __launch_bounds__(1024) k_pp_deposit_routes(uint8_t *tau, int n, int K, int m,
                                                            long long steps,
                                                            const int32_t *routes,
                                                            const int32_t *lens, int ant,
                                                            float delta, PpTables tab) {
  __shared__ float s_deq[256], s_thr[256];
  __shared__ long long s_scan[1024];
  __shared__ int s_len[1024];
  __shared__ int s_zz;
  __shared__ long long s_carry;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwaves = nt >> 6;
  const int S = n + 1;
  if (tid == 0) {
    s_zz = 0;
    s_carry = 0;
  }
  pp_stage_tables(tab, s_deq, s_thr);
  const long long limit = STEP ? steps : (long long)K * S;
  for (int v0 = 0; v0 < K; v0 += nt) {
    const int cnt = K - v0 < nt ? K - v0 : nt;
    const int len = tid < cnt ? lens[STEP ? (size_t)ant * K + v0 + tid : (size_t)v0 + tid] : 1;
    s_len[tid] = len;
    /* inclusive scan of len - 1 (Hillis-Steele over the workgroup) */
    s_scan[tid] = (long long)len - 1;
    __syncthreads();
    for (int d = 1; d < nt; d <<= 1) {
      const long long add = tid >= d ? s_scan[tid - d] : 0;
      __syncthreads();
      s_scan[tid] += add;
      __syncthreads();
    }
    const long long carry = s_carry;
    for (int w = wave; w < cnt; w += nwaves) {
      const int L = s_len[w];
      const long long base =
          STEP ? carry + s_scan[w] - ((long long)s_len[w] - 1) : (long long)(v0 + w) * S;
      /* node at position i of the route, the depot at 0 */
      auto seq = [&](int i) -> int {
        if (i == 0) return 0;
        const long long e = base + i;
        if (e < 0 || e >= limit) return -1;
        return STEP ? routes[(size_t)e * m + ant] : routes[(size_t)e];
      };
      for (int pos = 1 + lane; pos < L; pos += 64) {
        const int u = seq(pos - 1), v = seq(pos);
        if ((unsigned)u > (unsigned)n || (unsigned)v > (unsigned)n) continue;
        if (u == 0 && v == 0) {
          atomicAdd(&s_zz, 1);
          continue;
        }
        const int um1 = pos >= 2 ? seq(pos - 2) : -1;
        const int vp1 = pos + 1 < L ? seq(pos + 1) : -1;
        if (u != 0 && v == 0 && um1 == 0) continue; /* done by the lane at pos-1 */
        uint32_t tuv, tvu;
        dep_edge(tau, S, u, v, u == 0 && vp1 == 0, delta, s_deq, s_thr, tuv, tvu);
        tau[(size_t)u * S + v] = (uint8_t)tuv;
        tau[(size_t)v * S + u] = (uint8_t)tvu;
      }
    }
    __syncthreads();
    if (tid == 0) s_carry = carry + s_scan[cnt - 1];
    __syncthreads();
  }
  if (tid == 0 && s_zz > 0) {
    uint32_t xq = tau[0];
    for (int r = 0; r < s_zz; ++r) {
      const uint32_t nq = aco_quantize_thr(s_thr, s_deq[xq] + delta);
      xq = nq > xq ? nq : xq;
    }
    tau[0] = (uint8_t)xq;
  }
}
template __global__ void k_pp_deposit_routes<true>(uint8_t *, int, int, int, long long,
                                                   const int32_t *, const int32_t *, int, float,
                                                   PpTables);
template __global__ void k_pp_deposit_routes<false>(uint8_t *, int, int, int, long long,
                                                    const int32_t *, const int32_t *, int, float,
                                                    PpTables);

/* k_route_costs on caller buffers: a tour with a length outside 0..stride or
 * a node outside 0..n gets cost -1. */
__global__ void __launch_bounds__(256) k_pp_route_costs(const int32_t *x, const int32_t *y, int n,
                                                        const int32_t *tours,
                                                        const int32_t *lens, int count,
                                                        int stride, int real, long long *costs) {
  const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= count) return;
  const int32_t *tour = tours + (size_t)t * stride;
  const int len = lens[t];
  if (len < 0 || len > stride) {
    if (lane == 0) costs[t] = -1;
    return;
  }
  long long acc = 0;
  bool bad = false;
  for (int p = lane; p + 1 < len; p += 64) {
    const int a = tour[p], b = tour[p + 1];
    if ((unsigned)a > (unsigned)n || (unsigned)b > (unsigned)n) {
      bad = true;
      continue;
    }
    acc += edge_units_r(real, make_int2(x[a], y[a]), make_int2(x[b], y[b]));
  }
  if (len == 1 && lane == 0 && (unsigned)tour[0] > (unsigned)n) bad = true;
  acc = wave_sum_i64(acc);
  bad = ballot64(bad) != 0;
  if (lane == 0) costs[t] = bad ? -1 : acc;
}
