/*
 * aco_kernels.h - device data layout shared by the gfx950 kernels and the
 * host side of the C ABI (internal; the public boundary is include/aco_hip.h).
 *
 * HBM layout of a context (all pools 256-byte aligned):
 *   xy      int2   [sum(n_i+1) padded]   integer coordinates per node
 *   dem     uint16 [same]                demand per node
 *   dscore  float  [same]                depot-close score 2*(1/d(i,0))^beta
 *   rec     uint32 [sum((n_i+1)*k + 64)] candidate records idx | nint(d) << 16
 *   cidx    uint16 [same]                candidate indices (rec & 0xFFFF); 64 zero
 *                  entries after each instance's table let every lane of a
 *                  wavefront read row cur in bounds
 *   slot_of uint8  per instance (n+1)^2   1 + slot of v among u's candidates, 0 if
 *                  v is not one (lets a deposit patch tau_cand in one load)
 *   score   float  per colony (n+1)*k+64 taupow[tau_cand] * eta[d], refreshed by
 *                  the evaporation kernel and patched by the deposits
 *   fscore  float  per colony (n+1) rows x S (S = n+1 rounded up to 16):
 *                  taupow[tau[u][v]] * eta[d(u,v)] for every pair, the weights
 *                  of the all-customer fallback roulette, kept in step with tau
 *                  by the same kernels (optional: NULL when too large)
 *   pher    uint8  per colony: (n+1) x S pheromone matrix (row stride S, the
 *                  rows of the reference's (n+1)^2 matrix padded to 16 bytes;
 *                  the get/set entry points convert), then the (n+1)*k
 *                  candidate-aligned copy tau_cand[i][t] == tau[i][cand[i][t]],
 *                  16-byte padded
 *   tours   int32  [colonies][ants][stride] step-interleaved routes
 *   len/unvis int32, cost int64 [colonies*ants]
 *   best    int32  [colonies][stride]   global-best tour per colony
 *   tables  deq[256], thr[255], taupow[256], eta[code] (code = nint distance,
 *           or in real mode the bin of the float distance)
 */
#ifndef ACO_KERNELS_H
#define ACO_KERNELS_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "aco_hip.h"
#include "aco_inline.h"

struct InstDesc {
  int n, K, k, Q;
  int node_off;        /* into xy/dem/dscore (multiple of 4) */
  int row_base;        /* first kNN row of this instance */
  long long total_demand;
  long long cand_off;  /* into rec and cidx (multiple of 8) */
  long long slot_off;  /* bytes into slot_of: (n+1)^2, slot+1 of v in row u or 0 */
  uint32_t seed;
  int fs_stride;       /* S = n+1 rounded up to 16: fscore row stride and the
                          per-ant state length in LDS */
  int learned;         /* candidate slots weigh ceta[slot] instead of eta[d] */
  float scale;         /* larger side of the coordinate bounding box (>= 1) */
};

struct ColonyDesc {
  int inst;
  int colony_in_inst;  /* Philox ctr[3] */
  int m;
  int stride;          /* tour stride (multiple of 64) */
  long long pher_off;  /* bytes into pher (multiple of 256) */
  long long tcand_off; /* bytes into pher (multiple of 16) */
  long long pher_bytes;/* whole block, multiple of 16 */
  long long tour_off;  /* int32 elements into tours */
  int ant_off;         /* into len/cost/unvis */
  int best_off;        /* int32 elements into best */
  long long score_off; /* float elements into score (multiple of 4) */
  long long mat_bytes; /* (n+1)^2 rounded to 16: the matrix part of the block */
  long long fs_off;    /* float elements into fscore (multiple of 16) */
};

/* Stop rules of a run (aco_cuda.cu:322-399): fixed epochs, stagnation count,
 * an epoch cap and the significant-improvement threshold. */
struct RunRules {
  int fixed_epochs;
  int stagnation;
  int max_epochs;
  int pad;
  double min_rel;
};

struct DevCtx {
  const InstDesc *inst;
  const ColonyDesc *col;
  AcoHipColonyState *state;
  const int *row_inst;
  const int2 *xy;
  const uint16_t *dem;
  float *dscore;
  uint32_t *rec;
  uint16_t *cidx;      /* candidate index per slot (rec & 0xFFFF), per instance */
  uint8_t *slot_of;    /* per instance (n+1)^2: 1 + slot of v in row u, 0 if none */
  float *score;        /* taupow[tau_cand] * eta[d] per slot, per colony */
  float *fscore;       /* per colony (n+1) x S all-pairs weights, NULL = off */
  const float *ceta;   /* learned heuristic per candidate slot (indexed like rec) */
  uint8_t *pher;
  int32_t *tours;
  int32_t *len;
  long long *cost;
  int32_t *unvis;
  int32_t *best;
  const float *deq;
  const float *thr;
  const float *taupow;
  const float *eta;
  int eta_len;
  int eta_lds;         /* the eta table is staged in LDS (else read from global) */
  int dist_real;       /* ACO_HIP_DIST_REAL: float distances, costs in 2^-23 units */
  uint32_t q_evap;
  float qdep;
  const RunRules *rules; /* stop rules of the current run, read at each epoch's end (device
                          * memory, so a captured epoch graph outlives a change of rules) */
  void *hist;          /* HistEntry[colonies][ACO_HIST] per-epoch progress ring */
  unsigned long long *stats; /* [colonies][ACO_STATS] counters, NULL = off */
  int *done;           /* per-colony block counter of k_evap_update (kept at 0) */
  int *feas;           /* per colony: 1 + the last epoch with a feasible ant */
};

/* counters per colony: construct [0] decisions, [1] fallbacks, [2] fallback
 * cycles, [3] ant cycles, [4] ants, [5] staging cycles (clock64 = shader
 * clock); update kernel [6] updates, [7] evaporation ticks, [8] selection
 * ticks, [9] deposit ticks (wall_clock64), [10] ~(first block start) of the
 * running update (reset to 0 by its last block) */
#define ACO_STATS 16
/* colonies whose evaporation offsets k_evaporate_flat stages in LDS */
#define EVAP_LDS_COLONIES 1024

#define ACO_HIST 64
struct HistEntry {
  int epoch;
  int iter_since_best;
  int improved;
  int halted;
  long long global_best;
};

static inline __host__ __device__ size_t round16(size_t x) { return (x + 15) & ~(size_t)15; }

/* Squared length of an edge (coordinates in [0, 32767]: exact in 32 bits). */
static inline __device__ uint32_t edge_sq(int2 a, int2 b) {
  return aco_sq15(a.x - b.x) + aco_sq15(a.y - b.y);
}

/* Distance code of an edge, the index of its (1/d)^beta in the eta table:
 * the nint distance, or in real mode the bin of the float distance. */
template <bool REAL>
static inline __device__ uint32_t edge_code_t(int2 a, int2 b) {
  const uint32_t s = edge_sq(a, b);
  return REAL ? aco_real_eta_bin(aco_real_dist_s(s)) : (uint32_t)aco_nint_sqrt_u32(s);
}
static inline __device__ uint32_t edge_code_r(int real, int2 a, int2 b) {
  return real ? edge_code_t<true>(a, b) : edge_code_t<false>(a, b);
}
static inline __device__ uint32_t edge_code(const DevCtx &P, int2 a, int2 b) {
  return edge_code_r(P.dist_real, a, b);
}

/* Cost of an edge in cost units: the nint distance, or in real mode the
 * float distance times 2^23 (exact). */
static inline __device__ long long edge_units_r(int real, int2 a, int2 b) {
  const uint32_t s = edge_sq(a, b);
  return real ? (long long)aco_real_units(aco_real_dist_s(s)) : (long long)aco_nint_sqrt_u32(s);
}
static inline __device__ long long edge_units(const DevCtx &P, int2 a, int2 b) {
  return edge_units_r(P.dist_real, a, b);
}

/* Tour length L of a cost in cost units, as the float the deposits divide
 * by: (float)cost, times 2^-23 in real mode (exact scaling). */
static inline __host__ __device__ float cost_len(int dist_real, long long cost) {
  return dist_real ? (float)cost * (float)ACO_HIP_REAL_COST_UNIT : (float)cost;
}

/* Heuristic factor of candidate slot e (0 <= e < (n+1)k) of instance I: the
 * learned prior when set, else (1/d)^beta of the slot's distance code. */
static inline __device__ float slot_eta(const DevCtx &P, const InstDesc &I, size_t e,
                                        uint32_t rec) {
  return I.learned ? P.ceta[I.cand_off + e] : P.eta[rec >> 16];
}

/* is_significant_improvement (src/cuda/aco_cuda.cu:83-94), ACO_EPS = 1e-9. */
static inline __host__ __device__ int aco_significant(double prev, double now, double min_rel) {
  const double eps = 1e-9;
  if (prev >= 1.7976931348623157e308 || now >= 1.7976931348623157e308) return now < prev;
  if (now >= prev - eps) return 0;
  const double rel = (prev - now) / (prev > eps ? prev : eps);
  return rel + eps >= min_rel;
}

/* kernels (aco_kernels.hip; construction: aco_construct.h, instantiated in
 * aco_construct_fsm.hip for FSM = true and aco_construct_tab.hip for false) */
__global__ void k_knn(DevCtx P, int total_rows, int knn_key, float depot_w);
/* quantization tables passed by value (2 KB of kernel arguments) */
struct PpTables {
  float deq[256];
  float thr[255];
};
/* kernels of the plain-pointer entry points (caller-owned buffers, the
 * reference's pheromone layout: (n+1)^2 bytes, row stride n+1) */
__global__ void k_pp_fill(uint8_t *tau, size_t bytes, int q);
__global__ void k_pp_evaporate(uint8_t *tau, size_t bytes, int delta, int q_min);
__global__ void k_pp_knn(const int32_t *x, const int32_t *y, int n, int k, int knn_key, int real,
                         int32_t *cand, int32_t *code);
__global__ void k_pp_deposit_tour(uint8_t *tau, int n, const int32_t *tour, int len, float delta,
                                  PpTables tab);
template <bool STEP>
__global__ void k_pp_deposit_routes(uint8_t *tau, int n, int K, int m, long long steps,
                                    const int32_t *routes, const int32_t *lens, int ant,
                                    float delta, PpTables tab);
__global__ void k_pp_route_costs(const int32_t *x, const int32_t *y, int n, const int32_t *tours,
                                 const int32_t *lens, int count, int stride, int real,
                                 long long *costs);
__global__ void k_fill_tau(uint8_t *pher, size_t bytes, uint32_t word);
template <bool STAGE, bool STATS, int ROWS, bool XCD, bool FSM>
__global__ void k_construct(DevCtx P, int groups_per_colony, int W, const int *cols);
template <bool STAGE, bool STATS, bool XCD, bool FSM, bool RING, bool B8>
__global__ void __launch_bounds__(B8 ? 1024 : 512) k_construct_pair(DevCtx P, int groups_per_colony,
                                                                   int W, const int *cols);
__global__ void k_evaporate(DevCtx P, int mat_blocks, int from_ants);
__global__ void k_evaporate_flat(DevCtx P, const long long *item_off, const long long *slot_off,
                                 int ncol, int eta_lds, const int *cmap);
__global__ void k_refresh_scores(DevCtx P);
__global__ void k_update(DevCtx P, int mode);
__global__ void k_evap_update(DevCtx P, int mat_blocks);
__global__ void k_select_deposit(DevCtx P, int stage_eta, const int *cmap);

/* LDS bytes of k_select_deposit: both tours (int32) and the position of
 * every customer in each (u16). */
static inline __host__ __device__ size_t sel_dep_lds(int n1, int stride) {
  return (size_t)stride * 8 + (size_t)((n1 + 7) & ~7) * 4;
}
__global__ void k_build_slots(DevCtx P, int total_rows);
__global__ void k_deposit_tour(DevCtx P, int colony, const int32_t *tour, int len, float delta);
__global__ void k_route_costs(DevCtx P, int instance, const int32_t *tours, const int32_t *lens,
                              int count, int stride, long long *out);

#endif
