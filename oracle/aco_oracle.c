/*
 * aco_oracle.c - CPU oracle of the ACO-CVRP hot path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, the smoke check and the benchmark's
 * CPU-baseline leg load this code, and only as the checker.  The product
 * path (the HIP library) never calls it.
 *
 * It restates, in plain C, the reference CUDA back-end's per-epoch algorithm
 * (src/cuda/aco_cuda.cu:188-454 driving src/cuda/aco_cuda_kernels.cu) with the
 * changes the MI355X build makes on purpose:
 *   D1 weighted demand: a candidate is eligible when load + d_j <= Q and the
 *      close rule uses the remaining demand;
 *   D2 TSPLIB nearest-integer distances with integer route costs, or (real
 *      mode) the reference's unrounded distances as floats, with route costs
 *      kept as exact sums of the float lengths in units of 2^-23 and the
 *      heuristic (1/d)^beta tabled over bins of relative width 2^-10;
 *   D8 Philox4x32-10 draws keyed by (decision, ant, epoch, colony);
 *   exact roulette: weights are summed in a fixed, documented order (the
 *      order of the wavefront scan) and selection takes the first prefix above
 *      the threshold, so a roulette never "misses" (aco_cuda_kernels.cu:313-321);
 *   dequantize / tau^alpha / eta^beta / quantize go through host tables.
 * Every function cites the reference lines it follows.  It rebuilds its own
 * tables and never includes product headers.
 *
 * Compile with -ffp-contract=off: every float operation below is a single
 * IEEE operation, in the order written.
 */
#include "aco_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define VISITED 0xFFFFu
#define FB_V 16            /* customers per lane in the fallback scan */
#define FB_CHUNK (64 * FB_V)

struct OracleColony {
  int n, K, Q, k, m;
  int real; /* real-distance mode */
  int32_t *x, *y, *dem;
  int64_t total_demand;
  float alpha, beta, rho, tau0, Qdep, depot_w;
  uint32_t seed, colony;
  /* tables */
  float log_min, step;
  uint8_t q_tau0, q_evap;
  float deq[256], thr[255], taupow[256];
  float *eta_lut; /* (1/d)^beta by distance code 0..dmax */
  int dmax;       /* largest distance code */
  /* candidates */
  uint32_t *rec;  /* idx | d << 16, (n+1)*k */
  float *ceta;    /* learned heuristic per candidate slot, NULL = (1/d)^beta */
  float *dscore;  /* depot close score per node */
  /* pheromone */
  uint8_t *tau;      /* (n+1)^2 */
  uint8_t *tau_cand; /* (n+1)*k */
  /* ants */
  int stride;
  int32_t *tours, *len, *unvisited;
  int64_t *cost, *fallbacks;
  /* colony state */
  OracleState st;
  int32_t *best_tour;
  int stagnation;
  double min_rel;
  int fixed_epochs;
};

/* ---- primitives ---------------------------------------------------------- */

/* Philox4x32-10, Salmon et al. SC'11: round = two 32x32->64 multiplies,
 * Weyl key schedule. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t a = (uint64_t)0xD2511F53u * c0;
    uint64_t b = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)b;
    uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)a;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

/* TSPLIB nint(sqrt(dx^2+dy^2)) by exact integer square root. */
int32_t oracle_nint_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2) {
  int64_t dx = x1 - x2, dy = y1 - y2;
  uint64_t s = (uint64_t)(dx * dx + dy * dy);
  uint64_t r = (uint64_t)sqrt((double)s);
  while (r * r > s) --r;
  while ((r + 1) * (r + 1) <= s) ++r;
  /* sqrt(s) >= r + 1/2  <=>  4s >= 4r^2 + 4r + 1  <=>  s > r^2 + r */
  return (int32_t)(r + ((s > r * r + r) ? 1 : 0));
}

static uint64_t sq_dist(const OracleColony *o, int a, int b) {
  int64_t dx = o->x[a] - o->x[b], dy = o->y[a] - o->y[b];
  return (uint64_t)(dx * dx + dy * dy);
}

/* ---- real mode (instance_parser.c:17-19, aco_cuda_kernels.cu:53-57) ------ */

/* The reference's unrounded distance as a float: the double root (glibc's
 * sqrt is correctly rounded) rounded once to float. */
static float real_dist_s(uint64_t s) { return (float)sqrt((double)s); }

/* Heuristic-table code of a float distance: 0 for d = 0, else 1 + its bin
 * among 1024 bins per binade of the float's bit pattern. */
static uint32_t real_bin(float d) {
  if (!(d >= 1.0f)) return 0;
  uint32_t bits;
  memcpy(&bits, &d, 4);
  return 1u + ((bits - 0x3F800000u) >> 13);
}

/* Distance code of an edge: nint distance, or the bin of the float distance. */
static uint32_t dist_code(const OracleColony *o, int a, int b) {
  if (o->real) return real_bin(real_dist_s(sq_dist(o, a, b)));
  return (uint32_t)oracle_nint_dist(o->x[a], o->y[a], o->x[b], o->y[b]);
}

/* Cost of an edge: nint distance, or the float distance in units of 2^-23. */
static int64_t edge_cost(const OracleColony *o, int a, int b) {
  if (o->real) return (int64_t)((double)real_dist_s(sq_dist(o, a, b)) * 8388608.0);
  return oracle_nint_dist(o->x[a], o->y[a], o->x[b], o->y[b]);
}

/* Tour length the deposits divide by. */
static float cost_len(const OracleColony *o, int64_t c) {
  return o->real ? (float)c * (1.0f / 8388608.0f) : (float)c;
}

/* Inclusive prefix sums of up to 64 lane values in the order of the
 * wavefront scan: four in-row Kogge-Stone passes (shift 1, 2, 4, 8 inside each
 * row of 16 lanes, lanes shifted in from outside the row add nothing), then
 * row r (r = 1..R-1, R = ceil(nlanes/16)) adds the last prefix of row r-1.
 * Plays the role of cuda_warp_sum/cuda_warp_prefix_sum
 * (aco_cuda_kernels.cu:64-85), with one order for total and prefixes. */
void oracle_scan_rows(const float *w, int nlanes, float *prefix) {
  float cur[64], nxt[64];
  for (int l = 0; l < 64; ++l) cur[l] = (l < nlanes) ? w[l] : 0.0f;
  for (int off = 1; off < 16; off <<= 1) {
    for (int l = 0; l < 64; ++l) nxt[l] = ((l & 15) >= off) ? cur[l] + cur[l - off] : cur[l];
    memcpy(cur, nxt, sizeof(cur));
  }
  int rows = (nlanes + 15) / 16;
  for (int r = 1; r < rows; ++r) {
    float c = cur[16 * r - 1];
    for (int l = 16 * r; l < 16 * r + 16; ++l) cur[l] = cur[l] + c;
  }
  memcpy(prefix, cur, sizeof(float) * 64);
}

/* ---- tables (aco_cuda.cu:209-215,242; aco_cuda_kernels.cu:25-45,207-212) -- */

static void build_tables(OracleColony *o) {
  float lmin = logf(0.0001f), lmax = logf(100.0f);
  float step = (lmax - lmin) / 255.0f;
  o->log_min = lmin;
  o->step = step;
  float lr = logf(1.0f - o->rho);
  o->q_evap = (uint8_t)fmaxf(1.0f, roundf(-lr / step));
  o->q_tau0 = (uint8_t)fmaxf(0.0f, fminf(255.0f, roundf((logf(o->tau0) - lmin) / step)));
  for (int q = 0; q < 256; ++q) {
    /* dequantize_tau: expf(q*step + log_min), the product-sum fused */
    o->deq[q] = (float)exp((double)fmaf((float)q, step, lmin));
    o->taupow[q] = (o->alpha == 1.0f) ? o->deq[q] : (float)pow((double)o->deq[q], (double)o->alpha);
  }
  for (int b = 0; b < 255; ++b) {
    /* quantize_tau rounds (log tau - log_min)/step: level >= b+1 iff the
     * ideal value is >= b + 1/2.  Smallest float meeting that, in double. */
    double target = (double)lmin + ((double)b + 0.5) * (double)step;
    float f = (float)exp(target);
    while (log((double)f) < target) f = nextafterf(f, INFINITY);
    while (log((double)nextafterf(f, 0.0f)) >= target) f = nextafterf(f, 0.0f);
    o->thr[b] = f;
  }
}

uint8_t oracle_quantize(const OracleColony *o, float tau) {
  if (!(tau > 0.0f)) return 0;
  int q = 0;
  while (q < 255 && tau >= o->thr[q]) ++q; /* linear count, thresholds increase */
  return (uint8_t)q;
}

/* ---- candidate lists (aco_cuda_kernels.cu:164-214, aco_parallel.c:218-250) */

typedef struct {
  uint64_t key;
} KeyItem;

static int cmp_key(const void *a, const void *b) {
  uint64_t x = ((const KeyItem *)a)->key, y = ((const KeyItem *)b)->key;
  return (x > y) - (x < y);
}

static void build_candidates(OracleColony *o, int knn_key) {
  int n = o->n, k = o->k;
  KeyItem *items = (KeyItem *)malloc(sizeof(KeyItem) * (size_t)(n + 1));
  for (int i = 0; i <= n; ++i) {
    int cnt = 0;
    for (int j = 1; j <= n; ++j) {
      if (j == i) continue;
      uint64_t s = sq_dist(o, i, j);
      uint64_t hi;
      if (knn_key == 1) {
        float fd = (float)sqrt((double)s);
        uint32_t bits;
        memcpy(&bits, &fd, 4);
        hi = bits;
      } else {
        hi = s;
      }
      items[cnt++].key = (hi << 24) | (uint64_t)j;
    }
    qsort(items, (size_t)cnt, sizeof(KeyItem), cmp_key);
    for (int t = 0; t < k; ++t) {
      uint32_t idx = 0, d = 0;
      if (t < cnt) {
        idx = (uint32_t)(items[t].key & 0xFFFFFFu);
        d = dist_code(o, i, (int)idx);
      }
      o->rec[(size_t)i * k + t] = idx | (d << 16);
    }
  }
  free(items);
  for (int i = 0; i <= n; ++i) o->dscore[i] = o->depot_w * o->eta_lut[dist_code(o, i, 0)];
}

/* ---- lifecycle ----------------------------------------------------------- */

OracleColony *oracle_create(int n, int K, int Q, const int32_t *x, const int32_t *y,
                            const int32_t *demand, int cand_k, int ants, float alpha, float beta,
                            float rho, float tau0, float Qdep, float depot_w, int knn_key,
                            uint32_t seed, uint32_t colony, int dist_mode) {
  if (n <= 0 || K <= 0 || Q <= 0 || ants <= 0 || cand_k <= 0) return NULL;
  OracleColony *o = (OracleColony *)calloc(1, sizeof(*o));
  o->real = dist_mode == 1;
  o->n = n; o->K = K; o->Q = Q; o->m = ants;
  o->k = cand_k > n ? n : cand_k; /* aco_cuda.cu:198-199 clamps k to n */
  if (o->k > 64) o->k = 64;
  o->alpha = alpha; o->beta = beta; o->rho = rho; o->tau0 = tau0;
  o->Qdep = Qdep; o->depot_w = depot_w; o->seed = seed; o->colony = colony;
  o->x = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  o->y = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  o->dem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  memcpy(o->x, x, sizeof(int32_t) * (size_t)(n + 1));
  memcpy(o->y, y, sizeof(int32_t) * (size_t)(n + 1));
  memcpy(o->dem, demand, sizeof(int32_t) * (size_t)(n + 1));
  o->total_demand = 0;
  int32_t xmin = x[0], xmax = x[0], ymin = y[0], ymax = y[0];
  for (int i = 0; i <= n; ++i) {
    if (i) o->total_demand += demand[i];
    if (x[i] < xmin) xmin = x[i];
    if (x[i] > xmax) xmax = x[i];
    if (y[i] < ymin) ymin = y[i];
    if (y[i] > ymax) ymax = y[i];
  }
  /* the bounding box's diagonal has the largest distance code */
  if (o->real) {
    const uint64_t dx = (uint64_t)(xmax - xmin), dy = (uint64_t)(ymax - ymin);
    o->dmax = (int)real_bin(real_dist_s(dx * dx + dy * dy));
  } else {
    o->dmax = oracle_nint_dist(xmin, ymin, xmax, ymax);
  }
  build_tables(o);
  o->eta_lut = (float *)malloc(sizeof(float) * (size_t)(o->dmax + 1));
  for (int d = 0; d <= o->dmax; ++d) {
    double base;
    if (d == 0) {
      base = (double)1e-9f;
    } else if (o->real) {
      /* the middle float of bin d: bits 1.0f + (d-1)*2^13 + 2^12 */
      const uint32_t bits = 0x3F800000u + ((uint32_t)(d - 1) << 13) + (1u << 12);
      float c;
      memcpy(&c, &bits, 4);
      base = (double)c;
    } else {
      base = (double)d;
    }
    o->eta_lut[d] = (float)pow(base, -(double)beta);
  }
  o->rec = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1) * o->k);
  o->dscore = (float *)malloc(sizeof(float) * (size_t)(n + 1));
  build_candidates(o, knn_key);
  size_t nn = (size_t)(n + 1) * (size_t)(n + 1);
  o->tau = (uint8_t *)malloc(nn);
  o->tau_cand = (uint8_t *)malloc((size_t)(n + 1) * o->k);
  memset(o->tau, o->q_tau0, nn); /* kernel_init_tau fills every byte */
  memset(o->tau_cand, o->q_tau0, (size_t)(n + 1) * o->k);
  o->stride = ((n + K + 1) + 63) / 64 * 64;
  o->tours = (int32_t *)calloc((size_t)ants * o->stride, sizeof(int32_t));
  o->len = (int32_t *)calloc((size_t)ants, sizeof(int32_t));
  o->unvisited = (int32_t *)calloc((size_t)ants, sizeof(int32_t));
  o->cost = (int64_t *)calloc((size_t)ants, sizeof(int64_t));
  o->fallbacks = (int64_t *)calloc((size_t)ants, sizeof(int64_t));
  o->best_tour = (int32_t *)calloc((size_t)o->stride, sizeof(int32_t));
  o->st.best_ant = -1;
  o->st.iter_best_cost = INT64_MAX;
  o->st.global_best_cost = INT64_MAX;
  o->min_rel = 1e-3;
  return o;
}

void oracle_destroy(OracleColony *o) {
  if (!o) return;
  free(o->x); free(o->y); free(o->dem); free(o->eta_lut); free(o->rec); free(o->dscore);
  free(o->tau); free(o->tau_cand); free(o->tours); free(o->len); free(o->unvisited);
  free(o->cost); free(o->fallbacks); free(o->best_tour); free(o->ceta);
  free(o);
}

/* Learned heuristic prior (the seam of aco_cuda_kernels.cu:295): candidate
 * slot (i, t) weighs taupow[tau_cand] * eta[i*k + t] instead of
 * taupow * (1/d)^beta.  The fallback keeps (1/d)^beta.  NULL restores it. */
void oracle_set_candidate_eta(OracleColony *o, const float *eta) {
  free(o->ceta);
  o->ceta = NULL;
  if (!eta) return;
  const size_t cnt = (size_t)(o->n + 1) * o->k;
  o->ceta = (float *)malloc(sizeof(float) * cnt);
  memcpy(o->ceta, eta, sizeof(float) * cnt);
}

int oracle_cand_k(const OracleColony *o) { return o->k; }
int oracle_tour_stride(const OracleColony *o) { return o->stride; }

void oracle_get_candidates(const OracleColony *o, int32_t *idx, int32_t *dist, float *eta) {
  size_t total = (size_t)(o->n + 1) * o->k;
  for (size_t e = 0; e < total; ++e) {
    uint32_t r = o->rec[e];
    uint32_t j = r & 0xFFFFu, d = r >> 16;
    if (idx) idx[e] = (int32_t)j;
    if (dist) dist[e] = (int32_t)d;
    if (eta) eta[e] = j ? o->eta_lut[d] : 0.0f;
  }
}

void oracle_get_tables(const OracleColony *o, float *deq, float *thr, float *taupow,
                       float *eta_lut, int *eta_len, uint8_t *q_tau0, uint8_t *q_evap) {
  if (deq) memcpy(deq, o->deq, sizeof(o->deq));
  if (thr) memcpy(thr, o->thr, sizeof(o->thr));
  if (taupow) memcpy(taupow, o->taupow, sizeof(o->taupow));
  if (eta_lut) memcpy(eta_lut, o->eta_lut, sizeof(float) * (size_t)(o->dmax + 1));
  if (eta_len) *eta_len = o->dmax + 1;
  if (q_tau0) *q_tau0 = o->q_tau0;
  if (q_evap) *q_evap = o->q_evap;
}

static void refresh_tau_cand(OracleColony *o) {
  for (int i = 0; i <= o->n; ++i)
    for (int t = 0; t < o->k; ++t)
      o->tau_cand[(size_t)i * o->k + t] =
          o->tau[(size_t)i * (o->n + 1) + (o->rec[(size_t)i * o->k + t] & 0xFFFFu)];
}

void oracle_get_tau(const OracleColony *o, uint8_t *tau, uint8_t *tau_cand) {
  if (tau) memcpy(tau, o->tau, (size_t)(o->n + 1) * (o->n + 1));
  if (tau_cand) memcpy(tau_cand, o->tau_cand, (size_t)(o->n + 1) * o->k);
}

void oracle_set_tau(OracleColony *o, const uint8_t *tau) {
  memcpy(o->tau, tau, (size_t)(o->n + 1) * (o->n + 1));
  refresh_tau_cand(o);
}

void oracle_set_epoch(OracleColony *o, int epoch) { o->st.epoch = epoch; }

/* New stop rules start a new run of the colony: its halted flag is
 * re-evaluated under them, as the product's run loop does when a colony is
 * continued by a later run (the reference's solve is one-shot). */
void oracle_set_stop_rules(OracleColony *o, int stagnation_epochs, double min_rel,
                           int fixed_epochs) {
  o->stagnation = stagnation_epochs;
  o->min_rel = min_rel;
  o->fixed_epochs = fixed_epochs;
  OracleState *st = &o->st;
  if (fixed_epochs > 0) st->halted = st->epoch >= fixed_epochs;
  else st->halted = stagnation_epochs > 0 && st->iter_since_best >= stagnation_epochs;
}

/* ---- tour construction (aco_cuda_kernels.cu:447-562) --------------------- */

/* Fallback roulette over every unvisited customer that fits, weight
 * tau^alpha * eta^beta (select_global_fallback, aco_cuda_kernels.cu:336-430,
 * restated as an exact two-pass roulette).  Node j = c*1024 + l*16 + i
 * (0 <= j <= n) belongs to chunk c, lane l, slot i; the depot is visited from
 * the start and never carries weight.  Returns the node or -1 and the total. */
static int fallback_select(const OracleColony *o, const uint16_t *state, int cur, int room,
                           float u, float *total_out) {
  const int n = o->n;
  const uint8_t *trow = o->tau + (size_t)cur * (n + 1);
  int nch = (n + 1 + FB_CHUNK - 1) / FB_CHUNK;
  float *Rc = (float *)malloc(sizeof(float) * (size_t)nch);
  float *Cc = (float *)malloc(sizeof(float) * (size_t)nch);
  float lanes[64], P[64], w[64][FB_V];
  float R = 0.0f;
  for (int c = 0; c < nch; ++c) {
    for (int l = 0; l < 64; ++l) {
      float s = 0.0f;
      for (int i = 0; i < FB_V; ++i) {
        int j = c * FB_CHUNK + l * FB_V + i;
        float wt = 0.0f;
        if (j <= n && (int)state[j] <= room)
          wt = o->taupow[trow[j]] * o->eta_lut[dist_code(o, cur, j)];
        s = s + wt;
      }
      lanes[l] = s;
    }
    oracle_scan_rows(lanes, 64, P);
    Cc[c] = P[63];
    R = R + Cc[c];
    Rc[c] = R;
  }
  *total_out = R;
  if (!(R > 0.0f)) {
    free(Rc); free(Cc);
    return -1;
  }
  float thr = u * R;
  int cs = -1;
  for (int c = 0; c < nch; ++c)
    if (Rc[c] > thr) { cs = c; break; }
  if (cs < 0)
    for (int c = nch - 1; c >= 0; --c)
      if (Cc[c] > 0.0f) { cs = c; break; }
  float base = cs > 0 ? Rc[cs - 1] : 0.0f;
  float thr_c = thr - base;
  for (int l = 0; l < 64; ++l) {
    float s = 0.0f;
    for (int i = 0; i < FB_V; ++i) {
      int j = cs * FB_CHUNK + l * FB_V + i;
      float wt = 0.0f;
      if (j <= n && (int)state[j] <= room)
        wt = o->taupow[trow[j]] * o->eta_lut[dist_code(o, cur, j)];
      w[l][i] = wt;
      s = s + wt;
    }
    lanes[l] = s;
  }
  oracle_scan_rows(lanes, 64, P);
  int ls = -1;
  for (int l = 0; l < 64; ++l)
    if (lanes[l] > 0.0f && P[l] > thr_c) { ls = l; break; }
  if (ls < 0)
    for (int l = 63; l >= 0; --l)
      if (lanes[l] > 0.0f) { ls = l; break; }
  float thr_l = thr_c - (ls > 0 ? P[ls - 1] : 0.0f);
  int is = -1;
  float p = 0.0f;
  for (int i = 0; i < FB_V; ++i) {
    p = p + w[ls][i];
    if (w[ls][i] > 0.0f && p > thr_l) { is = i; break; }
  }
  if (is < 0)
    for (int i = FB_V - 1; i >= 0; --i)
      if (w[ls][i] > 0.0f) { is = i; break; }
  free(Rc); free(Cc);
  return cs * FB_CHUNK + ls * FB_V + is;
}

int oracle_construct_ant(const OracleColony *o, int ant, int epoch, int32_t *tour,
                         int64_t *cost_out, int32_t *unvisited_out, int64_t *fallbacks_out) {
  const int n = o->n, K = o->K, Q = o->Q, k = o->k;
  uint16_t *state = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n + 1));
  /* kernel_reset_ant_state (aco_cuda_kernels.cu:227-257): depot pre-visited */
  state[0] = VISITED;
  for (int j = 1; j <= n; ++j) state[j] = (uint16_t)o->dem[j];
  int64_t rem_dem = o->total_demand;
  int rem_cnt = n;
  uint32_t s = 0; /* decision counter: Philox ctr[0] */
  int pos = 0;
  int64_t fallbacks = 0;
  tour[0] = 0;
  const uint32_t key[2] = {o->seed, 0x5CA1AB1Eu};
  for (int veh = 0; veh < K; ++veh) {
    int cur = 0;
    int load = 0;
    for (;;) {
      /* loop head of aco_cuda_kernels.cu:492 */
      if (rem_cnt <= 0 || load >= Q) break;
      const int room = Q - load;
      uint32_t ctr[4] = {s, (uint32_t)ant, (uint32_t)epoch, o->colony}, r[4];
      oracle_philox4x32_10(ctr, key, r);
      ++s;
      /* candidate roulette: select_from_candidates (aco_cuda_kernels.cu:271-322) */
      float w[64], P[64];
      int jj[64], dd[64];
      for (int t = 0; t < 64; ++t) {
        w[t] = 0.0f;
        jj[t] = 0;
        dd[t] = 0;
      }
      for (int t = 0; t < k; ++t) {
        uint32_t rec = o->rec[(size_t)cur * k + t];
        int j = (int)(rec & 0xFFFFu);
        int d = (int)(rec >> 16);
        jj[t] = j;
        dd[t] = state[j];
        if ((int)state[j] <= room) {
          uint8_t q = o->tau_cand[(size_t)cur * k + t];
          w[t] = o->taupow[q] * (o->ceta ? o->ceta[(size_t)cur * k + t] : o->eta_lut[d]);
        }
      }
      oracle_scan_rows(w, k, P);
      /* the total: the prefix at the last lane of the last 16-lane row the
       * list spans, the weights past k being 0 (for k a multiple of 16 the
       * prefix of the last candidate); the rows' in-row scans fix its
       * grouping, so it can differ from P[k - 1] in the last bit */
      float T = P[16 * ((k + 15) / 16) - 1];
      int move = -1, mdem = 0;
      if (T > 0.0f) {
        float thr = u01(r[0]) * T;
        int sel = -1;
        for (int t = 0; t < k; ++t)
          if (w[t] > 0.0f && P[t] > thr) { sel = t; break; }
        if (sel < 0)
          for (int t = k - 1; t >= 0; --t)
            if (w[t] > 0.0f) { sel = t; break; }
        move = jj[sel];
        mdem = dd[sel];
      } else {
        ++fallbacks;
        move = fallback_select(o, state, cur, room, u01(r[0]), &T);
        if (move > 0) mdem = state[move];
      }
      /* depot-close roulette (aco_cuda_kernels.cu:501-520) on remaining demand */
      int close_allowed = rem_dem <= (int64_t)(K - veh - 1) * Q;
      float dsc = close_allowed ? o->dscore[cur] : 0.0f;
      if (move > 0 && dsc > 0.0f) {
        float thr2 = u01(r[1]) * (T + dsc);
        if (thr2 >= T) move = 0;
      }
      if (move <= 0) break;
      tour[++pos] = move;
      state[move] = VISITED;
      load += mdem;
      rem_dem -= mdem;
      --rem_cnt;
      cur = move;
    }
    tour[++pos] = 0; /* route close (aco_cuda_kernels.cu:547-552) */
  }
  int len = pos + 1;
  int64_t cost = 0;
  for (int p = 0; p + 1 < len; ++p) cost += edge_cost(o, tour[p], tour[p + 1]);
  *cost_out = cost;
  *unvisited_out = rem_cnt;
  if (fallbacks_out) *fallbacks_out = fallbacks;
  free(state);
  return len;
}

void oracle_construct(OracleColony *o, int nthreads) {
  const int epoch = o->st.epoch;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int a = 0; a < o->m; ++a) {
    o->len[a] = oracle_construct_ant(o, a, epoch, o->tours + (size_t)a * o->stride,
                                     &o->cost[a], &o->unvisited[a], &o->fallbacks[a]);
  }
  (void)nthreads;
}

void oracle_get_tours(const OracleColony *o, int32_t *tours, int32_t *len, int64_t *cost,
                      int32_t *unvisited) {
  if (tours) memcpy(tours, o->tours, sizeof(int32_t) * (size_t)o->m * o->stride);
  if (len) memcpy(len, o->len, sizeof(int32_t) * (size_t)o->m);
  if (cost) memcpy(cost, o->cost, sizeof(int64_t) * (size_t)o->m);
  if (unvisited) memcpy(unvisited, o->unvisited, sizeof(int32_t) * (size_t)o->m);
}

int64_t oracle_fallback_count(const OracleColony *o) {
  int64_t t = 0;
  for (int a = 0; a < o->m; ++a) t += o->fallbacks[a];
  return t;
}

/* ---- best selection and stop rules (aco_cuda.cu:104-126, 331-399) -------- */

static int significant(double prev, double now, double min_rel) {
  const double eps = 1e-9;
  if (prev >= DBL_MAX || now >= DBL_MAX) return now < prev;
  if (now >= prev - eps) return 0;
  return (prev - now) / fmax(prev, eps) + eps >= min_rel;
}

int oracle_significant(double prev, double now, double min_rel) {
  return significant(prev, now, min_rel);
}

/* Replaces the ants' results of the current epoch (costs, unvisited counts,
 * tour lengths; any may be NULL) so that the best selection and stop rules
 * can be driven with chosen values. */
void oracle_set_ant_results(OracleColony *o, const int64_t *cost, const int32_t *unvisited,
                            const int32_t *len) {
  if (cost) memcpy(o->cost, cost, sizeof(int64_t) * (size_t)o->m);
  if (unvisited) memcpy(o->unvisited, unvisited, sizeof(int32_t) * (size_t)o->m);
  if (len) memcpy(o->len, len, sizeof(int32_t) * (size_t)o->m);
}

void oracle_select_best(OracleColony *o) {
  OracleState *st = &o->st;
  if (st->halted) {
    st->found = 0;
    return;
  }
  int best = -1, feas = 0;
  for (int a = 0; a < o->m; ++a) {
    if (o->unvisited[a] != 0) continue;
    ++feas;
    if (best < 0 || o->cost[a] < o->cost[best]) best = a; /* ties -> lower ant */
  }
  st->feasible_ants = feas;
  st->found = best >= 0;
  st->best_ant = best;
  st->improved = 0;
  if (best >= 0) {
    int64_t c = o->cost[best];
    st->iter_best_cost = c;
    if (c < st->global_best_cost) {
      double prev = st->global_best_cost == INT64_MAX ? DBL_MAX : (double)st->global_best_cost;
      st->improved = 1;
      if (significant(prev, (double)c, o->min_rel)) st->iter_since_best = 0;
      else st->iter_since_best++;
      st->global_best_cost = c;
      st->best_len = o->len[best];
      memcpy(o->best_tour, o->tours + (size_t)best * o->stride, sizeof(int32_t) * (size_t)o->stride);
    } else {
      st->iter_since_best++;
    }
    const float qeff = o->Qdep * 100.0f;
    st->dep_iter = (0.3f * qeff) / cost_len(o, c);
    st->dep_best = (0.7f * qeff) / cost_len(o, st->global_best_cost);
  } else {
    st->iter_best_cost = INT64_MAX;
    st->iter_since_best++;
  }
  st->epoch++;
  if (o->fixed_epochs > 0) st->halted = st->epoch >= o->fixed_epochs;
  else if (o->stagnation > 0 && st->iter_since_best >= o->stagnation) st->halted = 1;
}

/* ---- pheromone update (aco_cuda_kernels.cu:145-155, 592-668) ------------- */

void oracle_evaporate(OracleColony *o) {
  if (!o->st.found) return;
  const uint8_t d = o->q_evap, qmin = 0;
  size_t nn = (size_t)(o->n + 1) * (o->n + 1);
  for (size_t e = 0; e < nn; ++e) {
    uint8_t q = o->tau[e];
    o->tau[e] = (q > qmin + d) ? (uint8_t)(q - d) : qmin;
  }
  size_t nc = (size_t)(o->n + 1) * o->k;
  for (size_t e = 0; e < nc; ++e) {
    uint8_t q = o->tau_cand[e];
    o->tau_cand[e] = (q > qmin + d) ? (uint8_t)(q - d) : qmin;
  }
}

static void patch_cand(OracleColony *o, int u, int v) {
  for (int t = 0; t < o->k; ++t)
    if ((int)(o->rec[(size_t)u * o->k + t] & 0xFFFFu) == v)
      o->tau_cand[(size_t)u * o->k + t] = o->tau[(size_t)u * (o->n + 1) + v];
}

/* Per edge in tour order: new = quantize(dequantize(tau[u][v]) + delta);
 * tau[u][v] = max(tau[u][v], new); tau[v][u] = max(tau[v][u], new)
 * (kernel_deposit_flat_solution, aco_cuda_kernels.cu:652-667). */
void oracle_deposit_tour(OracleColony *o, const int32_t *tour, int len, float delta) {
  const size_t W = (size_t)o->n + 1;
  for (int p = 0; p + 1 < len; ++p) {
    int u = tour[p], v = tour[p + 1];
    if ((unsigned)u > (unsigned)o->n || (unsigned)v > (unsigned)o->n) continue;
    uint8_t a = o->tau[(size_t)u * W + v];
    uint8_t nq = oracle_quantize(o, o->deq[a] + delta);
    if (o->tau[(size_t)u * W + v] < nq) o->tau[(size_t)u * W + v] = nq;
    if (o->tau[(size_t)v * W + u] < nq) o->tau[(size_t)v * W + u] = nq;
    patch_cand(o, u, v);
    patch_cand(o, v, u);
  }
}

/* The reference's per-route deposits, each route in turn and its edges in
 * order (the K threads of kernel_deposit_solution / kernel_deposit_flat_solution,
 * aco_cuda_kernels.cu:592-668, one after the other).  step_major: node pos
 * of route v is routes[(off_v + pos) * m + ant], off_v = sum over u < v of
 * (lens[ant*K + u] - 1) (:603-612); else routes[v * (n+1) + pos] with lens[v]
 * (:650-656).  Every route starts from the depot (prev = 0).  A position
 * outside the buffer (steps rows, or K rows of n+1) or a node outside 0..n
 * reads as "no node": edges touching it deposit nothing. */
void oracle_deposit_routes(OracleColony *o, const int32_t *routes, const int32_t *lens, int K,
                           int m, long long steps, int ant, float delta, int step_major) {
  const size_t W = (size_t)o->n + 1;
  const long long limit = step_major ? steps : (long long)K * (long long)W;
  long long off = 0;
  for (int v = 0; v < K; ++v) {
    const int len = step_major ? lens[(size_t)ant * K + v] : lens[v];
    const long long base = step_major ? off : (long long)v * (long long)W;
    int prev = 0;
    for (int pos = 1; pos < len; ++pos) {
      const long long e = base + pos;
      int cur = -1;
      if (e >= 0 && e < limit) cur = step_major ? routes[e * m + ant] : routes[e];
      if ((unsigned)cur > (unsigned)o->n) cur = -1;
      if (prev >= 0 && cur >= 0) {
        const uint8_t a = o->tau[(size_t)prev * W + cur];
        const uint8_t nq = oracle_quantize(o, o->deq[a] + delta);
        if (o->tau[(size_t)prev * W + cur] < nq) o->tau[(size_t)prev * W + cur] = nq;
        if (o->tau[(size_t)cur * W + prev] < nq) o->tau[(size_t)cur * W + prev] = nq;
        patch_cand(o, prev, cur);
        patch_cand(o, cur, prev);
      }
      prev = cur;
    }
    off += (long long)len - 1;
  }
}

/* kernel_evaporate_tau (aco_cuda_kernels.cu:145-155) with any delta and
 * q_min: every byte of the matrix, q > q_min + delta ? q - delta : q_min. */
void oracle_evaporate_levels(OracleColony *o, uint8_t delta, uint8_t q_min) {
  const size_t nn = (size_t)(o->n + 1) * (o->n + 1);
  for (size_t e = 0; e < nn; ++e) {
    const uint8_t q = o->tau[e];
    o->tau[e] = (q > q_min + delta) ? (uint8_t)(q - delta) : q_min;
  }
}

void oracle_deposit_iter(OracleColony *o) {
  if (!o->st.found) return;
  int b = o->st.best_ant;
  oracle_deposit_tour(o, o->tours + (size_t)b * o->stride, o->len[b], o->st.dep_iter);
}

void oracle_deposit_best(OracleColony *o) {
  if (!o->st.found) return;
  oracle_deposit_tour(o, o->best_tour, o->st.best_len, o->st.dep_best);
}

/* One epoch of aco_cuda.cu:331-429: construct, best, evaporate, two deposits. */
void oracle_epochs(OracleColony *o, int count, int nthreads) {
  for (int e = 0; e < count; ++e) {
    if (o->st.halted) break;
    oracle_construct(o, nthreads);
    oracle_select_best(o);
    oracle_evaporate(o);
    oracle_deposit_iter(o);
    oracle_deposit_best(o);
  }
}

void oracle_get_state(const OracleColony *o, OracleState *st) { *st = o->st; }

int oracle_get_best(const OracleColony *o, int32_t *tour, int32_t *len, int64_t *cost) {
  if (o->st.global_best_cost == INT64_MAX) return 0;
  if (tour) memcpy(tour, o->best_tour, sizeof(int32_t) * (size_t)o->st.best_len);
  if (len) *len = o->st.best_len;
  if (cost) *cost = o->st.global_best_cost;
  return 1;
}

int64_t oracle_route_cost(const OracleColony *o, const int32_t *tour, int len) {
  int64_t c = 0;
  for (int p = 0; p + 1 < len; ++p) c += edge_cost(o, tour[p], tour[p + 1]);
  return c;
}
