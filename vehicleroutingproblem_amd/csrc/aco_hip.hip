/*
 * aco_hip.hip - host side of the C ABI declared in include/aco_hip.h.
 *
 * Replaces the CUDA host driver aco_vrp_cuda_with_capacity
 * (src/cuda/aco_cuda.cu:188-454): the context owns every device pool, every
 * launch is checked, the iteration-best / global-best / stagnation logic runs
 * on the device (no per-epoch device-to-host copy of ant summaries or routes),
 * and epochs are launched in chunks, optionally replayed from a hipGraph.
 */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <vector>

#include "aco_hip.h"
#include "aco_host.h"
#include "aco_inline.h"
#include "aco_kernels.h"
#include "aco_ctx.h"

/* ---- helpers --------------------------------------------------------------- */

int set_err(AcoHipContext *c, int code, const char *fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static size_t r16(size_t x) { return (x + 15) & ~(size_t)15; }
static size_t r256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T>
static int dalloc(AcoHipContext *c, T **p, size_t bytes) {
  bytes = r256(bytes ? bytes : 1) + 256;
  HIPC(c, hipMalloc((void **)p, bytes));
  HIPC(c, hipMemset(*p, 0, bytes));
  return ACO_HIP_OK;
}

static void free_pools(AcoHipContext *c) {
  void *ptrs[] = {c->d_inst, c->d_col,   c->d_state,  c->d_row_inst, c->d_xy,   c->d_dem,
                  c->d_dscore, c->d_rec, c->d_pher,   c->d_tours,    c->d_len,  c->d_unvis,
                  c->d_best, c->d_cost,  c->d_deq,    c->d_thr,      c->d_taupow, c->d_eta,
                  c->d_scratch_tour, c->d_hist, c->d_stats, c->d_cidx, c->d_score, c->d_done, c->d_slot,
                  c->d_fscore, c->d_ceta, c->d_rules};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->d_inst = nullptr; c->d_col = nullptr; c->d_state = nullptr; c->d_row_inst = nullptr;
  c->d_xy = nullptr; c->d_dem = nullptr; c->d_dscore = nullptr; c->d_rec = nullptr;
  c->d_pher = nullptr; c->d_tours = nullptr; c->d_len = nullptr; c->d_unvis = nullptr;
  c->d_best = nullptr; c->d_cost = nullptr; c->d_deq = nullptr; c->d_thr = nullptr;
  c->d_taupow = nullptr; c->d_eta = nullptr; c->d_scratch_tour = nullptr; c->d_hist = nullptr;
  c->d_stats = nullptr; c->d_cidx = nullptr; c->d_score = nullptr; c->d_done = nullptr;
  c->d_slot = nullptr; c->d_fscore = nullptr; c->d_ceta = nullptr; c->d_rules = nullptr;
  if (c->d_colmap) (void)hipFree(c->d_colmap);
  c->d_colmap = nullptr;
  if (c->d_feas) (void)hipFree(c->d_feas);
  if (c->d_item_off) (void)hipFree(c->d_item_off);
  if (c->d_slot_off) (void)hipFree(c->d_slot_off);
  if (c->d_boff) (void)hipFree(c->d_boff);
  c->d_feas = nullptr;
  c->d_item_off = c->d_slot_off = c->d_boff = nullptr;
  for (hipStream_t st : c->bstreams) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : c->bevents) (void)hipEventDestroy(ev);
  c->bstreams.clear();
  c->bevents.clear();
  c->buckets.clear();
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  c->loaded = false;
}

/* Row stride S of the all-pairs weight matrix and length of the per-ant
 * state array: n+1 rounded up to 16. */
static int fs_stride_of(int n1) { return (n1 + 15) & ~15; }

/* LDS bytes of one construct workgroup for an instance of n1 nodes (layout
 * of k_construct and k_construct_pair): staged candidate rows (+64 padding
 * entries), depot scores and (unless b8) coordinates; the computed-fallback
 * tables (taupow, and eta unless eta_len == 0) when the weight matrix is
 * off; per wave a state array of S u16 and a tour buffer of `stride` u16,
 * or with `pair` two state arrays (of S bytes each with b8) and (unless
 * `ring`: the pair then stores its tours in 128-byte chunks from registers)
 * two tour buffers, and a 128-byte store sink. */
static size_t lds_for(int n1, int k, int eta_len, int W, int stage, int fs_on, int stride,
                      int pair, int ring = 0, int b8 = 0) {
  const size_t state = (size_t)fs_stride_of(n1) * (b8 ? 1 : 2);
  const size_t tour = r16((size_t)stride * 2);
  size_t b = (size_t)W * ((pair ? 2 * state + (ring ? 0 : 2 * tour) : state + tour) + 128);
  if (!fs_on) b += 256 * 4 + r16((size_t)eta_len * 4);
  if (stage) {
    const size_t ent = (size_t)n1 * k + 64;
    b += r16(ent * 2) + r16(ent * 4) + r16((size_t)n1 * 4) + (b8 ? 0 : r16((size_t)n1 * 8));
  }
  return b;
}


/* ---- timing ---------------------------------------------------------------- */

enum { T_CONSTRUCT, T_BEST, T_EVAP, T_DEPOSIT, T_KNN };

static hipEvent_t get_event(AcoHipContext *c) {
  if (!c->free_events.empty()) {
    hipEvent_t e = c->free_events.back();
    c->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static int resolve_timing(AcoHipContext *c) {
  if (c->pending.empty()) return ACO_HIP_OK;
  HIPC(c, hipStreamSynchronize(c->stream));
  for (TimingRec &r : c->pending) {
    float ms = 0.0f;
    HIPC(c, hipEventElapsedTime(&ms, r.a, r.b));
    for (hipEvent_t e : r.more) {
      float m2 = 0.0f;
      HIPC(c, hipEventElapsedTime(&m2, r.a, e));
      ms = std::max(ms, m2);
      c->free_events.push_back(e);
    }
    switch (r.kind) {
      case T_CONSTRUCT: c->t.construct_ms += ms; c->t.construct_launches++; break;
      case T_BEST: c->t.best_ms += ms; c->t.best_launches++; break;
      case T_EVAP: c->t.evaporate_ms += ms; c->t.evaporate_launches++; break;
      case T_DEPOSIT: c->t.deposit_ms += ms; c->t.deposit_launches++; break;
      default: c->t.knn_ms += ms; c->t.knn_launches++; break;
    }
    c->free_events.push_back(r.a);
    c->free_events.push_back(r.b);
  }
  c->pending.clear();
  return ACO_HIP_OK;
}

struct Timed {
  AcoHipContext *c;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(AcoHipContext *cc, int k) : c(cc), kind(k) {
    if (c->timing) {
      a = get_event(c);
      b = get_event(c);
      if (a && b) (void)hipEventRecord(a, c->stream);
    }
  }
  ~Timed() {
    if (c->timing && a && b) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({kind, a, b, {}});
      if (c->pending.size() >= 4096) (void)resolve_timing(c);
    }
  }
};

/* ---- context ---------------------------------------------------------------- */

extern "C" int aco_hip_create(AcoHipContext **out, int device) {
  if (!out) return ACO_HIP_ERR_INVALID_INPUT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return ACO_HIP_ERR_BACKEND;
  if (device < 0 || device >= count) return ACO_HIP_ERR_INVALID_INPUT;
  AcoHipContext *c = new AcoHipContext();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ACO_HIP_ERR_BACKEND;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    c->num_cu = cus;
  /* device address of the staging array of the allocation-free entry point
   * (looked up once per device; a failure only disables that entry point) */
  (void)noalloc_attach_device(device);
  *out = c;
  return ACO_HIP_OK;
}

extern "C" void aco_hip_destroy(AcoHipContext *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (TimingRec &r : c->pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
    for (hipEvent_t e : r.more) (void)hipEventDestroy(e);
  }
  for (hipEvent_t e : c->free_events) (void)hipEventDestroy(e);
  free_pools(c);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char *aco_hip_last_error(const AcoHipContext *c) { return c ? c->err : "null context"; }

extern "C" int aco_hip_sync(AcoHipContext *c) {
  if (!c) return ACO_HIP_ERR_INVALID_INPUT;
  HIPC(c, hipStreamSynchronize(c->stream));
  return ACO_HIP_OK;
}

static int env_int(const char *name, int dflt) {
  const char *s = getenv(name);
  return (s && *s) ? atoi(s) : dflt;
}

/* Ants per workgroup and LDS staging of a bucket for the given ants per
 * wavefront (pair) and state width (b8: byte states).
 * Staging preference: the requested wavefronts per workgroup (W_req,
 * clamped to 1..8, 1..16 with byte states, and the ant count), staged rows
 * with the eta table, staged
 * rows without it, then (unless staging was forced) global rows, then fewer
 * wavefronts per workgroup. */
static int shape_for(AcoHipContext *c, LaunchBucket &lb, int W_req, int want_stage, int pair,
                     int b8) {
  int kmax = 1;
  for (int i : lb.insts) kmax = std::max(kmax, c->inst[i].k);
  const int apw = pair ? 2 : 1;
  const int wmax = b8 ? 16 : 8;
  const int wcap = std::max(1, std::min(wmax, (c->prm.ants + apw - 1) / apw));
  int W = std::max(1, std::min(wcap, W_req));
  int stage = want_stage == 0 ? 0 : 1;
  int eta_in_lds = 1;
  int ring = 0;
  size_t need = 0;
  /* pairs keep their tours in LDS when the workgroups a CU holds at once
   * (the bucket's share of the device, at most 32 wavefronts) all fit with
   * them; otherwise in registers, stored out in 128-byte chunks */
  auto recompute = [&]() {
    auto need_with = [&](int rg) {
      size_t b = 0;
      for (int i : lb.insts)
        b = std::max(b, lds_for(c->inst_n1[i], c->inst[i].k, eta_in_lds ? c->eta_len : 0, W,
                                stage, c->fs_on, c->max_stride, pair, rg, b8));
      return b;
    };
    ring = 0;
    need = need_with(0);
    if (pair) {
      const long long groups = (long long)lb.cols.size() * lb.lanes * ((c->prm.ants + 2 * W - 1) / (2 * W));
      const long long per_cu = std::max(1LL, std::min((groups + c->num_cu - 1) / c->num_cu,
                                                      (long long)(32 / W)));
      if (c->ring_mode > 0 || (c->ring_mode < 0 && (long long)need * per_cu > (long long)LDS_LIMIT)) {
        ring = 1;
        need = need_with(1);
      }
    }
  };
  recompute();
  if (need > LDS_LIMIT) {
    eta_in_lds = 0;
    recompute();
  }
  /* automatic staging: staged rows cap a CU at LDS_LIMIT / need workgroups;
   * when the bucket's workgroups need more than one such round over the
   * device, rows read from global memory (L2-resident, no LDS to share) let
   * many more wavefronts run side by side, which hides each decision's
   * latency (n=1000, 32 colonies of 256 ants: 1.09 ms per epoch against
   * 2.47 ms staged) */
  if (want_stage < 0 && need <= LDS_LIMIT) {
    const long long groups = (long long)lb.cols.size() * lb.lanes * ((c->prm.ants + W * apw - 1) / (W * apw));
    const long long per_round = (long long)c->num_cu * (long long)(LDS_LIMIT / need);
    if (groups > per_round) need = LDS_LIMIT + 1; /* take the global-row branch below */
  }
  if (need > LDS_LIMIT && want_stage < 0) {
    stage = 0;
    eta_in_lds = 1;
    recompute();
    if (need > LDS_LIMIT) {
      eta_in_lds = 0;
      recompute();
    }
  }
  while (need > LDS_LIMIT && W > 1) {
    --W;
    recompute();
  }
  if (need > LDS_LIMIT)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT,
                   "instance too large for the LDS layout (%zu bytes needed, stage=%d)", need,
                   stage);
  lb.W = W;
  lb.stage = stage;
  lb.eta_lds = eta_in_lds;
  lb.rows = (kmax + 15) / 16;
  lb.pair = pair;
  lb.ring = ring;
  lb.b8 = b8;
  lb.groups = (c->prm.ants + W * apw - 1) / (W * apw);
  lb.lds = need;
  return ACO_HIP_OK;
}

/* The bucket's shape.  want_pair: 1 two ants per wavefront when the
 * candidate lists allow it (k <= 32), 0 one, -1 automatic: two when the
 * launch has more ants than the device has SIMDs (the wavefronts then share
 * SIMDs and the launch is bound by the instructions issued per ant, which a
 * pair halves); one otherwise (each ant's chain of decisions is the bound).
 * Pairs keep byte states (bucket demands all <= 254, ACO_HIP_STATE8 not 0)
 * when asked to (ACO_HIP_STATE8=1) or when the u16 layout's workgroups
 * need more than one round of the device's LDS: byte states halve the
 * per-ant LDS and so raise the wavefronts resident at once.  Otherwise u16:
 * the byte states' capped room costs one instruction per decision. */
static int choose_shape(AcoHipContext *c, LaunchBucket &lb, int W_req, int want_stage,
                        int want_pair = -1) {
  int kmax = 1, dmax = 0;
  for (int i : lb.insts) {
    kmax = std::max(kmax, c->inst[i].k);
    dmax = std::max(dmax, c->inst_dmax[i]);
  }
  int pair = 0;
  if (kmax <= 32) {
    if (want_pair < 0) want_pair = c->pair_mode;
    if (want_pair > 0) pair = 1;
    else if (want_pair < 0)
      pair = (long long)lb.cols.size() * lb.lanes * c->prm.ants > 4LL * c->num_cu ? 1 : 0;
  }
  const bool b8_ok = pair && dmax <= 254 && c->b8_mode != 0;
  if (b8_ok && c->b8_mode > 0) return shape_for(c, lb, W_req, want_stage, pair, 1);
  LaunchBucket t16 = lb;
  const int rc = shape_for(c, t16, W_req, want_stage, pair, 0);
  if (b8_ok) {
    bool widen = rc != ACO_HIP_OK;
    if (!widen) {
      const long long groups = (long long)t16.cols.size() * t16.lanes * t16.groups;
      widen = groups > (long long)c->num_cu * (long long)(LDS_LIMIT / t16.lds);
    }
    LaunchBucket t8 = lb;
    if (widen && shape_for(c, t8, W_req, want_stage, pair, 1) == ACO_HIP_OK) {
      lb = t8;
      c->err[0] = 0;
      return ACO_HIP_OK;
    }
  }
  if (rc) return rc;
  lb = t16;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_load(AcoHipContext *c, const AcoHipInstance *in, int num,
                            const AcoHipParams *p) {
  if (!c || !in || num <= 0 || !p) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "null argument");
  if (p->cand_k < 1 || p->cand_k > 64)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "cand_k must be in [1, 64]");
  if (p->ants < 1 || p->ants > (1 << 20))
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "ants must be in [1, 2^20]");
  if (p->colonies < 1) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "colonies must be >= 1");
  /* colony indices, per-colony grids and state arrays are int-sized */
  if ((long long)num * p->colonies > (1LL << 20))
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT,
                   "instances x colonies must be at most 2^20 (got %lld)",
                   (long long)num * p->colonies);
  if (!(p->rho > 0.0f && p->rho < 1.0f))
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "rho must be in (0, 1)");
  if (!(p->tau0 > 0.0f)) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "tau0 must be > 0");
  if (p->knn_key != ACO_HIP_KNN_EXACT && p->knn_key != ACO_HIP_KNN_REF_F32)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "unknown knn_key");
  if (p->dist_mode != ACO_HIP_DIST_NINT && p->dist_mode != ACO_HIP_DIST_REAL)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "unknown dist_mode");
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipStreamSynchronize(c->stream));
  free_pools(c);
  c->prm = *p;
  c->num_inst = num;
  c->num_col = num * p->colonies;
  c->inst.assign(num, InstDesc{});
  c->inst_n1.assign(num, 0);
  c->inst_dmax.assign(num, 0);
  int dmax = 0;
  size_t node_off = 0, cand_off = 0, slot_off = 0;
  int row_base = 0;
  for (int i = 0; i < num; ++i) {
    const AcoHipInstance &a = in[i];
    /* K <= 65535 keeps the close rule's remaining capacity (K-v-1)*Q within
     * int32 on the device (65534 * 32767 < 2^31) */
    if (a.n < 1 || a.n > 65534 || a.K < 1 || a.K > 65535 || a.capacity < 1 ||
        a.capacity > 32767 || !a.x || !a.y || !a.demand)
      return set_err(c, ACO_HIP_ERR_INVALID_INPUT,
                     "instance %d: need 1 <= n <= 65534, 1 <= K <= 65535, "
                     "1 <= capacity <= 32767", i);
    if (a.demand[0] != 0)
      return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "instance %d: depot demand must be 0", i);
    int32_t xmin = a.x[0], xmax = a.x[0], ymin = a.y[0], ymax = a.y[0];
    long long total = 0;
    int dmax_i = 0;
    for (int j = 0; j <= a.n; ++j) {
      if (a.x[j] < 0 || a.x[j] > 32767 || a.y[j] < 0 || a.y[j] > 32767)
        return set_err(c, ACO_HIP_ERR_INVALID_INPUT,
                       "instance %d: coordinates must be integers in [0, 32767]", i);
      if (a.demand[j] < 0 || a.demand[j] > 32767)
        return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "instance %d: demand out of [0, 32767]", i);
      total += j ? a.demand[j] : 0;
      dmax_i = std::max(dmax_i, (int)a.demand[j]);
      xmin = std::min(xmin, a.x[j]); xmax = std::max(xmax, a.x[j]);
      ymin = std::min(ymin, a.y[j]); ymax = std::max(ymax, a.y[j]);
    }
    /* largest distance code of the instance: that of its bounding box's
     * diagonal (codes grow with the distance) */
    if (p->dist_mode == ACO_HIP_DIST_REAL) {
      const uint32_t dx = (uint32_t)(xmax - xmin), dy = (uint32_t)(ymax - ymin);
      dmax = std::max(dmax, (int)aco_real_eta_bin(aco_real_dist_s(dx * dx + dy * dy)));
    } else {
      dmax = std::max(dmax, (int)aco_nint_dist(xmin, ymin, xmax, ymax));
    }
    InstDesc &d = c->inst[i];
    d.scale = (float)std::max(1, std::max(xmax - xmin, ymax - ymin));
    d.n = a.n;
    d.K = a.K;
    if (a.cand_k < 0 || a.cand_k > 64)
      return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "instance %d: cand_k must be in [0, 64]", i);
    d.k = std::min(a.cand_k > 0 ? a.cand_k : p->cand_k, a.n);
    d.Q = a.capacity;
    d.node_off = (int)node_off;
    d.row_base = row_base;
    d.total_demand = total;
    d.cand_off = (long long)cand_off;
    d.slot_off = (long long)slot_off;
    slot_off += r256(((size_t)a.n + 1) * ((size_t)a.n + 1));
    d.seed = a.seed;
    d.fs_stride = fs_stride_of(a.n + 1);
    c->inst_n1[i] = a.n + 1;
    c->inst_dmax[i] = dmax_i;
    node_off += ((size_t)a.n + 1 + 3) & ~(size_t)3;
    /* 64 padding entries: every lane of a wavefront reads row n in bounds */
    cand_off += (((size_t)a.n + 1) * d.k + 64 + 7) & ~(size_t)7;
    row_base += a.n + 1;
  }
  c->nodes_total = node_off;
  c->slot_total = slot_off;
  c->rec_total = cand_off;
  c->total_rows = row_base;

  /* tables */
  aco_tau_tables_build(&c->tables, p->alpha, p->rho, p->tau0);
  c->eta_len = dmax + 1;
  c->eta_lut.assign((size_t)c->eta_len, 0.0f);
  if (p->dist_mode == ACO_HIP_DIST_REAL) aco_eta_table_build_real(c->eta_lut.data(), c->eta_len, p->beta);
  else aco_eta_table_build(c->eta_lut.data(), dmax, p->beta);

  /* colonies */
  c->col.assign(c->num_col, ColonyDesc{});
  size_t pher = 0, tours = 0, best = 0, score = 0, fs = 0;
  int ants = 0, max_stride = 0;
  size_t max_vec = 0, max_slots = 0, max_items = 0;
  for (int i = 0; i < num; ++i) {
    const InstDesc &d = c->inst[i];
    const size_t n1 = (size_t)d.n + 1;
    const int stride = (int)((((size_t)d.n + d.K + 1) + 63) & ~(size_t)63);
    for (int cc = 0; cc < p->colonies; ++cc) {
      ColonyDesc &cd = c->col[(size_t)i * p->colonies + cc];
      cd.inst = i;
      cd.colony_in_inst = cc;
      cd.m = p->ants;
      cd.stride = stride;
      cd.pher_off = (long long)pher;
      /* pheromone rows padded to the stride S (16-byte aligned rows) */
      const size_t mat = n1 * (size_t)d.fs_stride;
      cd.tcand_off = (long long)(pher + mat);
      cd.pher_bytes = (long long)(mat + r16(n1 * d.k));
      cd.tour_off = (long long)tours;
      cd.ant_off = ants;
      cd.best_off = (int)best;
      cd.score_off = (long long)score;
      cd.mat_bytes = (long long)mat;
      cd.fs_off = (long long)fs;
      fs += n1 * (size_t)d.fs_stride;
      score += (n1 * d.k + 64 + 3) & ~(size_t)3;
      max_slots = std::max(max_slots, n1 * d.k);
      max_items = std::max(max_items, n1 * (size_t)(d.fs_stride / 16));
      pher += r256((size_t)cd.pher_bytes);
      tours += (size_t)p->ants * stride;
      best += stride;
      ants += p->ants;
      max_stride = std::max(max_stride, stride);
      max_vec = std::max(max_vec, (size_t)cd.mat_bytes / 16);
    }
  }
  c->pher_total = pher;
  c->tours_total = tours;
  c->best_total = best;
  c->ants_total = ants;
  c->max_stride = max_stride;
  c->max_mat_vec = (int)std::min(max_vec, (size_t)1 << 30);
  c->max_slots = (int)std::min(max_slots, (size_t)1 << 30);
  c->max_fs_items = (int)std::min(max_items, (size_t)1 << 30);
  c->score_total = score;
  c->fs_total = fs;

  /* all-pairs fallback weights: on unless disabled or above 16 GiB */
  int want_fs = p->fallback_matrix;
  if (want_fs < 0) want_fs = env_int("ACO_HIP_FALLBACK_MATRIX", -1);
  const size_t fs_bytes = sizeof(float) * fs;
  if (want_fs < 0) want_fs = fs_bytes <= ((size_t)16 << 30) ? 1 : 0;
  c->fs_on = want_fs ? 1 : 0;

  /* launch shapes: instances are grouped by node count into size buckets
   * (n+1 up to 256, 512, 1024, ..., one bucket unless ACO_HIP_BUCKETS=0 merges
   * them); each bucket gets its own ants per workgroup and LDS staging and
   * its own construct launch */
  int max_k = 0;
  for (int i = 0; i < num; ++i) max_k = std::max(max_k, c->inst[i].k);
  c->rows = (max_k + 15) / 16;
  {
    const bool split = env_int("ACO_HIP_BUCKETS", 1) != 0;
    std::vector<int> cls(num);
    for (int i = 0; i < num; ++i) {
      int lo = 256;
      while (split && lo < c->inst_n1[i] && lo < (1 << 17)) lo <<= 1;
      cls[i] = split ? lo : 0;
    }
    std::vector<int> order(num);
    for (int i = 0; i < num; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cls[a] < cls[b]; });
    for (int q = 0; q < num; ++q) {
      const int i = order[q];
      if (c->buckets.empty() || cls[i] != c->buckets.back().n1_hi) {
        LaunchBucket nb;
        nb.n1_hi = cls[i];
        nb.n1_lo = 1 << 30;
        c->buckets.push_back(nb);
      }
      LaunchBucket &lb = c->buckets.back();
      lb.insts.push_back(i);
      lb.n1_lo = std::min(lb.n1_lo, c->inst_n1[i]);
      for (int cc = 0; cc < p->colonies; ++cc) lb.cols.push_back(i * p->colonies + cc);
    }
    /* lanes: the size class of a context with one class, when its launch
     * holds more than 8 ants per CU and at least two instances, is split into
     * two lanes (ACO_HIP_LANES forces the count for every class), each a
     * bucket with its own chain of construct, evaporation and deposit on its
     * own stream.  A lane's update kernels, which leave most of the device
     * idle, then overlap the other lane's construction.  Several classes
     * already overlap their chains (and more streams than the device's
     * hardware queues gain nothing), so they are not split. */
    const int lanes_env = env_int("ACO_HIP_LANES", -1);
    {
      std::vector<LaunchBucket> split_b;
      const bool one_class = c->buckets.size() == 1;
      for (const LaunchBucket &lb : c->buckets) {
        const int ni = (int)lb.insts.size();
        const long long ants = (long long)lb.cols.size() * p->ants;
        int L = lanes_env > 0 ? lanes_env : (one_class && ants > 8LL * c->num_cu ? 2 : 1);
        L = std::max(1, std::min(L, ni));
        for (int l = 0; l < L; ++l) {
          LaunchBucket nb;
          nb.n1_hi = lb.n1_hi;
          nb.n1_lo = 1 << 30;
          nb.lanes = L;
          nb.lane = l;
          for (int q = l * ni / L; q < (l + 1) * ni / L; ++q) {
            const int i = lb.insts[q];
            nb.insts.push_back(i);
            nb.n1_lo = std::min(nb.n1_lo, c->inst_n1[i]);
            for (int cc = 0; cc < p->colonies; ++cc) nb.cols.push_back(i * p->colonies + cc);
          }
          split_b.push_back(nb);
        }
      }
      c->buckets.swap(split_b);
    }
    int off = 0;
    for (LaunchBucket &lb : c->buckets) {
      int hi = 0;
      for (int i : lb.insts) hi = std::max(hi, c->inst_n1[i]);
      lb.n1_hi = hi; /* the range actually present */
      lb.col_off = off;
      off += (int)lb.cols.size();
    }
  }
  c->pair_mode = env_int("ACO_HIP_PAIR", -1);
  c->ring_mode = env_int("ACO_HIP_TOUR_RING", -1);
  c->b8_mode = env_int("ACO_HIP_STATE8", -1);
  c->xcd_mode = env_int("ACO_HIP_XCD", -1);
  const int W0 = p->ants_per_group > 0 ? p->ants_per_group : env_int("ACO_HIP_ANTS_PER_GROUP", 4);
  int want_stage = p->stage_lds;
  if (want_stage < 0) want_stage = env_int("ACO_HIP_STAGE_LDS", -1);
  for (LaunchBucket &lb : c->buckets) {
    int rc = choose_shape(c, lb, W0, want_stage);
    if (rc) return rc;
  }
  /* epochs as evaporate + multi-workgroup select/deposit while both tours
   * and their position tables fit in LDS */
  size_t sd = 0;
  for (int i = 0; i < num; ++i) sd = std::max(sd, sel_dep_lds(c->inst_n1[i], c->max_stride));
  c->sd_eta = c->eta_len <= 8192;
  if (c->sd_eta) sd += r16((size_t)c->eta_len * 4);
  c->sd_lds = (sd <= 131072 && c->max_stride <= 65536) ? sd : 0;
  if (env_int("ACO_HIP_FUSED_UPDATE", 0)) c->sd_lds = 0; /* force the fused update kernel */
  /* force the context-wide update kernels after all buckets (instead of one
   * construct/evaporate/deposit chain per bucket) */
  c->serial_updates = env_int("ACO_HIP_SERIAL_UPDATES", 0) != 0;
  /* workgroups per colony: enough for 256 pair events each, but about 2048 in
   * all (a batch of many colonies fills the chip with one or a few each, and
   * every workgroup stages both tours) */
  c->sd_blocks = std::max(
      1, std::min({64, (2 * c->max_stride + 255) / 256, (2048 + c->num_col - 1) / c->num_col}));
  c->W = c->buckets[0].W;
  c->stage = c->buckets[0].stage;
  c->eta_lds = c->buckets[0].eta_lds;
  c->groups = c->buckets[0].groups;
  c->lds_bytes = c->buckets[0].lds;

  /* device pools */
  int rc;
  if ((rc = dalloc(c, &c->d_inst, sizeof(InstDesc) * num))) return rc;
  if ((rc = dalloc(c, &c->d_col, sizeof(ColonyDesc) * c->num_col))) return rc;
  if ((rc = dalloc(c, &c->d_state, sizeof(AcoHipColonyState) * c->num_col))) return rc;
  if ((rc = dalloc(c, &c->d_hist, sizeof(HistEntry) * HIST * c->num_col))) return rc;
  if ((rc = dalloc(c, &c->d_stats, sizeof(unsigned long long) * ACO_STATS * c->num_col)))
    return rc;
  if ((rc = dalloc(c, &c->d_done, sizeof(int) * c->num_col))) return rc;
  if ((rc = dalloc(c, &c->d_rules, sizeof(RunRules)))) return rc;
  if ((rc = dalloc(c, &c->d_colmap, sizeof(int) * c->num_col))) return rc;
  if ((rc = dalloc(c, &c->d_feas, sizeof(int) * c->num_col))) return rc;
  if ((rc = dalloc(c, &c->d_item_off, sizeof(long long) * (c->num_col + 1)))) return rc;
  if ((rc = dalloc(c, &c->d_slot_off, sizeof(long long) * (c->num_col + 1)))) return rc;
  {
    std::vector<long long> io(c->num_col + 1, 0), so(c->num_col + 1, 0);
    for (int q = 0; q < c->num_col; ++q) {
      const ColonyDesc &cd = c->col[q];
      const InstDesc &d = c->inst[cd.inst];
      const long long n1 = d.n + 1;
      io[q + 1] = io[q] + (c->fs_on ? n1 * (d.fs_stride / 16) : cd.mat_bytes / 16);
      so[q + 1] = so[q] + n1 * d.k;
    }
    c->flat_total = io[c->num_col] + so[c->num_col];
    HIPC(c, hipMemcpy(c->d_item_off, io.data(), sizeof(long long) * io.size(),
                      hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->d_slot_off, so.data(), sizeof(long long) * so.size(),
                      hipMemcpyHostToDevice));
    /* the same offsets per size bucket, over the bucket's colony list */
    std::vector<long long> bo;
    for (LaunchBucket &lb : c->buckets) {
      const int nc = (int)lb.cols.size();
      lb.off_at = bo.size();
      long long ia = 0, sa = 0;
      std::vector<long long> bi(nc + 1, 0), bs(nc + 1, 0);
      for (int q = 0; q < nc; ++q) {
        const int col = lb.cols[q];
        ia += io[col + 1] - io[col];
        sa += so[col + 1] - so[col];
        bi[q + 1] = ia;
        bs[q + 1] = sa;
      }
      lb.flat_total = ia + sa;
      lb.sd_blocks = std::max(1, std::min({64, (2 * c->max_stride + 255) / 256,
                                           (2048 + nc - 1) / std::max(nc, 1)}));
      bo.insert(bo.end(), bi.begin(), bi.end());
      bo.insert(bo.end(), bs.begin(), bs.end());
    }
    if ((rc = dalloc(c, &c->d_boff, sizeof(long long) * bo.size()))) return rc;
    HIPC(c, hipMemcpy(c->d_boff, bo.data(), sizeof(long long) * bo.size(), hipMemcpyHostToDevice));
  }
  {
    std::vector<int> cm;
    for (const LaunchBucket &lb : c->buckets) cm.insert(cm.end(), lb.cols.begin(), lb.cols.end());
    HIPC(c, hipMemcpy(c->d_colmap, cm.data(), sizeof(int) * cm.size(), hipMemcpyHostToDevice));
  }
  for (size_t b = 1; b < c->buckets.size(); ++b) {
    hipStream_t st = nullptr;
    HIPC(c, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    c->bstreams.push_back(st);
  }
  for (size_t b = 0; b < c->buckets.size(); ++b) {
    hipEvent_t ev = nullptr;
    HIPC(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    c->bevents.push_back(ev);
  }
  if ((rc = dalloc(c, &c->d_row_inst, sizeof(int) * (size_t)c->total_rows))) return rc;
  /* +16 entries: 16-node vector reads of the last instance stay in bounds */
  if ((rc = dalloc(c, &c->d_xy, sizeof(int2) * (c->nodes_total + 16)))) return rc;
  if ((rc = dalloc(c, &c->d_dem, sizeof(uint16_t) * c->nodes_total))) return rc;
  if ((rc = dalloc(c, &c->d_dscore, sizeof(float) * c->nodes_total))) return rc;
  if ((rc = dalloc(c, &c->d_rec, sizeof(uint32_t) * c->rec_total))) return rc;
  if ((rc = dalloc(c, &c->d_cidx, sizeof(uint16_t) * c->rec_total))) return rc;
  if ((rc = dalloc(c, &c->d_ceta, sizeof(float) * c->rec_total))) return rc;
  if ((rc = dalloc(c, &c->d_slot, c->slot_total))) return rc;
  if ((rc = dalloc(c, &c->d_score, sizeof(float) * c->score_total))) return rc;
  if (c->fs_on && (rc = dalloc(c, &c->d_fscore, sizeof(float) * c->fs_total))) return rc;
  if ((rc = dalloc(c, &c->d_pher, c->pher_total))) return rc;
  if ((rc = dalloc(c, &c->d_tours, sizeof(int32_t) * c->tours_total))) return rc;
  if ((rc = dalloc(c, &c->d_len, sizeof(int32_t) * (size_t)ants))) return rc;
  if ((rc = dalloc(c, &c->d_unvis, sizeof(int32_t) * (size_t)ants))) return rc;
  if ((rc = dalloc(c, &c->d_cost, sizeof(long long) * (size_t)ants))) return rc;
  if ((rc = dalloc(c, &c->d_best, sizeof(int32_t) * c->best_total))) return rc;
  if ((rc = dalloc(c, &c->d_scratch_tour, sizeof(int32_t) * (size_t)max_stride))) return rc;
  if ((rc = dalloc(c, &c->d_deq, sizeof(float) * 256))) return rc;
  if ((rc = dalloc(c, &c->d_thr, sizeof(float) * 256))) return rc;
  if ((rc = dalloc(c, &c->d_taupow, sizeof(float) * 256))) return rc;
  if ((rc = dalloc(c, &c->d_eta, sizeof(float) * (size_t)c->eta_len))) return rc;

  std::vector<int2> xy(c->nodes_total, make_int2(0, 0));
  std::vector<uint16_t> dem(c->nodes_total, 0);
  std::vector<int> row_inst((size_t)c->total_rows);
  for (int i = 0; i < num; ++i) {
    const InstDesc &d = c->inst[i];
    for (int j = 0; j <= d.n; ++j) {
      xy[(size_t)d.node_off + j] = make_int2(in[i].x[j], in[i].y[j]);
      dem[(size_t)d.node_off + j] = (uint16_t)in[i].demand[j];
      row_inst[(size_t)d.row_base + j] = i;
    }
  }
  HIPC(c, hipMemcpy(c->d_inst, c->inst.data(), sizeof(InstDesc) * num, hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_col, c->col.data(), sizeof(ColonyDesc) * c->num_col,
                    hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_row_inst, row_inst.data(), sizeof(int) * row_inst.size(),
                    hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_xy, xy.data(), sizeof(int2) * xy.size(), hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_dem, dem.data(), sizeof(uint16_t) * dem.size(), hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_deq, c->tables.deq, sizeof(float) * 256, hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_thr, c->tables.thr, sizeof(float) * 255, hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_taupow, c->tables.taupow, sizeof(float) * 256, hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_eta, c->eta_lut.data(), sizeof(float) * c->eta_lut.size(),
                    hipMemcpyHostToDevice));

  DevCtx &P = c->P;
  P.inst = c->d_inst;
  P.col = c->d_col;
  P.state = c->d_state;
  P.row_inst = c->d_row_inst;
  P.xy = c->d_xy;
  P.dem = c->d_dem;
  P.dscore = c->d_dscore;
  P.rec = c->d_rec;
  P.cidx = c->d_cidx;
  P.slot_of = c->d_slot;
  P.score = c->d_score;
  P.fscore = c->fs_on ? c->d_fscore : nullptr;
  P.feas = c->d_feas;
  P.ceta = c->d_ceta;
  P.pher = c->d_pher;
  P.tours = c->d_tours;
  P.len = c->d_len;
  P.cost = c->d_cost;
  P.unvis = c->d_unvis;
  P.best = c->d_best;
  P.deq = c->d_deq;
  P.thr = c->d_thr;
  P.taupow = c->d_taupow;
  P.eta = c->d_eta;
  P.eta_len = c->eta_len;
  P.eta_lds = c->eta_lds;
  P.dist_real = p->dist_mode == ACO_HIP_DIST_REAL ? 1 : 0;
  P.q_evap = c->tables.q_evap;
  P.qdep = p->Q;
  c->rules = RunRules{0, 0, 0, 0, 1e-3}; /* no stop rule outside a run */
  HIPC(c, hipMemcpy(c->d_rules, &c->rules, sizeof(RunRules), hipMemcpyHostToDevice));
  P.rules = c->d_rules;
  P.hist = (void *)c->d_hist;
  P.stats = nullptr;
  P.done = c->d_done;

#define ACO_VARIANTS(R, F)                                                                      \
  (const void *)k_construct<true, false, R, false, F>,                                         \
      (const void *)k_construct<false, false, R, false, F>,                                    \
      (const void *)k_construct<true, true, R, false, F>,                                      \
      (const void *)k_construct<false, true, R, false, F>,                                     \
      (const void *)k_construct<true, false, R, true, F>,                                      \
      (const void *)k_construct<false, false, R, true, F>
#define ACO_PAIR_VARIANTS(F, RG, B8)                                                            \
  (const void *)k_construct_pair<true, false, false, F, RG, B8>,                               \
      (const void *)k_construct_pair<false, false, false, F, RG, B8>,                          \
      (const void *)k_construct_pair<true, true, false, F, RG, B8>,                            \
      (const void *)k_construct_pair<false, true, false, F, RG, B8>,                           \
      (const void *)k_construct_pair<true, false, true, F, RG, B8>,                            \
      (const void *)k_construct_pair<false, false, true, F, RG, B8>
  const void *variants[] = {ACO_VARIANTS(1, true),  ACO_VARIANTS(2, true),  ACO_VARIANTS(3, true),
                            ACO_VARIANTS(4, true),  ACO_VARIANTS(1, false), ACO_VARIANTS(2, false),
                            ACO_VARIANTS(3, false), ACO_VARIANTS(4, false),
                            ACO_PAIR_VARIANTS(true, false, false), ACO_PAIR_VARIANTS(false, false, false),
                            ACO_PAIR_VARIANTS(true, true, false), ACO_PAIR_VARIANTS(false, true, false),
                            ACO_PAIR_VARIANTS(true, false, true), ACO_PAIR_VARIANTS(false, false, true),
                            ACO_PAIR_VARIANTS(true, true, true), ACO_PAIR_VARIANTS(false, true, true)};
#undef ACO_VARIANTS
#undef ACO_PAIR_VARIANTS
  for (const void *f : variants)
    HIPC(c, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_LIMIT));
  HIPC(c, hipFuncSetAttribute((const void *)k_select_deposit,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
  /* colony offsets (<= 16 KB) + taupow + an eta table of up to 64 KB */
  HIPC(c, hipFuncSetAttribute((const void *)k_evaporate_flat,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 98304));
  c->loaded = true;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_num_colonies(const AcoHipContext *c) { return c ? c->num_col : 0; }

extern "C" int aco_hip_num_buckets(const AcoHipContext *c) {
  return (c && c->loaded) ? (int)c->buckets.size() : 0;
}

static void bucket_info(const LaunchBucket &lb, AcoHipBucketInfo *info) {
  info->n1_min = lb.n1_lo;
  info->n1_max = lb.n1_hi;
  info->instances = (int)lb.insts.size();
  info->colonies = (int)lb.cols.size();
  info->ants_per_group = lb.W;
  info->stage_lds = lb.stage;
  info->lds_bytes = (int)lb.lds;
  info->tuned_ms = lb.tuned_ms;
  info->ants_per_wave = lb.pair ? 2 : 1;
  info->state_bytes = lb.pair && lb.b8 ? 1 : 2;
  info->lanes = lb.lanes;
}

extern "C" int aco_hip_get_bucket(const AcoHipContext *c, int b, AcoHipBucketInfo *info) {
  if (!c || !c->loaded || !info || b < 0 || b >= (int)c->buckets.size())
    return ACO_HIP_ERR_INVALID_INPUT;
  bucket_info(c->buckets[b], info);
  return ACO_HIP_OK;
}

static void drop_graph(AcoHipContext *c) {
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
}

static void sync_shape0(AcoHipContext *c) {
  c->W = c->buckets[0].W;
  c->stage = c->buckets[0].stage;
  c->eta_lds = c->buckets[0].eta_lds;
  c->groups = c->buckets[0].groups;
  c->lds_bytes = c->buckets[0].lds;
}

extern "C" int aco_hip_set_bucket_shape(AcoHipContext *c, int b, int apg, int stage) {
  return aco_hip_set_bucket_layout(c, b, apg, stage, 0);
}

extern "C" int aco_hip_set_bucket_layout(AcoHipContext *c, int b, int apg, int stage, int apw) {
  if (!c || !c->loaded) return ACO_HIP_ERR_INVALID_INPUT;
  if (b < 0 || b >= (int)c->buckets.size() || apg < 1 || apg > 16 || stage < -1 || stage > 1 ||
      apw < 0 || apw > 2)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "bad bucket shape request");
  LaunchBucket trial = c->buckets[b];
  int rc = choose_shape(c, trial, apg, stage, apw == 0 ? -1 : apw - 1);
  if (rc) return rc;
  trial.tuned_ms = 0.0f;
  c->buckets[b] = trial;
  sync_shape0(c);
  drop_graph(c);
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_shape(const AcoHipContext *c, int *cand_k, int *ants, int *stride,
                                 int *apg, int *stage) {
  if (!c || !c->loaded) return ACO_HIP_ERR_INVALID_INPUT;
  if (cand_k) *cand_k = c->inst[0].k;
  if (ants) *ants = c->prm.ants;
  if (stride) *stride = c->col[0].stride;
  if (apg) *apg = c->W;
  if (stage) *stage = c->stage;
  return ACO_HIP_OK;
}


/* ---- kernels ----------------------------------------------------------------- */

extern "C" int aco_hip_build_candidates(AcoHipContext *c) {
  NEED_LOADED(c);
  const int blocks = (c->total_rows * 64 + 255) / 256;
  {
    Timed tm(c, T_KNN);
    hipLaunchKernelGGL(k_knn, dim3(blocks), dim3(256), 0, c->stream, c->P, c->total_rows,
                       c->prm.knn_key, c->prm.depot_close_weight);
  }
  LAUNCHED(c, "k_knn");
  hipLaunchKernelGGL(k_build_slots, dim3(blocks), dim3(256), 0, c->stream, c->P, c->total_rows);
  LAUNCHED(c, "k_build_slots");
  return launch_refresh(c);
}

extern "C" int aco_hip_init_tau(AcoHipContext *c) {
  NEED_LOADED(c);
  const uint32_t q = c->tables.q_tau0;
  const uint32_t word = q | (q << 8) | (q << 16) | (q << 24);
  const size_t vecs = c->pher_total / 16;
  const int blocks = (int)std::min<size_t>((vecs + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_tau, dim3(std::max(blocks, 1)), dim3(256), 0, c->stream, c->d_pher,
                     c->pher_total, word);
  LAUNCHED(c, "k_fill_tau");
  int rc = launch_refresh(c);
  if (rc) return rc;
  std::vector<AcoHipColonyState> st(c->num_col);
  for (AcoHipColonyState &s : st) {
    memset(&s, 0, sizeof(s));
    s.best_ant = -1;
    s.iter_best_cost = INT64_MAX;
    s.global_best_cost = INT64_MAX;
  }
  HIPC(c, hipMemcpyAsync(c->d_state, st.data(), sizeof(AcoHipColonyState) * st.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemsetAsync(c->d_feas, 0, sizeof(int) * c->num_col, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return ACO_HIP_OK;
}

/* One launch of the construction variant that matches the bucket's shape:
 * ANTS = 1 (k_construct, ROWS = R) or 2 (k_construct_pair, tours in
 * registers with RING, byte states with B8), FSM = fallback weights from the
 * all-pairs weight matrix.  Launches of several colonies keep each colony's
 * workgroups on one XCD (the counter variants keep the plain order). */
template <int ANTS, int R, bool FSM, bool RING = false, bool B8 = false>
static void launch_variant(const AcoHipContext *c, const LaunchBucket &lb, hipStream_t st,
                           bool stats) {
  DevCtx P = c->P;
  P.eta_lds = lb.eta_lds;
  const int *cols = c->d_colmap + lb.col_off;
  dim3 grid((unsigned)(lb.cols.size() * lb.groups)), block((unsigned)(64 * lb.W));
  const bool xcd = c->xcd_mode < 0 ? lb.cols.size() > 1 : c->xcd_mode > 0;
#define ACO_GO(ST, SS, X)                                                                       \
  do {                                                                                          \
    if constexpr (ANTS == 2)                                                                    \
      hipLaunchKernelGGL((k_construct_pair<ST, SS, X, FSM, RING, B8>), grid, block, lb.lds, st, P, \
                         lb.groups, lb.W, cols);                                                \
    else                                                                                        \
      hipLaunchKernelGGL((k_construct<ST, SS, R, X, FSM>), grid, block, lb.lds, st, P,          \
                         lb.groups, lb.W, cols);                                                \
  } while (0)
  if (lb.stage && !stats && xcd)
    ACO_GO(true, false, true);
  else if (!stats && xcd)
    ACO_GO(false, false, true);
  else if (lb.stage && !stats)
    ACO_GO(true, false, false);
  else if (!stats)
    ACO_GO(false, false, false);
  else if (lb.stage)
    ACO_GO(true, true, false);
  else
    ACO_GO(false, true, false);
#undef ACO_GO
}

template <bool FSM>
static void launch_bucket_f(const AcoHipContext *c, const LaunchBucket &lb, hipStream_t st,
                            bool stats) {
  if (lb.pair) {
    if (lb.ring && lb.b8)
      launch_variant<2, 1, FSM, true, true>(c, lb, st, stats);
    else if (lb.b8)
      launch_variant<2, 1, FSM, false, true>(c, lb, st, stats);
    else if (lb.ring)
      launch_variant<2, 1, FSM, true>(c, lb, st, stats);
    else
      launch_variant<2, 1, FSM, false>(c, lb, st, stats);
    return;
  }
  switch (lb.rows) {
    case 1: launch_variant<1, 1, FSM>(c, lb, st, stats); break;
    case 2: launch_variant<1, 2, FSM>(c, lb, st, stats); break;
    case 3: launch_variant<1, 3, FSM>(c, lb, st, stats); break;
    default: launch_variant<1, 4, FSM>(c, lb, st, stats); break;
  }
}

static void launch_bucket(AcoHipContext *c, const LaunchBucket &lb, hipStream_t st, bool stats) {
  if (c->P.fscore)
    launch_bucket_f<true>(c, lb, st, stats);
  else
    launch_bucket_f<false>(c, lb, st, stats);
}

/* One construct launch per size bucket: bucket 0 on the context stream, the
 * others on their own streams forked from it and joined back, so the
 * buckets run side by side (also inside a captured graph). */
static int launch_construct(AcoHipContext *c) {
  const bool stats = c->P.stats != nullptr;
  const size_t nb = c->buckets.size();
  if (nb == 1) {
    Timed tm(c, T_CONSTRUCT);
    launch_bucket(c, c->buckets[0], c->stream, stats);
  } else {
    /* the buckets' launches run concurrently on their own streams; timed,
     * the phase lasts from bucket 0's start to the latest end, each end
     * recorded on its own stream (events around the fork and the join
     * would add their cross-stream waits to it) */
    HIPC(c, hipEventRecord(c->bevents[0], c->stream));
    for (size_t b = 1; b < nb; ++b) HIPC(c, hipStreamWaitEvent(c->bstreams[b - 1], c->bevents[0], 0));
    const bool timed = c->timing;
    TimingRec rec{T_CONSTRUCT, nullptr, nullptr, {}};
    if (timed) {
      rec.a = get_event(c);
      rec.b = get_event(c);
      if (!rec.a || !rec.b) return set_err(c, ACO_HIP_ERR_BACKEND, "hipEventCreate failed");
      HIPC(c, hipEventRecord(rec.a, c->stream));
    }
    launch_bucket(c, c->buckets[0], c->stream, stats);
    if (timed) HIPC(c, hipEventRecord(rec.b, c->stream));
    for (size_t b = 1; b < nb; ++b) {
      launch_bucket(c, c->buckets[b], c->bstreams[b - 1], stats);
      if (timed) {
        hipEvent_t e = get_event(c);
        if (!e) return set_err(c, ACO_HIP_ERR_BACKEND, "hipEventCreate failed");
        HIPC(c, hipEventRecord(e, c->bstreams[b - 1]));
        rec.more.push_back(e);
      }
    }
    for (size_t b = 1; b < nb; ++b) {
      HIPC(c, hipEventRecord(c->bevents[b], c->bstreams[b - 1]));
      HIPC(c, hipStreamWaitEvent(c->stream, c->bevents[b], 0));
    }
    if (timed) {
      c->pending.push_back(std::move(rec));
      if (c->pending.size() >= 4096) {
        const int rc = resolve_timing(c);
        if (rc) return rc;
      }
    }
  }
  LAUNCHED(c, "k_construct");
  return ACO_HIP_OK;
}

extern "C" int aco_hip_autotune(AcoHipContext *c, int launches, AcoHipBucketInfo *info,
                                int max_buckets) {
  NEED_LOADED(c);
  if (launches < 1) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "launches must be >= 1");
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipStreamSynchronize(c->stream));
  {
    std::vector<AcoHipColonyState> st(c->num_col);
    HIPC(c, hipMemcpy(st.data(), c->d_state, sizeof(AcoHipColonyState) * c->num_col,
                      hipMemcpyDeviceToHost));
    for (const AcoHipColonyState &x : st)
      if (x.halted) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "autotune needs running colonies");
  }
  hipEvent_t a = nullptr, z = nullptr;
  HIPC(c, hipEventCreate(&a));
  HIPC(c, hipEventCreate(&z));
  const bool stats = c->P.stats != nullptr;
  for (size_t b = 0; b < c->buckets.size(); ++b) {
    LaunchBucket best = c->buckets[b];
    float best_ms = -1.0f;
    for (int pair = 0; pair <= 1; ++pair)
    for (int stage = 1; stage >= 0; --stage)
      for (int W = 1; W <= 16; W <<= 1) {
        if (W * (pair + 1) > c->prm.ants + pair) continue;
        LaunchBucket trial = c->buckets[b];
        if (choose_shape(c, trial, W, stage, pair) != ACO_HIP_OK) continue;
        /* does not fit as asked (or no pairs for lists above 32) */
        if (trial.W != W || trial.stage != stage || trial.pair != pair) continue;
        /* a lane is timed alone, but runs beside its class's other lanes:
         * staged rows only where all of them fit one LDS round */
        if (trial.stage && trial.lanes > 1 &&
            (long long)trial.cols.size() * trial.lanes * trial.groups >
                (long long)c->num_cu * (long long)(LDS_LIMIT / trial.lds))
          continue;
        launch_bucket(c, trial, c->stream, stats);          /* warm caches */
        HIPC(c, hipEventRecord(a, c->stream));
        for (int r = 0; r < launches; ++r) launch_bucket(c, trial, c->stream, stats);
        HIPC(c, hipEventRecord(z, c->stream));
        LAUNCHED(c, "k_construct");
        HIPC(c, hipEventSynchronize(z));
        float ms = 0.0f;
        HIPC(c, hipEventElapsedTime(&ms, a, z));
        ms /= (float)launches;
        if (best_ms < 0.0f || ms < best_ms) {
          best_ms = ms;
          best = trial;
        }
      }
    if (best_ms >= 0.0f) {
      best.tuned_ms = best_ms;
      c->buckets[b] = best;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(z);
  sync_shape0(c);
  drop_graph(c);
  if (info)
    for (int b = 0; b < (int)c->buckets.size() && b < max_buckets; ++b)
      bucket_info(c->buckets[b], info + b);
  return ACO_HIP_OK;
}

static int launch_update(AcoHipContext *c, int mode) {
  {
    Timed tm(c, (mode & 1) ? T_BEST : T_DEPOSIT);
    hipLaunchKernelGGL(k_update, dim3(c->num_col), dim3(1024), 0, c->stream, c->P, mode);
  }
  LAUNCHED(c, "k_update");
  return ACO_HIP_OK;
}

static int launch_evap(AcoHipContext *c, int from_ants) {
  const int mat = c->fs_on ? c->max_fs_items : c->max_mat_vec;
  const int xm = std::max(1, std::min((mat + 255) / 256, 1024));
  const int xs = std::max(1, std::min((c->max_slots + 255) / 256, 512));
  {
    Timed tm(c, T_EVAP);
    hipLaunchKernelGGL(k_evaporate, dim3(xm + xs, c->num_col), dim3(256), 0, c->stream, c->P,
                       xm, from_ants);
  }
  LAUNCHED(c, "k_evaporate");
  return ACO_HIP_OK;
}

static int launch_evap_update(AcoHipContext *c) {
  const int mat = c->fs_on ? c->max_fs_items : c->max_mat_vec;
  const int xm = std::max(1, std::min((mat + 1023) / 1024, 512));
  const int xs = std::max(1, std::min((c->max_slots + 1023) / 1024, 256));
  {
    Timed tm(c, T_EVAP);
    hipLaunchKernelGGL(k_evap_update, dim3(xm + xs, c->num_col), dim3(1024), 0, c->stream, c->P,
                       xm);
  }
  LAUNCHED(c, "k_evap_update");
  return ACO_HIP_OK;
}

/* Flat evaporation of `ncol` colonies (offsets item_off/slot_off; colony
 * list cmap, or all colonies in order when null) on stream st. */
static void enqueue_evap_flat(AcoHipContext *c, hipStream_t st, const long long *item_off,
                              const long long *slot_off, int ncol, const int *cmap,
                              long long total) {
  /* the colony offsets (up to EVAP_LDS_COLONIES colonies), the taupow
   * table, and the eta table while it is at most 64 KB, in LDS.  Every
   * workgroup stages the tables once, so each walks a long item range: past
   * 8 KB of eta the grid is two workgroups of 512 threads per CU (what fits
   * the LDS; 16 wavefronts to cover the memory latency), else 16 of 256
   * (measured on the n=1000 workload over 32 seeds: 52 us per launch against
   * 64 us at 64 workgroups per CU and 66 us at 4; the 512-instance batch's
   * epoch is within 1.5% from 16 to 128). */
  const int eta_lds = c->eta_len <= 16384 ? 1 : 0;
  const bool big_eta = eta_lds && c->eta_len > 2048;
  const long long cap = (big_eta ? 2LL : 16LL) * c->num_cu;
  const int threads = big_eta ? 512 : 256;
  const long long blocks = std::max(1LL, std::min((total + threads - 1) / threads, cap));
  const size_t off_lds =
      ncol <= EVAP_LDS_COLONIES ? 2 * sizeof(long long) * ((size_t)ncol + 1) : 0;
  const size_t lds = off_lds + 256 * 4 + (eta_lds ? (size_t)c->eta_len * 4 : 0);
  hipLaunchKernelGGL(k_evaporate_flat, dim3((unsigned)blocks), dim3(threads), lds, st, c->P, item_off,
                     slot_off, ncol, eta_lds, cmap);
}

static int launch_evap_flat(AcoHipContext *c) {
  {
    Timed tm(c, T_EVAP);
    enqueue_evap_flat(c, c->stream, c->d_item_off, c->d_slot_off, c->num_col, nullptr,
                      c->flat_total);
  }
  LAUNCHED(c, "k_evaporate_flat");
  return ACO_HIP_OK;
}

static int launch_select_deposit(AcoHipContext *c) {
  {
    Timed tm(c, T_BEST);
    hipLaunchKernelGGL(k_select_deposit, dim3(c->sd_blocks, c->num_col), dim3(256), c->sd_lds,
                       c->stream, c->P, c->sd_eta, nullptr);
  }
  LAUNCHED(c, "k_select_deposit");
  return ACO_HIP_OK;
}

/* `count` epochs of a multi-bucket context as independent chains: every
 * bucket (a size class, or a lane of one) constructs, evaporates and
 * deposits its own colonies on its own stream (bucket 0 on the context
 * stream), forked from the context stream once and joined back after the
 * last epoch, so a chain runs ahead of the others as far as it is faster.  A
 * bucket's update kernels then run while other buckets construct.  Colonies
 * never share state, so the result is the same as with the context-wide
 * update kernels. */
static int launch_epoch_pipelined(AcoHipContext *c, int count) {
  const bool stats = c->P.stats != nullptr;
  const size_t nb = c->buckets.size();
  HIPC(c, hipEventRecord(c->bevents[0], c->stream));
  for (size_t b = 1; b < nb; ++b) HIPC(c, hipStreamWaitEvent(c->bstreams[b - 1], c->bevents[0], 0));
  for (int e = 0; e < count; ++e)
    for (size_t b = 0; b < nb; ++b) {
      const LaunchBucket &lb = c->buckets[b];
      hipStream_t st = b == 0 ? c->stream : c->bstreams[b - 1];
      const int nc = (int)lb.cols.size();
      const int *cmap = c->d_colmap + lb.col_off;
      /* the other lanes of a class start after lane 0's first construct, so
       * that the lanes' update phases fall in each other's constructions */
      const size_t b0 = b - (size_t)lb.lane;
      if (e == 0 && lb.lane > 0) HIPC(c, hipStreamWaitEvent(st, c->bevents[b0], 0));
      launch_bucket(c, lb, st, stats);
      if (e == 0 && lb.lane == 0 && lb.lanes > 1) HIPC(c, hipEventRecord(c->bevents[b0], st));
      enqueue_evap_flat(c, st, c->d_boff + lb.off_at, c->d_boff + lb.off_at + nc + 1, nc, cmap,
                        lb.flat_total);
      hipLaunchKernelGGL(k_select_deposit, dim3(lb.sd_blocks, nc), dim3(256), c->sd_lds, st,
                         c->P, c->sd_eta, cmap);
    }
  for (size_t b = 1; b < nb; ++b) {
    HIPC(c, hipEventRecord(c->bevents[b], c->bstreams[b - 1]));
    HIPC(c, hipStreamWaitEvent(c->stream, c->bevents[b], 0));
  }
  LAUNCHED(c, "k_construct/k_evaporate_flat/k_select_deposit");
  return ACO_HIP_OK;
}

int launch_refresh(AcoHipContext *c) {
  const int work = std::max(c->max_slots, c->fs_on ? c->max_fs_items : 0);
  const int xs = std::max(1, std::min((work + 255) / 256, 1024));
  hipLaunchKernelGGL(k_refresh_scores, dim3(xs, c->num_col), dim3(256), 0, c->stream, c->P);
  LAUNCHED(c, "k_refresh_scores");
  return ACO_HIP_OK;
}

extern "C" int aco_hip_construct(AcoHipContext *c) {
  NEED_LOADED(c);
  return launch_construct(c);
}
extern "C" int aco_hip_select_best(AcoHipContext *c) {
  NEED_LOADED(c);
  return launch_update(c, 1);
}
extern "C" int aco_hip_evaporate(AcoHipContext *c) {
  NEED_LOADED(c);
  return launch_evap(c, 0);
}
extern "C" int aco_hip_deposit_iter(AcoHipContext *c) {
  NEED_LOADED(c);
  return launch_update(c, 2);
}
extern "C" int aco_hip_deposit_best(AcoHipContext *c) {
  NEED_LOADED(c);
  return launch_update(c, 4);
}

static int launch_epochs(AcoHipContext *c, int count) {
  int rc;
  for (int e = 0; e < count; ++e) {
    /* construct, then evaporation (it reads feasibility from the ants),
     * then best selection + both deposits over many workgroups - one such
     * chain per size bucket when there are several (unless timing, which
     * measures each phase over all buckets, or ACO_HIP_SERIAL_UPDATES asks
     * for the context-wide order); instances whose tours do not fit the LDS
     * tables of that kernel use the fused evaporate + single-workgroup
     * update kernel instead */
    if (c->sd_lds && c->buckets.size() > 1 && !c->timing && !c->serial_updates) {
      if ((rc = launch_epoch_pipelined(c, count))) return rc;
      break;
    }
    if ((rc = launch_construct(c))) return rc;
    if (c->sd_lds) {
      if ((rc = launch_evap_flat(c))) return rc;
      if ((rc = launch_select_deposit(c))) return rc;
    } else {
      if ((rc = launch_evap_update(c))) return rc;
    }
  }
  return ACO_HIP_OK;
}

extern "C" int aco_hip_epochs(AcoHipContext *c, int count) {
  NEED_LOADED(c);
  if (count < 0) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "negative epoch count");
  return launch_epochs(c, count);
}

/* ---- epoch loop ---------------------------------------------------------------- */

static int capture_chunk(AcoHipContext *c, int chunk) {
  if (c->graph && c->graph_chunk == chunk) return ACO_HIP_OK;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const bool timing = c->timing;
  c->timing = false;
  int rc = launch_epochs(c, chunk);
  c->timing = timing;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_err(c, ACO_HIP_ERR_BACKEND, "hipStreamEndCapture: %s",
                                      hipGetErrorString(e));
  e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_err(c, ACO_HIP_ERR_BACKEND, "hipGraphInstantiate: %s",
                                      hipGetErrorString(e));
  c->graph_chunk = chunk;
  return ACO_HIP_OK;
}

/* A colony cost in the units of the printed cost: the nint length, or the
 * real length in real mode. */
static double printed_cost(const AcoHipContext *c, long long cost) {
  return c->P.dist_real ? (double)cost * ACO_HIP_REAL_COST_UNIT : (double)cost;
}

static void print_progress(int epoch, double elapsed, double max_rt, int since, int max_stag,
                           long long best, double unit) {
  char tr[32], er[32], bc[64];
  if (max_rt > 0.0) snprintf(tr, sizeof(tr), "%.1fs", std::max(0.0, max_rt - elapsed));
  else snprintf(tr, sizeof(tr), "senza limite");
  if (max_stag > 0) {
    int rem = max_stag - since;
    snprintf(er, sizeof(er), "%d", rem > 0 ? rem : 0);
    if (rem <= 0) snprintf(tr, sizeof(tr), "0.0s");
  } else {
    snprintf(er, sizeof(er), "senza limite");
  }
  if (best != INT64_MAX) snprintf(bc, sizeof(bc), "%.3f", (double)best * unit);
  else snprintf(bc, sizeof(bc), "non disponibile");
  printf("[progress][hip] epoca=%d | epoche_rimanenti_prima_stop=%s | tempo_trascorso=%.1fs | "
         "tempo_rimanente=%s | best_cost=%s\n",
         epoch, er, elapsed, tr, bc);
  fflush(stdout);
}

extern "C" int aco_hip_run(AcoHipContext *c, const AcoHipRunConfig *cfg, AcoHipRunStats *stats) {
  NEED_LOADED(c);
  if (!cfg) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "null run config");
  HIPC(c, hipSetDevice(c->device));
  RunRules R{};
  R.fixed_epochs = cfg->fixed_epochs > 0 ? cfg->fixed_epochs : 0;
  R.stagnation = cfg->stagnation_epochs > 0 ? cfg->stagnation_epochs : 0;
  R.max_epochs = cfg->max_epochs > 0 ? cfg->max_epochs : 0;
  R.min_rel = cfg->min_rel_improvement > 0.0 ? cfg->min_rel_improvement : 1e-3;
  if (!R.fixed_epochs && !R.stagnation && !R.max_epochs && !(cfg->timeout_seconds > 0))
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT,
                   "no stop rule: set a timeout, stagnation, fixed or max epochs");
  /* the rules live in device memory, so an epoch graph captured by an earlier run with
   * the same chunk length stays valid */
  c->rules = R;
  HIPC(c, hipMemcpyAsync(c->d_rules, &c->rules, sizeof(RunRules), hipMemcpyHostToDevice,
                         c->stream));
  int chunk = cfg->chunk_epochs > 0 ? std::min(cfg->chunk_epochs, HIST) : 8;
  const bool use_graph = cfg->use_graph && !c->timing;
  std::vector<AcoHipColonyState> st(c->num_col);
  std::vector<HistEntry> hist(HIST);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(st.data(), c->d_state, sizeof(AcoHipColonyState) * c->num_col,
                    hipMemcpyDeviceToHost));
  /* a new run re-evaluates the halted flag under its own stop rules */
  for (AcoHipColonyState &s : st) {
    if (R.fixed_epochs > 0) s.halted = s.epoch >= R.fixed_epochs;
    else s.halted = R.stagnation > 0 && s.iter_since_best >= R.stagnation;
    if (R.max_epochs > 0 && s.epoch >= R.max_epochs) s.halted = 1;
  }
  HIPC(c, hipMemcpy(c->d_state, st.data(), sizeof(AcoHipColonyState) * c->num_col,
                    hipMemcpyHostToDevice));
  const int start_epoch = st[0].epoch;
  if (cfg->print_progress) {
    printf("HIP Solver starting... (N=%d, K=%d, M=%d)\n", c->inst[0].n, c->inst[0].K,
           c->prm.ants);
    fflush(stdout);
  }
  const double t0 = now_s();
  double next_beat = cfg->progress_interval_seconds > 0 ? t0 + cfg->progress_interval_seconds : 0;
  int printed_upto = start_epoch;
  int rc;
  bool halted0 = true;
  for (const AcoHipColonyState &s : st) halted0 = halted0 && s.halted;
  for (; !halted0;) {
    double t = now_s();
    if (cfg->timeout_seconds > 0.0 && (t - t0) > cfg->timeout_seconds) break;
    if (use_graph) {
      if ((rc = capture_chunk(c, chunk))) return rc;
      HIPC(c, hipGraphLaunch(c->graph, c->stream));
    } else {
      if ((rc = launch_epochs(c, chunk))) return rc;
    }
    HIPC(c, hipMemcpyAsync(st.data(), c->d_state, sizeof(AcoHipColonyState) * c->num_col,
                           hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemcpyAsync(hist.data(), c->d_hist, sizeof(HistEntry) * HIST,
                           hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    const double now = now_s();
    const double elapsed = now - t0;
    if (cfg->print_progress) {
      for (int e = printed_upto + 1; e <= st[0].epoch; ++e) {
        const HistEntry &h = hist[(e - 1) % HIST];
        if (h.epoch != e) continue;
        const int reached_time = cfg->timeout_seconds > 0.0 && elapsed >= cfg->timeout_seconds;
        const int reached_stag = h.halted;
        if (e == 1 || e % 10 == 0 || h.improved || reached_time || reached_stag)
          print_progress(e, elapsed, cfg->timeout_seconds, h.iter_since_best,
                         cfg->stagnation_epochs, h.global_best,
                         c->P.dist_real ? ACO_HIP_REAL_COST_UNIT : 1.0);
      }
      printed_upto = st[0].epoch;
    }
    if (next_beat > 0 && now >= next_beat) {
      fprintf(stderr, "[hip] elapsed %.1fs, iter %d, best %.3f\n", elapsed, st[0].epoch,
              st[0].global_best_cost == INT64_MAX ? 0.0 : printed_cost(c, st[0].global_best_cost));
      next_beat = now + cfg->progress_interval_seconds;
    }
    bool all_halted = true;
    for (const AcoHipColonyState &s : st) all_halted = all_halted && s.halted;
    if (all_halted) break;
  }
  const double secs = now_s() - t0;
  if (stats) {
    stats->epochs = st[0].epoch - start_epoch;
    stats->seconds = secs;
    long long tours = 0;
    for (const AcoHipColonyState &s : st) tours += (long long)(s.epoch - start_epoch) * c->prm.ants;
    stats->ant_tours_per_s = secs > 0 ? (double)tours / secs : 0.0;
  }
  return resolve_timing(c);
}

/* ---- downloads / uploads ------------------------------------------------------ */

static int check_inst(AcoHipContext *c, int i) {
  if (i < 0 || i >= c->num_inst) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "bad instance %d", i);
  return ACO_HIP_OK;
}
static int check_col(AcoHipContext *c, int i) {
  if (i < 0 || i >= c->num_col) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "bad colony %d", i);
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_candidates(AcoHipContext *c, int inst, int32_t *idx, int32_t *dist,
                                      float *eta) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_inst(c, inst))) return rc;
  const InstDesc &d = c->inst[inst];
  const size_t cnt = (size_t)(d.n + 1) * d.k;
  std::vector<uint32_t> rec(cnt);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(rec.data(), c->d_rec + d.cand_off, cnt * 4, hipMemcpyDeviceToHost));
  for (size_t e = 0; e < cnt; ++e) {
    const uint32_t j = rec[e] & 0xFFFFu, dd = rec[e] >> 16;
    if (idx) idx[e] = (int32_t)j;
    if (dist) dist[e] = (int32_t)dd;
    if (eta) eta[e] = j ? c->eta_lut[dd] : 0.0f;
  }
  /* a learned heuristic replaces the distance term: report the one in use */
  if (eta && d.learned)
    HIPC(c, hipMemcpy(eta, c->d_ceta + d.cand_off, cnt * sizeof(float), hipMemcpyDeviceToHost));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_set_candidate_eta(AcoHipContext *c, int inst, const float *eta) {
  NEED_LOADED(c);
  if (inst < 0 || inst >= c->num_inst)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "instance %d out of range", inst);
  InstDesc &d = c->inst[inst];
  const size_t cnt = ((size_t)d.n + 1) * d.k;
  if (eta) {
    for (size_t e = 0; e < cnt; ++e)
      if (!(eta[e] >= 0.0f && eta[e] <= 3.0e38f))
        return set_err(c, ACO_HIP_ERR_INVALID_INPUT,
                       "candidate heuristic %zu is not a finite value >= 0", e);
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipMemcpy(c->d_ceta + d.cand_off, eta, sizeof(float) * cnt, hipMemcpyHostToDevice));
  } else {
    HIPC(c, hipStreamSynchronize(c->stream));
  }
  return ctx_set_learned(c, inst, eta ? 1 : 0);
}

int ctx_set_learned(AcoHipContext *c, int inst, int learned) {
  InstDesc &d = c->inst[inst];
  d.learned = learned ? 1 : 0;
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(c->d_inst + inst, &d, sizeof(InstDesc), hipMemcpyHostToDevice));
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  int rc = launch_refresh(c);
  if (rc) return rc;
  HIPC(c, hipStreamSynchronize(c->stream));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_tau(AcoHipContext *c, int colony, uint8_t *tau, uint8_t *tau_cand) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  const ColonyDesc &cd = c->col[colony];
  const InstDesc &d = c->inst[cd.inst];
  const size_t n1 = (size_t)d.n + 1;
  HIPC(c, hipStreamSynchronize(c->stream));
  if (tau)
    HIPC(c, hipMemcpy2D(tau, n1, c->d_pher + cd.pher_off, (size_t)d.fs_stride, n1, n1,
                        hipMemcpyDeviceToHost));
  if (tau_cand)
    HIPC(c, hipMemcpy(tau_cand, c->d_pher + cd.tcand_off, n1 * d.k, hipMemcpyDeviceToHost));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_set_tau(AcoHipContext *c, int colony, const uint8_t *tau) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  if (!tau) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "null tau");
  const ColonyDesc &cd = c->col[colony];
  const InstDesc &d = c->inst[cd.inst];
  const size_t n1 = (size_t)d.n + 1, cnt = n1 * d.k;
  std::vector<uint32_t> rec(cnt);
  std::vector<uint8_t> tc(cnt);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(rec.data(), c->d_rec + d.cand_off, cnt * 4, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n1; ++i)
    for (int t = 0; t < d.k; ++t) tc[i * d.k + t] = tau[i * n1 + (rec[i * d.k + t] & 0xFFFFu)];
  HIPC(c, hipMemcpy2D(c->d_pher + cd.pher_off, (size_t)d.fs_stride, tau, n1, n1, n1,
                      hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->d_pher + cd.tcand_off, tc.data(), cnt, hipMemcpyHostToDevice));
  HIPC(c, hipMemset(c->d_feas + colony, 0, sizeof(int)));
  int rc2 = launch_refresh(c);
  if (rc2) return rc2;
  HIPC(c, hipStreamSynchronize(c->stream));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_scores(AcoHipContext *c, int colony, float *score) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  if (!score) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "null score");
  const ColonyDesc &cd = c->col[colony];
  const InstDesc &d = c->inst[cd.inst];
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(score, c->d_score + cd.score_off, sizeof(float) * (size_t)(d.n + 1) * d.k,
                    hipMemcpyDeviceToHost));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_fallback_weights(AcoHipContext *c, int colony, float *weights) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  if (!weights) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "null weights");
  if (!c->fs_on) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "fallback weight matrix is off");
  const ColonyDesc &cd = c->col[colony];
  const InstDesc &d = c->inst[cd.inst];
  const size_t n1 = (size_t)d.n + 1;
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy2D(weights, n1 * sizeof(float), c->d_fscore + cd.fs_off,
                      (size_t)d.fs_stride * sizeof(float), n1 * sizeof(float), n1,
                      hipMemcpyDeviceToHost));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_tours(AcoHipContext *c, int colony, int32_t *tours, int32_t *len,
                                 int64_t *cost, int32_t *unvisited) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  const ColonyDesc &cd = c->col[colony];
  HIPC(c, hipStreamSynchronize(c->stream));
  if (tours)
    HIPC(c, hipMemcpy(tours, c->d_tours + cd.tour_off, sizeof(int32_t) * (size_t)cd.m * cd.stride,
                      hipMemcpyDeviceToHost));
  if (len)
    HIPC(c, hipMemcpy(len, c->d_len + cd.ant_off, sizeof(int32_t) * cd.m, hipMemcpyDeviceToHost));
  if (cost)
    HIPC(c, hipMemcpy(cost, c->d_cost + cd.ant_off, sizeof(int64_t) * cd.m, hipMemcpyDeviceToHost));
  if (unvisited)
    HIPC(c, hipMemcpy(unvisited, c->d_unvis + cd.ant_off, sizeof(int32_t) * cd.m,
                      hipMemcpyDeviceToHost));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_colony(AcoHipContext *c, int colony, AcoHipColonyState *st) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  if (!st) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "null state");
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(st, c->d_state + colony, sizeof(*st), hipMemcpyDeviceToHost));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_set_epoch(AcoHipContext *c, int colony, int epoch) {
  AcoHipColonyState st;
  int rc = aco_hip_get_colony(c, colony, &st);
  if (rc) return rc;
  st.epoch = epoch;
  HIPC(c, hipMemcpy(c->d_state + colony, &st, sizeof(st), hipMemcpyHostToDevice));
  HIPC(c, hipMemset(c->d_feas + colony, 0, sizeof(int))); /* no feasible ant seen at this epoch */
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_best(AcoHipContext *c, int colony, int32_t *tour, int32_t *len,
                                int64_t *cost) {
  AcoHipColonyState st;
  int rc = aco_hip_get_colony(c, colony, &st);
  if (rc) return rc;
  if (st.global_best_cost == INT64_MAX)
    return set_err(c, ACO_HIP_ERR_NO_SOLUTION, "no feasible tour found");
  if (tour)
    HIPC(c, hipMemcpy(tour, c->d_best + c->col[colony].best_off, sizeof(int32_t) * st.best_len,
                      hipMemcpyDeviceToHost));
  if (len) *len = st.best_len;
  if (cost) *cost = st.global_best_cost;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_deposit_tour(AcoHipContext *c, int colony, const int32_t *tour, int len,
                                    float delta) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  const ColonyDesc &cd = c->col[colony];
  const InstDesc &d = c->inst[cd.inst];
  if (!tour || len < 2 || len > cd.stride)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "tour length must be in [2, %d]", cd.stride);
  for (int p = 0; p < len; ++p)
    if (tour[p] < 0 || tour[p] > d.n)
      return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "tour node out of range at %d", p);
  /* the kernel gives each edge its own thread: only a depot-delimited tour
   * visiting each customer at most once has no other shared entries than
   * those it orders itself (single-customer and empty routes) */
  if (tour[0] != 0 || tour[len - 1] != 0)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "tour must start and end at the depot");
  {
    std::vector<unsigned char> seen((size_t)d.n + 1, 0);
    for (int p = 0; p < len; ++p) {
      const int v = tour[p];
      if (v == 0) continue;
      if (seen[v]) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "customer %d repeats at %d", v, p);
      seen[v] = 1;
    }
  }
  HIPC(c, hipMemcpyAsync(c->d_scratch_tour, tour, sizeof(int32_t) * len, hipMemcpyHostToDevice,
                         c->stream));
  hipLaunchKernelGGL(k_deposit_tour, dim3(1), dim3(1024), 0, c->stream, c->P, colony,
                     c->d_scratch_tour, len, delta);
  LAUNCHED(c, "k_deposit_tour");
  HIPC(c, hipStreamSynchronize(c->stream));
  return ACO_HIP_OK;
}

int check_route_cost_args(AcoHipContext *c, int inst, const int32_t *tours, const int32_t *len,
                          int count, int stride, const int64_t *cost) {
  int rc;
  if ((rc = check_inst(c, inst))) return rc;
  if (count <= 0) return ACO_HIP_OK;
  if (!tours || !len || !cost || stride <= 0)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "bad route-cost arguments");
  const int n = c->inst[inst].n;
  for (int t = 0; t < count; ++t) {
    if (len[t] < 0 || len[t] > stride)
      return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "tour %d length out of range", t);
    for (int p = 0; p < len[t]; ++p)
      if (tours[(size_t)t * stride + p] < 0 || tours[(size_t)t * stride + p] > n)
        return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "tour %d node out of range", t);
  }
  return ACO_HIP_OK;
}

extern "C" int aco_hip_route_costs(AcoHipContext *c, int inst, const int32_t *tours,
                                   const int32_t *len, int count, int stride, int64_t *cost) {
  NEED_LOADED(c);
  int rc = check_route_cost_args(c, inst, tours, len, count, stride, cost);
  if (rc || count <= 0) return rc;
  int32_t *dt = nullptr, *dl = nullptr;
  long long *dc = nullptr;
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipMalloc(&dt, sizeof(int32_t) * (size_t)count * stride));
  HIPC(c, hipMalloc(&dl, sizeof(int32_t) * count));
  HIPC(c, hipMalloc(&dc, sizeof(long long) * count));
  HIPC(c, hipMemcpy(dt, tours, sizeof(int32_t) * (size_t)count * stride, hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(dl, len, sizeof(int32_t) * count, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_route_costs, dim3((count * 64 + 255) / 256), dim3(256), 0, c->stream, c->P,
                     inst, dt, dl, count, stride, dc);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(cost, dc, sizeof(long long) * count, hipMemcpyDeviceToHost);
  (void)hipFree(dt);
  (void)hipFree(dl);
  (void)hipFree(dc);
  if (e != hipSuccess)
    return set_err(c, ACO_HIP_ERR_BACKEND, "route costs: %s", hipGetErrorString(e));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_tables(AcoHipContext *c, float *deq, float *thr, float *taupow,
                                  float *eta_lut, int *eta_len, uint8_t *q_tau0, uint8_t *q_evap) {
  NEED_LOADED(c);
  if (deq) memcpy(deq, c->tables.deq, sizeof(float) * 256);
  if (thr) memcpy(thr, c->tables.thr, sizeof(float) * 255);
  if (taupow) memcpy(taupow, c->tables.taupow, sizeof(float) * 256);
  if (eta_lut) memcpy(eta_lut, c->eta_lut.data(), sizeof(float) * c->eta_lut.size());
  if (eta_len) *eta_len = c->eta_len;
  if (q_tau0) *q_tau0 = c->tables.q_tau0;
  if (q_evap) *q_evap = c->tables.q_evap;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_set_stats(AcoHipContext *c, int enable) {
  NEED_LOADED(c);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemset(c->d_stats, 0, sizeof(unsigned long long) * ACO_STATS * c->num_col));
  c->P.stats = enable ? c->d_stats : nullptr;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_stats(AcoHipContext *c, int colony, AcoHipStats *st) {
  NEED_LOADED(c);
  int rc;
  if ((rc = check_col(c, colony))) return rc;
  if (!st) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "null stats");
  unsigned long long v[ACO_STATS];
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(v, c->d_stats + (size_t)colony * ACO_STATS, sizeof(v), hipMemcpyDeviceToHost));
  st->decisions = v[0];
  st->fallbacks = v[1];
  st->fallback_cycles = v[2];
  st->ant_cycles = v[3];
  st->ants = v[4];
  st->staging_cycles = v[5];
  st->updates = v[6];
  st->evap_ticks = v[7];
  st->select_ticks = v[8];
  st->deposit_ticks = v[9];
  int khz = 0;
  HIPC(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
  st->wall_clock_hz = (uint64_t)khz * 1000u;
  return ACO_HIP_OK;
}

extern "C" int aco_hip_set_timing(AcoHipContext *c, int enable) {
  if (!c) return ACO_HIP_ERR_INVALID_INPUT;
  int rc = resolve_timing(c);
  if (rc) return rc;
  c->timing = enable != 0;
  memset(&c->t, 0, sizeof(c->t));
  return ACO_HIP_OK;
}

extern "C" int aco_hip_get_timing(AcoHipContext *c, AcoHipTiming *t) {
  if (!c || !t) return ACO_HIP_ERR_INVALID_INPUT;
  int rc = resolve_timing(c);
  if (rc) return rc;
  *t = c->t;
  return ACO_HIP_OK;
}

/* ---- one-call solver ------------------------------------------------------------ */

extern "C" int aco_vrp_hip_with_capacity(int n, int K, int capacity, int m, const int32_t *x,
                                         const int32_t *y, const int32_t *demand, double alpha,
                                         double beta, double rho, double tau0, double Q,
                                         unsigned int seed, int dist_mode, int32_t *best_tour,
                                         int32_t *best_len, int64_t *best_cost) {
  if (n <= 0 || K <= 0 || !x || !y || !demand || !best_tour || !best_len || !best_cost)
    return ACO_HIP_ERR_INVALID_INPUT;
  AcoRuntimeConfig rt;
  aco_runtime_config_load_env(&rt);
  if (m == 0) m = 256; /* aco_cuda.cu:206 */
  AcoHipContext *c = nullptr;
  int rc = aco_hip_create(&c, env_int("ACO_HIP_DEVICE", 0));
  if (rc) {
    fprintf(stderr, "HIP device unavailable\n");
    return rc;
  }
  AcoHipInstance in{n, K, capacity, x, y, demand, seed};
  AcoHipParams p{};
  p.cand_k = std::min(env_int("ACO_HIP_CAND_K", 32), n); /* aco_cuda.cu:198-199 */
  p.ants = m;
  p.colonies = 1;
  p.alpha = (float)alpha;
  p.beta = (float)beta;
  p.rho = (float)rho;
  p.tau0 = (float)tau0;
  p.Q = (float)Q;
  p.depot_close_weight = 2.0f;
  p.knn_key = env_int("ACO_HIP_KNN_KEY", ACO_HIP_KNN_EXACT);
  p.dist_mode = dist_mode;
  p.ants_per_group = env_int("ACO_HIP_ANTS_PER_GROUP", 0);
  p.stage_lds = env_int("ACO_HIP_STAGE_LDS", -1);
  if ((rc = aco_hip_load(c, &in, 1, &p)) || (rc = aco_hip_build_candidates(c)) ||
      (rc = aco_hip_init_tau(c))) {
    fprintf(stderr, "HIP solver setup failed: %s\n", aco_hip_last_error(c));
    aco_hip_destroy(c);
    return rc;
  }
  AcoHipRunConfig cfg{};
  cfg.timeout_seconds = rt.timeout_seconds;
  cfg.stagnation_epochs = rt.stagnation_epochs;
  cfg.min_rel_improvement = rt.min_rel_improvement;
  cfg.fixed_epochs = rt.fixed_epochs;
  cfg.max_epochs = env_int("ACO_HIP_MAX_EPOCHS", 0);
  if (!cfg.fixed_epochs && !cfg.stagnation_epochs && !(cfg.timeout_seconds > 0) &&
      !cfg.max_epochs)
    cfg.stagnation_epochs = 100; /* bounded default when no stop rule is configured */
  cfg.progress_interval_seconds = rt.progress_interval_seconds;
  cfg.print_progress = 1;
  cfg.chunk_epochs = env_int("ACO_HIP_CHUNK", 8);
  cfg.use_graph = env_int("ACO_HIP_GRAPH", 1);
  AcoHipRunStats stats{};
  rc = aco_hip_run(c, &cfg, &stats);
  if (rc) {
    fprintf(stderr, "HIP solver failed: %s\n", aco_hip_last_error(c));
    aco_hip_destroy(c);
    return rc;
  }
  rc = aco_hip_get_best(c, 0, best_tour, best_len, best_cost);
  aco_hip_destroy(c);
  return rc;
}
