/*
 * aco_ctx.h - the context object behind the C ABI (internal; shared by the
 * solver host code aco_hip.hip and the learned-prior host code aco_prior.hip).
 */
#ifndef ACO_CTX_H
#define ACO_CTX_H

#include <hip/hip_runtime.h>

#include <vector>

#include "aco_hip.h"
#include "aco_host.h"
#include "aco_kernels.h"

#define LDS_LIMIT 163840u
#define HIST ACO_HIST

struct TimingRec {
  int kind;
  hipEvent_t a, b;
  /* extra end events of a phase of concurrent launches (b is the first):
   * the phase lasts from a to the latest of them */
  std::vector<hipEvent_t> more;
};

/* One size class of instances and its construct launch shape. */
struct LaunchBucket {
  int n1_lo = 0, n1_hi = 0;  /* node-count range of the class */
  std::vector<int> insts;    /* instances of the class */
  std::vector<int> cols;     /* their colonies */
  int col_off = 0;           /* offset of `cols` in d_colmap */
  int W = 4, stage = 1, eta_lds = 1, rows = 1, groups = 1;
  int pair = 0;              /* two ants per wavefront (k_construct_pair) */
  int ring = 0;              /* pairs: tours in registers, stored in 128-byte chunks */
  int b8 = 0;                /* pairs: byte states (every demand of the bucket <= 254) */
  int lanes = 1, lane = 0;   /* lanes of its size class (concurrent chains), and its index */
  size_t lds = 0;
  float tuned_ms = 0.0f;     /* construct time measured by the autotuner, 0 = not tuned */
  size_t off_at = 0;         /* its item offsets at d_boff[off_at], slot offsets after them */
  long long flat_total = 0;  /* its evaporation items + slots */
  int sd_blocks = 1;         /* select/deposit workgroups per colony */
};

struct AcoHipContext {
  int device = 0;
  int num_cu = 256; /* compute units of the device */
  hipStream_t stream = nullptr;
  char err[512] = {0};
  bool loaded = false;
  AcoHipParams prm{};
  int num_inst = 0, num_col = 0, W = 4, stage = 1, groups = 1, total_rows = 0, eta_lds = 1;
  int rows = 1;        /* 16-lane rows of the widest candidate list (k_construct ROWS) */
  int pair_mode = -1;  /* ACO_HIP_PAIR: -1 automatic, 0 never, 1 whenever k <= 32 */
  int ring_mode = -1;  /* ACO_HIP_TOUR_RING: pairs' tours -1 automatic, 0 in LDS, 1 in registers */
  int b8_mode = -1;    /* ACO_HIP_STATE8: pairs' byte states -1 / 1 when demands allow, 0 never */
  int xcd_mode = -1;   /* ACO_HIP_XCD: XCD-aware workgroup order -1 with several colonies, 0, 1 */
  int fs_on = 0;       /* all-pairs fallback weight matrix allocated */
  size_t sd_lds = 0;   /* dynamic LDS of k_select_deposit, 0 = fused k_evap_update path */
  int sd_eta = 0;      /* k_select_deposit stages the eta table in LDS */
  int sd_blocks = 1;   /* workgroups of k_select_deposit per colony */
  size_t lds_bytes = 0;
  std::vector<InstDesc> inst;
  std::vector<ColonyDesc> col;
  std::vector<int> inst_n1;
  std::vector<int> inst_dmax; /* largest customer demand per instance */
  AcoTauTables tables{};
  std::vector<float> eta_lut;
  int eta_len = 0;
  size_t pher_total = 0, tours_total = 0, best_total = 0, rec_total = 0, nodes_total = 0;
  size_t score_total = 0, slot_total = 0, fs_total = 0;
  int ants_total = 0, max_stride = 0, max_mat_vec = 0, max_slots = 0, max_fs_items = 0;
  /* device pools */
  InstDesc *d_inst = nullptr;
  ColonyDesc *d_col = nullptr;
  AcoHipColonyState *d_state = nullptr;
  int *d_row_inst = nullptr;
  int2 *d_xy = nullptr;
  uint16_t *d_dem = nullptr;
  float *d_dscore = nullptr;
  uint32_t *d_rec = nullptr;
  uint16_t *d_cidx = nullptr;
  uint8_t *d_slot = nullptr;
  float *d_score = nullptr;
  float *d_fscore = nullptr;
  float *d_ceta = nullptr;
  uint8_t *d_pher = nullptr;
  int32_t *d_tours = nullptr, *d_len = nullptr, *d_unvis = nullptr, *d_best = nullptr;
  long long *d_cost = nullptr;
  float *d_deq = nullptr, *d_thr = nullptr, *d_taupow = nullptr, *d_eta = nullptr;
  int32_t *d_scratch_tour = nullptr;
  HistEntry *d_hist = nullptr;
  unsigned long long *d_stats = nullptr;
  int *d_done = nullptr;
  RunRules *d_rules = nullptr; /* the current run's stop rules (read by the select kernels) */
  RunRules rules{};            /* host copy of *d_rules */
  DevCtx P{};
  /* timing */
  bool timing = false;
  bool serial_updates = false; /* ACO_HIP_SERIAL_UPDATES=1: context-wide update kernels */
  std::vector<TimingRec> pending;
  std::vector<hipEvent_t> free_events;
  AcoHipTiming t{};
  /* graph */
  hipGraphExec_t graph = nullptr;
  int graph_chunk = 0;
  /* size buckets: one construct launch each, forked over streams */
  std::vector<LaunchBucket> buckets;
  int *d_colmap = nullptr;
  int *d_feas = nullptr;                   /* per colony: 1 + last epoch with a feasible ant */
  long long *d_item_off = nullptr, *d_slot_off = nullptr; /* flat evaporation prefix offsets */
  long long *d_boff = nullptr; /* the same per size bucket: item then slot offsets */
  long long flat_total = 0;
  std::vector<hipStream_t> bstreams; /* streams of buckets 1.. */
  std::vector<hipEvent_t> bevents;   /* fork event, then one join event per extra bucket */
};

/* ---- helpers shared by the host files ---------------------------------------- */

int set_err(AcoHipContext *c, int code, const char *fmt, ...);
/* Recompute every colony's candidate scores (and all-pairs weights). */
int launch_refresh(AcoHipContext *c);
/* Switch instance `inst` to (learned != 0) or from the per-slot heuristic
 * already in d_ceta, and refresh the scores (stream-ordered, synchronous). */
int ctx_set_learned(AcoHipContext *c, int inst, int learned);
/* Argument checks shared by the route-cost entry points (0 = OK). */
int check_route_cost_args(AcoHipContext *c, int inst, const int32_t *tours, const int32_t *len,
                          int count, int stride, const int64_t *cost);
/* Looks up the device address of the allocation-free entry point's staging
 * array for `device` (current); 0 on success. */
int noalloc_attach_device(int device);

#define HIPC(ctx, call)                                                                         \
  do {                                                                                          \
    hipError_t e__ = (call);                                                                    \
    if (e__ != hipSuccess)                                                                      \
      return set_err(ctx, ACO_HIP_ERR_BACKEND, "%s failed at %s:%d: %s", #call, __FILE__,      \
                     __LINE__, hipGetErrorString(e__));                                         \
  } while (0)

#define LAUNCHED(ctx, name)                                                                     \
  do {                                                                                          \
    hipError_t e__ = hipGetLastError();                                                         \
    if (e__ != hipSuccess)                                                                      \
      return set_err(ctx, ACO_HIP_ERR_BACKEND, "launch of %s failed: %s", name,                 \
                     hipGetErrorString(e__));                                                   \
  } while (0)

#define NEED_LOADED(c)                                                                          \
  do {                                                                                          \
    if (!(c)) return ACO_HIP_ERR_INVALID_INPUT;                                                 \
    if (!(c)->loaded) return set_err(c, ACO_HIP_ERR_INVALID_INPUT, "no instance loaded");       \
  } while (0)

#endif
