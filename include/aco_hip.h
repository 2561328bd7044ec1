/*
 * aco_hip.h - C ABI of the gfx950 ACO-CVRP kernels.
 *
 * This header replaces the CUDA-only kernel boundary of the reference,
 * include/aco_cuda_kernels.h:82-181 (seven __global__ kernels and their
 * `launch_*` host launchers, C++ linkage, caller-owned cudaMalloc buffers,
 * void returns).  Here:
 *   - every entry point is extern "C" and includable from C11;
 *   - no HIP or CUDA type appears in a signature (opaque context, plain
 *     pointers and sizes);
 *   - the context owns every device buffer; the caller owns host buffers;
 *   - every call returns an int status whose values 0..4 mirror AcoStatus
 *     (include/aco.h:82-88); aco_hip_last_error() gives the message.
 *
 * A context holds a batch of instances (one instance is a batch of one).
 * Each instance runs `colonies` independent colonies of `ants` ants.  Colony
 * c of instance i has index i*colonies + c in the per-colony calls.
 *
 * Data conventions (identical on host, device and in the CPU oracle):
 *   - nodes 0..n, node 0 is the depot; integer coordinates in [0, 32767];
 *   - distances (AcoHipParams.dist_mode): TSPLIB EUC_2D nearest-integer
 *     nint(sqrt(dx^2+dy^2)) (ACO_HIP_DIST_NINT), or the reference's unrounded
 *     distance as the float nearest to sqrt(dx^2+dy^2) (ACO_HIP_DIST_REAL,
 *     src/common/instance_parser.c:17-19, src/cuda/aco_cuda_kernels.cu:53-57);
 *   - costs are int64 in cost units: the nint length, or in real mode the
 *     exact sum of the float edge lengths times 2^23 (real cost = cost *
 *     ACO_HIP_REAL_COST_UNIT);
 *   - pheromone is the reference's log-quantized uint8 format
 *     (src/cuda/aco_cuda_kernels.cu:25-45, src/cuda/aco_cuda.cu:209-242):
 *     an (n+1)x(n+1) byte matrix with row stride n+1;
 *   - a tour is the reference's step-interleaved route sequence
 *     (src/cuda/aco_cuda.cu:138-168): 0, c.., 0, c.., 0, ... with K+1 depot
 *     entries (consecutive routes share their boundary depot) and
 *     length = visited customers + K + 1.
 */
#ifndef ACO_HIP_H
#define ACO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; 0..4 are AcoStatus (include/aco.h:82-88). */
enum {
  ACO_HIP_OK = 0,
  ACO_HIP_ERR_INVALID_INPUT = 1,
  ACO_HIP_ERR_ALLOCATION = 2,
  ACO_HIP_ERR_NO_SOLUTION = 3,
  ACO_HIP_ERR_BACKEND = 4
};

/* Candidate-list ordering keys. */
enum {
  /* exact integer key (dx^2+dy^2, index): ties go to the lower index */
  ACO_HIP_KNN_EXACT = 0,
  /* the reference OpenMP key (float)sqrt((double)(dx^2+dy^2)) with strict
   * '<' insertion, ties to the lower index (src/openmp-mpi/aco_parallel.c:218-250) */
  ACO_HIP_KNN_REF_F32 = 1
};

/* Distance of an edge, the objective and the heuristic 1/d. */
enum {
  ACO_HIP_DIST_NINT = 0, /* TSPLIB EUC_2D nearest integer, integer costs */
  ACO_HIP_DIST_REAL = 1  /* unrounded (float nearest to the root), the reference's metric */
};
#define ACO_HIP_REAL_COST_UNIT (1.0 / 8388608.0)

/* Heuristic term used for candidate edges: (1/d)^beta or a learned prior. */
enum {
  ACO_HIP_ETA_DISTANCE = 0, /* eta^beta = (1/d)^beta, reference aco_cuda_kernels.cu:207-212 */
  ACO_HIP_ETA_LEARNED = 1   /* per-candidate code set by aco_hip_set_candidate_eta() */
};

typedef struct AcoHipContext AcoHipContext;

/* One CVRP instance (host memory, read during aco_hip_load only). */
typedef struct {
  int n;                 /* customers; nodes are 0..n */
  int K;                 /* vehicles (routes per tour) */
  int capacity;          /* Q, 1..32767 */
  const int32_t *x;      /* n+1 coordinates, 0..32767 */
  const int32_t *y;      /* n+1 coordinates, 0..32767 */
  const int32_t *demand; /* n+1 demands, demand[0] == 0, customers 0..32767 */
  uint32_t seed;         /* solver seed of this instance (Philox key) */
  int cand_k;            /* candidate list length of this instance, 0 = AcoHipParams.cand_k */
} AcoHipInstance;

/* Solver parameters shared by every instance of a context.
 * Reference values: alpha 1, beta 2, rho 0.5, tau0 1, Q 1 (src/cuda/main_vrp.cu:53-57),
 * depot close weight 2 (src/cuda/aco_cuda.cu:240), cand_k 32 (aco_cuda.cu:198). */
typedef struct {
  int cand_k;            /* candidate list length k, 1..64 (clamped to n) */
  int ants;              /* ants per colony m, >= 1 */
  int colonies;          /* colonies per instance, >= 1 */
  float alpha;           /* pheromone exponent (any real value) */
  float beta;            /* heuristic exponent (any real value) */
  float rho;             /* evaporation factor in (0, 1) */
  float tau0;            /* initial pheromone */
  float Q;               /* deposit factor; deposits use 100*Q (aco_cuda.cu:228) */
  float depot_close_weight;
  int knn_key;           /* ACO_HIP_KNN_* */
  int ants_per_group;    /* ants (wavefronts) per workgroup, 0 = auto */
  int stage_lds;         /* -1 auto, 0 read candidate rows from global memory, 1 stage in LDS */
  int fallback_matrix;   /* -1 auto, 0 off, 1 on: keep the all-pairs fallback weights
                            taupow[tau[u][v]]*eta[d(u,v)] as a float matrix per colony
                            ((n+1) x roundup(n+1,16) floats) instead of recomputing them */
  int dist_mode;         /* ACO_HIP_DIST_*: real mode optimises unrounded lengths, with
                            (1/d)^beta taken from a table binned at relative width 2^-10 */
} AcoHipParams;

/* Device-side state of one colony, read back with aco_hip_get_colony(). */
typedef struct {
  int32_t epoch;            /* epochs executed */
  int32_t iter_since_best;  /* reference iter_since_best (aco_cuda.cu:322-399) */
  int32_t halted;           /* stop rule reached; later epochs are no-ops */
  int32_t found;            /* last epoch had a feasible ant */
  int32_t best_ant;         /* last epoch's iteration-best ant or -1 */
  int32_t improved;         /* last epoch improved the global best */
  int32_t best_len;         /* length of the global-best tour */
  int32_t feasible_ants;    /* feasible ants in the last epoch */
  int64_t iter_best_cost;   /* last epoch's best feasible cost or INT64_MAX */
  int64_t global_best_cost; /* best feasible cost so far or INT64_MAX */
  float dep_iter;           /* last iteration-best deposit 0.3*100Q/L_it */
  float dep_best;           /* last global-best deposit 0.7*100Q/L* */
} AcoHipColonyState;

/* Stop rules of the epoch loop (src/common/aco_config.c:33-73). */
typedef struct {
  double timeout_seconds;     /* 0 = off; checked between epoch chunks */
  int stagnation_epochs;      /* 0 = off */
  double min_rel_improvement; /* fraction, reference default 1e-3 */
  int fixed_epochs;           /* >0: run exactly this many epochs */
  int max_epochs;             /* >0: hard cap (safety), 0 = none */
  double progress_interval_seconds; /* stderr heartbeat, 0 = off */
  int print_progress;         /* print the reference's [progress] lines */
  int chunk_epochs;           /* epochs per launch batch / graph replay, 0 = auto */
  int use_graph;              /* capture a chunk of epochs in a hipGraph */
} AcoHipRunConfig;

typedef struct {
  int epochs;               /* epochs executed by the first colony */
  double seconds;           /* wall seconds of the epoch loop */
  double ant_tours_per_s;   /* sum over colonies of ants*epochs / seconds */
} AcoHipRunStats;

/* Per-kernel timing from hipEvents recorded on the context stream. */
typedef struct {
  double construct_ms;  int construct_launches;
  double best_ms;       int best_launches;      /* update kernel incl. fused deposits */
  double evaporate_ms;  int evaporate_launches;
  double deposit_ms;    int deposit_launches;
  double knn_ms;        int knn_launches;
} AcoHipTiming;

/* ---- context ------------------------------------------------------------ */
int aco_hip_create(AcoHipContext **out, int device);
void aco_hip_destroy(AcoHipContext *ctx);
const char *aco_hip_last_error(const AcoHipContext *ctx);
int aco_hip_sync(AcoHipContext *ctx);

/* Uploads instances, builds the quantization and heuristic tables and
 * allocates every pool.  Replaces the cudaMalloc/cudaMemcpy block of
 * src/cuda/aco_cuda.cu:209-286. */
int aco_hip_load(AcoHipContext *ctx, const AcoHipInstance *inst, int num_instances,
                 const AcoHipParams *params);
int aco_hip_num_colonies(const AcoHipContext *ctx);
/* Effective per-context sizes after clamping: k, ants, tour stride, ants/group, lds flag. */
int aco_hip_get_shape(const AcoHipContext *ctx, int *cand_k, int *ants, int *tour_stride,
                      int *ants_per_group, int *stage_lds);

/* ---- size buckets and autotuning (no reference counterpart: the reference
 *      solves one instance per launch) ---------------------------------------
 * Instances of a context are grouped by node count (n+1 up to 256, 512,
 * 1024, ...; ACO_HIP_BUCKETS=0 keeps one group); a context with a single
 * group whose launch holds more than 8 ants per CU splits it into two lanes
 * (ACO_HIP_LANES forces the count for every group).  Each bucket (a group or
 * a lane) has its own construct launch shape and its own chain of launches,
 * and the chains run concurrently. */
typedef struct {
  int n1_min, n1_max;    /* node counts present in the bucket */
  int instances, colonies;
  int ants_per_group;    /* wavefronts (ants) per workgroup */
  int stage_lds;         /* candidate rows staged in LDS (1) or read from global memory (0) */
  int lds_bytes;         /* LDS per workgroup */
  float tuned_ms;        /* construct time of the chosen shape (autotuner), 0 = not tuned */
  int ants_per_wave;     /* 1, or 2: two ants per wavefront (candidate lists <= 32) */
  int state_bytes;       /* bytes per node of an ant's LDS state: 2, or 1 for pairs whose
                          * instances' demands are all <= 254 */
  int lanes;             /* buckets of its size class running as concurrent chains */
} AcoHipBucketInfo;
int aco_hip_num_buckets(const AcoHipContext *ctx);
int aco_hip_get_bucket(const AcoHipContext *ctx, int bucket, AcoHipBucketInfo *info);
/* Force a bucket's shape (stage_lds -1 = automatic); the effective shape may
 * have fewer ants per workgroup when LDS does not fit the request (at most 8
 * wavefronts per workgroup, 16 for pairs with byte states).  Ants per
 * wavefront stay automatic. */
int aco_hip_set_bucket_shape(AcoHipContext *ctx, int bucket, int ants_per_group, int stage_lds);
/* The same with the ants per wavefront: 0 automatic, 1, or 2 (two ants per
 * wavefront, for candidate lists of at most 32; 1 is used above that).  With
 * two, ants_per_group counts wavefronts, each holding two ants. */
int aco_hip_set_bucket_layout(AcoHipContext *ctx, int bucket, int ants_per_group, int stage_lds,
                              int ants_per_wave);
/* Time `launches` construct launches of every bucket for each wavefronts-
 * per-workgroup in {1, 2, 4, 8, 16}, each staging mode that fits and one or two
 * ants per wavefront, apply the fastest per bucket and report them (info:
 * max_buckets entries).  Colonies must not be halted; pheromone and colony
 * state are left unchanged. */
int aco_hip_autotune(AcoHipContext *ctx, int launches, AcoHipBucketInfo *info, int max_buckets);

/* ---- kernels (one launch each, asynchronous on the context stream) ------ */
/* kNN candidate lists: replaces launch_build_candidate_lists
 * (include/aco_cuda_kernels.h:156, kernel aco_cuda_kernels.cu:164-214). */
int aco_hip_build_candidates(AcoHipContext *ctx);
/* Fill every pheromone byte with q(tau0) and reset the colony state:
 * replaces launch_init_tau (include/aco_cuda_kernels.h:151). */
int aco_hip_init_tau(AcoHipContext *ctx);
/* Reset ants and build one tour per ant for the current device epoch:
 * replaces launch_reset_ant_state + launch_construct_solutions
 * (include/aco_cuda_kernels.h:161,166; kernels aco_cuda_kernels.cu:227-562). */
int aco_hip_construct(AcoHipContext *ctx);
/* Iteration best, global best, stagnation and stop bookkeeping on device:
 * replaces select_iter_best_host and the host block of aco_cuda.cu:351-399. */
int aco_hip_select_best(AcoHipContext *ctx);
/* Saturating quantized evaporation: replaces launch_evaporate_tau
 * (include/aco_cuda_kernels.h:171, kernel aco_cuda_kernels.cu:145-155). */
int aco_hip_evaporate(AcoHipContext *ctx);
/* Iteration-best deposit: replaces launch_deposit_solution (aco_cuda_kernels.h:176). */
int aco_hip_deposit_iter(AcoHipContext *ctx);
/* Global-best deposit: replaces launch_deposit_flat_solution (aco_cuda_kernels.h:181). */
int aco_hip_deposit_best(AcoHipContext *ctx);
/* `count` full epochs, two launches each: construct, then one pheromone
 * update launch whose blocks evaporate (deciding from the ants whether this
 * epoch found a feasible tour) and whose last block to finish runs
 * select_best, deposit_iter and deposit_best.  Same result as the reference
 * order (select, evaporate, deposit, deposit). */
int aco_hip_epochs(AcoHipContext *ctx, int count);

/* ---- epoch loop: replaces aco_vrp_cuda_with_capacity (aco_cuda.cu:188-454) - */
int aco_hip_run(AcoHipContext *ctx, const AcoHipRunConfig *cfg, AcoHipRunStats *stats);

/* ---- downloads / uploads (synchronous) ---------------------------------- */
/* (n+1)*k candidate indices, distance codes and heuristic values (any may be
 * NULL).  The code is the nint distance, or in real mode the heuristic-table
 * index of the float distance (aco_real_eta_bin); the heuristic is eta^beta
 * of the distance, or the learned heuristic once one is set. */
int aco_hip_get_candidates(AcoHipContext *ctx, int instance, int32_t *idx, int32_t *dist,
                           float *eta);
/* Learned heuristic prior (the seam of aco_cuda_kernels.cu:295, where eta_beta
 * enters the candidate weight): candidate slot (i, t) of the instance weighs
 * taupow[tau_cand] * eta[i*k + t] instead of taupow * (1/d)^beta.  eta has
 * (n+1)*k finite values >= 0 (row-major like aco_hip_get_candidates); NULL
 * restores (1/d)^beta.  The all-customer fallback keeps (1/d)^beta.  Scores of
 * every colony of the instance are refreshed; call before the next epoch. */
int aco_hip_set_candidate_eta(AcoHipContext *ctx, int instance, const float *eta);
/* (n+1)^2 pheromone bytes and (n+1)*k candidate-aligned copy (either may be NULL). */
int aco_hip_get_tau(AcoHipContext *ctx, int colony, uint8_t *tau, uint8_t *tau_cand);
int aco_hip_set_tau(AcoHipContext *ctx, int colony, const uint8_t *tau);
/* (n+1)*k candidate scores taupow[tau_cand] * eta^beta, the weights the
 * construction samples (the reference OpenMP score matrix, aco_parallel.c:535-540). */
int aco_hip_get_scores(AcoHipContext *ctx, int colony, float *score);
/* (n+1)*(n+1) fallback weights taupow[tau[u][v]] * eta^beta(d(u,v)) of the
 * all-pairs matrix (the weights select_global_fallback computes per step,
 * aco_cuda_kernels.cu:336-430).  ACO_HIP_ERR_INVALID_INPUT when the matrix is
 * off (AcoHipParams.fallback_matrix). */
int aco_hip_get_fallback_weights(AcoHipContext *ctx, int colony, float *weights);
/* tours: ants*tour_stride int32; len, unvisited: ants int32; cost: ants int64. */
int aco_hip_get_tours(AcoHipContext *ctx, int colony, int32_t *tours, int32_t *len,
                      int64_t *cost, int32_t *unvisited);
int aco_hip_get_colony(AcoHipContext *ctx, int colony, AcoHipColonyState *st);
int aco_hip_set_epoch(AcoHipContext *ctx, int colony, int epoch);
int aco_hip_get_best(AcoHipContext *ctx, int colony, int32_t *tour, int32_t *len,
                     int64_t *cost);
/* Deposit a host tour on one colony (sequential edge semantics of
 * kernel_deposit_flat_solution, aco_cuda_kernels.cu:637-668).  The tour
 * starts and ends at the depot and visits each customer at most once; other
 * tours are refused with ACO_HIP_ERR_INVALID_INPUT. */
int aco_hip_deposit_tour(AcoHipContext *ctx, int colony, const int32_t *tour, int len,
                         float delta);
/* Route-cost reduction of host tours of instance `instance` (int64 nint sums);
 * the reference accumulates it in-kernel (aco_cuda_kernels.cu:540,550). */
int aco_hip_route_costs(AcoHipContext *ctx, int instance, const int32_t *tours,
                        const int32_t *len, int count, int stride, int64_t *cost);
/* Route costs without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as aco_hip_route_costs, but the tours, lengths and
 * costs are staged in a fixed device array of ACO_HIP_NOALLOC_SCRATCH_BYTES
 * that each GPU holds until the process ends, instead of buffers allocated
 * per call.  A batch needing more staging than that (count*stride*4 +
 * count*4 bytes, each rounded up to 256, + count*8) is refused with
 * ACO_HIP_ERR_INVALID_INPUT.  All calls on one GPU share its array: calls
 * that overlap on one GPU (several host threads) race.  Synchronous. */
#define ACO_HIP_NOALLOC_SCRATCH_BYTES (4u << 20)
int aco_hip_route_costs_noalloc(AcoHipContext *ctx, int instance, const int32_t *tours,
                                const int32_t *len, int count, int stride, int64_t *cost);
/* Host copies of the shared tables: deq[256], thr[255], taupow[256], eta[eta_len]. */
int aco_hip_get_tables(AcoHipContext *ctx, float *deq, float *thr, float *taupow,
                       float *eta_lut, int *eta_len, uint8_t *q_tau0, uint8_t *q_evap);

/* Device-side construction counters (the reference declares CudaIterStats,
 * include/aco_cuda_kernels.h:63-74, but never fills it). */
typedef struct {
  uint64_t decisions;        /* roulette decisions (moves and route closes) */
  uint64_t fallbacks;        /* decisions that needed the full-row fallback */
  uint64_t fallback_cycles;  /* shader clocks spent in the fallback */
  uint64_t ant_cycles;       /* shader clocks from ant start to tour end */
  uint64_t ants;             /* tours counted */
  uint64_t staging_cycles;   /* shader clocks per workgroup spent staging LDS, summed */
  /* pheromone-update kernel, in constant-rate wall-clock ticks (wall_clock_hz) */
  uint64_t updates;          /* epochs whose update ran the best selection */
  uint64_t evap_ticks;       /* first block start -> last block done evaporating */
  uint64_t select_ticks;     /* best selection and stop rules */
  uint64_t deposit_ticks;    /* iteration-best + global-best deposits */
  uint64_t wall_clock_hz;    /* tick rate of the three fields above */
} AcoHipStats;

/* Enable (1) / disable (0) the counters; enabling clears them. */
int aco_hip_set_stats(AcoHipContext *ctx, int enable);
/* Counters summed since enabling, for one colony. */
int aco_hip_get_stats(AcoHipContext *ctx, int colony, AcoHipStats *st);

/* ---- learned edge prior ----------------------------------------------------
 * No reference counterpart: the edge-scoring network that replaces eta_beta
 * at the seam of aco_cuda_kernels.cu:295 (see aco_hip_set_candidate_eta).
 * Network 8 -> 64 -> 64 -> 1 (ReLU), BF16 matrices on MFMA with FP32
 * accumulation, FP32 master parameters, Adam, mean binary cross-entropy
 * against "candidate edge joins consecutive nodes of an elite tour".
 * Parameter vector (ACO_PRIOR_PARAMS floats): W1[64][8], b1[64], W2[64][64],
 * b2[64], w3[64], b3.  Features per candidate slot (i, t), j = cand[i][t],
 * L = larger side of the coordinate bounding box: (x_i-x_0)/L, (y_i-y_0)/L,
 * (x_j-x_0)/L, (y_j-y_0)/L, demand_j/Q, nint d(i,j)/L, cosine of the angle
 * at i between i->j and i->depot, t/k; each rounded to BF16. */
#define ACO_PRIOR_FEATURES 8
#define ACO_PRIOR_PARAMS 4801
typedef struct AcoHipPrior AcoHipPrior;
/* Glorot-uniform weights from Philox(seed), zero biases. */
int aco_hip_prior_create(AcoHipPrior **out, int device, uint32_t seed);
void aco_hip_prior_destroy(AcoHipPrior *prior);
const char *aco_hip_prior_last_error(const AcoHipPrior *prior);
int aco_hip_prior_get_params(AcoHipPrior *prior, float *params);
/* Also resets the optimiser state. */
int aco_hip_prior_set_params(AcoHipPrior *prior, const float *params);
/* (n+1)*k rows of ACO_PRIOR_FEATURES BF16 features (raw uint16) of an instance. */
int aco_hip_edge_features(AcoHipContext *ctx, int instance, uint16_t *features);
/* Logits of `count` edges given as BF16 feature rows. */
int aco_hip_prior_forward(AcoHipPrior *prior, const uint16_t *features, int64_t count,
                          float *logits);
/* Gradient of the mean BCE (ACO_PRIOR_PARAMS floats) and the mean BCE itself. */
int aco_hip_prior_gradient(AcoHipPrior *prior, const uint16_t *features, const uint8_t *labels,
                           int64_t count, float *grad, float *loss);
/* Training set: append every candidate edge of an instance, labelled from an
 * elite tour of it (len entries, depot-delimited routes). */
int aco_hip_prior_add_instance(AcoHipPrior *prior, AcoHipContext *ctx, int instance,
                               const int32_t *tour, int len);
int aco_hip_prior_clear(AcoHipPrior *prior);
int64_t aco_hip_prior_size(const AcoHipPrior *prior);
/* `steps` full-batch Adam steps (beta 0.9/0.999) over the training set;
 * loss_out (may be NULL) = mean BCE of the last step's forward pass. */
int aco_hip_prior_train(AcoHipPrior *prior, int steps, float lr, float *loss_out);
/* Score an instance's candidate edges and install
 * eta = (1/d)^beta * gain * sigmoid(logit) as its learned heuristic. */
int aco_hip_prior_apply(AcoHipPrior *prior, AcoHipContext *ctx, int instance, float gain);

/* ---- timing ------------------------------------------------------------- */
int aco_hip_set_timing(AcoHipContext *ctx, int enable);
int aco_hip_get_timing(AcoHipContext *ctx, AcoHipTiming *t);

/* ---- one-call solver: one instance, one colony, weighted demand and the
 *      distance mode of choice (the extern "C" form of
 *      aco_vrp_cuda_with_capacity, src/cuda/aco_cuda.cu:188-454, whose exact
 *      C++ prototype is in aco_cuda_compat.h).  Writes the best tour (capacity
 *      n+K+1), its length and its cost in cost units.  Stop rules come from
 *      the ACO_SOLVER_* environment like the reference. ------------------- */
int aco_vrp_hip_with_capacity(int n, int K, int capacity, int m, const int32_t *x,
                              const int32_t *y, const int32_t *demand, double alpha,
                              double beta, double rho, double tau0, double Q,
                              unsigned int seed, int dist_mode, int32_t *best_tour,
                              int32_t *best_len, int64_t *best_cost);

#ifdef __cplusplus
}
#endif

#endif /* ACO_HIP_H */
