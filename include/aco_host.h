/*
 * aco_host.h - host-side C surface kept from the reference.
 *
 *  - .vrp parser (include/instance_parser.h:19-44, src/common/instance_parser.c:226-388),
 *    extended to weighted DEMAND_SECTION and integer coordinates;
 *  - .vrp writer and seeded Philox instance generators (uniform and clustered);
 *  - Solution store, validation and cost (include/solution.h:32-77,
 *    src/common/solution.c:126-234), validation extended with capacity;
 *  - runtime configuration from ACO_SOLVER_* (include/aco_config.h:4-18,
 *    src/common/aco_config.c:5-73);
 *  - CLI helpers and output format (include/cli_common.h:10-15, src/common/cli_common.c);
 *  - quantization / heuristic tables shared by the HIP path;
 *  - integer TSPLIB distance.
 */
#ifndef ACO_HOST_H
#define ACO_HOST_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- instance ----------------------------------------------------------- */
typedef struct {
  int n;            /* customers; nodes 0..n */
  int K;            /* VEHICLES */
  int capacity;     /* CAPACITY */
  int32_t *x;       /* n+1 */
  int32_t *y;       /* n+1 */
  int32_t *demand;  /* n+1 */
  int unit_demand;  /* every customer demand is 1 */
  char name[128];
} VrpInstance;

/* Loads a TSPLIB CVRP file.  Follows vrp_load_tsplib_euc2d_coords
 * (src/common/instance_parser.c:226-388): DIMENSION gives n = dim-1, VEHICLES,
 * CAPACITY and DEMAND_SECTION are required, depot demand must be 0,
 * NODE_COORD_SECTION rows are positional.  Unlike the reference, any
 * non-negative customer demand is accepted unless require_unit_demand is set
 * (then demand != 1 fails as at instance_parser.c:357-366).  Coordinates must
 * be integers in [0, 32767].  Returns 0 on success and prints the reason on
 * stderr otherwise. */
int vrp_read(const char *path, VrpInstance *out, int require_unit_demand);
int vrp_write(const char *path, const VrpInstance *inst);
void vrp_free(VrpInstance *inst);
int vrp_alloc(VrpInstance *inst, int n);

/* Seeded generators (Philox4x32-10 keyed by seed).
 * Uniform: depot (c/2,c/2), customers integer U{0..side}^2, demands U{dmin..dmax},
 * Q = ceil(slack_pct/100 * sum(d) / K).
 * Clustered: `clusters` centres U[0,side]^2, points centre + N(0, sigma^2) per
 * axis, clipped to [0,side] and rounded; depot at the box centre. */
int vrp_generate_uniform(VrpInstance *out, int n, int K, int side, int dmin, int dmax,
                         int slack_pct, uint32_t seed);
int vrp_generate_clustered(VrpInstance *out, int n, int K, int side, int clusters,
                           double sigma, int dmin, int dmax, int slack_pct, uint32_t seed);
/* Q = ceil(slack_pct/100 * total / K), exact integer arithmetic. */
int vrp_capacity_for(long long total_demand, int K, int slack_pct);

/* TSPLIB EUC_2D nearest integer distance, exact for |dx|,|dy| <= 32767. */
int vrp_nint_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2);
/* Unrounded Euclidean distance (the reference's euc_2d, instance_parser.c:17-19). */
double vrp_real_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2);

/* ---- solution ------------------------------------------------------------ */
typedef struct {
  int *nodes;
  int len;
  int cap;
} Route;

typedef struct {
  Route *routes;
  int K;
  int route_cap;
  int *nodes_storage;
} Solution;

Solution *solution_create(int K, int n);
void solution_free(Solution *s);
void solution_reset(Solution *s);
/* Splits a step-interleaved tour (0,c..,0,c..,0) into K routes, each with
 * both depots (the layout of copy_ant_to_solution, src/cuda/aco_cuda.cu:138-168).
 * Returns 1 on success, 0 if the tour does not hold exactly K routes. */
int solution_from_tour(Solution *s, const int32_t *tour, int len);
/* Structural checks of solution_validate (src/common/solution.c:161-234). */
bool solution_validate(const Solution *s, int n, int K, char *err, size_t err_len);
/* solution_validate plus per-route load <= capacity. */
bool solution_validate_capacity(const Solution *s, const VrpInstance *inst, char *err,
                                size_t err_len);
/* Integer nint route cost (exact). */
int64_t solution_cost_nint(const Solution *s, const VrpInstance *inst);
/* Real (unrounded, double) route cost, the reference's solution_cost. */
double solution_cost_real(const Solution *s, const VrpInstance *inst);
/* Real-mode cost in the solver's units (float distances x 2^23, exact). */
int64_t solution_cost_real_units(const Solution *s, const VrpInstance *inst);

/* ---- runtime configuration ---------------------------------------------- */
typedef struct {
  double timeout_seconds;
  int stagnation_epochs;
  double min_rel_improvement;
  int fixed_epochs;
  int candidate_k;
  int ants;
  unsigned int seed;
  double progress_interval_seconds;
} AcoRuntimeConfig;

void aco_runtime_config_defaults(AcoRuntimeConfig *config);
/* Same variables, defaults and clamps as aco_config.c:33-73. */
void aco_runtime_config_load_env(AcoRuntimeConfig *config);
double aco_parse_min_rel_improvement_percent(const char *s, double default_fraction);
/* Reference improvement test (aco_cuda.cu:83-94), eps 1e-9. */
int aco_is_significant_improvement(double prev_best, double new_best, double min_rel);

/* ---- CLI helpers (src/common/cli_common.c) -------------------------------- */
int cli_parse_int_arg(const char *s, int *out);
unsigned int cli_parse_uint_arg(const char *s, int *ok);
void cli_print_solution_routes(const Solution *best, int K);
void cli_print_solution_cost(double best_cost);
int cli_validate_solution_or_report(const Solution *best, int n, int K, double best_cost);
const char *aco_status_string(int status);

/* ---- quantization and heuristic tables ----------------------------------- */
typedef struct {
  float log_tau_min;    /* logf(1e-4f) */
  float log_tau_step;   /* (logf(100) - logf(1e-4)) / 255 */
  uint8_t q_tau0;       /* clamp(roundf((logf(tau0) - log_min)/step)) */
  uint8_t q_evap;       /* max(1, roundf(-logf(1-rho)/step)) */
  float deq[256];       /* exp(q*step + log_min) */
  float thr[255];       /* smallest tau quantizing to q+1 */
  float taupow[256];    /* deq[q]^alpha */
} AcoTauTables;

/* Parameters of src/cuda/aco_cuda.cu:209-215,242 (host float math). */
void aco_tau_tables_build(AcoTauTables *t, float alpha, float rho, float tau0);
/* (1/d)^beta for d = 0..dmax, d = 0 uses 1/1e-9 like 1/(0+CUDA_EPS). */
void aco_eta_table_build(float *eta, int dmax, float beta);
/* Real mode: eta^beta by heuristic-table index (aco_real_eta_bin of the
 * float distance), index 0 = distance 0; len = largest index + 1. */
void aco_eta_table_build_real(float *eta, int len, float beta);
/* out[i] = the float nearest to sqrt(s[i]), with the kernels' arithmetic. */
void aco_real_dist_batch(const uint32_t *s, int64_t count, float *out);
/* Quantize through the threshold table (number of thresholds <= tau). */
uint8_t aco_tau_quantize(const AcoTauTables *t, float tau);
/* Host entry to the shared Philox4x32-10 (include/aco_inline.h). */
void aco_philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* ACO_HOST_H */
