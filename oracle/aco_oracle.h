/*
 * aco_oracle.h - CPU oracle of the ACO-CVRP hot path.  TEST INFRASTRUCTURE:
 * only tests, the smoke check and the benchmark's CPU-baseline leg use it,
 * always as the checker and never as the product path.
 */
#ifndef ACO_ORACLE_H
#define ACO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleColony OracleColony;

typedef struct {
  int32_t epoch, iter_since_best, halted, found, best_ant, improved, best_len, feasible_ants;
  int64_t iter_best_cost, global_best_cost;
  float dep_iter, dep_best;
} OracleState;

/* dist_mode: 0 = TSPLIB nint distances and integer costs, 1 = real
 * distances (floats), costs in units of 2^-23. */
OracleColony *oracle_create(int n, int K, int Q, const int32_t *x, const int32_t *y,
                            const int32_t *demand, int cand_k, int ants, float alpha, float beta,
                            float rho, float tau0, float Qdep, float depot_w, int knn_key,
                            uint32_t seed, uint32_t colony, int dist_mode);
void oracle_destroy(OracleColony *o);
int oracle_cand_k(const OracleColony *o);
int oracle_tour_stride(const OracleColony *o);
void oracle_get_candidates(const OracleColony *o, int32_t *idx, int32_t *dist, float *eta);
void oracle_get_tables(const OracleColony *o, float *deq, float *thr, float *taupow,
                       float *eta_lut, int *eta_len, uint8_t *q_tau0, uint8_t *q_evap);
void oracle_get_tau(const OracleColony *o, uint8_t *tau, uint8_t *tau_cand);
void oracle_set_tau(OracleColony *o, const uint8_t *tau);
void oracle_set_epoch(OracleColony *o, int epoch);
/* Per-candidate-slot heuristic ((n+1)*k values) replacing (1/d)^beta in the
 * candidate roulette; NULL restores the distance heuristic. */
void oracle_set_candidate_eta(OracleColony *o, const float *eta);
void oracle_set_stop_rules(OracleColony *o, int stagnation_epochs, double min_rel,
                           int fixed_epochs);
/* One ant's tour (tour capacity >= tour_stride). Returns the tour length. */
int oracle_construct_ant(const OracleColony *o, int ant, int epoch, int32_t *tour,
                         int64_t *cost, int32_t *unvisited, int64_t *fallbacks);
/* Every ant of the colony for the current epoch (OpenMP over ants). */
void oracle_construct(OracleColony *o, int nthreads);
void oracle_get_tours(const OracleColony *o, int32_t *tours, int32_t *len, int64_t *cost,
                      int32_t *unvisited);
void oracle_select_best(OracleColony *o);
/* is_significant_improvement (src/cuda/aco_cuda.cu:83-94). */
int oracle_significant(double prev, double now, double min_rel);
/* Test hook: replace the current epoch's ant results (NULL keeps a field). */
void oracle_set_ant_results(OracleColony *o, const int64_t *cost, const int32_t *unvisited,
                            const int32_t *len);
void oracle_evaporate(OracleColony *o);
void oracle_deposit_tour(OracleColony *o, const int32_t *tour, int len, float delta);
/* Per-route deposits of the reference's route layouts on the colony's
 * matrix (step_major 1: kernel_deposit_solution's routes[steps][m] and
 * lens[m][K]; 0: kernel_deposit_flat_solution's routes[K][n+1] and lens[K]). */
void oracle_deposit_routes(OracleColony *o, const int32_t *routes, const int32_t *lens, int K,
                           int m, long long steps, int ant, float delta, int step_major);
/* kernel_evaporate_tau with explicit delta and q_min (whole matrix). */
void oracle_evaporate_levels(OracleColony *o, uint8_t delta, uint8_t q_min);
void oracle_deposit_iter(OracleColony *o);
void oracle_deposit_best(OracleColony *o);
void oracle_epochs(OracleColony *o, int count, int nthreads);
void oracle_get_state(const OracleColony *o, OracleState *st);
int oracle_get_best(const OracleColony *o, int32_t *tour, int32_t *len, int64_t *cost);
int64_t oracle_route_cost(const OracleColony *o, const int32_t *tour, int len);
int64_t oracle_fallback_count(const OracleColony *o);

/* Primitive restatements exposed for known-answer tests. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int32_t oracle_nint_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2);
void oracle_scan_rows(const float *w, int nlanes, float *prefix);
uint8_t oracle_quantize(const OracleColony *o, float tau);

#ifdef __cplusplus
}
#endif

#endif
