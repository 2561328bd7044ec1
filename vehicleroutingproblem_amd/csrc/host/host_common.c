/*
 * Solution store, validation, route cost, ACO_SOLVER_* configuration and CLI
 * output of the HIP solver.  Behaviour and text follow the reference's
 * src/common/solution.c, src/common/aco_config.c and src/common/cli_common.c
 * (cited per function); validation adds the capacity check the reference
 * leaves to offline pyvrp validation (tools/python/validate_pyvrp.py:30-107).
 */
#include "aco_host.h"
#include "aco_inline.h"

#include <errno.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- solution (src/common/solution.c:45-137) ---------------------------- */

Solution *solution_create(int K, int n) {
  if (K <= 0 || n < 0) return NULL;
  Solution *s = (Solution *)calloc(1, sizeof(*s));
  if (!s) return NULL;
  s->K = K;
  s->route_cap = n + 2; /* both depots plus every customer */
  s->routes = (Route *)calloc((size_t)K, sizeof(Route));
  s->nodes_storage = (int *)calloc((size_t)K * (size_t)s->route_cap, sizeof(int));
  if (!s->routes || !s->nodes_storage) {
    solution_free(s);
    return NULL;
  }
  for (int i = 0; i < K; ++i) {
    s->routes[i].nodes = s->nodes_storage + (size_t)i * (size_t)s->route_cap;
    s->routes[i].cap = s->route_cap;
    s->routes[i].len = 0;
  }
  return s;
}

void solution_free(Solution *s) {
  if (!s) return;
  free(s->nodes_storage);
  free(s->routes);
  free(s);
}

void solution_reset(Solution *s) {
  for (int i = 0; i < s->K; ++i) s->routes[i].len = 0;
}

int solution_from_tour(Solution *s, const int32_t *tour, int len) {
  solution_reset(s);
  if (len < 2 || tour[0] != 0 || tour[len - 1] != 0) return 0;
  int v = 0;
  Route *r = &s->routes[0];
  r->nodes[r->len++] = 0;
  for (int p = 1; p < len; ++p) {
    int node = tour[p];
    if (r->len >= r->cap) return 0;
    r->nodes[r->len++] = node;
    if (node == 0) {
      ++v;
      if (p == len - 1) break;
      if (v >= s->K) return 0;
      r = &s->routes[v];
      r->nodes[r->len++] = 0;
    }
  }
  return v == s->K;
}

static void set_err(char *err, size_t err_len, const char *msg) {
  if (err && err_len) snprintf(err, err_len, "%s", msg);
}

/* Checks and messages of solution_validate (src/common/solution.c:161-234). */
bool solution_validate(const Solution *s, int n, int K, char *err, size_t err_len) {
  if (!s) {
    set_err(err, err_len, "solution is NULL");
    return false;
  }
  if (s->K != K) {
    set_err(err, err_len, "solution K does not match expected K");
    return false;
  }
  if (!s->routes) {
    set_err(err, err_len, "solution routes are NULL");
    return false;
  }
  if (n < 0 || K < 0) {
    set_err(err, err_len, "invalid n or K");
    return false;
  }
  unsigned char *seen = (unsigned char *)calloc((size_t)n + 1, 1);
  if (!seen) {
    set_err(err, err_len, "allocation failure in solution_validate");
    return false;
  }
  bool ok = true;
  for (int i = 0; i < s->K && ok; ++i) {
    const Route *r = &s->routes[i];
    if (r->len < 2) {
      set_err(err, err_len, "route length < 2");
      ok = false;
      break;
    }
    if (r->nodes[0] != 0 || r->nodes[r->len - 1] != 0) {
      set_err(err, err_len, "route must start/end at depot 0");
      ok = false;
      break;
    }
    for (int t = 0; t < r->len; ++t) {
      int node = r->nodes[t];
      if (node < 0 || node > n) {
        set_err(err, err_len, "route contains out-of-range node");
        ok = false;
        break;
      }
      if (node == 0) {
        if (t != 0 && t != r->len - 1) {
          set_err(err, err_len, "route contains depot inside the path");
          ok = false;
          break;
        }
        continue;
      }
      if (seen[node]) {
        set_err(err, err_len, "customer visited more than once");
        ok = false;
        break;
      }
      seen[node] = 1;
    }
  }
  for (int node = 1; ok && node <= n; ++node) {
    if (!seen[node]) {
      set_err(err, err_len, "customer not visited");
      ok = false;
    }
  }
  free(seen);
  if (ok) set_err(err, err_len, "");
  return ok;
}

bool solution_validate_capacity(const Solution *s, const VrpInstance *inst, char *err,
                                size_t err_len) {
  if (!solution_validate(s, inst->n, inst->K, err, err_len)) return false;
  for (int i = 0; i < s->K; ++i) {
    long long load = 0;
    const Route *r = &s->routes[i];
    for (int t = 0; t < r->len; ++t) load += inst->demand[r->nodes[t]];
    if (load > inst->capacity) {
      set_err(err, err_len, "route load exceeds vehicle capacity");
      return false;
    }
  }
  return true;
}

int64_t solution_cost_nint(const Solution *s, const VrpInstance *inst) {
  int64_t cost = 0;
  for (int i = 0; i < s->K; ++i) {
    const Route *r = &s->routes[i];
    for (int t = 0; t + 1 < r->len; ++t) {
      int u = r->nodes[t], v = r->nodes[t + 1];
      cost += vrp_nint_dist(inst->x[u], inst->y[u], inst->x[v], inst->y[v]);
    }
  }
  return cost;
}

/* Real-mode route cost in the solver's units: the sum over edges of the
 * float distance times 2^23 (exact; aco_real_units). */
int64_t solution_cost_real_units(const Solution *s, const VrpInstance *inst) {
  int64_t cost = 0;
  for (int i = 0; i < s->K; ++i) {
    const Route *r = &s->routes[i];
    for (int t = 0; t + 1 < r->len; ++t) {
      int u = r->nodes[t], v = r->nodes[t + 1];
      int32_t dx = inst->x[u] - inst->x[v], dy = inst->y[u] - inst->y[v];
      cost += aco_real_units(aco_real_dist_s((uint32_t)(dx * dx) + (uint32_t)(dy * dy)));
    }
  }
  return cost;
}

/* solution_cost (src/common/solution.c:126-137) on unrounded distances. */
double solution_cost_real(const Solution *s, const VrpInstance *inst) {
  double cost = 0.0;
  for (int i = 0; i < s->K; ++i) {
    const Route *r = &s->routes[i];
    for (int t = 0; t + 1 < r->len; ++t) {
      int u = r->nodes[t], v = r->nodes[t + 1];
      cost += vrp_real_dist(inst->x[u], inst->y[u], inst->x[v], inst->y[v]);
    }
  }
  return cost;
}

/* ---- configuration (src/common/aco_config.c:5-73) ----------------------- */

double aco_parse_min_rel_improvement_percent(const char *s, double default_fraction) {
  if (!s || !*s) return default_fraction;
  double percent = atof(s);
  if (percent <= 0.0) return default_fraction;
  return percent / 100.0;
}

void aco_runtime_config_defaults(AcoRuntimeConfig *c) {
  if (!c) return;
  c->timeout_seconds = 0.0;
  c->stagnation_epochs = 0;
  c->min_rel_improvement = 1e-3;
  c->fixed_epochs = 0;
  c->candidate_k = 0;
  c->ants = 0;
  c->seed = 1234u;
  c->progress_interval_seconds = 10.0;
}

void aco_runtime_config_load_env(AcoRuntimeConfig *c) {
  if (!c) return;
  aco_runtime_config_defaults(c);
  const char *s_timeout = getenv("ACO_SOLVER_TIMEOUT_SECONDS");
  const char *s_stag = getenv("ACO_SOLVER_STAGNATION_EPOCHS");
  const char *s_rel = getenv("ACO_SOLVER_MIN_REL_IMPROVEMENT");
  const char *s_fixed = getenv("ACO_SOLVER_FIXED_EPOCHS");
  const char *s_prog = getenv("ACO_SOLVER_PROGRESS_INTERVAL_SECONDS");
  c->timeout_seconds = (s_timeout && *s_timeout) ? atof(s_timeout) : 0.0;
  c->stagnation_epochs = (s_stag && *s_stag) ? atoi(s_stag) : 0;
  c->min_rel_improvement = aco_parse_min_rel_improvement_percent(s_rel, 1e-3);
  c->fixed_epochs = (s_fixed && *s_fixed) ? atoi(s_fixed) : 0;
  c->progress_interval_seconds = (s_prog && *s_prog) ? atof(s_prog) : 10.0;
  if (c->stagnation_epochs < 0) c->stagnation_epochs = 0;
  if (c->min_rel_improvement <= 0.0) c->min_rel_improvement = 1e-3;
  if (c->fixed_epochs < 0) c->fixed_epochs = 0;
  if (c->progress_interval_seconds < 0.0) c->progress_interval_seconds = 0.0;
}

/* is_significant_improvement (src/cuda/aco_cuda.cu:83-94), ACO_EPS = 1e-9. */
int aco_is_significant_improvement(double prev_best, double new_best, double min_rel) {
  const double eps = 1e-9;
  if (prev_best >= DBL_MAX || new_best >= DBL_MAX) return new_best < prev_best;
  if (new_best >= prev_best - eps) return 0;
  double rel_gain = (prev_best - new_best) / fmax(prev_best, eps);
  return rel_gain + eps >= min_rel;
}

/* ---- CLI (src/common/cli_common.c:8-85) --------------------------------- */

int cli_parse_int_arg(const char *s, int *out) {
  if (!s || !out) return 0;
  char *end = NULL;
  errno = 0;
  long v = strtol(s, &end, 10);
  if (errno != 0 || end == s || *end != '\0') return 0;
  if (v < 0 || v > 100000000L) return 0;
  *out = (int)v;
  return 1;
}

unsigned int cli_parse_uint_arg(const char *s, int *ok) {
  if (!s || !ok) {
    if (ok) *ok = 0;
    return 0u;
  }
  char *end = NULL;
  errno = 0;
  unsigned long v = strtoul(s, &end, 10);
  if (errno != 0 || end == s || *end != '\0' || v > 0xFFFFFFFFUL) {
    *ok = 0;
    return 0u;
  }
  *ok = 1;
  return (unsigned int)v;
}

void cli_print_solution_routes(const Solution *best, int K) {
  for (int i = 0; i < K; ++i) {
    const Route *r = &best->routes[i];
    printf("Route %d:", i + 1);
    for (int t = 0; t < r->len; ++t) {
      if (r->nodes[t] != 0) printf(" %d", r->nodes[t]);
    }
    printf("\n");
  }
}

void cli_print_solution_cost(double best_cost) {
  printf("Cost: %.3f\n", best_cost);
  printf("best cost: %.3f\n", best_cost);
}

int cli_validate_solution_or_report(const Solution *best, int n, int K, double best_cost) {
  char err[160];
  if (!isfinite(best_cost) || best_cost <= 0.0) {
    fprintf(stderr, "invalid solution cost: %.12g\n", best_cost);
    return 0;
  }
  if (!solution_validate(best, n, K, err, sizeof(err))) {
    fprintf(stderr, "invalid solution: %s\n", err);
    return 0;
  }
  return 1;
}

/* aco_status_string (src/common/aco_shared.c:8-23). */
const char *aco_status_string(int status) {
  switch (status) {
    case 0: return "ok";
    case 1: return "invalid input";
    case 2: return "allocation failure";
    case 3: return "no solution";
    case 4: return "backend failure";
    default: return "unknown status";
  }
}
