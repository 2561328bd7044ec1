/*
 * TSPLIB CVRP reader/writer and seeded instance generators.
 *
 * Reader semantics follow vrp_load_tsplib_euc2d_coords
 * (src/common/instance_parser.c:226-388): keyword lines are matched by prefix,
 * the value is read after an optional ':'; DIMENSION gives n = dim-1 and must
 * precede the sections; NODE_COORD_SECTION and DEMAND_SECTION rows are read
 * positionally (ids ignored); VEHICLES, CAPACITY and DEMAND_SECTION are
 * required; depot demand must be 0.  Extensions: weighted demands are
 * accepted (the reference rejects demand != 1), coordinates must be integers
 * in [0, 32767] (nint distances are exact there).
 */
#include "aco_host.h"
#include "aco_inline.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define VRP_COORD_MAX 32767

int vrp_alloc(VrpInstance *inst, int n) {
  memset(inst, 0, sizeof(*inst));
  if (n <= 0) return 1;
  inst->n = n;
  inst->x = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  inst->y = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  inst->demand = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  if (!inst->x || !inst->y || !inst->demand) {
    vrp_free(inst);
    return 2;
  }
  return 0;
}

void vrp_free(VrpInstance *inst) {
  if (!inst) return;
  free(inst->x);
  free(inst->y);
  free(inst->demand);
  inst->x = inst->y = inst->demand = NULL;
  inst->n = 0;
}

/* Integer after the keyword: "KEY : v", "KEY: v" or "KEY v". */
static int keyword_int(const char *line, size_t keylen, int *out) {
  const char *p = line + keylen;
  const char *colon = strchr(p, ':');
  if (colon) p = colon + 1;
  while (*p && isspace((unsigned char)*p)) ++p;
  char *end = NULL;
  long v = strtol(p, &end, 10);
  if (end == p) return 0;
  *out = (int)v;
  return 1;
}

static int keyword_str(const char *line, size_t keylen, char *out, size_t cap) {
  const char *p = line + keylen;
  const char *colon = strchr(p, ':');
  if (colon) p = colon + 1;
  while (*p && isspace((unsigned char)*p)) ++p;
  size_t len = strcspn(p, "\r\n");
  while (len > 0 && isspace((unsigned char)p[len - 1])) --len;
  if (len >= cap) len = cap - 1;
  memcpy(out, p, len);
  out[len] = '\0';
  return 1;
}

int vrp_read(const char *path, VrpInstance *out, int require_unit_demand) {
  FILE *f = fopen(path, "r");
  if (!f) {
    perror("fopen");
    return 1;
  }
  char line[512];
  int n = 0, K = 0, capacity = 0;
  int have_coords = 0, have_demands = 0;
  char name[128] = "";
  int32_t *x = NULL, *y = NULL, *dem = NULL;
  int rc = 1;

  while (fgets(line, sizeof(line), f)) {
    if (strncmp(line, "NAME", 4) == 0) {
      keyword_str(line, 4, name, sizeof(name));
    } else if (strncmp(line, "DIMENSION", 9) == 0) {
      int dim = 0;
      if (keyword_int(line, 9, &dim)) {
        /* the sections already read were sized by the earlier DIMENSION */
        if ((x || dem) && dim - 1 != n) {
          fprintf(stderr, "DIMENSION changes after a section was read\n");
          goto done;
        }
        n = dim - 1;
      }
    } else if (strncmp(line, "VEHICLES", 8) == 0) {
      keyword_int(line, 8, &K);
    } else if (strncmp(line, "CAPACITY", 8) == 0) {
      keyword_int(line, 8, &capacity);
    } else if (strncmp(line, "NODE_COORD_SECTION", 18) == 0) {
      if (n <= 0) {
        fprintf(stderr, "DIMENSION must be defined before NODE_COORD_SECTION\n");
        goto done;
      }
      free(x);
      free(y);
      x = (int32_t *)malloc(((size_t)n + 1) * sizeof(int32_t));
      y = (int32_t *)malloc(((size_t)n + 1) * sizeof(int32_t));
      if (!x || !y) goto done;
      for (int i = 0; i <= n; ++i) {
        int id;
        double xd, yd;
        if (fscanf(f, "%d %lf %lf", &id, &xd, &yd) != 3) {
          fprintf(stderr, "Failed to read coordinates at index %d\n", i);
          goto done;
        }
        if (xd != floor(xd) || yd != floor(yd) || xd < 0 || yd < 0 || xd > VRP_COORD_MAX ||
            yd > VRP_COORD_MAX) {
          fprintf(stderr, "Coordinates must be integers in [0, %d] (index %d) in %s\n",
                  VRP_COORD_MAX, i, path);
          goto done;
        }
        x[i] = (int32_t)xd;
        y[i] = (int32_t)yd;
      }
      have_coords = 1;
    } else if (strncmp(line, "DEMAND_SECTION", 14) == 0) {
      if (n <= 0) {
        fprintf(stderr, "DIMENSION must be defined before DEMAND_SECTION\n");
        goto done;
      }
      free(dem);
      dem = (int32_t *)malloc(((size_t)n + 1) * sizeof(int32_t));
      if (!dem) goto done;
      for (int i = 0; i <= n; ++i) {
        int id, d;
        if (fscanf(f, "%d %d", &id, &d) != 2) {
          fprintf(stderr, "Failed to read demands at index %d\n", i);
          goto done;
        }
        dem[i] = d;
      }
      have_demands = 1;
    }
  }

  if (n <= 0 || !have_coords) {
    fprintf(stderr, "Invalid or missing DIMENSION/NODE_COORD_SECTION\n");
    goto done;
  }
  if (K <= 0 || capacity <= 0 || !have_demands) {
    fprintf(stderr, "Missing VEHICLES/CAPACITY/DEMAND_SECTION in %s\n", path);
    goto done;
  }
  if (dem[0] != 0) {
    fprintf(stderr, "Depot demand must be 0 in %s\n", path);
    goto done;
  }
  int unit = 1;
  for (int i = 1; i <= n; ++i) {
    if (dem[i] != 1) unit = 0;
    if (dem[i] < 0 || dem[i] > VRP_COORD_MAX) {
      fprintf(stderr, "Customer demand out of range [0, %d] in %s\n", VRP_COORD_MAX, path);
      goto done;
    }
  }
  if (require_unit_demand && !unit) {
    fprintf(stderr, "Customer demand must be 1 for all customers in %s\n", path);
    goto done;
  }
  if (capacity > VRP_COORD_MAX) {
    fprintf(stderr, "CAPACITY must be in [1, %d] in %s\n", VRP_COORD_MAX, path);
    goto done;
  }

  memset(out, 0, sizeof(*out));
  out->n = n;
  out->K = K;
  out->capacity = capacity;
  out->x = x;
  out->y = y;
  out->demand = dem;
  out->unit_demand = unit;
  snprintf(out->name, sizeof(out->name), "%s", name);
  x = y = dem = NULL;
  rc = 0;

done:
  free(x);
  free(y);
  free(dem);
  fclose(f);
  return rc;
}

/* Writes the layout of the reference instances (tab-separated rows, 1-based
 * ids, DEPOT_SECTION 1 -1, EOF), which both readers accept. */
int vrp_write(const char *path, const VrpInstance *inst) {
  FILE *f = fopen(path, "w");
  if (!f) {
    perror("fopen");
    return 1;
  }
  fprintf(f, "NAME: %s\n", inst->name[0] ? inst->name : "generated");
  fprintf(f, "TYPE: CVRP\n");
  fprintf(f, "DIMENSION: %d\n", inst->n + 1);
  fprintf(f, "VEHICLES: %d\n", inst->K);
  fprintf(f, "EDGE_WEIGHT_TYPE: EUC_2D\n");
  fprintf(f, "CAPACITY: %d\n", inst->capacity);
  fprintf(f, "NODE_COORD_SECTION\n");
  for (int i = 0; i <= inst->n; ++i) fprintf(f, "%d\t%d\t%d\n", i + 1, inst->x[i], inst->y[i]);
  fprintf(f, "DEMAND_SECTION\n");
  for (int i = 0; i <= inst->n; ++i) fprintf(f, "%d\t%d\n", i + 1, inst->demand[i]);
  fprintf(f, "DEPOT_SECTION\n1\n-1\nEOF\n");
  return fclose(f) == 0 ? 0 : 1;
}

/* ---- generators --------------------------------------------------------- */

/* Stream of Philox words keyed by (seed, stream id). */
typedef struct {
  uint32_t seed, stream, counter, buf[4];
  int pos;
} GenStream;

static void gs_init(GenStream *g, uint32_t seed, uint32_t stream) {
  g->seed = seed;
  g->stream = stream;
  g->counter = 0;
  g->pos = 4;
}

static uint32_t gs_next(GenStream *g) {
  if (g->pos == 4) {
    aco_philox4x32_10(g->counter++, g->stream, 0x6E6E6E6Eu, 0u, g->seed, 0x9E4A7C15u, g->buf);
    g->pos = 0;
  }
  return g->buf[g->pos++];
}

/* Integer uniform in [lo, hi] by multiply-shift (Lemire): deterministic. */
static int gs_range(GenStream *g, int lo, int hi) {
  uint64_t span = (uint64_t)(hi - lo) + 1u;
  return lo + (int)(((uint64_t)gs_next(g) * span) >> 32);
}

/* Uniform double in (0,1) from 53 bits. */
static double gs_unit(GenStream *g) {
  uint64_t a = gs_next(g) >> 5, b = gs_next(g) >> 6;
  return ((double)(a * 67108864u + b) + 0.5) * (1.0 / 9007199254740992.0);
}

int vrp_capacity_for(long long total, int K, int slack_pct) {
  long long num = (long long)slack_pct * total;
  long long den = 100LL * K;
  long long q = (num + den - 1) / den;
  return (int)(q < 1 ? 1 : q);
}

static int finish_demands(VrpInstance *out, GenStream *g, int dmin, int dmax, int K,
                          int slack_pct) {
  long long total = 0;
  out->demand[0] = 0;
  int unit = 1;
  for (int i = 1; i <= out->n; ++i) {
    out->demand[i] = gs_range(g, dmin, dmax);
    total += out->demand[i];
    if (out->demand[i] != 1) unit = 0;
  }
  out->K = K;
  out->capacity = vrp_capacity_for(total, K, slack_pct);
  out->unit_demand = unit;
  return out->capacity <= VRP_COORD_MAX ? 0 : 1;
}

int vrp_generate_uniform(VrpInstance *out, int n, int K, int side, int dmin, int dmax,
                         int slack_pct, uint32_t seed) {
  if (n <= 0 || K <= 0 || side <= 0 || side > VRP_COORD_MAX || dmin < 0 || dmax < dmin)
    return 1;
  if (vrp_alloc(out, n)) return 2;
  GenStream gc, gd;
  gs_init(&gc, seed, 1u);
  gs_init(&gd, seed, 2u);
  out->x[0] = side / 2;
  out->y[0] = side / 2;
  for (int i = 1; i <= n; ++i) {
    out->x[i] = gs_range(&gc, 0, side);
    out->y[i] = gs_range(&gc, 0, side);
  }
  snprintf(out->name, sizeof(out->name), "uniform_n%d_k%d_s%u", n, K, seed);
  return finish_demands(out, &gd, dmin, dmax, K, slack_pct);
}

int vrp_generate_clustered(VrpInstance *out, int n, int K, int side, int clusters,
                           double sigma, int dmin, int dmax, int slack_pct, uint32_t seed) {
  if (n <= 0 || K <= 0 || side <= 0 || side > VRP_COORD_MAX || clusters <= 0 || dmin < 0 ||
      dmax < dmin)
    return 1;
  if (vrp_alloc(out, n)) return 2;
  GenStream gc, gd, gp;
  gs_init(&gc, seed, 3u);
  gs_init(&gp, seed, 4u);
  gs_init(&gd, seed, 2u);
  double *cx = (double *)malloc(sizeof(double) * (size_t)clusters);
  double *cy = (double *)malloc(sizeof(double) * (size_t)clusters);
  if (!cx || !cy) {
    free(cx);
    free(cy);
    vrp_free(out);
    return 2;
  }
  for (int c = 0; c < clusters; ++c) {
    cx[c] = gs_unit(&gc) * side;
    cy[c] = gs_unit(&gc) * side;
  }
  out->x[0] = side / 2;
  out->y[0] = side / 2;
  const double two_pi = 6.283185307179586;
  for (int i = 1; i <= n; ++i) {
    int c = gs_range(&gp, 0, clusters - 1);
    /* Box-Muller pair */
    double u1 = gs_unit(&gp), u2 = gs_unit(&gp);
    double r = sqrt(-2.0 * log(u1));
    double px = cx[c] + sigma * r * cos(two_pi * u2);
    double py = cy[c] + sigma * r * sin(two_pi * u2);
    if (px < 0) px = 0;
    if (px > side) px = side;
    if (py < 0) py = 0;
    if (py > side) py = side;
    out->x[i] = (int32_t)floor(px + 0.5);
    out->y[i] = (int32_t)floor(py + 0.5);
  }
  free(cx);
  free(cy);
  snprintf(out->name, sizeof(out->name), "clustered_n%d_k%d_s%u", n, K, seed);
  return finish_demands(out, &gd, dmin, dmax, K, slack_pct);
}
