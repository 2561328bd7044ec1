/* The oracle under AddressSanitizer / UBSan (built and run by
 * tests/test_sanitizers.py): seeded random instances, including
 * fallback-heavy (short candidate lists), capacity-infeasible and
 * two-chunk (n > 1024) ones, three epochs each, in both distance modes. */
#include <stdio.h>
#include <stdlib.h>
#include "aco_oracle.h"
static unsigned rs = 12345;
static int rnd(int m) { rs = rs * 1103515245u + 12345u; return (int)((rs >> 8) % (unsigned)m); }
int main(void) {
  const int sizes[][4] = {{5, 2, 20, 8}, {100, 8, 15, 64}, {1030, 9, 12, 16}, {300, 10, 32, 32},
                          {50, 2, 4, 16}, {2100, 12, 6, 8}};
  for (int t = 0; t < 12; ++t) {
    const int real = t >= 6;
    const int n = sizes[t % 6][0], K = sizes[t % 6][1], k = sizes[t % 6][2], m = sizes[t % 6][3];
    int32_t *x = malloc(sizeof(int32_t) * (n + 1)), *y = malloc(sizeof(int32_t) * (n + 1));
    int32_t *d = malloc(sizeof(int32_t) * (n + 1));
    long tot = 0;
    for (int i = 0; i <= n; ++i) { x[i] = rnd(1000); y[i] = rnd(1000); d[i] = i ? 1 + rnd(10) : 0; tot += d[i]; }
    int Q = (int)((tot * 11 / 10 + K - 1) / K);
    if (t % 6 == 4) Q = 3; /* capacity-short: infeasible */
    OracleColony *o = oracle_create(n, K, Q, x, y, d, k, m, 1.0f, 2.0f, 0.1f, 1.0f, 1.0f, 2.0f, 0, 7u + t, 0u,
                                    real);
    if (!o) { printf("create failed %d\n", t); return 1; }
    oracle_epochs(o, 3, 1);
    OracleState st;
    oracle_get_state(o, &st);
    printf("case %d n=%d real=%d epoch=%d best=%lld feasible_ants=%d\n", t, n, real, st.epoch,
           (long long)st.global_best_cost, st.feasible_ants);
    oracle_destroy(o);
    free(x); free(y); free(d);
  }
  return 0;
}
