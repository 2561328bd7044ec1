/*
 * aco_construct_tab.hip - construction kernels whose fallback roulette
 * reads its weights from the pheromone row and the taupow / eta tables.
 */
#include "aco_construct.h"

#define ACO_CONSTRUCT_VARIANTS(R)                                                          \
  template __global__ void k_construct<true, false, R, false, false>(DevCtx, int, int, const int *);  \
  template __global__ void k_construct<false, false, R, false, false>(DevCtx, int, int, const int *); \
  template __global__ void k_construct<true, true, R, false, false>(DevCtx, int, int, const int *);   \
  template __global__ void k_construct<false, true, R, false, false>(DevCtx, int, int, const int *);  \
  template __global__ void k_construct<true, false, R, true, false>(DevCtx, int, int, const int *);   \
  template __global__ void k_construct<false, false, R, true, false>(DevCtx, int, int, const int *);
ACO_CONSTRUCT_VARIANTS(1)
ACO_CONSTRUCT_VARIANTS(2)
ACO_CONSTRUCT_VARIANTS(3)
ACO_CONSTRUCT_VARIANTS(4)
#undef ACO_CONSTRUCT_VARIANTS

#define ACO_PAIR_VARIANTS(RG, B8)                                                          \
  template __global__ void k_construct_pair<true, false, false, false, RG, B8>(DevCtx, int, int, const int *);  \
  template __global__ void k_construct_pair<false, false, false, false, RG, B8>(DevCtx, int, int, const int *); \
  template __global__ void k_construct_pair<true, true, false, false, RG, B8>(DevCtx, int, int, const int *);   \
  template __global__ void k_construct_pair<false, true, false, false, RG, B8>(DevCtx, int, int, const int *);  \
  template __global__ void k_construct_pair<true, false, true, false, RG, B8>(DevCtx, int, int, const int *);   \
  template __global__ void k_construct_pair<false, false, true, false, RG, B8>(DevCtx, int, int, const int *);
ACO_PAIR_VARIANTS(false, false)
ACO_PAIR_VARIANTS(true, false)
ACO_PAIR_VARIANTS(false, true)
ACO_PAIR_VARIANTS(true, true)
#undef ACO_PAIR_VARIANTS
