/*
 * ref_omp_l3_driver.c - the reference's OpenMP solver (src/openmp-mpi/
 * aco_parallel.c, included below as it lies) with its L3-size probe
 * (get_l3_cache_size, aco_parallel.c:145-159) answered from the environment
 * variable REF_L3_SIZE (same text format as the sysfs file, e.g. "224K")
 * instead of /sys.  The probe sets the reference's candidate width,
 * k = 0.7 * L3 / ((n+1) * 8) clamped to [16, 512] (tune_candidate_k,
 * aco_parallel.c:167-174): "224K" gives k = 20 at n = 1000, the width this
 * solver uses, so the two can be timed at equal candidate width.  Without
 * REF_L3_SIZE the sysfs file is read as in the shipped binary.  Linked with
 * the reference's src/main.c by oracle/ref.mk; a timed CPU baseline only.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static FILE *ref_l3_probe_open(const char *path, const char *mode);
#define fopen ref_l3_probe_open
#include "aco_parallel.c"
#undef fopen

static FILE *ref_l3_probe_open(const char *path, const char *mode) {
  const char *fixed = getenv("REF_L3_SIZE");
  if (fixed && strcmp(path, "/sys/devices/system/cpu/cpu0/cache/index3/size") == 0)
    return fmemopen((void *)fixed, strlen(fixed), mode);
  return fopen(path, mode);
}
