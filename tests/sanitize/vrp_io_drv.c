/* The .vrp reader and writer under AddressSanitizer / UBSan (built and run
 * by tests/test_sanitizers.py): every file named on the command line is
 * read and, when valid, written and read back.  Exit status 1 when a round
 * trip disagrees. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "aco_host.h"
int main(int argc, char **argv) {
  int bad = 0;
  const char *rt = getenv("VRP_RT_PATH") ? getenv("VRP_RT_PATH") : "rt.vrp";
  for (int i = 1; i < argc; ++i) {
    VrpInstance v;
    memset(&v, 0, sizeof v);
    int rc = vrp_read(argv[i], &v, 0);
    printf("%s rc=%d n=%d\n", argv[i], rc, rc ? -1 : v.n);
    if (!rc) {
      rc = vrp_write(rt, &v);
      VrpInstance w;
      memset(&w, 0, sizeof w);
      if (rc || vrp_read(rt, &w, 0) || w.n != v.n) bad = 1;
      vrp_free(&w);
      vrp_free(&v);
    }
  }
  return bad;
}
