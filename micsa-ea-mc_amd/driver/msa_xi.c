/*
 * msa_xi.c -- offline overlap correlation function and correlation lengths
 * from the conf files a run wrote (the analysis behind the reference's
 * FIGURES/fig2: C4(r), xi_12, xi_23; the reference ships no code for it).
 *
 * usage: msa_xi L nr BITdir [BITdir ...]
 *
 * Every BITdir (one sample, <dir>I%03d/BIT%03d/) contributes its replica
 * pairs (0,1), (2,3), ... read from R%02d/conf; all confs must hold the same
 * iteration.  Prints "r C4(r)" for r = 0..L/2, then a line
 * "# now <t> terms <n> rc <rc> xi12 <x> xi23 <x>".
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "msa_host.h"

int main(int argc, char **argv)
{
  if (argc < 4) {
    fprintf(stderr, "usage: %s L nr BITdir [BITdir ...]\n", argv[0]);
    return 1;
  }
  const int L = atoi(argv[1]), R = atoi(argv[2]);
  if (L < 8 || L % 8 || R < 2 || R > MSA_MAXR) {
    fprintf(stderr, "L must be a multiple of 8 and 2 <= nr <= %d\n", MSA_MAXR);
    return 1;
  }
  const int nr = L / 2 + 1;
  const uint64_t V8 = (uint64_t)L * L * L / 8;
  uint8_t *ja = (uint8_t *)malloc(V8), *jb = (uint8_t *)malloc(V8);
  int64_t *diff = (int64_t *)calloc(3 * (size_t)nr, sizeof(int64_t));
  int64_t *sum = (int64_t *)calloc(3 * (size_t)nr, sizeof(int64_t));
  double *c4 = (double *)malloc(sizeof(double) * nr);
  if (!ja || !jb || !diff || !sum || !c4) {
    fprintf(stderr, "out of memory\n");
    return 1;
  }
  char e[512], path[4096];
  long long now = -1;
  int64_t terms = 0;
  for (int a = 3; a < argc; a++)
    for (int p = 0; 2 * p + 1 < R; p++) {
      uint8_t *jj[2] = {ja, jb};
      for (int h = 0; h < 2; h++) {
        snprintf(path, sizeof path, "%s/R%02d/conf", argv[a], 2 * p + h);
        msa_sdata d;
        double beta;
        msa_prng pr;
        msa_xo256 xo;
        int32_t en;
        if (msa_conf_read(path, L, &d, &beta, &pr, &xo, jj[h], &en, e, sizeof e)) {
          fprintf(stderr, "%s\n", e);
          return 1;
        }
        if (now >= 0 && (long long)d.now != now) {
          fprintf(stderr, "%s holds iteration %llu, others %lld\n", path,
                  (unsigned long long)d.now, now);
          return 1;
        }
        now = (long long)d.now;
      }
      if (msa_overlap_counts(L, ja, jb, diff)) return 1;
      for (int k = 0; k < 3 * nr; k++) sum[k] += diff[k];
      terms++;
    }
  msa_c4_from_counts(L, terms, sum, c4);
  double xi12, xi23;
  const int rc = msa_xi_integrals(nr, c4, &xi12, &xi23);
  for (int r = 0; r < nr; r++) printf("%d %.10g\n", r, c4[r]);
  printf("# now %lld terms %lld rc %d xi12 %.10g xi23 %.10g\n", now, (long long)terms, rc, xi12,
         xi23);
  return 0;
}
