/*
 * msa_corr.c -- offline observables from conf files: the replica-overlap
 * correlation function and the integral correlation lengths (the quantities
 * of the reference's figures, FIGURES/fig2/fig2.gpt:129-212, whose analysis
 * code the reference does not contain).
 *
 *   q_x = s^a_x s^b_x,  C4(r) = (1/3V) sum_mu sum_x q_x q_{x + r e_mu}
 *   I_k = integral_0^rc r^k C4(r) dr,  xi_{k,k+1} = I_{k+1} / I_k
 *
 * The counts are exact integers (bonds where q differs at distance r); the
 * integrals use the trapezoid rule on r = 0, 1, ..., up to rc = L/2 or the
 * last r before C4 first drops to zero or below.
 */
#include <stdlib.h>
#include <string.h>

#include "msa_host.h"

/* diff[mu * (L/2 + 1) + r] = #{x : q_x != q_{x + r e_mu}}, r = 0..L/2, for two
 * conf spin payloads (bit s & 7 of byte s >> 3, s = x + L (y + L z)). */
int msa_overlap_counts(int L, const uint8_t *janus_a, const uint8_t *janus_b, int64_t *diff)
{
  if (L < 2 || L % 8) return -1;
  const int64_t V = (int64_t)L * L * L, nr = L / 2 + 1;
  uint8_t *q = (uint8_t *)malloc((size_t)V);
  if (!q) return -1;
  for (int64_t s = 0; s < V; s++)
    q[s] = (uint8_t)(((janus_a[s >> 3] ^ janus_b[s >> 3]) >> (s & 7)) & 1);
  memset(diff, 0, sizeof(int64_t) * 3 * (size_t)nr);
  for (int mu = 0; mu < 3; mu++) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 1; r < nr; r++) {
      int64_t cnt = 0;
      for (int z = 0; z < L; z++)
        for (int y = 0; y < L; y++) {
          const uint8_t *row = q + (int64_t)L * (y + (int64_t)L * z);
          const uint8_t *nb;
          if (mu == 0) {
            for (int x = 0; x < L; x++) cnt += row[x] ^ row[(x + r) % L];
            continue;
          }
          nb = mu == 1 ? q + (int64_t)L * (((y + r) % L) + (int64_t)L * z)
                       : q + (int64_t)L * (y + (int64_t)L * ((z + r) % L));
          for (int x = 0; x < L; x++) cnt += row[x] ^ nb[x];
        }
      diff[mu * nr + r] = cnt;
    }
  }
  free(q);
  return 0;
}

/* C4(r) from counts summed over any number of (sample, replica pair) terms:
 * c4[r] = 1 - 2 sum_mu diff / (3 V nterms). */
void msa_c4_from_counts(int L, int64_t nterms, const int64_t *diff, double *c4)
{
  const int nr = L / 2 + 1;
  const double V = (double)L * L * L;
  for (int r = 0; r < nr; r++) {
    const double d = (double)(diff[r] + diff[nr + r] + diff[2 * nr + r]);
    c4[r] = 1.0 - 2.0 * d / (3.0 * V * (double)nterms);
  }
}

/* I_k by the trapezoid rule over r = 0 .. rc and the ratios xi12, xi23;
 * returns rc. */
int msa_xi_integrals(int nr, const double *c4, double *xi12, double *xi23)
{
  int rc = nr - 1;
  for (int r = 1; r < nr; r++)
    if (c4[r] <= 0) {
      rc = r - 1;
      break;
    }
  double I[4] = {0, 0, 0, 0};
  for (int k = 1; k <= 3; k++)
    for (int r = 0; r < rc; r++) {
      double a = c4[r], b = c4[r + 1], pa = 1, pb = 1;
      for (int j = 0; j < k; j++) {
        pa *= r;
        pb *= r + 1;
      }
      I[k] += 0.5 * (pa * a + pb * b);
    }
  *xi12 = I[1] > 0 ? I[2] / I[1] : 0;
  *xi23 = I[2] > 0 ? I[3] / I[2] : 0;
  return rc;
}
