/*
 * msa_tables.c -- quad-precision tables and annealing schedules.
 *
 * Heat-bath LUT: the table of PROGRAM/input/LUT/preparation.c:265-388 (written
 * there for q = e^{-4 beta}), here for a general Bernoulli parameter q so the
 * demon refresh can draw its two bit planes from it.  Thresholds:
 * get_thresholds, PROGRAM/src/ini.c:680-710.
 */
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

#include "msa_host.h"

#define TWO64Q 18446744073709551616.Q

/* cumulative thresholds of the 256 bytes of eight iid Bernoulli(q1) bits:
 * width of byte n = floor(2^64 q1^k (1-q1)^(8-k)), k = popcount(n); the last
 * threshold is 2^64-1. */
static void lut_from_q(__float128 q1, msa_lut *out)
{
  const __float128 q0 = 1.0Q - q1;
  uint64_t width[9], cum[256];
  for (int k = 0; k <= 8; k++) width[k] = (uint64_t)(TWO64Q * (powq(q0, 8 - k) * powq(q1, k)));
  cum[0] = width[0];
  for (int n = 1; n < 255; n++) cum[n] = cum[n - 1] + width[__builtin_popcount(n)];
  cum[255] = ~0ULL;
  for (int n = 0; n < 256; n++) {
    out->umbrales[2 * n] = (uint32_t)(cum[n] >> 32);
    out->umbrales[2 * n + 1] = (uint32_t)cum[n];
  }
  /* presearch: for each 10-bit prefix m, the first byte whose threshold lies
   * above the smallest / the largest 64-bit value with that prefix.  Both
   * queries grow with m, so one forward pointer per query suffices. */
  const uint64_t tail = (1ULL << 54) - 1;
  int lo = 0, hi = 0;
  uint16_t *pre = (uint16_t *)out->prebusqueda;
  for (uint64_t m = 0; m < 1024; m++) {
    const uint64_t a = m << 54, b = a | tail;
    while (lo < 255 && !(a < cum[lo])) lo++;
    while (hi < 255 && !(b < cum[hi])) hi++;
    pre[m] = (uint16_t)(lo | (hi << 8));
  }
}

void msa_lut_beta(double beta, msa_lut *out)
{
  lut_from_q(expq(-4.0Q * (__float128)beta), out);
}

void msa_lut_demon(double beta, msa_lut *plane0, msa_lut *plane1)
{
  const __float128 p = expq(-4.0Q * (__float128)beta);
  const __float128 p2 = p * p;
  lut_from_q(p / (1.0Q + p), plane0);
  lut_from_q(p2 / (1.0Q + p2), plane1);
}

void msa_thresholds(double beta, uint64_t thr[3])
{
  const __float128 f = expq(-4.0Q * (__float128)beta);
  __float128 w[4] = {1.0Q, 0, 0, 0}, z = 1.0Q;
  for (int d = 1; d < 4; d++) {
    w[d] = w[d - 1] * f;
    z += w[d];
  }
  for (int d = 0; d < 4; d++) w[d] /= z;
  /* P(D >= 3), P(D >= 2), P(D >= 1) */
  const __float128 c3 = w[3], c2 = w[2] + c3, c1 = w[1] + c2;
  thr[0] = (uint64_t)(TWO64Q * c3);
  thr[1] = (uint64_t)(TWO64Q * c2);
  thr[2] = (uint64_t)(TWO64Q * c1);
}

/* ------------------------------------------------------------ schedule -- */

int64_t msa_sched_count(const msa_sched *s, int64_t niter)
{
  if (niter <= 0) return 0;
  if (s->refresh_every > 0) return (niter - 1) / s->refresh_every + 1;
  if (s->refresh_every == -1) {
    int64_t n = 1;                                  /* t = 0 */
    for (int64_t t = 1; t < niter; t <<= 1) n++;    /* t = 1, 2, 4, ... */
    return n;
  }
  return 0;
}

int64_t msa_sched_time(const msa_sched *s, int64_t k)
{
  if (s->refresh_every > 0) return k * s->refresh_every;
  return k == 0 ? 0 : (int64_t)1 << (k - 1);
}

double msa_sched_beta(const msa_sched *s, int64_t k)
{
  if (s->T0 <= 0) return s->beta_fixed;
  int64_t tk = msa_sched_time(s, k);
  if (tk > s->nsweeps) tk = s->nsweeps;   /* past the end of the anneal T stays at T1 */
  const double T = s->T0 + (s->T1 - s->T0) * (double)tk / (double)s->nsweeps;
  return 1.0 / T;
}

int msa_sched_tables_resume(const msa_sched *s, int64_t k0, int64_t nref, int R, msa_prng *pr,
                            msa_xo256 *xo, uint32_t jobid, msa_u32x4 *ctr, msa_lut *luts,
                            int32_t *lut_idx)
{
  if (R < 1 || R > MSA_MAXR || nref < 0 || k0 < 0) return -1;
  /* refreshes at the same inverse temperature share one LUT pair (a schedule
   * is monotone, so equal betas are consecutive): a fixed-beta run, or the
   * part of a run past the end of its anneal, builds one pair */
  double *beta = (double *)malloc(sizeof(double) * (size_t)(nref > 0 ? nref : 1));
  if (!beta) return -1;
  int64_t npair = 0;
  for (int64_t k = 0; k < nref; k++) {
    const uint64_t t = (uint64_t)msa_sched_time(s, k);
    for (int r = 0; r < R; r++) {
      if (k >= k0)
        msa_fill_counter(&pr[r], &xo[r], t, 0, jobid, &ctr[k * R + r]);
      else
        memset(&ctr[k * R + r], 0, sizeof(msa_u32x4));
    }
    const double b = msa_sched_beta(s, k);
    if (npair == 0 || b != beta[npair - 1]) beta[npair++] = b;
    lut_idx[k] = (int32_t)(npair - 1);
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t n = 0; n < npair; n++) msa_lut_demon(beta[n], &luts[2 * n], &luts[2 * n + 1]);
  free(beta);
  return (int)npair;
}

int msa_sched_tables_from(const msa_sched *s, int64_t nref, int R, msa_prng *pr, msa_xo256 *xo,
                          uint32_t jobid, msa_u32x4 *ctr, msa_lut *luts, int32_t *lut_idx)
{
  return msa_sched_tables_resume(s, 0, nref, R, pr, xo, jobid, ctr, luts, lut_idx);
}

void msa_gen_skip(msa_prng *pr, msa_xo256 *xo, int R, int64_t n)
{
  msa_u32x4 c;
  for (int64_t k = 0; k < n; k++)
    for (int r = 0; r < R; r++) msa_fill_counter(&pr[r], &xo[r], 0, 0, 0, &c);
}

int msa_sched_tables(const msa_sched *s, int64_t nref, int R, const uint64_t *seed_MC,
                     uint32_t jobid, msa_u32x4 *ctr, uint32_t *keys, msa_lut *luts,
                     int32_t *lut_idx)
{
  if (R < 1 || R > MSA_MAXR || nref < 0) return -1;
  msa_prng pr[MSA_MAXR];
  msa_xo256 xo[MSA_MAXR];
#pragma omp parallel for schedule(static)
  for (int r = 0; r < R; r++) msa_cpu_generators(&pr[r], &xo[r], &keys[r], seed_MC[r]);
  return msa_sched_tables_from(s, nref, R, pr, xo, jobid, ctr, luts, lut_idx);
}

void msa_thermal_counters(msa_prng *pr, msa_xo256 *xo, int R, int64_t nterm, uint32_t jobid,
                          msa_u32x4 *ctr)
{
  for (int64_t t = 0; t < nterm; t++)
    for (uint32_t p = 0; p < 2; p++)
      for (int r = 0; r < R; r++)
        msa_fill_counter(&pr[r], &xo[r], (uint64_t)(2 * t + p), p, jobid, &ctr[(t * 2 + p) * R + r]);
}

/* ------------------------------------------------- shared grid tables -- */

void msa_lut_grid(double tgrid, int32_t nlo, int32_t nlut, msa_lut *luts)
{
#pragma omp parallel for schedule(dynamic, 8)
  for (int32_t n = 0; n < nlut; n++)
    msa_lut_demon(1.0 / (tgrid * (double)(nlo + n)), &luts[2 * n], &luts[2 * n + 1]);
}

int msa_sched_tables_grid(const msa_sched *s, int64_t nref, int R, const uint64_t *seed_MC,
                          uint32_t jobid, double tgrid, int32_t nlo, int32_t nlut,
                          msa_u32x4 *ctr, uint32_t *keys, int32_t *lut_idx)
{
  if (R < 1 || R > MSA_MAXR || nref < 0 || tgrid <= 0 || nlut < 1) return -1;
  msa_prng pr[MSA_MAXR];
  msa_xo256 xo[MSA_MAXR];
  for (int r = 0; r < R; r++) msa_cpu_generators(&pr[r], &xo[r], &keys[r], seed_MC[r]);
  for (int64_t k = 0; k < nref; k++) {
    const uint64_t t = (uint64_t)msa_sched_time(s, k);
    for (int r = 0; r < R; r++) msa_fill_counter(&pr[r], &xo[r], t, 0, jobid, &ctr[k * R + r]);
    const double T = 1.0 / msa_sched_beta(s, k);
    int64_t n = (int64_t)floor(T / tgrid + 0.5) - nlo;
    if (n < 0) n = 0;
    if (n >= nlut) n = nlut - 1;
    lut_idx[k] = (int32_t)n;
  }
  return 0;
}
