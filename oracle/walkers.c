/*
 * walkers.c -- oracle restatement of the reference's demon ("walker")
 * initialisation, InitWalkers (PROGRAM/src/ini.c:712-781), for one replica of
 * one sample.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference fills TOTAL_WALKERS_WORDS = MSC_V * nc words per replica,
 * word iword at CPUwalker[iword * NR + ir] (ini.c:750-775), each from 64 draws
 * in bit order.  A draw steps xoshiro256++ and then the PR+LCG generator,
 * e = PR+LCG final + xoshiro final, and the demon is the number of
 * thresholds e lies below (thresholds from get_thresholds, ini.c:680-710).
 * Bit 0 of the demon goes to walker0, bit 1 to walker1.  The generators are
 * left advanced, as the reference keeps them for the next refresh.
 */
#include "oracle.h"

void orc_init_walkers(orc_prlcg *pr, orc_xoshiro *xo, const uint64_t thr[3], uint64_t nwords,
                      uint64_t *w0, uint64_t *w1)
{
  for (uint64_t iw = 0; iw < nwords; iw++) {
    uint64_t d0 = 0, d1 = 0;
    for (int bit = 0; bit < 64; bit++) {
      orc_xoshiro_step(xo);
      orc_prlcg_step(pr);
      const uint64_t e = pr->out + xo->out;
      const uint64_t w = (uint64_t)(e < thr[0]) + (uint64_t)(e < thr[1]) + (uint64_t)(e < thr[2]);
      d0 |= (w & 1u) << bit;
      d1 |= ((w >> 1) & 1u) << bit;
    }
    w0[iw] = d0;
    w1[iw] = d1;
  }
}
