/*
 * lut.c -- oracle restatement of the quad-precision tables: the heat-bath
 * LUT of PROGRAM/input/LUT/preparation.c:265-388 and the demon thresholds of
 * PROGRAM/src/ini.c:680-710.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The LUT generator is written for a general Bernoulli parameter q1: the
 * reference builds it for q1 = e^{-4 beta} (preparation.c:282-284); the demon
 * refresh uses q1 = p/(1+p) for bit plane 0 and q1 = p^2/(1+p^2) for bit
 * plane 1, with p = e^{-4 beta}, which makes D = D0 + 2 D1 heat-bath
 * distributed, P(D) ~ p^D for D in 0..3.
 */
#include <quadmath.h>
#include <string.h>
#include "oracle.h"

#define TWO64Q 18446744073709551616.Q

/* prepara_LUT, preparation.c:265-388, with qprob1 = q1, qprob0 = 1 - q1 */
void orc_lut_build_q(const void *q1p, orc_lut *out)
{
  __float128 q1 = *(const __float128 *)q1p;
  __float128 q0 = 1.0Q - q1;
  uint64_t caso[9], cum[256];
  for (int i = 0; i < 9; i++) {
    __float128 p = powq(q0, 8 - i) * powq(q1, i);
    caso[i] = (uint64_t)(TWO64Q * p);
  }
  cum[0] = caso[0];
  for (int i = 1; i < 255; i++) cum[i] = cum[i - 1] + caso[__builtin_popcount(i)];
  cum[255] = ~0ULL;
  for (int i = 0; i < 256; i++) {
    out->umbrales[2 * i] = (uint32_t)(cum[i] >> 32);
    out->umbrales[2 * i + 1] = (uint32_t)(cum[i] & 0xFFFFFFFFULL);
  }
  /* presearch over the 10 leading bits, preparation.c:340-376 */
  const uint64_t tail = (1ULL << (64 - 10)) - 1;
  uint32_t tmin[1024], tmax[1024];
  for (uint64_t msb = 0; msb < 1024; msb++) {
    uint64_t t = msb << (64 - 10);
    int i;
    for (i = 0; i < 256; i++) if (t < cum[i]) break;
    tmin[msb] = (uint32_t)(i > 255 ? 255 : i);
    t |= tail;
    for (i = 0; i < 256; i++) if (t < cum[i]) break;
    tmax[msb] = (uint32_t)(i > 255 ? 255 : i);
  }
  for (int i = 0; i < 1024; i += 2)
    out->prebusqueda[i >> 1] = tmin[i] | (tmax[i] << 8) | (tmin[i + 1] << 16) | (tmax[i + 1] << 24);
}

void orc_lut_build_beta(double beta, orc_lut *out)
{
  __float128 q1 = expq(-4.0Q * (__float128)beta);
  orc_lut_build_q(&q1, out);
}

void orc_lut_build_demon(double beta, orc_lut *lut_d0, orc_lut *lut_d1)
{
  __float128 p = expq(-4.0Q * (__float128)beta);
  __float128 a = p / (1.0Q + p);
  __float128 p2 = p * p;
  __float128 b = p2 / (1.0Q + p2);
  orc_lut_build_q(&a, lut_d0);
  orc_lut_build_q(&b, lut_d1);
}

/* get_thresholds, ini.c:680-710 */
void orc_thresholds(double beta, uint64_t thr[3])
{
  __float128 f = expq(-4.0Q * (__float128)beta);
  __float128 prob[4], norm;
  prob[0] = 1.0Q;
  norm = prob[0];
  for (int i = 1; i < 4; i++) { prob[i] = prob[i - 1] * f; norm += prob[i]; }
  for (int i = 0; i < 4; i++) prob[i] /= norm;
  __float128 acc[3];
  acc[0] = prob[3];
  acc[1] = prob[2] + acc[0];
  acc[2] = prob[1] + acc[1];
  for (int i = 0; i < 3; i++) thr[i] = (uint64_t)(TWO64Q * acc[i]);
}
