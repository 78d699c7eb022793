/*
 * philox.c -- oracle restatement of the reference counter-based bits:
 * Philox4x32-10, the lsb mixing, the LUT bisection and the 32-bit word of
 * eight-per-byte Bernoulli bits.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows PROGRAM/src/main.cu:546-702.
 */
#include "oracle.h"

/* _update_state / _update_key / _philox_4x32_10, main.cu:562-595:
 * ten rounds, nine key bumps between them. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t v0 = ctr[0], v1 = ctr[1], v2 = ctr[2], v3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; round++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * v0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * v2;
    uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    v0 = hi1 ^ v1 ^ k0;
    v1 = lo1;
    v2 = hi0 ^ v3 ^ k1;
    v3 = lo0;
    if (round < 9) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
  }
  out[0] = v0; out[1] = v1; out[2] = v2; out[3] = v3;
}

/* R < umbral[n] as a 64-bit comparison split in two 32-bit halves,
 * the predicate inside _bisection, main.cu:615-630 */
static inline int below(const orc_lut *lut, uint32_t n, uint32_t rmsb, uint32_t rlsb)
{
  uint32_t um = lut->umbrales[2 * n], ul = lut->umbrales[2 * n + 1];
  return (rmsb == um) ? (rlsb < ul) : (rmsb < um);
}

static inline uint32_t presearch(const orc_lut *lut, uint32_t idx)
{
  uint32_t w = lut->prebusqueda[idx >> 1];
  return (idx & 1) ? (w >> 16) : (w & 0xFFFFu);
}

/* _bisection, main.cu:615-630: returns the byte n with
 * umbral[n-1] <= R < umbral[n] (umbral[-1] = 0, umbral[255] = 2^64-1). */
uint32_t orc_bisection(const orc_lut *lut, uint32_t rmsb, uint32_t rlsb)
{
  uint32_t msb = rmsb >> (32 - 10);
  uint32_t lo, hi;
  if (below(lut, 254, rmsb, rlsb)) {
    uint32_t e = presearch(lut, msb);
    lo = e & 255;
    hi = e >> 8;
  } else {
    lo = hi = 255;
  }
  while (hi - lo > 1) {
    uint32_t mid = (hi + lo) >> 1;
    if (below(lut, mid, rmsb, rlsb)) hi = mid; else lo = mid;
  }
  return below(lut, lo, rmsb, rlsb) ? lo : hi;
}

/* _obten_aleatorio_contador + _get_rnd_word, main.cu:597-644.  The counter's
 * .x is incremented after the Philox call (main.cu:606).  For rot == 0 the
 * reference's (bb<<0)|(bb>>32) evaluates to bb on every target in use; the
 * shift by 32 is avoided here. */
uint32_t orc_rnd_word(const orc_lut *lut, uint32_t ctr[4], const uint32_t key[2])
{
  uint32_t v[4], lsb[4];
  orc_philox4x32_10(ctr, key, v);
  ctr[0]++;
  uint32_t a = v[0] ^ v[2];
  uint32_t b = v[1] ^ v[3];
  lsb[0] = ((v[1] << 17) | (v[0] >> 15)) ^ a ^ ((a << 21) | (b >> 11));
  lsb[1] = ((v[0] << 17) | (v[1] >> 15)) ^ b ^ (b << 21);
  lsb[2] = (a << 28) | (b >> 4);
  lsb[3] = (b << 28) | (a >> 4);
  uint32_t rot = (lsb[0] + v[1] + lsb[2] + v[3]) >> 27;
  uint32_t w = orc_bisection(lut, v[0], lsb[0]);
  w |= orc_bisection(lut, v[1], lsb[1]) << 8;
  w |= orc_bisection(lut, v[2], lsb[2]) << 16;
  w |= orc_bisection(lut, v[3], lsb[3]) << 24;
  if (rot) w = (w << rot) | (w >> (32 - rot));
  return w;
}
