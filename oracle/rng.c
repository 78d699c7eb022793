/*
 * rng.c -- oracle restatement of the reference host generators.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows PROGRAM/src/random_generator.cu and PROGRAM/include/random_generator.h.
 * Seeds are never 0 here: the reference maps seed 0 to /dev/urandom
 * (random_generator.cu:11-16), which no test can reproduce.
 */
#include "oracle.h"

#define LCG_A 3202034522624059733ULL   /* random_generator.cu:37, random_generator.h:77 */

/* Init_luxury, random_generator.cu:27-55 */
void orc_lux_init(orc_lux *g, uint64_t seed)
{
  g->ip = 128;
  g->ip1 = (uint8_t)(g->ip - 10);
  g->ip2 = (uint8_t)(g->ip - 24);
  for (int i = 0; i < 1111; i++) seed = seed * LCG_A + 1;
  for (int j = g->ip2; j < g->ip; j++) {
    uint64_t v = 0;
    for (int i = 0; i < 3; i++) {
      seed = seed * LCG_A + 1;
      v |= (seed >> 56) << (8 * i);
    }
    g->x[j] = (uint32_t)v;
  }
  seed = seed * LCG_A + 1;
  g->borrow = (seed >> 63) ? 1 : 0;
  for (int i = 0; i < 1111; i++) (void)orc_lux_next(g);
}

/* luxury, random_generator.cu:57-75: 389 SWB steps per returned value; the
 * value returned is the slot at the write pointer after the last step. */
uint32_t orc_lux_next(orc_lux *g)
{
  for (int i = 0; i < 389; i++) {
    int32_t d = (int32_t)(g->x[g->ip1] - g->x[g->ip2] - (uint32_t)g->borrow);
    g->x[g->ip] = (uint32_t)((d + 16777216) & 16777215);
    g->borrow = d < 0;
    g->ip++; g->ip1++; g->ip2++;
  }
  return g->x[g->ip];
}

/* Luescher_PRNG, random_generator.cu:77-91: top byte of 8 luxury draws */
uint64_t orc_lux_u64(orc_lux *g)
{
  uint64_t r = 0;
  for (int j = 0; j < 8; j++) r |= ((uint64_t)(orc_lux_next(g) >> 16)) << (8 * j);
  return r;
}

/* _actualiza_aleatorio_HQ_escalar, random_generator.h:71-76 */
void orc_prlcg_step(orc_prlcg *g)
{
  g->ira[g->ip] = g->ira[g->ip1] + g->ira[g->ip2];
  g->pr = g->ira[g->ip] ^ g->ira[g->ip3];
  g->zseed = g->zseed * LCG_A + 1;
  g->out = g->zseed + g->pr;
  g->ip++; g->ip1++; g->ip2++; g->ip3++;
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* _actualiza_xoshiro256pp, random_generator.h:58-67 (Vigna's xoshiro256++) */
void orc_xoshiro_step(orc_xoshiro *g)
{
  uint64_t *s = g->s;
  g->out = rotl64(s[0] + s[3], 23) + s[0];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
}

static void prlcg_pointers(orc_prlcg *g)
{
  g->ip = 128;
  g->ip1 = (uint8_t)(g->ip - 24);
  g->ip2 = (uint8_t)(g->ip - 55);
  g->ip3 = (uint8_t)(g->ip - 61);
}

/* Init_Rand_HQ_64bits, random_generator.cu:160-200 */
void orc_prlcg_init_hq(orc_prlcg *g, uint64_t seed)
{
  for (int i = 0; i < 256; i++) g->ira[i] = 0;
  g->pr = g->out = 0;
  g->zseed = seed;
  for (int i = 0; i < 111; i++) g->zseed = g->zseed * LCG_A + 1;
  for (int k = 128 - 61; k < 128; k++) {
    g->ira[k] = 0;
    for (int i = 0; i < 8; i++) {
      g->zseed = g->zseed * LCG_A + 1;
      g->ira[k] |= (g->zseed >> 56) << (8 * i);
    }
  }
  prlcg_pointers(g);
  for (int i = 0; i < 1111; i++) {          /* lagged sum only, no LCG */
    g->ira[g->ip] = g->ira[g->ip1] + g->ira[g->ip2];
    g->ip++; g->ip1++; g->ip2++; g->ip3++;
  }
  for (int i = 0; i < 1111; i++) orc_prlcg_step(g);
}

/* Inicia_generadores_CPU, random_generator.cu:94-158 */
void orc_init_cpu_generators(orc_prlcg *pr, orc_xoshiro *xo, uint32_t *key, uint64_t seed)
{
  orc_lux lx;
  orc_lux_init(&lx, seed);
  for (int i = 0; i < 1111; i++) (void)orc_lux_next(&lx);

  for (int i = 0; i < 256; i++) pr->ira[i] = 0;
  pr->pr = pr->out = 0;
  pr->zseed = orc_lux_u64(&lx);
  for (int k = 128 - 61; k < 128; k++) pr->ira[k] = orc_lux_u64(&lx);
  prlcg_pointers(pr);
  for (int i = 0; i < 2222; i++) orc_prlcg_step(pr);

  uint64_t st[4] = {0, 0, 0, 0};
  while (!(st[0] | st[1] | st[2] | st[3])) {
    for (int j = 0; j < 4; j++) {
      st[j] = 0;
      for (int i = 0; i < 64; i++)
        st[j] |= (((uint64_t)orc_lux_next(&lx) >> 23) & 1ULL) << i;
    }
  }
  for (int j = 0; j < 4; j++) xo->s[j] = st[j];
  xo->out = 0;

  uint32_t k = 0;
  for (int i = 0; i < 32; i++) k |= ((orc_lux_next(&lx) >> 23) & 1U) << i;
  *key = k;
}

/* generate_seeds_from_one, random_generator.cu:202-234.  The draw at line 213
 * is overwritten at line 215, so the skip length is always 109. */
void orc_seeds_from_one(uint64_t seed, uint64_t *out, int n)
{
  orc_lux lx;
  orc_lux_init(&lx, seed);
  for (int i = 0; i < 1109; i++) (void)orc_lux_next(&lx);
  for (int k = 0; k < n; k++) {
    (void)orc_lux_u64(&lx);
    uint32_t sel = (uint32_t)((LCG_A + 1) >> 61);
    int skip = sel == 0 ? 709 : sel == 1 ? 109 : sel == 2 ? 1373 : 2099;
    for (int i = 0; i < skip; i++) (void)orc_lux_next(&lx);
    uint64_t v = orc_lux_u64(&lx);
    out[k] = v * 6364136223846793005ULL + 14426950408889634071ULL;
  }
}

/* ini_random, random_generator.cu:238-256 (JANUS build) */
void orc_janus_init(orc_janus *g, uint64_t seed)
{
  for (int i = 0; i < 256; i++) g->ira[i] = 0;
  g->zseed = seed;
  for (int i = 0; i < 11; i++) g->zseed = g->zseed * LCG_A + 1;
  g->ip = 128;
  g->ip1 = (uint8_t)(g->ip - 24);
  g->ip2 = (uint8_t)(g->ip - 55);
  g->ip3 = (uint8_t)(g->ip - 61);
  for (int i = g->ip3; i < g->ip; i++) {
    g->zseed = g->zseed * LCG_A + 1;
    g->ira[i] = g->zseed;
  }
  for (int i = 0; i < 1111; i++) {         /* PRRANDOM, value discarded */
    g->ira[g->ip] = g->ira[g->ip1] + g->ira[g->ip2];
    g->ip++; g->ip1++; g->ip2++; g->ip3++;
  }
}

/* HIRANDOM = PRRANDOM + CGRANDOM, random_generator.h:83-85 */
uint64_t orc_janus_next(orc_janus *g)
{
  uint64_t v = g->ira[g->ip1] + g->ira[g->ip2];
  g->ira[g->ip] = v;
  uint64_t pr = v ^ g->ira[g->ip3];
  g->ip++; g->ip1++; g->ip2++; g->ip3++;
  g->zseed = g->zseed * LCG_A + 1;
  return pr + g->zseed;
}

/* _fill_bits, fill_bits.h:47-68, for one replica: iteration counter in .x
 * (15 bits above the internal-call and parity bits) and .y (25 bits), 52 bits
 * of host entropy spread over .x/.y/.z/.w, the job id in the low 14 bits of .w. */
void orc_fill_counter(orc_prlcg *pr, orc_xoshiro *xo, uint64_t iter, uint64_t parity,
                      uint32_t jobid, uint32_t c[4])
{
  orc_xoshiro_step(xo);
  orc_prlcg_step(pr);
  uint64_t e = pr->out + xo->out;
  e >>= 64 - 52;
  c[1] = (uint32_t)(iter >> 15);
  c[0] = (uint32_t)(((iter & 32767ULL) << 3) | ((parity & 1ULL) << 2));
  c[0] |= (uint32_t)((e & 0x3FFFULL) << 18);
  e >>= 14;
  c[1] |= (uint32_t)((e & 0x7FULL) << 25);
  e >>= 7;
  c[2] = (uint32_t)((e & 0x1FFFULL) << 19);
  e >>= 13;
  c[3] = ((uint32_t)(e << 14)) | jobid;
}
