/*
 * msa_rng.c -- host generators producing the same streams as the reference's
 * PROGRAM/src/random_generator.cu, so the same seeds give the same couplings,
 * spins, keys and Philox counters.  Seeds must be non-zero: the reference
 * replaces 0 by /dev/urandom (random_generator.cu:11-16), the driver does the
 * same before calling in here.
 */
#include <string.h>

#include "msa_host.h"

static const uint64_t MUL = 3202034522624059733ULL;   /* 64-bit LCG multiplier */

static inline uint64_t lcg(uint64_t v) { return v * MUL + 1; }
static inline uint64_t rol(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

/* Init_luxury, random_generator.cu:27-55: 24 words of 24 bits from the top
 * bytes of an LCG, borrow from the LCG's top bit, 1111 warm-up draws. */
void msa_swb_seed(msa_swb *g, uint64_t seed)
{
  memset(g, 0, sizeof(*g));
  g->p = 128;
  g->p10 = 118;
  g->p24 = 104;
  for (int i = 0; i < 1111; i++) seed = lcg(seed);
  for (int slot = 104; slot < 128; slot++) {
    uint32_t w = 0;
    for (int byte = 0; byte < 3; byte++) {
      seed = lcg(seed);
      w |= (uint32_t)(seed >> 56) << (8 * byte);
    }
    g->x[slot] = w;
  }
  seed = lcg(seed);
  g->borrow = (int32_t)(seed >> 63);
  for (int i = 0; i < 1111; i++) (void)msa_swb_next(g);
}

/* luxury, random_generator.cu:57-75: 389 subtract-with-borrow steps
 * x[p] = x[p-10] - x[p-24] - borrow (mod 2^24), then the word at p. */
uint32_t msa_swb_next(msa_swb *g)
{
  for (int step = 0; step < 389; step++) {
    int32_t d = (int32_t)(g->x[g->p10] - g->x[g->p24] - (uint32_t)g->borrow);
    g->x[g->p] = (uint32_t)(d + (1 << 24)) & ((1u << 24) - 1);
    g->borrow = d < 0;
    g->p++;
    g->p10++;
    g->p24++;
  }
  return g->x[g->p];
}

/* Luescher_PRNG, random_generator.cu:77-91 */
uint64_t msa_swb_u64(msa_swb *g)
{
  uint64_t v = 0;
  for (int byte = 0; byte < 8; byte++) v |= (uint64_t)(msa_swb_next(g) >> 16) << (8 * byte);
  return v;
}

static void prng_reset_pointers(msa_prng *g)
{
  g->p = 128;
  g->p24 = 104;
  g->p55 = 73;
  g->p61 = 67;
}

/* _actualiza_aleatorio_HQ_escalar, random_generator.h:71-76 */
uint64_t msa_prng_next(msa_prng *g)
{
  const uint64_t v = g->ring[g->p24] + g->ring[g->p55];
  g->ring[g->p] = v;
  const uint64_t pr = v ^ g->ring[g->p61];
  g->lcg = lcg(g->lcg);
  g->out = g->lcg + pr;
  g->p++;
  g->p24++;
  g->p55++;
  g->p61++;
  return g->out;
}

/* The lagged sum x_m = x_{m-24} + x_{m-55} (mod 2^64) is linear: with the
 * state (x_{m-55} .. x_{m-1}) as the basis t^0 .. t^54, x_{m-55+k} is the
 * state weighted by the coefficients of t^k mod (t^55 - t^31 - 1).  Products
 * of two such polynomials, reduced by t^i = t^(i-24) + t^(i-55) for i >= 55. */
#define LF_N 55
static void lf_mulmod(const uint64_t *a, const uint64_t *b, uint64_t *out)
{
  uint64_t p[2 * LF_N - 1];
  memset(p, 0, sizeof p);
  for (int i = 0; i < LF_N; i++)
    if (a[i])
      for (int j = 0; j < LF_N; j++) p[i + j] += a[i] * b[j];
  for (int i = 2 * LF_N - 2; i >= LF_N; i--) {
    p[i - 24] += p[i];
    p[i - 55] += p[i];
  }
  memcpy(out, p, sizeof(uint64_t) * LF_N);
}

/* t * a mod (t^55 - t^31 - 1) */
static void lf_shift(uint64_t *a)
{
  const uint64_t top = a[LF_N - 1];
  memmove(a + 1, a, sizeof(uint64_t) * (LF_N - 1));
  a[0] = top;
  a[31] += top;
}

/* msa_prng_next applied n times, by jump-ahead for large n: the lagged sum by
 * t^k mod its characteristic polynomial (above), the LCG v -> MUL v + 1 as an
 * affine map raised to the n-th power; the result equals stepping (the
 * 62 values the output needs, x_{m+n-62} .. x_{m+n-1}, are all recurrence
 * values once n >= 7, since they then start at or after x_{m-55}). */
void msa_prng_skip(msa_prng *g, uint64_t n)
{
  if (n < 64) {
    for (uint64_t i = 0; i < n; i++) (void)msa_prng_next(g);
    return;
  }
  uint64_t st[LF_N];                         /* x_{m-55} .. x_{m-1} */
  for (int j = 0; j < LF_N; j++) st[j] = g->ring[(uint8_t)(g->p - LF_N + j)];
  /* c = t^(n - 7) mod P: x_{m+n-62} = x_{m-55 + (n-7)} */
  uint64_t c[LF_N], base[LF_N];
  memset(c, 0, sizeof c);
  memset(base, 0, sizeof base);
  c[0] = 1;
  base[1] = 1;
  for (uint64_t e = n - 7; e; e >>= 1) {
    if (e & 1) lf_mulmod(c, base, c);
    if (e > 1) lf_mulmod(base, base, base);
  }
  uint64_t x[62];
  for (int k = 0; k < 62; k++) {
    uint64_t v = 0;
    for (int j = 0; j < LF_N; j++) v += c[j] * st[j];
    x[k] = v;
    lf_shift(c);
  }
  /* affine LCG power: v -> A v + B */
  uint64_t A = 1, B = 0, a = MUL, b = 1;
  for (uint64_t e = n; e; e >>= 1) {
    if (e & 1) {
      B = a * B + b;
      A = a * A;
    }
    b = a * b + b;
    a = a * a;
  }
  g->lcg = A * g->lcg + B;
  g->p = (uint8_t)(g->p + n);
  g->p24 = (uint8_t)(g->p24 + n);
  g->p55 = (uint8_t)(g->p55 + n);
  g->p61 = (uint8_t)(g->p61 + n);
  for (int k = 0; k < 62; k++) g->ring[(uint8_t)(g->p - 62 + k)] = x[k];
  g->out = g->lcg + (x[61] ^ x[0]);
}

/* _actualiza_xoshiro256pp, random_generator.h:58-67 */
uint64_t msa_xo256_next(msa_xo256 *g)
{
  uint64_t *s = g->s;
  g->out = rol(s[0] + s[3], 23) + s[0];
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rol(s[3], 45);
  return g->out;
}

/* Init_Rand_HQ_64bits, random_generator.cu:160-200 */
void msa_prng_seed_hq(msa_prng *g, uint64_t seed)
{
  memset(g, 0, sizeof(*g));
  g->lcg = seed;
  for (int i = 0; i < 111; i++) g->lcg = lcg(g->lcg);
  for (int k = 67; k < 128; k++) {
    uint64_t w = 0;
    for (int byte = 0; byte < 8; byte++) {
      g->lcg = lcg(g->lcg);
      w |= (g->lcg >> 56) << (8 * byte);
    }
    g->ring[k] = w;
  }
  prng_reset_pointers(g);
  for (int i = 0; i < 1111; i++) {           /* warm-up of the lagged sum alone */
    g->ring[g->p] = g->ring[g->p24] + g->ring[g->p55];
    g->p++;
    g->p24++;
    g->p55++;
    g->p61++;
  }
  for (int i = 0; i < 1111; i++) (void)msa_prng_next(g);
}

/* Inicia_generadores_CPU, random_generator.cu:94-158: a Luescher stream seeds
 * the PR+LCG generator, then the xoshiro256++ state and the 32-bit Philox key
 * from bit 23 of successive Luescher words. */
void msa_cpu_generators(msa_prng *pr, msa_xo256 *xo, uint32_t *key, uint64_t seed)
{
  msa_swb sw;
  msa_swb_seed(&sw, seed);
  for (int i = 0; i < 1111; i++) (void)msa_swb_next(&sw);
  memset(pr, 0, sizeof(*pr));
  pr->lcg = msa_swb_u64(&sw);
  for (int k = 67; k < 128; k++) pr->ring[k] = msa_swb_u64(&sw);
  prng_reset_pointers(pr);
  for (int i = 0; i < 2222; i++) (void)msa_prng_next(pr);

  uint64_t st[4] = {0, 0, 0, 0};
  while (!(st[0] | st[1] | st[2] | st[3]))
    for (int j = 0; j < 4; j++) {
      st[j] = 0;
      for (int bit = 0; bit < 64; bit++) st[j] |= (uint64_t)((msa_swb_next(&sw) >> 23) & 1u) << bit;
    }
  memcpy(xo->s, st, sizeof(st));
  xo->out = 0;
  uint32_t k = 0;
  for (int bit = 0; bit < 32; bit++) k |= ((msa_swb_next(&sw) >> 23) & 1u) << bit;
  *key = k;
}

/* generate_seeds_from_one, random_generator.cu:202-234.  The 3-bit selector
 * is computed from a constant there (line 215 overwrites the draw of line
 * 213), so every seed skips 109 Luescher words; the draw itself still
 * advances the stream. */
void msa_seeds_from_one(uint64_t seed, uint64_t *out, int n)
{
  msa_swb sw;
  msa_swb_seed(&sw, seed);
  for (int i = 0; i < 1109; i++) (void)msa_swb_next(&sw);
  for (int k = 0; k < n; k++) {
    (void)msa_swb_u64(&sw);
    for (int i = 0; i < 109; i++) (void)msa_swb_next(&sw);
    out[k] = msa_swb_u64(&sw) * 6364136223846793005ULL + 14426950408889634071ULL;
  }
}

/* ini_random, random_generator.cu:238-256 */
void msa_janus_seed(msa_prng *g, uint64_t seed)
{
  memset(g, 0, sizeof(*g));
  g->lcg = seed;
  for (int i = 0; i < 11; i++) g->lcg = lcg(g->lcg);
  prng_reset_pointers(g);
  for (int k = 67; k < 128; k++) {
    g->lcg = lcg(g->lcg);
    g->ring[k] = g->lcg;
  }
  for (int i = 0; i < 1111; i++) {
    g->ring[g->p] = g->ring[g->p24] + g->ring[g->p55];
    g->p++;
    g->p24++;
    g->p55++;
    g->p61++;
  }
}

/* HIRANDOM = PRRANDOM + CGRANDOM, random_generator.h:83-85 */
uint64_t msa_janus_next(msa_prng *g) { return msa_prng_next(g); }

/* _fill_bits (fill_bits.h:47-68): 52 bits of host entropy (PR + xoshiro,
 * top bits) and the iteration counter placed around the fields the kernel
 * fills: .x = entropy[13:0]<<18 | iter[14:0]<<3 | parity<<2 | internal call,
 * .y = entropy[20:14]<<25 | iter>>15, .z = entropy[33:21]<<19 (+ thread id),
 * .w = entropy[51:34]<<14 | jobid. */
void msa_fill_counter(msa_prng *pr, msa_xo256 *xo, uint64_t iter, uint32_t parity,
                      uint32_t jobid, msa_u32x4 *c)
{
  (void)msa_xo256_next(xo);
  (void)msa_prng_next(pr);
  uint64_t e = (pr->out + xo->out) >> 12;
  c->x = (uint32_t)((e & 0x3FFFu) << 18) | (uint32_t)((iter & 0x7FFFu) << 3) |
         ((parity & 1u) << 2);
  c->y = (uint32_t)(((e >> 14) & 0x7Fu) << 25) | (uint32_t)(iter >> 15);
  c->z = (uint32_t)(((e >> 21) & 0x1FFFu) << 19);
  c->w = (uint32_t)((e >> 34) << 14) | jobid;
}
