/*
 * msa_walk.c -- host side of the reference's walker streams (InitWalkers,
 * PROGRAM/src/ini.c:712-781): their generators and the conversion between a
 * generator pair and the stream state the device runs (include/msa.h).
 */
#include <string.h>

#include "msa_host.h"

/* Init_Random, ini.c:53-58: seeds_C from generate_seeds_from_one(seed_u, ...),
 * stream (i, r) from Inicia_generadores_CPU(seeds_C[i * NR + r]). */
void msa_walker_generators(uint64_t seed_u, int n, int R, msa_prng *pr, msa_xo256 *xo)
{
  uint64_t seeds[64 * MSA_MAXR];
  msa_seeds_from_one(seed_u, seeds, n * R);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n * R; k++) {
    uint32_t key;
    msa_cpu_generators(&pr[k], &xo[k], &key, seeds[k]);
  }
}

/* The last 61 lagged-sum values sit at ring[(p - 61 + m) & 255], m = 0..60
 * (the write pointer p holds the oldest value). */
void msa_walker_stream_of(const msa_prng *pr, const msa_xo256 *xo, msa_walker_stream *st)
{
  for (int m = 0; m < 61; m++) st->ring[m] = pr->ring[(uint8_t)(pr->p - 61 + m)];
  st->lcg = pr->lcg;
  memcpy(st->xo, xo->s, sizeof st->xo);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rotr(uint64_t x, int k) { return (x >> k) | (x << (64 - k)); }

/* Back: the 256-entry ring from the recurrence run backwards,
 * x_k = x_{k+55} - x_{k+31}; the last outputs from the last draw (PR+LCG
 * final = lcg + (x_{n-1} ^ x_{n-62})) and from the xoshiro state one step
 * back (the update inverted: s1 ^ s1 << 17 is undone by a shift series). */
void msa_walker_stream_to(const msa_walker_stream *st, uint8_t p, msa_prng *pr, msa_xo256 *xo)
{
  uint64_t x[256];                       /* x[m] = x_{n-256+m} */
  for (int m = 0; m < 61; m++) x[195 + m] = st->ring[m];
  for (int m = 194; m >= 0; m--) x[m] = x[m + 55] - x[m + 31];
  memset(pr, 0, sizeof *pr);
  for (int m = 0; m < 256; m++) pr->ring[(uint8_t)(p + m)] = x[m];
  pr->p = p;
  pr->p24 = (uint8_t)(p - 24);
  pr->p55 = (uint8_t)(p - 55);
  pr->p61 = (uint8_t)(p - 61);
  pr->lcg = st->lcg;
  pr->out = st->lcg + (x[255] ^ x[194]);
  const uint64_t *s = st->xo;
  const uint64_t s3a = rotr(s[3], 45);
  const uint64_t s0 = s[0] ^ s3a;
  const uint64_t y = s[1] ^ s[2];
  const uint64_t s1 = y ^ (y << 17) ^ (y << 34) ^ (y << 51);
  const uint64_t s2 = (s[1] ^ s1) ^ s0;
  const uint64_t s3 = s3a ^ s1;
  memcpy(xo->s, s, sizeof xo->s);
  xo->out = rotl(s0 + s3, 23) + s0;
  (void)s2;
}
