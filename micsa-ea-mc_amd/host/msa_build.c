/*
 * msa_build.c -- reference-mode construction of couplings and spins in the
 * sample-per-lane layout (64 samples per uint64 word, canonical site order).
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "msa_host.h"

/* Init_J, JANUS branch (ini.c:457-471): per sample, ini_random(seeds_J[b])
 * and HIRANDOM>>63 for the x, y, z bonds of each site in lexicographic order.
 * The 64 per-lane seeds come from generate_seeds_from_one(seed_J, 64), the
 * path of io.c:794-796; the JANUS override that would give all samples one
 * seed (io.c:797-799) is not applied.  Lanes are independent, so they are
 * generated in parallel: a task owns one byte of every word (eight lanes) over
 * one range of sites, its streams jumped ahead to the range's first draw. */
void msa_build_couplings(int L, uint64_t seed_J, uint64_t *J)
{
  const uint64_t V = (uint64_t)L * L * L;
  uint64_t seeds[64];
  msa_seeds_from_one(seed_J, seeds, 64);
  memset(J, 0, 3 * V * sizeof(uint64_t));
  const int nseg = V >= 4096 ? 8 : 1;
#pragma omp parallel for schedule(dynamic)
  for (int task = 0; task < 8 * nseg; task++) {
    /* tasks running together take different site ranges (no shared lines) */
    const int seg = task % nseg, byte = task / nseg;
    const uint64_t s0 = V * seg / nseg, s1 = V * (seg + 1) / nseg;
    msa_prng g[8];
    for (int j = 0; j < 8; j++) {
      msa_janus_seed(&g[j], seeds[8 * byte + j]);
      msa_prng_skip(&g[j], 3 * s0);
    }
    for (uint64_t s = s0; s < s1; s++)
      for (int d = 0; d < 3; d++) {
        uint64_t bits = 0;
        for (int j = 0; j < 8; j++) bits |= (msa_janus_next(&g[j]) >> 63) << j;
        /* each thread owns one byte of every word */
        ((uint8_t *)&J[d * V + s])[byte] = (uint8_t)bits;
      }
  }
}

/* Words [w0, w1) of a stream drawn `per` times per word (lane b < 64 the
 * top bit of draw b), for words in (replica, site) order; in parallel, each
 * chunk's generator jumped ahead to its first draw. */
static void spins_from_stream(const msa_prng *g0, uint64_t nw, int per, uint64_t *spin)
{
  int nchunk = 1;
#ifdef _OPENMP
  nchunk = 4 * omp_get_max_threads();
#endif
  if ((uint64_t)nchunk > nw / 256) nchunk = nw / 256 > 0 ? (int)(nw / 256) : 1;
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < nchunk; c++) {
    const uint64_t w0 = nw * c / nchunk, w1 = nw * (c + 1) / nchunk;
    msa_prng g = *g0;
    msa_prng_skip(&g, w0 * (uint64_t)per);
    for (uint64_t w = w0; w < w1; w++) {
      uint64_t v = 0;
      for (int b = 0; b < per; b++) {
        const uint64_t bit = msa_prng_next(&g) >> 63;
        if (b < 64) v |= bit << b;
      }
      spin[w] = v;
    }
  }
}

/* Init_u (ini.c:275-299): one Init_Rand_HQ_64bits(seed_u) stream drawn in
 * (replica, site, sample) order, spin = top bit; the per-site sample slots
 * (MAXNSAMPLES = 2 there, header.h:20) are the 64 lanes here. */
void msa_build_spins(int L, int R, uint64_t seed_u, uint64_t *spin)
{
  msa_prng g;
  msa_prng_seed_hq(&g, seed_u);
  spins_from_stream(&g, (uint64_t)R * L * L * L, 64, spin);
}

/* The JANUS build's Init_J with the override of io.c:797-799: every sample
 * uses one seed, so every lane of a J word is the same bit. */
void msa_build_couplings_one(int L, uint64_t seed, uint64_t *J)
{
  const uint64_t V = (uint64_t)L * L * L;
  msa_prng g0;
  msa_janus_seed(&g0, seed);
  const int nseg = V >= 4096 ? 16 : 1;
#pragma omp parallel for schedule(dynamic)
  for (int seg = 0; seg < nseg; seg++) {
    const uint64_t s0 = V * seg / nseg, s1 = V * (seg + 1) / nseg;
    msa_prng g = g0;
    msa_prng_skip(&g, 3 * s0);
    for (uint64_t s = s0; s < s1; s++)
      for (int d = 0; d < 3; d++) J[d * V + s] = (msa_janus_next(&g) >> 63) ? ~0ull : 0ull;
  }
}

/* Init_u (ini.c:275-299) with `nslot` sample slots per site (MAXNSAMPLES,
 * header.h:20): one Init_Rand_HQ_64bits(seed_u) stream drawn in (replica,
 * site, slot) order; lane b < nslot is slot b, the other lanes are zero. */
void msa_build_spins_slots(int L, int R, uint64_t seed_u, int nslot, uint64_t *spin)
{
  msa_prng g;
  msa_prng_seed_hq(&g, seed_u);
  spins_from_stream(&g, (uint64_t)R * L * L * L, nslot, spin);
}
