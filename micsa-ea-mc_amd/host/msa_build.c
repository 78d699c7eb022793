/*
 * msa_build.c -- reference-mode construction of couplings and spins in the
 * sample-per-lane layout (64 samples per uint64 word, canonical site order).
 */
#include <stdlib.h>
#include <string.h>

#include "msa_host.h"

/* Init_J, JANUS branch (ini.c:457-471): per sample, ini_random(seeds_J[b])
 * and HIRANDOM>>63 for the x, y, z bonds of each site in lexicographic order.
 * The 64 per-lane seeds come from generate_seeds_from_one(seed_J, 64), the
 * path of io.c:794-796; the JANUS override that would give all samples one
 * seed (io.c:797-799) is not applied.  Lanes are independent, so they are
 * generated in parallel and merged per bit. */
void msa_build_couplings(int L, uint64_t seed_J, uint64_t *J)
{
  const uint64_t V = (uint64_t)L * L * L;
  uint64_t seeds[64];
  msa_seeds_from_one(seed_J, seeds, 64);
  memset(J, 0, 3 * V * sizeof(uint64_t));
  /* eight lanes per pass write disjoint bits of the same byte lane group */
#pragma omp parallel for schedule(static)
  for (int byte = 0; byte < 8; byte++) {
    msa_prng g[8];
    for (int j = 0; j < 8; j++) msa_janus_seed(&g[j], seeds[8 * byte + j]);
    for (uint64_t s = 0; s < V; s++)
      for (int d = 0; d < 3; d++) {
        uint64_t bits = 0;
        for (int j = 0; j < 8; j++) bits |= (msa_janus_next(&g[j]) >> 63) << j;
        /* each thread owns one byte of every word */
        ((uint8_t *)&J[d * V + s])[byte] = (uint8_t)bits;
      }
  }
}

/* Init_u (ini.c:275-299): one Init_Rand_HQ_64bits(seed_u) stream drawn in
 * (replica, site, sample) order, spin = top bit; the per-site sample slots
 * (MAXNSAMPLES = 2 there, header.h:20) are the 64 lanes here. */
void msa_build_spins(int L, int R, uint64_t seed_u, uint64_t *spin)
{
  const uint64_t V = (uint64_t)L * L * L;
  msa_prng g;
  msa_prng_seed_hq(&g, seed_u);
  for (int r = 0; r < R; r++)
    for (uint64_t s = 0; s < V; s++) {
      uint64_t w = 0;
      for (int b = 0; b < 64; b++) w |= (msa_prng_next(&g) >> 63) << b;
      spin[(uint64_t)r * V + s] = w;
    }
}

/* The JANUS build's Init_J with the override of io.c:797-799: every sample
 * uses one seed, so every lane of a J word is the same bit. */
void msa_build_couplings_one(int L, uint64_t seed, uint64_t *J)
{
  const uint64_t V = (uint64_t)L * L * L;
  msa_prng g;
  msa_janus_seed(&g, seed);
  for (uint64_t s = 0; s < V; s++)
    for (int d = 0; d < 3; d++) J[d * V + s] = (msa_janus_next(&g) >> 63) ? ~0ull : 0ull;
}

/* Init_u (ini.c:275-299) with `nslot` sample slots per site (MAXNSAMPLES,
 * header.h:20): one Init_Rand_HQ_64bits(seed_u) stream drawn in (replica,
 * site, slot) order; lane b < nslot is slot b, the other lanes are zero. */
void msa_build_spins_slots(int L, int R, uint64_t seed_u, int nslot, uint64_t *spin)
{
  const uint64_t V = (uint64_t)L * L * L;
  msa_prng g;
  msa_prng_seed_hq(&g, seed_u);
  for (int r = 0; r < R; r++)
    for (uint64_t s = 0; s < V; s++) {
      uint64_t w = 0;
      for (int b = 0; b < nslot; b++) {
        const uint64_t v = msa_prng_next(&g) >> 63;
        if (b < 64) w |= v << b;
      }
      spin[(uint64_t)r * V + s] = w;
    }
}
