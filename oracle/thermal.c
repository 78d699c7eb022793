/*
 * thermal.c -- the reference's canonical Metropolis thermalization
 * (Monte_Carlo with flag = 0, PROGRAM/src/main.cu:202-252): per colour, one
 * d_gen_rndbits_cuda launch (main.cu:647-702) then the update kernel with
 * rand_h != 0 (main.cu:945-958).  Random words are indexed per MSC word and
 * replica, not per sample, so every sample sees the same random bits.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include "oracle.h"

/* d_gen_rndbits_cuda, main.cu:647-702, with NBETAS = 1 (ibeta = 0): thread
 * tid < n*R, n = V/128 MSC words of one colour, replica ir = tid % R
 * (main.cu:664-665), counter ctr[ir] with .z += tid (main.cu:676), key
 * {tid, key[ir]}, six _get_rnd_word calls on the heat-bath table:
 * b0 = w0<<32 | w1, b1 = w2<<32 | w3, b2 = w4<<32 | w5 (main.cu:683-696),
 * stored at out[k*n*R + tid] (main.cu:698-700). */
void orc_gen_rndbits(int L, int R, const orc_lut *lut, const uint32_t *ctr, const uint32_t *key,
                     uint64_t *out)
{
  const uint64_t n = (uint64_t)L * L * L / 128, T = n * (uint64_t)R;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)T; t++) {
    const uint32_t tid = (uint32_t)t;
    const int ir = (int)(t % R);
    uint32_t c[4] = {ctr[4 * ir], ctr[4 * ir + 1], ctr[4 * ir + 2] + tid, ctr[4 * ir + 3]};
    uint32_t k[2] = {tid, key[ir]};
    uint64_t w[6];
    for (int j = 0; j < 6; j++) w[j] = orc_rnd_word(lut, c, k);
    out[t] = w[0] << 32 | w[1];
    out[T + t] = w[2] << 32 | w[3];
    out[2 * T + t] = w[4] << 32 | w[5];
  }
}

/* The Metropolis branch, main.cu:945-958: R = id1 + 2 id2 counts the
 * leading ones of (b0, b1, b2), so P(R >= k) = p^k with p the table's bit
 * probability; the flip algebra is taken as written there. */
uint64_t orc_metropolis_flip(uint64_t j1, uint64_t j2, uint64_t j3, uint64_t j4, uint64_t b0,
                             uint64_t b1, uint64_t b2)
{
  b1 = b0 & b1;
  b2 = b1 & b2;
  const uint64_t id2 = b1;
  const uint64_t id1 = (b0 ^ b1) | b2;
  const uint64_t j2ORj4 = j2 | j4;
  return (j1 & j3 & id1) | ((j2ORj4 | id2) & (j1 | j3 | id1)) | ((j2 & j4) | (j2ORj4 & id2));
}

static inline uint64_t site_of(int L, int x, int y, int z)
{
  return (uint64_t)x + (uint64_t)L * ((uint64_t)y + (uint64_t)L * (uint64_t)z);
}
static inline int wrap(int a, int L) { a %= L; return a < 0 ? a + L : a; }

/* One colour of the thermal update in the lane layout.  Site (x, y, z) sits
 * in the MSC word of half index (a + ML be + ML^2 g) / 2 at bit
 * bx + 4 by + 16 bz (packing_u, ini.c:317-337; ML = L/4); its three random
 * bits are that bit of the words of (half index, replica), the same for
 * every sample (rand_h is offset by site and replica only, main.cu:877-879). */
void orc_thermal_halfsweep(orc_lattice *lat, int colour, const uint64_t *rnd)
{
  const int L = lat->L, R = lat->R, ML = L / 4;
  const uint64_t V = lat->V, T = V / 128 * (uint64_t)R;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < L; z++)
    for (int y = 0; y < L; y++)
      for (int x = (colour ^ y ^ z) & 1; x < L; x += 2) {
        const int bx = x / ML, by = y / ML, bz = z / ML;
        const int ms = (x - bx * ML) + ML * ((y - by * ML) + ML * (z - bz * ML));
        const int bit = bx + 4 * by + 16 * bz;
        const uint64_t s = site_of(L, x, y, z);
        const uint64_t n[6] = {
          site_of(L, wrap(x + 1, L), y, z), site_of(L, x, wrap(y + 1, L), z),
          site_of(L, x, y, wrap(z + 1, L)), site_of(L, wrap(x - 1, L), y, z),
          site_of(L, x, wrap(y - 1, L), z), site_of(L, x, y, wrap(z - 1, L))};
        const uint64_t jb[6] = {lat->J[s], lat->J[V + s], lat->J[2 * V + s],
                                lat->J[n[3]], lat->J[V + n[4]], lat->J[2 * V + n[5]]};
        for (int r = 0; r < R; r++) {
          uint64_t *S = lat->S + (uint64_t)r * V;
          const uint64_t w = (uint64_t)(ms / 2) * R + r;
          const uint64_t b0 = 0 - ((rnd[w] >> bit) & 1);
          const uint64_t b1 = 0 - ((rnd[T + w] >> bit) & 1);
          const uint64_t b2 = 0 - ((rnd[2 * T + w] >> bit) & 1);
          const uint64_t sv = S[s];
          uint64_t f0 = sv ^ jb[0] ^ S[n[0]], f1 = sv ^ jb[1] ^ S[n[1]], f2 = sv ^ jb[2] ^ S[n[2]];
          uint64_t k1 = f0 ^ f1, k2 = f0 & f1;
          const uint64_t j1 = k1 ^ f2, j2 = k2 ^ (k1 & f2);
          f0 = S[n[3]] ^ jb[3] ^ sv; f1 = S[n[4]] ^ jb[4] ^ sv; f2 = S[n[5]] ^ jb[5] ^ sv;
          k1 = f0 ^ f1; k2 = f0 & f1;
          const uint64_t j3 = k1 ^ f2, j4 = k2 ^ (k1 & f2);
          S[s] = sv ^ orc_metropolis_flip(j1, j2, j3, j4, b0, b1, b2);
        }
      }
}

/* Monte_Carlo with flag = 0 (main.cu:202-252): bits from counter ctr0
 * (parity 0) then the colour-0 update, bits from ctr1 (parity 1) then the
 * colour-1 update.  ctr0/ctr1 are [R][4], key [R]. */
void orc_thermal_iteration(orc_lattice *lat, const orc_lut *lut, const uint32_t *ctr0,
                           const uint32_t *ctr1, const uint32_t *key)
{
  const uint64_t T = lat->V / 128 * (uint64_t)lat->R;
  uint64_t *rnd = (uint64_t *)malloc(3 * T * sizeof(uint64_t));
  orc_gen_rndbits(lat->L, lat->R, lut, ctr0, key, rnd);
  orc_thermal_halfsweep(lat, 0, rnd);
  orc_gen_rndbits(lat->L, lat->R, lut, ctr1, key, rnd);
  orc_thermal_halfsweep(lat, 1, rnd);
  free(rnd);
}
