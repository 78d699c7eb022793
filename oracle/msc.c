/*
 * msc.c -- literal transcription of the reference kernels in the reference's
 * own multispin layout (64 sites of ONE sample per word, 4x4x4 blocks,
 * bit = bx + 4 by + 16 bz), executed sequentially on the CPU.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Tables:  Init_MSC_neighbours ini.c:119-146, Init_MSC_index_and_rotate
 *          ini.c:148-258, packing_u ini.c:301-345, packing_J ini.c:521-578.
 * Kernels: d_oneMCstepBN_multisample main.cu:847-1080 (ISI branch),
 *          d_Walk main.cu:1084-1193, d_computeEne main.cu:745-815,
 *          Rotate* main.cu:706-742, Monte_Carlo main.cu:254-306.
 * Every grid thread of those kernels touches a disjoint set of words (the
 * walk swaps along a bijection), so a sequential loop over threads gives the
 * GPU result exactly.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

struct orc_msc {
  int L, R, nc, ML, MS, MV, HALF;
  int off[4], side[3];
  unsigned *neigh;                 /* MSC_neigh[MV*6] */
  unsigned *bianchi, *neri;        /* [HALF] */
  int *vicB, *vicN;                /* [HALF*6] half-index of neighbour */
  uint64_t *JB, *JN;               /* [HALF*6] */
  uint8_t *rotB, *rotN;            /* [HALF] */
  uint64_t *spin[2];               /* [HALF*R] per colour */
  uint64_t *w0[2], *w1[2];         /* [nc*HALF*R] per colour */
};

#define MASK_E 0x1111111111111111ULL
#define MASK_N 0x000F000F000F000FULL
#define MASK_U 0x000000000000FFFFULL
#define MASK_O 0x8888888888888888ULL
#define MASK_S 0xF000F000F000F000ULL
#define MASK_D 0xFFFF000000000000ULL
static inline uint64_t rotE(uint64_t v) { return (v & MASK_E) << 3 | (v & ~MASK_E) >> 1; }
static inline uint64_t rotN(uint64_t v) { return (v & MASK_N) << 12 | (v & ~MASK_N) >> 4; }
static inline uint64_t rotU(uint64_t v) { return (v & MASK_U) << 48 | (v & ~MASK_U) >> 16; }
static inline uint64_t rotO(uint64_t v) { return (v & MASK_O) >> 3 | (v & ~MASK_O) << 1; }
static inline uint64_t rotS(uint64_t v) { return (v & MASK_S) >> 12 | (v & ~MASK_S) << 4; }
static inline uint64_t rotD(uint64_t v) { return (v & MASK_D) >> 48 | (v & ~MASK_D) << 16; }

orc_msc *orc_msc_create(int L, int R, int nc)
{
  if (L % 8 != 0 || R < 1 || nc < 1 || nc > 6) return NULL;
  orc_msc *m = (orc_msc *)calloc(1, sizeof(orc_msc));
  m->L = L; m->R = R; m->nc = nc;
  m->ML = L / 4; m->MS = m->ML * m->ML; m->MV = m->MS * m->ML; m->HALF = m->MV / 2;
  const int MV = m->MV, HALF = m->HALF;
  m->neigh = (unsigned *)calloc((size_t)MV * 6, sizeof(unsigned));
  m->bianchi = (unsigned *)calloc(HALF, sizeof(unsigned));
  m->neri = (unsigned *)calloc(HALF, sizeof(unsigned));
  m->vicB = (int *)calloc((size_t)HALF * 6, sizeof(int));
  m->vicN = (int *)calloc((size_t)HALF * 6, sizeof(int));
  m->JB = (uint64_t *)calloc((size_t)HALF * 6, sizeof(uint64_t));
  m->JN = (uint64_t *)calloc((size_t)HALF * 6, sizeof(uint64_t));
  m->rotB = (uint8_t *)calloc(HALF, 1);
  m->rotN = (uint8_t *)calloc(HALF, 1);
  for (int c = 0; c < 2; c++) {
    m->spin[c] = (uint64_t *)calloc((size_t)HALF * R, sizeof(uint64_t));
    m->w0[c] = (uint64_t *)calloc((size_t)HALF * R * nc, sizeof(uint64_t));
    m->w1[c] = (uint64_t *)calloc((size_t)HALF * R * nc, sizeof(uint64_t));
  }

  /* Init_MSC_neighbours, ini.c:119-146 */
  int vol = 1;
  for (int d = 0; d < 2; d++) { m->side[d] = m->ML; m->off[d] = vol; vol *= m->side[d]; }
  m->side[2] = m->ML; m->off[2] = vol; vol *= m->side[2];
  m->off[3] = vol;
  for (int pc = 0; pc < MV; pc++) {
    int ix[3];
    for (int d = 0; d < 3; d++) ix[d] = (pc % m->off[d + 1]) / m->off[d];
    for (int d = 0; d < 3; d++) {
      int iy[3] = {ix[0], ix[1], ix[2]};
      iy[d] = (ix[d] + 1) % m->side[d];
      int pn = iy[0] + iy[1] * m->off[1] + iy[2] * m->off[2];
      m->neigh[6 * pc + d] = pn;
      m->neigh[6 * pn + d + 3] = pc;
    }
  }

  /* Init_MSC_index_and_rotate, ini.c:148-258: greedy colouring from site 0 */
  int *eoro = (int *)malloc(sizeof(int) * MV);
  for (int i = 0; i < MV; i++) eoro[i] = -1;
  int nb = 0, nn = 0;
  for (int pc = 0; pc < MV; pc++) {
    int skip = 0;
    for (int d = 0; d < 6; d++) if (eoro[m->neigh[6 * pc + d]] != -1) { skip = 1; break; }
    if (!skip) { if (nb == HALF) { free(eoro); orc_msc_destroy(m); return NULL; }
                 m->bianchi[nb++] = pc; eoro[pc] = pc; }
  }
  for (int i = 0; i < MV; i++)
    if (eoro[i] == -1) { if (nn == HALF) { free(eoro); orc_msc_destroy(m); return NULL; }
                         m->neri[nn++] = i; }
  free(eoro);
  if (nb != nn) { orc_msc_destroy(m); return NULL; }
  for (int k = 0; k < HALF; k++)
    for (int d = 0; d < 6; d++) {
      unsigned v = m->neigh[6 * m->bianchi[k] + d];
      unsigned lo = 0, hi = nn;
      while (lo < hi) { unsigned md = (lo + hi) / 2; if (m->neri[md] < v) lo = md + 1; else hi = md; }
      if (m->neri[lo] != v) { orc_msc_destroy(m); return NULL; }
      m->vicB[6 * k + d] = lo;
      v = m->neigh[6 * m->neri[k] + d];
      lo = 0; hi = nb;
      while (lo < hi) { unsigned md = (lo + hi) / 2; if (m->bianchi[md] < v) lo = md + 1; else hi = md; }
      if (m->bianchi[lo] != v) { orc_msc_destroy(m); return NULL; }
      m->vicN[6 * k + d] = lo;
    }
  for (int k = 0; k < HALF; k++) {
    uint8_t rb = 0, rn = 0;
    for (int d = 0; d < 6; d++) {
      unsigned want = (unsigned)(d < 3 ? m->off[d] : -m->off[d - 3]);
      if (m->neri[m->vicB[6 * k + d]] - m->bianchi[k] != want) rb |= (uint8_t)(1u << d);
      if (m->bianchi[m->vicN[6 * k + d]] - m->neri[k] != want) rn |= (uint8_t)(1u << d);
    }
    m->rotB[k] = rb; m->rotN[k] = rn;
  }
  return m;
}

void orc_msc_destroy(orc_msc *m)
{
  if (!m) return;
  free(m->neigh); free(m->bianchi); free(m->neri); free(m->vicB); free(m->vicN);
  free(m->JB); free(m->JN); free(m->rotB); free(m->rotN);
  for (int c = 0; c < 2; c++) { free(m->spin[c]); free(m->w0[c]); free(m->w1[c]); }
  free(m);
}

/* real site -> (colour, half index, bit), packing_u ini.c:317-337 */
static void msc_place(const orc_msc *m, int x, int y, int z, int *colour, int *half, int *bit,
                      int *msc_site)
{
  int bz = z / m->ML, g = z - bz * m->ML;
  int by = y / m->ML, be = y - by * m->ML;
  int bx = x / m->ML, a = x - bx * m->ML;
  *colour = (x ^ y ^ z) & 1;
  *bit = bx + 4 * by + 16 * bz;
  *msc_site = a + m->ML * be + g * m->MS;
  *half = *msc_site / 2;
}

/* packing_u (ini.c:301-345) for spins and the same placement for demon words
 * (walker word ic*HALF*NR + half*NR + r, main.cu:868-872, 884-887);
 * packing_J (ini.c:521-578): each bond bit is set at both endpoints. */
void orc_msc_pack(orc_msc *m, const uint8_t *spin, const uint8_t *jx, const uint8_t *jy,
                  const uint8_t *jz, const uint8_t *d0, const uint8_t *d1)
{
  const int L = m->L, R = m->R, HALF = m->HALF, nc = m->nc;
  const size_t V = (size_t)L * L * L;
  for (int c = 0; c < 2; c++) {
    memset(m->spin[c], 0, sizeof(uint64_t) * HALF * R);
    memset(m->w0[c], 0, sizeof(uint64_t) * HALF * R * nc);
    memset(m->w1[c], 0, sizeof(uint64_t) * HALF * R * nc);
  }
  uint64_t *J = (uint64_t *)calloc((size_t)m->MV * 6, sizeof(uint64_t));
  size_t s = 0;
  for (int z = 0; z < L; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++, s++) {
        int col, half, bit, ms;
        msc_place(m, x, y, z, &col, &half, &bit, &ms);
        for (int r = 0; r < R; r++) {
          m->spin[col][half * R + r] |= (uint64_t)(spin[r * V + s] & 1) << bit;
          for (int c = 0; c < nc; c++) {
            size_t w = (size_t)c * HALF * R + (size_t)half * R + r;
            m->w0[col][w] |= (uint64_t)(d0[((size_t)r * nc + c) * V + s] & 1) << bit;
            m->w1[col][w] |= (uint64_t)(d1[((size_t)r * nc + c) * V + s] & 1) << bit;
          }
        }
        const uint8_t *jj[3] = {jx, jy, jz};
        for (int d = 0; d < 3; d++) {
          int p = 6 * ms + d;
          unsigned pn = m->neigh[p];
          J[p] |= (uint64_t)(jj[d][s] & 1) << bit;
          J[6 * pn + d + 3] |= (uint64_t)(jj[d][s] & 1) << bit;
        }
      }
  for (int k = 0; k < HALF; k++)
    for (int d = 0; d < 6; d++) {
      m->JB[6 * k + d] = J[6 * m->bianchi[k] + d];
      m->JN[6 * k + d] = J[6 * m->neri[k] + d];
    }
  free(J);
}

void orc_msc_unpack(const orc_msc *m, uint8_t *spin, uint8_t *d0, uint8_t *d1)
{
  const int L = m->L, R = m->R, HALF = m->HALF, nc = m->nc;
  const size_t V = (size_t)L * L * L;
  size_t s = 0;
  for (int z = 0; z < L; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++, s++) {
        int col, half, bit, ms;
        msc_place(m, x, y, z, &col, &half, &bit, &ms);
        for (int r = 0; r < R; r++) {
          spin[r * V + s] = (uint8_t)((m->spin[col][half * R + r] >> bit) & 1);
          for (int c = 0; c < nc; c++) {
            size_t w = (size_t)c * HALF * R + (size_t)half * R + r;
            d0[((size_t)r * nc + c) * V + s] = (uint8_t)((m->w0[col][w] >> bit) & 1);
            d1[((size_t)r * nc + c) * V + s] = (uint8_t)((m->w1[col][w] >> bit) & 1);
          }
        }
      }
}

/* The words a walkers file holds (copy_from_GPU_for_write main.cu:479-487,
 * write_walkers io.c:276-288): walker0 colour 0, colour 1, then walker1
 * colour 0, colour 1, each nc*HALF*R words. */
void orc_msc_walker_words(const orc_msc *m, uint64_t *out)
{
  const size_t n = (size_t)m->nc * m->HALF * m->R;
  memcpy(out, m->w0[0], n * sizeof(uint64_t));
  memcpy(out + n, m->w0[1], n * sizeof(uint64_t));
  memcpy(out + 2 * n, m->w1[0], n * sizeof(uint64_t));
  memcpy(out + 3 * n, m->w1[1], n * sizeof(uint64_t));
}

/* d_oneMCstepBN_multisample, ISI branch (main.cu:847-1080), one colour */
static void msc_mc(orc_msc *m, int dir, int ic)
{
  const int R = m->R, HALF = m->HALF;
  uint64_t *newS = m->spin[dir], *oldS = m->spin[dir ? 0 : 1];
  const uint64_t *J = dir ? m->JN : m->JB;
  const int *nbr = dir ? m->vicN : m->vicB;
  const uint8_t *rot = dir ? m->rotN : m->rotB;
  uint64_t *W0 = m->w0[dir] + (size_t)ic * HALF * R, *W1 = m->w1[dir] + (size_t)ic * HALF * R;
  for (int ws = 0; ws < HALF; ws++)
    for (int wr = 0; wr < R; wr++) {
      uint64_t s = newS[ws * R + wr];
      const uint64_t *j = J + 6 * ws;
      uint64_t c0 = oldS[nbr[6 * ws + 0] * R + wr];
      uint64_t c1 = oldS[nbr[6 * ws + 1] * R + wr];
      uint64_t c2 = oldS[nbr[6 * ws + 2] * R + wr];
      uint64_t c3 = oldS[nbr[6 * ws + 3] * R + wr] ^ j[3];
      uint64_t c4 = oldS[nbr[6 * ws + 4] * R + wr] ^ j[4];
      uint64_t c5 = oldS[nbr[6 * ws + 5] * R + wr] ^ j[5];
      uint8_t r = rot[ws];
      if (r & 0x01) c0 = rotE(c0);
      if (r & 0x02) c1 = rotN(c1);
      if (r & 0x04) c2 = rotU(c2);
      if (r & 0x08) c3 = rotO(c3);
      if (r & 0x10) c4 = rotS(c4);
      if (r & 0x20) c5 = rotD(c5);
      uint64_t f0 = s ^ j[0] ^ c0, f1 = s ^ j[1] ^ c1, f2 = s ^ j[2] ^ c2;
      uint64_t k1 = f0 ^ f1, k2 = f0 & f1, j1 = k1 ^ f2, k3 = k1 & f2, j2 = k2 ^ k3;
      f0 = c3 ^ s; f1 = c4 ^ s; f2 = c5 ^ s;
      k1 = f0 ^ f1; k2 = f0 & f1; k3 = k1 & f2;
      uint64_t j3 = k1 ^ f2, j4 = k2 ^ k3;
      uint64_t flip;
      size_t w = (size_t)ws * R + wr;
      orc_isi_update(j1, j2, j3, j4, &W0[w], &W1[w], &flip);
      newS[w] = s ^ flip;
    }
}

/* d_oneMCstepBN_multisample with rand_h != 0 (main.cu:847-958): the
 * Metropolis branch, one colour; rand_h is offset by ibeta*(n*nrep) +
 * tidModNNRep = word*R + replica (main.cu:877-879) and read at strides of
 * the generator's thread count (here n*R). */
static void msc_mc_metropolis(orc_msc *m, int dir, const uint64_t *rnd)
{
  const int R = m->R, HALF = m->HALF;
  const size_t T = (size_t)HALF * R;
  uint64_t *newS = m->spin[dir], *oldS = m->spin[dir ? 0 : 1];
  const uint64_t *J = dir ? m->JN : m->JB;
  const int *nbr = dir ? m->vicN : m->vicB;
  const uint8_t *rot = dir ? m->rotN : m->rotB;
  for (int ws = 0; ws < HALF; ws++)
    for (int wr = 0; wr < R; wr++) {
      uint64_t s = newS[ws * R + wr];
      const uint64_t *j = J + 6 * ws;
      uint64_t c0 = oldS[nbr[6 * ws + 0] * R + wr];
      uint64_t c1 = oldS[nbr[6 * ws + 1] * R + wr];
      uint64_t c2 = oldS[nbr[6 * ws + 2] * R + wr];
      uint64_t c3 = oldS[nbr[6 * ws + 3] * R + wr] ^ j[3];
      uint64_t c4 = oldS[nbr[6 * ws + 4] * R + wr] ^ j[4];
      uint64_t c5 = oldS[nbr[6 * ws + 5] * R + wr] ^ j[5];
      uint8_t r = rot[ws];
      if (r & 0x01) c0 = rotE(c0);
      if (r & 0x02) c1 = rotN(c1);
      if (r & 0x04) c2 = rotU(c2);
      if (r & 0x08) c3 = rotO(c3);
      if (r & 0x10) c4 = rotS(c4);
      if (r & 0x20) c5 = rotD(c5);
      uint64_t f0 = s ^ j[0] ^ c0, f1 = s ^ j[1] ^ c1, f2 = s ^ j[2] ^ c2;
      uint64_t k1 = f0 ^ f1, k2 = f0 & f1, j1 = k1 ^ f2, k3 = k1 & f2, j2 = k2 ^ k3;
      f0 = c3 ^ s; f1 = c4 ^ s; f2 = c5 ^ s;
      k1 = f0 ^ f1; k2 = f0 & f1; k3 = k1 & f2;
      uint64_t j3 = k1 ^ f2, j4 = k2 ^ k3;
      const size_t w = (size_t)ws * R + wr;
      newS[w] = s ^ orc_metropolis_flip(j1, j2, j3, j4, rnd[w], rnd[T + w], rnd[2 * T + w]);
    }
}

/* Monte_Carlo flag=0, main.cu:202-252: d_gen_rndbits_cuda (orc_gen_rndbits)
 * with the parity-0 counter, the colour-0 update, then the same for parity 1. */
void orc_msc_thermal(orc_msc *m, const orc_lut *lut, const uint32_t *ctr0, const uint32_t *ctr1,
                     const uint32_t *key)
{
  uint64_t *rnd = (uint64_t *)malloc(3 * (size_t)m->HALF * m->R * sizeof(uint64_t));
  orc_gen_rndbits(m->L, m->R, lut, ctr0, key, rnd);
  msc_mc_metropolis(m, 0, rnd);
  orc_gen_rndbits(m->L, m->R, lut, ctr1, key, rnd);
  msc_mc_metropolis(m, 1, rnd);
  free(rnd);
}

/* d_Walk, main.cu:1084-1193, one colour pass over all layers */
static void msc_walk(orc_msc *m, int dir)
{
  const int R = m->R, HALF = m->HALF;
  const int *nbr = dir ? m->vicN : m->vicB;
  const uint8_t *rot = dir ? m->rotN : m->rotB;
  for (int ws = 0; ws < HALF; ws++)
    for (int wr = 0; wr < R; wr++) {
      uint8_t r = rot[ws];
      for (int ic = 0; ic < m->nc; ic++) {
        size_t base = (size_t)ic * HALF * R;
        uint64_t *W0 = m->w0[dir] + base, *W1 = m->w1[dir] + base;
        uint64_t *O0 = m->w0[dir ? 0 : 1] + base, *O1 = m->w1[dir ? 0 : 1] + base;
        size_t own = (size_t)ws * R + wr;
        size_t pn = (size_t)nbr[6 * ws + ic] * R + wr;
        uint64_t C0 = W0[own], C1 = W1[own], s0 = O0[pn], s1 = O1[pn];
        if (r & (1u << ic)) {
          switch (ic) {
          case 0: s0 = rotE(s0); s1 = rotE(s1); C0 = rotO(C0); C1 = rotO(C1); break;
          case 1: s0 = rotN(s0); s1 = rotN(s1); C0 = rotS(C0); C1 = rotS(C1); break;
          case 2: s0 = rotU(s0); s1 = rotU(s1); C0 = rotD(C0); C1 = rotD(C1); break;
          case 3: s0 = rotO(s0); s1 = rotO(s1); C0 = rotE(C0); C1 = rotE(C1); break;
          case 4: s0 = rotS(s0); s1 = rotS(s1); C0 = rotN(C0); C1 = rotN(C1); break;
          default: s0 = rotD(s0); s1 = rotD(s1); C0 = rotU(C0); C1 = rotU(C1); break;
          }
        }
        W0[own] = s0; W1[own] = s1;
        O0[pn] = C0; O1[pn] = C1;
      }
    }
}

/* Monte_Carlo flag=1, main.cu:254-306 */
void orc_msc_monte_carlo(orc_msc *m)
{
  for (int ic = 0; ic < m->nc; ic++) { msc_mc(m, 0, ic); msc_mc(m, 1, ic); }
  msc_walk(m, 0);
  msc_walk(m, 1);
}

/* d_computeEne, main.cu:745-815 (dir = 0, summed over the colour-0 words) */
void orc_msc_energy(const orc_msc *m, int32_t *e)
{
  const int R = m->R, HALF = m->HALF;
  const uint64_t *newS = m->spin[0], *oldS = m->spin[1];
  for (int wr = 0; wr < R; wr++) {
    int64_t tot = 0;
    for (int ws = 0; ws < HALF; ws++) {
      const uint64_t *j = m->JB + 6 * ws;
      uint64_t s = newS[ws * R + wr];
      uint8_t r = m->rotB[ws];
      int le = 0;
      for (int d = 0; d < 6; d++) {
        uint64_t c = oldS[m->vicB[6 * ws + d] * R + wr];
        if (d >= 3) c ^= j[d];
        if (r & (1u << d)) {
          switch (d) {
          case 0: c = rotE(c); break; case 1: c = rotN(c); break; case 2: c = rotU(c); break;
          case 3: c = rotO(c); break; case 4: c = rotS(c); break; default: c = rotD(c); break;
          }
        }
        uint64_t u = d < 3 ? (s ^ j[d] ^ c) : (c ^ s);
        le += __builtin_popcountll(u);
      }
      tot += 6 * 64 - 2 * le;
    }
    e[wr] = (int32_t)tot;
  }
}

/* calculate_scalar_energy, med.c:85-114 */
int orc_msc_scalar_energy(int L, const uint8_t *spin, const uint8_t *jx, const uint8_t *jy,
                          const uint8_t *jz)
{
  int e = 0;
  size_t s = 0;
  for (int z = 0; z < L; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++, s++) {
        size_t px = (size_t)((x + 1) % L) + (size_t)L * (y + (size_t)L * z);
        size_t py = (size_t)x + (size_t)L * (((y + 1) % L) + (size_t)L * z);
        size_t pz = (size_t)x + (size_t)L * (y + (size_t)L * ((z + 1) % L));
        e += spin[s] ^ jx[s] ^ spin[px];
        e += spin[s] ^ jy[s] ^ spin[py];
        e += spin[s] ^ jz[s] ^ spin[pz];
      }
  return 3 * L * L * L - 2 * e;
}

/* Structural facts the reference kernels rely on: colour-0 ("bianchi") MSC
 * sites are those with even block parity, and bianchi_index[k] is 2k or 2k+1,
 * so the packed word of half index k is the site the kernel updates.
 * Returns 0 when all hold. */
int orc_msc_check_tables(const orc_msc *m)
{
  for (int k = 0; k < m->HALF; k++) {
    unsigned b = m->bianchi[k], n = m->neri[k];
    if (b / 2 != (unsigned)k || n / 2 != (unsigned)k) return 1;
    int a = b % m->ML, be = (b / m->ML) % m->ML, g = b / m->MS;
    if ((a + be + g) & 1) return 2;
    a = n % m->ML; be = (n / m->ML) % m->ML; g = n / m->MS;
    if (!((a + be + g) & 1)) return 3;
  }
  return 0;
}
