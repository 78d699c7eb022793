/*
 * lattice.c -- oracle restatement of the micSA update in the sample-per-lane
 * layout: one uint64 word per (site, replica, field), bit b = sample b.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The per-bit rule is the reference kernel's (PROGRAM/src/main.cu:847-1080,
 * ISI branch) and the demon walk is d_Walk (main.cu:1084-1193), both written
 * over real-space periodic neighbours instead of the reference's 4x4x4
 * in-word blocks.  orc_msc_* (msc.c) is the literal MSC-layout transcription
 * the tests compare this against.
 *
 * Site s = x + L*(y + L*z); colour (x+y+z)&1, colour 0 = the reference's
 * "bianchi" set (ini.c:160-176 starts it at site 0).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static inline uint64_t site_of(int L, int x, int y, int z)
{
  return (uint64_t)x + (uint64_t)L * ((uint64_t)y + (uint64_t)L * (uint64_t)z);
}
static inline int wrap(int a, int L) { a %= L; return a < 0 ? a + L : a; }

/* ISI boolean algebra, main.cu:1037-1074.  (j1,j2) and (j3,j4) are 2-bit
 * sums of unsatisfied bonds in the + and - directions. */
void orc_isi_update(uint64_t j1, uint64_t j2, uint64_t j3, uint64_t j4,
                    uint64_t *d0, uint64_t *d1, uint64_t *flip)
{
  uint64_t D0 = *d0, D1 = *d1;
  uint64_t x13 = j1 ^ j3, x24 = j2 ^ j4;
  uint64_t sub = (x24 & (~j1 & ~j3)) | (~j2 & ~j4);
  uint64_t s0 = ~x13;
  uint64_t s1 = ~x24 & (x13 | (s0 & ~(j1 ^ j2)));
  s1 = (~sub & s1) | (sub & (s1 ^ s0));
  uint64_t n0 = D0 ^ s0;
  uint64_t k1 = D0 & s0;
  uint64_t k2 = D1 ^ s1;
  uint64_t n1 = k1 ^ k2;
  uint64_t k3 = D1 & s1;
  uint64_t carry = (k3 | (k2 & k1)) ^ sub;
  uint64_t f = ~carry;
  *d0 = (f & n0) | (~f & D0);
  *d1 = (f & n1) | (~f & D1);
  *flip = f;
}

/* local field, main.cu:898-944: f_d = s ^ J_d ^ s_nb, then bit-sliced
 * 3-input sums over (+x,+y,+z) -> (j1,j2) and (-x,-y,-z) -> (j3,j4). */
static inline void local_field(uint64_t s, const uint64_t nb[6], const uint64_t jb[6],
                               uint64_t *j1, uint64_t *j2, uint64_t *j3, uint64_t *j4)
{
  uint64_t f0 = s ^ jb[0] ^ nb[0], f1 = s ^ jb[1] ^ nb[1], f2 = s ^ jb[2] ^ nb[2];
  uint64_t k1 = f0 ^ f1, k2 = f0 & f1;
  *j1 = k1 ^ f2;
  *j2 = k2 ^ (k1 & f2);
  f0 = nb[3] ^ jb[3] ^ s; f1 = nb[4] ^ jb[4] ^ s; f2 = nb[5] ^ jb[5] ^ s;
  k1 = f0 ^ f1; k2 = f0 & f1;
  *j3 = k1 ^ f2;
  *j4 = k2 ^ (k1 & f2);
}

static void halfsweep_plane(orc_lattice *lat, int colour, int layer, int z)
{
  const int L = lat->L;
  const uint64_t V = lat->V;
  for (int y = 0; y < L; y++)
    for (int x = (colour ^ y ^ z) & 1; x < L; x += 2) {
      uint64_t s = site_of(L, x, y, z);
      uint64_t n[6] = {
        site_of(L, wrap(x + 1, L), y, z), site_of(L, x, wrap(y + 1, L), z),
        site_of(L, x, y, wrap(z + 1, L)), site_of(L, wrap(x - 1, L), y, z),
        site_of(L, x, wrap(y - 1, L), z), site_of(L, x, y, wrap(z - 1, L))};
      uint64_t jb[6] = {lat->J[s], lat->J[V + s], lat->J[2 * V + s],
                        lat->J[n[3]], lat->J[V + n[4]], lat->J[2 * V + n[5]]};
      for (int r = 0; r < lat->R; r++) {
        const uint64_t *S = lat->S + (uint64_t)r * V;
        uint64_t nb[6];
        for (int d = 0; d < 6; d++) nb[d] = S[n[d]];
        uint64_t j1, j2, j3, j4, flip;
        local_field(S[s], nb, jb, &j1, &j2, &j3, &j4);
        uint64_t off = ((uint64_t)r * lat->nc + layer) * V + s;
        orc_isi_update(j1, j2, j3, j4, &lat->D0[off], &lat->D1[off], &flip);
        lat->S[(uint64_t)r * V + s] = S[s] ^ flip;
      }
    }
}

/* one colour of a sweep with demon layer `layer`: main.cu:255-284 */
void orc_halfsweep(orc_lattice *lat, int colour, int layer)
{
  for (int z = 0; z < lat->L; z++) halfsweep_plane(lat, colour, layer, z);
}

static const int DIRV[6][3] = {{1,0,0},{0,1,0},{0,0,1},{-1,0,0},{0,-1,0},{0,0,-1}};

/* d_Walk, main.cu:1084-1193, for one layer ic: the colour-0 pass swaps the
 * demon at every colour-0 site x with the one at x+e_ic, the colour-1 pass
 * then swaps every colour-1 site y with y+e_ic.  Net: colour-0 demons move by
 * +2e_ic, colour-1 demons by -2e_ic (e_3..e_5 = -x,-y,-z). */
void orc_walk(orc_lattice *lat, int ic)
{
  const int L = lat->L;
  const uint64_t V = lat->V;
  uint64_t *tmp = (uint64_t *)malloc(V * sizeof(uint64_t));
  for (int r = 0; r < lat->R; r++)
    for (int plane = 0; plane < 2; plane++) {
      uint64_t *D = (plane ? lat->D1 : lat->D0) + ((uint64_t)r * lat->nc + ic) * V;
      memcpy(tmp, D, V * sizeof(uint64_t));
      for (int z = 0; z < L; z++)
        for (int y = 0; y < L; y++)
          for (int x = 0; x < L; x++) {
            int sgn = ((x + y + z) & 1) ? -1 : 1;   /* colour 1 moves backwards */
            int xs = wrap(x - 2 * sgn * DIRV[ic][0], L);
            int ys = wrap(y - 2 * sgn * DIRV[ic][1], L);
            int zs = wrap(z - 2 * sgn * DIRV[ic][2], L);
            D[site_of(L, x, y, z)] = tmp[site_of(L, xs, ys, zs)];
          }
    }
  free(tmp);
}

/* Demon refresh of one layer from Philox + LUT bits.  Word (layer c, site s)
 * of replica r uses thread id tid = c*V + s the way d_gen_rndbits_cuda
 * (main.cu:647-702) uses its thread index: counter ctr[r] with .z += tid,
 * key {tid, key[r]}, four calls with .x incremented after each; calls 0,1
 * (LUT of plane 0) give D0 = w0<<32 | w1, calls 2,3 (LUT of plane 1) give
 * D1 = w2<<32 | w3. */
void orc_refresh_layer(orc_lattice *lat, int layer, const uint32_t *ctr, const uint32_t *key,
                       const orc_lut *lut_d0, const orc_lut *lut_d1)
{
  const uint64_t V = lat->V;
  for (int r = 0; r < lat->R; r++)
    for (uint64_t s = 0; s < V; s++) {
      uint32_t tid = (uint32_t)((uint64_t)layer * V + s);
      uint32_t c[4] = {ctr[4 * r], ctr[4 * r + 1], ctr[4 * r + 2] + tid, ctr[4 * r + 3]};
      uint32_t k[2] = {tid, key[r]};
      uint64_t w0 = orc_rnd_word(lut_d0, c, k);
      uint64_t w1 = orc_rnd_word(lut_d0, c, k);
      uint64_t w2 = orc_rnd_word(lut_d1, c, k);
      uint64_t w3 = orc_rnd_word(lut_d1, c, k);
      uint64_t off = ((uint64_t)r * lat->nc + layer) * V + s;
      lat->D0[off] = (w0 << 32) | w1;
      lat->D1[off] = (w2 << 32) | w3;
    }
}

/* One micSA iteration (main.cu:137-150 with Monte_Carlo flag=1, main.cu:254-306):
 * optional refresh of every layer, nc sweeps of two colours, then the walk. */
void orc_iteration(orc_lattice *lat, int do_refresh, const uint32_t *ctr, const uint32_t *key,
                   const orc_lut *lut_d0, const orc_lut *lut_d1)
{
  if (do_refresh)
    for (int c = 0; c < lat->nc; c++) orc_refresh_layer(lat, c, ctr, key, lut_d0, lut_d1);
  for (int ic = 0; ic < lat->nc; ic++) {
    orc_halfsweep(lat, 0, ic);
    orc_halfsweep(lat, 1, ic);
  }
  for (int ic = 0; ic < lat->nc; ic++) orc_walk(lat, ic);
}

/* Same iteration, planes of one colour spread over threads (the CPU baseline).
 * Within a colour every site is independent, so the result is identical. */
void orc_iteration_omp(orc_lattice *lat, int do_refresh, const uint32_t *ctr, const uint32_t *key,
                       const orc_lut *lut_d0, const orc_lut *lut_d1, int nthreads)
{
  const uint64_t V = lat->V;
  if (do_refresh) {
    for (int c = 0; c < lat->nc; c++)
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int64_t s = 0; s < (int64_t)V; s++)
        for (int r = 0; r < lat->R; r++) {
          uint32_t tid = (uint32_t)((uint64_t)c * V + (uint64_t)s);
          uint32_t cc[4] = {ctr[4 * r], ctr[4 * r + 1], ctr[4 * r + 2] + tid, ctr[4 * r + 3]};
          uint32_t k[2] = {tid, key[r]};
          uint64_t w0 = orc_rnd_word(lut_d0, cc, k);
          uint64_t w1 = orc_rnd_word(lut_d0, cc, k);
          uint64_t w2 = orc_rnd_word(lut_d1, cc, k);
          uint64_t w3 = orc_rnd_word(lut_d1, cc, k);
          uint64_t off = ((uint64_t)r * lat->nc + c) * V + (uint64_t)s;
          lat->D0[off] = (w0 << 32) | w1;
          lat->D1[off] = (w2 << 32) | w3;
        }
  }
  for (int ic = 0; ic < lat->nc; ic++)
    for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int z = 0; z < lat->L; z++) halfsweep_plane(lat, colour, ic, z);
    }
  for (int ic = 0; ic < lat->nc; ic++) orc_walk(lat, ic);
}

static inline void add_lanes(int32_t *acc, uint64_t w, int32_t weight)
{
  while (w) {
    int b = __builtin_ctzll(w);
    acc[b] += weight;
    w &= w - 1;
  }
}

/* Integer energy per (replica, sample): 3V - 2 * #unsatisfied bonds, the
 * quantity d_computeEne sums (main.cu:745-815) and calculate_scalar_energy
 * recomputes (med.c:85-114). */
void orc_energy(const orc_lattice *lat, int32_t *e)
{
  const int L = lat->L;
  const uint64_t V = lat->V;
  for (int r = 0; r < lat->R; r++) {
    int32_t cnt[64] = {0};
    const uint64_t *S = lat->S + (uint64_t)r * V;
    for (int z = 0; z < L; z++)
      for (int y = 0; y < L; y++)
        for (int x = 0; x < L; x++) {
          uint64_t s = site_of(L, x, y, z);
          add_lanes(cnt, S[s] ^ lat->J[s] ^ S[site_of(L, wrap(x + 1, L), y, z)], 1);
          add_lanes(cnt, S[s] ^ lat->J[V + s] ^ S[site_of(L, x, wrap(y + 1, L), z)], 1);
          add_lanes(cnt, S[s] ^ lat->J[2 * V + s] ^ S[site_of(L, x, y, wrap(z + 1, L))], 1);
        }
    for (int b = 0; b < 64; b++) e[r * 64 + b] = (int32_t)(3 * V) - 2 * cnt[b];
  }
}

/* Demon energy per (replica, sample): sum over layers and sites of D0 + 2 D1,
 * Measure_Energy_Clerics, med.c:55-83. */
void orc_demon_energy(const orc_lattice *lat, int32_t *ed)
{
  const uint64_t V = lat->V;
  for (int r = 0; r < lat->R; r++) {
    int32_t cnt[64] = {0};
    for (int c = 0; c < lat->nc; c++)
      for (uint64_t s = 0; s < V; s++) {
        uint64_t off = ((uint64_t)r * lat->nc + c) * V + s;
        add_lanes(cnt, lat->D0[off], 1);
        add_lanes(cnt, lat->D1[off], 2);
      }
    memcpy(ed + r * 64, cnt, sizeof(cnt));
  }
}

/* Overlaps between replicas ra and rb per sample: q counts sites with
 * differing spins, ql counts bonds whose product s_x s_y differs (J cancels).
 * q = 1 - 2 q/V, q_link = 1 - 2 ql/(3V).  No reference counterpart exists
 * (the reference measures energies only, med.c:3-83). */
void orc_overlap(const orc_lattice *lat, int ra, int rb, int32_t *q, int32_t *ql)
{
  const int L = lat->L;
  const uint64_t V = lat->V;
  const uint64_t *A = lat->S + (uint64_t)ra * V, *B = lat->S + (uint64_t)rb * V;
  int32_t cq[64] = {0}, cl[64] = {0};
  for (int z = 0; z < L; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++) {
        uint64_t s = site_of(L, x, y, z);
        uint64_t nb[3] = {site_of(L, wrap(x + 1, L), y, z), site_of(L, x, wrap(y + 1, L), z),
                          site_of(L, x, y, wrap(z + 1, L))};
        add_lanes(cq, A[s] ^ B[s], 1);
        for (int d = 0; d < 3; d++) add_lanes(cl, (A[s] ^ A[nb[d]]) ^ (B[s] ^ B[nb[d]]), 1);
      }
  memcpy(q, cq, sizeof(cq));
  memcpy(ql, cl, sizeof(cl));
}

/* Init_u, ini.c:275-299, with the per-site sample slots widened from
 * MAXNSAMPLES=2 to the 64 lanes, all lanes selected: one PR+LCG stream
 * (Init_Rand_HQ_64bits(seed_u)), order replica, site, lane; bit = top bit. */
void orc_init_spins_ref(orc_lattice *lat, uint64_t seed_u)
{
  orc_prlcg g;
  orc_prlcg_init_hq(&g, seed_u);
  for (int r = 0; r < lat->R; r++)
    for (uint64_t s = 0; s < lat->V; s++) {
      uint64_t w = 0;
      for (int b = 0; b < 64; b++) {
        orc_prlcg_step(&g);
        w |= (g.out >> 63) << b;
      }
      lat->S[(uint64_t)r * lat->V + s] = w;
    }
}

/* Init_J JANUS branch, ini.c:457-471: ini_random(seeds_J[b]) then three
 * HIRANDOM>>63 per site (x, y, z) in lexicographic site order.  Lane b takes
 * seeds_J[b] of generate_seeds_from_one(seed_J, 64) (io.c:794-796); the JANUS
 * override that gives every sample the same seed (io.c:797-799) is not applied,
 * so the 64 lanes are independent samples. */
void orc_init_J_ref(orc_lattice *lat, uint64_t seed_J)
{
  uint64_t seeds[64];
  orc_seeds_from_one(seed_J, seeds, 64);
  const uint64_t V = lat->V;
  memset(lat->J, 0, 3 * V * sizeof(uint64_t));
  for (int b = 0; b < 64; b++) {
    orc_janus g;
    orc_janus_init(&g, seeds[b]);
    for (uint64_t s = 0; s < V; s++)
      for (int d = 0; d < 3; d++) lat->J[d * V + s] |= (orc_janus_next(&g) >> 63) << b;
  }
}

/* Counter-based construction for the batched and served forms: J word (d,s)
 * = Philox({s, d, 1, 0}, {seed_lo, seed_hi}), spin word (r,s) =
 * Philox({s, r, 2, 0}, key); word = out[0]<<32 | out[1]. */
void orc_init_philox(orc_lattice *lat, uint64_t seed)
{
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (uint64_t s = 0; s < lat->V; s++) {
    for (int d = 0; d < 3; d++) {
      uint32_t c[4] = {(uint32_t)s, (uint32_t)d, 1u, 0u}, o[4];
      orc_philox4x32_10(c, key, o);
      lat->J[(uint64_t)d * lat->V + s] = ((uint64_t)o[0] << 32) | o[1];
    }
    for (int r = 0; r < lat->R; r++) {
      uint32_t c[4] = {(uint32_t)s, (uint32_t)r, 2u, 0u}, o[4];
      orc_philox4x32_10(c, key, o);
      lat->S[(uint64_t)r * lat->V + s] = ((uint64_t)o[0] << 32) | o[1];
    }
  }
}
