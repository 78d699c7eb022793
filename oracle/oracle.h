/*
 * oracle.h -- CPU restatement of the reference micSA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so.  The product
 * (libmsa_hip.so, libmsa_host.so, the msa_anneal driver) never links it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference checkout root).  How each function is pinned:
 *   orc_lux_*, orc_prlcg_*, orc_xoshiro_*, orc_seed_*, orc_janus_*
 *       golden streams produced by PROGRAM/src/random_generator.cu itself,
 *       compiled by oracle/Makefile into oracle/_ref (tests/golden/rng_*.txt);
 *       xoshiro256++ also against Vigna's published reference output.
 *   orc_philox4x32_10   Random123 known-answer vectors.
 *   orc_lut_build_*     the reference's committed LUT_for_PRNG_nbits10_NB1.bin,
 *                       plus PROGRAM/input/LUT/preparation.c built into
 *                       oracle/_ref for other betas.
 *   orc_bisection       exhaustive endpoint check of all 256 CDF intervals
 *                       (the check preparation.c:87-217 performs).
 *   orc_isi_update      exhaustive closed form over all 64 (n+, n-, D) cases.
 *   orc_halfsweep/walk  the literal MSC-layout transcription (orc_msc_*) of
 *                       main.cu:847-1193 on the same configuration.
 *   orc_energy          orc_msc_energy (main.cu:745-815) and the invariants
 *                       E(ferromagnet)=3V, E+4*sum(D) conserved.
 *   orc_overlap         brute force per site and per bond.
 *   orc_fill_counter    golden counters from _fill_bits (fill_bits.h:47-68)
 *                       driven by the reference RNG in oracle/_ref.
 *   orc_refresh_layer   parity unpinned against any reference output (the
 *                       reference has no device refresh); its pieces
 *                       (Philox, counter, LUT, bisection) are pinned as above
 *                       and its law is checked statistically.
 *   orc_metropolis_flip exhaustive closed form: flip iff n_u + R >= 3 with R
 *                       the leading ones of (b0, b1, b2).
 *   orc_gen_rndbits     built from pinned pieces (counter, Philox, LUT,
 *                       bisection) with the thread indexing of
 *                       main.cu:647-702; its law is checked statistically.
 *   orc_thermal_*       the MSC-layout transcription orc_msc_thermal
 *                       (main.cu:202-252, 847-958) on the same state.
 */
#ifndef MSA_ORACLE_H
#define MSA_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- host generators (random_generator.cu / .h) ---------- */

typedef struct {            /* 24-bit subtract-with-borrow, lags 10/24 */
  uint32_t x[256];
  int borrow;
  uint8_t ip, ip1, ip2;
} orc_lux;

typedef struct {            /* Parisi-Rapuano (+24,+55, xor -61) + 64-bit LCG */
  uint64_t ira[256];
  uint64_t pr, zseed, out;
  uint8_t ip, ip1, ip2, ip3;
} orc_prlcg;

typedef struct {
  uint64_t s[4];
  uint64_t out;
} orc_xoshiro;

void     orc_lux_init(orc_lux *g, uint64_t seed);
uint32_t orc_lux_next(orc_lux *g);
uint64_t orc_lux_u64(orc_lux *g);
void     orc_prlcg_step(orc_prlcg *g);
void     orc_xoshiro_step(orc_xoshiro *g);
void     orc_prlcg_init_hq(orc_prlcg *g, uint64_t seed);
void     orc_init_cpu_generators(orc_prlcg *pr, orc_xoshiro *xo, uint32_t *key,
                                 uint64_t seed);
void     orc_seeds_from_one(uint64_t seed, uint64_t *out, int n);

typedef struct {            /* the JANUS build's global PR+LCG (HIRANDOM) */
  uint64_t ira[256];
  uint64_t zseed;
  uint8_t ip, ip1, ip2, ip3;
} orc_janus;

void     orc_janus_init(orc_janus *g, uint64_t seed);
uint64_t orc_janus_next(orc_janus *g);

/* ---------------- Philox + LUT heat-bath bits (main.cu:546-702) -------- */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

typedef struct {            /* s_lut_heat_bath, header.h:113-116 (NUMBITSPREBUSQUEDAS=10) */
  uint32_t umbrales[512];   /* 128 uint4 = 256 x (msb32, lsb32) */
  uint32_t prebusqueda[512];/* 128 uint4 = 1024 x uint16 (min | max<<8) */
} orc_lut;

void     orc_lut_build_q(const void *q1_f128, orc_lut *out);   /* q1 is a __float128 */
void     orc_lut_build_beta(double beta, orc_lut *out);
void     orc_lut_build_demon(double beta, orc_lut *lut_d0, orc_lut *lut_d1);
void     orc_thresholds(double beta, uint64_t thr[3]);
uint32_t orc_bisection(const orc_lut *lut, uint32_t rmsb, uint32_t rlsb);
uint32_t orc_rnd_word(const orc_lut *lut, uint32_t ctr[4], const uint32_t key[2]);

void orc_fill_counter(orc_prlcg *pr, orc_xoshiro *xo, uint64_t iter, uint64_t parity,
                      uint32_t jobid, uint32_t ctr_out[4]);

/* ---------------- lattice, sample-per-lane layout ---------------------- */

typedef struct {
  int L, R, nc;
  uint64_t V;
  uint64_t *J;   /* [3][V]      bond s -> s+e_d, bit b = sample b          */
  uint64_t *S;   /* [R][V]      spins                                       */
  uint64_t *D0;  /* [R][nc][V]  demon bit plane 0                           */
  uint64_t *D1;  /* [R][nc][V]  demon bit plane 1                           */
} orc_lattice;

void orc_isi_update(uint64_t j1, uint64_t j2, uint64_t j3, uint64_t j4,
                    uint64_t *d0, uint64_t *d1, uint64_t *flip);
void orc_halfsweep(orc_lattice *lat, int colour, int layer);
void orc_walk(orc_lattice *lat, int layer);
void orc_refresh_layer(orc_lattice *lat, int layer, const uint32_t *ctr /*[R][4]*/,
                       const uint32_t *key /*[R]*/, const orc_lut *lut_d0,
                       const orc_lut *lut_d1);
/* one full iteration; ctr/key/luts may be NULL when no refresh happens */
void orc_iteration(orc_lattice *lat, int do_refresh, const uint32_t *ctr,
                   const uint32_t *key, const orc_lut *lut_d0, const orc_lut *lut_d1);
void orc_iteration_omp(orc_lattice *lat, int do_refresh, const uint32_t *ctr,
                       const uint32_t *key, const orc_lut *lut_d0, const orc_lut *lut_d1,
                       int nthreads);

void orc_energy(const orc_lattice *lat, int32_t *e /*[R][64]*/);
void orc_demon_energy(const orc_lattice *lat, int32_t *ed /*[R][64]*/);
void orc_overlap(const orc_lattice *lat, int ra, int rb, int32_t *q /*[64]*/,
                 int32_t *ql /*[64]*/);

/* canonical Metropolis thermalization, Monte_Carlo flag = 0 (thermal.c) */
void     orc_gen_rndbits(int L, int R, const orc_lut *lut, const uint32_t *ctr /*[R][4]*/,
                         const uint32_t *key /*[R]*/, uint64_t *out /*[3][V/128*R]*/);
uint64_t orc_metropolis_flip(uint64_t j1, uint64_t j2, uint64_t j3, uint64_t j4, uint64_t b0,
                             uint64_t b1, uint64_t b2);
void     orc_thermal_halfsweep(orc_lattice *lat, int colour, const uint64_t *rnd);
void     orc_thermal_iteration(orc_lattice *lat, const orc_lut *lut, const uint32_t *ctr0,
                               const uint32_t *ctr1, const uint32_t *key);

/* state construction (ini.c:275-299, ini.c:457-471, io.c:794-796) */
void orc_init_spins_ref(orc_lattice *lat, uint64_t seed_u);
void orc_init_J_ref(orc_lattice *lat, uint64_t seed_J);
/* counter-based construction used by the batched/served forms */
void orc_init_philox(orc_lattice *lat, uint64_t seed);

/* ---------------- reference MSC layout transcription (one sample) ------ */

typedef struct orc_msc orc_msc;
orc_msc *orc_msc_create(int L, int R, int nc);
void     orc_msc_destroy(orc_msc *m);
/* real-space one-sample arrays: spin[R][V], Jx/Jy/Jz[V], d0/d1[R][nc][V] (0/1 bytes) */
void     orc_msc_pack(orc_msc *m, const uint8_t *spin, const uint8_t *jx, const uint8_t *jy,
                      const uint8_t *jz, const uint8_t *d0, const uint8_t *d1);
void     orc_msc_unpack(const orc_msc *m, uint8_t *spin, uint8_t *d0, uint8_t *d1);
void     orc_msc_walker_words(const orc_msc *m, uint64_t *out /*[4*nc*HALF*R]*/);
/* InitWalkers (ini.c:712-781) for one replica: nwords = MSC_V * nc words,
 * w0/w1 in iword order; advances pr and xo by 64 * nwords draws. */
void     orc_init_walkers(orc_prlcg *pr, orc_xoshiro *xo, const uint64_t thr[3], uint64_t nwords,
                          uint64_t *w0, uint64_t *w1);
void     orc_msc_monte_carlo(orc_msc *m);       /* main.cu:254-306 (micro branch) */
void     orc_msc_thermal(orc_msc *m, const orc_lut *lut, const uint32_t *ctr0,
                         const uint32_t *ctr1, const uint32_t *key);   /* main.cu:202-252 */
void     orc_msc_energy(const orc_msc *m, int32_t *e /*[R]*/);
int      orc_msc_scalar_energy(int L, const uint8_t *spin, const uint8_t *jx,
                               const uint8_t *jy, const uint8_t *jz);
int      orc_msc_check_tables(const orc_msc *m);

#ifdef __cplusplus
}
#endif
#endif
