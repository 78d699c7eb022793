/*
 * msa_host.h -- host C library of the micSA MI355X path (libmsa_host.so):
 * the reference's host generators and quad-precision tables, state
 * construction, annealing schedules and the reference's file formats.
 * Pure C (gcc + libquadmath), no GPU code; the driver links it with
 * libmsa_hip.so (include/msa.h).
 */
#ifndef MSA_HOST_H
#define MSA_HOST_H

#include <stdint.h>
#include <stdio.h>

#include "msa.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------- generators --
 * Same streams as PROGRAM/src/random_generator.cu (host generators of the
 * reference); state kept in structs instead of file globals. */

typedef struct {            /* Luescher 24-bit subtract-with-borrow, lags 10/24 */
  uint32_t x[256];
  int32_t borrow;
  uint8_t p, p10, p24;
} msa_swb;

typedef struct {            /* Parisi-Rapuano lagged sum (+24,+55, xor -61) + LCG */
  uint64_t ring[256];
  uint64_t lcg, out;
  uint8_t p, p24, p55, p61;
} msa_prng;

typedef struct { uint64_t s[4], out; } msa_xo256;

/* Init_luxury / luxury / Luescher_PRNG, random_generator.cu:27-91 */
void     msa_swb_seed(msa_swb *g, uint64_t seed);
uint32_t msa_swb_next(msa_swb *g);
uint64_t msa_swb_u64(msa_swb *g);
/* _actualiza_aleatorio_HQ_escalar, random_generator.h:71-76 */
uint64_t msa_prng_next(msa_prng *g);
/* n draws of msa_prng_next skipped by jump-ahead (same state as stepping) */
void     msa_prng_skip(msa_prng *g, uint64_t n);
/* _actualiza_xoshiro256pp, random_generator.h:58-67 */
uint64_t msa_xo256_next(msa_xo256 *g);
/* Init_Rand_HQ_64bits, random_generator.cu:160-200 */
void     msa_prng_seed_hq(msa_prng *g, uint64_t seed);
/* Inicia_generadores_CPU, random_generator.cu:94-158 */
void     msa_cpu_generators(msa_prng *pr, msa_xo256 *xo, uint32_t *key, uint64_t seed);
/* generate_seeds_from_one, random_generator.cu:202-234 (its skip is always 109) */
void     msa_seeds_from_one(uint64_t seed, uint64_t *out, int n);
/* ini_random + HIRANDOM, random_generator.cu:238-256, random_generator.h:83-85 */
void     msa_janus_seed(msa_prng *g, uint64_t seed);
uint64_t msa_janus_next(msa_prng *g);

/* _fill_bits, fill_bits.h:47-68, for one replica: advances (xo, pr) once. */
void msa_fill_counter(msa_prng *pr, msa_xo256 *xo, uint64_t iter, uint32_t parity,
                      uint32_t jobid, msa_u32x4 *ctr);

/* -------------------------------------------------------------- tables -- */

/* Heat-bath table for eight iid Bernoulli(q) bits per byte (preparation.c:265-388),
 * for q = e^{-4 beta} (the reference's table). */
void msa_lut_beta(double beta, msa_lut *out);
/* The pair used by the demon refresh: plane 0 with q = p/(1+p), plane 1 with
 * q = p^2/(1+p^2), p = e^{-4 beta}, so that D = D0 + 2 D1 has P(D) ~ p^D. */
void msa_lut_demon(double beta, msa_lut *plane0, msa_lut *plane1);
/* get_thresholds, ini.c:680-710 */
void msa_thresholds(double beta, uint64_t thr[3]);

/* ------------------------------------------------------ construction -- */

/* Couplings of the 64 sample lanes: lane b is ini_random(seeds_J[b]) with
 * three HIRANDOM>>63 per site (x,y,z) (Init_J JANUS branch, ini.c:457-471),
 * seeds_J = generate_seeds_from_one(seed_J, 64) (io.c:794-796).
 * J is [3][V] canonical order. */
void msa_build_couplings(int L, uint64_t seed_J, uint64_t *J);
/* Spins: Init_u (ini.c:275-299) with the 64 lanes as the sample slots:
 * one Init_Rand_HQ_64bits(seed_u) stream, replica, site, lane order. */
void msa_build_spins(int L, int R, uint64_t seed_u, uint64_t *spin);
/* The reference's default (JANUS) build: all samples share the J of one
 * seed (io.c:797-799 overrides seeds_J), lanes identical. */
void msa_build_couplings_one(int L, uint64_t seed, uint64_t *J);
/* Init_u with nslot sample slots per site (MAXNSAMPLES = 2 in the reference):
 * lane b < nslot holds slot b's sample, other lanes zero. */
void msa_build_spins_slots(int L, int R, uint64_t seed_u, int nslot, uint64_t *spin);

/* ------------------------------------------------------------ schedule -- */

typedef struct {
  double T0, T1;            /* linear anneal T(t) = T0 + (T1-T0) t / nsweeps */
  int64_t nsweeps;          /* schedule length N */
  int64_t refresh_every;    /* > 0 uniform cadence, -1 = t = 0 and powers of two */
  double beta_fixed;        /* used when T0 <= 0: every refresh at this beta */
} msa_sched;

/* Number of refreshes in iterations 0 .. niter-1 and their iterations. */
int64_t msa_sched_count(const msa_sched *s, int64_t niter);
int64_t msa_sched_time(const msa_sched *s, int64_t k);
double  msa_sched_beta(const msa_sched *s, int64_t k);

/* Build the refresh tables for refreshes 0..nref-1: counters ctr[k*R + r]
 * from per-replica generators seeded with seed_MC[r] (Init_Random,
 * ini.c:49-51) filled as _fill_bits at iteration t_k, parity 0; keys[r];
 * LUT pairs luts[2n], luts[2n+1] for the distinct inverse temperatures of
 * the refreshes in order, and lut_idx[k] = the pair of refresh k (refreshes
 * at equal beta share one pair; luts must hold room for nref pairs).
 * Returns the number of pairs written, or -1 on bad arguments.  The LUTs
 * are built in parallel (OpenMP), the counters sequentially.  Past the end
 * of the anneal (t > nsweeps) the temperature stays at T1. */
int msa_sched_tables(const msa_sched *s, int64_t nref, int R, const uint64_t *seed_MC,
                     uint32_t jobid, msa_u32x4 *ctr, uint32_t *keys, msa_lut *luts,
                     int32_t *lut_idx);

/* The same tables continuing from given generator states pr[R], xo[R] (for
 * instance those a thermal phase left, msa_thermal_counters); the caller
 * holds the keys. */
int msa_sched_tables_from(const msa_sched *s, int64_t nref, int R, msa_prng *pr, msa_xo256 *xo,
                          uint32_t jobid, msa_u32x4 *ctr, msa_lut *luts, int32_t *lut_idx);

/* The same for a run resumed at refresh k0 (k0 = msa_sched_count(s, now)):
 * pr/xo hold the generators after the draws of refreshes 0..k0-1; counters of
 * those refreshes are zero, as a run from iteration now never reads them. */
int msa_sched_tables_resume(const msa_sched *s, int64_t k0, int64_t nref, int R, msa_prng *pr,
                            msa_xo256 *xo, uint32_t jobid, msa_u32x4 *ctr, msa_lut *luts,
                            int32_t *lut_idx);

/* Advance each replica's generators by n _fill_bits draws. */
void msa_gen_skip(msa_prng *pr, msa_xo256 *xo, int R, int64_t n);

/* Counters of the canonical thermalization (Monte_Carlo with flag = 0,
 * main.cu:202-229): for t = 0 .. nterm-1 and parity p = 0, 1, _fill_bits at
 * iteration 2t + p with parity p for every replica, ctr[(t*2 + p)*R + r].
 * Advances each replica's generators 2*nterm times. */
void msa_thermal_counters(msa_prng *pr, msa_xo256 *xo, int R, int64_t nterm, uint32_t jobid,
                          msa_u32x4 *ctr);

/* Shared-table form for batches and serving: temperatures are rounded to the
 * grid T = tgrid * n, n = nlo .. nlo+nlut-1, and every lattice indexes one
 * table of LUT pairs built once (msa_lut_grid).  The refresh at iteration t
 * uses n = round(T(t) / tgrid) clamped to the grid. */
void msa_lut_grid(double tgrid, int32_t nlo, int32_t nlut, msa_lut *luts);
int  msa_sched_tables_grid(const msa_sched *s, int64_t nref, int R, const uint64_t *seed_MC,
                           uint32_t jobid, double tgrid, int32_t nlo, int32_t nlut,
                           msa_u32x4 *ctr, uint32_t *keys, int32_t *lut_idx);

/* ------------------------------------------------------ file formats -- */

#define MSA_MAXSTR 1024
#define MSA_MAXR 16

/* s_data (header.h:98-111) read from input.in (io.c:464-502): output dir,
 * maxtime, now, ntimes, nterm, nc, nr, L, flag, seed_J, seed_u, seed_MC[nr]. */
typedef struct {
  char dir[MSA_MAXSTR];
  unsigned long long maxtime, now;
  int ntimes, nterm, nc, nr, l, flag;
  uint64_t seed_J, seed_u, seed_MC[MSA_MAXR];
} msa_input;

int  msa_read_input(const char *path, msa_input *in, char *err, int errlen);
void msa_print_input(FILE *f, const msa_input *in);              /* print_data, io.c:518-532 */
int  msa_read_betas(const char *path, double *beta, char *err, int errlen);   /* io.c:534-558 */
int  msa_read_lut_file(const char *path, double beta, msa_lut *lut, char *err, int errlen);
int  msa_read_times(const char *path, int n, unsigned long long *times, char *err, int errlen);
int  msa_read_samples(const char *path, int n, int *samples, char *err, int errlen);
/* annealing schedule file: T0 T1 nsweeps refresh_every (one per line);
 * T0 = T1 = 0 keeps the inverse temperature of beta.dat */
int  msa_read_schedule(const char *path, msa_sched *s, char *err, int errlen);
/* changentimes.in (stop, io.c:930-968): "maxtime [t_1 t_2 ...]", the new end
 * of the run and, when given, the new strictly increasing list of
 * measurement times.  0: no file; 1: read (*times malloc'd, NULL when only
 * maxtime is given); 2: no number in the file (removed); -1: bad file. */
int  msa_read_changentimes(const char *path, unsigned long long *maxtime,
                           unsigned long long **times, int *ntimes, char *err, int errlen);

/* Mean and error over replicas as Measures does (med.c:13-51):
 * mean of x_r, sqrt((<x^2> - <x>^2) / (n - 1)). */
void msa_mean_err(const double *x, int n, double *mean, double *err);

/* ----------------------------------------------------- walker streams --
 * Host side of the reference's demon streams (InitWalkers, ini.c:712-781),
 * run on the device by msa_set_walker_streams (include/msa.h). */

/* Init_Random (ini.c:53-58): stream (i, r) is Inicia_generadores_CPU of
 * seeds_C[i*R + r], seeds_C = generate_seeds_from_one(seed_u, ...);
 * pr[i*R + r], xo[i*R + r] for i < n. */
void msa_walker_generators(uint64_t seed_u, int n, int R, msa_prng *pr, msa_xo256 *xo);
/* A generator pair as the device's stream state, and back: p is the ring
 * pointer of the rebuilt PR+LCG generator (a refresh of V*nc draws, V a
 * multiple of 256, leaves the pointer where it was); the rebuilt generators
 * continue the same streams and carry the last outputs as `out`. */
void msa_walker_stream_of(const msa_prng *pr, const msa_xo256 *xo, msa_walker_stream *st);
void msa_walker_stream_to(const msa_walker_stream *st, uint8_t p, msa_prng *pr, msa_xo256 *xo);

/* ------------------------------------------------- offline observables --
 * From the spins of conf files (two replicas of one sample): q_x = s^a s^b,
 * C4(r) = <q_x q_{x + r e_mu}> over sites and the three axes, and the
 * integral correlation lengths xi_{k,k+1} = I_{k+1}/I_k, I_k = int r^k C4. */

/* diff[mu*(L/2+1) + r] = sites where q_x != q_{x + r e_mu}, r = 0..L/2 (exact
 * integers; L a multiple of 8).  Returns 0, -1 on bad L. */
int  msa_overlap_counts(int L, const uint8_t *janus_a, const uint8_t *janus_b, int64_t *diff);
/* C4(r), r = 0..L/2, from counts summed over nterms (sample, pair) terms. */
void msa_c4_from_counts(int L, int64_t nterms, const int64_t *diff, double *c4);
/* Trapezoid integrals up to rc = the last r before C4 <= 0 (at most nr-1);
 * returns rc. */
int  msa_xi_integrals(int nr, const double *c4, double *xi12, double *xi23);

/* -------------------------------------------------------- checkpoints --
 * The reference's checkpoint files (io.c:17-393) with NBETAS = 1 and NR = R
 * replicas, byte for byte; L must be a multiple of 8 (its multispin layout).
 * Functions return 0, or -1 with a message in err. */

/* s_data as stored in conf files (header.h:98-111, seed_MC[16]): 184 bytes */
typedef struct {
  uint64_t maxtime, now;
  int32_t ntimes, nterm, nc, nr, l, flag;
  uint64_t seed_J, seed_u, seed_MC[MSA_MAXR];
} msa_sdata;

/* icds (header.h:69-78), the head of walkers files: 40 bytes */
typedef struct {
  uint32_t first, isample, nrep, walkers_per_site, lv, t;
  uint64_t seed;
  double beta;
} msa_icds;

#define MSA_PRNG_REF_BYTES 2080   /* sizeof(s_aleatorio_HQ_64bits), random_generator.h:29-36 */
#define MSA_XO_REF_BYTES 48       /* sizeof(s_xoshiro256pp), random_generator.h:39-44 */

/* Generator states as the reference's structs store them. */
void msa_prng_to_ref(const msa_prng *g, uint8_t *out);
void msa_prng_from_ref(const uint8_t *in, msa_prng *g);
void msa_xo256_to_ref(const msa_xo256 *g, uint8_t *out);
void msa_xo256_from_ref(const uint8_t *in, msa_xo256 *g);

uint64_t msa_conf_bytes(int L);                      /* io.c:22-25 */
uint64_t msa_walkers_bytes(int L, int R, int nc);    /* io.c:260-262 */
uint64_t msa_msc_words(int L, int R, int nc);        /* demon words of one sample */

/* write_conf (io.c:17-96) for one (sample, replica): <rep_dir>/conf.%016llu
 * (now), hard link <rep_dir>/conf, link count and size checked.  janus: the
 * V/8 spin bytes (msa_janus_from_lanes); energy = 3V - 2 * unsatisfied. */
int msa_conf_write(const char *rep_dir, uint64_t now, const msa_sdata *d, double beta,
                   const msa_prng *pr, const msa_xo256 *xo, int L, const uint8_t *janus,
                   int32_t energy, char *err, int errlen);
/* read_conf's parse of one file (io.c:110-196) after its size check. */
int msa_conf_read(const char *path, int L, msa_sdata *d, double *beta, msa_prng *pr,
                  msa_xo256 *xo, uint8_t *janus, int32_t *energy, char *err, int errlen);

/* write_walkers (io.c:254-323): <sample_dir>/walkers.%016llu (now), hard link
 * <sample_dir>/walkers; walkers.<prev> is removed when prev != 0.
 * pr[R], xo[R]: the demon-refresh generators; w: msa_msc_words words. */
int msa_walkers_write(const char *sample_dir, uint64_t now, uint64_t prev, const msa_icds *ic,
                      int L, int R, int nc, const msa_prng *pr, const msa_xo256 *xo,
                      const uint64_t *w, char *err, int errlen);
/* read_walkers' parse of one file (io.c:325-393) after its size check. */
int msa_walkers_read(const char *path, int L, int R, int nc, msa_icds *ic, msa_prng *pr,
                     msa_xo256 *xo, uint64_t *w, char *err, int errlen);

/* Spins of lanes lanes[0..n-1] in conf order (Janus_packing_for_write,
 * io.c:212-231): janus[(i*R + r)*V/8 + (s >> 3)] bit s & 7.  _to_ overwrites
 * only the listed lanes of spin[R][V]. */
void msa_janus_from_lanes(int L, int R, const uint64_t *spin, int n, const int *lanes,
                          uint8_t *janus);
void msa_janus_to_lanes(int L, int R, const uint8_t *janus, int n, const int *lanes,
                        uint64_t *spin);
/* Demons d0/d1 [R][nc][V] of the listed lanes as the walkers file's words
 * (packing_u placement, ini.c:301-345; word c*HALF*R + half*R + r of each
 * colour, main.cu:868-872): w[i * msa_msc_words + ...], plane 0 colours 0, 1
 * then plane 1 colours 0, 1.  _to_ overwrites only the listed lanes. */
void msa_msc_from_lanes(int L, int R, int nc, const uint64_t *d0, const uint64_t *d1, int n,
                        const int *lanes, uint64_t *w);
void msa_msc_to_lanes(int L, int R, int nc, const uint64_t *w, int n, const int *lanes,
                      uint64_t *d0, uint64_t *d1);

/* calculate_scalar_energy (med.c:85-113) of every lane: e[r*64 + b]. */
void msa_lane_energy(int L, int R, const uint64_t *J, const uint64_t *spin, int32_t *e);
/* countJ of Init_J (ini.c:457-511): bonds with J bit 1, cnt[64]. */
void msa_lane_countJ(int L, const uint64_t *J, int32_t *cnt);

#ifdef __cplusplus
}
#endif
#endif
