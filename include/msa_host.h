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
 * LUT pairs luts[2k], luts[2k+1]; lut_idx[k] = k.  Tables are built in
 * parallel over k for the LUTs (OpenMP), sequentially for the counters. */
int msa_sched_tables(const msa_sched *s, int64_t nref, int R, const uint64_t *seed_MC,
                     uint32_t jobid, msa_u32x4 *ctr, uint32_t *keys, msa_lut *luts,
                     int32_t *lut_idx);

/* The same tables continuing from given generator states pr[R], xo[R] (for
 * instance those a thermal phase left, msa_thermal_counters); the caller
 * holds the keys. */
int msa_sched_tables_from(const msa_sched *s, int64_t nref, int R, msa_prng *pr, msa_xo256 *xo,
                          uint32_t jobid, msa_u32x4 *ctr, msa_lut *luts, int32_t *lut_idx);

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
/* annealing schedule file: T0 T1 nsweeps refresh_every (one per line) */
int  msa_read_schedule(const char *path, msa_sched *s, char *err, int errlen);

/* Mean and error over replicas as Measures does (med.c:13-51):
 * mean of x_r, sqrt((<x^2> - <x>^2) / (n - 1)). */
void msa_mean_err(const double *x, int n, double *mean, double *err);

#ifdef __cplusplus
}
#endif
#endif
