/*
 * msa_anneal.c -- the annealing driver: the reference's main() (PROGRAM/src/main.cu:10-195)
 * with its command line, input files and evol_*.txt output, running the
 * sweeps through libmsa_hip.so (include/msa.h).
 *
 * usage: msa_anneal isample nbits beta.dat input.in LUT list_times device
 *                   [max_time [list_samples [schedule.in]]]
 *
 *   isample      job id, < 2^14 (main.cu:70-73); enters every Philox counter
 *   nbits        number of sample lanes written to BIT%03d/ (1..64); all 64
 *                lanes are always simulated
 *   beta.dat     one inverse temperature (io.c:534-558)
 *   input.in     the reference's input file (io.c:464-502); nr = replicas (1..16)
 *   LUT          the reference's heat-bath table file, checked against beta.dat
 *                (io.c:561-591)
 *   list_times   measurement iterations (io.c:626-655)
 *   device       GPU index
 *   max_time     wall-clock budget in seconds, 0 = none (timing.c:35-58)
 *   list_samples lanes to write (io.c:593-624)
 *   schedule.in  linear anneal "T0 T1 nsweeps refresh_every"; without it the
 *                demons are refreshed at fixed beta (beta.dat) at iterations
 *                0, 1, 2, 4, 8, ... as the reference does (main.cu:140-142)
 *
 * Differences from the reference, by design: demons are refreshed on the GPU
 * from Philox4x32-10 + LUT bits (no host threads); restart from checkpoints
 * (flag = 0) is not part of this path and is rejected with an error.  The
 * canonical Metropolis thermalization (nterm > 0, main.cu:124-128) runs on
 * the GPU with the LUT file's heat-bath table, its counters drawn from the
 * per-replica generators before those of the demon refreshes, as the
 * reference's _fill_bits calls consume the same generators.
 */
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "msa.h"
#include "msa_host.h"

static void die(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fputc('\n', stderr);
  exit(1);
}

#define CHK(call)                                                               \
  do {                                                                          \
    int rc_ = (call);                                                           \
    if (rc_ != MSA_OK) die("%s: %s", #call, msa_strerror(rc_));                 \
  } while (0)

static uint64_t urandom64(void)
{
  uint64_t v = 0;
  FILE *f = fopen("/dev/urandom", "r");
  if (!f || fread(&v, sizeof v, 1, f) != 1) die("cannot read /dev/urandom");
  fclose(f);
  return v;
}

/* comprueba_semilla, random_generator.cu:11-16 */
static uint64_t nonzero_seed(uint64_t s) { return s ? s : urandom64(); }

static double now_s(clockid_t id)
{
  struct timespec ts;
  clock_gettime(id, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void print_help(const char *prog)
{
  printf("Usage: %s isample nbits beta.dat input.in LUT list_times device "
         "[max_time [list_samples [schedule.in]]]\n", prog);
  exit(0);
}

int main(int argc, char **argv)
{
  if (argc > 1 && (!strcmp(argv[1], "-h") || !strcmp(argv[1], "--help"))) print_help(argv[0]);
  if (argc < 8 || argc > 11)
    die("Usage: %s isample nbits beta.dat input.in LUT list_times device "
        "[max_time [list_samples [schedule.in]]]", argv[0]);
  const int sample = atoi(argv[1]);
  const int nbits = atoi(argv[2]);
  const char *f_beta = argv[3], *f_input = argv[4], *f_lut = argv[5], *f_times = argv[6];
  const int device = atoi(argv[7]);
  const unsigned long long max_time = argc > 8 ? strtoull(argv[8], NULL, 10) : 0;
  const char *f_samples = argc > 9 ? argv[9] : "";
  const char *f_sched = argc > 10 ? argv[10] : NULL;
  char e[512];

  if (nbits < 1 || nbits > 64) die("Error: nbits out of range, 1 <= nbits <= 64");
  if (sample < 0) die("Error: sample out of range, 0 <= sample");
  if (sample > (1 << 14) - 1) die("Our Counter-based PRNG  cannot accomodate more than %d samples", 1 << 14);

  msa_input in;
  if (msa_read_input(f_input, &in, e, sizeof e)) die("%s", e);
  double beta;
  if (msa_read_betas(f_beta, &beta, e, sizeof e)) die("%s", e);
  msa_lut lut_file;
  if (msa_read_lut_file(f_lut, beta, &lut_file, e, sizeof e)) die("%s", e);
  unsigned long long *times = (unsigned long long *)calloc(in.ntimes > 0 ? in.ntimes : 1, sizeof *times);
  if (in.ntimes > 0 && msa_read_times(f_times, in.ntimes, times, e, sizeof e)) die("%s", e);
  int samples[64];
  if (msa_read_samples(f_samples, nbits, samples, e, sizeof e)) die("%s", e);
  for (int i = 0; i < nbits; i++)
    if (samples[i] < 0 || samples[i] > 63) die("sample lane %d outside 0..63", samples[i]);
  if (in.nterm < 0) die("nterm=%d must not be negative", in.nterm);
  if (in.nterm > 0 && ((in.l & 0xffff) % 8)) die("nterm=%d: thermalization needs L to be a multiple of 8", in.nterm);
  if (in.flag == 0 || in.flag < -2 || in.flag > 2) die("flag=%d: only new runs (flag 1 or 2) are supported", in.flag);
  if (in.nc < 1 || in.nc > 6) die("nc=%d outside 1..6", in.nc);
  const int L = in.l & 0xffff, R = in.nr;

  printf("Path: %s\n", in.dir);
  msa_print_input(stdout, &in);
  fflush(stdout);

  msa_sched sch;
  memset(&sch, 0, sizeof sch);
  if (f_sched) {
    if (msa_read_schedule(f_sched, &sch, e, sizeof e)) die("%s", e);
  } else {
    sch.refresh_every = -1;
    sch.beta_fixed = beta;
  }

  /* output tree: dir/I%03d/BIT%03d/ (create_sample_path, io.c:657-677) */
  char dir[MSA_MAXSTR + 16], bitdir[64][MSA_MAXSTR + 32];
  snprintf(dir, sizeof dir, "%sI%03d/", in.dir, sample);
  mkdir(dir, 0775);
  for (int i = 0; i < nbits; i++) {
    snprintf(bitdir[i], sizeof bitdir[i], "%sBIT%03d/", dir, samples[i]);
    mkdir(bitdir[i], 0775);
  }
  printf("Dirsample = %s\n", dir);

  const uint64_t seed_J = nonzero_seed(in.seed_J), seed_u = nonzero_seed(in.seed_u);
  uint64_t seed_MC[MSA_MAXR];
  for (int r = 0; r < R; r++) seed_MC[r] = nonzero_seed(in.seed_MC[r]);

  const uint64_t V = (uint64_t)L * L * L;
  uint64_t *J = (uint64_t *)malloc(3 * V * sizeof(uint64_t));
  uint64_t *S = (uint64_t *)malloc((uint64_t)R * V * sizeof(uint64_t));
  if (!J || !S) die("Problems allocating host lattice");
  printf("Initializating J's and Spins ...\n");
  fflush(stdout);
  msa_build_couplings(L, seed_J, J);
  msa_build_spins(L, R, seed_u, S);

  msa_geom g = {L, R, in.nc};
  msa_ctx *ctx = NULL;
  CHK(msa_create(device, &g, &ctx));
  CHK(msa_upload_state(ctx, J, S, NULL, NULL));
  free(J);
  free(S);

  const int64_t niter = (int64_t)in.maxtime;
  const int64_t nref = msa_sched_count(&sch, niter > 0 ? niter : 1);
  msa_u32x4 *ctr = (msa_u32x4 *)malloc(sizeof(msa_u32x4) * (size_t)(nref * R));
  msa_lut *luts = (msa_lut *)malloc(sizeof(msa_lut) * 2 * (size_t)nref);
  int32_t *lut_idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)nref);
  uint32_t keys[MSA_MAXR];
  if (!ctr || !luts || !lut_idx) die("Problems allocating refresh tables (%lld)", (long long)nref);
  msa_prng pr[MSA_MAXR];
  msa_xo256 xo[MSA_MAXR];
  for (int r = 0; r < R; r++) msa_cpu_generators(&pr[r], &xo[r], &keys[r], seed_MC[r]);
  if (in.nterm > 0) {
    /* THERMALIZATION, main.cu:124-128: Monte_Carlo(flag = 0) nterm times */
    printf("Thermalization: %d Metropolis iterations at beta = %.6f\n", in.nterm, beta);
    fflush(stdout);
    msa_u32x4 *tctr = (msa_u32x4 *)malloc(sizeof(msa_u32x4) * 2 * (size_t)in.nterm * R);
    if (!tctr) die("Problems allocating thermal counters");
    msa_thermal_counters(pr, xo, R, in.nterm, (uint32_t)sample, tctr);
    CHK(msa_run_thermal(ctx, in.nterm, &lut_file, tctr, keys));
    free(tctr);
  }
  msa_sched_tables_from(&sch, nref, R, pr, xo, (uint32_t)sample, ctr, luts, lut_idx);
  CHK(msa_set_luts(ctx, (int32_t)nref, luts));
  CHK(msa_set_refresh(ctx, (int32_t)nref, ctr, keys, lut_idx));
  free(luts);
  free(ctr);
  free(lut_idx);

  int32_t *unsat = (int32_t *)calloc((size_t)R * 64, sizeof(int32_t));
  int32_t *demon = (int32_t *)calloc((size_t)R * 64, sizeof(int32_t));
  int32_t *q = (int32_t *)calloc((size_t)(R / 2 + 1) * 64, sizeof(int32_t));
  int32_t *ql = (int32_t *)calloc((size_t)(R / 2 + 1) * 64, sizeof(int32_t));
  double xs[MSA_MAXR], xc[MSA_MAXR];

  double wall0 = now_s(CLOCK_MONOTONIC), cpu0 = now_s(CLOCK_PROCESS_CPUTIME_ID);
  const double wall_start = wall0;
  unsigned long long iter = in.now, prev = in.now;
  int it = 0;
  while (it < in.ntimes && times[it] <= in.now) it++;
  while (iter < in.maxtime && it < in.ntimes) {
    const unsigned long long target = times[it] < in.maxtime ? times[it] : in.maxtime;
    CHK(msa_run(ctx, iter, (int64_t)(target - iter), sch.refresh_every));
    iter = target;
    if (iter != times[it]) break;
    CHK(msa_measure(ctx, unsat, demon, R >= 2 ? q : NULL, R >= 2 ? ql : NULL));

    /* measure_time, timing.c:3-33: picoseconds per spin update over the interval */
    const double wall1 = now_s(CLOCK_MONOTONIC), cpu1 = now_s(CLOCK_PROCESS_CPUTIME_ID);
    const double spins = (double)V * R * 64.0 * in.nc * (double)(iter - prev);
    printf("%4.0fs (%.3f ps/s, %.0f%%): ", wall1 - wall0,
           (wall1 - wall0) * 1e12 / spins, (cpu1 - cpu0) / (wall1 - wall0) * 100.0);
    wall0 = wall1;
    cpu0 = cpu1;
    prev = iter;

    /* write_measures, io.c:395-462 */
    printf("Block = %llu\n", iter);
    for (int i = 0; i < nbits; i++) {
      const int b = samples[i];
      double es, ees, ec, eec;
      for (int r = 0; r < R; r++) {
        xs[r] = -(double)(3 * (int64_t)V - 2 * (int64_t)unsat[r * 64 + b]) / (double)V;
        xc[r] = (double)demon[r * 64 + b] / ((double)V * in.nc);
      }
      msa_mean_err(xs, R, &es, &ees);
      msa_mean_err(xc, R, &ec, &eec);
      if (R < 2) ees = eec = 0.0;
      const double Bc = 0.25 * log(1. + 1. / ec);
      printf("Sample = %d\n", b);
      printf("ibeta=%d: E_s=%10.8g E_c=%10.8g T_cl=%10.8g \n", 0, es, ec, 1. / Bc);
      char path[sizeof bitdir[0] + 64];
      FILE *f;
      snprintf(path, sizeof path, "%sevol_E_spins.txt", bitdir[i]);
      if (!(f = fopen(path, "a"))) die("Problems opening %s", path);
      fprintf(f, "%llu %14.10g %14.10g\n", iter, es, ees);
      fclose(f);
      snprintf(path, sizeof path, "%sevol_E_walkers.txt", bitdir[i]);
      if (!(f = fopen(path, "a"))) die("Problems opening %s", path);
      fprintf(f, "%llu %14.10g %14.10g\n", iter, ec, eec);
      fclose(f);
      snprintf(path, sizeof path, "%sevol_T_walkers.txt", bitdir[i]);
      if (!(f = fopen(path, "a"))) die("Problems opening %s", path);
      fprintf(f, "%llu %14.10g \n", iter, 1. / Bc);
      fclose(f);
      /* exact integer observables (no reference counterpart) */
      snprintf(path, sizeof path, "%sevol_E_int.txt", bitdir[i]);
      if (!(f = fopen(path, "a"))) die("Problems opening %s", path);
      fprintf(f, "%llu", iter);
      for (int r = 0; r < R; r++)
        fprintf(f, " %lld %d", (long long)(3 * (int64_t)V - 2 * (int64_t)unsat[r * 64 + b]),
                demon[r * 64 + b]);
      fprintf(f, "\n");
      fclose(f);
      if (R >= 2) {
        snprintf(path, sizeof path, "%sevol_overlap.txt", bitdir[i]);
        if (!(f = fopen(path, "a"))) die("Problems opening %s", path);
        fprintf(f, "%llu", iter);
        for (int p = 0; 2 * p + 1 < R; p++)
          fprintf(f, " %14.10g %14.10g", 1.0 - 2.0 * q[p * 64 + b] / (double)V,
                  1.0 - 2.0 * ql[p * 64 + b] / (3.0 * (double)V));
        fprintf(f, "\n");
        fclose(f);
      }
    }
    fflush(stdout);
    it++;
    if (max_time && now_s(CLOCK_MONOTONIC) - wall_start > (double)max_time) break;
  }
  CHK(msa_sync(ctx));
  printf("Simulation I=%d, nbits=%d finished at iteration %llu\n", sample, nbits, iter);
  msa_destroy(ctx);
  free(unsat);
  free(demon);
  free(q);
  free(ql);
  free(times);
  return 0;
}
