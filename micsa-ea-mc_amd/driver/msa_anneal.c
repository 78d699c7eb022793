/*
 * msa_anneal.c -- the annealing driver: the reference's main() (PROGRAM/src/main.cu:10-195)
 * with its command line, input files, evol_*.txt output and checkpoints,
 * running the sweeps through libmsa_hip.so (include/msa.h).
 *
 * usage: msa_anneal isample nbits beta.dat input.in LUT list_times device
 *                   [max_time [list_samples [schedule.in]]]
 *
 *   isample      job id, < 2^14 (main.cu:70-73); enters every Philox counter
 *   nbits        number of sample lanes written to BIT%03d/ (1..64); all 64
 *                lanes are always simulated
 *   beta.dat     one inverse temperature (io.c:534-558)
 *   input.in     the reference's input file (io.c:464-502); nr = replicas (1..16);
 *                flag 1 new run, 0 restart from the checkpoints, 2 restart
 *                when checkpoints exist, else new run (io.c:680-752)
 *   LUT          the reference's heat-bath table file, checked against beta.dat
 *                (io.c:561-591)
 *   list_times   measurement iterations (io.c:626-655)
 *   device       GPU index
 *   max_time     wall-clock budget in seconds, 0 = none (timing.c:35-58)
 *   list_samples lanes to write (io.c:593-624)
 *   schedule.in  linear anneal "T0 T1 nsweeps refresh_every" (T0 = T1 = 0:
 *                fixed beta from beta.dat), with the Philox demon refresh on
 *                all 64 lanes
 *
 * Without schedule.in the run is the reference program's own (default JANUS
 * build, NBETAS = 1, MAXNSAMPLES = 2) bit for bit: all samples share the J of
 * the JANUS seed (io.c:797-799), spins from Init_u with two sample slots
 * (ini.c:275-299), demons refreshed at iterations 0, 1, 2, 4, ...
 * (main.cu:140-142) from the reference's walker streams (InitWalkers,
 * ini.c:712-781, on the device), a restart repeats the thermalization as the
 * reference does (main.cu:124-128).  Its limits apply: nbits <= 2, samples
 * 0 and 1, L a multiple of 8.
 *
 * Files under <dir>I%03d/ as the reference writes them: master_seedJ.in,
 * run.log, the `running` lock, `error`; BIT%03d/seeds.in, evol_*.txt and
 * walkers.%016llu (+ link walkers); BIT%03d/R%02d/conf.%016llu (+ link conf).
 * Checkpoints are written at every measurement time when L is a multiple
 * of 8 (the reference's multispin layout, which the walkers file uses).
 *
 * With schedule.in, by design: demons are refreshed on the GPU from
 * Philox4x32-10 + LUT bits, so the generators stored in conf and walkers
 * files are the per-replica generators those refreshes draw their counters
 * from, at the checkpoint's iteration, and a restart does not repeat the
 * thermalization, so it continues bit for bit as if it had not stopped.  Signals
 * (SIGTERM, SIGINT, SIGUSR1) stop the run at the next chunk boundary, after
 * the device has finished its queued work, with the reference's error file
 * and lock removal (io.c:970-1014).  The canonical Metropolis thermalization
 * (nterm > 0) runs on the GPU with the LUT file's heat-bath table, its
 * counters drawn from the per-replica generators before those of the demon
 * refreshes, as the reference's _fill_bits calls consume the same generators.
 */
#include <errno.h>
#include <math.h>
#include <signal.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "msa.h"
#include "msa_host.h"

#define PATHLEN (MSA_MAXSTR + 64)

static char g_dir[PATHLEN];        /* <dir>I%03d/ once created */
static int g_locked = 0;           /* this process holds <dir>running */
static const msa_input *g_in = NULL;
static volatile sig_atomic_t g_signal = 0;
static char g_host[100];
static int g_pid;

/* delete_running, io.c:1086-1092 (only a lock this process made) */
static void delete_running(void)
{
  if (!g_locked) return;
  char name[PATHLEN + 16];
  snprintf(name, sizeof name, "%srunning", g_dir);
  remove(name);
  g_locked = 0;
}

static void flush_checkpoint(int dying);

/* create_error + print_and_exit (io.c:996-1014, 1114-1122): the message goes
 * to stderr and, once the sample directory exists, to <dir>error with the
 * run's data; the lock is released.  A checkpoint whose files are still to be
 * written is written first, as the reference writes its confs before it can
 * stop. */
static void die(const char *fmt, ...)
{
  flush_checkpoint(1);
  char msg[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  fprintf(stderr, "%s\n", msg);
  if (g_dir[0]) {
    char name[PATHLEN + 16];
    snprintf(name, sizeof name, "%serror", g_dir);
    FILE *f = fopen(name, "w");
    if (f) {
      fprintf(f, "%s\n", msg);
      if (g_in) msa_print_input(f, g_in);
      fclose(f);
    }
    if (g_in) msa_print_input(stderr, g_in);
  }
  delete_running();
  exit(1);
}

#define CHK(call)                                                               \
  do {                                                                          \
    int rc_ = (call);                                                           \
    if (rc_ != MSA_OK) die("%s: %s", #call, msa_strerror(rc_));                 \
  } while (0)

static void *xmalloc(size_t n, const char *what)
{
  void *p = malloc(n ? n : 1);
  if (!p) die("Problems allocating %s (%zu bytes)", what, n);
  return p;
}

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

/* check_time, timing.c:35-58: stop when the next interval, estimated as the
 * last one scaled by the ratio of their lengths (integer ratio) plus 10%,
 * would overrun the budget.  The first call only starts the clock. */
static int check_time(unsigned long long max_time, unsigned long long dt0,
                      unsigned long long dt1)
{
  static time_t ini = 0;
  static unsigned total = 0;
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  if (ini > 0 && max_time != 0) {
    const unsigned elapsed = (unsigned)(ts.tv_sec - ini);
    total += elapsed;
    const unsigned next = (unsigned)((double)elapsed * (double)(dt0 ? dt1 / dt0 : 0) * 1.1);
    if ((double)max_time - total - next <= 0) return 1;
  }
  ini = ts.tv_sec;
  return 0;
}

static void on_signal(int s) { g_signal = s; }

/* handle_sigterm, io.c:977-982: the handler only records the signal */
static void handle_signals(void)
{
  struct sigaction sa;
  memset(&sa, 0, sizeof sa);
  sa.sa_handler = on_signal;
  sigemptyset(&sa.sa_mask);
  sigaction(SIGTERM, &sa, NULL);
  sigaction(SIGINT, &sa, NULL);
  sigaction(SIGUSR1, &sa, NULL);
}

/* stop_run, io.c:984-994 */
static void check_signal(void)
{
  const int s = g_signal;
  if (!s) return;
  die("Received %s\nAborting programma (SIGNAL=%d)",
      s == SIGTERM ? "SIGTERM" : s == SIGINT ? "SIGINT" : "SIGUSR1", s);
}

static void writelog(FILE *f)
{
  const time_t t = time(NULL);
  fprintf(f, "%s %5d %10ld %s", g_host, g_pid, (long)t, ctime(&t));
}

/* create_running, io.c:1016-1047: <dir>running.<host>.<pid> hard-linked to
 * <dir>running; an existing lock stops the run. */
static void create_running(void)
{
  char name[PATHLEN + 160], lk[PATHLEN + 16];
  gethostname(g_host, sizeof g_host - 1);
  g_host[sizeof g_host - 1] = 0;
  g_pid = (int)getpid();
  snprintf(name, sizeof name, "%srunning.%s.%d", g_dir, g_host, g_pid);
  snprintf(lk, sizeof lk, "%srunning", g_dir);
  FILE *f = fopen(name, "w");
  if (!f) die("Unable to open %s", name);
  writelog(f);
  fclose(f);
  if (link(name, lk) == -1) {
    remove(name);
    fprintf(stderr, "%s exists\n", lk);
    exit(1);
  }
  remove(name);
  g_locked = 1;
  snprintf(name, sizeof name, "%srun.log", g_dir);
  if (!(f = fopen(name, "a"))) die("Unable to open %s", name);
  writelog(f);
  fclose(f);
}

static void print_help(const char *prog)
{
  printf("Usage: %s isample nbits beta.dat input.in LUT list_times device "
         "[max_time [list_samples [schedule.in]]]\n", prog);
  exit(0);
}

/* Two generator states are the same when their reference images are. */
static int gens_equal(const msa_prng *a, const msa_xo256 *ax, const msa_prng *b,
                      const msa_xo256 *bx)
{
  uint8_t ia[MSA_PRNG_REF_BYTES], ib[MSA_PRNG_REF_BYTES];
  uint8_t xa[MSA_XO_REF_BYTES], xb[MSA_XO_REF_BYTES];
  msa_prng_to_ref(a, ia);
  msa_prng_to_ref(b, ib);
  msa_xo256_to_ref(ax, xa);
  msa_xo256_to_ref(bx, xb);
  return !memcmp(ia, ib, sizeof ia) && !memcmp(xa, xb, sizeof xa);
}

/* The state of the run the checkpoint files describe. */
typedef struct {
  int L, R, nc, nbits;
  const int *samples;
  char (*bitdir)[PATHLEN];
  char (*repdir)[MSA_MAXR][PATHLEN + 8];
  const msa_input *in;
  double beta;
  int refmode;                  /* the reference's own workload (no schedule.in) */
  uint8_t walker_p;             /* ring pointer of the walker generators */
  uint64_t seeds_J[64], seeds_C[64];
} run_files;

/* io.c:797-799 (JANUS build): every sample's J seed */
#define JANUS_SEED_J 9164166177945067407ULL

/* check_and_prepare_simulation_backup, io.c:690-752: counts the conf files of
 * the listed samples and resolves flag 2.  Returns the resolved flag and the
 * iteration the confs hold. */
static int prepare_backup(const run_files *rf, int flag, int can_ckpt, unsigned long long *now)
{
  int with = 0, without = 0;
  *now = 0;
  int seen = 0;
  for (int i = 0; i < rf->nbits; i++) {
    int there = 0;
    for (int r = 0; r < rf->R; r++) {
      char name[PATHLEN + 16];
      snprintf(name, sizeof name, "%sconf", rf->repdir[i][r]);
      FILE *f = fopen(name, "rb");
      if (!f) continue;
      uint64_t head[2] = {0, 0};
      const size_t got = fread(head, sizeof head[0], 2, f);
      fclose(f);
      if (got != 2) die("Problems reading %s", name);
      if (seen && head[1] != *now) die("Different number of iterations in samples from %s", g_dir);
      *now = head[1];
      seen = 1;
      there++;
    }
    if (there > 0 && there != rf->R)
      die("There are less conf files than it must be in %s (%d/%d)", rf->bitdir[i], there, rf->R);
    if (there) with++;
    else without++;
  }
  if (with && without) die("Some samples have conf files, others not, in %s", g_dir);
  if (flag == 2) flag = with ? 0 : 1;
  if (flag == 1 && with)
    die("There are some confs in %s.\nIf you want to make a new run, please delete the "
        "configuration manually", g_dir);
  if (flag == 0 && !with) die("There is not %sconf, it is impossible to do a backup",
                              rf->repdir[0][0]);
  if (flag == 0 && !can_ckpt) die("Checkpoints need L to be a multiple of 8");
  return flag;
}

/* get_seeds, io.c:754-871.  Returns 1 when no seed file exists (a new run,
 * which then writes them). */
static int get_seeds(run_files *rf, int flag, uint64_t *seed_J, uint64_t *seed_u,
                     uint64_t *seed_MC, int32_t *countJ0)
{
  int missing = 0;
  char name[PATHLEN + 32];
  snprintf(name, sizeof name, "%smaster_seedJ.in", g_dir);
  FILE *f = fopen(name, "r");
  if (!f) {
    if (flag == 1) missing++;
    else die("There is not file %s, and it must exist", name);
  } else {
    unsigned long long sJ;
    if (fscanf(f, "%llu", &sJ) != 1) die("Nothing in file %s", name);
    fclose(f);
    if (*seed_J > 0 && *seed_J != sJ) die("Master seedJ in input file and %s differ", name);
    *seed_J = sJ;
  }
  *seed_J = nonzero_seed(*seed_J);
  msa_seeds_from_one(*seed_J, rf->seeds_J, 64);
  if (rf->refmode)
    for (int b = 0; b < 64; b++) rf->seeds_J[b] = JANUS_SEED_J;
  for (int i = 0; i < rf->nbits; i++) {
    countJ0[i] = 0;
    snprintf(name, sizeof name, "%sseeds.in", rf->bitdir[i]);
    if (!(f = fopen(name, "r"))) {
      if (flag == 1) {
        missing++;
        continue;
      }
      die("There is not file %s, and it must exist", name);
    }
    unsigned long long sJ, su, sMC[MSA_MAXR];
    int n = fscanf(f, "%llu %llu", &sJ, &su);
    for (int r = 0; r < rf->R; r++) n += fscanf(f, " %llu", &sMC[r]);
    n += fscanf(f, " %d", &countJ0[i]);
    fclose(f);
    if (n != rf->R + 3) die("File %s must have %d fields", name, rf->R + 3);
    if (rf->seeds_J[rf->samples[i]] != sJ)
      die("Seed J generated and %s differ (bit=%d)", name, rf->samples[i]);
    if (*seed_u > 0 && *seed_u != su) die("Seed u in input and %s differ", name);
    *seed_u = su;
    for (int r = 0; r < rf->R; r++) {
      if (seed_MC[r] > 0 && seed_MC[r] != sMC[r]) die("Seed MC in input and %s differ", name);
      seed_MC[r] = sMC[r];
    }
  }
  if (missing == rf->nbits + 1) return 1;
  if (missing > 0) die("Some samples have file seeds.in, others not have");
  return 0;
}

/* write_seeds_in_file, io.c:873-911 */
static void write_seeds(const run_files *rf, uint64_t seed_J, uint64_t seed_u,
                        const uint64_t *seed_MC, const int32_t *countJ0)
{
  char name[PATHLEN + 32];
  snprintf(name, sizeof name, "%smaster_seedJ.in", g_dir);
  FILE *f = fopen(name, "w");
  if (!f) die("Impossible to open seed file to write");
  fprintf(f, "%llu\n", (unsigned long long)seed_J);
  fclose(f);
  for (int i = 0; i < rf->nbits; i++) {
    snprintf(name, sizeof name, "%sseeds.in", rf->bitdir[i]);
    if (!(f = fopen(name, "w"))) die("Impossible to open seed file to write");
    fprintf(f, "%llu %llu", (unsigned long long)rf->seeds_J[rf->samples[i]],
            (unsigned long long)seed_u);
    for (int r = 0; r < rf->R; r++) fprintf(f, " %llu", (unsigned long long)seed_MC[r]);
    fprintf(f, " %d\n", countJ0[i]);
    fclose(f);
  }
}

static void fill_sdata(const run_files *rf, msa_sdata *d, unsigned long long now, int flag,
                       uint64_t seed_u, const uint64_t *seed_MC)
{
  memset(d, 0, sizeof *d);
  d->maxtime = rf->in->maxtime;
  d->now = now;
  d->ntimes = rf->in->ntimes;
  d->nterm = rf->in->nterm;
  d->nc = rf->nc;
  d->nr = rf->R;
  d->l = rf->in->l;
  d->flag = flag;
  d->seed_u = seed_u;
  for (int r = 0; r < rf->R; r++) d->seed_MC[r] = seed_MC[r];
}

static void fill_icds(const run_files *rf, int i, msa_icds *ic)
{
  memset(ic, 0, sizeof *ic);
  ic->first = 0;
  ic->isample = (uint32_t)i;                 /* p2icds[g]->isample = g, ini.c:636 */
  ic->nrep = (uint32_t)rf->R;
  ic->walkers_per_site = (uint32_t)rf->nc;
  ic->lv = (uint32_t)rf->L;
  ic->seed = rf->seeds_C[rf->samples[i]];    /* ini.c:640 */
  ic->beta = rf->beta;
}

/* backup = read_conf + read_walkers (io.c:98-210, 325-393, 913-928): the
 * listed lanes' spins and demons overwrite S, d0, d1; the generators of the
 * conf files are returned; every stored energy is checked against the spins. */
static unsigned long long read_checkpoint(const run_files *rf, const uint64_t *J, uint64_t *S,
                                          uint64_t *d0, uint64_t *d1, msa_prng *pr,
                                          msa_xo256 *xo, msa_prng *wpr, msa_xo256 *wxo)
{
  const int L = rf->L, R = rf->R, nc = rf->nc, n = rf->nbits;
  const uint64_t V = (uint64_t)L * L * L, V8 = V / 8, nw = msa_msc_words(L, R, nc);
  uint8_t *jan = (uint8_t *)xmalloc((size_t)(V8 * R * n), "conf spins");
  int32_t *stored = (int32_t *)xmalloc(sizeof(int32_t) * R * n, "conf energies");
  char e[512], name[PATHLEN + 32];
  unsigned long long now = 0;
  for (int i = 0; i < n; i++)
    for (int r = 0; r < R; r++) {
      snprintf(name, sizeof name, "%sconf", rf->repdir[i][r]);
      msa_sdata d;
      double b;
      msa_prng p;
      msa_xo256 x;
      if (msa_conf_read(name, L, &d, &b, &p, &x, jan + ((uint64_t)i * R + r) * V8,
                        &stored[i * R + r], e, sizeof e))
        die("%s\nProblems reading confs", e);
      if (d.nr != R) die("d.nr = %d != NR = %d", d.nr, R);
      if (d.l != rf->in->l) die("l=%d != L=%d", d.l, rf->in->l);
      if (d.nc != nc || d.nterm != rf->in->nterm) die("Different data from input and %s", name);
      if (d.seed_J != rf->seeds_J[rf->samples[i]])
        die("Different seed_J generated and read in %s", name);
      if (i == 0 && r == 0) now = d.now;
      else if (d.now != now) die("Different number of iterations in %s", name);
      if (fabs(b - rf->beta) > 1e-12)
        die("Different beta (ibeta=0, %.12g != %.12g)", b, rf->beta);
      if (i == 0) {
        pr[r] = p;
        xo[r] = x;
      } else if (!gens_equal(&p, &x, &pr[r], &xo[r])) {
        die("Generators of replica %d in %s differ from those of the first sample", r, name);
      }
    }
  msa_janus_to_lanes(L, R, jan, n, rf->samples, S);
  int32_t *en = (int32_t *)xmalloc(sizeof(int32_t) * R * 64, "energies");
  msa_lane_energy(L, R, J, S, en);
  for (int i = 0; i < n; i++)
    for (int r = 0; r < R; r++)
      if (en[r * 64 + rf->samples[i]] != stored[i * R + r])
        die("Energy read and calculate differ in ibit=%d ir=%d ibeta=0:escalar=%d, read=%d\n"
            "Problems reading confs", rf->samples[i], r, en[r * 64 + rf->samples[i]],
            stored[i * R + r]);
  free(en);
  free(stored);
  free(jan);

  uint64_t *w = (uint64_t *)xmalloc(sizeof(uint64_t) * nw * n, "walkers");
  msa_prng *wp = (msa_prng *)xmalloc(sizeof(msa_prng) * R, "walker generators");
  msa_xo256 *wx = (msa_xo256 *)xmalloc(sizeof(msa_xo256) * R, "walker generators");
  for (int i = 0; i < n; i++) {
    snprintf(name, sizeof name, "%swalkers", rf->bitdir[i]);
    msa_icds ic, want;
    if (msa_walkers_read(name, L, R, nc, &ic, wp, wx, w + (uint64_t)i * nw, e, sizeof e))
      die("%s\nProblems reading walkers", e);
    fill_icds(rf, i, &want);
    if (ic.isample != want.isample || ic.nrep != want.nrep ||
        ic.walkers_per_site != want.walkers_per_site || ic.lv != want.lv ||
        ic.seed != want.seed || fabs(ic.beta - want.beta) > 1e-6)
      die("Different icds from simulation and %s", name);
    for (int r = 0; r < R; r++) {
      if (rf->refmode) {                /* the walker streams of this sample */
        wpr[i * R + r] = wp[r];
        wxo[i * R + r] = wx[r];
      } else if (!gens_equal(&wp[r], &wx[r], &pr[r], &xo[r])) {
        die("Generators of replica %d in %s differ from those of the conf files", r, name);
      }
    }
  }
  msa_msc_to_lanes(L, R, nc, w, n, rf->samples, d0, d1);
  free(w);
  free(wp);
  free(wx);
  return now;
}

/* write_walkers + write_conf (io.c:17-96, 254-323) for the listed lanes at
 * iteration `now`; energies e[R][64]; pr/xo the generators at `now`. */
static void write_checkpoint(const run_files *rf, unsigned long long now,
                             unsigned long long prev, int flag, uint64_t seed_u,
                             const uint64_t *seed_MC, const uint64_t *S, const uint64_t *d0,
                             const uint64_t *d1, const int32_t *energy, const msa_prng *pr,
                             const msa_xo256 *xo, const msa_prng *wpr, const msa_xo256 *wxo)
{
  const int L = rf->L, R = rf->R, nc = rf->nc, n = rf->nbits;
  const uint64_t V8 = (uint64_t)L * L * L / 8, nw = msa_msc_words(L, R, nc);
  char e[512];
  uint64_t *w = (uint64_t *)xmalloc(sizeof(uint64_t) * nw * n, "walkers");
  msa_msc_from_lanes(L, R, nc, d0, d1, n, rf->samples, w);
  for (int i = 0; i < n; i++) {
    msa_icds ic;
    fill_icds(rf, i, &ic);
    if (msa_walkers_write(rf->bitdir[i], now, prev, &ic, L, R, nc, wpr ? wpr + i * R : pr,
                          wxo ? wxo + i * R : xo, w + (uint64_t)i * nw, e, sizeof e))
      die("%s\nProblems writing walkers", e);
  }
  free(w);
  uint8_t *jan = (uint8_t *)xmalloc((size_t)(V8 * R * n), "conf spins");
  msa_janus_from_lanes(L, R, S, n, rf->samples, jan);
  msa_sdata d;
  fill_sdata(rf, &d, now, flag, seed_u, seed_MC);
  for (int i = 0; i < n; i++) {
    d.seed_J = rf->seeds_J[rf->samples[i]];
    for (int r = 0; r < R; r++)
      if (msa_conf_write(rf->repdir[i][r], now, &d, rf->beta, &pr[r], &xo[r], L,
                         jan + ((uint64_t)i * R + r) * V8, energy[r * 64 + rf->samples[i]], e,
                         sizeof e))
        die("%s\nProblems writing confs", e);
  }
  free(jan);
}

/* A checkpoint whose state copy is queued (msa_download_state_async) and
 * whose files are not yet written: the loop completes it after queuing the
 * next sweeps, so the copy and the file work overlap the device, and before
 * any exit (die, a signal, the end of the run). */
static struct {
  int pending;
  msa_ctx *ctx;
  const run_files *rf;
  unsigned long long now, prev;
  int flag;
  uint64_t seed_u;
  const uint64_t *seed_MC, *S, *d0, *d1;
  int32_t energy[MSA_MAXR * 64];
  msa_prng pr[MSA_MAXR];
  msa_xo256 xo[MSA_MAXR];
  const msa_prng *wpr;
  const msa_xo256 *wxo;
} g_ck;

static void flush_checkpoint(int dying)
{
  if (!g_ck.pending) return;
  g_ck.pending = 0;
  const int rc = msa_download_wait(g_ck.ctx);
  if (rc != MSA_OK) {
    if (dying) {
      fprintf(stderr, "the checkpoint at iteration %llu was not written: %s\n", g_ck.now,
              msa_strerror(rc));
      return;
    }
    die("msa_download_wait: %s", msa_strerror(rc));
  }
  write_checkpoint(g_ck.rf, g_ck.now, g_ck.prev, g_ck.flag, g_ck.seed_u, g_ck.seed_MC, g_ck.S,
                   g_ck.d0, g_ck.d1, g_ck.energy, g_ck.pr, g_ck.xo, g_ck.wpr, g_ck.wxo);
}

/* The demon refresh of iterations 0 .. niter-1 on the device: the
 * reference's walker streams (refmode; st = their states now) with
 * get_thresholds at beta.dat, or the Philox stream's counters and LUT pairs
 * (from the generators at the run's start, refreshes before k_start unused). */
static void set_refresh(msa_ctx *ctx, const msa_sched *sch, int refmode, int64_t niter, int R,
                        double beta, int nbits, const int *samples, const msa_walker_stream *st,
                        int64_t k_start, const msa_prng *base_pr, const msa_xo256 *base_xo,
                        uint32_t sample, const uint32_t *keys)
{
  const int64_t nref = msa_sched_count(sch, niter > 0 ? niter : 1);
  if (refmode) {
    /* refresh_walkers (main.cu:340-377): refresh k takes the next segment of
     * each walker stream, with get_thresholds at beta.dat (ini.c:680-710) */
    uint64_t *thr = (uint64_t *)xmalloc(sizeof(uint64_t) * 3 * (size_t)nref, "thresholds");
    msa_thresholds(beta, thr);
    for (int64_t k = 1; k < nref; k++) memcpy(thr + 3 * k, thr, 3 * sizeof(uint64_t));
    CHK(msa_set_walker_streams(ctx, nbits, samples, st, (int32_t)nref, thr));
    free(thr);
    return;
  }
  msa_prng pr[MSA_MAXR];
  msa_xo256 xo[MSA_MAXR];
  memcpy(pr, base_pr, sizeof(msa_prng) * R);
  memcpy(xo, base_xo, sizeof(msa_xo256) * R);
  msa_u32x4 *ctr = (msa_u32x4 *)xmalloc(sizeof(msa_u32x4) * (size_t)(nref * R), "refresh counters");
  msa_lut *luts = (msa_lut *)xmalloc(sizeof(msa_lut) * 2 * (size_t)nref, "refresh tables");
  int32_t *lut_idx = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)nref, "refresh indices");
  /* one LUT pair per distinct beta: a fixed-beta run builds one */
  const int npair = msa_sched_tables_resume(sch, k_start, nref, R, pr, xo, sample, ctr, luts, lut_idx);
  if (npair < 1) die("Problems building the refresh tables (nref=%lld)", (long long)nref);
  CHK(msa_set_luts(ctx, npair, luts));
  CHK(msa_set_refresh(ctx, (int32_t)nref, ctr, keys, lut_idx));
  free(luts);
  free(ctr);
  free(lut_idx);
}

/* stop (io.c:930-968): <dir>changentimes.in, when present, sets a new end of
 * the run and, when it lists them, new measurement times.  The reference's
 * reader has no defined behaviour (an uninitialised pointer); here the file
 * is "maxtime [t_1 t_2 ...]" (msa_read_changentimes), read before the run
 * and after every measurement, and kept (so it is read again each time); a
 * file without a number is removed.  Returns 1 when it changed anything. */
static int change_times(msa_input *in, unsigned long long **times)
{
  char name[PATHLEN + 32], e[512];
  snprintf(name, sizeof name, "%schangentimes.in", g_dir);
  unsigned long long maxtime = in->maxtime, *nt = NULL;
  int n = 0;
  const int rc = msa_read_changentimes(name, &maxtime, &nt, &n, e, sizeof e);
  if (rc < 0) die("%s", e);
  if (rc == 2) {
    printf(" File '%s' exists but holds no number: removed\n", name);
    fflush(stdout);
  }
  if (rc != 1) return 0;
  int changed = maxtime != in->maxtime;
  in->maxtime = maxtime;
  if (n > 0) {
    changed = changed || n != in->ntimes || memcmp(nt, *times, sizeof *nt * (size_t)n);
    free(*times);
    *times = nt;
    in->ntimes = n;
  }
  if (changed) {
    printf("%s: maxtime %llu, %d measurement times\n", name, in->maxtime, in->ntimes);
    fflush(stdout);
  }
  return changed;
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
  check_time(max_time, 1, 0);

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
  if (in.flag < 0 || in.flag > 2) die("data.flag is not recognize (flag=%d; 0 restart, 1 new, 2 either)", in.flag);
  if (in.nc < 1 || in.nc > 6) die("nc=%d outside 1..6", in.nc);
  const int L = in.l & 0xffff, R = in.nr;
  const int can_ckpt = (L % 8 == 0);
  const int refmode = f_sched == NULL;
  if (refmode) {
    /* the reference's limits (header.h:20, main.cu:64-65, ini.c:281-295) */
    if (nbits > 2) die("Error: nbits out of range, 0 < bit < %d", 2);
    for (int i = 0; i < nbits; i++)
      if (samples[i] > 1 || (i > 0 && samples[i] <= samples[i - 1]))
        die("Problems generating spins (samples must be 0 and/or 1, increasing)");
    if (L % 8) die("L must be a multiple of 8");
  }

  printf("Path: %s\n", in.dir);
  msa_print_input(stdout, &in);
  fflush(stdout);

  msa_sched sch;
  memset(&sch, 0, sizeof sch);
  if (f_sched) {
    if (msa_read_schedule(f_sched, &sch, e, sizeof e)) die("%s", e);
    if (sch.T0 <= 0) sch.beta_fixed = beta;
  } else {
    sch.refresh_every = -1;
    sch.beta_fixed = beta;
  }

  /* output tree: dir/I%03d/BIT%03d/R%02d/ (create_sample_path, io.c:657-677) */
  static char bitdir[64][PATHLEN];
  static char repdir[64][MSA_MAXR][PATHLEN + 8];
  snprintf(g_dir, sizeof g_dir, "%sI%03d/", in.dir, sample);
  mkdir(g_dir, 0775);
  for (int i = 0; i < nbits; i++) {
    snprintf(bitdir[i], sizeof bitdir[i], "%sBIT%03d/", g_dir, samples[i]);
    mkdir(bitdir[i], 0775);
    for (int r = 0; r < R; r++) {
      snprintf(repdir[i][r], sizeof repdir[i][r], "%sR%02d/", bitdir[i], r);
      if (can_ckpt) mkdir(repdir[i][r], 0775);
    }
  }
  printf("Dirsample = %s\n", g_dir);
  g_in = &in;

  run_files rf;
  memset(&rf, 0, sizeof rf);
  rf.L = L;
  rf.R = R;
  rf.nc = in.nc;
  rf.nbits = nbits;
  rf.samples = samples;
  rf.bitdir = bitdir;
  rf.repdir = repdir;
  rf.in = &in;
  rf.beta = beta;
  rf.refmode = refmode;

  unsigned long long conf_now = 0;
  const int flag = can_ckpt ? prepare_backup(&rf, in.flag, can_ckpt, &conf_now)
                            : (in.flag == 0 ? (die("Checkpoints need L to be a multiple of 8"), 0) : 1);
  uint64_t seed_J = in.seed_J, seed_u = in.seed_u;
  uint64_t seed_MC[MSA_MAXR];
  for (int r = 0; r < R; r++) seed_MC[r] = in.seed_MC[r];
  int32_t countJ0[64];
  const int new_seeds = get_seeds(&rf, flag, &seed_J, &seed_u, seed_MC, countJ0);

  printf("Betas and LUT have been read correctly.\n");
  printf("Initializating J's and Spins ...\n");
  fflush(stdout);
  handle_signals();
  create_running();

  /* Init_Random, ini.c:33-59 */
  seed_u = nonzero_seed(seed_u);
  for (int r = 0; r < R; r++) seed_MC[r] = nonzero_seed(seed_MC[r]);
  msa_seeds_from_one(seed_u, rf.seeds_C, 64);

  const uint64_t V = (uint64_t)L * L * L;
  /* the state arrays go to and from the device at the start and at every
   * measurement (checkpoints): page-locked where the runtime gives it, so
   * those transfers run at the host link's rate */
  const size_t nJ = 3 * V * sizeof(uint64_t), nS = (size_t)R * V * sizeof(uint64_t);
  const size_t nD = (size_t)R * in.nc * V * sizeof(uint64_t);
  uint64_t *J = (uint64_t *)xmalloc(nJ, "couplings");
  uint64_t *S = (uint64_t *)msa_host_alloc((int64_t)nS);
  uint64_t *D0 = S ? (uint64_t *)msa_host_alloc((int64_t)nD) : NULL;
  uint64_t *D1 = D0 ? (uint64_t *)msa_host_alloc((int64_t)nD) : NULL;
  const int pinned = D1 != NULL;
  if (!pinned) {
    msa_host_free(S);
    msa_host_free(D0);
    S = (uint64_t *)xmalloc(nS, "spins");
    D0 = (uint64_t *)xmalloc(nD, "demons");
    D1 = (uint64_t *)xmalloc(nD, "demons");
  }
  memset(D0, 0, nD);
  memset(D1, 0, nD);
  if (refmode) {
    msa_build_couplings_one(L, JANUS_SEED_J, J);
    msa_build_spins_slots(L, R, seed_u, 2, S);
  } else {
    msa_build_couplings(L, seed_J, J);
    msa_build_spins(L, R, seed_u, S);
  }

  /* countJ check of Init_J, ini.c:500-511 */
  int32_t countJ[64];
  msa_lane_countJ(L, J, countJ);
  for (int i = 0; i < nbits; i++) {
    const int32_t c = countJ[samples[i]];
    if ((flag == 0 || countJ0[i] != 0) && c != countJ0[i])
      die("countJ and countJ0 differ (%d != %d --- ibit=%d)\nProblems generating J's", c,
          countJ0[i], samples[i]);
    countJ0[i] = c;
  }

  msa_prng pr[MSA_MAXR];
  msa_xo256 xo[MSA_MAXR];
  uint32_t keys[MSA_MAXR];
  for (int r = 0; r < R; r++) msa_cpu_generators(&pr[r], &xo[r], &keys[r], seed_MC[r]);

  /* the walker streams of the reference's demon refresh (Init_Random,
   * ini.c:53-58), per listed sample and replica */
  msa_prng *wpr = NULL;
  msa_xo256 *wxo = NULL;
  if (refmode) {
    wpr = (msa_prng *)xmalloc(sizeof(msa_prng) * nbits * R, "walker generators");
    wxo = (msa_xo256 *)xmalloc(sizeof(msa_xo256) * nbits * R, "walker generators");
    msa_walker_generators(seed_u, nbits, R, wpr, wxo);
    rf.walker_p = wpr[0].p;
  }

  /* the reference reads changentimes.in while restoring confs (io.c:149)
   * and once more before the run (main.cu:107) */
  change_times(&in, &times);
  unsigned long long start = in.now;
  if (flag == 0) {
    start = read_checkpoint(&rf, J, S, D0, D1, pr, xo, wpr, wxo);
    msa_print_input(stdout, &in);
    fflush(stdout);
    if (start >= in.maxtime) die("There are %llu iterations in %s\nProblems in backup", start, g_dir);
  }
  if (new_seeds) write_seeds(&rf, seed_J, seed_u, seed_MC, countJ0);

  msa_geom g = {L, R, in.nc};
  msa_ctx *ctx = NULL;
  CHK(msa_create(device, &g, &ctx));
  {
    /* the sweep kernel, its tile and the refresh mode for this geometry,
     * timed on a scratch lattice (every choice gives the same bits) */
    int32_t km = 0, kty = 0, ktz = 0, rmode = 0;
    float kms = 0;
    CHK(msa_autotune(ctx, 16, &km, &kty, &ktz, &kms));
    CHK(msa_get_refresh_mode(ctx, &rmode));
    printf("Kernel: mode %d, tile %d x %d, %.4f ms per sweep without refresh; refresh mode %d\n",
           km, kty, ktz, kms, rmode);
    fflush(stdout);
  }
  CHK(msa_upload_state(ctx, J, S, flag == 0 ? D0 : NULL, flag == 0 ? D1 : NULL));

  if (in.nterm > 0 && (flag == 1 || refmode)) {
    /* THERMALIZATION, main.cu:124-128: Monte_Carlo(flag = 0) nterm times, on a
     * restart too in the reference's own workload */
    printf("Thermalization: %d Metropolis iterations at beta = %.6f\n", in.nterm, beta);
    fflush(stdout);
    msa_u32x4 *tctr = (msa_u32x4 *)xmalloc(sizeof(msa_u32x4) * 2 * (size_t)in.nterm * R,
                                           "thermal counters");
    msa_thermal_counters(pr, xo, R, in.nterm, (uint32_t)sample, tctr);
    CHK(msa_run_thermal(ctx, in.nterm, &lut_file, tctr, keys));
    free(tctr);
  } else if (in.nterm > 0) {
    printf("Restart: the checkpoint holds the thermalized state; thermalization is not repeated\n");
  }

  /* the generators at `start`; those at a later iteration T follow by
   * skipping the draws of the refreshes in [start, T) */
  const int64_t k_start = msa_sched_count(&sch, (int64_t)start);
  msa_prng base_pr[MSA_MAXR];
  msa_xo256 base_xo[MSA_MAXR];
  memcpy(base_pr, pr, sizeof pr);
  memcpy(base_xo, xo, sizeof xo);

  msa_walker_stream *wst = NULL;
  if (refmode) {
    wst = (msa_walker_stream *)xmalloc(sizeof(msa_walker_stream) * nbits * R, "walker streams");
    for (int k = 0; k < nbits * R; k++) msa_walker_stream_of(&wpr[k], &wxo[k], &wst[k]);
  }
  set_refresh(ctx, &sch, refmode, (int64_t)in.maxtime, R, beta, nbits, samples, wst, k_start,
              base_pr, base_xo, (uint32_t)sample, keys);
  unsigned long long tables_to = in.maxtime;   /* iterations the refresh tables cover */

  int32_t *unsat = (int32_t *)calloc((size_t)R * 64, sizeof(int32_t));
  int32_t *demon = (int32_t *)calloc((size_t)R * 64, sizeof(int32_t));
  int32_t *q = (int32_t *)calloc((size_t)(R / 2 + 1) * 64, sizeof(int32_t));
  int32_t *ql = (int32_t *)calloc((size_t)(R / 2 + 1) * 64, sizeof(int32_t));
  int32_t *energy = (int32_t *)calloc((size_t)R * 64, sizeof(int32_t));
  double xs[MSA_MAXR], xc[MSA_MAXR];

  /* iterations per chunk between signal checks: about 2^38 spin updates
   * (tens of milliseconds), at most 65536 iterations */
  const double per_iter = (double)V * R * 64.0 * in.nc;
  int64_t chunk = per_iter >= 2.7e11 ? 1 : (int64_t)(2.7e11 / per_iter);
  if (chunk > 65536) chunk = 65536;

  double wall0 = now_s(CLOCK_MONOTONIC), cpu0 = now_s(CLOCK_PROCESS_CPUTIME_ID);
  unsigned long long iter = start, prev = start;
  int it = 0;
  while (it < in.ntimes && times[it] <= start) it++;
  while (iter < in.maxtime && it < in.ntimes) {
    const unsigned long long target = times[it] < in.maxtime ? times[it] : in.maxtime;
    while (iter < target) {
      const int64_t n = (int64_t)(target - iter) < chunk ? (int64_t)(target - iter) : chunk;
      CHK(msa_run(ctx, iter, n, sch.refresh_every));
      iter += (unsigned long long)n;
      flush_checkpoint(0);                /* the previous checkpoint's files */
      CHK(msa_sync(ctx));
      check_signal();
    }
    if (iter != times[it]) break;
    CHK(msa_measure(ctx, unsat, demon, R >= 2 ? q : NULL, R >= 2 ? ql : NULL));

    /* measure_time, timing.c:3-33: picoseconds per spin update over the interval */
    const double wall1 = now_s(CLOCK_MONOTONIC), cpu1 = now_s(CLOCK_PROCESS_CPUTIME_ID);
    const double spins = (double)V * R * 64.0 * in.nc * (double)(iter - prev);
    printf("%4.0fs (%.3f ps/s, %.0f%%): ", wall1 - wall0,
           (wall1 - wall0) * 1e12 / spins, (cpu1 - cpu0) / (wall1 - wall0) * 100.0);
    wall0 = wall1;
    cpu0 = cpu1;
    const unsigned long long dt0 = iter - prev;
    prev = iter;

    /* write_measures, io.c:395-462 */
    printf("Block = %llu\n", iter);
    for (int r = 0; r < R; r++)
      for (int b = 0; b < 64; b++) energy[r * 64 + b] = (int32_t)(3 * (int64_t)V - 2 * (int64_t)unsat[r * 64 + b]);
    for (int i = 0; i < nbits; i++) {
      const int b = samples[i];
      double es, ees, ec, eec;
      for (int r = 0; r < R; r++) {
        xs[r] = -(double)energy[r * 64 + b] / (double)V;
        xc[r] = (double)demon[r * 64 + b] / ((double)V * in.nc);
      }
      msa_mean_err(xs, R, &es, &ees);
      msa_mean_err(xc, R, &ec, &eec);
      if (R < 2) ees = eec = 0.0;
      const double Bc = 0.25 * log(1. + 1. / ec);
      printf("Sample = %d\n", b);
      printf("ibeta=%d: E_s=%10.8g E_c=%10.8g T_cl=%10.8g \n", 0, es, ec, 1. / Bc);
      char path[PATHLEN + 64];
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
      for (int r = 0; r < R; r++) fprintf(f, " %d %d", energy[r * 64 + b], demon[r * 64 + b]);
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

    if (can_ckpt) {
      /* copy_from_GPU_for_write (main.cu:473-490), write_walkers and
       * write_conf, completed once the next sweeps are queued
       * (flush_checkpoint) */
      CHK(msa_download_state_async(ctx, S, D0, D1));
      msa_prng *gpr = g_ck.pr;
      msa_xo256 *gxo = g_ck.xo;
      memcpy(gpr, base_pr, sizeof g_ck.pr);
      memcpy(gxo, base_xo, sizeof g_ck.xo);
      if (refmode) {
        /* the thermal generators are not drawn after the thermalization; the
         * walker generators are the streams after the last refresh */
        msa_walker_stream *st = (msa_walker_stream *)xmalloc(
            sizeof(msa_walker_stream) * nbits * R, "walker streams");
        CHK(msa_get_walker_streams(ctx, st));
        for (int k = 0; k < nbits * R; k++) msa_walker_stream_to(&st[k], rf.walker_p, &wpr[k], &wxo[k]);
        free(st);
      } else {
        msa_gen_skip(gpr, gxo, R, msa_sched_count(&sch, (int64_t)iter) - k_start);
      }
      g_ck.ctx = ctx;
      g_ck.rf = &rf;
      g_ck.now = iter;
      g_ck.prev = it > 0 ? times[it - 1] : 0;
      g_ck.flag = flag;
      g_ck.seed_u = seed_u;
      g_ck.seed_MC = seed_MC;
      g_ck.S = S;
      g_ck.d0 = D0;
      g_ck.d1 = D1;
      memcpy(g_ck.energy, energy, sizeof(int32_t) * 64 * (size_t)R);
      g_ck.wpr = wpr;
      g_ck.wxo = wxo;
      g_ck.pending = 1;
    }
    check_signal();
    if (change_times(&in, &times)) {
      /* a new list: continue with its first time after this one; a later
       * end needs refresh tables that reach it */
      it = 0;
      while (it < in.ntimes && times[it] <= iter) it++;
      if (in.maxtime > tables_to) {
        if (refmode) CHK(msa_get_walker_streams(ctx, wst));
        set_refresh(ctx, &sch, refmode, (int64_t)in.maxtime, R, beta, nbits, samples, wst, k_start,
                    base_pr, base_xo, (uint32_t)sample, keys);
        tables_to = in.maxtime;
      }
    } else {
      ++it;
    }
    if (check_time(max_time, dt0, it < in.ntimes ? times[it] - iter : 0) || it >= in.ntimes)
      break;
  }
  flush_checkpoint(0);
  CHK(msa_sync(ctx));
  delete_running();
  printf("Simulation I=%d, nbits=%d finished at iteration %llu\n", sample, nbits, iter);
  fflush(stdout);
  msa_destroy(ctx);
  free(unsat);
  free(demon);
  free(q);
  free(ql);
  free(energy);
  free(times);
  free(J);
  if (pinned) {
    msa_host_free(S);
    msa_host_free(D0);
    msa_host_free(D1);
  } else {
    free(S);
    free(D0);
    free(D1);
  }
  free(wpr);
  free(wxo);
  free(wst);
  return 0;
}
