/*
 * msa_serve.c -- the served form: a continuous-batching request loop on one
 * GPU (no reference counterpart: the reference runs one sample set per
 * process, PROGRAM/src/main.cu:33-50).
 *
 * usage: msa_serve requests.txt results.txt [--max-bytes N] [--epoch N]
 *                  [--block B] [--workers W] [--state-dir DIR] [--solo-min-L N]
 *                  [--solo-min-iter N] [--solo-share X]
 *
 * requests.txt   one request per line: rid L R niter T0 T1 seed arrival_s
 *                (arrival in seconds after the start of the stream; '#' starts
 *                a comment line).  Each request is a micSA anneal of an L^3
 *                lattice, 64 +-J samples per lane, R replicas, nc = 1, state
 *                built by msa_init_philox(seed), T0 -> T1 linearly over niter
 *                sweeps with the demon refresh every sweep, temperatures on the
 *                2^-10 grid of the shared LUT table (T in [0.5, 2.5]), refresh
 *                counters from seed_MC[r] = seed + r with jobid = rid mod 2^14.
 * results.txt    one line per request, in finishing order: rid L R niter
 *                arrival admit finish latency, then the 64 per-lane integer
 *                energies 3V - 2 n_unsat of each replica.
 * --max-bytes    device memory the admitted lattices may hold (default 8 GiB)
 * --epoch        iterations per batch run between admissions (default 64)
 * --block        threads per block of the batch kernel; 0 = the autotuner
 *                (time 256, 512 and 1024 on one request of each lattice size
 *                of the stream, 8 iterations each; default)
 * --workers      host threads building refresh tables (default 4)
 * --state-dir    write each request's final spins and demons (canonical order,
 *                spins [R][V] then d0 and d1 [R][V], uint64) to
 *                DIR/req_<rid>.bin
 * --solo-min-L   requests with L >= N run one at a time as plain lattices
 *                (msa_run with the autotuner's kernel and refresh mode, the
 *                same grid tables) beside the batch, which takes the others;
 *                0 = every request in the batch (default 128: at L = 128
 *                the plain form, its refresh drawn ahead, runs 1.4 times the
 *                batch kernel's rate; at L = 96 the two are within 5 %)
 * --solo-min-iter  only requests of at least N sweeps go solo (default 1024:
 *                a solo request pays its own start -- tables, a new draw
 *                pipeline -- which short requests do not amortise)
 * --solo-share   an eligible request goes solo while the solo lane's pending
 *                work (sites x sweeps of its unfinished requests) is at most X
 *                times the batch's, else to the batch, so that one lane does
 *                not finish long after the other (default 2; 0 = every
 *                eligible request solo)
 *
 * A request is admitted, in arrival order, once it has arrived, its tables
 * are built and its footprint fits next to the lattices on the device; every
 * admitted lattice advances in the same launches (msa_batch_*); one epoch
 * stays queued on the device while the previous epoch's finished requests are
 * measured and retired.  Solo requests (--solo-min-L) wait in their own FIFO
 * and run on a plain context per lattice geometry (built, autotuned and given
 * the grid tables before the stream starts), in chunks of iterations queued
 * whenever the previous chunk has run, so the loop keeps serving the batch;
 * their device memory is not counted against --max-bytes.  Prints a JSON
 * summary line on stdout.
 */
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "msa.h"
#include "msa_host.h"

#define TGRID (1.0 / 1024)
#define T_LO 0.5
#define T_HI 2.5

typedef struct {
  int rid, L, R;
  int64_t niter;
  double T0, T1, arrival;
  uint64_t seed;
  /* tables, built by a worker */
  msa_u32x4 *ctr;
  uint32_t keys[MSA_MAXR];
  int32_t *idx;
  int64_t nref;
  volatile int ready;       /* 1 built, -1 failed */
  /* outcome */
  double admit, finish;
  int32_t *energy;          /* [R][64] */
} request;

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void fail(const char *fmt, ...) __attribute__((format(printf, 1, 2), noreturn));
static void fail(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "msa_serve: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  exit(1);
}

#define CHK(call)                                                              \
  do {                                                                         \
    int rc_ = (call);                                                          \
    if (rc_ != MSA_OK) fail("%s: %s", #call, msa_strerror(rc_));               \
  } while (0)

static const int32_t g_nlo = 512;                 /* round(T_LO / TGRID) */
static const int32_t g_nlut = 2560 - 512 + 1;     /* round(T_HI / TGRID) - nlo + 1 */

/* The refresh tables of one request (msa_sched_tables_grid): counters,
 * keys and indices into the shared grid table. */
static int build_tables(request *q)
{
  msa_sched s = {q->T0, q->T1, q->niter, 1, 0.0};
  q->nref = msa_sched_count(&s, q->niter);
  const int64_t n = q->nref > 0 ? q->nref : 1;
  q->ctr = (msa_u32x4 *)malloc(sizeof(msa_u32x4) * (size_t)(n * q->R));
  q->idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  if (!q->ctr || !q->idx) return -1;
  uint64_t seeds[MSA_MAXR];
  for (int r = 0; r < q->R; r++) seeds[r] = q->seed + (uint64_t)r;
  return msa_sched_tables_grid(&s, q->nref, q->R, seeds, (uint32_t)(q->rid % 16384), TGRID, g_nlo,
                               g_nlut, q->ctr, q->keys, q->idx);
}

static void free_tables(request *q)
{
  free(q->ctr);
  free(q->idx);
  q->ctr = NULL;
  q->idx = NULL;
}

static msa_job job_of(const request *q)
{
  msa_job j;
  memset(&j, 0, sizeof j);
  j.geom.L = q->L;
  j.geom.R = q->R;
  j.geom.nc = 1;
  j.seed = q->seed;
  j.niter = q->niter;
  j.refresh_every = 1;
  j.nref = (int32_t)q->nref;
  j.ctr = q->ctr;
  j.key = q->keys;
  j.lut_idx = q->idx;
  return j;
}

/* ------------------------------------------------ table-building workers -- */
typedef struct {
  pthread_mutex_t mu;
  pthread_cond_t cv;
  int *todo;          /* request indices, FIFO */
  int head, tail, stop;
  request *reqs;
} pool;

static void *worker(void *arg)
{
  pool *p = (pool *)arg;
  for (;;) {
    pthread_mutex_lock(&p->mu);
    while (p->head == p->tail && !p->stop) pthread_cond_wait(&p->cv, &p->mu);
    if (p->head == p->tail && p->stop) {
      pthread_mutex_unlock(&p->mu);
      return NULL;
    }
    const int i = p->todo[p->head++];
    pthread_mutex_unlock(&p->mu);
    const int rc = build_tables(&p->reqs[i]);
    __atomic_store_n(&p->reqs[i].ready, rc == 0 ? 1 : -1, __ATOMIC_RELEASE);
  }
}

static void submit(pool *p, int i)
{
  pthread_mutex_lock(&p->mu);
  p->todo[p->tail++] = i;
  pthread_cond_signal(&p->cv);
  pthread_mutex_unlock(&p->mu);
}

static int ready_of(const request *q) { return __atomic_load_n(&q->ready, __ATOMIC_ACQUIRE); }

/* --------------------------------------------------------- block tuning -- */
/* Threads per block of the batch kernel: each candidate timed on a scratch
 * batch holding one request of each lattice size of the stream (64
 * iterations, T 2.0 -> 0.7), 8 iterations after 2 of warm-up. */
static int tune_block(const msa_lut *luts, const request *reqs, int n)
{
  int sizes[64], ns = 0;
  for (int i = 0; i < n && ns < 64; i++) {
    int seen = 0;
    for (int k = 0; k < ns; k++) seen |= sizes[k] == reqs[i].L;
    if (!seen) sizes[ns++] = reqs[i].L;
  }
  request sample[64];
  memset(sample, 0, sizeof sample);
  for (int k = 0; k < ns; k++) {
    sample[k].rid = k;
    sample[k].L = sizes[k];
    sample[k].R = 1;
    sample[k].niter = 64;
    sample[k].T0 = 2.0;
    sample[k].T1 = 0.7;
    sample[k].seed = 7 + (uint64_t)k;
    if (build_tables(&sample[k])) fail("tables of the tuning sample");
  }
  const int cand[3] = {256, 512, 1024};
  int best = 1024;
  double tbest = 1e30;
  for (int c = 0; c < 3; c++) {
    msa_batch *b = NULL;
    CHK(msa_batch_create(0, 0, &b));
    CHK(msa_batch_set_luts(b, g_nlut, luts));
    CHK(msa_batch_set_shape(b, cand[c], 8));
    for (int k = 0; k < ns; k++) {
      int32_t slot;
      msa_job j = job_of(&sample[k]);
      CHK(msa_batch_add(b, &j, &slot));
    }
    CHK(msa_batch_run(b, 2));
    CHK(msa_batch_sync(b));
    const double t0 = now_s();
    CHK(msa_batch_run(b, 8));
    CHK(msa_batch_sync(b));
    const double dt = now_s() - t0;
    CHK(msa_batch_destroy(b));
    if (dt < tbest) {
      tbest = dt;
      best = cand[c];
    }
  }
  for (int k = 0; k < ns; k++) free_tables(&sample[k]);
  return best;
}

/* ---------------------------------------------------------- solo lane -- */
/* Plain contexts of the solo requests, one per (L, R), reused. */
typedef struct {
  int L, R;
  msa_ctx *ctx;
} solo_ctx;

#define SOLO_CHUNK 2048   /* iterations queued at a time */

static msa_ctx *solo_ctx_of(solo_ctx *sc, int nsc, int L, int R)
{
  for (int k = 0; k < nsc; k++)
    if (sc[k].L == L && sc[k].R == R) return sc[k].ctx;
  return NULL;
}

/* A finished request's per-lane energies, its state file and its results
 * line; from its plain context (ctx) or its batch slot. */
static void write_result(FILE *out, request *q, const char *state_dir, msa_ctx *ctx, msa_batch *b,
                         int slot)
{
  int32_t *unsat = q->energy;
  if (ctx) CHK(msa_measure(ctx, unsat, NULL, NULL, NULL));
  else CHK(msa_batch_measure(b, slot, unsat, NULL, NULL, NULL));
  const int64_t V = (int64_t)q->L * q->L * q->L;
  for (int i = 0; i < q->R * 64; i++) unsat[i] = (int32_t)(3 * V - 2 * (int64_t)unsat[i]);
  if (state_dir) {
    uint64_t *st = (uint64_t *)malloc(sizeof(uint64_t) * 3 * (size_t)(q->R * V));
    if (!st) fail("out of memory");
    if (ctx) CHK(msa_download_state(ctx, st, st + q->R * V, st + 2 * q->R * V));
    else CHK(msa_batch_download_state(b, slot, st, st + q->R * V, st + 2 * q->R * V));
    char path[4096];
    snprintf(path, sizeof path, "%s/req_%d.bin", state_dir, q->rid);
    FILE *sf = fopen(path, "wb");
    if (!sf || fwrite(st, sizeof(uint64_t), 3 * (size_t)(q->R * V), sf) != 3 * (size_t)(q->R * V))
      fail("cannot write %s", path);
    fclose(sf);
    free(st);
  }
  fprintf(out, "%d %d %d %lld %.6f %.6f %.6f %.6f", q->rid, q->L, q->R, (long long)q->niter,
          q->arrival, q->admit, q->finish, q->finish - q->arrival);
  for (int i = 0; i < q->R * 64; i++) fprintf(out, " %d", q->energy[i]);
  fprintf(out, "\n");
}

/* A request's work: sites x replicas x sweeps. */
static double work_of(const request *q)
{
  return (double)q->R * q->L * q->L * q->L * (double)q->niter;
}

/* ------------------------------------------------------------------ main -- */
static int cmp_double(const void *a, const void *b)
{
  const double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : x > y;
}

static double percentile(double *v, int n, double p)
{
  /* linear interpolation between closest ranks (numpy's default) */
  qsort(v, (size_t)n, sizeof *v, cmp_double);
  const double pos = p / 100.0 * (n - 1);
  const int lo = (int)floor(pos), hi = lo + 1 < n ? lo + 1 : lo;
  return v[lo] + (v[hi] - v[lo]) * (pos - lo);
}

static int cmp_int(const void *a, const void *b)
{
  const int x = *(const int *)a, y = *(const int *)b;
  return x < y ? -1 : x > y;
}

static int cmp_arrival(const void *a, const void *b)
{
  const request *x = (const request *)a, *y = (const request *)b;
  return x->arrival < y->arrival ? -1 : x->arrival > y->arrival;
}

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: %s requests.txt results.txt [--max-bytes N] [--epoch N] [--block B] "
                    "[--workers W] [--state-dir DIR] [--solo-min-L N] [--solo-min-iter N] "
                    "[--solo-share X]\n", argv[0]);
    return 1;
  }
  int64_t max_bytes = (int64_t)8 << 30;
  int epoch = 64, block = 0, workers = 4;
  const char *state_dir = NULL;
  int solo_min_L = 128;
  int64_t solo_min_iter = 1024;
  double solo_share = 2.0;
  for (int a = 3; a < argc; a++) {
    const char *opt = argv[a];
    if (a + 1 >= argc) fail("option %s needs a value", opt);
    const char *v = argv[++a];
    if (!strcmp(opt, "--max-bytes")) max_bytes = strtoll(v, NULL, 10);
    else if (!strcmp(opt, "--epoch")) epoch = atoi(v);
    else if (!strcmp(opt, "--block")) block = atoi(v);
    else if (!strcmp(opt, "--workers")) workers = atoi(v);
    else if (!strcmp(opt, "--state-dir")) state_dir = v;
    else if (!strcmp(opt, "--solo-min-L")) solo_min_L = atoi(v);
    else if (!strcmp(opt, "--solo-min-iter")) solo_min_iter = strtoll(v, NULL, 10);
    else if (!strcmp(opt, "--solo-share")) solo_share = atof(v);
    else fail("unknown option %s", opt);
  }
  if (epoch < 1 || workers < 1 || max_bytes < 1 || solo_min_L < 0 || solo_min_iter < 0 ||
      solo_share < 0)
    fail("bad option value");

  /* requests */
  FILE *f = fopen(argv[1], "r");
  if (!f) fail("cannot open %s", argv[1]);
  int cap = 1024, n = 0;
  request *reqs = (request *)calloc((size_t)cap, sizeof *reqs);
  char line[512];
  while (reqs && fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    if (n == cap) {
      cap *= 2;
      reqs = (request *)realloc(reqs, sizeof *reqs * (size_t)cap);
      if (!reqs) break;
      memset(reqs + n, 0, sizeof *reqs * (size_t)(cap - n));
    }
    request *q = &reqs[n];
    long long niter;
    unsigned long long seed;
    if (sscanf(line, "%d %d %d %lld %lf %lf %llu %lf", &q->rid, &q->L, &q->R, &niter, &q->T0, &q->T1,
               &seed, &q->arrival) != 8)
      fail("%s: bad request line %d", argv[1], n + 1);
    q->niter = niter;
    q->seed = seed;
    if (q->L < 4 || q->L % 2 || q->R < 1 || q->R > MSA_MAXR || q->niter < 1 || q->rid < 0)
      fail("request %d: L=%d R=%d niter=%lld", q->rid, q->L, q->R, niter);
    n++;
  }
  fclose(f);
  if (!reqs) fail("out of memory");
  if (n == 0) fail("no requests in %s", argv[1]);
  {
    /* request ids name the results and state files: they must be unique */
    int *ids = (int *)malloc(sizeof(int) * (size_t)n);
    if (!ids) fail("out of memory");
    for (int i = 0; i < n; i++) ids[i] = reqs[i].rid;
    qsort(ids, (size_t)n, sizeof *ids, cmp_int);
    for (int i = 0; i + 1 < n; i++)
      if (ids[i] == ids[i + 1]) fail("request id %d appears twice", ids[i]);
    free(ids);
  }
  qsort(reqs, (size_t)n, sizeof *reqs, cmp_arrival);
  for (int i = 0; i < n; i++) {
    reqs[i].energy = (int32_t *)calloc((size_t)reqs[i].R * 64, sizeof(int32_t));
    if (!reqs[i].energy) fail("out of memory");
  }
  FILE *out = fopen(argv[2], "w");
  if (!out) fail("cannot open %s", argv[2]);

  /* the shared temperature-grid LUT table */
  msa_lut *luts = (msa_lut *)malloc(sizeof(msa_lut) * 2 * (size_t)g_nlut);
  if (!luts) fail("out of memory");
  msa_lut_grid(TGRID, g_nlo, g_nlut, luts);
  if (block == 0) block = tune_block(luts, reqs, n);

  msa_batch *b = NULL;
  CHK(msa_batch_create(0, max_bytes, &b));
  CHK(msa_batch_set_luts(b, g_nlut, luts));
  CHK(msa_batch_set_shape(b, block, epoch));

  /* the solo lane's contexts, ready before the stream starts */
  solo_ctx sc[64];
  int nsc = 0;
  for (int i = 0; i < n && solo_min_L > 0; i++) {
    const request *q = &reqs[i];
    if (q->L < solo_min_L || solo_ctx_of(sc, nsc, q->L, q->R)) continue;
    if (nsc == 64) fail("more than 64 solo lattice geometries");
    msa_geom g = {q->L, q->R, 1};
    sc[nsc].L = q->L;
    sc[nsc].R = q->R;
    CHK(msa_create(0, &g, &sc[nsc].ctx));
    CHK(msa_autotune(sc[nsc].ctx, 16, NULL, NULL, NULL, NULL));
    CHK(msa_set_luts(sc[nsc].ctx, g_nlut, luts));
    nsc++;
  }
  int *sq = (int *)malloc(sizeof(int) * (size_t)n), sqh = 0, sqt = 0;
  int solo_cur = -1;
  int64_t solo_enq = 0;
  msa_ctx *solo = NULL;
  if (!sq) fail("out of memory");

  pool p;
  memset(&p, 0, sizeof p);
  pthread_mutex_init(&p.mu, NULL);
  pthread_cond_init(&p.cv, NULL);
  p.todo = (int *)malloc(sizeof(int) * (size_t)n);
  p.reqs = reqs;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)workers);
  if (!p.todo || !th) fail("out of memory");
  for (int w = 0; w < workers; w++) pthread_create(&th[w], NULL, worker, &p);

  /* slot -> request index */
  int nslot_cap = 256;
  int *slot_req = (int *)malloc(sizeof(int) * (size_t)nslot_cap);
  for (int s = 0; s < nslot_cap; s++) slot_req[s] = -1;
  int *queue = (int *)malloc(sizeof(int) * (size_t)n), qh = 0, qt = 0;
  int *finishing = (int *)malloc(sizeof(int) * (size_t)nslot_cap), nfin = 0;
  int *newfin = (int *)malloc(sizeof(int) * (size_t)nslot_cap);
  int64_t *tp = (int64_t *)malloc(sizeof(int64_t) * (size_t)nslot_cap);
  int64_t *te = (int64_t *)malloc(sizeof(int64_t) * (size_t)nslot_cap);
  if (!slot_req || !queue || !finishing || !newfin || !tp || !te) fail("out of memory");
  int nactive = 0, done = 0, nxt = 0, max_active = 0, max_slot = 0;
  int64_t max_bytes_used = 0;
  double updates = 0;
  double solo_pending = 0, batch_pending = 0;   /* work of each lane's unfinished requests */
  const double t_start = now_s();
  double last_note = t_start;

  while (done < n) {
    const double now = now_s() - t_start;
    while (nxt < n && reqs[nxt].arrival <= now) {
      submit(&p, nxt);
      const request *r = &reqs[nxt];
      const double w = work_of(r);
      if (solo_min_L > 0 && r->L >= solo_min_L && r->niter >= solo_min_iter &&
          (solo_share == 0 || solo_pending <= solo_share * batch_pending)) {
        sq[sqt++] = nxt;
        solo_pending += w;
      } else {
        queue[qt++] = nxt;
        batch_pending += w;
      }
      nxt++;
    }
    /* the solo lane: finish the running request, queue its next chunk, or
     * start the next one */
    if (solo_cur >= 0) {
      request *q = &reqs[solo_cur];
      int32_t idle = 0;
      CHK(msa_query(solo, &idle));
      if (idle && solo_enq < q->niter) {
        const int64_t c = q->niter - solo_enq < SOLO_CHUNK ? q->niter - solo_enq : SOLO_CHUNK;
        CHK(msa_run(solo, (uint64_t)solo_enq, c, 1));
        solo_enq += c;
      } else if (idle) {
        q->finish = now_s() - t_start;
        write_result(out, q, state_dir, solo, NULL, -1);
        free_tables(q);
        solo_pending -= work_of(q);
        done++;
        updates += 64.0 * q->R * (double)q->L * q->L * q->L * (double)q->niter;
        solo_cur = -1;
      }
    }
    if (solo_cur < 0 && sqh < sqt) {
      request *q = &reqs[sq[sqh]];
      const int rd = ready_of(q);
      if (rd < 0) fail("request %d: refresh tables could not be built", q->rid);
      if (rd > 0) {
        solo = solo_ctx_of(sc, nsc, q->L, q->R);
        CHK(msa_init_philox(solo, q->seed));
        CHK(msa_set_refresh(solo, (int32_t)q->nref, q->ctr, q->keys, q->idx));
        q->admit = now_s() - t_start;
        solo_cur = sq[sqh++];
        const int64_t c = q->niter < SOLO_CHUNK ? q->niter : SOLO_CHUNK;
        CHK(msa_run(solo, 0, c, 1));
        solo_enq = c;
      }
    }
    while (qh < qt) {                                /* FIFO admission by footprint */
      request *q = &reqs[queue[qh]];
      int rd = ready_of(q);
      if (rd == 0 && nactive > 0) break;             /* admitted at a later epoch */
      while (rd == 0) {
        usleep(100);
        rd = ready_of(q);
      }
      if (rd < 0) fail("request %d: refresh tables could not be built", q->rid);
      msa_job j = job_of(q);
      int32_t slot;
      const int rc = msa_batch_add(b, &j, &slot);
      if (rc == MSA_ERR_NOMEM) {
        if (nactive == 0) fail("request %d alone exceeds the memory limit", q->rid);
        break;
      }
      CHK(rc);
      free_tables(q);
      if (slot >= nslot_cap) {
        const int old = nslot_cap;
        while (slot >= nslot_cap) nslot_cap *= 2;
        slot_req = (int *)realloc(slot_req, sizeof(int) * (size_t)nslot_cap);
        finishing = (int *)realloc(finishing, sizeof(int) * (size_t)nslot_cap);
        newfin = (int *)realloc(newfin, sizeof(int) * (size_t)nslot_cap);
        tp = (int64_t *)realloc(tp, sizeof(int64_t) * (size_t)nslot_cap);
        te = (int64_t *)realloc(te, sizeof(int64_t) * (size_t)nslot_cap);
        if (!slot_req || !finishing || !newfin || !tp || !te) fail("out of memory");
        for (int s = old; s < nslot_cap; s++) slot_req[s] = -1;
      }
      slot_req[slot] = queue[qh];
      if (slot + 1 > max_slot) max_slot = slot + 1;
      q->admit = now_s() - t_start;
      qh++;
      nactive++;
    }
    if (nactive > max_active) max_active = nactive;
    const int64_t used = msa_batch_bytes_in_use(b);
    if (used > max_bytes_used) max_bytes_used = used;
    int running = 0;
    for (int s = 0; s < max_slot; s++) {
      if (slot_req[s] < 0) continue;
      int fin = 0;
      for (int k = 0; k < nfin; k++) fin |= finishing[k] == s;
      running += !fin;
    }
    if (!running && !nfin) {
      if (solo_cur >= 0 || sqh < sqt) {
        usleep(200);                               /* the solo lane runs */
      } else if (nxt < n) {
        const double wait = reqs[nxt].arrival - (now_s() - t_start);
        if (wait > 0) usleep((useconds_t)(1e6 * (wait < 0.01 ? wait : 0.01)));
      }
      continue;
    }
    /* one epoch stays queued on the device while the previous one's finished
     * requests are measured and retired */
    int nnew = 0;
    if (running) {
      CHK(msa_batch_run(b, epoch));
      CHK(msa_batch_planned(b, max_slot, tp, te));
      for (int s = 0; s < max_slot; s++) {
        if (slot_req[s] < 0) continue;
        int fin = 0;
        for (int k = 0; k < nfin; k++) fin |= finishing[k] == s;
        if (!fin && tp[s] >= te[s]) newfin[nnew++] = s;
      }
      CHK(msa_batch_wait_previous(b));
    } else {
      CHK(msa_batch_sync(b));
    }
    for (int k = 0; k < nfin; k++) {                 /* retire */
      const int s = finishing[k];
      request *q = &reqs[slot_req[s]];
      q->finish = now_s() - t_start;
      write_result(out, q, state_dir, NULL, b, s);
      CHK(msa_batch_remove(b, s));
      batch_pending -= work_of(q);
      slot_req[s] = -1;
      nactive--;
      done++;
      updates += 64.0 * q->R * (double)q->L * q->L * q->L * (double)q->niter;
      const double t = now_s();
      if (done % 1000 == 0 || t - last_note >= 30.0) {
        last_note = t;
        fprintf(stderr, "msa_serve: %d of %d requests finished\n", done, n);
      }
    }
    memcpy(finishing, newfin, sizeof(int) * (size_t)nnew);
    nfin = nnew;
  }
  const double makespan = now_s() - t_start;
  fclose(out);
  pthread_mutex_lock(&p.mu);
  p.stop = 1;
  pthread_cond_broadcast(&p.cv);
  pthread_mutex_unlock(&p.mu);
  for (int w = 0; w < workers; w++) pthread_join(th[w], NULL);
  CHK(msa_batch_destroy(b));
  for (int k = 0; k < nsc; k++) CHK(msa_destroy(sc[k].ctx));

  double *lat = (double *)malloc(sizeof(double) * (size_t)n);
  if (!lat) fail("out of memory");
  for (int i = 0; i < n; i++) lat[i] = reqs[i].finish - reqs[i].arrival;
  const double p50 = percentile(lat, n, 50), p99 = percentile(lat, n, 99);
  printf("{\"requests\": %d, \"makespan_s\": %.6f, \"spin_updates_per_ps\": %.6f, "
         "\"latency_p50_s\": %.6f, \"latency_p99_s\": %.6f, \"max_concurrent\": %d, "
         "\"max_bytes_in_use\": %lld, \"epoch\": %d, \"block\": %d, \"solo_requests\": %d, "
         "\"solo_min_L\": %d, \"solo_min_iter\": %lld, \"solo_share\": %g}\n",
         n, makespan, updates / makespan / 1e12, p50, p99, max_active, (long long)max_bytes_used,
         epoch, block, sqt, solo_min_L, (long long)solo_min_iter, solo_share);
  for (int i = 0; i < n; i++) free(reqs[i].energy);
  free(reqs);
  free(luts);
  free(lat);
  free(p.todo);
  free(th);
  free(slot_req);
  free(queue);
  free(finishing);
  free(newfin);
  free(tp);
  free(te);
  free(sq);
  return 0;
}
