/*
 * msa_io.c -- the reference's input formats (PROGRAM/src/io.c:464-655) and
 * the replica statistics of Measures (PROGRAM/src/med.c:13-51).
 */
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "msa_host.h"

static int err(char *buf, int len, const char *fmt, ...)
{
  if (buf && len > 0) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, (size_t)len, fmt, ap);
    va_end(ap);
  }
  return -1;
}

/* read_input (io.c:464-502): line 1 the output directory, then one value per
 * line: maxtime, now (%llu); ntimes, nterm, nc, nr, L, flag (%u); seed_J,
 * seed_u, seed_MC[0..nr-1] (%llu).  Only the first token of a line counts,
 * so lines may carry comments after the value. */
int msa_read_input(const char *path, msa_input *in, char *e, int elen)
{
  FILE *f = fopen(path, "r");
  if (!f) return err(e, elen, "Error: No file %s", path);
  char line[MSA_MAXSTR + 1];
  memset(in, 0, sizeof(*in));
  int ok = 1;
  if (!fgets(line, MSA_MAXSTR, f) || sscanf(line, "%1023s", in->dir) != 1) ok = 0;
  unsigned long long *ull[2] = {&in->maxtime, &in->now};
  for (int j = 0; ok && j < 2; j++)
    if (!fgets(line, MSA_MAXSTR, f) || sscanf(line, "%llu", ull[j]) != 1) ok = 0;
  int *iv[6] = {&in->ntimes, &in->nterm, &in->nc, &in->nr, &in->l, &in->flag};
  for (int j = 0; ok && j < 6; j++)
    if (!fgets(line, MSA_MAXSTR, f) || sscanf(line, "%d", iv[j]) != 1) ok = 0;
  if (ok && (in->nr < 1 || in->nr > MSA_MAXR)) {
    fclose(f);
    return err(e, elen, "d.nr = %d outside 1..%d", in->nr, MSA_MAXR);
  }
  unsigned long long v;
  if (ok && fgets(line, MSA_MAXSTR, f) && sscanf(line, "%llu", &v) == 1) in->seed_J = v; else ok = 0;
  if (ok && fgets(line, MSA_MAXSTR, f) && sscanf(line, "%llu", &v) == 1) in->seed_u = v; else ok = 0;
  for (int r = 0; ok && r < in->nr; r++) {
    if (fgets(line, MSA_MAXSTR, f) && sscanf(line, "%llu", &v) == 1) in->seed_MC[r] = v;
    else ok = 0;
  }
  fclose(f);
  if (!ok) return err(e, elen, "Bad data input in %s", path);
  /* check_data, io.c:504-516 */
  if (in->nterm < 0 || in->nc < 0 || in->ntimes < 0) return err(e, elen, "Bad data input");
  return 0;
}

/* print_data, io.c:518-532 */
void msa_print_input(FILE *f, const msa_input *d)
{
  fprintf(f, "maxtime      %16llu\n", d->maxtime);
  fprintf(f, "now          %16llu\n", d->now);
  fprintf(f, "ntimes       %16d\n", d->ntimes);
  fprintf(f, "nterm        %16d\n", d->nterm);
  fprintf(f, "nr           %16d\n", d->nr);
  fprintf(f, "nc           %16d\n", d->nc);
  fprintf(f, "L            %16d\n", d->l & 0xffff);
  fprintf(f, "M            %16d\n", d->l >> 16);
  fprintf(f, "flag         %16d\n", d->flag);
  fprintf(f, "seed_J       %016llx\n", (unsigned long long)d->seed_J);
  fprintf(f, "seed_u       %016llx\n", (unsigned long long)d->seed_u);
  for (int r = 0; r < d->nr; r++)
    fprintf(f, "seed_MC[%d]   %016llx\n", r, (unsigned long long)d->seed_MC[r]);
}

/* read_betas with NBETAS = 1 (io.c:534-558) */
int msa_read_betas(const char *path, double *beta, char *e, int elen)
{
  FILE *f = fopen(path, "rt");
  if (!f) return err(e, elen, "File %s doesn't exist.", path);
  int n = fscanf(f, "%lf", beta);
  fclose(f);
  if (n != 1) return err(e, elen, "File %s only has 0 lines (it needs 1)", path);
  return 0;
}

/* read_lut (io.c:561-591): a double beta that must match beta.in within
 * 1e-14, then the table. */
int msa_read_lut_file(const char *path, double beta, msa_lut *lut, char *e, int elen)
{
  FILE *f = fopen(path, "rb");
  if (!f) return err(e, elen, "Problems opening %s", path);
  double b = 0;
  if (fread(&b, sizeof b, 1, f) != 1) { fclose(f); return err(e, elen, "Problems reading betas"); }
  if (fabs(beta - b) > 1e-14) {
    fclose(f);
    return err(e, elen, "betas[0]=%.14g from beta file, betas[0]=%.14g from %s", beta, b, path);
  }
  if (fread(lut, sizeof(*lut), 1, f) != 1) { fclose(f); return err(e, elen, "Problems reading lut_heat_bath"); }
  fclose(f);
  return 0;
}

/* read_list_times (io.c:626-655): n increasing times */
int msa_read_times(const char *path, int n, unsigned long long *t, char *e, int elen)
{
  FILE *f = fopen(path, "rt");
  if (!f) return err(e, elen, "File %s doesn't exist.", path);
  for (int i = 0; i < n; i++)
    if (fscanf(f, "%llu ", &t[i]) != 1) {
      fclose(f);
      return err(e, elen, "File %s only has %d lines (it needs %d)", path, i, n);
    }
  fclose(f);
  for (int i = 0; i + 1 < n; i++)
    if (t[i] >= t[i + 1]) return err(e, elen, "List of times must be sorted!");
  return 0;
}

/* The run-time change of a run's end and measurement list (stop, io.c:930-968).
 * The reference's reader uses an uninitialised pointer for the first value
 * and then reads that value's count of times, so it has no defined
 * behaviour to keep; this one reads the first value as the new maxtime and
 * every further value as the new list of measurement times, which must be
 * strictly increasing.  Returns 0 when the file does not exist, 1 when read
 * (*times = a malloc'd list of *ntimes values, or NULL with *ntimes = 0 when
 * the file holds the maxtime alone), 2 when it holds no number (it is
 * removed, as the reference removes it), -1 on a bad value or an unsorted
 * list. */
int msa_read_changentimes(const char *path, unsigned long long *maxtime, unsigned long long **times,
                          int *ntimes, char *e, int elen)
{
  *times = NULL;
  *ntimes = 0;
  FILE *f = fopen(path, "r");
  if (!f) return 0;
  unsigned long long v;
  if (fscanf(f, "%llu", &v) != 1) {
    fclose(f);
    remove(path);
    return 2;
  }
  *maxtime = v;
  int cap = 0, n = 0;
  unsigned long long *t = NULL;
  while (fscanf(f, "%llu", &v) == 1) {
    if (n == cap) {
      cap = cap ? 2 * cap : 64;
      unsigned long long *nt = (unsigned long long *)realloc(t, sizeof *t * (size_t)cap);
      if (!nt) {
        free(t);
        fclose(f);
        return err(e, elen, "Problems allocating memory for list_times");
      }
      t = nt;
    }
    t[n++] = v;
  }
  const int bad = !feof(f);
  fclose(f);
  if (bad) {
    free(t);
    return err(e, elen, "File %s: value %d is not a number", path, n + 2);
  }
  for (int i = 0; i + 1 < n; i++)
    if (t[i] >= t[i + 1]) {
      free(t);
      return err(e, elen, "List of times %s is not sorted.\nList of times must be sorted!", path);
    }
  *times = t;
  *ntimes = n;
  return 1;
}

/* read_list_samples (io.c:593-624): empty path = 0..n-1, else n increasing ids */
int msa_read_samples(const char *path, int n, int *s, char *e, int elen)
{
  if (!path || !path[0]) {
    for (int i = 0; i < n; i++) s[i] = i;
    return 0;
  }
  FILE *f = fopen(path, "rt");
  if (!f) return err(e, elen, "File %s doesn't exist.", path);
  for (int i = 0; i < n; i++)
    if (fscanf(f, "%d ", &s[i]) != 1) {
      fclose(f);
      return err(e, elen, "File %s only has %d lines (it needs %d)", path, i, n);
    }
  fclose(f);
  for (int i = 0; i + 1 < n; i++)
    if (s[i] >= s[i + 1]) return err(e, elen, "List of samples must be sorted!");
  return 0;
}

/* schedule.in: T0, T1, nsweeps, refresh_every -- one value per line */
int msa_read_schedule(const char *path, msa_sched *s, char *e, int elen)
{
  FILE *f = fopen(path, "rt");
  if (!f) return err(e, elen, "File %s doesn't exist.", path);
  long long n = 0, every = 0;
  int ok = fscanf(f, "%lf %lf %lld %lld", &s->T0, &s->T1, &n, &every) == 4;
  fclose(f);
  if (!ok) return err(e, elen, "%s needs T0 T1 nsweeps refresh_every", path);
  const int fixed = s->T0 == 0 && s->T1 == 0;
  if ((!fixed && (s->T0 <= 0 || s->T1 <= 0)) || n <= 0 || (every <= 0 && every != -1))
    return err(e, elen, "bad schedule in %s", path);
  s->nsweeps = n;
  s->refresh_every = every;
  s->beta_fixed = 0;
  return 0;
}

void msa_mean_err(const double *x, int n, double *mean, double *errv)
{
  double a = 0, a2 = 0;
  for (int i = 0; i < n; i++) {
    a += x[i];
    a2 += x[i] * x[i];
  }
  a /= n;
  a2 /= n;
  *mean = a;
  *errv = sqrt((a2 - a * a) / (double)(n - 1));
}
