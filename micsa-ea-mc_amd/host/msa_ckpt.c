/*
 * msa_ckpt.c -- the reference's checkpoint files (PROGRAM/src/io.c:17-393)
 * written from and read into the sample-per-lane layout:
 *
 *   <rep>/conf.%016llu + hard link <rep>/conf     one (sample, replica)
 *     s_data (184 B) | beta | PR+LCG state (2080 B) | xoshiro state (48 B) |
 *     clone map (1 B) | spins, 1 bit per site lexicographic (V/8 B) | energy
 *   <sample>/walkers.%016llu + hard link walkers  one sample
 *     icds (40 B) | NR PR+LCG states | NR xoshiro states |
 *     demon plane 0 (colour 0, colour 1) | demon plane 1 (colour 0, colour 1)
 *     in the reference's multispin words (64 sites of the sample per word)
 *
 * The layout conversions are 64 x 64 bit transposes: a block of 64 lane words
 * (64 sites, every sample) becomes 64 sample words (64 sites, one sample).
 */
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "msa_host.h"

static int fail(char *buf, int len, const char *fmt, ...)
{
  if (buf && len > 0) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, (size_t)len, fmt, ap);
    va_end(ap);
  }
  return -1;
}

/* --------------------------------------------------- generator images -- */

/* s_aleatorio_HQ_64bits (random_generator.h:29-36): ira[256], PR, zseed,
 * final, ip, ip1, ip2, ip3, 4 padding bytes.  PR is the lagged-sum term of
 * the last draw: final = zseed + PR (random_generator.h:62-67). */
void msa_prng_to_ref(const msa_prng *g, uint8_t *out)
{
  uint64_t w[259];
  memcpy(w, g->ring, sizeof g->ring);
  w[256] = g->out - g->lcg;
  w[257] = g->lcg;
  w[258] = g->out;
  memcpy(out, w, sizeof w);
  const uint8_t ip[8] = {g->p, g->p24, g->p55, g->p61, 0, 0, 0, 0};
  memcpy(out + sizeof w, ip, sizeof ip);
}

void msa_prng_from_ref(const uint8_t *in, msa_prng *g)
{
  uint64_t w[259];
  memcpy(w, in, sizeof w);
  memcpy(g->ring, w, sizeof g->ring);
  g->lcg = w[257];
  g->out = w[258];
  g->p = in[sizeof w];
  g->p24 = in[sizeof w + 1];
  g->p55 = in[sizeof w + 2];
  g->p61 = in[sizeof w + 3];
}

/* s_xoshiro256pp (random_generator.h:39-44): s[4], t, final.  t is a
 * temporary of the update (s[1] << 17 of the previous state), not state:
 * it is written as 0 and ignored on reading. */
void msa_xo256_to_ref(const msa_xo256 *g, uint8_t *out)
{
  const uint64_t w[6] = {g->s[0], g->s[1], g->s[2], g->s[3], 0, g->out};
  memcpy(out, w, sizeof w);
}

void msa_xo256_from_ref(const uint8_t *in, msa_xo256 *g)
{
  uint64_t w[6];
  memcpy(w, in, sizeof w);
  memcpy(g->s, w, sizeof g->s);
  g->out = w[5];
}

/* ------------------------------------------------------------- sizes -- */

#define SDATA_BYTES 184
#define ICDS_BYTES 40

/* write_conf's predicted_size, io.c:22-25 (NBETAS = 1) */
uint64_t msa_conf_bytes(int L)
{
  const uint64_t V = (uint64_t)L * L * L;
  return SDATA_BYTES + sizeof(double) + MSA_PRNG_REF_BYTES + MSA_XO_REF_BYTES + 1 + V / 8 +
         sizeof(int32_t);
}

/* multispin demon words of one sample: 2 planes x 2 colours x nc x HALF x R */
uint64_t msa_msc_words(int L, int R, int nc)
{
  const uint64_t V = (uint64_t)L * L * L;
  return 2 * (V / 64) * (uint64_t)R * (uint64_t)nc;
}

/* write_walkers' predicted_size, io.c:260-262 */
uint64_t msa_walkers_bytes(int L, int R, int nc)
{
  return ICDS_BYTES + (uint64_t)R * (MSA_PRNG_REF_BYTES + MSA_XO_REF_BYTES) +
         sizeof(uint64_t) * msa_msc_words(L, R, nc);
}

static void sdata_image(const msa_sdata *d, uint8_t *out)
{
  uint8_t *p = out;
  memcpy(p, &d->maxtime, 8);
  memcpy(p + 8, &d->now, 8);
  const int32_t iv[6] = {d->ntimes, d->nterm, d->nc, d->nr, d->l, d->flag};
  memcpy(p + 16, iv, sizeof iv);
  uint64_t sv[2 + MSA_MAXR];
  sv[0] = d->seed_J;
  sv[1] = d->seed_u;
  memcpy(sv + 2, d->seed_MC, sizeof d->seed_MC);
  memcpy(p + 40, sv, sizeof sv);
}

static void sdata_parse(const uint8_t *p, msa_sdata *d)
{
  memcpy(&d->maxtime, p, 8);
  memcpy(&d->now, p + 8, 8);
  int32_t iv[6];
  memcpy(iv, p + 16, sizeof iv);
  d->ntimes = iv[0];
  d->nterm = iv[1];
  d->nc = iv[2];
  d->nr = iv[3];
  d->l = iv[4];
  d->flag = iv[5];
  uint64_t sv[2 + MSA_MAXR];
  memcpy(sv, p + 40, sizeof sv);
  d->seed_J = sv[0];
  d->seed_u = sv[1];
  memcpy(d->seed_MC, sv + 2, sizeof d->seed_MC);
}

/* ------------------------------------------------------- file commit -- */

/* Write `n` bytes to <dir>/<base>.%016llu and make <dir>/<base> a hard link
 * to it, then check the link count and size (io.c:64-85, 290-311). */
static int commit_file(const char *dir, const char *base, uint64_t tag, const uint8_t *buf,
                       uint64_t n, char *e, int elen)
{
  char name_t[MSA_MAXSTR + 64], name[MSA_MAXSTR + 64];
  snprintf(name_t, sizeof name_t, "%s/%s.%016llu", dir, base, (unsigned long long)tag);
  snprintf(name, sizeof name, "%s/%s", dir, base);
  FILE *f = fopen(name_t, "wb");
  if (!f) return fail(e, elen, "Problems opening %s", name_t);
  const size_t w = fwrite(buf, 1, (size_t)n, f);
  if (fclose(f) != 0 || w != (size_t)n) return fail(e, elen, "Problems writing %s", name_t);
  remove(name);
  if (link(name_t, name)) return fail(e, elen, "Problems in hard-link to %s", name_t);
  struct stat st;
  if (stat(name, &st)) return fail(e, elen, "Problems in stat of %s", name);
  if (st.st_nlink != 2) return fail(e, elen, "The number of hard-links of %s is not 2", name);
  if ((uint64_t)st.st_size != n)
    return fail(e, elen, "There are only %lld bytes in %s (must be %llu)", (long long)st.st_size,
                name, (unsigned long long)n);
  return 0;
}

/* Read exactly `n` bytes of `path` after checking its size (io.c:121-125, 342-348). */
static uint8_t *load_file(const char *path, uint64_t n, char *e, int elen)
{
  struct stat st;
  if (stat(path, &st)) {
    fail(e, elen, "There is no file %s", path);
    return NULL;
  }
  if ((uint64_t)st.st_size != n) {
    fail(e, elen, "There are only %lld bytes in %s (must be %llu)", (long long)st.st_size, path,
         (unsigned long long)n);
    return NULL;
  }
  uint8_t *buf = (uint8_t *)malloc((size_t)n);
  FILE *f = fopen(path, "rb");
  if (!buf || !f) {
    free(buf);
    if (f) fclose(f);
    fail(e, elen, "Problems opening %s", path);
    return NULL;
  }
  const size_t r = fread(buf, 1, (size_t)n, f);
  fclose(f);
  if (r != (size_t)n) {
    free(buf);
    fail(e, elen, "Problems reading %s", path);
    return NULL;
  }
  return buf;
}

/* ------------------------------------------------------------- confs -- */

/* write_conf, io.c:17-96, for one (sample, replica): the clone map of the
 * single temperature is 0 and the energy is 3V - 2 * (unsatisfied bonds). */
int msa_conf_write(const char *rep_dir, uint64_t now, const msa_sdata *d, double beta,
                   const msa_prng *pr, const msa_xo256 *xo, int L, const uint8_t *janus,
                   int32_t energy, char *e, int elen)
{
  const uint64_t n = msa_conf_bytes(L), V8 = (uint64_t)L * L * L / 8;
  uint8_t *buf = (uint8_t *)malloc((size_t)n);
  if (!buf) return fail(e, elen, "Problems allocating %llu bytes", (unsigned long long)n);
  uint8_t *p = buf;
  sdata_image(d, p);
  p += SDATA_BYTES;
  memcpy(p, &beta, sizeof beta);
  p += sizeof beta;
  msa_prng_to_ref(pr, p);
  p += MSA_PRNG_REF_BYTES;
  msa_xo256_to_ref(xo, p);
  p += MSA_XO_REF_BYTES;
  *p++ = 0;
  memcpy(p, janus, (size_t)V8);
  p += V8;
  memcpy(p, &energy, sizeof energy);
  const int rc = commit_file(rep_dir, "conf", now, buf, n, e, elen);
  free(buf);
  return rc;
}

/* read_conf's parse of one file, io.c:110-196 (the caller checks the fields
 * against its run and the energy against the spins). */
int msa_conf_read(const char *path, int L, msa_sdata *d, double *beta, msa_prng *pr,
                  msa_xo256 *xo, uint8_t *janus, int32_t *energy, char *e, int elen)
{
  const uint64_t n = msa_conf_bytes(L), V8 = (uint64_t)L * L * L / 8;
  uint8_t *buf = load_file(path, n, e, elen);
  if (!buf) return -1;
  const uint8_t *p = buf;
  sdata_parse(p, d);
  p += SDATA_BYTES;
  memcpy(beta, p, sizeof *beta);
  p += sizeof *beta;
  msa_prng_from_ref(p, pr);
  p += MSA_PRNG_REF_BYTES;
  msa_xo256_from_ref(p, xo);
  p += MSA_XO_REF_BYTES;
  const uint8_t clone = *p++;
  if (clone != 0) {
    free(buf);
    return fail(e, elen, "There is an error in the permutations in %s", path);
  }
  memcpy(janus, p, (size_t)V8);
  p += V8;
  memcpy(energy, p, sizeof *energy);
  free(buf);
  return 0;
}

/* ----------------------------------------------------------- walkers -- */

/* write_walkers, io.c:254-323: the demon words w (msa_msc_words) of one
 * sample; the file of the previous measurement (prev != 0) is removed. */
int msa_walkers_write(const char *sample_dir, uint64_t now, uint64_t prev, const msa_icds *ic,
                      int L, int R, int nc, const msa_prng *pr, const msa_xo256 *xo,
                      const uint64_t *w, char *e, int elen)
{
  const uint64_t n = msa_walkers_bytes(L, R, nc), nw = msa_msc_words(L, R, nc);
  uint8_t *buf = (uint8_t *)malloc((size_t)n);
  if (!buf) return fail(e, elen, "Problems allocating %llu bytes", (unsigned long long)n);
  uint8_t *p = buf;
  const uint32_t iv[6] = {ic->first, ic->isample, ic->nrep, ic->walkers_per_site, ic->lv, ic->t};
  memcpy(p, iv, sizeof iv);
  memcpy(p + 24, &ic->seed, 8);
  memcpy(p + 32, &ic->beta, 8);
  p += ICDS_BYTES;
  for (int r = 0; r < R; r++, p += MSA_PRNG_REF_BYTES) msa_prng_to_ref(&pr[r], p);
  for (int r = 0; r < R; r++, p += MSA_XO_REF_BYTES) msa_xo256_to_ref(&xo[r], p);
  memcpy(p, w, (size_t)(nw * sizeof(uint64_t)));
  int rc = commit_file(sample_dir, "walkers", now, buf, n, e, elen);
  free(buf);
  if (rc == 0 && prev) {
    char old[MSA_MAXSTR + 64];
    snprintf(old, sizeof old, "%s/walkers.%016llu", sample_dir, (unsigned long long)prev);
    remove(old);
  }
  return rc;
}

/* read_walkers' parse of one file, io.c:325-393 */
int msa_walkers_read(const char *path, int L, int R, int nc, msa_icds *ic, msa_prng *pr,
                     msa_xo256 *xo, uint64_t *w, char *e, int elen)
{
  const uint64_t n = msa_walkers_bytes(L, R, nc), nw = msa_msc_words(L, R, nc);
  uint8_t *buf = load_file(path, n, e, elen);
  if (!buf) return -1;
  const uint8_t *p = buf;
  uint32_t iv[6];
  memcpy(iv, p, sizeof iv);
  ic->first = iv[0];
  ic->isample = iv[1];
  ic->nrep = iv[2];
  ic->walkers_per_site = iv[3];
  ic->lv = iv[4];
  ic->t = iv[5];
  memcpy(&ic->seed, p + 24, 8);
  memcpy(&ic->beta, p + 32, 8);
  p += ICDS_BYTES;
  for (int r = 0; r < R; r++, p += MSA_PRNG_REF_BYTES) msa_prng_from_ref(p, &pr[r]);
  for (int r = 0; r < R; r++, p += MSA_XO_REF_BYTES) msa_xo256_from_ref(p, &xo[r]);
  memcpy(w, p, (size_t)(nw * sizeof(uint64_t)));
  free(buf);
  return 0;
}

/* ------------------------------------------------- layout conversion -- */

/* In-place transpose of a 64 x 64 bit matrix, a[i] bit j <-> a[j] bit i:
 * six rounds that swap the off-diagonal blocks of side 32, 16, ..., 1. */
static void transpose64(uint64_t a[64])
{
  uint64_t m = 0x00000000FFFFFFFFULL;
  for (int j = 32; j; j >>= 1, m ^= m << j)
    for (int k = 0; k < 64; k = ((k | j) + 1) & ~j) {
      const uint64_t t = ((a[k] >> j) ^ a[k | j]) & m;
      a[k] ^= t << j;
      a[k | j] ^= t;
    }
}

/* Janus_packing_for_write (io.c:212-231): spin of site s = x + L(y + Lz) at
 * bit s&7 of byte s>>3.  janus[(i*R + r) * V/8 ...] for listed lane lanes[i].
 * Needs V % 64 == 0 (L a multiple of 4). */
void msa_janus_from_lanes(int L, int R, const uint64_t *spin, int n, const int *lanes,
                          uint8_t *janus)
{
  const uint64_t V = (uint64_t)L * L * L, V8 = V / 8, nb = V / 64;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < (int64_t)(nb * R); q++) {
    const int r = (int)(q / (int64_t)nb);
    const uint64_t blk = (uint64_t)q % nb;
    uint64_t a[64];
    memcpy(a, spin + (uint64_t)r * V + 64 * blk, sizeof a);
    transpose64(a);
    for (int i = 0; i < n; i++)
      memcpy(janus + ((uint64_t)i * R + r) * V8 + 8 * blk, &a[lanes[i]], 8);
  }
}

/* Janus_unpacking_for_read (io.c:233-250) into the listed lanes of `spin`;
 * the other lanes keep their bits. */
void msa_janus_to_lanes(int L, int R, const uint8_t *janus, int n, const int *lanes,
                        uint64_t *spin)
{
  const uint64_t V = (uint64_t)L * L * L, V8 = V / 8, nb = V / 64;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < (int64_t)(nb * R); q++) {
    const int r = (int)(q / (int64_t)nb);
    const uint64_t blk = (uint64_t)q % nb;
    uint64_t a[64];
    uint64_t *dst = spin + (uint64_t)r * V + 64 * blk;
    memcpy(a, dst, sizeof a);
    transpose64(a);
    for (int i = 0; i < n; i++)
      memcpy(&a[lanes[i]], janus + ((uint64_t)i * R + r) * V8 + 8 * blk, 8);
    transpose64(a);
    memcpy(dst, a, sizeof a);
  }
}

/* The multispin placement of packing_u (ini.c:301-345): with ML = L/4, site
 * (x, y, z) = (a + ML bx, e + ML by, g + ML bz) goes to word a + ML e + ML^2 g
 * of the sample, bit bx + 4 by + 16 bz; the word's colour is (a + e + g) & 1
 * and its half index word / 2.  Demon words of layer c, replica r are at
 * c*HALF*R + half*R + r of their colour (main.cu:868-872), and the file holds
 * plane 0 (colour 0, colour 1) then plane 1 (colour 0, colour 1). */
static void msc_block_sites(int L, int ML, int a, int e, int g, uint64_t site[64])
{
  for (int k = 0; k < 64; k++) {
    const int bx = k & 3, by = (k >> 2) & 3, bz = k >> 4;
    site[k] = (uint64_t)(a + ML * bx) + (uint64_t)L * ((uint64_t)(e + ML * by) +
                                                        (uint64_t)L * (uint64_t)(g + ML * bz));
  }
}

void msa_msc_from_lanes(int L, int R, int nc, const uint64_t *d0, const uint64_t *d1, int n,
                        const int *lanes, uint64_t *w)
{
  const int ML = L / 4;
  const uint64_t V = (uint64_t)L * L * L, MV = V / 64, HALF = MV / 2;
  const uint64_t col_words = (uint64_t)nc * HALF * R, nw = msa_msc_words(L, R, nc);
#pragma omp parallel for schedule(static)
  for (int64_t ms = 0; ms < (int64_t)MV; ms++) {
    const int a = (int)(ms % ML), e = (int)((ms / ML) % ML), g = (int)(ms / ((int64_t)ML * ML));
    const int col = (a + e + g) & 1;
    const uint64_t half = (uint64_t)ms / 2;
    uint64_t site[64], blk[64];
    msc_block_sites(L, ML, a, e, g, site);
    for (int plane = 0; plane < 2; plane++) {
      const uint64_t *src = plane ? d1 : d0;
      for (int r = 0; r < R; r++)
        for (int c = 0; c < nc; c++) {
          const uint64_t *sl = src + ((uint64_t)r * nc + c) * V;
          for (int k = 0; k < 64; k++) blk[k] = sl[site[k]];
          transpose64(blk);
          const uint64_t off = (uint64_t)(2 * plane + col) * col_words +
                               (uint64_t)c * HALF * R + half * R + (uint64_t)r;
          for (int i = 0; i < n; i++) w[(uint64_t)i * nw + off] = blk[lanes[i]];
        }
    }
  }
}

void msa_msc_to_lanes(int L, int R, int nc, const uint64_t *w, int n, const int *lanes,
                      uint64_t *d0, uint64_t *d1)
{
  const int ML = L / 4;
  const uint64_t V = (uint64_t)L * L * L, MV = V / 64, HALF = MV / 2;
  const uint64_t col_words = (uint64_t)nc * HALF * R, nw = msa_msc_words(L, R, nc);
#pragma omp parallel for schedule(static)
  for (int64_t ms = 0; ms < (int64_t)MV; ms++) {
    const int a = (int)(ms % ML), e = (int)((ms / ML) % ML), g = (int)(ms / ((int64_t)ML * ML));
    const int col = (a + e + g) & 1;
    const uint64_t half = (uint64_t)ms / 2;
    uint64_t site[64], blk[64];
    msc_block_sites(L, ML, a, e, g, site);
    for (int plane = 0; plane < 2; plane++) {
      uint64_t *dst = plane ? d1 : d0;
      for (int r = 0; r < R; r++)
        for (int c = 0; c < nc; c++) {
          uint64_t *sl = dst + ((uint64_t)r * nc + c) * V;
          for (int k = 0; k < 64; k++) blk[k] = sl[site[k]];
          transpose64(blk);
          const uint64_t off = (uint64_t)(2 * plane + col) * col_words +
                               (uint64_t)c * HALF * R + half * R + (uint64_t)r;
          for (int i = 0; i < n; i++) blk[lanes[i]] = w[(uint64_t)i * nw + off];
          transpose64(blk);
          for (int k = 0; k < 64; k++) sl[site[k]] = blk[k];
        }
    }
  }
}

/* ------------------------------------------------- per-lane integrity -- */

/* Bit-sliced per-lane counters: plane i holds bit i of each lane's count. */
static inline void vadd(uint64_t *pl, uint64_t x)
{
  for (int i = 0; x && i < 40; i++) {
    const uint64_t c = pl[i] & x;
    pl[i] ^= x;
    x = c;
  }
}

static void vcounts(const uint64_t *pl, int64_t *cnt)
{
  for (int b = 0; b < 64; b++) {
    int64_t v = 0;
    for (int i = 0; i < 40; i++) v |= (int64_t)((pl[i] >> b) & 1) << i;
    cnt[b] = v;
  }
}

/* calculate_scalar_energy (med.c:85-113) for every lane and replica:
 * e[r*64 + b] = 3V - 2 * #(s ^ J ^ s_neighbour) over the +x, +y, +z bonds. */
void msa_lane_energy(int L, int R, const uint64_t *J, const uint64_t *spin, int32_t *e)
{
  const uint64_t V = (uint64_t)L * L * L;
  for (int r = 0; r < R; r++) {
    const uint64_t *S = spin + (uint64_t)r * V;
    int64_t tot[64] = {0};
#pragma omp parallel
    {
      uint64_t pl[40] = {0};
#pragma omp for schedule(static)
      for (int z = 0; z < L; z++)
        for (int y = 0; y < L; y++)
          for (int x = 0; x < L; x++) {
            const uint64_t s = (uint64_t)x + (uint64_t)L * ((uint64_t)y + (uint64_t)L * z);
            const uint64_t sx = s - x + (uint64_t)((x + 1) % L);
            const uint64_t sy = s + (uint64_t)L * ((uint64_t)((y + 1) % L) - (uint64_t)y);
            const uint64_t sz = s + (uint64_t)L * L * ((uint64_t)((z + 1) % L) - (uint64_t)z);
            vadd(pl, S[s] ^ J[s] ^ S[sx]);
            vadd(pl, S[s] ^ J[V + s] ^ S[sy]);
            vadd(pl, S[s] ^ J[2 * V + s] ^ S[sz]);
          }
      int64_t cnt[64];
      vcounts(pl, cnt);
#pragma omp critical
      for (int b = 0; b < 64; b++) tot[b] += cnt[b];
    }
    for (int b = 0; b < 64; b++) e[r * 64 + b] = (int32_t)(3 * (int64_t)V - 2 * tot[b]);
  }
}

/* countJ of Init_J (ini.c:457-511): per lane, the bonds whose J bit is 1. */
void msa_lane_countJ(int L, const uint64_t *J, int32_t *cnt)
{
  const uint64_t n = 3 * (uint64_t)L * L * L;
  int64_t tot[64] = {0};
#pragma omp parallel
  {
    uint64_t pl[40] = {0};
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) vadd(pl, J[i]);
    int64_t c[64];
    vcounts(pl, c);
#pragma omp critical
    for (int b = 0; b < 64; b++) tot[b] += c[b];
  }
  for (int b = 0; b < 64; b++) cnt[b] = (int32_t)tot[b];
}
