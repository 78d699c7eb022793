// ref_rng_dump.cc -- TEST INFRASTRUCTURE ONLY.
//
// Prints golden streams from the reference's own host generators: this file
// is linked against PROGRAM/src/random_generator.cu compiled where it lies in
// the reference checkout (oracle/Makefile, target `ref`), and it expands the
// reference's own macros _actualiza_aleatorio_HQ_escalar,
// _actualiza_xoshiro256pp (random_generator.h:58-76), HIRANDOM
// (random_generator.h:83-85) and _fill_bits (fill_bits.h:47-68).
// Output goes to stdout; tests/golden/gen_golden.py turns it into fixtures.
//
// usage: ref_rng_dump lux|lux64|cpu|hq|seeds|janus SEED N
//        ref_rng_dump fill SEED0 NREP ITER PARITY JOBID NCALLS
#include "header.h"
#include "fill_bits.h"
#include <stdlib.h>

// Init_luxury/luxury keep their state in random_generator.cu's file globals.
extern void Init_luxury(uint64_t);

int main(int argc, char **argv)
{
  if (argc < 4) { fprintf(stderr, "usage\n"); return 2; }
  const char *what = argv[1];
  unsigned long long seed = strtoull(argv[2], NULL, 0);
  if (!strcmp(what, "fill")) {
    if (argc != 8) { fprintf(stderr, "usage fill\n"); return 2; }
    int nrep = atoi(argv[3]);
    unsigned long long iter0 = strtoull(argv[4], NULL, 0);
    uint64_t parity = strtoull(argv[5], NULL, 0);
    uint32_t mi_jobID = (uint32_t)strtoul(argv[6], NULL, 0);
    int ncalls = atoi(argv[7]);
    s_aleatorio_HQ_64bits random_PRC[NR];
    s_xoshiro256pp random_xoshiro256pp[NR];
    uint32_t keys[NR];
    s_time s_time_and_entropy;
    uint64_t entropy, temporal, iter;
    int ir;
    for (ir = 0; ir < NR; ir++)
      Inicia_generadores_CPU(&random_PRC[ir], &random_xoshiro256pp[ir], &keys[ir], seed + ir);
    for (ir = 0; ir < nrep; ir++) printf("key %d %u\n", ir, keys[ir]);
    for (int k = 0; k < ncalls; k++) {
      iter = iter0 + (uint64_t)k;
      _fill_bits;
      for (int r = 0; r < nrep; r++)
        printf("ctr %d %d %u %u %u %u\n", k, r, s_time_and_entropy.vec[r].x,
               s_time_and_entropy.vec[r].y, s_time_and_entropy.vec[r].z,
               s_time_and_entropy.vec[r].w);
    }
    (void)temporal; (void)entropy;
    return 0;
  }
  int n = atoi(argv[3]);
  if (!strcmp(what, "lux")) {
    Init_luxury(seed);
    for (int i = 0; i < n; i++) printf("%u\n", luxury());
  } else if (!strcmp(what, "lux64")) {
    Init_luxury(seed);
    for (int i = 0; i < n; i++) printf("%llu\n", (unsigned long long)Luescher_PRNG());
  } else if (!strcmp(what, "cpu")) {
    s_aleatorio_HQ_64bits pr;
    s_xoshiro256pp xo;
    uint32_t key;
    Inicia_generadores_CPU(&pr, &xo, &key, seed);
    printf("key %u\n", key);
    for (int i = 0; i < n; i++) {
      _actualiza_aleatorio_HQ_escalar(pr);
      _actualiza_xoshiro256pp(xo);
      printf("%llu %llu\n", (unsigned long long)pr.final, (unsigned long long)xo.final);
    }
  } else if (!strcmp(what, "hq")) {
    s_aleatorio_HQ_64bits pr;
    Init_Rand_HQ_64bits(&pr, seed);
    for (int i = 0; i < n; i++) {
      _actualiza_aleatorio_HQ_escalar(pr);
      printf("%llu\n", (unsigned long long)pr.final);
    }
  } else if (!strcmp(what, "seeds")) {
    uint64_t s = seed;
    uint64_t *list = (uint64_t *)malloc(sizeof(uint64_t) * n);
    generate_seeds_from_one(&s, list, n);
    for (int i = 0; i < n; i++) printf("%llu\n", (unsigned long long)list[i]);
    free(list);
  } else if (!strcmp(what, "janus")) {
    ini_random(seed);
    for (int i = 0; i < n; i++) printf("%llu\n", (unsigned long long)HIRANDOM);
  } else {
    fprintf(stderr, "unknown stream %s\n", what);
    return 2;
  }
  return 0;
}
