/*
 * msa.h -- C ABI of libmsa_hip.so, the device layer of the micSA path on
 * MI355X (gfx950).  Plain pointers and sizes, no HIP or C++ types.
 *
 * This library replaces the CUDA-aware host functions of the reference
 * (PROGRAM/src/main.cu, prototypes in PROGRAM/include/header.h:329-340).
 * Each entry point names the reference function(s) it stands in for.
 *
 * Data model (sample-per-lane): every uint64 word holds one site of 64
 * independent samples, bit b = sample b.  Host-side arrays use the canonical
 * site order s = x + L*(y + L*z):
 *   J     [3][V]        bond from s to s+e_d, d = x,y,z (one bit per sample)
 *   spin  [R][V]        R replicas share J
 *   d0,d1 [R][nc][V]    the two bit planes of the 2-bit demons (D = d0 + 2 d1)
 * On the device the same words are stored colour-split: the sites of colour
 * c = (x+y+z)&1 packed as h = x/2 + (L/2)(y + L z), one plane per colour.
 *
 * Every call returns MSA_OK (0) or a negative msa_status; msa_strerror()
 * describes the last error of the calling thread.  A context is not
 * thread-safe: one host thread drives it, as one thread drives the
 * reference's device (main.cu:137-180).
 */
#ifndef MSA_H
#define MSA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum msa_status {
  MSA_OK = 0,
  MSA_ERR_ARG = -1,     /* invalid argument / geometry */
  MSA_ERR_HIP = -2,     /* HIP runtime error */
  MSA_ERR_NOMEM = -3,   /* device allocation failed */
  MSA_ERR_STATE = -4,   /* call out of order (e.g. run before tables are set) */
  MSA_ERR_RANGE = -5    /* refresh or LUT index outside the uploaded tables */
};

typedef struct msa_ctx msa_ctx;

typedef struct {
  int32_t L;    /* linear size, even, 4..1024 (reference: compile-time L, header.h:38-51) */
  int32_t R;    /* replicas sharing J, 1..16 (reference: NR = 16, header.h:17) */
  int32_t nc;   /* demon layers, 1..6 (reference: data.nc, input.in:6, main.cu:1082) */
} msa_geom;

typedef struct { uint32_t x, y, z, w; } msa_u32x4;   /* Philox counter (CUDA uint4) */

/* Heat-bath table for eight iid Bernoulli(q) bits per byte, byte-identical to
 * s_lut_heat_bath with NUMBITSPREBUSQUEDAS = 10 (header.h:113-116):
 * 256 x (msb32, lsb32) cumulative thresholds, then 1024 x uint16 presearch
 * entries (min | max << 8).  This is the payload of the reference LUT file
 * after its leading double (io.c:561-591). */
typedef struct {
  uint32_t umbrales[512];
  uint32_t prebusqueda[512];
} msa_lut;

/* Allocate device state for one lattice on `device` and create the
 * context's stream.  Replaces CopyConstOnDevice (main.cu:379-402) and the
 * allocations of CopyDataOnDevice / mmcuda (main.cu:404-471, 1209-1289). */
int msa_create(int device, const msa_geom *geom, msa_ctx **out);

/* Release everything.  Replaces FreeDevice (main.cu:508-530). */
int msa_destroy(msa_ctx *ctx);

/* Upload couplings, spins and demons (canonical order, see above); d0/d1 may
 * be NULL for all-zero demons.  Replaces the uploads of CopyDataOnDevice
 * (main.cu:404-471) and of refresh_walkers (main.cu:355-362). */
int msa_upload_state(msa_ctx *ctx, const uint64_t *J, const uint64_t *spin,
                     const uint64_t *d0, const uint64_t *d1);

/* Download spins and demons (any pointer may be NULL) and couplings.
 * Replaces copy_from_GPU_for_write (main.cu:473-490). */
int msa_download_state(msa_ctx *ctx, uint64_t *spin, uint64_t *d0, uint64_t *d1);
int msa_download_couplings(msa_ctx *ctx, uint64_t *J);

/* msa_download_state without waiting: the arrays receive the state as the
 * work queued so far leaves it, and hold it once msa_download_wait returns;
 * work queued after this call (msa_run ...) runs meanwhile.  The context keeps
 * a device staging buffer of the requested arrays' size (8 R V bytes for the
 * spins, 16 R nc V for the demons) until msa_destroy; a call waits for the
 * copy of the call before it.  The copy overlaps the device only into
 * page-locked arrays (msa_host_alloc).  The driver uses it for its
 * checkpoints (copy_from_GPU_for_write, main.cu:473-490, waits). */
int msa_download_state_async(msa_ctx *ctx, uint64_t *spin, uint64_t *d0, uint64_t *d1);
int msa_download_wait(msa_ctx *ctx);

/* Counter-based construction on the device: J word (d,s) and spin word (r,s)
 * are Philox4x32-10 outputs of counter {s, d|r, 1|2, 0} and key
 * {seed_lo, seed_hi}; demons are cleared.  No reference counterpart: the
 * reference constructs state on the host (ini.c:275-299, 425-519); this is
 * the construction used by the batched and served forms. */
int msa_init_philox(msa_ctx *ctx, uint64_t seed);

/* Upload `nlut` pairs of heat-bath tables (luts[2*i] for demon bit plane 0,
 * luts[2*i+1] for plane 1).  Replaces the LUT upload of CopyConstOnDevice
 * (main.cu:396-398), generalised to one pair per temperature step. */
int msa_set_luts(msa_ctx *ctx, int32_t nlut, const msa_lut *luts);

/* Upload the refresh schedule: for refresh k (0 <= k < nref) and replica r,
 * the Philox counter ctr[k*R + r] (built like _fill_bits, fill_bits.h:47-68),
 * the per-replica keys key[r] (seed_keys, random_generator.cu:150-156) and
 * the LUT pair index lut_idx[k].  No single reference counterpart: the
 * reference refreshes demons on host threads (refresh_walkers main.cu:340-377,
 * InitWalkers ini.c:712-781); here the refresh runs on the device, fused
 * into the first sweep that uses each demon layer. */
int msa_set_refresh(msa_ctx *ctx, int32_t nref, const msa_u32x4 *ctr, const uint32_t *key,
                    const int32_t *lut_idx);

/* Run iterations t0 .. t0+niter-1.  One iteration = optional demon refresh
 * of every layer, nc sweeps (colour 0 then colour 1, sweep ic using layer ic)
 * and the demon walk (main.cu:137-150, Monte_Carlo main.cu:254-306 with
 * flag = 1).  Refresh cadence:
 *   refresh_every  > 0 : refresh when t % refresh_every == 0, k = t / refresh_every
 *   refresh_every == 0 : never
 *   refresh_every == -1: at t = 0 and powers of two (main.cu:140-142),
 *                        k = 0 for t = 0, k = log2(t) + 1 otherwise
 * Asynchronous on the context stream.  Replaces the Monte_Carlo launches. */
int msa_run(msa_ctx *ctx, uint64_t t0, int64_t niter, int64_t refresh_every);

/* Same as msa_run, then wait and report the elapsed device time of the
 * launches (HIP events on the context stream) in *ms. */
int msa_run_timed(msa_ctx *ctx, uint64_t t0, int64_t niter, int64_t refresh_every, float *ms);

/* Canonical Metropolis thermalization: niter iterations of Monte_Carlo with
 * flag = 0 (main.cu:202-252, the thermal loop main.cu:124-128).  Iteration t
 * generates the random words of parity 0 (d_gen_rndbits_cuda main.cu:647-702,
 * counter ctr[(t*2 + 0)*R + r]) and updates colour 0 with the Metropolis
 * branch (main.cu:945-958), then the same for parity 1 / colour 1.  `lut` is
 * the heat-bath table of the thermal beta (dev_lut_heat_bath[0]); key[R] as
 * for msa_set_refresh; demons are not touched.  The random bits of a site are
 * those of its word in the reference's MSC layout (packing_u ini.c:317-337),
 * shared by all 64 samples, so L must be a multiple of 8.  Blocks until done. */
int msa_run_thermal(msa_ctx *ctx, int64_t niter, const msa_lut *lut, const msa_u32x4 *ctr,
                    const uint32_t *key);

/* msa_run_thermal without allocating device memory: the same iterations,
 * bit-identical results, with the search table, counters, keys and thermal
 * random words in a fixed device array of 16 MiB that the library holds on
 * every GPU it runs on (set up by msa_create, kept until the process ends).
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Returns MSA_ERR_RANGE when a call needs more than the array
 * (about 24 V R / 128 + 32 R niter bytes plus 40 KiB).  All calls on one GPU
 * share the array: calls that overlap on one GPU (other host threads, other
 * contexts of the same device) race.  Blocks until done. */
int msa_run_thermal_noalloc(msa_ctx *ctx, int64_t niter, const msa_lut *lut, const msa_u32x4 *ctr,
                            const uint32_t *key);

/* Integer observables per sample lane (each array [..][64], NULL to skip):
 *   unsat[R][64]   unsatisfied bonds; energy h_Ener = 3V - 2*unsat
 *                  (d_computeEne main.cu:745-815, Measure_Energy_Spins main.cu:315-336)
 *   demon[R][64]   sum over layers and sites of D0 + 2 D1
 *                  (Measure_Energy_Clerics med.c:55-83, on the device here)
 *   q[R/2][64]     sites where replicas 2p and 2p+1 differ
 *   qlink[R/2][64] bonds where s_x s_y differs between replicas 2p and 2p+1
 * q and qlink have no reference counterpart.  Blocks until done. */
int msa_measure(msa_ctx *ctx, int32_t *unsat, int32_t *demon, int32_t *q, int32_t *qlink);

/* msa_measure in two halves: msa_measure_async queues the measurement of the
 * state as the work queued so far leaves it and returns; msa_measure_wait
 * waits for it and fills the arrays as msa_measure does.  Work queued in
 * between (msa_run ...) runs meanwhile, so a loop can queue its next sweeps
 * before it collects the counts.  One measurement may be queued at a time
 * (MSA_ERR_STATE otherwise); msa_measure in between is allowed.  The first
 * msa_measure_async of a context makes a small page-locked array it keeps.
 * No reference counterpart (Measure_Energy_Spins, main.cu:315-336, waits). */
int msa_measure_async(msa_ctx *ctx);
int msa_measure_wait(msa_ctx *ctx, int32_t *unsat, int32_t *demon, int32_t *q, int32_t *qlink);

/* Demon refresh from the reference's walker streams (InitWalkers,
 * ini.c:712-781, run on host threads there, refresh_walkers main.cu:340-377):
 * per (sample, replica) one sequential stream of the PR+LCG and xoshiro256++
 * generators, 64 draws per multispin word, MSC_V * nc words per refresh, demon
 * = number of thresholds the sum of the two outputs lies below.  The state of
 * a stream: ring = the last 61 values of the PR lagged sum (x_{n-61} ..
 * x_{n-1}), lcg = the LCG state, xo = the xoshiro256++ state. */
typedef struct {
  uint64_t ring[61];
  uint64_t lcg;
  uint64_t xo[4];
} msa_walker_stream;

/* Replace the Philox refresh of msa_run by the walker streams: n streams per
 * replica, lanes[i] the lane of stream i, st[i*R + r] their states, thr[k][3]
 * the thresholds of refresh k (get_thresholds, ini.c:680-710).  Refresh k
 * writes every demon (all layers, both colours) of the listed lanes from the
 * next segment of each stream and zeros in the other lanes; refresh indices
 * follow msa_run's cadence.  L must be a multiple of 8.  n = 0 returns to
 * the Philox refresh.  Bit-identical to the reference's walkers. */
int msa_set_walker_streams(msa_ctx *ctx, int32_t n, const int32_t *lanes,
                           const msa_walker_stream *st, int32_t nref, const uint64_t *thr);
/* The streams' current states (after the refreshes run so far), st[n*R]. */
int msa_get_walker_streams(msa_ctx *ctx, msa_walker_stream *st);

/* Kernel choice for sweeps without a refresh (no reference counterpart; the
 * autotuner and the batched form set it).  mode 0: two kernels per sweep,
 * one per colour.  mode 1: one kernel per sweep whose blocks own ty rows x
 * tz planes (ty, tz dividing L; 0,0 = built-in default).  mode 2: the same
 * tiles, each word loaded once per block (register plane window, LDS
 * exchange) and tiles placed per XCD.  mode 3 (nc = 1 only): two
 * refresh-free iterations per kernel (temporal blocking) on ty x tz tiles
 * (ty + 6 thread rows, one replica per block); a single refresh-free
 * iteration left between refreshes uses the mode 2 kernel (or mode 0 where
 * mode 2 does not exist for R).  mode 4 (lattices whose state fits one CU:
 * L^3/2 <= 2048 sites of a colour and (8 + 16 nc (+ 12 when L^3/2 > 1024))
 * L^3 bytes <= 160 KiB: L = 16 with one layer, 14 up to two, 12 up to five,
 * 10 and below up to six; the default for them): up to 64 iterations per launch, one
 * block per replica with the lattice in LDS; refresh sweeps whose demons are
 * drawn ahead (refresh modes 1 and 2) run in the same launches.  Sweeps with
 * a refresh drawn inside the sweep (refresh mode 0) always use the
 * two-kernel form.  Results are identical in every mode; only the speed
 * differs. */
int msa_set_tiling(msa_ctx *ctx, int32_t mode, int32_t ty, int32_t tz);
int msa_get_tiling(msa_ctx *ctx, int32_t *mode, int32_t *ty, int32_t *tz);

/* How msa_run performs the sweeps with a Philox refresh (no reference
 * counterpart; the reference draws its refresh on the host, main.cu:340-377).
 * mode 0: the refresh is drawn inside the sweep's demon load (two kernels per
 * sweep, one per colour).  mode 1: the demons of the coming refreshes are
 * drawn ahead, up to 64 refreshes per launch, by a separate kernel on a
 * second stream of the context (it overlaps the sweeps before them, also
 * across msa_run calls; at most 512 MiB of device memory, allocated on the
 * first such run), and each refresh sweep runs the refresh-free kernels
 * reading them; a refresh sweep followed by a refresh-free one runs as the
 * msa_set_tiling sweep pair.  mode 2: as mode 1, and two refresh sweeps in a
 * row run as one sweep pair of at most 512 threads per block (nc = 1, L up to
 * 144; elsewhere as mode 1).  mode -1 (the default): the mode msa_autotune
 * measured fastest for the geometry, or before any autotune 2 where the
 * pair's tile has at least 6 rows, 1 where it exists or L <= 48, else 0.
 * Results are identical in every mode; msa_get_refresh_mode gives the mode in
 * effect (0, 1 or 2).  Walker-stream refreshes are not affected. */
int msa_set_refresh_mode(msa_ctx *ctx, int32_t mode);
int msa_get_refresh_mode(msa_ctx *ctx, int32_t *mode);

/* Autotuner (no reference counterpart): times `sweeps` sweeps of every
 * kernel choice valid for this geometry on a scratch lattice of the same
 * geometry (the context's own state is not touched), applies the fastest and
 * remembers it for the process; later calls for the same (L, R, nc) reuse it.
 * It then times sweeps with a refresh each in refresh modes 0, 1 and 2 with
 * that kernel choice and makes the fastest the context's default refresh
 * mode (msa_set_refresh_mode(-1)).  Outputs may be NULL. */
int msa_autotune(msa_ctx *ctx, int32_t sweeps, int32_t *mode, int32_t *ty, int32_t *tz,
                 float *ms_per_sweep);

/* ------------------------------------------------------------ batches --
 * The batched form (no reference counterpart; the reference farms samples
 * out to separate processes, main.cu:33-50): many lattices of uneven L, each
 * with its own length and refresh schedule, advance together -- every launch
 * covers all unfinished lattices and a captured HIP graph replays `chunk`
 * iterations at a time.  Each lattice's results are bit-identical to msa_run
 * on the same lattice with the same tables.  Batched lattices have nc = 1. */
typedef struct msa_batch msa_batch;

typedef struct {
  msa_geom geom;            /* nc must be 1 */
  uint64_t seed;            /* state built as msa_init_philox(seed) */
  int64_t niter;            /* iterations this lattice runs */
  int64_t refresh_every;    /* cadence as in msa_run */
  int32_t nref;             /* refresh tables as in msa_set_refresh ... */
  const msa_u32x4 *ctr;     /* [nref][R] */
  const uint32_t *key;      /* [R] */
  const int32_t *lut_idx;   /* [nref], into the batch's shared LUT table */
} msa_job;

/* max_bytes bounds the device memory of the admitted lattices (<= 0: none). */
int msa_batch_create(int device, int64_t max_bytes, msa_batch **out);
int msa_batch_destroy(msa_batch *b);
/* Shared LUT pairs, before the first msa_batch_add. */
int msa_batch_set_luts(msa_batch *b, int32_t nlut, const msa_lut *luts);
/* Admit a lattice; MSA_ERR_NOMEM when its footprint (msa_batch_job_bytes)
 * does not fit under max_bytes next to those admitted. */
int msa_batch_add(msa_batch *b, const msa_job *job, int32_t *slot);
int64_t msa_batch_job_bytes(const msa_job *job);
int64_t msa_batch_bytes_in_use(msa_batch *b);
/* Replace the constructed state of a lattice that has not run yet. */
int msa_batch_upload_state(msa_batch *b, int32_t slot, const uint64_t *J, const uint64_t *spin,
                           const uint64_t *d0, const uint64_t *d1);
/* Launch shape: threads per block (128, 256, 512, 1024) and iterations per graph. */
int msa_batch_set_shape(msa_batch *b, int32_t block, int32_t chunk);
/* Advance every unfinished lattice by up to niter iterations (asynchronous:
 * the launches are enqueued on the batch's stream). */
int msa_batch_run(msa_batch *b, int64_t niter);
int msa_batch_sync(msa_batch *b);
/* Wait for the work enqueued before the latest msa_batch_run, so a caller can
 * keep one run queued on the device while it handles the previous one. */
int msa_batch_wait_previous(msa_batch *b);
/* Iteration reached and length of slots 0..nslots-1 (-1 for empty slots),
 * read from the device after a sync. */
int msa_batch_progress(msa_batch *b, int32_t nslots, int64_t *t, int64_t *tend);
/* The same counts as planned by the host: the iterations enqueued so far,
 * which the device reaches when it finishes that work (no sync). */
int msa_batch_planned(msa_batch *b, int32_t nslots, int64_t *t, int64_t *tend);
/* Per-lane observables and state of a slot.  For a lattice that has reached
 * its length these wait only for the launches that finished it (on the
 * lattice's own stream), not for later batch work; otherwise for the batch. */
int msa_batch_measure(msa_batch *b, int32_t slot, int32_t *unsat, int32_t *demon, int32_t *q,
                      int32_t *qlink);
int msa_batch_download_state(msa_batch *b, int32_t slot, uint64_t *spin, uint64_t *d0, uint64_t *d1);
/* Retire a lattice.  Its device buffers are kept (up to 256 lattices) and
 * reused by a later admission of the same geometry, so a serving loop does no
 * device allocation in steady state; lattices kept this way count against
 * max_bytes together with those admitted, and are freed first when a new
 * geometry needs room. */
int msa_batch_remove(msa_batch *b, int32_t slot);

/* HBM copy probe (no reference counterpart): copies `bytes` between two
 * fresh device buffers `iters` times with a plain streaming kernel and
 * reports the mean time per copy; calibrates bandwidth counters and gives
 * the practical bandwidth ceiling (2*bytes moved per copy). */
int msa_copy_probe(int device, int64_t bytes, int32_t iters, float *ms_per_copy);

/* Wait for all work on the context stream. */
int msa_sync(msa_ctx *ctx);

/* Page-locked host memory (NULL on failure) for the host arrays of the
 * state transfers: msa_upload_state / msa_download_state on such arrays run
 * at the host link's full rate instead of through the runtime's staging
 * copies.  No reference counterpart (its host arrays are malloc'd,
 * ini.c:34-60); the driver keeps its checkpoint arrays here. */
void *msa_host_alloc(int64_t bytes);
void msa_host_free(void *p);

/* Without waiting: *done = 1 when all work queued on the context stream has
 * run (what msa_sync would wait for), else 0.  No reference counterpart: for
 * host loops that drive several lattices (the serving loop). */
int msa_query(msa_ctx *ctx, int32_t *done);

/* Bytes of device memory held by a context of this geometry (for admission
 * by HBM footprint). */
int64_t msa_footprint_bytes(const msa_geom *geom);

/* Description of the last error on the calling thread (never NULL). */
const char *msa_strerror(int status);

#ifdef __cplusplus
}
#endif
#endif
