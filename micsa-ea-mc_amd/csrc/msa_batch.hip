// msa_batch.hip -- the batched form: many lattices of uneven L advanced in
// the same kernel launches (see msa_batch_* in include/msa.h).
//
// Every lattice is a full msa_ctx (its own ping-pong state and refresh
// tables); the batch adds a device array of lattice descriptors, a split of
// the tiles of every active lattice over persistent blocks, and a captured
// HIP graph of `chunk` iterations (two launches each: colour 0, colour 1).  Each
// lattice keeps its iteration counter t on the device; the last block of a
// lattice to finish colour 1 advances it, so the captured graph can be
// replayed any number of times and lattices that reach their length stop by
// themselves.  Integer results are those of msa_run on the same lattice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "msa_common.hpp"
#include "msa_ctx.hpp"

struct BLat {
  uint64_t *sp[2];
  const uint64_t *jp;
  ulonglong2 *dm[2];
  msa::Geo g;
  const uint4 *ctr;
  const uint32_t *key;
  const msa::FineLut *luts;
  const int32_t *lut_idx;
  int64_t every;
  uint64_t t, tend;       // t: the host's copy (the kernels use BProg)
  uint32_t ntiles;        // tiles of this lattice
  int32_t nref;
  uint32_t bx, nr, nxc;   // a tile is bx site pairs of nr rows (y, z); nxc tiles per row block
  uint32_t mLh, mL;       // reciprocals for n / Lh and n / L (div_recip)
};

// The progress of a lattice, the only batch state the kernels write: its
// iteration counter and the tiles of colour 1 finished in this iteration.
// Kept apart from the read-only descriptors so that those load into scalar
// registers.
struct BProg {
  uint64_t t;
  uint32_t done, pad;
};

// A run of tiles tile0 .. tile0+ntile-1 of lattice `lat`.  Block b of a
// launch works through segments seg_start[b] .. seg_start[b+1]-1.
struct BSeg {
  uint32_t lat, tile0, ntile, pad;
};

// floor(n / d) for n * d < 2^32 with m = floor(2^32 / d) + 1: the error of
// n * m / 2^32 over n / d is below n / 2^32 < 1 / d, so the floor is exact.
static inline uint32_t div_recip_of(uint32_t d) { return (uint32_t)((1ull << 32) / d + 1); }
__device__ __forceinline__ uint32_t div_recip(uint32_t n, uint32_t m) { return __umulhi(n, m); }

// One colour of one sweep for every lattice of the batch: the update of
// k_halfsweep (d_oneMCstepBN_multisample + fused walk + fused refresh) for
// nc = 1, with each lattice's own geometry, buffers, counters and tables.
// Blocks are persistent: the host splits the tiles of all unfinished lattices
// into one contiguous run per block (consecutive runs on one XCD), so a
// block's run covers one lattice or the end of one and the start of the next;
// it stages a lattice's two search tables in LDS when its run reaches that
// lattice and the lattice refreshes this iteration.  Index arithmetic is
// 32-bit (every batched lattice has L^3 < 2^31, checked on admission) and
// divisions by L/2 and L use per-lattice reciprocals.  Held to four waves
// per SIMD (128 VGPRs), the occupancy the tables' LDS allows anyway.  R1:
// every lattice of the launch has one replica (the host checks), so the
// replica loop is a single pass and needs fewer registers.
template <int BS, bool R1>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4)))
k_batch_halfsweep(const BLat *__restrict__ lats, BProg *prog, const BSeg *__restrict__ segs,
                  const uint32_t *__restrict__ seg_start, int colour) {
  extern __shared__ uint4 s_tab[];
  const msa::FineLut *tabs = reinterpret_cast<const msa::FineLut *>(s_tab);
  const msa::LutView t0 = msa::lut_view(tabs), t1 = msa::lut_view(tabs + 1);
  const msa::FineLut *staged = nullptr;
  const uint32_t s_end = seg_start[blockIdx.x + 1];
  for (uint32_t s = seg_start[blockIdx.x]; s < s_end; ++s) {
    const BSeg it = segs[s];
    const BLat *lp = lats + it.lat;
    BProg *pp = prog + it.lat;
    // t moves only once every tile of the lattice is counted, this run's
    // included, so every thread of the block reads the same value
    const uint64_t tv = pp->t;
    const uint64_t t = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(tv >> 32)) << 32) |
                       __builtin_amdgcn_readfirstlane((uint32_t)tv);
    if (t >= lp->tend) continue;                   // whole block: lattice finished
    const msa::Geo g = lp->g;
    const int64_t k = refresh_index(t, lp->every);
    const bool refresh = k >= 0 && k < lp->nref;
    if (refresh) {
      const msa::FineLut *src = lp->luts + 2 * (size_t)lp->lut_idx[k];
      if (src != staged) {
        __syncthreads();                           // the previous tables are no longer read
        const uint4 *s16 = reinterpret_cast<const uint4 *>(src);
        const uint32_t n16 = 2 * sizeof(msa::FineLut) / 16;
        for (uint32_t j = threadIdx.x; j < n16; j += BS) s_tab[j] = s16[j];
        __syncthreads();
        staged = src;
      }
    }
    const int par = (int)(t & 1u);
    const uint64_t *A = lp->sp[par];
    uint64_t *B = lp->sp[par ^ 1];
    const ulonglong2 *dmi = lp->dm[par];
    ulonglong2 *dmo = lp->dm[par ^ 1];
    const uint32_t bx = lp->bx, nr = lp->nr, nxc = lp->nxc, Lh = g.Lh, L = g.L;
    const uint32_t rloc = bx == Lh ? div_recip(threadIdx.x, lp->mLh) : threadIdx.x / bx;
    const uint32_t xin = threadIdx.x - rloc * bx;
    const int col = colour;
    const uint32_t H = (uint32_t)g.H;
    const uint64_t *Jc = lp->jp + (uint64_t)col * 3 * g.H;
    const uint64_t *Jo = lp->jp + (uint64_t)(col ^ 1) * 3 * g.H;
    const uint64_t *Sn = col ? B : A;
    const uint4 *ctr = lp->ctr + (refresh ? (size_t)k * g.R : 0);
    for (uint32_t tt = 0; tt < it.ntile; ++tt) {
      const uint32_t tile = it.tile0 + tt;
      const uint32_t rbk = nxc == 1 ? tile : tile / nxc;
      const uint32_t i = (tile - rbk * nxc) * bx + xin;
      const uint32_t row = rbk * nr + rloc;
      if (rloc >= nr || i >= Lh || row >= L * L) continue;
      const uint32_t z = div_recip(row, lp->mL), y = row - z * L;
      const uint32_t off = (y + z + (uint32_t)col) & 1u;
      const uint32_t x = 2u * i + off;
      const uint32_t ip = off ? msa::wrapi((int)i + 1, Lh) : i;
      const uint32_t im = off ? i : msa::wrapi((int)i - 1, Lh);
      const uint32_t yp = msa::wrapi((int)y + 1, L), ym = msa::wrapi((int)y - 1, L);
      const uint32_t zp = msa::wrapi((int)z + 1, L), zm = msa::wrapi((int)z - 1, L);
      const uint32_t rb = Lh * row;
      const uint32_t h = i + rb;
      const uint32_t hxp = ip + rb, hxm = im + rb;
      const uint32_t hyp = i + Lh * (yp + L * z), hym = i + Lh * (ym + L * z);
      const uint32_t hzp = i + Lh * (y + L * zp), hzm = i + Lh * (y + L * zm);
      const uint32_t hw = (uint32_t)msa::walk_dest(g, col, 1, 0, 0, i, y, z);   // layer 0 walks along +x
      const uint32_t tid = x + L * row;
      uint64_t J0, J1, J2, J3, J4, J5;
      for (uint32_t r = 0; r < (R1 ? 1u : g.R); ++r) {
        const uint64_t *So = Sn + (uint64_t)(2 * r + (col ^ 1)) * g.H;
        const ulonglong2 *Di = dmi + ((uint64_t)r * 2 + col) * g.H;
        ulonglong2 *Do = dmo + ((uint64_t)r * 2 + col) * g.H;
        uint64_t sown, n0, n1, n2, n3, n4, n5;
        auto load = [&]() {
          if (r == 0) {
            J0 = Jc[h]; J1 = Jc[H + h]; J2 = Jc[2 * H + h];
            J3 = Jo[hxm]; J4 = Jo[H + hym]; J5 = Jo[2 * H + hzm];
          }
          sown = A[(uint64_t)(2 * r + col) * g.H + h];
          n0 = So[hxp]; n1 = So[hyp]; n2 = So[hzp]; n3 = So[hxm]; n4 = So[hym]; n5 = So[hzm];
        };
        uint64_t D0, D1;
        if (refresh) {
          // the site's loads are issued first and land while the refresh
          // (Philox and the table search, VALU-bound) runs
          load();
          const uint4 c = ctr[r];
          const uint32_t key = lp->key[r];
          D0 = msa::refresh_plane(t0, c, key, tid, 0);
          D1 = msa::refresh_plane(t1, c, key, tid, 1);
        } else {
          const ulonglong2 d = Di[h];
          D0 = d.x;
          D1 = d.y;
          load();
        }
        const uint64_t sn = msa::site_update(sown, n0, n1, n2, n3, n4, n5, J0, J1, J2, J3, J4, J5,
                                             D0, D1);
        B[(uint64_t)(2 * r + col) * g.H + h] = sn;
        Do[hw] = make_ulonglong2(D0, D1);
      }
    }
    if (colour == 1) {
      __syncthreads();
      if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&pp->done, it.ntile) + it.ntile == lp->ntiles) {   // the lattice's last run
          pp->done = 0;
          pp->t = t + 1;
          __threadfence();
        }
      }
    }
  }
}

static int batch_lds_attr();

struct msa_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t max_bytes = 0, used_bytes = 0;
  std::vector<msa_ctx *> slot;          // nullptr = free
  std::vector<int64_t> slot_bytes;
  std::vector<int32_t *> slot_lut_idx;  // device copies
  std::vector<int32_t> slot_idx_cap;    // their capacity
  // retired lattices kept for reuse by admissions of the same geometry
  // (no allocation on the device); their bytes count against max_bytes
  struct Idle {
    msa_ctx *c;
    int32_t *idx;
    int32_t cap;
    int64_t bytes;
    bool quiet;   // finished and its last launches completed: no launch reads its buffers
  };
  std::vector<Idle> idle;
  int64_t idle_bytes = 0;
  std::vector<BLat> h_lats;
  std::vector<char> fresh;              // added since the device table was last pushed
  BLat *d_lats = nullptr;
  BProg *d_prog = nullptr;
  size_t cap = 0;
  BSeg *d_segs = nullptr;                // the blocks' runs of tiles
  uint32_t *d_seg_start = nullptr;      // [nblocks + 1]
  uint32_t nblocks = 0, segs_cap = 0, starts_cap = 0;
  uint64_t ntiles = 0;                  // tiles of the unfinished lattices in the split
  // pinned staging of the descriptor, progress and split uploads; ev_up marks
  // the end of the last upload (the staging is not rewritten before it)
  char *stage = nullptr;
  size_t stage_cap = 0;
  hipEvent_t ev_up = nullptr;
  bool graph_stale = true;              // launch arguments changed since the capture
  // per slot: an event recorded after the launches in which the lattice
  // reached its length (ev_set), so its results can be read on its own
  // stream without waiting for later batch work
  std::vector<hipEvent_t> slot_ev;
  std::vector<char> slot_ev_set;
  hipEvent_t ev_prev = nullptr;         // end of the work enqueued before the last msa_batch_run
  msa::FineLut *luts = nullptr;         // shared search tables, 2 per pair
  int32_t nlut = 0;
  int block = 256;
  bool r1 = false;                      // launches use the one-replica kernel
  int ncu = 256;                        // compute units of the device
  int32_t chunk = 16;                   // iterations per captured graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool dirty = true;
};

static int batch_clear_graph(msa_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  return MSA_OK;
}

// Bring the host copy of the iteration counters up to date from the device
// table for every lattice that is in that table (not the ones added since it
// was pushed; a slot can be reused, and its new lattice may even get the old
// one's buffers back from the allocator, so identity is tracked explicitly).
static int batch_pull(msa_batch *b) {
  if (!b->d_prog || !b->cap) return MSA_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<BProg> cur(b->cap);
  HIPCHK(hipMemcpyAsync(cur.data(), b->d_prog, sizeof(BProg) * b->cap, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (size_t s = 0; s < std::min(b->slot.size(), b->cap); ++s)
    if (b->slot[s] && !b->fresh[s]) b->h_lats[s].t = cur[s].t;
  return MSA_OK;
}

extern "C" int msa_batch_create(int device, int64_t max_bytes, msa_batch **out) {
  if (!out) return msa_fail(MSA_ERR_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return msa_fail(MSA_ERR_ARG, "device=%d count=%d", device, ndev);
  HIPCHK(hipSetDevice(device));
  int rc = batch_lds_attr();
  if (rc) return rc;
  msa_batch *b = new msa_batch();
  b->device = device;
  if (hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || b->ncu < 1)
    b->ncu = 256;
  b->max_bytes = max_bytes > 0 ? max_bytes : INT64_MAX;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_up, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_prev, hipEventDisableTiming) != hipSuccess) {
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return msa_fail(MSA_ERR_HIP, "stream/event creation failed");
  }
  *out = b;
  return MSA_OK;
}

extern "C" int msa_batch_destroy(msa_batch *b) {
  if (!b) return MSA_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  batch_clear_graph(b);
  for (size_t s = 0; s < b->slot.size(); ++s) {
    if (b->slot[s]) msa_destroy(b->slot[s]);
    (void)hipFree(b->slot_lut_idx[s]);
  }
  for (auto &e : b->idle) {
    msa_destroy(e.c);
    (void)hipFree(e.idx);
  }
  (void)hipFree(b->d_lats);
  (void)hipFree(b->d_prog);
  (void)hipFree(b->d_segs);
  (void)hipFree(b->d_seg_start);
  (void)hipFree(b->luts);
  (void)hipHostFree(b->stage);
  if (b->ev_up) (void)hipEventDestroy(b->ev_up);
  if (b->ev_prev) (void)hipEventDestroy(b->ev_prev);
  for (hipEvent_t e : b->slot_ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return MSA_OK;
}

extern "C" int msa_batch_set_luts(msa_batch *b, int32_t nlut, const msa_lut *luts) {
  if (!b || !luts || nlut < 1) return msa_fail(MSA_ERR_ARG, "need at least one LUT pair");
  for (const msa_ctx *c : b->slot)
    if (c) return msa_fail(MSA_ERR_STATE, "set the shared LUTs before adding lattices");
  for (int32_t k = 0; k < 2 * nlut; ++k)
    if (!lut_valid(luts[k])) return msa_fail(MSA_ERR_ARG, "LUT %d is malformed", k);
  HIPCHK(hipSetDevice(b->device));
  (void)hipFree(b->luts);
  b->luts = nullptr;
  b->nlut = 0;
  const int rc = msa_build_fine_luts(b->stream, nlut, luts, &b->luts);
  if (rc) return rc;
  b->nlut = nlut;
  return MSA_OK;
}

static int64_t job_bytes(const msa_job *j) {
  return msa_footprint_bytes(&j->geom) + (int64_t)j->nref * (16 * j->geom.R + 4) + 64 * j->geom.R;
}

extern "C" int64_t msa_batch_job_bytes(const msa_job *job) {
  if (!job) return -1;
  return job_bytes(job);
}

extern "C" int msa_batch_add(msa_batch *b, const msa_job *j, int32_t *slot_out) {
  if (!b || !j || !slot_out) return msa_fail(MSA_ERR_ARG, "NULL argument");
  if (j->geom.nc != 1) return msa_fail(MSA_ERR_ARG, "batched lattices use one demon layer (nc=%d)", j->geom.nc);
  if ((uint64_t)j->geom.L * j->geom.L * j->geom.L >= (1ull << 31))
    return msa_fail(MSA_ERR_ARG, "batched lattices need L^3 < 2^31 (L=%d)", j->geom.L);
  if (!b->luts) return msa_fail(MSA_ERR_STATE, "set the shared LUTs first");
  if (j->niter < 0) return msa_fail(MSA_ERR_ARG, "niter=%lld", (long long)j->niter);
  if (j->refresh_every < -1) return msa_fail(MSA_ERR_ARG, "refresh_every=%lld", (long long)j->refresh_every);
  if (j->niter > 0) {
    // every refresh the run will need must be in the table
    int64_t kmax = -1;
    if (j->refresh_every > 0) kmax = (j->niter - 1) / j->refresh_every;
    else if (j->refresh_every == -1) kmax = refresh_index(j->niter <= 1 ? 0 : 1ull << (63 - __builtin_clzll((uint64_t)j->niter - 1)), -1);
    if (kmax >= j->nref) return msa_fail(MSA_ERR_RANGE, "run of %lld iterations needs %lld refreshes, %d given",
                                         (long long)j->niter, (long long)kmax + 1, j->nref);
  }
  const int64_t bytes = job_bytes(j);
  if (bytes < 0) return msa_fail(MSA_ERR_ARG, "bad geometry");
  if (b->used_bytes + bytes > b->max_bytes)
    return msa_fail(MSA_ERR_NOMEM, "admission: %lld bytes in use + %lld > limit %lld",
                    (long long)b->used_bytes, (long long)bytes, (long long)b->max_bytes);
  // an idle lattice of the same geometry, else make room among the idle ones;
  // nothing of the batch may still run when an idle lattice's buffers are
  // rewritten or freed (launches enqueued before its removal may still read it)
  HIPCHK(hipSetDevice(b->device));
  msa_ctx *c = nullptr;
  int32_t *d_idx = nullptr;
  int32_t idx_cap = 0;
  for (size_t k = 0; k < b->idle.size(); ++k) {
    const msa_geom &g = b->idle[k].c->geom;
    if (g.L == j->geom.L && g.R == j->geom.R && g.nc == j->geom.nc) {
      if (!b->idle[k].quiet) HIPCHK(hipStreamSynchronize(b->stream));
      c = b->idle[k].c;
      d_idx = b->idle[k].idx;
      idx_cap = b->idle[k].cap;
      b->idle_bytes -= b->idle[k].bytes;
      b->idle.erase(b->idle.begin() + (long)k);
      break;
    }
  }
  while (!c && !b->idle.empty() && b->used_bytes + b->idle_bytes + bytes > b->max_bytes) {
    HIPCHK(hipStreamSynchronize(b->stream));
    msa_destroy(b->idle.back().c);
    (void)hipFree(b->idle.back().idx);
    b->idle_bytes -= b->idle.back().bytes;
    b->idle.pop_back();
  }
  int rc = MSA_OK;
  if (!c && (rc = msa_create(b->device, &j->geom, &c))) return rc;
  auto drop = [&](int code) {
    msa_destroy(c);
    (void)hipFree(d_idx);
    return code;
  };
  if ((rc = msa_init_philox(c, j->seed)) || (rc = msa_ctx_share_luts(c, b->luts, b->nlut)) ||
      (j->nref > 0 && (rc = msa_set_refresh(c, j->nref, j->ctr, j->key, j->lut_idx))))
    return drop(rc);
  if (j->nref > 0) {
    if (j->nref > idx_cap) {
      (void)hipFree(d_idx);
      d_idx = nullptr;
      idx_cap = 0;
      if (hipMalloc(&d_idx, sizeof(int32_t) * (size_t)j->nref) != hipSuccess)
        return drop(msa_fail(MSA_ERR_NOMEM, "refresh index table"));
      idx_cap = j->nref;
    }
    if (hipMemcpyAsync(d_idx, j->lut_idx, sizeof(int32_t) * (size_t)j->nref, hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return drop(msa_fail(MSA_ERR_HIP, "refresh index upload"));
  }
  size_t s = 0;
  while (s < b->slot.size() && b->slot[s]) ++s;
  if (s == b->slot.size()) {
    b->slot.push_back(nullptr);
    b->slot_bytes.push_back(0);
    b->slot_lut_idx.push_back(nullptr);
    b->slot_idx_cap.push_back(0);
    b->h_lats.push_back(BLat{});
    b->fresh.push_back(0);
    b->slot_ev.push_back(nullptr);
    b->slot_ev_set.push_back(0);
  }
  b->slot_ev_set[s] = 0;
  b->fresh[s] = 1;
  b->slot[s] = c;
  b->slot_bytes[s] = bytes;
  b->slot_lut_idx[s] = d_idx;
  b->slot_idx_cap[s] = idx_cap;
  b->used_bytes += bytes;
  BLat &L = b->h_lats[s];
  L = BLat{};
  L.sp[0] = c->sp[0];
  L.sp[1] = c->sp[1];
  L.jp = c->jp;
  L.dm[0] = c->dm[0];
  L.dm[1] = c->dm[1];
  L.g = c->g;
  L.ctr = c->ctr;
  L.key = c->key;
  L.luts = b->luts;
  L.lut_idx = d_idx;
  L.every = j->refresh_every;
  L.t = 0;
  L.tend = (uint64_t)j->niter;
  L.nref = j->nref;
  b->dirty = true;
  *slot_out = (int32_t)s;
  return MSA_OK;
}

extern "C" int msa_batch_upload_state(msa_batch *b, int32_t slot, const uint64_t *J,
                                      const uint64_t *spin, const uint64_t *d0, const uint64_t *d1) {
  if (!b || slot < 0 || (size_t)slot >= b->slot.size() || !b->slot[slot])
    return msa_fail(MSA_ERR_ARG, "no lattice in slot %d", slot);
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  int rc = batch_pull(b);
  if (rc) return rc;
  const uint64_t t = b->h_lats[slot].t;
  if (t != 0) return msa_fail(MSA_ERR_STATE, "slot %d already ran %llu iterations", slot, (unsigned long long)t);
  return msa_upload_state(b->slot[slot], J, spin, d0, d1);
}

// The kernel of a block size and replica form.
static const void *batch_kernel(int block, bool r1) {
  switch (block) {
    case 128: return r1 ? reinterpret_cast<const void *>(&k_batch_halfsweep<128, true>)
                        : reinterpret_cast<const void *>(&k_batch_halfsweep<128, false>);
    case 512: return r1 ? reinterpret_cast<const void *>(&k_batch_halfsweep<512, true>)
                        : reinterpret_cast<const void *>(&k_batch_halfsweep<512, false>);
    case 1024: return r1 ? reinterpret_cast<const void *>(&k_batch_halfsweep<1024, true>)
                         : reinterpret_cast<const void *>(&k_batch_halfsweep<1024, false>);
    default: return r1 ? reinterpret_cast<const void *>(&k_batch_halfsweep<256, true>)
                       : reinterpret_cast<const void *>(&k_batch_halfsweep<256, false>);
  }
}

template <int BS>
static void batch_launch_bs(msa_batch *b, int col) {
  const dim3 grid(b->nblocks), block(BS);
  const size_t lds = kRefreshLds;
  const BLat *L = b->d_lats;
  BProg *P = b->d_prog;
  const BSeg *sg = b->d_segs;
  const uint32_t *st = b->d_seg_start;
  if (b->r1)
    hipLaunchKernelGGL((k_batch_halfsweep<BS, true>), grid, block, lds, b->stream, L, P, sg, st, col);
  else
    hipLaunchKernelGGL((k_batch_halfsweep<BS, false>), grid, block, lds, b->stream, L, P, sg, st, col);
}

static void batch_launch(msa_batch *b, int col) {
  switch (b->block) {
    case 128: batch_launch_bs<128>(b, col); break;
    case 512: batch_launch_bs<512>(b, col); break;
    case 1024: batch_launch_bs<1024>(b, col); break;
    default: batch_launch_bs<256>(b, col); break;
  }
}

// The batch kernels stage two search tables (68 KiB) in dynamic LDS.
static int batch_lds_attr() {
  for (int block : {128, 256, 512, 1024})
    for (bool r1 : {false, true})
      HIPCHK(hipFuncSetAttribute(batch_kernel(block, r1), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kRefreshLds));
  return MSA_OK;
}

// Blocks of the batch kernel with `block` threads that one CU holds at once
// (registers and the tables' LDS), at least 1.
static int batch_resident(int block, bool r1) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel(block, r1), block, kRefreshLds) !=
          hipSuccess || n < 1)
    n = 1;
  return n;
}

// The work split of one launch: the tiles of all unfinished lattices, in slot
// order, cut into `nblocks` contiguous runs of equal length (to one tile).
// Block b runs on XCD b mod kXcd, so the runs are dealt out such that each
// XCD takes consecutive runs: neighbouring tiles, which read each other's
// boundary rows, then share an L2.  A run is split where it crosses from one
// lattice to the next (segments).
constexpr uint32_t kXcd = 8;

static void batch_split(const std::vector<uint32_t> &tiles, const std::vector<uint32_t> &lat,
                        uint32_t nblocks, std::vector<BSeg> &segs, std::vector<uint32_t> &start) {
  uint64_t total = 0;
  std::vector<uint64_t> first(tiles.size() + 1, 0);   // first global tile of each lattice
  for (size_t j = 0; j < tiles.size(); ++j) first[j + 1] = first[j] + tiles[j];
  total = first.back();
  // position of block b in XCD-major order
  std::vector<uint32_t> rank(nblocks);
  {
    uint32_t pos = 0;
    for (uint32_t x = 0; x < kXcd; ++x)
      for (uint32_t bb = x; bb < nblocks; bb += kXcd) rank[bb] = pos++;
  }
  segs.clear();
  start.assign(nblocks + 1, 0);
  for (uint32_t bb = 0; bb < nblocks; ++bb) {
    start[bb] = (uint32_t)segs.size();
    const uint64_t lo = total * rank[bb] / nblocks, hi = total * (rank[bb] + 1) / nblocks;
    // lattice holding global tile lo
    size_t j = (size_t)(std::upper_bound(first.begin(), first.end(), lo) - first.begin()) - 1;
    for (uint64_t g = lo; g < hi; ++j) {
      const uint64_t end = std::min<uint64_t>(hi, first[j + 1]);
      if (end > g) segs.push_back(BSeg{lat[j], (uint32_t)(g - first[j]), (uint32_t)(end - g), 0});
      g = end;
    }
  }
  start[nblocks] = (uint32_t)segs.size();
}

// Device arrays of at least n elements; a reallocation invalidates the
// captured graph (its launches hold the old addresses) and waits for the
// launches already enqueued, which may still read the old array.
template <class T>
static int batch_grow(msa_batch *b, T **arr, uint32_t *cap, size_t n, const char *what) {
  if (n <= *cap) return MSA_OK;
  const size_t want = std::max<size_t>(n, 2 * (size_t)*cap);
  HIPCHK(hipStreamSynchronize(b->stream));
  (void)hipFree(*arr);
  *arr = nullptr;
  *cap = 0;
  if (hipMalloc(arr, sizeof(T) * want) != hipSuccess) return msa_fail(MSA_ERR_NOMEM, "%s", what);
  *cap = (uint32_t)want;
  b->graph_stale = true;
  return MSA_OK;
}

// Push descriptors, progress and the work split, and capture the graph when
// its launch arguments changed.  Nothing waits for the device: the uploads
// go on the batch stream, after the launches already enqueued, and the
// progress pushed for a running lattice is the host's count of the
// iterations enqueued for it, which those launches reach exactly.  The grid
// is the number of blocks the CUs hold at once whatever the batch holds
// (blocks without a run exit), so admissions and retirements reuse the
// captured graph.
static int batch_prepare(msa_batch *b) {
  if (!b->dirty) return MSA_OK;
  const size_t n = b->slot.size();
  std::vector<uint32_t> tiles, lat;     // unfinished lattices: tile counts and slots
  uint64_t total = 0;
  for (size_t s = 0; s < n; ++s) {
    if (!b->slot[s]) continue;
    BLat &L = b->h_lats[s];
    const uint32_t Lh = L.g.Lh, rows = L.g.L * L.g.L, bs = (uint32_t)b->block;
    L.bx = Lh <= bs ? Lh : bs;
    L.nr = Lh <= bs ? bs / Lh : 1;
    L.nxc = (Lh + L.bx - 1) / L.bx;
    L.mLh = div_recip_of(Lh);
    L.mL = div_recip_of(L.g.L);
    L.ntiles = ((rows + L.nr - 1) / L.nr) * L.nxc;
    if (L.t < L.tend) {
      tiles.push_back(L.ntiles);
      lat.push_back((uint32_t)s);
      total += L.ntiles;
    }
  }
  // the one-replica kernel when every lattice of the split has R = 1
  bool r1 = true;
  for (uint32_t s : lat) r1 = r1 && b->h_lats[s].g.R == 1;
  if (r1 != b->r1) b->graph_stale = true;
  b->r1 = r1;
  const uint32_t nblocks = (uint32_t)((uint64_t)b->ncu * (uint64_t)batch_resident(b->block, r1));
  if (nblocks != b->nblocks) b->graph_stale = true;
  std::vector<BSeg> segs;
  std::vector<uint32_t> start;
  batch_split(tiles, lat, nblocks, segs, start);
  if (n > b->cap) {                     // descriptors and progress share one capacity
    const size_t want = std::max<size_t>(n, 2 * b->cap);
    HIPCHK(hipStreamSynchronize(b->stream));
    (void)hipFree(b->d_lats);
    (void)hipFree(b->d_prog);
    b->d_lats = nullptr;
    b->d_prog = nullptr;
    b->cap = 0;
    if (hipMalloc(&b->d_lats, sizeof(BLat) * want) != hipSuccess ||
        hipMalloc(&b->d_prog, sizeof(BProg) * want) != hipSuccess)
      return msa_fail(MSA_ERR_NOMEM, "descriptors");
    b->cap = want;
    b->graph_stale = true;
  }
  int rc = MSA_OK;
  if ((rc = batch_grow(b, &b->d_segs, &b->segs_cap, std::max<size_t>(segs.size(), 1), "work split")) ||
      (rc = batch_grow(b, &b->d_seg_start, &b->starts_cap, start.size(), "work split")))
    return rc;
  // stage everything in pinned memory once the previous upload has left it
  const size_t o_prog = sizeof(BLat) * n, o_segs = o_prog + sizeof(BProg) * n;
  const size_t o_start = o_segs + sizeof(BSeg) * segs.size(), bytes = o_start + sizeof(uint32_t) * start.size();
  HIPCHK(hipEventSynchronize(b->ev_up));
  if (bytes > b->stage_cap) {
    (void)hipHostFree(b->stage);
    b->stage = nullptr;
    b->stage_cap = 0;
    if (hipHostMalloc(reinterpret_cast<void **>(&b->stage), 2 * bytes) != hipSuccess)
      return msa_fail(MSA_ERR_NOMEM, "upload staging");
    b->stage_cap = 2 * bytes;
  }
  BProg *prog = reinterpret_cast<BProg *>(b->stage + o_prog);
  for (size_t s = 0; s < n; ++s) prog[s] = BProg{b->h_lats[s].t, 0, 0};
  if (n) std::memcpy(b->stage, b->h_lats.data(), sizeof(BLat) * n);
  if (!segs.empty()) std::memcpy(b->stage + o_segs, segs.data(), sizeof(BSeg) * segs.size());
  std::memcpy(b->stage + o_start, start.data(), sizeof(uint32_t) * start.size());
  if (n) {
    HIPCHK(hipMemcpyAsync(b->d_lats, b->stage, sizeof(BLat) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_prog, prog, sizeof(BProg) * n, hipMemcpyHostToDevice, b->stream));
  }
  if (!segs.empty())
    HIPCHK(hipMemcpyAsync(b->d_segs, b->stage + o_segs, sizeof(BSeg) * segs.size(), hipMemcpyHostToDevice,
                          b->stream));
  HIPCHK(hipMemcpyAsync(b->d_seg_start, b->stage + o_start, sizeof(uint32_t) * start.size(),
                        hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipEventRecord(b->ev_up, b->stream));
  for (size_t s = 0; s < n; ++s) b->fresh[s] = 0;
  b->nblocks = nblocks;
  b->ntiles = total;
  if (b->graph_stale || !b->gexec) {
    HIPCHK(hipStreamSynchronize(b->stream));   // the old graph may still run
    batch_clear_graph(b);
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int32_t it = 0; it < b->chunk; ++it)
      for (int col = 0; col < 2; ++col) batch_launch(b, col);
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    b->graph_stale = false;
  }
  b->dirty = false;
  return MSA_OK;
}

extern "C" int msa_batch_set_shape(msa_batch *b, int32_t block, int32_t chunk) {
  if (!b) return msa_fail(MSA_ERR_ARG, "NULL batch");
  if (block != 128 && block != 256 && block != 512 && block != 1024)
    return msa_fail(MSA_ERR_ARG, "block=%d (128, 256, 512 or 1024)", block);
  if (chunk < 1 || chunk > 4096) return msa_fail(MSA_ERR_ARG, "chunk=%d", chunk);
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->block != block) {
    // tile counts change: progress is kept by batch_prepare, done counters are reset there
    b->block = block;
  }
  b->chunk = chunk;
  b->dirty = true;
  b->graph_stale = true;
  return MSA_OK;
}

extern "C" int msa_batch_run(msa_batch *b, int64_t niter) {
  if (!b) return msa_fail(MSA_ERR_ARG, "NULL batch");
  if (niter < 0) return msa_fail(MSA_ERR_ARG, "niter=%lld", (long long)niter);
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipEventRecord(b->ev_prev, b->stream));
  // Every unfinished lattice advances one iteration per iteration, so the
  // host knows where each reaches its length: the run is cut there and the
  // work list rebuilt without the finished lattices (their blocks would only
  // exit, and the remaining tiles are regrouped over the CUs).
  int64_t left = niter;
  while (left > 0) {
    if (b->dirty) {
      int rc = batch_prepare(b);
      if (rc) return rc;
    }
    if (!b->ntiles) break;
    int64_t step = left;
    for (size_t s = 0; s < b->slot.size(); ++s)
      if (b->slot[s] && b->h_lats[s].t < b->h_lats[s].tend)
        step = std::min<int64_t>(step, (int64_t)(b->h_lats[s].tend - b->h_lats[s].t));
    int64_t k = step;
    while (k >= b->chunk) {
      HIPCHK(hipGraphLaunch(b->gexec, b->stream));
      k -= b->chunk;
    }
    for (int64_t it = 0; it < k; ++it)
      for (int col = 0; col < 2; ++col) batch_launch(b, col);
    HIPCHK(hipGetLastError());
    for (size_t s = 0; s < b->slot.size(); ++s) {
      BLat &L = b->h_lats[s];
      if (!b->slot[s] || L.t >= L.tend) continue;
      L.t += (uint64_t)step;
      if (L.t >= L.tend) {
        b->dirty = true;
        if (!b->slot_ev[s] && hipEventCreateWithFlags(&b->slot_ev[s], hipEventDisableTiming) != hipSuccess)
          return msa_fail(MSA_ERR_HIP, "event creation failed");
        HIPCHK(hipEventRecord(b->slot_ev[s], b->stream));
        b->slot_ev_set[s] = 1;
      }
    }
    left -= step;
  }
  return MSA_OK;
}

extern "C" int msa_batch_sync(msa_batch *b) {
  if (!b) return msa_fail(MSA_ERR_ARG, "NULL batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  return MSA_OK;
}

extern "C" int msa_batch_wait_previous(msa_batch *b) {
  if (!b) return msa_fail(MSA_ERR_ARG, "NULL batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipEventSynchronize(b->ev_prev));
  return MSA_OK;
}

extern "C" int msa_batch_planned(msa_batch *b, int32_t nslots, int64_t *t, int64_t *tend) {
  if (!b || nslots < 0) return msa_fail(MSA_ERR_ARG, "bad argument");
  for (int32_t s = 0; s < nslots; ++s) {
    const bool live = (size_t)s < b->slot.size() && b->slot[s];
    if (t) t[s] = live ? (int64_t)b->h_lats[s].t : -1;
    if (tend) tend[s] = live ? (int64_t)b->h_lats[s].tend : -1;
  }
  return MSA_OK;
}

extern "C" int msa_batch_progress(msa_batch *b, int32_t nslots, int64_t *t, int64_t *tend) {
  if (!b || nslots < 0) return msa_fail(MSA_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(b->device));
  int rc = batch_pull(b);
  if (rc) return rc;
  for (int32_t s = 0; s < nslots; ++s) {
    const bool live = (size_t)s < b->slot.size() && b->slot[s];
    if (t) t[s] = live ? (int64_t)b->h_lats[s].t : -1;
    if (tend) tend[s] = live ? (int64_t)b->h_lats[s].tend : -1;
  }
  return MSA_OK;
}

// the slot's context with its buffer parity matching the device counter
static int batch_ctx(msa_batch *b, int32_t slot, msa_ctx **out) {
  if (!b || slot < 0 || (size_t)slot >= b->slot.size() || !b->slot[slot])
    return msa_fail(MSA_ERR_ARG, "no lattice in slot %d", slot);
  HIPCHK(hipSetDevice(b->device));
  msa_ctx *c = b->slot[slot];
  if (b->slot_ev_set[slot]) {
    // finished: its state is final once the launches that finished it are
    // done, and later batch work does not touch it
    HIPCHK(hipStreamWaitEvent(c->stream, b->slot_ev[slot], 0));
  } else {
    int rc = batch_pull(b);
    if (rc) return rc;
  }
  const uint64_t t = b->h_lats[slot].t;
  c->scur = c->cur = (int)(t & 1u);
  *out = c;
  return MSA_OK;
}

extern "C" int msa_batch_measure(msa_batch *b, int32_t slot, int32_t *unsat, int32_t *demon,
                                 int32_t *q, int32_t *qlink) {
  msa_ctx *c = nullptr;
  int rc = batch_ctx(b, slot, &c);
  return rc ? rc : msa_measure(c, unsat, demon, q, qlink);
}

extern "C" int msa_batch_download_state(msa_batch *b, int32_t slot, uint64_t *spin, uint64_t *d0,
                                        uint64_t *d1) {
  msa_ctx *c = nullptr;
  int rc = batch_ctx(b, slot, &c);
  return rc ? rc : msa_download_state(c, spin, d0, d1);
}

extern "C" int msa_batch_remove(msa_batch *b, int32_t slot) {
  if (!b || slot < 0 || (size_t)slot >= b->slot.size() || !b->slot[slot])
    return msa_fail(MSA_ERR_ARG, "no lattice in slot %d", slot);
  HIPCHK(hipSetDevice(b->device));
  if (b->idle.size() < 256) {           // keep it for a later admission of its geometry
    // quiet when it finished and the launches that finished it completed
    const bool quiet = b->slot_ev_set[slot] && hipEventQuery(b->slot_ev[slot]) == hipSuccess;
    b->idle.push_back({b->slot[slot], b->slot_lut_idx[slot], b->slot_idx_cap[slot], b->slot_bytes[slot], quiet});
    b->idle_bytes += b->slot_bytes[slot];
  } else {
    HIPCHK(hipStreamSynchronize(b->stream));   // enqueued launches may still read it
    msa_destroy(b->slot[slot]);
    (void)hipFree(b->slot_lut_idx[slot]);
  }
  b->slot[slot] = nullptr;
  b->slot_lut_idx[slot] = nullptr;
  b->used_bytes -= b->slot_bytes[slot];
  b->slot_bytes[slot] = 0;
  b->h_lats[slot] = BLat{};
  b->fresh[slot] = 0;
  b->slot_ev_set[slot] = 0;
  b->dirty = true;
  return MSA_OK;
}

extern "C" int64_t msa_batch_bytes_in_use(msa_batch *b) { return b ? b->used_bytes : -1; }
