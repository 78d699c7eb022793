// msa_ctx.hpp -- the lattice context shared by the translation units of
// libmsa_hip.so (msa_hip.hip: single lattice; msa_batch.hip: batches).
#pragma once
#include <vector>

#include "msa_common.hpp"

// The drawn-ahead refresh pipeline of a context (msa_hip.hip, launch_run).
struct Predraw {
  int64_t kb = 0;           // refresh index of chunk 0
  int64_t drawn = 0;        // chunks whose draw is queued
  int64_t waited = -1;      // last chunk the sweep stream waits for
  int64_t released = 0;     // chunks whose readers are all queued (e_used recorded)
  bool live = false;        // the drawn chunks match the current tables
};

struct msa_ctx {
  int device = 0;
  msa_geom geom{};
  msa::Geo g{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t *sp[2] = {nullptr, nullptr};   // spins, ping-pong
  uint64_t *jp = nullptr;
  ulonglong2 *dm[2] = {nullptr, nullptr}; // demons, ping-pong
  int scur = 0;                           // spin buffer holding the current state
  int cur = 0;                            // demon buffer holding the current state
  msa::FineLut *luts = nullptr;           // [2 * nlut] search tables (plane 0, plane 1)
  bool luts_owned = true;                 // false: a batch's shared table
  int32_t nlut = 0;
  uint4 *ctr = nullptr;
  uint32_t *key = nullptr;
  int32_t nref = 0;
  int32_t ctr_cap = 0;                    // refreshes the ctr buffer holds
  std::vector<int32_t> lut_idx;           // host copy, indexes luts
  int32_t *meas = nullptr;                // counts [2R + 2(R/2)][64]
  int32_t *meas_host = nullptr;           // page-locked copy of the counts
  hipEvent_t e_meas = nullptr;            // after the copy of the last queued measurement
  bool meas_busy = false;                 // a measurement is queued and not yet waited for
  void *slab = nullptr;                   // the one allocation holding sp, jp, dm and meas
  uint64_t *th_rnd = nullptr;             // thermal random words [3][V/128 * R]
  uint32_t hnr = 1, hgrid = 1;            // two-pass kernel: rows per tile, persistent blocks
  uint32_t hrs = 1;                       // its replica groups per block
  int fused = 0;                          // sweeps without refresh: 0 two-pass, 1/2 one-pass kernel
  uint32_t fty = 0, ftz = 0, fny = 0;     // its tile: rows, planes, thread rows
  int fyhalo = 1;
  int pair = 0;                           // refresh-free sweeps two per pass (k_sweep_pair)
  int res = 0;                            // iterations in LDS-resident launches (k_sweep_resident)
  uint32_t pty = 0, ptz = 0;              // its tile: rows, planes
  // Philox refresh sweeps: 0 = the refresh fused into the two-pass kernel's
  // demon load; 1 = demons drawn ahead by k_refresh_draw on a second stream
  // and read by the refresh-free kernels; 2 = as 1, two refresh sweeps in a
  // row as one D2 sweep pair; -1 = chosen by geometry
  int predraw = -1;
  uint32_t qty = 0, qtz = 0;              // tile of the D2 sweep pair (0: none)
  int tuned_rmode = -1;                   // refresh mode msa_autotune chose (-1: none)
  hipStream_t dstream = nullptr;
  hipEvent_t e_drawn[2] = {nullptr, nullptr}, e_used[2] = {nullptr, nullptr};
  ulonglong2 *pre = nullptr;              // [2 pK][R nc V] drawn demons (two halves)
  uint32_t pK = 0;                        // refreshes per half
  uint32_t ncu = 0;                       // compute units of the device
  Predraw pd;
  // msa_download_state_async: the state relaid out to canonical order on
  // `stream`, then copied to the host on xstream while later work runs
  void *xfer = nullptr;
  size_t xfer_bytes = 0;
  hipStream_t xstream = nullptr;
  hipEvent_t e_snap = nullptr, e_xfer = nullptr;
  bool xfer_busy = false;                 // a copy is queued and not yet waited for
  // demon refresh from the reference's walker streams (msa_walkers.hip)
  int32_t w_n = 0;                        // streams per replica, 0 = Philox refresh
  int w_cur = 0;                          // w_st buffer holding the current states
  msa_walker_stream *w_st[2] = {nullptr, nullptr};
  void *w_cs = nullptr;                   // per-chunk start states
  int32_t *w_slot = nullptr;              // [64] stream of each lane
  uint64_t *w_lf = nullptr, *w_lcg = nullptr, *w_xo = nullptr;   // jump tables
  uint32_t w_W = 0, w_nchunk = 0;         // words per chunk, chunks per segment
  std::vector<uint64_t> w_thr;            // [nref][3] thresholds
};


// Point the context at search tables it does not own (a batch's shared table).
int msa_ctx_share_luts(msa_ctx *c, msa::FineLut *luts, int32_t nlut);

// Upload nlut LUT pairs and build their search tables (2 * nlut msa::FineLut)
// in new device memory *out; synchronous on `st`.
int msa_build_fine_luts(hipStream_t st, int32_t nlut, const msa_lut *host, msa::FineLut **out);

// Walker-stream refresh k of the current demon buffer (msa_walkers.hip).
int walkers_refresh(msa_ctx *c, int64_t k);
void walkers_free(msa_ctx *c);

// Dynamic LDS of the refresh kernels (two search tables).
constexpr size_t kRefreshLds = 2 * sizeof(msa::FineLut);
