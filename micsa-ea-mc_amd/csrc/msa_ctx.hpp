// msa_ctx.hpp -- the lattice context shared by the translation units of
// libmsa_hip.so (msa_hip.hip: single lattice; msa_batch.hip: batches).
#pragma once
#include <vector>

#include "msa_common.hpp"

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
  msa_lut *luts = nullptr;
  bool luts_owned = true;                 // false: a batch's shared table
  int32_t nlut = 0;
  uint4 *ctr = nullptr;
  uint32_t *key = nullptr;
  int32_t nref = 0;
  std::vector<int32_t> lut_idx;           // host copy, indexes luts
  int32_t *meas = nullptr;                // [64]
  dim3 blk{1, 1, 1}, grd{1, 1, 1};        // two-pass kernel launch
  int fused = 0;                          // sweeps without refresh: 0 two-pass, 1/2 one-pass kernel
  uint32_t fty = 0, ftz = 0, fny = 0;     // its tile: rows, planes, thread rows
  int fyhalo = 1;
};


// Point the context at a LUT table it does not own (a batch's shared table).
int msa_ctx_share_luts(msa_ctx *c, msa_lut *luts, int32_t nlut);
