// msa_common.hpp -- host-side helpers shared by the translation units of
// libmsa_hip.so (error reporting, layout maps, table checks, cadence).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>

#include "msa.h"
#include "msa_device.hpp"

// Record the message for msa_strerror() and return `code`.
int msa_fail(int code, const char *fmt, ...);

#define HIPCHK(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return msa_fail(MSA_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                      __FILE__, __LINE__);                                                \
  } while (0)

// canonical site s = x + L*(y + L*z) -> colour-split index h, colour *col
static inline uint64_t split_index(const msa::Geo &g, uint64_t s, uint32_t *col) {
  const uint32_t x = (uint32_t)(s % g.L);
  const uint64_t yz = s / g.L;
  const uint32_t y = (uint32_t)(yz % g.L), z = (uint32_t)(yz / g.L);
  *col = (x + y + z) & 1u;
  return (x >> 1) + (uint64_t)g.Lh * (y + (uint64_t)g.L * z);
}

// A heat-bath table is usable when its thresholds are non-decreasing, the
// last one is 2^64-1 and every presearch entry has min <= max <= 255.
static inline bool lut_valid(const msa_lut &t) {
  uint64_t prev = 0;
  for (int n = 0; n < 256; ++n) {
    const uint64_t u = ((uint64_t)t.umbrales[2 * n] << 32) | t.umbrales[2 * n + 1];
    if (u < prev) return false;
    prev = u;
  }
  if (prev != ~0ull) return false;
  const uint16_t *pre = reinterpret_cast<const uint16_t *>(t.prebusqueda);
  for (int m = 0; m < 1024; ++m)
    if ((pre[m] & 255) > (pre[m] >> 8)) return false;
  return true;
}

// Refresh index of iteration t, -1 when t does not refresh (see msa_run).
__host__ __device__ static inline int64_t refresh_index(uint64_t t, int64_t every) {
  if (every > 0) return (t % (uint64_t)every == 0) ? (int64_t)(t / (uint64_t)every) : -1;
  if (every == -1) {
    if (t == 0) return 0;
    if (t & (t - 1)) return -1;
    int64_t k = 1;
    while (t >>= 1) ++k;
    return k;   // log2(t) + 1
  }
  return -1;
}
