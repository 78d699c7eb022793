// msa_hip.hip -- kernels and C ABI of libmsa_hip.so (see include/msa.h).
//
// MI355X / gfx950 only.  Integer and bitwise work: the sweep is bound by HBM
// bandwidth, the fused demon refresh by integer VALU throughput.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "msa.h"
#include "msa_device.hpp"

using msa::Geo;

// ------------------------------------------------------------ errors ------
static thread_local char g_err[512] = "no error";

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(MSA_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                                 \
  } while (0)

// ----------------------------------------------------------- kernels ------
struct SweepArgs {
  uint64_t *sp;
  const uint64_t *jp;
  const ulonglong2 *dm_in;
  ulonglong2 *dm_out;
  Geo g;
  int layer, colour;
  int refresh;            // 1: generate demons instead of reading them
  const uint4 *ctr;       // [R] counters of this refresh
  const uint32_t *key;    // [R]
  const msa_lut *lut;     // [2] plane-0 and plane-1 tables of this refresh
};

// Lattice direction e_l for demon layer l (main.cu:1125-1175): +x,+y,+z,-x,-y,-z.
__device__ __forceinline__ void layer_dir(int l, int &dx, int &dy, int &dz) {
  const int sgn = l < 3 ? 1 : -1;
  const int a = l % 3;
  dx = a == 0 ? sgn : 0;
  dy = a == 1 ? sgn : 0;
  dz = a == 2 ? sgn : 0;
}

// One colour of a sweep with demon layer `layer` (d_oneMCstepBN_multisample,
// main.cu:847-1080, ISI branch), with the demon walk of that layer fused into
// the demon store (d_Walk, main.cu:1084-1193: colour-0 demons move +2e_l,
// colour-1 demons -2e_l) and, when `refresh` is set, the heat-bath refresh of
// the layer fused into the demon load.  Demons are read from dm_in and written
// to dm_out, so the walk's permutation never races with a load.
// Thread = one site of the colour, all R replicas (J is loaded once).
__global__ void __launch_bounds__(256) k_halfsweep(SweepArgs a) {
  __shared__ uint32_t s_lut[2 * 1024];
  const Geo g = a.g;
  if (a.refresh) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(a.lut);
    const uint32_t nthr = blockDim.x * blockDim.y;
    for (uint32_t k = threadIdx.y * blockDim.x + threadIdx.x; k < 2048; k += nthr) s_lut[k] = src[k];
    __syncthreads();
  }
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
  const uint32_t z = blockIdx.z;
  if (i >= g.Lh || y >= g.L) return;

  const int col = a.colour;
  const uint32_t off = (y + z + (uint32_t)col) & 1u;
  const uint32_t x = 2u * i + off;
  const uint32_t ip = off ? msa::wrapi((int)i + 1, g.Lh) : i;      // site x+1
  const uint32_t im = off ? i : msa::wrapi((int)i - 1, g.Lh);      // site x-1
  const uint32_t yp = msa::wrapi((int)y + 1, g.L), ym = msa::wrapi((int)y - 1, g.L);
  const uint32_t zp = msa::wrapi((int)z + 1, g.L), zm = msa::wrapi((int)z - 1, g.L);
  const uint64_t row = (uint64_t)g.Lh * (y + (uint64_t)g.L * z);
  const uint64_t h = i + row;
  const uint64_t hxp = ip + row, hxm = im + row;
  const uint64_t hyp = i + (uint64_t)g.Lh * (yp + (uint64_t)g.L * z);
  const uint64_t hym = i + (uint64_t)g.Lh * (ym + (uint64_t)g.L * z);
  const uint64_t hzp = i + (uint64_t)g.Lh * (y + (uint64_t)g.L * zp);
  const uint64_t hzm = i + (uint64_t)g.Lh * (y + (uint64_t)g.L * zm);

  const uint64_t *Jc = a.jp + (uint64_t)col * 3 * g.H;
  const uint64_t *Jo = a.jp + (uint64_t)(col ^ 1) * 3 * g.H;
  const uint64_t J0 = Jc[h], J1 = Jc[g.H + h], J2 = Jc[2 * g.H + h];
  const uint64_t J3 = Jo[hxm], J4 = Jo[g.H + hym], J5 = Jo[2 * g.H + hzm];

  // destination of this colour's demon after the walk
  int dx, dy, dz;
  layer_dir(a.layer, dx, dy, dz);
  const int sg = col ? -2 : 2;
  const uint32_t wi = msa::wrapi((int)i + (sg / 2) * dx, g.Lh);
  const uint32_t wy = msa::wrapi((int)y + sg * dy, g.L);
  const uint32_t wz = msa::wrapi((int)z + sg * dz, g.L);
  const uint64_t hw = wi + (uint64_t)g.Lh * (wy + (uint64_t)g.L * wz);

  msa::LutView t0{reinterpret_cast<const uint2 *>(s_lut),
                  reinterpret_cast<const uint16_t *>(s_lut + 512)};
  msa::LutView t1{reinterpret_cast<const uint2 *>(s_lut + 1024),
                  reinterpret_cast<const uint16_t *>(s_lut + 1536)};
  const uint32_t tid = (uint32_t)((uint64_t)a.layer * g.V + x + (uint64_t)g.L * (y + (uint64_t)g.L * z));

  for (uint32_t r = 0; r < g.R; ++r) {
    uint64_t *Sc = a.sp + (uint64_t)(2 * r + col) * g.H;
    const uint64_t *So = a.sp + (uint64_t)(2 * r + (col ^ 1)) * g.H;
    const uint64_t s = Sc[h];
    uint64_t j1, j2, j3, j4;
    msa::local_field(s, So[hxp], So[hyp], So[hzp], So[hxm], So[hym], So[hzm], J0, J1, J2, J3, J4,
                     J5, j1, j2, j3, j4);
    const uint64_t dbase = ((uint64_t)(r * g.nc + a.layer) * 2 + col) * g.H;
    uint64_t D0, D1;
    if (a.refresh) {
      msa::refresh_demon(t0, t1, a.ctr[r], a.key[r], tid, D0, D1);
    } else {
      const ulonglong2 d = a.dm_in[dbase + h];
      D0 = d.x;
      D1 = d.y;
    }
    const uint64_t flip = msa::isi_update(j1, j2, j3, j4, D0, D1);
    Sc[h] = s ^ flip;
    a.dm_out[dbase + hw] = make_ulonglong2(D0, D1);
  }
}

struct MeasArgs {
  const uint64_t *sp;
  const uint64_t *jp;
  const ulonglong2 *dm;
  Geo g;
  int what;       // 0 unsat bonds of replica ra, 1 demon energy of ra, 2 q(ra,rb), 3 qlink(ra,rb)
  int ra, rb;
  uint32_t per_thread;
  int32_t *out;   // [64]
};

// Per-lane integer observables.  Every bond touches exactly one colour-0
// site, so the six bonds of the colour-0 sites cover the lattice once
// (the sum d_computeEne forms, main.cu:745-815).  Counts accumulate in
// bit-sliced counters and are transposed to lanes with wavefront ballots.
__global__ void __launch_bounds__(256) k_measure(MeasArgs a) {
  const Geo g = a.g;
  msa::VCount<10> cnt;
  cnt.clear();
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
  for (uint32_t k = 0; k < a.per_thread; ++k) {
    const uint64_t h = gid + k * nthr;
    if (h >= g.H) break;
    const uint32_t i = (uint32_t)(h % g.Lh);
    const uint64_t yz = h / g.Lh;
    const uint32_t y = (uint32_t)(yz % g.L), z = (uint32_t)(yz / g.L);
    if (a.what == 1) {
      for (uint32_t l = 0; l < g.nc; ++l)
        for (int c = 0; c < 2; ++c) {
          const ulonglong2 d = a.dm[((uint64_t)(a.ra * g.nc + l) * 2 + c) * g.H + h];
          cnt.add(d.x, 0);
          cnt.add(d.y, 1);
        }
      continue;
    }
    if (a.what == 2) {
      for (int c = 0; c < 2; ++c)
        cnt.add(a.sp[(uint64_t)(2 * a.ra + c) * g.H + h] ^ a.sp[(uint64_t)(2 * a.rb + c) * g.H + h]);
      continue;
    }
    const uint32_t off = (y + z) & 1u;   // colour-0 site
    const uint32_t ip = off ? msa::wrapi((int)i + 1, g.Lh) : i;
    const uint32_t im = off ? i : msa::wrapi((int)i - 1, g.Lh);
    const uint32_t yp = msa::wrapi((int)y + 1, g.L), ym = msa::wrapi((int)y - 1, g.L);
    const uint32_t zp = msa::wrapi((int)z + 1, g.L), zm = msa::wrapi((int)z - 1, g.L);
    const uint64_t row = (uint64_t)g.Lh * (y + (uint64_t)g.L * z);
    const uint64_t nb[6] = {ip + row, i + (uint64_t)g.Lh * (yp + (uint64_t)g.L * z),
                            i + (uint64_t)g.Lh * (y + (uint64_t)g.L * zp), im + row,
                            i + (uint64_t)g.Lh * (ym + (uint64_t)g.L * z),
                            i + (uint64_t)g.Lh * (y + (uint64_t)g.L * zm)};
    const uint64_t *Sa0 = a.sp + (uint64_t)(2 * a.ra) * g.H, *Sa1 = Sa0 + g.H;
    if (a.what == 0) {
      const uint64_t *J0 = a.jp, *J1 = a.jp + 3 * g.H;
      const uint64_t s = Sa0[h];
      for (int d = 0; d < 3; ++d) {
        cnt.add(s ^ J0[d * g.H + h] ^ Sa1[nb[d]]);
        cnt.add(s ^ J1[d * g.H + nb[d + 3]] ^ Sa1[nb[d + 3]]);
      }
    } else {
      const uint64_t *Sb0 = a.sp + (uint64_t)(2 * a.rb) * g.H, *Sb1 = Sb0 + g.H;
      const uint64_t sa = Sa0[h], sb = Sb0[h];
      for (int d = 0; d < 6; ++d) cnt.add((sa ^ Sa1[nb[d]]) ^ (sb ^ Sb1[nb[d]]));
    }
  }
  const uint32_t tot = cnt.wave_total();
  if (tot) atomicAdd(a.out + __lane_id(), (int32_t)tot);
}

// Counter-based construction (msa_init_philox): thread = canonical site.
__global__ void __launch_bounds__(256) k_init_philox(uint64_t *sp, uint64_t *jp, Geo g,
                                                     uint32_t k0, uint32_t k1) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  const uint32_t x = (uint32_t)(s % g.L);
  const uint64_t yz = s / g.L;
  const uint32_t y = (uint32_t)(yz % g.L), z = (uint32_t)(yz / g.L);
  const uint32_t col = (x + y + z) & 1u;
  const uint64_t h = (x >> 1) + (uint64_t)g.Lh * (y + (uint64_t)g.L * z);
  for (uint32_t d = 0; d < 3; ++d) {
    uint32_t v0 = (uint32_t)s, v1 = d, v2 = 1u, v3 = 0u;
    msa::philox10(v0, v1, v2, v3, k0, k1);
    jp[(uint64_t)(col * 3 + d) * g.H + h] = ((uint64_t)v0 << 32) | v1;
  }
  for (uint32_t r = 0; r < g.R; ++r) {
    uint32_t v0 = (uint32_t)s, v1 = r, v2 = 2u, v3 = 0u;
    msa::philox10(v0, v1, v2, v3, k0, k1);
    sp[(uint64_t)(2 * r + col) * g.H + h] = ((uint64_t)v0 << 32) | v1;
  }
}

// ----------------------------------------------------------- context ------
struct msa_ctx {
  int device = 0;
  msa_geom geom{};
  Geo g{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t *sp = nullptr, *jp = nullptr;
  ulonglong2 *dm[2] = {nullptr, nullptr};
  int cur = 0;                       // demon buffer holding the current state
  msa_lut *luts = nullptr;
  int32_t nlut = 0;
  uint4 *ctr = nullptr;
  uint32_t *key = nullptr;
  int32_t nref = 0;
  std::vector<int32_t> lut_idx;      // host copy, indexes luts
  int32_t *meas = nullptr;           // [64]
  dim3 blk{1, 1, 1}, grd{1, 1, 1};
};

static int check_geom(const msa_geom *g) {
  if (!g) return fail(MSA_ERR_ARG, "geometry is NULL");
  if (g->L < 4 || g->L > 1024 || (g->L & 1))
    return fail(MSA_ERR_ARG, "L=%d must be even and in 4..1024", g->L);
  if (g->R < 1 || g->R > 16) return fail(MSA_ERR_ARG, "R=%d must be in 1..16", g->R);
  if (g->nc < 1 || g->nc > 6) return fail(MSA_ERR_ARG, "nc=%d must be in 1..6", g->nc);
  const uint64_t V = (uint64_t)g->L * g->L * g->L;
  if ((uint64_t)g->nc * V > 0xFFFFFFFFull) return fail(MSA_ERR_ARG, "nc*V exceeds 2^32");
  return MSA_OK;
}

extern "C" int64_t msa_footprint_bytes(const msa_geom *geom) {
  if (check_geom(geom) != MSA_OK) return -1;
  const int64_t V = (int64_t)geom->L * geom->L * geom->L;
  return 8 * V * (3 + geom->R) + 2 * 16 * V * geom->R * geom->nc + 256;
}

// Block shape for the half-sweep: x spans a whole row of one colour when it
// fits a wavefront, else the largest divisor of L/2 not above 64.
static void pick_launch(msa_ctx *c) {
  const uint32_t Lh = c->g.Lh;
  uint32_t bx = Lh;
  if (bx > 64)
    for (bx = 64; bx > 1; --bx)
      if (Lh % bx == 0) break;
  uint32_t by = 256 / bx;
  if (by > c->g.L) by = c->g.L;
  if (by < 1) by = 1;
  c->blk = dim3(bx, by, 1);
  c->grd = dim3((Lh + bx - 1) / bx, (c->g.L + by - 1) / by, c->g.L);
}

extern "C" int msa_create(int device, const msa_geom *geom, msa_ctx **out) {
  if (!out) return fail(MSA_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = check_geom(geom);
  if (rc) return rc;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(MSA_ERR_ARG, "device=%d count=%d", device, ndev);
  HIPCHK(hipSetDevice(device));
  msa_ctx *c = new msa_ctx();
  c->device = device;
  c->geom = *geom;
  c->g.L = geom->L;
  c->g.Lh = geom->L / 2;
  c->g.R = geom->R;
  c->g.nc = geom->nc;
  c->g.V = (uint64_t)geom->L * geom->L * geom->L;
  c->g.H = c->g.V / 2;
  const uint64_t V = c->g.V;
  auto cleanup = [&](int code) { msa_destroy(c); return code; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return cleanup(fail(MSA_ERR_HIP, "stream/event creation failed"));
  if (hipMalloc(&c->sp, 8 * V * geom->R) != hipSuccess ||
      hipMalloc(&c->jp, 8 * V * 3) != hipSuccess ||
      hipMalloc(&c->dm[0], 16 * V * geom->R * geom->nc) != hipSuccess ||
      hipMalloc(&c->dm[1], 16 * V * geom->R * geom->nc) != hipSuccess ||
      hipMalloc(&c->meas, 64 * sizeof(int32_t)) != hipSuccess)
    return cleanup(fail(MSA_ERR_NOMEM, "device allocation of %lld bytes failed",
                        (long long)msa_footprint_bytes(geom)));
  if (hipMemset(c->dm[0], 0, 16 * V * geom->R * geom->nc) != hipSuccess ||
      hipMemset(c->sp, 0, 8 * V * geom->R) != hipSuccess ||
      hipMemset(c->jp, 0, 8 * V * 3) != hipSuccess)
    return cleanup(fail(MSA_ERR_HIP, "initial clear failed"));
  pick_launch(c);
  *out = c;
  return MSA_OK;
}

extern "C" int msa_destroy(msa_ctx *c) {
  if (!c) return MSA_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->sp);
  (void)hipFree(c->jp);
  (void)hipFree(c->dm[0]);
  (void)hipFree(c->dm[1]);
  (void)hipFree(c->luts);
  (void)hipFree(c->ctr);
  (void)hipFree(c->key);
  (void)hipFree(c->meas);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MSA_OK;
}

// canonical <-> colour-split index maps
static inline uint64_t split_index(const Geo &g, uint64_t s, uint32_t *col) {
  const uint32_t x = (uint32_t)(s % g.L);
  const uint64_t yz = s / g.L;
  const uint32_t y = (uint32_t)(yz % g.L), z = (uint32_t)(yz / g.L);
  *col = (x + y + z) & 1u;
  return (x >> 1) + (uint64_t)g.Lh * (y + (uint64_t)g.L * z);
}

extern "C" int msa_upload_state(msa_ctx *c, const uint64_t *J, const uint64_t *spin,
                                const uint64_t *d0, const uint64_t *d1) {
  if (!c || !J || !spin) return fail(MSA_ERR_ARG, "NULL argument");
  if ((d0 == nullptr) != (d1 == nullptr)) return fail(MSA_ERR_ARG, "d0 and d1 must both be given");
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint64_t V = g.V, H = g.H;
  std::vector<uint64_t> jb(3 * V), sb(g.R * V);
  for (uint64_t s = 0; s < V; ++s) {
    uint32_t col;
    const uint64_t h = split_index(g, s, &col);
    for (int d = 0; d < 3; ++d) jb[(col * 3 + d) * H + h] = J[d * V + s];
    for (uint32_t r = 0; r < g.R; ++r) sb[(2 * r + col) * H + h] = spin[r * V + s];
  }
  HIPCHK(hipMemcpyAsync(c->jp, jb.data(), 8 * 3 * V, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->sp, sb.data(), 8 * g.R * V, hipMemcpyHostToDevice, c->stream));
  c->cur = 0;
  if (d0) {
    std::vector<ulonglong2> db((uint64_t)g.R * g.nc * V);
    for (uint64_t s = 0; s < V; ++s) {
      uint32_t col;
      const uint64_t h = split_index(g, s, &col);
      for (uint32_t r = 0; r < g.R; ++r)
        for (uint32_t l = 0; l < g.nc; ++l) {
          const uint64_t src = ((uint64_t)r * g.nc + l) * V + s;
          db[(((uint64_t)r * g.nc + l) * 2 + col) * H + h] = make_ulonglong2(d0[src], d1[src]);
        }
    }
    HIPCHK(hipMemcpyAsync(c->dm[0], db.data(), 16 * db.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(hipMemsetAsync(c->dm[0], 0, 16ull * g.R * g.nc * V, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return MSA_OK;
}

extern "C" int msa_download_state(msa_ctx *c, uint64_t *spin, uint64_t *d0, uint64_t *d1) {
  if (!c) return fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint64_t V = g.V, H = g.H;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (spin) {
    std::vector<uint64_t> sb(g.R * V);
    HIPCHK(hipMemcpy(sb.data(), c->sp, 8 * g.R * V, hipMemcpyDeviceToHost));
    for (uint64_t s = 0; s < V; ++s) {
      uint32_t col;
      const uint64_t h = split_index(g, s, &col);
      for (uint32_t r = 0; r < g.R; ++r) spin[r * V + s] = sb[(2 * r + col) * H + h];
    }
  }
  if (d0 || d1) {
    std::vector<ulonglong2> db((uint64_t)g.R * g.nc * V);
    HIPCHK(hipMemcpy(db.data(), c->dm[c->cur], 16 * db.size(), hipMemcpyDeviceToHost));
    for (uint64_t s = 0; s < V; ++s) {
      uint32_t col;
      const uint64_t h = split_index(g, s, &col);
      for (uint32_t r = 0; r < g.R; ++r)
        for (uint32_t l = 0; l < g.nc; ++l) {
          const ulonglong2 v = db[(((uint64_t)r * g.nc + l) * 2 + col) * H + h];
          const uint64_t dst = ((uint64_t)r * g.nc + l) * V + s;
          if (d0) d0[dst] = v.x;
          if (d1) d1[dst] = v.y;
        }
    }
  }
  return MSA_OK;
}

extern "C" int msa_download_couplings(msa_ctx *c, uint64_t *J) {
  if (!c || !J) return fail(MSA_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint64_t V = g.V, H = g.H;
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<uint64_t> jb(3 * V);
  HIPCHK(hipMemcpy(jb.data(), c->jp, 8 * 3 * V, hipMemcpyDeviceToHost));
  for (uint64_t s = 0; s < V; ++s) {
    uint32_t col;
    const uint64_t h = split_index(g, s, &col);
    for (int d = 0; d < 3; ++d) J[d * V + s] = jb[(col * 3 + d) * H + h];
  }
  return MSA_OK;
}

extern "C" int msa_init_philox(msa_ctx *c, uint64_t seed) {
  if (!c) return fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  const uint64_t V = c->g.V;
  const uint32_t nb = (uint32_t)((V + 255) / 256);
  hipLaunchKernelGGL(k_init_philox, dim3(nb), dim3(256), 0, c->stream, c->sp, c->jp, c->g,
                     (uint32_t)seed, (uint32_t)(seed >> 32));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(c->dm[0], 0, 16ull * c->g.R * c->g.nc * V, c->stream));
  c->cur = 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

// A heat-bath table is usable when its thresholds are non-decreasing, the last
// one is 2^64-1 and every presearch entry has min <= max <= 255.
static bool lut_valid(const msa_lut &t) {
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

extern "C" int msa_set_luts(msa_ctx *c, int32_t nlut, const msa_lut *luts) {
  if (!c || !luts || nlut < 1) return fail(MSA_ERR_ARG, "need at least one LUT pair");
  for (int32_t k = 0; k < 2 * nlut; ++k)
    if (!lut_valid(luts[k])) return fail(MSA_ERR_ARG, "LUT %d is malformed", k);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(c->luts);
  c->luts = nullptr;
  c->nlut = 0;
  if (hipMalloc(&c->luts, sizeof(msa_lut) * 2 * (size_t)nlut) != hipSuccess)
    return fail(MSA_ERR_NOMEM, "LUT table of %d pairs", nlut);
  HIPCHK(hipMemcpy(c->luts, luts, sizeof(msa_lut) * 2 * (size_t)nlut, hipMemcpyHostToDevice));
  c->nlut = nlut;
  return MSA_OK;
}

extern "C" int msa_set_refresh(msa_ctx *c, int32_t nref, const msa_u32x4 *ctr,
                               const uint32_t *key, const int32_t *lut_idx) {
  if (!c || !ctr || !key || !lut_idx || nref < 1) return fail(MSA_ERR_ARG, "bad refresh table");
  for (int32_t k = 0; k < nref; ++k)
    if (lut_idx[k] < 0 || lut_idx[k] >= c->nlut)
      return fail(MSA_ERR_RANGE, "lut_idx[%d]=%d outside 0..%d (set LUTs first)", k, lut_idx[k],
                  c->nlut - 1);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(c->ctr);
  (void)hipFree(c->key);
  c->ctr = nullptr;
  c->key = nullptr;
  c->nref = 0;
  const size_t R = c->g.R;
  if (hipMalloc(&c->ctr, sizeof(uint4) * R * (size_t)nref) != hipSuccess ||
      hipMalloc(&c->key, sizeof(uint32_t) * R) != hipSuccess)
    return fail(MSA_ERR_NOMEM, "refresh table of %d entries", nref);
  HIPCHK(hipMemcpy(c->ctr, ctr, sizeof(uint4) * R * (size_t)nref, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->key, key, sizeof(uint32_t) * R, hipMemcpyHostToDevice));
  c->lut_idx.assign(lut_idx, lut_idx + nref);
  c->nref = nref;
  return MSA_OK;
}

// refresh index of iteration t, -1 when t does not refresh
static int64_t refresh_index(uint64_t t, int64_t every) {
  if (every > 0) return (t % (uint64_t)every == 0) ? (int64_t)(t / (uint64_t)every) : -1;
  if (every == -1) {
    if (t == 0) return 0;
    if (t & (t - 1)) return -1;
    return 64 - __builtin_clzll(t);   // log2(t) + 1
  }
  return -1;
}

static int launch_run(msa_ctx *c, uint64_t t0, int64_t niter, int64_t every) {
  if (niter < 0) return fail(MSA_ERR_ARG, "niter=%lld", (long long)niter);
  if (every < -1) return fail(MSA_ERR_ARG, "refresh_every=%lld", (long long)every);
  for (int64_t n = 0; n < niter; ++n) {
    const uint64_t t = t0 + (uint64_t)n;
    const int64_t k = refresh_index(t, every);
    if (k >= 0 && (k >= c->nref || c->nlut == 0))
      return fail(MSA_ERR_RANGE, "iteration %llu needs refresh %lld but %d are set",
                  (unsigned long long)t, (long long)k, c->nref);
    SweepArgs a{};
    a.sp = c->sp;
    a.jp = c->jp;
    a.dm_in = c->dm[c->cur];
    a.dm_out = c->dm[c->cur ^ 1];
    a.g = c->g;
    a.refresh = k >= 0;
    if (k >= 0) {
      a.ctr = c->ctr + (size_t)k * c->g.R;
      a.key = c->key;
      a.lut = c->luts + 2 * (size_t)c->lut_idx[k];
    }
    for (uint32_t l = 0; l < c->g.nc; ++l) {
      a.layer = (int)l;
      for (int col = 0; col < 2; ++col) {
        a.colour = col;
        hipLaunchKernelGGL(k_halfsweep, c->grd, c->blk, 0, c->stream, a);
      }
    }
    c->cur ^= 1;
  }
  HIPCHK(hipGetLastError());
  return MSA_OK;
}

extern "C" int msa_run(msa_ctx *c, uint64_t t0, int64_t niter, int64_t refresh_every) {
  if (!c) return fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  return launch_run(c, t0, niter, refresh_every);
}

extern "C" int msa_run_timed(msa_ctx *c, uint64_t t0, int64_t niter, int64_t refresh_every,
                             float *ms) {
  if (!c || !ms) return fail(MSA_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int rc = launch_run(c, t0, niter, refresh_every);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  HIPCHK(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return MSA_OK;
}

static int measure_one(msa_ctx *c, int what, int ra, int rb, int32_t *host_out) {
  const Geo &g = c->g;
  const uint32_t per_thread = 8;
  const uint64_t nthr = (g.H + per_thread - 1) / per_thread;
  const uint32_t nb = (uint32_t)((nthr + 255) / 256);
  MeasArgs a{};
  a.sp = c->sp;
  a.jp = c->jp;
  a.dm = c->dm[c->cur];
  a.g = g;
  a.what = what;
  a.ra = ra;
  a.rb = rb;
  a.per_thread = per_thread;
  a.out = c->meas;
  HIPCHK(hipMemsetAsync(c->meas, 0, 64 * sizeof(int32_t), c->stream));
  hipLaunchKernelGGL(k_measure, dim3(nb), dim3(256), 0, c->stream, a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host_out, c->meas, 64 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

extern "C" int msa_measure(msa_ctx *c, int32_t *unsat, int32_t *demon, int32_t *q,
                           int32_t *qlink) {
  if (!c) return fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  int rc;
  for (uint32_t r = 0; r < c->g.R; ++r) {
    if (unsat && (rc = measure_one(c, 0, (int)r, 0, unsat + 64 * r))) return rc;
    if (demon && (rc = measure_one(c, 1, (int)r, 0, demon + 64 * r))) return rc;
  }
  for (uint32_t p = 0; 2 * p + 1 < c->g.R; ++p) {
    if (q && (rc = measure_one(c, 2, (int)(2 * p), (int)(2 * p + 1), q + 64 * p))) return rc;
    if (qlink && (rc = measure_one(c, 3, (int)(2 * p), (int)(2 * p + 1), qlink + 64 * p))) return rc;
  }
  return MSA_OK;
}

extern "C" int msa_sync(msa_ctx *c) {
  if (!c) return fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

extern "C" const char *msa_strerror(int status) {
  if (status == MSA_OK) return "ok";
  return g_err;
}
