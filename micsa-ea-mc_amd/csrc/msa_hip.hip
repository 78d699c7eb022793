// msa_hip.hip -- kernels and C ABI of libmsa_hip.so (see include/msa.h).
//
// MI355X / gfx950 only.  Integer and bitwise work: the sweep is bound by HBM
// bandwidth, the fused demon refresh by integer VALU throughput.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "msa.h"
#include "msa_common.hpp"
#include "msa_ctx.hpp"
#include "msa_device.hpp"

using msa::Geo;

// ------------------------------------------------------------ errors ------
static thread_local char g_err[512] = "no error";

int msa_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// ----------------------------------------------------------- kernels ------
struct SweepArgs {
  const uint64_t *sp_own;   // this colour's spins before the update
  const uint64_t *sp_nb;    // the other colour's spins (old for colour 0, new for colour 1)
  uint64_t *sp_out;         // this colour's updated spins
  const uint64_t *jp;
  const ulonglong2 *dm_in;
  ulonglong2 *dm_out;
  Geo g;
  int layer, colour;
  uint32_t nr;              // rows (y, z) per tile
  uint32_t rs;              // replica groups (ONE form: R): a block is rs x nr x L/2 threads
  uint32_t mLh, mL, mG;     // div_recip reciprocals of L/2, L and nr L/2
  const uint4 *ctr;         // refresh: [R] counters of this refresh
  const uint32_t *key;      // refresh: [R]
  const msa::FineLut *lut;  // refresh: [2] plane-0 and plane-1 tables of this refresh
  const uint64_t *rnd;      // thermal: [3][rT] random words of k_gen_rndbits
  uint32_t rT;              // thermal: V/128 * R
};

enum HalfKind { kMicro = 0, kRefresh = 1, kThermal = 2 };

// floor(n / d) for n * d < 2^32 with m = floor(2^32 / d) + 1 (div_recip_of):
// n * m / 2^32 exceeds n / d by less than n / 2^32 < 1 / d, so the floor is exact.
static inline uint32_t div_recip_of(uint32_t d) { return (uint32_t)((1ull << 32) / d + 1); }
__device__ __forceinline__ uint32_t div_recip(uint32_t n, uint32_t m) { return __umulhi(n, m); }

// One colour of a sweep with demon layer `layer` (d_oneMCstepBN_multisample,
// main.cu:847-1080, ISI branch), with the demon walk of that layer fused into
// the demon store and, for kRefresh, the heat-bath refresh of the layer fused
// into the demon load.  Demons go dm_in -> dm_out and spins sp_own -> sp_out,
// so neither the walk's permutation nor the update races with a load.
// kThermal is the same kernel's Metropolis branch (main.cu:945-958) of the
// canonical thermalization: no demons; site (x, y, z) takes bit
// bx + 4 by + 16 bz of the random words of its MSC word
// (a + ML be + ML^2 g) / 2 and replica (packing_u ini.c:317-337, ML = L/4),
// the same bit for every sample, as rand_h is indexed in main.cu:877-879.
// Thread = one site of the colour and either all R replicas (J loaded once)
// or, in the ONE form, the one replica of its replica group: a block of the
// ONE form is R groups of whole wavefronts, so a small lattice still fills
// its CUs with waves (ONE with R = 1 is the plain one-replica kernel).  Blocks
// are persistent: each stages the two search tables of the refresh in LDS
// once (2 x 34 KiB) and then walks tiles of nr rows.  Used for every sweep
// with a refresh (VALU-bound: Philox and the table search), for the
// two-pass mode and for thermal sweeps.
template <int KIND, bool ONE>
__global__ void __launch_bounds__(1024) k_halfsweep(SweepArgs a) {
  constexpr bool REFRESH = KIND == kRefresh;
  extern __shared__ uint4 s_tab[];
  const Geo g = a.g;
  const msa::FineLut *tabs = reinterpret_cast<const msa::FineLut *>(s_tab);
  if (REFRESH) {
    const uint4 *src = reinterpret_cast<const uint4 *>(a.lut);
    const uint32_t n16 = 2 * sizeof(msa::FineLut) / 16;
    for (uint32_t k = threadIdx.x; k < n16; k += blockDim.x) s_tab[k] = src[k];
    __syncthreads();
  }
  const msa::LutView t0 = msa::lut_view(tabs), t1 = msa::lut_view(tabs + 1);
  const uint32_t Lh = g.Lh, L = g.L;
  // replica group: wave-uniform (groups are whole wavefronts), so replica
  // offsets stay scalar
  const uint32_t rq = ONE ? __builtin_amdgcn_readfirstlane(div_recip(threadIdx.x, a.mG)) : 0;
  const uint32_t tl = threadIdx.x - rq * (a.nr * Lh);
  const uint32_t rloc = div_recip(tl, a.mLh), i = tl - rloc * Lh;
  const uint32_t nrows = L * L, ntiles = (nrows + a.nr - 1) / a.nr;
  const int col = a.colour;
  int dx, dy, dz;
  msa::layer_dir(a.layer, dx, dy, dz);
  const uint64_t *Jc = a.jp + (uint64_t)col * 3 * g.H;
  const uint64_t *Jo = a.jp + (uint64_t)(col ^ 1) * 3 * g.H;
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t row = tile * a.nr + rloc;
    if (rloc >= a.nr || row >= nrows) continue;
    const uint32_t z = div_recip(row, a.mL), y = row - z * L;
    const uint32_t off = (y + z + (uint32_t)col) & 1u;
    const uint32_t x = 2u * i + off;
    const uint32_t ip = off ? msa::wrapi((int)i + 1, Lh) : i;      // site x+1
    const uint32_t im = off ? i : msa::wrapi((int)i - 1, Lh);      // site x-1
    const uint32_t yp = msa::wrapi((int)y + 1, L), ym = msa::wrapi((int)y - 1, L);
    const uint32_t zp = msa::wrapi((int)z + 1, L), zm = msa::wrapi((int)z - 1, L);
    const uint32_t H = (uint32_t)g.H;
    const uint32_t rb = Lh * row, h = i + rb, hxp = ip + rb, hxm = im + rb;
    const uint32_t hyp = i + Lh * (yp + L * z), hym = i + Lh * (ym + L * z);
    const uint32_t hzp = i + Lh * (y + L * zp), hzm = i + Lh * (y + L * zm);
    const uint64_t hw = msa::walk_dest(g, col, dx, dy, dz, i, y, z);
    const uint32_t tid = (uint32_t)((uint64_t)a.layer * g.V + x + (uint64_t)L * row);
    uint64_t J0, J1, J2, J3, J4, J5;
    if (KIND == kThermal) {
      const uint32_t ML = L >> 2;
      const uint32_t bx = (x >= ML) + (x >= 2 * ML) + (x >= 3 * ML);
      const uint32_t by = (y >= ML) + (y >= 2 * ML) + (y >= 3 * ML);
      const uint32_t bz = (z >= ML) + (z >= 2 * ML) + (z >= 3 * ML);
      const uint32_t ms = (x - bx * ML) + ML * ((y - by * ML) + ML * (z - bz * ML));
      const uint32_t bit = bx + 4 * by + 16 * bz;
      J0 = Jc[h]; J1 = Jc[H + h]; J2 = Jc[2 * H + h];
      J3 = Jo[hxm]; J4 = Jo[H + hym]; J5 = Jo[2 * H + hzm];
      for (uint32_t r = rq; r < (ONE ? rq + 1 : g.R); ++r) {
        const uint32_t w = (ms >> 1) * g.R + r;
        const uint64_t b0 = 0ull - ((a.rnd[w] >> bit) & 1ull);
        const uint64_t b1 = 0ull - ((a.rnd[a.rT + w] >> bit) & 1ull);
        const uint64_t b2 = 0ull - ((a.rnd[2 * (uint64_t)a.rT + w] >> bit) & 1ull);
        const uint64_t *So = a.sp_nb + (uint64_t)(2 * r + (col ^ 1)) * g.H;
        const uint64_t sown = a.sp_own[(uint64_t)(2 * r + col) * g.H + h];
        a.sp_out[(uint64_t)(2 * r + col) * g.H + h] =
            msa::metropolis_update(sown, So[hxp], So[hyp], So[hzp], So[hxm], So[hym], So[hzm], J0,
                                   J1, J2, J3, J4, J5, b0, b1, b2);
      }
      continue;
    }
    for (uint32_t r = rq; r < (ONE ? rq + 1 : g.R); ++r) {
      const uint64_t *So = a.sp_nb + (uint64_t)(2 * r + (col ^ 1)) * g.H;
      const uint64_t dbase = ((uint64_t)(r * g.nc + a.layer) * 2 + col) * g.H;
      uint64_t sown, n0, n1, n2, n3, n4, n5;
      auto load = [&]() {
        if (r == rq) {
          J0 = Jc[h]; J1 = Jc[H + h]; J2 = Jc[2 * H + h];
          J3 = Jo[hxm]; J4 = Jo[H + hym]; J5 = Jo[2 * H + hzm];
        }
        sown = a.sp_own[(uint64_t)(2 * r + col) * g.H + h];
        n0 = So[hxp]; n1 = So[hyp]; n2 = So[hzp]; n3 = So[hxm]; n4 = So[hym]; n5 = So[hzm];
      };
      uint64_t D0, D1;
      if (REFRESH) {
        // the site's loads are issued first and land while the refresh
        // (Philox and the table search, VALU-bound) runs
        load();
        D0 = msa::refresh_plane(t0, a.ctr[r], a.key[r], tid, 0);
        D1 = msa::refresh_plane(t1, a.ctr[r], a.key[r], tid, 1);
      } else {
        const ulonglong2 d = a.dm_in[dbase + h];
        D0 = d.x;
        D1 = d.y;
        load();
      }
      const uint64_t sn = msa::site_update(sown, n0, n1, n2, n3, n4, n5, J0, J1, J2, J3, J4, J5,
                                           D0, D1);
      a.sp_out[(uint64_t)(2 * r + col) * g.H + h] = sn;
      a.dm_out[dbase + hw] = make_ulonglong2(D0, D1);
    }
  }
}

// Refresh demons drawn ahead of their sweeps: the words k_halfsweep<kRefresh>
// draws in its demon load, written out as a demon buffer so that a sweep
// without a refresh reading it as its input demons performs the refresh
// sweep.  Block (bx, j) draws refresh k0 + j into out + j * R nc V: it stages
// that refresh's two search tables in LDS once, then walks the planes
// p = (r nc + l) 2 + c (wave-uniform: replica r, layer l, colour c) and, in
// each, rows of L/2 sites (thread = nr rows x L/2 sites, a grid-stride loop
// over the L^2 rows); the site's thread id is tid = l V + x + L (y + L z), as
// in the fused refresh.  Pure VALU work (Philox and the table search) and one
// streaming store per word: run on a second stream, it fills the issue slots
// the memory-bound sweeps leave idle.
constexpr int kDrawMax = 64;   // refreshes per draw launch (DrawArgs::lidx)
// dynamic LDS of a draw block: the two search tables, requested as more than
// half of a CU's 160 KiB so that a CU holds one draw block
constexpr size_t kDrawLds = 84 * 1024;
struct DrawArgs {
  ulonglong2 *out;              // [nk][R nc 2 H]
  const uint4 *ctr;             // counters of refresh k0: ctr[j R + r] is refresh k0 + j
  const uint32_t *key;          // [R]
  const msa::FineLut *luts;     // search tables, pair n at luts + 2 n
  Geo g;
  uint32_t nr;                  // rows per block
  uint32_t mLh, mL;             // div_recip reciprocals of L/2 and L
  int32_t lidx[kDrawMax];       // LUT pair of each refresh of the launch
};

__global__ void __launch_bounds__(768) k_refresh_draw(DrawArgs a) {
  extern __shared__ uint4 s_tab[];
  const Geo g = a.g;
  const uint32_t j = blockIdx.y;
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(a.luts + 2 * (size_t)a.lidx[j]);
    const uint32_t n16 = 2 * sizeof(msa::FineLut) / 16;
    for (uint32_t k = threadIdx.x; k < n16; k += blockDim.x) s_tab[k] = src[k];
    __syncthreads();
  }
  const msa::FineLut *tabs = reinterpret_cast<const msa::FineLut *>(s_tab);
  const msa::LutView t0 = msa::lut_view(tabs), t1 = msa::lut_view(tabs + 1);
  const uint32_t L = g.L, Lh = g.Lh, nrows = L * L;
  const uint32_t rloc = div_recip(threadIdx.x, a.mLh), i = threadIdx.x - rloc * Lh;
  if (rloc >= a.nr) return;
  const uint32_t np = 2 * g.R * g.nc;
  ulonglong2 *out = a.out + (uint64_t)j * np * g.H;
  for (uint32_t p = 0; p < np; ++p) {
    const uint32_t col = p & 1u, l = (p >> 1) % g.nc, r = (p >> 1) / g.nc;
    const uint4 c4 = a.ctr[(size_t)j * g.R + r];
    const uint32_t kr = a.key[r];
    ulonglong2 *o = out + (uint64_t)p * g.H;
    for (uint32_t row = blockIdx.x * a.nr + rloc; row < nrows; row += gridDim.x * a.nr) {
      const uint32_t z = div_recip(row, a.mL), y = row - z * L;
      const uint32_t x = 2u * i + ((y + z + col) & 1u);
      const uint32_t tid = (uint32_t)((uint64_t)l * g.V + x + (uint64_t)L * row);
      const uint64_t D0 = msa::refresh_plane(t0, c4, kr, tid, 0);
      const uint64_t D1 = msa::refresh_plane(t1, c4, kr, tid, 1);
      const msa::u64x2 v = {D0, D1};
      __builtin_nontemporal_store(v, reinterpret_cast<msa::u64x2 *>(o + i + (uint64_t)Lh * row));
    }
  }
}

// Search tables of heat-bath tables first .. ntab-1 (msa::FineLut), table
// first + blockIdx.y: thresholds copied, presearch entry b = the byte of the
// first R of bucket b, found by binary search for the smallest n with
// R < umbral[n] (255 when none), and the sentinel 255 after the last bucket.
__global__ void __launch_bounds__(256) k_fine_luts(const msa_lut *luts, uint32_t first,
                                                   uint32_t ntab, msa::FineLut *out) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t tab = first + blockIdx.y;
  constexpr uint32_t nb = 1u << msa::kPreBits;
  if (b >= nb + 16 || tab >= ntab) return;
  if (b >= nb) {
    out[tab].pre[b] = 255;
    return;
  }
  const uint32_t *u = luts[tab].umbrales;
  const uint64_t R = (uint64_t)b << (64 - msa::kPreBits);
  uint32_t lo = 0, hi = 255;
  while (lo < hi) {
    const uint32_t m = (lo + hi) >> 1;
    const uint64_t t = ((uint64_t)u[2 * m] << 32) | u[2 * m + 1];
    if (R < t) hi = m; else lo = m + 1;
  }
  out[tab].pre[b] = (uint8_t)lo;
  if (b < 256) out[tab].umb[b] = ((uint64_t)u[2 * b] << 32) | u[2 * b + 1];
}

// d_gen_rndbits_cuda (main.cu:647-702) with NBETAS = 1: thread tid < T =
// (V/128) R, replica tid % R (main.cu:664-665), counter ctr[ir] with
// .z += tid, key {tid, key[ir]}, six words from the heat-bath table:
// out[k T + tid] = w_{2k} << 32 | w_{2k+1} (main.cu:683-700).
__global__ void __launch_bounds__(256) k_gen_rndbits(const uint4 *ctr, const uint32_t *key,
                                                     const msa::FineLut *lut, uint32_t R,
                                                     uint32_t T, uint64_t *out) {
  __shared__ uint4 s_tab[sizeof(msa::FineLut) / 16];
  const uint4 *src = reinterpret_cast<const uint4 *>(lut);
  for (uint32_t k = threadIdx.x; k < sizeof(msa::FineLut) / 16; k += blockDim.x) s_tab[k] = src[k];
  __syncthreads();
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= T) return;
  const msa::LutView t = msa::lut_view(reinterpret_cast<const msa::FineLut *>(s_tab));
  const uint32_t ir = tid % R;
  const uint4 c4 = ctr[ir];
  uint32_t c[4] = {c4.x, c4.y, c4.z + tid, c4.w};
  uint64_t w[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) w[j] = msa::rnd_word(t, c, tid, key[ir]);
  out[tid] = (w[0] << 32) | w[1];
  out[T + tid] = (w[2] << 32) | w[3];
  out[2 * (uint64_t)T + tid] = (w[4] << 32) | w[5];
}

// Upload `ntab` heat-bath tables to raw (device) and build their search
// tables in out (device); asynchronous on `st`.
static int fine_luts_into(hipStream_t st, size_t ntab, const msa_lut *host, msa_lut *raw,
                          msa::FineLut *out) {
  if (hipMemcpyAsync(raw, host, sizeof(msa_lut) * ntab, hipMemcpyHostToDevice, st) != hipSuccess)
    return msa_fail(MSA_ERR_HIP, "LUT upload");
  // gridDim.y is at most 65535: launches of up to that many tables
  constexpr uint32_t kMaxY = 65535;
  for (size_t first = 0; first < ntab; first += kMaxY) {
    const uint32_t n = (uint32_t)(ntab - first < kMaxY ? ntab - first : kMaxY);
    hipLaunchKernelGGL(k_fine_luts, dim3((1u << msa::kPreBits) / 256 + 1, n), dim3(256), 0, st,
                       raw, (uint32_t)first, (uint32_t)ntab, out);
  }
  if (hipGetLastError() != hipSuccess) return msa_fail(MSA_ERR_HIP, "search table construction");
  return MSA_OK;
}

int msa_build_fine_luts(hipStream_t st, int32_t nlut, const msa_lut *host, msa::FineLut **out) {
  *out = nullptr;
  msa_lut *raw = nullptr;
  const size_t ntab = 2 * (size_t)nlut;
  if (hipMalloc(&raw, sizeof(msa_lut) * ntab) != hipSuccess)
    return msa_fail(MSA_ERR_NOMEM, "LUT upload of %d pairs", nlut);
  if (hipMalloc(out, sizeof(msa::FineLut) * ntab) != hipSuccess) {
    (void)hipFree(raw);
    return msa_fail(MSA_ERR_NOMEM, "search tables of %d pairs", nlut);
  }
  int rc = fine_luts_into(st, ntab, host, raw, *out);
  if (!rc && hipStreamSynchronize(st) != hipSuccess)
    rc = msa_fail(MSA_ERR_HIP, "search table construction");
  (void)hipFree(raw);
  if (rc) {
    (void)hipFree(*out);
    *out = nullptr;
  }
  return rc;
}

struct FusedArgs {
  const uint64_t *sp_in;
  uint64_t *sp_out;
  const uint64_t *jp;
  const ulonglong2 *dm_in;
  ulonglong2 *dm_out;
  Geo g;
  int layer;
  uint32_t ty, tz;   // rows per tile, planes per slab
  uint32_t ny;       // thread rows per block: ty + 2 (halo rows) or L (whole planes)
  int yhalo;
  int dm_store;      // 0: the walked demons are not written (the next sweep refreshes them)
};

// A whole sweep (both colours, one demon layer, walk fused) in one pass.
// Block = all L/2 site pairs of ny rows; it owns ty rows x tz planes and
// walks its slab plane by plane: at step s it updates the colour-0 sites of
// plane s, then the colour-1 sites of plane s-1, whose six colour-0
// neighbours (planes s-2, s-1, s) are all new by then.  Colour-0 values of
// the rows and planes just outside the tile are recomputed, not written
// (their owners write them), so blocks never wait for each other.  Spins and
// demons are read from the *_in buffers and written to the *_out buffers, so
// every global load sees the state before the sweep; J is read once per sweep
// from HBM, its reuse between the two colours comes from the caches.
// New colour-0 values of one plane are exchanged through LDS, three plane
// buffers deep so one barrier per step orders writer and readers.
template <int R>
__global__ void __launch_bounds__(1024) k_sweep_fused(FusedArgs a) {
  extern __shared__ uint64_t s_e[];
  const Geo g = a.g;
  const uint32_t Lh = g.Lh, L = g.L;
  const uint64_t H = g.H;
  const uint32_t i = threadIdx.x % Lh, row = threadIdx.x / Lh;
  const uint32_t y0 = blockIdx.x * a.ty, z0 = blockIdx.y * a.tz;
  const bool yh = a.yhalo != 0;
  const uint32_t y = yh ? msa::wrapi((int)(y0 + row) - 1, L) : row;
  const bool own = yh ? (row >= 1 && row <= a.ty) : true;
  const uint32_t rm = yh ? row - 1 : (row == 0 ? L - 1 : row - 1);
  const uint32_t rp = yh ? row + 1 : (row + 1 == L ? 0 : row + 1);
  const uint32_t ym = msa::wrapi((int)y - 1, L), yp = msa::wrapi((int)y + 1, L);
  const uint64_t pstride = (uint64_t)a.ny * Lh * R;
  int dx, dy, dz;
  msa::layer_dir(a.layer, dx, dy, dz);
  const uint64_t *J = a.jp;
  const uint64_t lbase0 = (uint64_t)(a.layer * 2 + 0) * H, lbase1 = (uint64_t)(a.layer * 2 + 1) * H;
  const uint64_t rstride = (uint64_t)g.nc * 2 * H;   // demon stride between replicas

  uint64_t Em2[R], Em1[R], E0[R];
#pragma unroll
  for (int r = 0; r < R; ++r) Em2[r] = Em1[r] = E0[r] = 0;

  const int nstep = (int)a.tz + 2;
  for (int k = 0; k < nstep; ++k) {
    const int s = (int)z0 - 1 + k;
    // ---------------- colour 0, plane z
    {
      const uint32_t z = msa::wrapi(s, L);
      const uint32_t off = (y + z) & 1u;
      const uint32_t ip = off ? msa::wrapi((int)i + 1, Lh) : i, im = off ? i : msa::wrapi((int)i - 1, Lh);
      const uint32_t zp = msa::wrapi((int)z + 1, L), zm = msa::wrapi((int)z - 1, L);
      const uint64_t rb = (uint64_t)Lh * (y + (uint64_t)L * z);
      const uint64_t h = i + rb;
      const uint64_t hxp = ip + rb, hxm = im + rb;
      const uint64_t hyp = i + (uint64_t)Lh * (yp + (uint64_t)L * z);
      const uint64_t hym = i + (uint64_t)Lh * (ym + (uint64_t)L * z);
      const uint64_t hzp = i + (uint64_t)Lh * (y + (uint64_t)L * zp);
      const uint64_t hzm = i + (uint64_t)Lh * (y + (uint64_t)L * zm);
      const uint64_t J0 = J[h], J1 = J[H + h], J2 = J[2 * H + h];
      const uint64_t J3 = J[3 * H + hxm], J4 = J[4 * H + hym], J5 = J[5 * H + hzm];
      const bool write = own && k >= 1 && k <= (int)a.tz;
      const uint64_t hw = msa::walk_dest(g, 0, dx, dy, dz, i, y, z);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint64_t *S0 = a.sp_in + (uint64_t)(2 * r) * H, *S1 = S0 + H;
        const ulonglong2 d = a.dm_in[r * rstride + lbase0 + h];
        uint64_t D0 = d.x, D1 = d.y;
        E0[r] = msa::site_update(S0[h], S1[hxp], S1[hyp], S1[hzp], S1[hxm], S1[hym], S1[hzm],
                                 J0, J1, J2, J3, J4, J5, D0, D1);
        s_e[(k % 3) * pstride + ((uint64_t)row * Lh + i) * R + r] = E0[r];
        if (write) {
          a.sp_out[(uint64_t)(2 * r) * H + h] = E0[r];
          a.dm_out[r * rstride + lbase0 + hw] = make_ulonglong2(D0, D1);
        }
      }
    }
    __syncthreads();
    // ---------------- colour 1, plane z-1
    if (own && k >= 2) {
      const uint32_t z = msa::wrapi(s - 1, L);
      const uint32_t off = (y + z + 1) & 1u;
      const uint32_t ip = off ? msa::wrapi((int)i + 1, Lh) : i, im = off ? i : msa::wrapi((int)i - 1, Lh);
      const uint32_t zm = msa::wrapi((int)z - 1, L);
      const uint64_t rb = (uint64_t)Lh * (y + (uint64_t)L * z);
      const uint64_t h = i + rb;
      const uint64_t hxm = im + rb;
      const uint64_t hym = i + (uint64_t)Lh * (ym + (uint64_t)L * z);
      const uint64_t hzm = i + (uint64_t)Lh * (y + (uint64_t)L * zm);
      const uint64_t J0 = J[3 * H + h], J1 = J[4 * H + h], J2 = J[5 * H + h];
      const uint64_t J3 = J[hxm], J4 = J[H + hym], J5 = J[2 * H + hzm];
      const uint64_t hw = msa::walk_dest(g, 1, dx, dy, dz, i, y, z);
      const uint64_t *Eb = s_e + ((k - 1) % 3) * pstride;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint64_t *S1 = a.sp_in + (uint64_t)(2 * r + 1) * H;
        const ulonglong2 d = a.dm_in[r * rstride + lbase1 + h];
        uint64_t D0 = d.x, D1 = d.y;
        const uint64_t nxp = Eb[((uint64_t)row * Lh + ip) * R + r];
        const uint64_t nxm = Eb[((uint64_t)row * Lh + im) * R + r];
        const uint64_t nyp = Eb[((uint64_t)rp * Lh + i) * R + r];
        const uint64_t nym = Eb[((uint64_t)rm * Lh + i) * R + r];
        const uint64_t sn = msa::site_update(S1[h], nxp, nyp, E0[r], nxm, nym, Em2[r],
                                             J0, J1, J2, J3, J4, J5, D0, D1);
        a.sp_out[(uint64_t)(2 * r + 1) * H + h] = sn;
        a.dm_out[r * rstride + lbase1 + hw] = make_ulonglong2(D0, D1);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      Em2[r] = Em1[r];
      Em1[r] = E0[r];
    }
  }
}

// Tile index of a block of the one-pass kernels.  Blocks are dispatched to
// the 8 XCDs round-robin (block b on XCD b mod 8); with `xcd` set, XCD x gets
// the x-th contiguous run of tiles in (plane slab, row tile) order, so tiles
// that share halo rows sit on the same XCD and meet in its L2.
__device__ __forceinline__ void tile_of(uint32_t b, uint32_t nty, uint32_t ntz, int xcd,
                                        uint32_t &ty_idx, uint32_t &tz_idx) {
  const uint32_t n = nty * ntz;
  uint32_t t = b;
  if (xcd && (n % 8u) == 0) t = (b % 8u) * (n / 8u) + b / 8u;
  ty_idx = t % nty;
  tz_idx = t / nty;
}

// One sweep in one pass, second form.  Same wavefront as k_sweep_fused, but
// every word a block needs is loaded from global memory once per block:
//   * odd-colour spins of its own column ride a three-plane register window
//     (z-1, z, z+1) and reach in-plane neighbours through one LDS plane;
//   * the thread's own J words of a plane stay in registers for the later
//     steps that use them (col-1 J for the odd update one step later,
//     col-0 J_z for the -z bond two steps later);
//   * new colour-0 values are exchanged through a double LDS buffer.
// Two barriers per step; the loads of each colour are issued right after the
// barrier that precedes it (all waves of the block then fetch the same lines
// together).  Demons are read and spins and demons written with the streaming
// hint: each is touched once per sweep, while spins and J are re-read by the
// neighbouring tiles' halos.  Tiles are placed XCD-aware (tile_of).  Offsets
// are 32-bit byte offsets from uniform bases (every buffer below 4 GiB,
// checked by the host).  At R = 4 the LDS bound (three planes in 64 KiB)
// already limits a block to 682 threads, so the launch bound is lowered to
// leave more registers.
template <int R>
__global__ void __launch_bounds__(R >= 4 ? 768 : 1024) k_sweep_fused2(FusedArgs a) {
  using msa::ldo;
  using msa::ldo_stream;
  using msa::sto_stream;
  extern __shared__ uint64_t s_lds[];
  // ahead of the refresh draws that may share the CU (they fill the issue
  // slots this latency-bound kernel leaves)
  __builtin_amdgcn_s_setprio(2);
  const Geo g = a.g;
  const uint32_t Lh = g.Lh, L = g.L;
  const uint64_t H = g.H;
  const uint32_t i = threadIdx.x % Lh, row = threadIdx.x / Lh;
  uint32_t tyi, tzi;
  tile_of(blockIdx.x, L / a.ty, L / a.tz, 1, tyi, tzi);
  const uint32_t y0 = tyi * a.ty, z0 = tzi * a.tz;
  const bool yh = a.yhalo != 0;
  const uint32_t y = yh ? msa::wrapi((int)(y0 + row) - 1, L) : row;
  const bool own = yh ? (row >= 1 && row <= a.ty) : true;
  const uint32_t rm = yh ? (row == 0 ? 0 : row - 1) : (row == 0 ? L - 1 : row - 1);
  const uint32_t rp = yh ? (row + 1 == a.ny ? row : row + 1) : (row + 1 == L ? 0 : row + 1);
  const uint32_t ym = msa::wrapi((int)y - 1, L), yp = msa::wrapi((int)y + 1, L);
  // LDS planes hold one sub-plane per replica, so the 8-byte words that a
  // wavefront reads for one replica are contiguous (no bank conflicts)
  const uint32_t pr = a.ny * Lh;                         // one replica's sub-plane
  const uint32_t slot = row * Lh + i;                    // this thread's LDS slot
  const uint32_t plane = pr * R;                         // one LDS plane
  uint64_t *S1p = s_lds;                                 // odd spins, plane s
  uint64_t *Eb = s_lds + plane;                          // new even, planes s-1 / s
  int dx, dy, dz;
  msa::layer_dir(a.layer, dx, dy, dz);
  const uint64_t *Jx0 = a.jp, *Jy0 = a.jp + H, *Jz0 = a.jp + 2 * H;
  const uint64_t *Jx1 = a.jp + 3 * H, *Jy1 = a.jp + 4 * H, *Jz1 = a.jp + 5 * H;
  const uint64_t rstride = (uint64_t)g.nc * 2 * H;
  const ulonglong2 *Din0 = a.dm_in + (uint64_t)(a.layer * 2) * H, *Din1 = Din0 + H;
  ulonglong2 *Dout0 = a.dm_out + (uint64_t)(a.layer * 2) * H, *Dout1 = Dout0 + H;
  const uint32_t LhL = Lh * L;
  const uint32_t yb = Lh * y, ymb = Lh * ym, ypb = Lh * yp;
  const bool hrow = yh && (row == 0 || row + 1 == a.ny);   // a halo row: one y neighbour outside
  const uint32_t hrb = row == 0 ? ymb : ypb;

  uint64_t Sm[R], S0w[R], Sp[R], Em2[R], Em1[R], E0[R];
  // prime: odd spins of planes z0-2, z0-1; col-1 J_z of plane z0-2
  const uint32_t hza = i + yb + LhL * msa::wrapi((int)z0 - 2, L);
  const uint32_t hzb = i + yb + LhL * msa::wrapi((int)z0 - 1, L);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint64_t *S1 = a.sp_in + (uint64_t)(2 * r + 1) * H;
    Sm[r] = ldo(S1, hza * 8u);
    S0w[r] = ldo(S1, hzb * 8u);
    Em2[r] = Em1[r] = E0[r] = 0;
  }
  uint64_t j1z_prev = ldo(Jz1, hza * 8u);
  uint64_t j1x_prev = 0, j1y_prev = 0;       // col-1 J of plane s-1
  uint64_t j0z_m1 = 0, j0z_m2 = 0;           // col-0 J_z of planes s-1, s-2

  const int nstep = (int)a.tz + 2;
  for (int k = 0; k < nstep; ++k) {
    const int s = (int)z0 - 1 + k;
    const uint32_t z = msa::wrapi(s, L), zo = msa::wrapi(s - 1, L), zn = msa::wrapi(s + 1, L);
    const uint32_t h = i + yb + LhL * z, ho = i + yb + LhL * zo, hn = i + yb + LhL * zn;
    // x parity of the colour-0 sites of plane s, (y + s) & 1, equals that of
    // the colour-1 sites of plane s-1, (y + s - 1 + 1) & 1 (L is even), so
    // both updates of a step share the x neighbours ip / im
    const uint32_t off = (y + z) & 1u;
    const uint32_t ip = off ? msa::wrapi((int)i + 1, Lh) : i, im = off ? i : msa::wrapi((int)i - 1, Lh);
    // ---- the step's own column: odd spins of plane s+1, J of plane s
#pragma unroll
    for (int r = 0; r < R; ++r) Sp[r] = ldo(a.sp_in + (uint64_t)(2 * r + 1) * H, hn * 8u);
    const uint64_t j0x = ldo(Jx0, h * 8u), j0y = ldo(Jy0, h * 8u), j0z = ldo(Jz0, h * 8u);
    const uint64_t j1x = ldo(Jx1, h * 8u), j1y = ldo(Jy1, h * 8u), j1z = ldo(Jz1, h * 8u);
    const uint64_t c0J3 = ldo(Jx1, (im + yb + LhL * z) * 8u), c0J4 = ldo(Jy1, (i + ymb + LhL * z) * 8u);
#pragma unroll
    for (int r = 0; r < R; ++r) S1p[r * pr + slot] = S0w[r];
    __syncthreads();
    // ---- colour 0, plane s
    {
      uint64_t S0v[R], Hv[R];
      ulonglong2 dv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        S0v[r] = ldo(a.sp_in + (uint64_t)(2 * r) * H, h * 8u);
        dv[r] = ldo_stream(Din0 + r * rstride, h * 16u);
        Hv[r] = hrow ? ldo(a.sp_in + (uint64_t)(2 * r + 1) * H, (i + hrb + LhL * z) * 8u) : 0;
      }
      const bool write = own && k >= 1 && k <= (int)a.tz;
      const uint32_t hw = (uint32_t)msa::walk_dest(g, 0, dx, dy, dz, i, y, z);
      uint64_t *Ew = Eb + (k & 1) * plane;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        uint64_t D0 = dv[r].x, D1 = dv[r].y;
        const uint64_t nxp = S1p[r * pr + row * Lh + ip];
        const uint64_t nxm = S1p[r * pr + row * Lh + im];
        const uint64_t nyp = yh && row + 1 == a.ny ? Hv[r] : S1p[r * pr + rp * Lh + i];
        const uint64_t nym = yh && row == 0 ? Hv[r] : S1p[r * pr + rm * Lh + i];
        E0[r] = msa::site_update(S0v[r], nxp, nyp, Sp[r], nxm, nym, Sm[r], j0x, j0y, j0z, c0J3,
                                 c0J4, j1z_prev, D0, D1);
        Ew[r * pr + slot] = E0[r];
        if (write) {
          sto_stream(a.sp_out + (uint64_t)(2 * r) * H, h * 8u, E0[r]);
          if (a.dm_store) sto_stream(Dout0 + r * rstride, hw * 16u, D0, D1);
        }
      }
    }
    __syncthreads();
    // ---- colour 1, plane s-1
    if (own && k >= 2) {
      const uint64_t c1J3 = ldo(Jx0, (im + yb + LhL * zo) * 8u);
      const uint64_t c1J4 = ldo(Jy0, (i + ymb + LhL * zo) * 8u);
      ulonglong2 dv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) dv[r] = ldo_stream(Din1 + r * rstride, ho * 16u);
      const uint32_t hw = (uint32_t)msa::walk_dest(g, 1, dx, dy, dz, i, y, zo);
      const uint64_t *Er = Eb + ((k - 1) & 1) * plane;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        uint64_t D0 = dv[r].x, D1 = dv[r].y;
        const uint64_t nxp = Er[r * pr + row * Lh + ip];
        const uint64_t nxm = Er[r * pr + row * Lh + im];
        const uint64_t nyp = Er[r * pr + rp * Lh + i];
        const uint64_t nym = Er[r * pr + rm * Lh + i];
        const uint64_t sn = msa::site_update(Sm[r], nxp, nyp, E0[r], nxm, nym, Em2[r], j1x_prev,
                                             j1y_prev, j1z_prev, c1J3, c1J4, j0z_m2, D0, D1);
        sto_stream(a.sp_out + (uint64_t)(2 * r + 1) * H, ho * 8u, sn);
        if (a.dm_store) sto_stream(Dout1 + r * rstride, hw * 16u, D0, D1);
      }
    }
    // ---- rotate the windows
#pragma unroll
    for (int r = 0; r < R; ++r) {
      Sm[r] = S0w[r];
      S0w[r] = Sp[r];
      Em2[r] = Em1[r];
      Em1[r] = E0[r];
    }
    j1x_prev = j1x;
    j1y_prev = j1y;
    j1z_prev = j1z;
    j0z_m2 = j0z_m1;
    j0z_m1 = j0z;
  }
}

struct PairArgs {
  const uint64_t *sp_in;
  uint64_t *sp_out;
  const uint64_t *jp;
  const ulonglong2 *dm_in;
  const ulonglong2 *dm_in2;   // D2: the second iteration's demons (drawn for its refresh)
  ulonglong2 *dm_out;
  Geo g;
  uint32_t ty, tz;   // own rows, own planes
  uint32_t ny;       // thread rows: ty + 6
  int dm_store;      // 0: the walked demons are not written (the next sweep refreshes them)
};

// Block b of the sweep-pair kernel -> (row tile, plane slab, replica).  XCD x
// (blocks b = x mod 8) gets the x-th contiguous run of (tile, replica) pairs,
// the replicas of a tile next to each other, so the blocks that share halo
// rows and J words run on one XCD at the same time.
__device__ __forceinline__ void tile_of_pair(uint32_t b, uint32_t nty, uint32_t ntiles, uint32_t R,
                                             uint32_t &tyi, uint32_t &tzi, uint32_t &r) {
  const uint32_t n = ntiles * R;
  const uint32_t t = (n % 8u) == 0 ? (b % 8u) * (n / 8u) + b / 8u : b;
  r = t % R;
  const uint32_t tile = t / R;
  tyi = tile % nty;
  tzi = tile / nty;
}

// Two consecutive sweeps of one replica, demon layer 0 (nc = 1), in one pass
// (temporal blocking): the lattice state is read and written once per two
// sweeps.  The z wavefront of k_sweep_fused2 carries four phases, step k at
// plane s = z0 - 3 + k:
//   A0  first sweep, colour 0, plane s      (all ny rows)
//   A1  first sweep, colour 1, plane s - 1  (rows 1 .. ny-2)
//   B0  second sweep, colour 0, plane s - 2 (rows 2 .. ny-3)
//   B1  second sweep, colour 1, plane s - 3 (rows 3 .. ny-4: the tile's own rows)
// Each phase needs the previous phase's values of its plane (in-plane
// neighbours, from LDS, written one step earlier) and of the planes above and
// below (the thread's own column, in registers).  Only the second sweep's
// results of the own rows and planes are written.  The first sweep's demons
// walk (layer 0: colour 0 by +1, colour 1 by -1 in the split index) to the
// neighbouring thread of the same row: they are written to a three-slot LDS
// ring as they are produced and read by the second sweep two steps later.
// J is loaded once, by the first sweep: every J word the second sweep needs
// is one the same thread loaded two to four steps earlier, kept in register
// windows.  The spin and result planes are double-buffered, so one barrier
// per step orders all writers and readers.
// A step's own colour-0 spin, halo spin and demons are loaded during the step
// before, right after the phase that last used their registers, so they are
// in flight through three phases and the barrier; the other loads are issued
// together before the barrier.  The demons are read with ordinary loads (a
// halo row's demons are read again by the neighbouring tile, which runs on the
// same XCD) and the results are written with the non-temporal hint.
// D2: the second iteration refreshes its demons -- they come from dm_in2 (the
// refresh's demons drawn ahead) instead of the first iteration's walked ones,
// each loaded during the step before the one that uses it; the walked-demon
// rings are not needed (88 instead of 184 bytes of LDS per thread) and blocks
// are at most 512 threads, so the kernel has room for the two more loads in
// flight.
template <bool D2>
__global__ void __launch_bounds__(D2 ? 512 : 1024) k_sweep_pair(PairArgs a) {
  using msa::ldo;
  using msa::ldo_stream;
  using msa::sto_stream;
  using msa::wrapi;
  extern __shared__ uint64_t s_lds[];
  // ahead of the refresh draws that may share the CU (they fill the issue
  // slots this latency-bound kernel leaves)
  __builtin_amdgcn_s_setprio(2);
  const Geo g = a.g;
  const uint32_t Lh = g.Lh, L = g.L, ny = a.ny;
  const uint64_t H = g.H;
  const uint32_t i = threadIdx.x % Lh, row = threadIdx.x / Lh;
  uint32_t tyi, tzi, r;
  tile_of_pair(blockIdx.x, L / a.ty, (L / a.ty) * (L / a.tz), g.R, tyi, tzi, r);
  const uint32_t y0 = tyi * a.ty, z0 = tzi * a.tz;
  const uint32_t y = wrapi((int)(y0 + row) - 3, L);
  const uint32_t ym = wrapi((int)y - 1, L), yp = wrapi((int)y + 1, L);
  const uint32_t rm = row == 0 ? 0 : row - 1, rp = row + 1 == ny ? row : row + 1;
  const bool rowA1 = row >= 1 && row + 1 < ny;
  const bool rowB0 = row >= 2 && row + 2 < ny;
  const bool rowB1 = row >= 3 && row + 3 < ny;
  const bool hrow = row == 0 || row + 1 == ny;
  const uint32_t P = ny * Lh, slot = row * Lh + i;
  uint64_t *S1p = s_lds, *A0b = s_lds + 2 * P, *A1b = s_lds + 5 * P, *B0b = s_lds + 8 * P;
  ulonglong2 *XD0 = reinterpret_cast<ulonglong2 *>(s_lds + 11 * P), *XD1 = XD0 + 3 * P;
  const uint64_t *S0in = a.sp_in + (uint64_t)(2 * r) * H, *S1in = S0in + H;
  uint64_t *S0out = a.sp_out + (uint64_t)(2 * r) * H, *S1out = S0out + H;
  const ulonglong2 *Din0 = a.dm_in + (uint64_t)(2 * r) * H, *Din1 = Din0 + H;
  ulonglong2 *Dout0 = a.dm_out + (uint64_t)(2 * r) * H, *Dout1 = Dout0 + H;
  const ulonglong2 *Din20 = D2 ? a.dm_in2 + (uint64_t)(2 * r) * H : nullptr;
  const ulonglong2 *Din21 = D2 ? Din20 + H : nullptr;
  const uint64_t *Jx0 = a.jp, *Jy0 = a.jp + H, *Jz0 = a.jp + 2 * H;
  const uint64_t *Jx1 = a.jp + 3 * H, *Jy1 = a.jp + 4 * H, *Jz1 = a.jp + 5 * H;
  const uint32_t LhL = Lh * L;
  const uint32_t yb = Lh * y, ymb = Lh * ym, hrb = row == 0 ? Lh * ym : Lh * yp;
  const uint32_t iw0 = i + 1 == Lh ? 0 : i + 1, iw1 = i == 0 ? Lh - 1 : i - 1;

  // odd input spins of the own column, planes s-1, s, s+1
  uint64_t Sm = ldo(S1in, (i + yb + LhL * wrapi((int)z0 - 4, L)) * 8u);
  uint64_t S0w = ldo(S1in, (i + yb + LhL * wrapi((int)z0 - 3, L)) * 8u);
  uint64_t Sp = 0;
  // loaded one step ahead, into the registers the step before has just
  // finished with: A0's own spin, its halo spin and A0's and A1's demons
  uint64_t S0v = ldo(S0in, (i + yb + LhL * wrapi((int)z0 - 3, L)) * 8u);
  uint64_t Hv = hrow ? ldo(S1in, (i + hrb + LhL * wrapi((int)z0 - 3, L)) * 8u) : 0;
  uint64_t E0 = 0, F0 = 0, G0 = 0;        // A0, A1, B0 results of the step (planes s, s-1, s-2)
  // J windows, [d] = loaded d+1 steps ago: the first sweep's own bonds
  // (colour 0 at plane s-1-d: x, y, z; colour 1: x, y, z), the colour-1 bonds
  // of A0's -x and -y neighbours and the colour-0 bonds of A1's
  uint64_t w0x[2] = {0, 0}, w0y[2] = {0, 0}, w0z[4] = {0, 0, 0, 0};
  uint64_t w1x[3] = {0, 0, 0}, w1y[3] = {0, 0, 0}, w1z[3] = {0, 0, 0};
  uint64_t wc0x[2] = {0, 0}, wc0y[2] = {0, 0}, wc1x[2] = {0, 0}, wc1y[2] = {0, 0};
  w1z[0] = ldo(Jz1, (i + yb + LhL * wrapi((int)z0 - 4, L)) * 8u);
  ulonglong2 dA0 = ldo(Din0, (i + yb + LhL * wrapi((int)z0 - 3, L)) * 16u);
  ulonglong2 dA1 = make_ulonglong2(0, 0);
  ulonglong2 dB0 = make_ulonglong2(0, 0), dB1 = make_ulonglong2(0, 0);   // D2
  // the prime loads complete before the loop, so that the loop's only wait
  // is the one after each barrier
  msa::wait_vmem();

  const int nstep = (int)a.tz + 6;
  for (int k = 0; k < nstep; ++k) {
    const int s = (int)z0 - 3 + k;
    const uint32_t z = wrapi(s, L), z1 = wrapi(s - 1, L), z2 = wrapi(s - 2, L), z3 = wrapi(s - 3, L);
    const uint32_t h = i + yb + LhL * z;
    const uint32_t h2 = i + yb + LhL * z2, h3 = i + yb + LhL * z3;
    const uint32_t off = (y + z) & 1u;
    const uint32_t ip = off ? (i + 1 == Lh ? 0 : i + 1) : i, im = off ? i : (i == 0 ? Lh - 1 : i - 1);
    const int cur = k & 1;
    // three-slot rings: the results of steps k, k-1 and k-2
    const uint32_t ring = (uint32_t)(k % 3), ring1 = (uint32_t)((k + 2) % 3), ring2 = (uint32_t)((k + 1) % 3);
    // ---- global loads of the step and the LDS write before the barrier
    Sp = ldo(S1in, (i + yb + LhL * wrapi(s + 1, L)) * 8u);
    const uint64_t j0x = ldo(Jx0, h * 8u), j0y = ldo(Jy0, h * 8u), j0z = ldo(Jz0, h * 8u);
    const uint64_t j1x = ldo(Jx1, h * 8u), j1y = ldo(Jy1, h * 8u), j1z = ldo(Jz1, h * 8u);
    const uint64_t c0J3 = ldo(Jx1, (im + yb + LhL * z) * 8u), c0J4 = ldo(Jy1, (i + ymb + LhL * z) * 8u);
    // A1's loads (its -x and -y colour-0 bonds, its demons) with the others
    const bool doA1 = rowA1 && k >= 2;
    const uint64_t c1J3 = doA1 ? ldo(Jx0, (im + yb + LhL * z1) * 8u) : 0;
    const uint64_t c1J4 = doA1 ? ldo(Jy0, (i + ymb + LhL * z1) * 8u) : 0;
    S1p[cur * P + slot] = S0w;
    __syncthreads();
    msa::wait_vmem();
    // ---- A0: first sweep, colour 0, plane s
    {
      uint64_t D0 = dA0.x, D1 = dA0.y;
      const uint64_t *Sl = S1p + cur * P;
      const uint64_t nxp = Sl[row * Lh + ip], nxm = Sl[row * Lh + im];
      const uint64_t nyp = row + 1 == ny ? Hv : Sl[rp * Lh + i];
      const uint64_t nym = row == 0 ? Hv : Sl[rm * Lh + i];
      E0 = msa::site_update(S0v, nxp, nyp, Sp, nxm, nym, Sm, j0x, j0y, j0z, c0J3, c0J4, w1z[0],
                            D0, D1);
      A0b[ring * P + slot] = E0;
      if (!D2) XD0[ring * P + row * Lh + iw0] = make_ulonglong2(D0, D1);   // walked, read by B0 at k+2
    }
    // the next step's A0 inputs, in flight through A1, B0, B1 and the barrier
    if (k + 1 < nstep) {
      const uint32_t zn = wrapi(s + 1, L);
      dA0 = ldo(Din0, (i + yb + LhL * zn) * 16u);
      S0v = ldo(S0in, (i + yb + LhL * zn) * 8u);
      Hv = hrow ? ldo(S1in, (i + hrb + LhL * zn) * 8u) : 0;
    }
    // ---- A1: first sweep, colour 1, plane s-1
    if (doA1) {
      uint64_t D0 = dA1.x, D1 = dA1.y;
      const uint64_t *El = A0b + ring1 * P;
      const uint64_t Em2 = A0b[ring2 * P + slot];   // A0 at plane s-2, own column
      F0 = msa::site_update(Sm, El[row * Lh + ip], El[rp * Lh + i], E0, El[row * Lh + im],
                            El[rm * Lh + i], Em2, w1x[0], w1y[0], w1z[0], c1J3, c1J4, w0z[1],
                            D0, D1);
      A1b[ring * P + slot] = F0;
      if (!D2) XD1[ring * P + row * Lh + iw1] = make_ulonglong2(D0, D1);   // walked, read by B1 at k+2
    }
    if (rowA1 && k + 1 >= 2 && k + 1 < nstep) dA1 = ldo(Din1, h * 16u);   // next step's A1

    // ---- B0: second sweep, colour 0, plane s-2 (J of A0 two steps ago)
    if (rowB0 && k >= 4) {
      const ulonglong2 d = D2 ? dB0 : XD0[ring2 * P + slot];
      uint64_t D0 = d.x, D1 = d.y;
      const uint64_t *Fl = A1b + ring1 * P;
      const uint64_t Em2 = A0b[ring2 * P + slot];   // own spin: A0 at plane s-2
      const uint64_t F2 = A1b[ring2 * P + slot];    // A1 at plane s-3, own column
      G0 = msa::site_update(Em2, Fl[row * Lh + ip], Fl[rp * Lh + i], F0, Fl[row * Lh + im],
                            Fl[rm * Lh + i], F2, w0x[1], w0y[1], w0z[1], wc0x[1], wc0y[1], w1z[2],
                            D0, D1);
      B0b[ring * P + slot] = G0;
      if (rowB1 && k >= 5 && k <= (int)a.tz + 4) {
        sto_stream(S0out, h2 * 8u, G0);
        if (a.dm_store) sto_stream(Dout0, (iw0 + yb + LhL * z2) * 16u, D0, D1);
      }
    }
    if (D2 && rowB0 && k + 1 >= 4 && k + 1 < nstep) dB0 = ldo(Din20, (i + yb + LhL * z1) * 16u);
    // ---- B1: second sweep, colour 1, plane s-3 (the tile's own rows and planes)
    if (rowB1 && k >= 6) {
      const ulonglong2 d = D2 ? dB1 : XD1[ring2 * P + slot];
      uint64_t D0 = d.x, D1 = d.y;
      const uint64_t *Gl = B0b + ring1 * P;
      const uint64_t F2 = A1b[ring2 * P + slot];    // own spin: A1 at plane s-3
      const uint64_t G2 = B0b[ring2 * P + slot];    // B0 at plane s-4, own column
      const uint64_t sn = msa::site_update(F2, Gl[row * Lh + ip], Gl[rp * Lh + i], G0,
                                           Gl[row * Lh + im], Gl[rm * Lh + i], G2, w1x[2], w1y[2],
                                           w1z[2], wc1x[1], wc1y[1], w0z[3], D0, D1);
      sto_stream(S1out, h3 * 8u, sn);
      if (a.dm_store) sto_stream(Dout1, (iw1 + yb + LhL * z3) * 16u, D0, D1);
    }
    if (D2 && rowB1 && k + 1 >= 6 && k + 1 < nstep) dB1 = ldo(Din21, h2 * 16u);
    // ---- rotate the windows
    Sm = S0w;
    S0w = Sp;
    w0x[1] = w0x[0]; w0x[0] = j0x;
    w0y[1] = w0y[0]; w0y[0] = j0y;
    w0z[3] = w0z[2]; w0z[2] = w0z[1]; w0z[1] = w0z[0]; w0z[0] = j0z;
    w1x[2] = w1x[1]; w1x[1] = w1x[0]; w1x[0] = j1x;
    w1y[2] = w1y[1]; w1y[1] = w1y[0]; w1y[0] = j1y;
    w1z[2] = w1z[1]; w1z[1] = w1z[0]; w1z[0] = j1z;
    wc0x[1] = wc0x[0]; wc0x[0] = c0J3;
    wc0y[1] = wc0y[0]; wc0y[0] = c0J4;
    wc1x[1] = wc1x[0]; wc1x[0] = c1J3;
    wc1y[1] = wc1y[0]; wc1y[0] = c1J4;
  }
}

// ------------------------------------------------- resident small lattices --
// A lattice whose spins and demons fit in one CU's LDS (8 V + 16 nc V bytes,
// and 12 V more for the colour-1 bond planes when a thread owns two words of
// a colour: L = 16 with one layer, 14 with two, 12 with five, 10 with six)
// runs up to kResMax iterations in one launch: one block per replica keeps
// the state on chip (spins and demons in LDS, each thread's own spins and
// bonds in registers, or half of its bonds in LDS when it owns two words),
// and the colours of each layer follow each other behind barriers, so an
// iteration costs two to four barriers per layer instead of launches.  An
// iteration with a refresh reads the demons drawn ahead for it (refresh
// modes 1 and 2), loaded one colour ahead so the load overlaps the colour
// before.  Same updates, walk and order as k_halfsweep.
constexpr int kResMax = 64;
constexpr size_t kResLds = 160 * 1024;

// One layer and x rows that tile a wavefront: the demons stay in registers
// (k_sweep_resident_x).
static bool res_x(const Geo &g) { return g.nc == 1 && 64 % g.Lh == 0; }
// LDS of a resident block: spins and all bond planes (res_x); else spins,
// demons and, with two words of a colour per thread (L/2 L^2 > 1024), the
// colour-1 bond planes.
static size_t res_lds(const Geo &g) {
  if (res_x(g)) return 32 * (size_t)g.V;
  return (8 + 16 * (size_t)g.nc + (g.H > 1024 ? 12 : 0)) * g.V;
}
// At most two words of each colour per thread (L/2 L^2 <= 2048: L up to 16)
// and the state within one CU's LDS.
static bool res_fits(const Geo &g) { return g.H <= 2048 && res_lds(g) <= kResLds; }

struct ResArgs {
  const uint64_t *sp_in;    // spins [R][2][H] before the first iteration
  uint64_t *sp_out;         // after the last (may be sp_in)
  const uint64_t *jp;
  const ulonglong2 *dm_in;  // demons [R][nc][2][H]
  ulonglong2 *dm_out;       // (may be dm_in)
  const ulonglong2 *pre;    // demons drawn ahead: slot s at pre + s R nc V
  Geo g;
  uint32_t n;               // iterations, at most kResMax
  uint32_t mLh, mL;         // div_recip reciprocals of L/2 and L
  int dm_load;              // read dm_in (0: the first iteration's refresh replaces them)
  int dm_store;             // write dm_out (0: the next iteration's refresh replaces them)
  int16_t slot[kResMax];    // slot of each iteration's drawn demons, -1: no refresh
};

// Site h of colour col: its neighbours' indices in the other colour's plane
// (the indexing of k_halfsweep) and its coordinates for the walk.
struct ResNb {
  uint32_t xp, yp, zp, xm, ym, zm, i, y, z, off;
};
__device__ __forceinline__ ResNb res_nb(const Geo &g, int col, uint32_t h, uint32_t mLh,
                                        uint32_t mL) {
  const uint32_t Lh = g.Lh, L = g.L;
  const uint32_t row = div_recip(h, mLh), i = h - row * Lh;
  const uint32_t z = div_recip(row, mL), y = row - z * L;
  const uint32_t off = (y + z + (uint32_t)col) & 1u;
  const uint32_t ip = off ? msa::wrapi((int)i + 1, Lh) : i;
  const uint32_t im = off ? i : msa::wrapi((int)i - 1, Lh);
  const uint32_t yp = msa::wrapi((int)y + 1, L), ym = msa::wrapi((int)y - 1, L);
  const uint32_t zp = msa::wrapi((int)z + 1, L), zm = msa::wrapi((int)z - 1, L);
  ResNb n;
  n.xp = ip + Lh * row;
  n.xm = im + Lh * row;
  n.yp = i + Lh * (yp + L * z);
  n.ym = i + Lh * (ym + L * z);
  n.zp = i + Lh * (y + L * zp);
  n.zm = i + Lh * (y + L * zm);
  n.i = i;
  n.y = y;
  n.z = z;
  n.off = off;
  return n;
}

// Thread tid owns words h = tid + k blockDim (k < W) of each colour.  With
// W = 2 the colour-1 bond planes are in LDS: a colour-0 site keeps its +x,
// +y, +z bonds in registers (its -x, -y, -z bonds are colour-1 words), a
// colour-1 site its -x, -y, -z bonds (colour-0 words).
template <int W>
struct ResState {
  static constexpr int NJ = W == 2 ? 3 : 6;
  uint64_t s[2][W];         // own spins
  uint64_t J[2][W][NJ];     // own bonds: +x, +y, +z here, -x, -y, -z at the neighbours
  ulonglong2 nx[W];         // the next colour's drawn demons
};

// The drawn demons of colour c, layer l, iteration m into st.nx.
template <int W>
__device__ __forceinline__ void res_fetch(const ResArgs &a, ResState<W> &st, uint32_t m,
                                          uint32_t l, int c) {
  const uint32_t H = (uint32_t)a.g.H, tid = threadIdx.x, bd = blockDim.x;
  const size_t pw = (size_t)a.g.R * a.g.nc * a.g.V;
  const ulonglong2 *p = a.pre + (size_t)a.slot[m] * pw +
                        ((size_t)(blockIdx.x * a.g.nc + l) * 2 + (uint32_t)c) * H;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const uint32_t h = tid + (uint32_t)k * bd;
    if (h < H) st.nx[k] = p[h];
  }
}

// Colour C of layer l in iteration m.
template <int W, int C>
__device__ __forceinline__ void res_colour(const ResArgs &a, ResState<W> &st, uint64_t *ls,
                                           ulonglong2 *ld, const uint64_t *lj, uint32_t m,
                                           uint32_t l) {
  const Geo &g = a.g;
  const uint32_t H = (uint32_t)g.H, tid = threadIdx.x, bd = blockDim.x;
  const bool drawn = a.slot[m] >= 0;
  ulonglong2 d[W];
  ulonglong2 *plane = ld + ((size_t)l * 2 + C) * H;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const uint32_t h = tid + (uint32_t)k * bd;
    d[k] = make_ulonglong2(0, 0);
    if (h < H) d[k] = drawn ? st.nx[k] : plane[h];
  }
  // the next colour's drawn demons, in flight through this colour's update
  {
    uint32_t m2 = m, l2 = l;
    int c2 = C + 1;
    if (c2 == 2) {
      c2 = 0;
      if (++l2 == g.nc) {
        l2 = 0;
        ++m2;
      }
    }
    if (m2 < a.n && a.slot[m2] >= 0) res_fetch(a, st, m2, l2, c2);
  }
  if (!drawn) __syncthreads();          // every read of the plane before the walk writes it
  int dx, dy, dz;
  msa::layer_dir((int)l, dx, dy, dz);
  const uint64_t *So = ls + (size_t)(C ^ 1) * H;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const uint32_t h = tid + (uint32_t)k * bd;
    if (h >= H) continue;
    const ResNb nb = res_nb(g, C, h, a.mLh, a.mL);
    uint64_t D0 = d[k].x, D1 = d[k].y;
    const uint64_t *Jk = st.J[C][k];
    uint64_t j[6];
    if constexpr (W != 2) {
#pragma unroll
      for (int q = 0; q < 6; ++q) j[q] = Jk[q];
    } else if constexpr (C == 0) {
      j[0] = Jk[0];
      j[1] = Jk[1];
      j[2] = Jk[2];
      j[3] = lj[nb.xm];
      j[4] = lj[H + nb.ym];
      j[5] = lj[2 * H + nb.zm];
    } else {
      j[0] = lj[h];
      j[1] = lj[H + h];
      j[2] = lj[2 * H + h];
      j[3] = Jk[0];
      j[4] = Jk[1];
      j[5] = Jk[2];
    }
    const uint64_t sn = msa::site_update(st.s[C][k], So[nb.xp], So[nb.yp], So[nb.zp], So[nb.xm],
                                         So[nb.ym], So[nb.zm], j[0], j[1], j[2], j[3], j[4], j[5],
                                         D0, D1);
    st.s[C][k] = sn;
    ls[(size_t)C * H + h] = sn;
    plane[(uint32_t)msa::walk_dest(g, C, dx, dy, dz, nb.i, nb.y, nb.z)] = make_ulonglong2(D0, D1);
  }
  __syncthreads();
}

template <int W>
__global__ void __launch_bounds__(1024) k_sweep_resident(ResArgs a) {
  extern __shared__ uint4 s_res[];
  const Geo &g = a.g;
  const uint32_t H = (uint32_t)g.H, nc = g.nc, tid = threadIdx.x, bd = blockDim.x;
  const uint32_t r = blockIdx.x;
  uint64_t *ls = reinterpret_cast<uint64_t *>(s_res);                  // spins [2][H]
  ulonglong2 *ld = reinterpret_cast<ulonglong2 *>(ls + 2 * (size_t)H);  // demons [nc][2][H]
  const size_t nd = (size_t)nc * 2 * H;
  uint64_t *lj = reinterpret_cast<uint64_t *>(ld + nd);                 // W = 2: J colour 1 [3][H]
  const uint64_t *spi = a.sp_in + 2 * (size_t)r * H;
  for (uint32_t q = tid; q < 2 * H; q += bd) ls[q] = spi[q];
  if constexpr (W == 2)
    for (uint32_t q = tid; q < 3 * H; q += bd) lj[q] = a.jp[3 * (size_t)H + q];
  if (a.dm_load) {
    const ulonglong2 *dmi = a.dm_in + (size_t)r * nd;
    for (uint32_t q = tid; q < nd; q += bd) ld[q] = dmi[q];
  }
  ResState<W> st;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const uint64_t *Jc = a.jp + (size_t)c * 3 * H, *Jo = a.jp + (size_t)(c ^ 1) * 3 * H;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t h = tid + (uint32_t)k * bd;
      st.s[c][k] = 0;
#pragma unroll
      for (int d = 0; d < ResState<W>::NJ; ++d) st.J[c][k][d] = 0;
      if (h >= H) continue;
      const ResNb nb = res_nb(g, c, h, a.mLh, a.mL);
      st.s[c][k] = spi[(size_t)c * H + h];
      if (W != 2 || c == 0) {
        st.J[c][k][0] = Jc[h];
        st.J[c][k][1] = Jc[H + h];
        st.J[c][k][2] = Jc[2 * H + h];
      }
      if (W != 2 || c == 1) {
        const int o = W == 2 ? 0 : 3;
        st.J[c][k][o] = Jo[nb.xm];
        st.J[c][k][o + 1] = Jo[H + nb.ym];
        st.J[c][k][o + 2] = Jo[2 * H + nb.zm];
      }
    }
  }
  if (a.n > 0 && a.slot[0] >= 0) res_fetch(a, st, 0, 0, 0);
  __syncthreads();
  for (uint32_t m = 0; m < a.n; ++m)
    for (uint32_t l = 0; l < nc; ++l) {
      res_colour<W, 0>(a, st, ls, ld, lj, m, l);
      res_colour<W, 1>(a, st, ls, ld, lj, m, l);
    }
  uint64_t *spo = a.sp_out + 2 * (size_t)r * H;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t h = tid + (uint32_t)k * bd;
      if (h < H) spo[(size_t)c * H + h] = st.s[c][k];
    }
  if (a.dm_store) {
    ulonglong2 *dmo = a.dm_out + (size_t)r * nd;
    for (uint32_t q = tid; q < nd; q += bd) dmo[q] = ld[q];
  }
}

// The resident form for one demon layer and x rows that tile a wavefront
// (L/2 divides 64: L = 4, 8, 16).  Layer 0's demons walk one word along x
// inside their row, so each thread keeps its words' demons in registers and
// the walk is a lane shift within the wavefront (ds_bpermute): no demon
// planes in LDS, one barrier per colour.  The LDS holds the spins and all six
// bond planes (32 V bytes).
__device__ __forceinline__ ulonglong2 res_lane_take(ulonglong2 v, int addr) {
  const auto take = [addr](uint64_t w) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)w);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(w >> 32));
    return ((uint64_t)hi << 32) | lo;
  };
  return make_ulonglong2(take(v.x), take(v.y));
}

template <int W>
struct ResXState {
  uint64_t s[2][W];         // own spins
  ulonglong2 dm[2][W];      // own demons (after the walk)
  ulonglong2 nx[W];         // the next colour's drawn demons
};

template <int W, int C>
__device__ __forceinline__ void resx_colour(const ResArgs &a, ResXState<W> &st, uint64_t *ls,
                                            const uint64_t *lj, uint32_t m, bool walk,
                                            int src) {
  const Geo &g = a.g;
  const uint32_t H = (uint32_t)g.H, tid = threadIdx.x, bd = blockDim.x;
  const bool drawn = a.slot[m] >= 0;
  ulonglong2 d[W];
#pragma unroll
  for (int k = 0; k < W; ++k) d[k] = drawn ? st.nx[k] : st.dm[C][k];
  {
    const uint32_t m2 = C == 0 ? m : m + 1;
    if (m2 < a.n && a.slot[m2] >= 0) {
      const size_t pw = (size_t)g.R * g.V;
      const ulonglong2 *p = a.pre + (size_t)a.slot[m2] * pw +
                            ((size_t)blockIdx.x * 2 + (C ^ 1)) * H;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        const uint32_t h = tid + (uint32_t)k * bd;
        if (h < H) st.nx[k] = p[h];
      }
    }
  }
  const uint64_t *So = ls + (size_t)(C ^ 1) * H;
  const uint64_t *Jc = lj + (size_t)C * 3 * H, *Jo = lj + (size_t)(C ^ 1) * 3 * H;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const uint32_t h = tid + (uint32_t)k * bd;
    if (h >= H) continue;
    const ResNb nb = res_nb(g, C, h, a.mLh, a.mL);
    uint64_t D0 = d[k].x, D1 = d[k].y;
    // one x neighbour is the other colour's word at the same index, which
    // this thread holds (off = 0: the +x one, off = 1: the -x one)
    const uint64_t near = st.s[C ^ 1][k], far = So[nb.off ? nb.xp : nb.xm];
    const uint64_t sn = msa::site_update(st.s[C][k], nb.off ? far : near, So[nb.yp], So[nb.zp],
                                         nb.off ? near : far, So[nb.ym], So[nb.zm], Jc[h],
                                         Jc[H + h], Jc[2 * H + h], Jo[nb.xm], Jo[H + nb.ym],
                                         Jo[2 * H + nb.zm], D0, D1);
    st.s[C][k] = sn;
    ls[(size_t)C * H + h] = sn;
    d[k] = make_ulonglong2(D0, D1);
  }
  // the walk: colour 0 moves +x, colour 1 -x (wrapping in the row), every
  // lane taking part; skipped where the next iteration's refresh replaces them
  if (walk) {
#pragma unroll
    for (int k = 0; k < W; ++k) st.dm[C][k] = res_lane_take(d[k], src);
  }
  __syncthreads();
}

template <int W>
__global__ void __launch_bounds__(1024) k_sweep_resident_x(ResArgs a) {
  extern __shared__ uint4 s_res[];
  const Geo &g = a.g;
  const uint32_t H = (uint32_t)g.H, Lh = g.Lh, tid = threadIdx.x, bd = blockDim.x;
  const uint32_t r = blockIdx.x;
  uint64_t *ls = reinterpret_cast<uint64_t *>(s_res);   // spins [2][H]
  uint64_t *lj = ls + 2 * (size_t)H;                     // J [2 colours][3][H]
  const uint64_t *spi = a.sp_in + 2 * (size_t)r * H;
  const ulonglong2 *dmi = a.dm_in + 2 * (size_t)r * H;
  for (uint32_t q = tid; q < 2 * H; q += bd) ls[q] = spi[q];
  for (uint32_t q = tid; q < 6 * H; q += bd) lj[q] = a.jp[q];
  ResXState<W> st;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t h = tid + (uint32_t)k * bd;
      st.s[c][k] = h < H ? spi[(size_t)c * H + h] : 0;
      st.dm[c][k] = (a.dm_load && h < H) ? dmi[(size_t)c * H + h] : make_ulonglong2(0, 0);
    }
  if (a.n > 0 && a.slot[0] >= 0) {
    const ulonglong2 *p = a.pre + (size_t)a.slot[0] * ((size_t)g.R * g.V) + (size_t)r * 2 * H;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t h = tid + (uint32_t)k * bd;
      if (h < H) st.nx[k] = p[h];
    }
  }
  // lane of the word each word's walked demon comes from: x - 1 for colour
  // 0, x + 1 for colour 1, in the same row of L/2 consecutive lanes
  const uint32_t lane = __lane_id(), i = tid & (Lh - 1);
  const int src0 = (int)(i ? lane - 1 : lane + Lh - 1) * 4;
  const int src1 = (int)(i + 1 < Lh ? lane + 1 : lane + 1 - Lh) * 4;
  __syncthreads();
  for (uint32_t m = 0; m < a.n; ++m) {
    const bool walk = m + 1 < a.n ? a.slot[m + 1] < 0 : a.dm_store != 0;
    resx_colour<W, 0>(a, st, ls, lj, m, walk, src0);
    resx_colour<W, 1>(a, st, ls, lj, m, walk, src1);
  }
  uint64_t *spo = a.sp_out + 2 * (size_t)r * H;
  ulonglong2 *dmo = a.dm_out + 2 * (size_t)r * H;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t h = tid + (uint32_t)k * bd;
      if (h >= H) continue;
      spo[(size_t)c * H + h] = st.s[c][k];
      if (a.dm_store) dmo[(size_t)c * H + h] = st.dm[c][k];
    }
}

struct MeasArgs {
  const uint64_t *sp;     // spins [2R][H]: replica r colour c in plane 2r + c
  const uint64_t *jp;     // couplings [2 colours][3 directions][H]
  const ulonglong2 *dm;   // demons [R][nc][2 colours][H]
  Geo g;
  uint32_t nr, mLh;       // rows of L/2 words per step; div_recip reciprocal of L/2
  uint32_t steps;         // row steps per block, at most 16 (the counters' width)
  int32_t *out;           // [R unsat | R demon | R/2 q | R/2 qlink][64]
};

// Every per-lane integer observable of the lattice in one pass.  Every bond
// touches exactly one colour-0 site, so the six bonds of the colour-0 sites
// cover the lattice once (the sum d_computeEne forms, main.cu:745-815); the
// demon energy is Σ D0 + 2 D1 over both colours and all layers
// (Measure_Energy_Clerics, med.c:55-83).  blockIdx.y = 2p + k: replicas 2p
// and 2p + 1, k = 0 the spin observables (unsatisfied bonds of each, q and
// q_link of the pair), k = 1 the demon energies.  The six bond words of a
// site are summed per bit lane by an adder tree, the sums accumulate in
// bit-sliced counters, which are transposed to lanes across the wavefront,
// summed over the block in LDS and added to the output once per block.  The
// block is nr rows of L/2 threads rounded up to whole wavefronts (the
// transpose reads every lane; the extra lanes count nothing).
__global__ void __launch_bounds__(512) k_measure(MeasArgs a) {
  __shared__ uint32_t bsum[4][64];
  const Geo g = a.g;
  const uint32_t Lh = g.Lh, L = g.L, rows = L * L;
  const uint64_t H = g.H;
  const uint32_t rloc = div_recip(threadIdx.x, a.mLh), i = threadIdx.x - rloc * Lh;
  const uint32_t p = blockIdx.y >> 1, ra = 2 * p, rb = ra + 1;
  const bool pair = rb < g.R, demons = (blockIdx.y & 1u) != 0;
  for (uint32_t k = threadIdx.x; k < 256; k += blockDim.x) (&bsum[0][0])[k] = 0;
  uint32_t tot[4] = {0, 0, 0, 0};
  if (!demons) {
    msa::VCount<7> cu0, cu1, cq, cl;
    cu0.clear();
    cu1.clear();
    cq.clear();
    cl.clear();
    const uint64_t *J0 = a.jp, *J1 = a.jp + 3 * H;
    const uint64_t *Sa0 = a.sp + (uint64_t)(2 * ra) * H, *Sa1 = Sa0 + H;
    const uint64_t *Sb0 = Sa1 + H, *Sb1 = Sb0 + H;
    for (uint32_t k = 0; k < a.steps; ++k) {
      const uint32_t row = (blockIdx.x * a.steps + k) * a.nr + rloc;
      if (rloc >= a.nr || row >= rows) continue;
      const uint32_t z = row / L, y = row - z * L;
      const uint32_t off = (y + z) & 1u;   // colour-0 site at x = 2i + off
      const uint32_t ip = off ? msa::wrapi((int)i + 1, Lh) : i;
      const uint32_t im = off ? i : msa::wrapi((int)i - 1, Lh);
      const uint32_t yp = msa::wrapi((int)y + 1, L), ym = msa::wrapi((int)y - 1, L);
      const uint32_t zp = msa::wrapi((int)z + 1, L), zm = msa::wrapi((int)z - 1, L);
      const uint64_t h = (uint64_t)row * Lh + i, rowb = (uint64_t)row * Lh;
      const uint64_t nb[6] = {ip + rowb, i + (uint64_t)Lh * (yp + (uint64_t)L * z),
                              i + (uint64_t)Lh * (y + (uint64_t)L * zp), im + rowb,
                              i + (uint64_t)Lh * (ym + (uint64_t)L * z),
                              i + (uint64_t)Lh * (y + (uint64_t)L * zm)};
      uint64_t jv[6], na[6];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        jv[d] = J0[d * H + h];
        jv[d + 3] = J1[d * H + nb[d + 3]];
      }
#pragma unroll
      for (int d = 0; d < 6; ++d) na[d] = Sa1[nb[d]];
      const uint64_t sa = Sa0[h];
      uint64_t w[6], b0, b1, b2;
#pragma unroll
      for (int d = 0; d < 6; ++d) w[d] = sa ^ jv[d] ^ na[d];
      msa::sum6(w, b0, b1, b2);
      cu0.add3(b0, b1, b2);
      if (pair) {
        const uint64_t sb = Sb0[h];
        const uint64_t q0 = sa ^ sb, q1 = Sa1[h] ^ Sb1[h];
        cq.add3(q0 ^ q1, q0 & q1, 0);
        uint64_t nbv[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) {
          nbv[d] = Sb1[nb[d]];
          w[d] = sb ^ jv[d] ^ nbv[d];
        }
        msa::sum6(w, b0, b1, b2);
        cu1.add3(b0, b1, b2);
#pragma unroll
        for (int d = 0; d < 6; ++d) w[d] = (sa ^ na[d]) ^ (sb ^ nbv[d]);
        msa::sum6(w, b0, b1, b2);
        cl.add3(b0, b1, b2);
      }
    }
    tot[0] = cu0.wave_total();
    tot[1] = cu1.wave_total();
    tot[2] = cq.wave_total();
    tot[3] = cl.wave_total();
  } else {
    msa::VCount<10> c0, c1;
    c0.clear();
    c1.clear();
    const ulonglong2 *Da = a.dm + (uint64_t)ra * g.nc * 2 * H, *Db = Da + (uint64_t)g.nc * 2 * H;
    for (uint32_t k = 0; k < a.steps; ++k) {
      const uint32_t row = (blockIdx.x * a.steps + k) * a.nr + rloc;
      if (rloc >= a.nr || row >= rows) continue;
      const uint64_t h = (uint64_t)row * Lh + i;
      // per layer, both colours: (x0 + x1) + 2 (y0 + y1) <= 6 as three bit-planes
      auto layer_sum = [](const ulonglong2 &u, const ulonglong2 &v, msa::VCount<10> &cnt) {
        const uint64_t s0 = u.x ^ v.x, k0 = u.x & v.x, s1 = u.y ^ v.y, k1 = u.y & v.y;
        cnt.add3(s0, k0 ^ s1, (k0 & s1) | k1);
      };
      for (uint32_t l = 0; l < g.nc; ++l) {
        layer_sum(Da[2 * l * H + h], Da[(2 * l + 1) * H + h], c0);
        if (pair) layer_sum(Db[2 * l * H + h], Db[(2 * l + 1) * H + h], c1);
      }
    }
    tot[0] = c0.wave_total();
    tot[1] = c1.wave_total();
  }
  __syncthreads();
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (tot[j]) atomicAdd(&bsum[j][lane], tot[j]);
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < 256; k += blockDim.x) {
    const uint32_t j = k >> 6, b = k & 63u, v = bsum[j][b];
    if (!v || (j > 0 && !pair) || (demons && j > 1)) continue;
    const uint32_t slot = demons ? g.R + ra + j
                                 : (j < 2 ? ra + j : 2 * g.R + (j == 2 ? 0 : g.R / 2) + p);
    atomicAdd(a.out + 64 * slot + b, (int32_t)v);
  }
}

// Counter-based construction (msa_init_philox): thread = canonical site.
// Layout conversion of state transfers on the device: canonical site order
// s = x + L (y + L z) (the host arrays) <-> the colour-split planes.  Thread
// per (plane q, site s); TO_CANON: out[q V + s] = in[(2 q + col) H + h]
// (spins and demons keep both colours of a replica / layer next to each
// other) or in[(col nplane + q) H + h] (cmajor: couplings), else the
// reverse.  Demons (16-byte words) go through two 8-byte arrays, d0 and d1
// (DEM).
template <bool TO_CANON, bool DEM>
__global__ void __launch_bounds__(256) k_relayout(const void *in, void *out0, void *out1, Geo g,
                                                  uint32_t nplane, int cmajor) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)nplane * g.V) return;
  const uint64_t q = gid / g.V, s = gid - q * g.V;
  const uint32_t x = (uint32_t)(s % g.L);
  const uint64_t yz = s / g.L;
  const uint32_t y = (uint32_t)(yz % g.L), z = (uint32_t)(yz / g.L);
  const uint32_t col = (x + y + z) & 1u;
  const uint64_t h = (x >> 1) + (uint64_t)g.Lh * (y + (uint64_t)g.L * z);
  // split index: colour minor (spins, demons) or major (couplings: plane
  // col 3 + d)
  const uint64_t sp = (cmajor ? (uint64_t)col * nplane + q : q * 2 + col) * g.H + h;
  if (DEM) {
    if (TO_CANON) {
      const ulonglong2 v = reinterpret_cast<const ulonglong2 *>(in)[sp];
      reinterpret_cast<uint64_t *>(out0)[gid] = v.x;
      reinterpret_cast<uint64_t *>(out1)[gid] = v.y;
    } else {
      const uint64_t a = reinterpret_cast<const uint64_t *>(in)[gid];
      const uint64_t b = reinterpret_cast<const uint64_t *>(out1)[gid];   // d1 (input)
      reinterpret_cast<ulonglong2 *>(out0)[sp] = make_ulonglong2(a, b);
    }
  } else {
    if (TO_CANON)
      reinterpret_cast<uint64_t *>(out0)[gid] = reinterpret_cast<const uint64_t *>(in)[sp];
    else
      reinterpret_cast<uint64_t *>(out0)[sp] = reinterpret_cast<const uint64_t *>(in)[gid];
  }
}

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
static int check_geom(const msa_geom *g) {
  if (!g) return msa_fail(MSA_ERR_ARG, "geometry is NULL");
  if (g->L < 4 || g->L > 1024 || (g->L & 1))
    return msa_fail(MSA_ERR_ARG, "L=%d must be even and in 4..1024", g->L);
  if (g->R < 1 || g->R > 16) return msa_fail(MSA_ERR_ARG, "R=%d must be in 1..16", g->R);
  if (g->nc < 1 || g->nc > 6) return msa_fail(MSA_ERR_ARG, "nc=%d must be in 1..6", g->nc);
  const uint64_t V = (uint64_t)g->L * g->L * g->L;
  if ((uint64_t)g->nc * V > 0xFFFFFFFFull) return msa_fail(MSA_ERR_ARG, "nc*V exceeds 2^32");
  return MSA_OK;
}

// Rows of the count buffer: R unsatisfied-bond, R demon, R/2 q, R/2 q_link.
static uint32_t meas_rows(uint32_t R) { return 2 * R + 2 * (R / 2); }

// Bytes of a context's state allocation: spins and demons twice (ping-pong),
// J, the measurement counts, each rounded up to 256 bytes.
static size_t state_bytes(const msa_geom *geom) {
  const size_t V = (size_t)geom->L * geom->L * geom->L;
  auto up = [](size_t n) { return (n + 255) & ~(size_t)255; };
  return 2 * up(8 * V * geom->R) + up(8 * V * 3) + 2 * up(16 * V * geom->R * geom->nc) +
         up(64 * sizeof(int32_t) * meas_rows(geom->R));
}

extern "C" int64_t msa_footprint_bytes(const msa_geom *geom) {
  if (check_geom(geom) != MSA_OK) return -1;
  return (int64_t)state_bytes(geom);
}

// Tile of the two-pass kernel: nr rows of L/2 site pairs per block, rs
// replica groups.  Blocks are persistent (as many as the CUs hold at once,
// by the occupancy of the refresh kernel with its two LDS tables) and walk
// tiles of nr rows.  The refresh is VALU work with long dependency chains and
// LDS searches, so the choice maximises the waves resident per CU: over the
// form with all R replicas per thread and the ONE form with the replicas on
// separate groups of whole wavefronts (fewer registers, and on small
// lattices twice the threads: L = 64, R = 2 fills a CU with 1024 threads
// where one thread per site gives 512), and over the rows per tile; ties go
// to the larger tile and then to the ONE form.
static void pick_half(msa_ctx *c) {
  const uint32_t Lh = c->g.Lh, rows = c->g.L * c->g.L, R = c->g.R;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu < 1)
    ncu = 256;
  double best = -1;
  c->hnr = 1;
  c->hrs = R == 1 ? 1 : 0;
  c->hgrid = 1;
  for (int form = 0; form < 2; ++form) {
    const bool one = form == 1 || R == 1;
    const uint32_t rs = form == 1 ? R : 1;
    if (form == 1 && R == 1) continue;
    uint32_t step = 1;                       // ONE form: a group is whole wavefronts
    if (form == 1)
      while ((step * Lh) % 64) ++step;
    for (uint32_t nr = step; (uint64_t)nr * Lh * rs <= 1024; nr += step) {
      const uint32_t threads = nr * Lh * rs;
      int bpc = 0;
      const void *f = one ? reinterpret_cast<const void *>(&k_halfsweep<kRefresh, true>)
                          : reinterpret_cast<const void *>(&k_halfsweep<kRefresh, false>);
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, f, (int)threads, kRefreshLds) != hipSuccess ||
          bpc < 1)
        bpc = 1;
      const uint64_t tiles = (rows + nr - 1) / nr;
      const uint64_t blocks = tiles < (uint64_t)ncu * bpc ? tiles : (uint64_t)ncu * bpc;
      const double waves = (double)blocks * threads / 64.0 / ncu;   // active lanes, in waves
      if (waves >= best) {
        best = waves;
        c->hnr = nr;
        c->hrs = rs;
        c->hgrid = (uint32_t)blocks;
      }
    }
  }
  if (c->hrs == 0) c->hrs = 1;
}

static bool fused_supported_R(uint32_t R) { return R == 1 || R == 2 || R == 4; }

// A one-pass tile is valid when ty and tz divide L, the block (L/2 pairs x
// thread rows) fits 1024 threads and its LDS (three planes of new colour-0
// values) fits 64 KiB.
static int fused_tile_ok(const Geo &g, uint32_t ty, uint32_t tz) {
  if (!fused_supported_R(g.R) || ty < 1 || tz < 1 || ty > g.L || tz > g.L) return 0;
  if (g.L % ty || g.L % tz) return 0;
  const uint32_t ny = ty == g.L ? g.L : ty + 2;
  if ((uint64_t)g.Lh * ny > 1024) return 0;
  if ((uint64_t)3 * ny * g.Lh * g.R * 8 > 65536) return 0;
  if (g.R >= 4 && (uint64_t)g.Lh * ny > 768) return 0;   // launch bound of the R = 4 kernels
  return 1;
}

// The second one-pass kernel addresses every buffer with 32-bit byte offsets.
static bool narrow_ok(const Geo &g) {
  const uint64_t lim = 1ull << 32;
  return 16ull * g.R * g.nc * g.V < lim && 24ull * g.V < lim && 8ull * g.R * g.V < lim;
}

// LDS of the sweep-pair kernel per thread slot: the odd input spins
// double-buffered, three three-slot rings of 8-byte results and two of
// 16-byte demons.
static size_t pair_lds(const Geo &g, uint32_t ty) { return (size_t)184 * (ty + 6) * g.Lh; }
constexpr size_t kPairLdsMax = 160 * 1024;
// The D2 form (second iteration refreshed): no demon rings, at most 512 threads.
static size_t pair_lds_d2(const Geo &g, uint32_t ty) { return (size_t)88 * (ty + 6) * g.Lh; }
static bool pair_d2_ok(const Geo &g, uint32_t ty) {
  return (uint64_t)(ty + 6) * g.Lh <= 512 && pair_lds_d2(g, ty) <= kPairLdsMax;
}
static void pick_pair_d2(const Geo &g, uint32_t &bty, uint32_t &btz);
static int refresh_mode(const msa_ctx *c);

// A sweep-pair tile is valid for one demon layer (its demons walk along x,
// inside a block's rows), 32-bit offsets, ty and tz dividing L and a block
// (L/2 pairs x (ty + 6) rows) within 1024 threads and the LDS of a CU.
static int pair_tile_ok(const Geo &g, uint32_t ty, uint32_t tz) {
  if (g.nc != 1 || !narrow_ok(g) || ty < 1 || tz < 1 || ty > g.L || tz > g.L) return 0;
  if (g.L % ty || g.L % tz) return 0;
  if ((uint64_t)g.Lh * (ty + 6) > 1024) return 0;
  return pair_lds(g, ty) <= kPairLdsMax;
}

static void set_fused(msa_ctx *c, int mode, uint32_t ty, uint32_t tz) {
  c->fused = mode;
  c->fty = ty;
  c->ftz = tz;
  c->fyhalo = ty != c->g.L;
  c->fny = c->fyhalo ? ty + 2 : c->g.L;
}

// Default one-pass tile: the most rows per block (fewest recomputed halo rows)
// with at least 256 blocks where possible, slabs of 8..16 planes.
static void pick_fused(msa_ctx *c) {
  c->fused = 0;
  const Geo &g = c->g;
  uint32_t best_ty = 0, best_tz = 0;
  double best = -1;
  for (uint32_t ty = 1; ty <= g.L; ++ty)
    for (uint32_t tz = 4; tz <= g.L && tz <= 32; ++tz) {
      if (!fused_tile_ok(g, ty, tz)) continue;
      const double blocks = (double)(g.L / ty) * (g.L / tz);
      const double ny = ty == g.L ? g.L : ty + 2.0;
      const double waste = (ny / ty) * ((tz + 2.0) / tz);   // recomputed colour-0 work
      const double fill = blocks >= 256 ? 1.0 : blocks / 256.0;
      const double score = fill / waste;
      if (score > best) { best = score; best_ty = ty; best_tz = tz; }
    }
  if (best_ty) set_fused(c, narrow_ok(g) ? 2 : 1, best_ty, best_tz);
}

static int thermal_scratch_init(int device);

extern "C" int msa_create(int device, const msa_geom *geom, msa_ctx **out) {
  if (!out) return msa_fail(MSA_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = check_geom(geom);
  if (rc) return rc;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return msa_fail(MSA_ERR_ARG, "device=%d count=%d", device, ndev);
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
    return cleanup(msa_fail(MSA_ERR_HIP, "stream/event creation failed"));
  // one allocation for the state (spins and demons ping-pong, J, the
  // measurement counts), each array at a 256-byte boundary
  const size_t nsp = 8 * V * geom->R, njp = 8 * V * 3, ndm = 16 * V * geom->R * geom->nc;
  auto up = [](size_t n) { return (n + 255) & ~(size_t)255; };
  if (hipMalloc(&c->slab, state_bytes(geom)) != hipSuccess)
    return cleanup(msa_fail(MSA_ERR_NOMEM, "device allocation of %lld bytes failed",
                        (long long)msa_footprint_bytes(geom)));
  {
    char *q = static_cast<char *>(c->slab);
    c->sp[0] = reinterpret_cast<uint64_t *>(q); q += up(nsp);
    c->sp[1] = reinterpret_cast<uint64_t *>(q); q += up(nsp);
    c->jp = reinterpret_cast<uint64_t *>(q); q += up(njp);
    c->dm[0] = reinterpret_cast<ulonglong2 *>(q); q += up(ndm);
    c->dm[1] = reinterpret_cast<ulonglong2 *>(q); q += up(ndm);
    c->meas = reinterpret_cast<int32_t *>(q);
  }
  // every transfer and clear of a context goes on its own (non-blocking)
  // stream: the legacy null stream does not order against it
  if (hipMemsetAsync(c->dm[0], 0, 16 * V * geom->R * geom->nc, c->stream) != hipSuccess ||
      hipMemsetAsync(c->sp[0], 0, 8 * V * geom->R, c->stream) != hipSuccess ||
      hipMemsetAsync(c->jp, 0, 8 * V * 3, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return cleanup(msa_fail(MSA_ERR_HIP, "initial clear failed"));
  // the refresh kernel stages two search tables (68 KiB) in LDS; the
  // sweep-pair kernel's buffers take up to a CU's whole LDS
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_halfsweep<kRefresh, true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRefreshLds) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_halfsweep<kRefresh, false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRefreshLds) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_refresh_draw),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDrawLds) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sweep_pair<false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPairLdsMax) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sweep_pair<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPairLdsMax) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sweep_resident<1>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResLds) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sweep_resident<2>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResLds) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sweep_resident_x<1>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResLds) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sweep_resident_x<2>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResLds) != hipSuccess)
    return cleanup(msa_fail(MSA_ERR_HIP, "kernel LDS attribute"));
  pick_half(c);
  pick_fused(c);
  pick_pair_d2(c->g, c->qty, c->qtz);
  // lattices that fit one CU run resident (each of the other forms is bound
  // by its launches there)
  if (res_fits(c->g)) {
    c->res = 1;
    c->fused = 0;
  }
  if ((rc = thermal_scratch_init(device))) return cleanup(rc);
  *out = c;
  return MSA_OK;
}

extern "C" int msa_destroy(msa_ctx *c) {
  if (!c) return MSA_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->dstream) (void)hipStreamSynchronize(c->dstream);
  walkers_free(c);
  (void)hipFree(c->slab);
  (void)hipFree(c->pre);
  for (int h = 0; h < 2; ++h) {
    if (c->e_drawn[h]) (void)hipEventDestroy(c->e_drawn[h]);
    if (c->e_used[h]) (void)hipEventDestroy(c->e_used[h]);
  }
  if (c->dstream) (void)hipStreamDestroy(c->dstream);
  if (c->xstream) {
    (void)hipStreamSynchronize(c->xstream);
    (void)hipStreamDestroy(c->xstream);
  }
  if (c->e_meas) (void)hipEventDestroy(c->e_meas);
  if (c->meas_host) (void)hipHostFree(c->meas_host);
  if (c->e_snap) (void)hipEventDestroy(c->e_snap);
  if (c->e_xfer) (void)hipEventDestroy(c->e_xfer);
  (void)hipFree(c->xfer);
  if (c->luts_owned) (void)hipFree(c->luts);
  (void)hipFree(c->ctr);
  (void)hipFree(c->key);
  (void)hipFree(c->th_rnd);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MSA_OK;
}

// Default sweep-pair tile: the fewest recomputed sites with at least 256
// blocks (one replica per block) where possible.
static int pick_pair(const Geo &g, uint32_t &bty, uint32_t &btz) {
  double best = -1;
  bty = btz = 0;
  for (uint32_t ty = 1; ty <= g.L; ++ty)
    for (uint32_t tz = 4; tz <= g.L && tz <= 64; ++tz) {
      if (!pair_tile_ok(g, ty, tz)) continue;
      const double blocks = (double)(g.L / ty) * (g.L / tz) * g.R;
      double work = 0;
      for (int p = 0; p < 4; ++p) work += (ty + 6.0 - 2 * p) * (tz + 6.0 - 2 * p);
      const double waste = work / (4.0 * ty * tz);
      const double fill = blocks >= 256 ? 1.0 : blocks / 256.0;
      const double score = fill / waste;
      if (score > best) { best = score; bty = ty; btz = tz; }
    }
  return bty != 0;
}

// Tile of the D2 sweep pair (a pair whose second iteration refreshes, with
// demons drawn ahead): the pick_pair score over the tiles of at most 512
// threads, ties to the taller tile; ty = 0 when the geometry has none.
static void pick_pair_d2(const Geo &g, uint32_t &bty, uint32_t &btz) {
  double best = -1;
  bty = btz = 0;
  for (uint32_t ty = g.L; ty >= 1; --ty)
    for (uint32_t tz = 4; tz <= g.L && tz <= 64; ++tz) {
      if (!pair_tile_ok(g, ty, tz) || !pair_d2_ok(g, ty)) continue;
      const double blocks = (double)(g.L / ty) * (g.L / tz) * g.R;
      double work = 0;
      for (int p = 0; p < 4; ++p) work += (ty + 6.0 - 2 * p) * (tz + 6.0 - 2 * p);
      const double waste = work / (4.0 * ty * tz);
      const double fill = blocks >= 256 ? 1.0 : blocks / 256.0;
      const double score = fill / waste;
      if (score > best) { best = score; bty = ty; btz = tz; }
    }
}

extern "C" int msa_set_tiling(msa_ctx *c, int32_t mode, int32_t ty, int32_t tz) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  if (mode == 4) {
    if (!res_fits(c->g))
      return msa_fail(MSA_ERR_ARG, "L=%u nc=%u does not fit one CU (L/2 L^2 <= 2048 and "
                      "%zu bytes of state <= %zu)", c->g.L, c->g.nc, res_lds(c->g), kResLds);
    c->res = 1;
    c->pair = 0;
    c->fused = 0;                         // refresh sweeps drawn inside the sweep: two-pass
    return MSA_OK;
  }
  c->res = 0;
  if (mode == 3) {
    uint32_t y = (uint32_t)ty, z = (uint32_t)tz;
    if (ty == 0 && tz == 0 && !pick_pair(c->g, y, z))
      return msa_fail(MSA_ERR_ARG, "no sweep-pair tile for L=%u R=%u nc=%u", c->g.L, c->g.R, c->g.nc);
    if (!pair_tile_ok(c->g, y, z))
      return msa_fail(MSA_ERR_ARG, "sweep-pair tile ty=%d tz=%d invalid for L=%u R=%u nc=%u", ty, tz,
                      c->g.L, c->g.R, c->g.nc);
    // a sweep left over between refreshes runs on the one-pass kernel where
    // it exists for this R (its default tile), else on the two-pass form
    pick_fused(c);
    c->pair = 1;
    c->pty = y;
    c->ptz = z;
    return MSA_OK;
  }
  c->pair = 0;
  if (mode == 0) {
    c->fused = 0;
    return MSA_OK;
  }
  if (mode != 1 && mode != 2)
    return msa_fail(MSA_ERR_ARG, "mode=%d (0 two-pass, 1 and 2 one-pass, 3 sweep pairs, 4 resident)", mode);
  if (mode == 2 && !narrow_ok(c->g))
    return msa_fail(MSA_ERR_ARG, "mode 2 needs every buffer below 4 GiB (L=%u R=%u nc=%u)", c->g.L,
                    c->g.R, c->g.nc);
  if (ty == 0 && tz == 0) {
    pick_fused(c);
    if (!c->fused) return msa_fail(MSA_ERR_ARG, "no one-pass tile for L=%u R=%u", c->g.L, c->g.R);
    c->fused = mode;
    return MSA_OK;
  }
  if (!fused_tile_ok(c->g, (uint32_t)ty, (uint32_t)tz))
    return msa_fail(MSA_ERR_ARG, "one-pass tile ty=%d tz=%d invalid for L=%u R=%u", ty, tz, c->g.L, c->g.R);
  set_fused(c, mode, (uint32_t)ty, (uint32_t)tz);
  return MSA_OK;
}

extern "C" int msa_get_tiling(msa_ctx *c, int32_t *mode, int32_t *ty, int32_t *tz) {
  if (!c || !mode || !ty || !tz) return msa_fail(MSA_ERR_ARG, "NULL argument");
  *mode = c->res ? 4 : (c->pair ? 3 : c->fused);
  *ty = (int32_t)(c->pair ? c->pty : c->fty);
  *tz = (int32_t)(c->pair ? c->ptz : c->ftz);
  return MSA_OK;
}

extern "C" int msa_set_refresh_mode(msa_ctx *c, int32_t mode) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  if (mode < -1 || mode > 2) return msa_fail(MSA_ERR_ARG, "refresh mode %d (-1 .. 2)", mode);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->dstream) HIPCHK(hipStreamSynchronize(c->dstream));
  c->pd.live = false;
  c->predraw = mode;
  return MSA_OK;
}

extern "C" int msa_get_refresh_mode(msa_ctx *c, int32_t *mode) {
  if (!c || !mode) return msa_fail(MSA_ERR_ARG, "NULL argument");
  *mode = refresh_mode(c);
  return MSA_OK;
}

// Host arrays in canonical order go through a device scratch buffer of the
// largest array's size, allocated for the call: copied as they are, then
// permuted by k_relayout on the device (the host does no per-site work).
static int relayout(msa_ctx *c, bool to_canon, bool dem, const void *in, void *out0, void *out1,
                    uint32_t nplane, int cmajor) {
  const uint64_t n = (uint64_t)nplane * c->g.V;
  const dim3 grid((uint32_t)((n + 255) / 256));
  if (to_canon && dem)
    hipLaunchKernelGGL((k_relayout<true, true>), grid, dim3(256), 0, c->stream, in, out0, out1, c->g, nplane, cmajor);
  else if (to_canon)
    hipLaunchKernelGGL((k_relayout<true, false>), grid, dim3(256), 0, c->stream, in, out0, out1, c->g, nplane, cmajor);
  else if (dem)
    hipLaunchKernelGGL((k_relayout<false, true>), grid, dim3(256), 0, c->stream, in, out0, out1, c->g, nplane, cmajor);
  else
    hipLaunchKernelGGL((k_relayout<false, false>), grid, dim3(256), 0, c->stream, in, out0, out1, c->g, nplane, cmajor);
  HIPCHK(hipGetLastError());
  return MSA_OK;
}

struct Scratch {
  void *p = nullptr;
  ~Scratch() { if (p) (void)hipFree(p); }
};

extern "C" int msa_upload_state(msa_ctx *c, const uint64_t *J, const uint64_t *spin,
                                const uint64_t *d0, const uint64_t *d1) {
  if (!c || !J || !spin) return msa_fail(MSA_ERR_ARG, "NULL argument");
  if ((d0 == nullptr) != (d1 == nullptr)) return msa_fail(MSA_ERR_ARG, "d0 and d1 must both be given");
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint64_t V = g.V, nd = (uint64_t)g.R * g.nc * V;
  uint64_t need = 24 * V;
  if (8 * g.R * V > need) need = 8 * g.R * V;
  if (d0 && 16 * nd > need) need = 16 * nd;
  HIPCHK(hipStreamSynchronize(c->stream));
  Scratch x;
  if (hipMalloc(&x.p, need) != hipSuccess) return msa_fail(MSA_ERR_NOMEM, "transfer scratch of %llu bytes", (unsigned long long)need);
  uint64_t *xb = static_cast<uint64_t *>(x.p);
  int rc;
  HIPCHK(hipMemcpyAsync(xb, J, 8 * 3 * V, hipMemcpyHostToDevice, c->stream));
  if ((rc = relayout(c, false, false, xb, c->jp, nullptr, 3, 1))) return rc;
  HIPCHK(hipMemcpyAsync(xb, spin, 8 * g.R * V, hipMemcpyHostToDevice, c->stream));
  if ((rc = relayout(c, false, false, xb, c->sp[0], nullptr, g.R, 0))) return rc;
  c->scur = 0;
  c->cur = 0;
  if (d0) {
    HIPCHK(hipMemcpyAsync(xb, d0, 8 * nd, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(xb + nd, d1, 8 * nd, hipMemcpyHostToDevice, c->stream));
    if ((rc = relayout(c, false, true, xb, c->dm[0], xb + nd, g.R * g.nc, 0))) return rc;
  } else {
    HIPCHK(hipMemsetAsync(c->dm[0], 0, 16 * nd, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

extern "C" int msa_download_state(msa_ctx *c, uint64_t *spin, uint64_t *d0, uint64_t *d1) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint64_t V = g.V, nd = (uint64_t)g.R * g.nc * V;
  uint64_t need = spin ? 8 * g.R * V : 0;
  if ((d0 || d1) && 16 * nd > need) need = 16 * nd;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!need) return MSA_OK;
  Scratch x;
  if (hipMalloc(&x.p, need) != hipSuccess) return msa_fail(MSA_ERR_NOMEM, "transfer scratch of %llu bytes", (unsigned long long)need);
  uint64_t *xb = static_cast<uint64_t *>(x.p);
  int rc;
  if (spin) {
    if ((rc = relayout(c, true, false, c->sp[c->scur], xb, nullptr, g.R, 0))) return rc;
    HIPCHK(hipMemcpyAsync(spin, xb, 8 * g.R * V, hipMemcpyDeviceToHost, c->stream));
  }
  if (d0 || d1) {
    if ((rc = relayout(c, true, true, c->dm[c->cur], xb, xb + nd, g.R * g.nc, 0))) return rc;
    if (d0) HIPCHK(hipMemcpyAsync(d0, xb, 8 * nd, hipMemcpyDeviceToHost, c->stream));
    if (d1) HIPCHK(hipMemcpyAsync(d1, xb + nd, 8 * nd, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

extern "C" int msa_download_wait(msa_ctx *c) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  if (!c->xfer_busy) return MSA_OK;
  HIPCHK(hipSetDevice(c->device));
  c->xfer_busy = false;
  HIPCHK(hipEventSynchronize(c->e_xfer));
  return MSA_OK;
}

// The state as the work queued so far leaves it: relaid out on the context's
// stream into a staging buffer the context keeps (spins, then d0, then d1),
// which a second stream copies to the host arrays while later work on the
// context's stream proceeds.
extern "C" int msa_download_state_async(msa_ctx *c, uint64_t *spin, uint64_t *d0, uint64_t *d1) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  int rc = msa_download_wait(c);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint64_t nS = (uint64_t)g.R * g.V, nd = (uint64_t)g.R * g.nc * g.V;
  const bool dem = d0 || d1;
  const size_t need = 8 * ((spin ? nS : 0) + (dem ? 2 * nd : 0));
  if (!need) return MSA_OK;
  if (c->xfer_bytes < need) {
    (void)hipFree(c->xfer);
    c->xfer = nullptr;
    c->xfer_bytes = 0;
    if (hipMalloc(&c->xfer, need) != hipSuccess)
      return msa_fail(MSA_ERR_NOMEM, "staging buffer of %zu bytes", need);
    c->xfer_bytes = need;
  }
  if (!c->xstream && hipStreamCreateWithFlags(&c->xstream, hipStreamNonBlocking) != hipSuccess)
    return msa_fail(MSA_ERR_HIP, "copy stream");
  if ((!c->e_snap && hipEventCreateWithFlags(&c->e_snap, hipEventDisableTiming) != hipSuccess) ||
      (!c->e_xfer && hipEventCreateWithFlags(&c->e_xfer, hipEventDisableTiming) != hipSuccess))
    return msa_fail(MSA_ERR_HIP, "copy events");
  uint64_t *xs = static_cast<uint64_t *>(c->xfer), *xd = xs + (spin ? nS : 0);
  if (spin && (rc = relayout(c, true, false, c->sp[c->scur], xs, nullptr, g.R, 0))) return rc;
  if (dem && (rc = relayout(c, true, true, c->dm[c->cur], xd, xd + nd, g.R * g.nc, 0))) return rc;
  HIPCHK(hipEventRecord(c->e_snap, c->stream));
  HIPCHK(hipStreamWaitEvent(c->xstream, c->e_snap, 0));
  if (spin) HIPCHK(hipMemcpyAsync(spin, xs, 8 * nS, hipMemcpyDeviceToHost, c->xstream));
  if (d0) HIPCHK(hipMemcpyAsync(d0, xd, 8 * nd, hipMemcpyDeviceToHost, c->xstream));
  if (d1) HIPCHK(hipMemcpyAsync(d1, xd + nd, 8 * nd, hipMemcpyDeviceToHost, c->xstream));
  HIPCHK(hipEventRecord(c->e_xfer, c->xstream));
  c->xfer_busy = true;
  return MSA_OK;
}

extern "C" int msa_download_couplings(msa_ctx *c, uint64_t *J) {
  if (!c || !J) return msa_fail(MSA_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  Scratch x;
  if (hipMalloc(&x.p, 24 * c->g.V) != hipSuccess) return msa_fail(MSA_ERR_NOMEM, "transfer scratch");
  const int rc = relayout(c, true, false, c->jp, x.p, nullptr, 3, 1);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(J, x.p, 24 * c->g.V, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

extern "C" int msa_init_philox(msa_ctx *c, uint64_t seed) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  const uint64_t V = c->g.V;
  const uint32_t nb = (uint32_t)((V + 255) / 256);
  hipLaunchKernelGGL(k_init_philox, dim3(nb), dim3(256), 0, c->stream, c->sp[0], c->jp, c->g,
                     (uint32_t)seed, (uint32_t)(seed >> 32));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(c->dm[0], 0, 16ull * c->g.R * c->g.nc * V, c->stream));
  c->scur = 0;
  c->cur = 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

extern "C" int msa_set_luts(msa_ctx *c, int32_t nlut, const msa_lut *luts) {
  if (!c || !luts || nlut < 1) return msa_fail(MSA_ERR_ARG, "need at least one LUT pair");
  for (int32_t k = 0; k < 2 * nlut; ++k)
    if (!lut_valid(luts[k])) return msa_fail(MSA_ERR_ARG, "LUT %d is malformed", k);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->dstream) HIPCHK(hipStreamSynchronize(c->dstream));
  c->pd.live = false;
  if (c->luts_owned) (void)hipFree(c->luts);
  c->luts = nullptr;
  c->luts_owned = true;
  c->nlut = 0;
  const int rc = msa_build_fine_luts(c->stream, nlut, luts, &c->luts);
  if (rc) return rc;
  c->nlut = nlut;
  return MSA_OK;
}

int msa_ctx_share_luts(msa_ctx *c, msa::FineLut *luts, int32_t nlut) {
  if (c->dstream) (void)hipStreamSynchronize(c->dstream);
  c->pd.live = false;
  if (c->luts_owned) (void)hipFree(c->luts);
  c->luts = luts;
  c->nlut = nlut;
  c->luts_owned = false;
  return MSA_OK;
}

extern "C" int msa_set_refresh(msa_ctx *c, int32_t nref, const msa_u32x4 *ctr,
                               const uint32_t *key, const int32_t *lut_idx) {
  if (!c || !ctr || !key || !lut_idx || nref < 1) return msa_fail(MSA_ERR_ARG, "bad refresh table");
  for (int32_t k = 0; k < nref; ++k)
    if (lut_idx[k] < 0 || lut_idx[k] >= c->nlut)
      return msa_fail(MSA_ERR_RANGE, "lut_idx[%d]=%d outside 0..%d (set LUTs first)", k, lut_idx[k],
                  c->nlut - 1);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->dstream) HIPCHK(hipStreamSynchronize(c->dstream));
  c->pd.live = false;
  c->nref = 0;
  const size_t R = c->g.R;
  if (nref > c->ctr_cap) {           // buffers are kept and reused when large enough
    (void)hipFree(c->ctr);
    c->ctr = nullptr;
    c->ctr_cap = 0;
    if (hipMalloc(&c->ctr, sizeof(uint4) * R * (size_t)nref) != hipSuccess)
      return msa_fail(MSA_ERR_NOMEM, "refresh table of %d entries", nref);
    c->ctr_cap = nref;
  }
  if (!c->key && hipMalloc(&c->key, sizeof(uint32_t) * R) != hipSuccess)
    return msa_fail(MSA_ERR_NOMEM, "refresh keys");
  HIPCHK(hipMemcpyAsync(c->ctr, ctr, sizeof(uint4) * R * (size_t)nref, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(c->key, key, sizeof(uint32_t) * R, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->lut_idx.assign(lut_idx, lut_idx + nref);
  c->nref = nref;
  return MSA_OK;
}

static void launch_fused(msa_ctx *c, const FusedArgs &fa) {
  const dim3 block(c->g.Lh * c->fny);
  const size_t lds = (size_t)3 * c->fny * c->g.Lh * c->g.R * 8;
  if (c->fused == 2) {
    const dim3 grid((c->g.L / c->fty) * (c->g.L / c->ftz));
    switch (c->g.R) {
      case 1: hipLaunchKernelGGL(k_sweep_fused2<1>, grid, block, lds, c->stream, fa); break;
      case 2: hipLaunchKernelGGL(k_sweep_fused2<2>, grid, block, lds, c->stream, fa); break;
      default: hipLaunchKernelGGL(k_sweep_fused2<4>, grid, block, lds, c->stream, fa); break;
    }
    return;
  }
  const dim3 grid(c->g.L / c->fty, c->g.L / c->ftz);
  switch (c->g.R) {
    case 1: hipLaunchKernelGGL(k_sweep_fused<1>, grid, block, lds, c->stream, fa); break;
    case 2: hipLaunchKernelGGL(k_sweep_fused<2>, grid, block, lds, c->stream, fa); break;
    default: hipLaunchKernelGGL(k_sweep_fused<4>, grid, block, lds, c->stream, fa); break;
  }
}

// Two refresh-free iterations of a one-layer lattice in one pass.  The pass
// reads the current buffers and writes the other ones, so each ping-pong
// index flips once for the two iterations.  The first iteration's demons
// come from dmi: the current buffer, or the demons drawn ahead for a refresh.
// The second iteration takes the first one's walked demons, or, when it
// refreshes (dmi2 not NULL, a tile pair_d2_ok accepts), the demons drawn for
// its refresh.
static void launch_pair(msa_ctx *c, const ulonglong2 *dmi, const ulonglong2 *dmi2, bool dm_store) {
  PairArgs pa{};
  pa.dm_store = dm_store;
  pa.sp_in = c->sp[c->scur];
  pa.sp_out = c->sp[c->scur ^ 1];
  pa.jp = c->jp;
  pa.dm_in = dmi;
  pa.dm_in2 = dmi2;
  pa.dm_out = c->dm[c->cur ^ 1];
  pa.g = c->g;
  pa.ty = dmi2 ? c->qty : c->pty;
  pa.tz = dmi2 ? c->qtz : c->ptz;
  pa.ny = pa.ty + 6;
  const dim3 grid((c->g.L / pa.ty) * (c->g.L / pa.tz) * c->g.R), block(c->g.Lh * pa.ny);
  if (dmi2)
    hipLaunchKernelGGL(k_sweep_pair<true>, grid, block, pair_lds_d2(c->g, pa.ty), c->stream, pa);
  else
    hipLaunchKernelGGL(k_sweep_pair<false>, grid, block, pair_lds(c->g, c->pty), c->stream, pa);
  c->scur ^= 1;
  c->cur ^= 1;
}

// Iterations of a resident launch: the current buffers in, the buffers the
// per-iteration launches would leave current out (the spin index flips nc
// times per iteration, the demon index once).
static void launch_resident(msa_ctx *c, ResArgs &ra) {
  const Geo &g = c->g;
  const uint32_t H = (uint32_t)g.H;
  const uint32_t flips = ra.n * g.nc;
  ra.sp_in = c->sp[c->scur];
  ra.sp_out = c->sp[c->scur ^ (flips & 1)];
  ra.jp = c->jp;
  ra.dm_in = c->dm[c->cur];
  ra.dm_out = c->dm[c->cur ^ (ra.n & 1)];
  ra.pre = c->pre;
  ra.g = g;
  ra.mLh = div_recip_of(g.Lh);
  ra.mL = div_recip_of(g.L);
  const size_t lds = res_lds(g);
  if (res_x(g)) {
    if (H <= 1024)
      hipLaunchKernelGGL(k_sweep_resident_x<1>, dim3(g.R), dim3((H + 63) / 64 * 64), lds, c->stream, ra);
    else
      hipLaunchKernelGGL(k_sweep_resident_x<2>, dim3(g.R), dim3(1024), lds, c->stream, ra);
  } else if (H <= 1024)
    hipLaunchKernelGGL(k_sweep_resident<1>, dim3(g.R), dim3((H + 63) / 64 * 64), lds, c->stream, ra);
  else
    hipLaunchKernelGGL(k_sweep_resident<2>, dim3(g.R), dim3(1024), lds, c->stream, ra);
  c->scur ^= flips & 1;
  c->cur ^= ra.n & 1;
}

// The refresh mode msa_run uses (0 fused, 1 drawn ahead, 2 drawn ahead with
// D2 pairs): the one set, or for -1 (the default) the one msa_autotune
// measured fastest for this geometry, else demons drawn ahead where that
// measured faster -- geometries with a D2 sweep-pair tile (L up to 144) and
// small lattices (L <= 48), whose fused refresh sweeps are bound by their
// launches.
static int refresh_mode(const msa_ctx *c) {
  if (c->predraw >= 0) return c->predraw;
  if (c->tuned_rmode >= 0) return c->tuned_rmode;
  // untuned: D2 pairs where their tile has at least 6 own rows (thinner tiles
  // recompute so much halo that single one-pass sweeps measured faster)
  if (c->qty >= 6) return 2;
  return (c->qty > 0 || c->g.L <= 48) ? 1 : 0;
}

// Refreshes before iteration t (the index of the first refresh at or after t).
static int64_t refreshes_before(uint64_t t, int64_t every) {
  if (every > 0) return (int64_t)((t + (uint64_t)every - 1) / (uint64_t)every);
  if (every == -1) {
    if (t == 0) return 0;
    int64_t k = 1;                        // refresh k >= 1 is at iteration 2^(k-1)
    while (k < 64 && ((uint64_t)1 << (k - 1)) < t) ++k;
    return k;
  }
  return 0;
}

// Demons drawn ahead (predraw mode): words of one refresh, R nc V.
static size_t pre_words(const msa_ctx *c) { return (size_t)c->g.R * c->g.nc * c->g.V; }

// Stream, events and the two halves of the drawn-demon buffer, made on the
// first run in predraw mode: each half holds pK refreshes, at most 256 MiB
// (at least one refresh) and kDrawMax.
static int predraw_setup(msa_ctx *c) {
  if (c->pre) return MSA_OK;
  const size_t per = pre_words(c) * sizeof(ulonglong2);
  size_t K = ((size_t)256 << 20) / per;
  if (K < 1) K = 1;
  if (K > (size_t)kDrawMax) K = kDrawMax;
  if (!c->dstream && hipStreamCreateWithFlags(&c->dstream, hipStreamNonBlocking) != hipSuccess)
    return msa_fail(MSA_ERR_HIP, "draw stream creation failed");
  for (int h = 0; h < 2; ++h)
    if ((!c->e_drawn[h] && hipEventCreateWithFlags(&c->e_drawn[h], hipEventDisableTiming) != hipSuccess) ||
        (!c->e_used[h] && hipEventCreateWithFlags(&c->e_used[h], hipEventDisableTiming) != hipSuccess))
      return msa_fail(MSA_ERR_HIP, "draw event creation failed");
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu < 1)
    ncu = 256;
  c->ncu = (uint32_t)ncu;
  if (hipMalloc(&c->pre, 2 * K * per) != hipSuccess) {
    c->pre = nullptr;
    return msa_fail(MSA_ERR_NOMEM, "drawn-demon buffer of %lld bytes", (long long)(2 * K * per));
  }
  c->pK = (uint32_t)K;
  return MSA_OK;
}

// Refreshes kb, kb + 1, ... drawn pK at a time into the two halves of c->pre
// (chunk j: refreshes kb + j pK ..., half j & 1), as far as the refresh
// tables reach.  The sweeps that read chunk j wait for its draw
// (e_drawn[j & 1]); once every kernel reading chunk j is queued, e_used[j & 1]
// is recorded after them, and the draw of chunk j + 2 into the same half
// waits for it.  Draws run one chunk ahead of the sweeps, so a draw overlaps
// the sweeps of the chunk before it, also across msa_run calls: a call whose
// first refresh lies in a chunk already drawn (or the next one) continues the
// pipeline (c->pd); new tables or a new mode end it.
static int draw_chunk(msa_ctx *c, const Predraw &pd, int64_t jj) {
  const int64_t k0 = pd.kb + jj * (int64_t)c->pK;
  const int64_t nk = c->nref - k0 < (int64_t)c->pK ? c->nref - k0 : (int64_t)c->pK;
  HIPCHK(hipStreamWaitEvent(c->dstream, c->e_used[jj & 1], 0));
  DrawArgs a{};
  a.out = c->pre + (size_t)(jj & 1) * c->pK * pre_words(c);
  a.ctr = c->ctr + (size_t)k0 * c->g.R;
  a.key = c->key;
  a.luts = c->luts;
  a.g = c->g;
  const uint32_t Lh = c->g.Lh, rows = c->g.L * c->g.L;
  a.nr = 768 / Lh < 1 ? 1 : 768 / Lh;
  if (a.nr > rows) a.nr = rows;
  a.mLh = div_recip_of(Lh);
  a.mL = div_recip_of(c->g.L);
  for (int64_t j = 0; j < nk; ++j) a.lidx[j] = c->lut_idx[(size_t)(k0 + j)];
  // one block per CU: blocks of at most 768 threads (three waves per SIMD,
  // 80 VGPRs each) with more than half of the CU's LDS, so a sweep block of
  // up to two waves per SIMD at 136 VGPRs still fits beside them; the CUs
  // split over the launch's refreshes
  const uint32_t nbr = (rows + a.nr - 1) / a.nr;
  uint32_t gx = (c->ncu + (uint32_t)nk - 1) / (uint32_t)nk;
  if (gx > nbr) gx = nbr;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(k_refresh_draw, dim3(gx, (uint32_t)nk), dim3(a.nr * Lh), kDrawLds,
                     c->dstream, a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->e_drawn[jj & 1], c->dstream));
  return MSA_OK;
}

// Queue the draws up to chunk `upto` (at most the last chunk the tables
// reach) whose half is free: chunk jj needs chunk jj - 2 released.
static int draw_ahead(msa_ctx *c, Predraw &pd, int64_t upto) {
  const int64_t chunks = (c->nref - pd.kb + c->pK - 1) / (int64_t)c->pK;
  if (upto >= chunks) upto = chunks - 1;
  while (pd.drawn <= upto && pd.drawn < pd.released + 2) {
    const int rc = draw_chunk(c, pd, pd.drawn);
    if (rc) return rc;
    ++pd.drawn;
  }
  return MSA_OK;
}

// The drawn demons of refresh k, the sweep stream made to wait for them.
static int drawn_demons(msa_ctx *c, Predraw &pd, int64_t k, const ulonglong2 **out) {
  const int64_t j = (k - pd.kb) / (int64_t)c->pK;
  int rc = draw_ahead(c, pd, j + 1);
  if (rc) return rc;
  if (pd.drawn <= j) return msa_fail(MSA_ERR_STATE, "refresh %lld: its draw is not queued", (long long)k);
  if (j > pd.waited) {
    HIPCHK(hipStreamWaitEvent(c->stream, c->e_drawn[j & 1], 0));
    pd.waited = j;
  }
  *out = c->pre + ((size_t)(j & 1) * c->pK + (size_t)(k - pd.kb - j * (int64_t)c->pK)) * pre_words(c);
  return MSA_OK;
}

// Every kernel reading refreshes below k_next is queued: release their
// chunks and draw ahead into the halves that frees.
static int release_before(msa_ctx *c, Predraw &pd, int64_t k_next) {
  const int64_t j = (k_next - pd.kb) / (int64_t)c->pK;
  while (pd.released < j && pd.released < pd.drawn) {
    HIPCHK(hipEventRecord(c->e_used[pd.released & 1], c->stream));
    ++pd.released;
  }
  return draw_ahead(c, pd, j + 1);
}

static int launch_run(msa_ctx *c, uint64_t t0, int64_t niter, int64_t every) {
  if (niter < 0) return msa_fail(MSA_ERR_ARG, "niter=%lld", (long long)niter);
  if (every < -1) return msa_fail(MSA_ERR_ARG, "refresh_every=%lld", (long long)every);
  const int rmode = refresh_mode(c);
  const bool pdraw = rmode >= 1 && c->w_n == 0 && c->nlut > 0;
  Predraw &pd = c->pd;
  int64_t ke = 0;                          // refreshes below ke are this call's
  if (pdraw) {
    const int64_t kb = refreshes_before(t0, every);
    ke = refreshes_before(t0 + (uint64_t)niter, every);
    if (ke > c->nref) ke = c->nref;        // the loop reports the missing refresh
    if (ke > kb) {
      int rc = predraw_setup(c);
      if (rc) return rc;
      const int64_t j0 = pd.live && kb >= pd.kb ? (kb - pd.kb) / (int64_t)c->pK : -1;
      if (j0 >= pd.released && j0 <= pd.drawn) {
        // the pipeline goes on: chunks before this call's first are done
        rc = release_before(c, pd, kb);
        if (rc) return rc;
      } else {
        pd = Predraw{};
        pd.kb = kb;
        pd.live = true;
        // both halves are free once the work queued so far has run
        HIPCHK(hipEventRecord(c->e_used[0], c->stream));
        HIPCHK(hipEventRecord(c->e_used[1], c->stream));
      }
    }
  }
  const bool d2 = pdraw && rmode == 2 && c->qty > 0;
  for (int64_t n = 0; n < niter; ++n) {
    const uint64_t t = t0 + (uint64_t)n;
    if (c->res) {
      // resident: the iterations from n on whose refresh, if any, is drawn
      // ahead (a walker-stream refresh only at the first; a refresh drawn
      // inside the sweep runs below)
      ResArgs ra{};
      int64_t m = n, klast = -1;
      for (; m < niter && m - n < kResMax; ++m) {
        const uint64_t tm = t0 + (uint64_t)m;
        int64_t k = refresh_index(tm, every);
        if (k >= 0 && c->w_n > 0) {
          if (m > n) break;
          const int rc = walkers_refresh(c, k);
          if (rc) return rc;
          k = -1;
        }
        if (k >= 0 && (!pdraw || k >= c->nref)) break;
        ra.slot[m - n] = -1;
        if (k >= 0) {
          const ulonglong2 *p = nullptr;
          const int rc = drawn_demons(c, pd, k, &p);
          if (rc) return rc;
          ra.slot[m - n] = (int16_t)((p - c->pre) / (ptrdiff_t)pre_words(c));
          klast = k;
        }
      }
      if (m > n) {
        ra.n = (uint32_t)(m - n);
        ra.dm_load = ra.slot[0] < 0;
        const int64_t kn = m < niter ? refresh_index(t0 + (uint64_t)m, every) : -1;
        ra.dm_store = !(pdraw && kn >= 0 && kn < ke);
        launch_resident(c, ra);
        if (klast >= 0) {
          const int rc = release_before(c, pd, klast + 1);
          if (rc) return rc;
        }
        n = m - 1;
        continue;
      }
    }
    int64_t k = refresh_index(t, every);
    if (k >= 0 && c->w_n > 0) {           // the reference's walker streams
      const int rc = walkers_refresh(c, k);
      if (rc) return rc;
      k = -1;
    }
    if (k >= 0 && (k >= c->nref || c->nlut == 0))
      return msa_fail(MSA_ERR_RANGE, "iteration %llu needs refresh %lld but %d are set",
                  (unsigned long long)t, (long long)k, c->nref);
    const ulonglong2 *dmi = c->dm[c->cur];
    const bool drawn = k >= 0 && pdraw;
    if (drawn) {
      // the refresh's demons, drawn ahead, are this sweep's input demons
      const int rc = drawn_demons(c, pd, k, &dmi);
      if (rc) return rc;
    }
    if ((k < 0 || drawn) && n + 1 < niter) {
      const int64_t k2 = refresh_index(t + 1, every);
      const bool next_drawn = k2 >= 0 && d2 && k2 < ke;
      if ((k2 < 0 && c->pair) || next_drawn) {
        const ulonglong2 *dmi2 = nullptr;
        if (next_drawn) {
          const int rc = drawn_demons(c, pd, k2, &dmi2);
          if (rc) return rc;
        }
        // walked demons the next iteration's drawn refresh replaces are not written
        const int64_t k3 = n + 2 < niter ? refresh_index(t + 2, every) : -1;
        launch_pair(c, dmi, dmi2, !(pdraw && k3 >= 0 && k3 < ke));
        if (drawn || next_drawn) {
          const int rc = release_before(c, pd, (next_drawn ? k2 : k) + 1);
          if (rc) return rc;
        }
        ++n;
        continue;
      }
    }
    // a refresh drawn ahead runs the refresh-free kernels
    const int64_t kf = drawn ? -1 : k;
    const int64_t kn = n + 1 < niter ? refresh_index(t + 1, every) : -1;
    const bool dm_store = !(pdraw && kn >= 0 && kn < ke);
    ulonglong2 *dmo = c->dm[c->cur ^ 1];
    for (uint32_t l = 0; l < c->g.nc; ++l) {
      const uint64_t *A = c->sp[c->scur];
      uint64_t *B = c->sp[c->scur ^ 1];
      if (kf < 0 && c->fused) {
        FusedArgs fa{};
        fa.sp_in = A;
        fa.sp_out = B;
        fa.jp = c->jp;
        fa.dm_in = dmi;
        fa.dm_out = dmo;
        fa.g = c->g;
        fa.layer = (int)l;
        fa.ty = c->fty;
        fa.tz = c->ftz;
        fa.ny = c->fny;
        fa.yhalo = c->fyhalo;
        fa.dm_store = dm_store;
        launch_fused(c, fa);
      } else {
        SweepArgs a{};
        a.jp = c->jp;
        a.dm_in = dmi;
        a.dm_out = dmo;
        a.g = c->g;
        a.layer = (int)l;
        a.nr = c->hnr;
        a.rs = c->hrs;
        a.mLh = div_recip_of(c->g.Lh);
        a.mL = div_recip_of(c->g.L);
        a.mG = div_recip_of(c->hnr * c->g.Lh);
        if (kf >= 0) {
          a.ctr = c->ctr + (size_t)kf * c->g.R;
          a.key = c->key;
          a.lut = c->luts + 2 * (size_t)c->lut_idx[kf];
        }
        const dim3 grid(c->hgrid), block(c->hnr * c->g.Lh * c->hrs);
        for (int col = 0; col < 2; ++col) {
          a.colour = col;
          a.sp_own = A;
          a.sp_nb = col ? B : A;
          a.sp_out = B;
          const bool one = c->hrs == c->g.R;   // also R = 1
          if (kf >= 0 && one)
            hipLaunchKernelGGL((k_halfsweep<kRefresh, true>), grid, block, kRefreshLds, c->stream, a);
          else if (kf >= 0)
            hipLaunchKernelGGL((k_halfsweep<kRefresh, false>), grid, block, kRefreshLds, c->stream, a);
          else if (one)
            hipLaunchKernelGGL((k_halfsweep<kMicro, true>), grid, block, 0, c->stream, a);
          else
            hipLaunchKernelGGL((k_halfsweep<kMicro, false>), grid, block, 0, c->stream, a);
        }
      }
      c->scur ^= 1;
    }
    c->cur ^= 1;
    if (drawn) {
      const int rc = release_before(c, pd, k + 1);
      if (rc) return rc;
    }
  }
  HIPCHK(hipGetLastError());
  return MSA_OK;
}

extern "C" int msa_run(msa_ctx *c, uint64_t t0, int64_t niter, int64_t refresh_every) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  return launch_run(c, t0, niter, refresh_every);
}

extern "C" int msa_run_timed(msa_ctx *c, uint64_t t0, int64_t niter, int64_t refresh_every,
                             float *ms) {
  if (!c || !ms) return msa_fail(MSA_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int rc = launch_run(c, t0, niter, refresh_every);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  HIPCHK(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return MSA_OK;
}

static int thermal_args_ok(msa_ctx *c, int64_t niter, const msa_lut *lut, const msa_u32x4 *ctr,
                           const uint32_t *key) {
  if (!c || !lut || (niter > 0 && (!ctr || !key))) return msa_fail(MSA_ERR_ARG, "NULL argument");
  if (niter < 0) return msa_fail(MSA_ERR_ARG, "niter=%lld", (long long)niter);
  if (c->g.L % 8) return msa_fail(MSA_ERR_ARG, "thermal sweeps use the MSC word map: L=%u must be a multiple of 8", c->g.L);
  if (!lut_valid(*lut)) return msa_fail(MSA_ERR_ARG, "heat-bath LUT is malformed");
  return MSA_OK;
}

// Monte_Carlo with flag = 0 (main.cu:202-252) niter times: per iteration and
// parity p, the random words of parity p (counters dctr[(t*2 + p)*R ..]) and
// the Metropolis update of colour p, on the context stream; then wait.
static int thermal_sweeps(msa_ctx *c, int64_t niter, const msa::FineLut *tab, const uint4 *dctr,
                          const uint32_t *dkey, uint64_t *rnd) {
  const Geo &g = c->g;
  const uint32_t T = (uint32_t)(g.V / 128 * g.R);
  SweepArgs a{};
  a.jp = c->jp;
  a.g = g;
  a.nr = c->hnr;
  a.rs = c->hrs;
  a.mLh = div_recip_of(g.Lh);
  a.mL = div_recip_of(g.L);
  a.mG = div_recip_of(c->hnr * g.Lh);
  a.rnd = rnd;
  a.rT = T;
  const dim3 grid(c->hgrid), block(c->hnr * g.Lh * c->hrs);
  int rc = MSA_OK;
  for (int64_t t = 0; t < niter && !rc; ++t) {
    const uint64_t *A = c->sp[c->scur];
    uint64_t *B = c->sp[c->scur ^ 1];
    for (int col = 0; col < 2; ++col) {
      // Monte_Carlo flag = 0 (main.cu:202-252): bits of parity col, then colour col
      hipLaunchKernelGGL(k_gen_rndbits, dim3((T + 255) / 256), dim3(256), 0, c->stream,
                         dctr + ((size_t)t * 2 + col) * g.R, dkey, tab, g.R, T, rnd);
      a.colour = col;
      a.sp_own = A;
      a.sp_nb = col ? B : A;
      a.sp_out = B;
      if (c->hrs == g.R)
        hipLaunchKernelGGL((k_halfsweep<kThermal, true>), grid, block, 0, c->stream, a);
      else
        hipLaunchKernelGGL((k_halfsweep<kThermal, false>), grid, block, 0, c->stream, a);
    }
    if (hipGetLastError() != hipSuccess) rc = msa_fail(MSA_ERR_HIP, "thermal launch");
    c->scur ^= 1;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = msa_fail(MSA_ERR_HIP, "thermal sweeps");
  return rc;
}

extern "C" int msa_run_thermal(msa_ctx *c, int64_t niter, const msa_lut *lut, const msa_u32x4 *ctr,
                               const uint32_t *key) {
  int rc = thermal_args_ok(c, niter, lut, ctr, key);
  if (rc || niter == 0) return rc;
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint32_t T = (uint32_t)(g.V / 128 * g.R);
  if (!c->th_rnd && hipMalloc(&c->th_rnd, 3ull * T * sizeof(uint64_t)) != hipSuccess)
    return msa_fail(MSA_ERR_NOMEM, "thermal random words");
  // the table twice: msa_build_fine_luts builds pairs
  const msa_lut pair[2] = {*lut, *lut};
  msa::FineLut *tab = nullptr;
  rc = msa_build_fine_luts(c->stream, 1, pair, &tab);
  if (rc) return rc;
  uint4 *dctr = nullptr;
  uint32_t *dkey = nullptr;
  const size_t nctr = (size_t)niter * 2 * g.R;
  if (hipMalloc(&dctr, sizeof(uint4) * nctr) != hipSuccess || hipMalloc(&dkey, sizeof(uint32_t) * g.R) != hipSuccess) {
    rc = msa_fail(MSA_ERR_NOMEM, "thermal counters");
  } else if (hipMemcpyAsync(dctr, ctr, sizeof(uint4) * nctr, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
             hipMemcpyAsync(dkey, key, sizeof(uint32_t) * g.R, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    rc = msa_fail(MSA_ERR_HIP, "thermal counter upload");
  } else {
    rc = thermal_sweeps(c, niter, tab, dctr, dkey, c->th_rnd);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dctr);
  (void)hipFree(dkey);
  (void)hipFree(tab);
  return rc;
}

// ------------------------------------------------- fixed device scratch --
// The device memory msa_run_thermal allocates on every call (the raw LUT and
// its search table, the counters and keys, the thermal random words) lives
// here for msa_run_thermal_noalloc, which calls no allocator.  Every GPU the
// library runs on holds its own copy until the process ends; host code never
// uses the symbol itself (on the host it is a stand-in, not a device
// address) but the address hipGetSymbolAddress gives for the current device,
// fetched once per device by msa_create (thermal_scratch_init).
constexpr size_t kScratchBytes = 16u << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];

// With `device` current: the address of this device's scratch, once.
static int thermal_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return msa_fail(MSA_ERR_ARG, "device %d: the fixed scratch covers devices 0..%d", device, kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return MSA_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return msa_fail(MSA_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
  g_scratch_dev[device] = p;
  return MSA_OK;
}

static size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

// Scratch layout of one call: raw LUT | search table | keys | counters | random words.
static size_t thermal_scratch_bytes(const Geo &g, int64_t niter) {
  const size_t T = g.V / 128 * g.R;
  return up256(sizeof(msa_lut)) + up256(sizeof(msa::FineLut)) + up256(4 * (size_t)g.R) +
         up256(16 * (size_t)niter * 2 * g.R) + up256(24 * T);
}

extern "C" int msa_run_thermal_noalloc(msa_ctx *c, int64_t niter, const msa_lut *lut,
                                       const msa_u32x4 *ctr, const uint32_t *key) {
  int rc = thermal_args_ok(c, niter, lut, ctr, key);
  if (rc || niter == 0) return rc;
  const size_t need = thermal_scratch_bytes(c->g, niter);
  if (need > kScratchBytes)
    return msa_fail(MSA_ERR_RANGE,
                    "msa_run_thermal_noalloc: L=%u R=%u niter=%lld needs %zu bytes of the fixed "
                    "scratch, which holds %zu", c->g.L, c->g.R, (long long)niter, need, kScratchBytes);
  HIPCHK(hipSetDevice(c->device));
  uint8_t *base;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = static_cast<uint8_t *>(g_scratch_dev[c->device]);
  }
  if (!base) return msa_fail(MSA_ERR_STATE, "the fixed scratch of device %d is not set up", c->device);
  const Geo &g = c->g;
  const size_t T = g.V / 128 * g.R;
  msa_lut *raw = reinterpret_cast<msa_lut *>(base);
  uint8_t *q = base + up256(sizeof(msa_lut));
  msa::FineLut *tab = reinterpret_cast<msa::FineLut *>(q);
  q += up256(sizeof(msa::FineLut));
  uint32_t *dkey = reinterpret_cast<uint32_t *>(q);
  q += up256(4 * (size_t)g.R);
  uint4 *dctr = reinterpret_cast<uint4 *>(q);
  q += up256(16 * (size_t)niter * 2 * g.R);
  uint64_t *rnd = reinterpret_cast<uint64_t *>(q);
  (void)T;
  // every scratch byte a kernel reads is written first in stream order
  // (LUT, table, keys, counters by the copies and k_fine_luts; the random
  // words by k_gen_rndbits), so nothing needs clearing
  HIPCHK(hipStreamSynchronize(c->stream));
  rc = fine_luts_into(c->stream, 1, lut, raw, tab);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(dkey, key, sizeof(uint32_t) * g.R, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dctr, ctr, sizeof(uint4) * (size_t)niter * 2 * g.R, hipMemcpyHostToDevice,
                        c->stream));
  return thermal_sweeps(c, niter, tab, dctr, dkey, rnd);
}

// Queue the clear of the counts and k_measure on the current state.
static int measure_launch(msa_ctx *c) {
  HIPCHK(hipSetDevice(c->device));
  const Geo &g = c->g;
  const uint32_t nrow = meas_rows(g.R), rows = g.L * g.L;
  const size_t nword = 64 * (size_t)nrow;
  MeasArgs a{};
  a.sp = c->sp[c->scur];
  a.jp = c->jp;
  a.dm = c->dm[c->cur];
  a.g = g;
  // blocks of about 256 threads (at most 512): more, smaller blocks per CU
  // hide the latency of the row loads better (L = 160: 103 instead of 131 us
  // per call with 480-thread blocks)
  a.nr = g.Lh < 256 ? 256 / g.Lh : 1;
  a.mLh = div_recip_of(g.Lh);
  a.out = c->meas;
  // rows per thread: as many as leave one block per CU and job (fewer
  // blocks finish their counters' transposes sooner), at most 16 (the
  // width of the counters)
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu < 1)
    ncu = 256;
  const uint32_t njob = 2 * ((g.R + 1) / 2), groups = (rows + a.nr - 1) / a.nr;
  const uint32_t fill = (uint32_t)ncu / njob > 0 ? (uint32_t)ncu / njob : 1;
  a.steps = groups / fill < 1 ? 1 : (groups / fill > 16 ? 16 : groups / fill);
  const uint32_t nb = (groups + a.steps - 1) / a.steps;
  const uint32_t threads = (a.nr * g.Lh + 63) / 64 * 64;   // whole wavefronts
  HIPCHK(hipMemsetAsync(c->meas, 0, nword * sizeof(int32_t), c->stream));
  hipLaunchKernelGGL(k_measure, dim3(nb, njob), dim3(threads), 0, c->stream, a);
  HIPCHK(hipGetLastError());
  return MSA_OK;
}

// The counts [2R + 2(R/2)][64] from `host` into the caller's arrays.
static void measure_out(const msa_ctx *c, const int32_t *host, int32_t *unsat, int32_t *demon,
                        int32_t *q, int32_t *qlink) {
  const size_t nR = 64 * (size_t)c->g.R, nP = 64 * (size_t)(c->g.R / 2);
  if (unsat) memcpy(unsat, host, nR * sizeof(int32_t));
  if (demon) memcpy(demon, host + nR, nR * sizeof(int32_t));
  if (q && nP) memcpy(q, host + 2 * nR, nP * sizeof(int32_t));
  if (qlink && nP) memcpy(qlink, host + 2 * nR + nP, nP * sizeof(int32_t));
}

// A queued measurement's counts go to a page-locked array of the context
// (made on the first such call), an event after their copy.  The device
// counts are cleared and rewritten only by later work on the same stream,
// so a synchronous msa_measure in between leaves the queued copy intact.
extern "C" int msa_measure_async(msa_ctx *c) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  if (c->meas_busy)
    return msa_fail(MSA_ERR_STATE, "a measurement is queued: msa_measure_wait first");
  HIPCHK(hipSetDevice(c->device));
  const size_t nword = 64 * (size_t)meas_rows(c->g.R);
  if (!c->meas_host &&
      hipHostMalloc(reinterpret_cast<void **>(&c->meas_host), nword * sizeof(int32_t),
                    hipHostMallocDefault) != hipSuccess) {
    c->meas_host = nullptr;
    return msa_fail(MSA_ERR_NOMEM, "page-locked measurement array");
  }
  if (!c->e_meas && hipEventCreateWithFlags(&c->e_meas, hipEventDisableTiming) != hipSuccess) {
    c->e_meas = nullptr;
    return msa_fail(MSA_ERR_HIP, "measurement event");
  }
  const int rc = measure_launch(c);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->meas_host, c->meas, nword * sizeof(int32_t), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipEventRecord(c->e_meas, c->stream));
  c->meas_busy = true;
  return MSA_OK;
}

extern "C" int msa_measure_wait(msa_ctx *c, int32_t *unsat, int32_t *demon, int32_t *q,
                                int32_t *qlink) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  if (!c->meas_busy) return msa_fail(MSA_ERR_STATE, "no measurement is queued");
  HIPCHK(hipSetDevice(c->device));
  c->meas_busy = false;
  HIPCHK(hipEventSynchronize(c->e_meas));
  measure_out(c, c->meas_host, unsat, demon, q, qlink);
  return MSA_OK;
}

extern "C" int msa_measure(msa_ctx *c, int32_t *unsat, int32_t *demon, int32_t *q,
                           int32_t *qlink) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  const int rc = measure_launch(c);
  if (rc) return rc;
  int32_t host[64 * (2 * 16 + 2 * 8)];   // R <= 16 (check_geom)
  const size_t nword = 64 * (size_t)meas_rows(c->g.R);
  HIPCHK(hipMemcpyAsync(host, c->meas, nword * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  measure_out(c, host, unsat, demon, q, qlink);
  return MSA_OK;
}

// ------------------------------------------------------- copy probe ------
// Streaming copy of n 16-byte words: each block moves 4 x 256 words, every
// lane four 16-byte non-temporal loads in flight, then four stores (the
// fastest of the copy forms measured on this part, 6.1 TB/s for 2 GiB).
__global__ void __launch_bounds__(256) k_copy16(const ulonglong2 *__restrict__ src,
                                                ulonglong2 *__restrict__ dst, uint64_t n) {
  using u64x2 = msa::u64x2;
  const uint64_t base = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  const u64x2 *s = reinterpret_cast<const u64x2 *>(src);
  u64x2 *d = reinterpret_cast<u64x2 *>(dst);
  u64x2 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + 256 * k < n) v[k] = __builtin_nontemporal_load(s + base + 256 * k);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + 256 * k < n) __builtin_nontemporal_store(v[k], d + base + 256 * k);
}

extern "C" int msa_copy_probe(int device, int64_t bytes, int32_t iters, float *ms_per_copy) {
  if (bytes < 4096 || iters < 1 || !ms_per_copy) return msa_fail(MSA_ERR_ARG, "bad copy probe arguments");
  HIPCHK(hipSetDevice(device));
  const uint64_t n = (uint64_t)bytes / 16;
  ulonglong2 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MSA_OK;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    rc = msa_fail(MSA_ERR_NOMEM, "copy probe buffers");
  } else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = msa_fail(MSA_ERR_HIP, "copy probe stream");
  } else {
    (void)hipMemsetAsync(a, 1, n * 16, st);
    const uint32_t grid = (uint32_t)((n + 1023) / 1024);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, a, b, n);   // warm-up
    (void)hipEventRecord(e0, st);
    for (int32_t it = 0; it < iters; ++it)
      hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, a, b, n);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) rc = msa_fail(MSA_ERR_HIP, "copy probe sync");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_copy = ms / (float)iters;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

// ---------------------------------------------------------- autotuner ----
// Choice per (L, R, nc), measured once per process and reused.
struct TuneEntry {
  uint32_t L, R, nc;
  int32_t mode, ty, tz;
  float ms;
  int32_t rmode;            // refresh mode of the sweeps with a refresh
  int32_t fmode, fty, ftz;  // the fastest one-pass kernel and tile (fmode 0: none)
  float fms;
  int32_t ff;               // 1: single sweeps of the pair mode on that tile, 0: the default
};
static std::vector<TuneEntry> g_tuned;

// A heat-bath table for timing only, shaped like the real ones: the byte of
// eight independent bits, each set with probability q, with thresholds
// 2^64 P(byte <= n) in double precision (the table search costs what it costs
// on real tables, whose thresholds gather in few presearch buckets); every
// presearch entry spans 0..255 (valid for msa_set_luts).
static void timing_lut(msa_lut *t, double q) {
  double cum = 0;
  for (int n = 0; n < 256; ++n) {
    const int k = __builtin_popcount((unsigned)n);
    double p = 1;
    for (int b = 0; b < 8; ++b) p *= b < k ? q : 1 - q;
    cum += p;
    const double v = cum * 18446744073709551616.0;
    const uint64_t u = n == 255 || v >= 18446744073709551615.0 ? ~0ull : (uint64_t)v;
    t->umbrales[2 * n] = (uint32_t)(u >> 32);
    t->umbrales[2 * n + 1] = (uint32_t)u;
  }
  for (int n = 1; n < 256; ++n) {           // non-decreasing after rounding
    const uint64_t a = ((uint64_t)t->umbrales[2 * n - 2] << 32) | t->umbrales[2 * n - 1];
    const uint64_t b = ((uint64_t)t->umbrales[2 * n] << 32) | t->umbrales[2 * n + 1];
    if (b < a) {
      t->umbrales[2 * n] = t->umbrales[2 * n - 2];
      t->umbrales[2 * n + 1] = t->umbrales[2 * n - 1];
    }
  }
  for (int m = 0; m < 512; ++m) t->prebusqueda[m] = 0xFF00FF00u;
}

// Times `sweeps` sweeps with a refresh each in refresh mode `rm` on scratch
// context t, whose refresh tables cover them.
static int time_refresh(msa_ctx *t, int32_t sweeps, int rm, float *per) {
  int rc = msa_set_refresh_mode(t, rm);
  if (rc) return rc;
  float ms = 0;
  if ((rc = msa_run(t, 0, 4, 1))) return rc;                    // warm-up
  if ((rc = msa_run_timed(t, 4, sweeps, 1, &ms))) return rc;
  *per = ms / (float)sweeps;
  return MSA_OK;
}

// The tuned refresh mode, and for the sweep-pair mode the one-pass kernel
// and tile of its single sweeps: the fastest one-pass candidate where that
// timed faster with the refresh, else the default (msa_set_tiling's).
static void apply_tuned(msa_ctx *c, const TuneEntry &e) {
  c->tuned_rmode = e.rmode;
  if (c->pair && e.ff && e.fmode > 0) set_fused(c, e.fmode, (uint32_t)e.fty, (uint32_t)e.ftz);
}

extern "C" int msa_autotune(msa_ctx *c, int32_t sweeps, int32_t *mode, int32_t *ty, int32_t *tz,
                            float *ms_per_sweep) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  if (sweeps < 1) sweeps = 8;
  for (const TuneEntry &e : g_tuned)
    if (e.L == c->g.L && e.R == c->g.R && e.nc == c->g.nc) {
      int rc = msa_set_tiling(c, e.mode, e.ty, e.tz);
      if (rc) return rc;
      apply_tuned(c, e);
      if (mode) *mode = e.mode;
      if (ty) *ty = e.ty;
      if (tz) *tz = e.tz;
      if (ms_per_sweep) *ms_per_sweep = e.ms;
      return MSA_OK;
    }
  // candidates: the two-pass form and both one-pass kernels on every valid
  // tile with slabs of at most 32 planes
  struct Cand { int32_t mode, ty, tz; };
  std::vector<Cand> cands{{0, 0, 0}};
  for (uint32_t y = 1; y <= c->g.L; ++y)
    for (uint32_t z = 2; z <= c->g.L && z <= 32; ++z)
      if (fused_tile_ok(c->g, y, z))
        for (int32_t m = 1; m <= (narrow_ok(c->g) ? 2 : 1); ++m)
          cands.push_back({m, (int32_t)y, (int32_t)z});
  for (uint32_t y = 1; y <= c->g.L; ++y)
    for (uint32_t z = 4; z <= c->g.L && z <= 64; ++z)
      if (pair_tile_ok(c->g, y, z)) cands.push_back({3, (int32_t)y, (int32_t)z});
  if (res_fits(c->g)) cands.push_back({4, 0, 0});
  msa_ctx *t = nullptr;
  int rc = msa_create(c->device, &c->geom, &t);
  if (rc) return rc;
  rc = msa_init_philox(t, 0xA070u);
  TuneEntry best{c->g.L, c->g.R, c->g.nc, 0, 0, 0, 1e30f, 0, 0, 0, 0, 1e30f, 0};
  for (const Cand &cd : cands) {
    if (rc) break;
    if ((rc = msa_set_tiling(t, cd.mode, cd.ty, cd.tz))) break;
    float ms = 0;
    if ((rc = msa_run(t, 0, 2, 0))) break;                      // warm-up
    if ((rc = msa_run_timed(t, 0, sweeps, 0, &ms))) break;
    const float per = ms / (float)sweeps;
    if (per < best.ms) {
      best.mode = cd.mode;
      best.ty = cd.ty;
      best.tz = cd.tz;
      best.ms = per;
    }
    if ((cd.mode == 1 || cd.mode == 2) && per < best.fms) {
      best.fmode = cd.mode;
      best.fty = cd.ty;
      best.ftz = cd.tz;
      best.fms = per;
    }
  }
  // refresh sweeps (a refresh every sweep) with that tiling: drawn inside the
  // sweep (0), drawn ahead (1), drawn ahead with D2 pairs (2)
  if (!rc) {
    // long enough for the draws to run ahead of the sweeps (several draws of
    // up to 64 refreshes); the modes take turns, twice, and each keeps its
    // faster run
    const int32_t rsweeps = sweeps > 256 ? sweeps : 256;
    const int32_t nref = rsweeps + 8;
    std::vector<msa_lut> lut(2);
    // the demon planes' bit probabilities near T = 1.4, mid-anneal
    timing_lut(&lut[0], 0.06);
    timing_lut(&lut[1], 0.004);
    std::vector<msa_u32x4> ctr((size_t)nref * c->g.R);
    for (size_t q = 0; q < ctr.size(); ++q) ctr[q] = msa_u32x4{(uint32_t)q, 7u, 0u, 11u};
    std::vector<uint32_t> key(c->g.R, 0x5EEDu);
    std::vector<int32_t> idx(nref, 0);
    rc = msa_set_luts(t, 1, lut.data());
    if (!rc) rc = msa_set_refresh(t, nref, ctr.data(), key.data(), idx.data());
    // variants: mode 0; mode 1 with the default and with the fastest
    // one-pass tile for its sweeps; mode 2 (whose sweeps run in pairs)
    float per_v[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    const bool two_ff = t->pair && best.fmode > 0;
    for (int round = 0; round < 2 && !rc; ++round)
      for (int v = 0; v < 4 && !rc; ++v) {
        if ((v == 2 && !two_ff) || (v == 3 && t->qty == 0)) continue;
        TuneEntry e = best;
        e.ff = v == 2;
        if ((rc = msa_set_tiling(t, best.mode, best.ty, best.tz))) break;
        apply_tuned(t, e);
        float per = 0;
        if ((rc = time_refresh(t, rsweeps, v == 0 ? 0 : (v == 3 ? 2 : 1), &per))) break;
        if (per < per_v[v]) per_v[v] = per;
      }
    int bv = 0;
    for (int v = 1; v < 4; ++v)
      if (per_v[v] < per_v[bv]) bv = v;
    best.rmode = bv == 0 ? 0 : (bv == 3 ? 2 : 1);
    best.ff = bv == 2;
  }
  msa_destroy(t);
  if (rc) return rc;
  g_tuned.push_back(best);
  if ((rc = msa_set_tiling(c, best.mode, best.ty, best.tz))) return rc;
  apply_tuned(c, best);
  if (mode) *mode = best.mode;
  if (ty) *ty = best.ty;
  if (tz) *tz = best.tz;
  if (ms_per_sweep) *ms_per_sweep = best.ms;
  return MSA_OK;
}

extern "C" int msa_sync(msa_ctx *c) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSA_OK;
}

extern "C" void *msa_host_alloc(int64_t bytes) {
  if (bytes <= 0) return nullptr;
  void *p = nullptr;
  if (hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
    msa_fail(MSA_ERR_NOMEM, "page-locked host allocation of %lld bytes", (long long)bytes);
    return nullptr;
  }
  return p;
}

extern "C" void msa_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

extern "C" int msa_query(msa_ctx *c, int32_t *done) {
  if (!c || !done) return msa_fail(MSA_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(c->device));
  const hipError_t e = hipStreamQuery(c->stream);
  if (e == hipErrorNotReady) {
    *done = 0;
    return MSA_OK;
  }
  HIPCHK(e);
  *done = 1;
  return MSA_OK;
}

extern "C" const char *msa_strerror(int status) {
  if (status == MSA_OK) return "ok";
  return g_err;
}
