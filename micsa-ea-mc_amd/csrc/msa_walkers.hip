// msa_walkers.hip -- the reference's demon initialisation stream on the
// device: InitWalkers (PROGRAM/src/ini.c:712-781) reproduced bit for bit.
//
// The reference refreshes the demons of each (sample, replica) from one
// sequential stream of two host generators (PR+LCG and xoshiro256++), 64
// draws per multispin word, MSC_V * nc words per refresh, on host threads.
// Here one refresh is two kernels:
//   k_walk_jump  the generator state at the start of every chunk of W words
//                of every stream, by jump-ahead from the stream's state at
//                the start of the refresh: the lagged sum
//                x_n = x_{n-24} + x_{n-55} (mod 2^64) is linear over Z/2^64,
//                so x_{N-55+m} = sum_i (z^m mod z^55 - z^31 - 1)_i x_{N-55+i};
//                the LCG jumps as an affine map; xoshiro256++'s state update
//                is linear over GF(2) and jumps by matrix powers;
//   k_walk_gen   one wavefront per (chunk, replica), lane = sample lane: each
//                lane runs its stream over the chunk (the last 64 lagged-sum
//                values in registers), and the 64 lanes' words are transposed
//                across the wavefront (DPP and swizzle stages) into the lane
//                layout (one word per site) and
//                stored at the sites of the reference's multispin word
//                (packing_u placement, ini.c:301-345).
// Lanes without a stream get zero demons.
#include <array>
#include <vector>

#include "msa_ctx.hpp"

namespace {

constexpr uint64_t kLcgA = 3202034522624059733ULL;   // random_generator.h:77
constexpr int kLagN = 55;                              // x_n = x_{n-24} + x_{n-55}
constexpr int kRing = 61;                              // PR = x_n ^ x_{n-61}

// ---------------------------------------------------------- host maths --

using Poly = std::array<uint64_t, kLagN>;   // coefficients of z^0 .. z^54

// a * b mod (z^55 - z^31 - 1) over Z/2^64; z^d = z^(d-24) + z^(d-55)
Poly poly_mul(const Poly &a, const Poly &b) {
  uint64_t t[2 * kLagN - 1] = {0};
  for (int i = 0; i < kLagN; ++i)
    if (a[i])
      for (int j = 0; j < kLagN; ++j) t[i + j] += a[i] * b[j];
  for (int d = 2 * kLagN - 2; d >= kLagN; --d) {
    t[d - 24] += t[d];
    t[d - 55] += t[d];
  }
  Poly r;
  for (int i = 0; i < kLagN; ++i) r[i] = t[i];
  return r;
}

Poly poly_zpow(uint64_t m) {
  Poly r{}, b{};
  r[0] = 1;
  b[1] = 1;
  while (m) {
    if (m & 1) r = poly_mul(r, b);
    b = poly_mul(b, b);
    m >>= 1;
  }
  return r;
}

// xoshiro256++ state update (random_generator.h:50-59) without the output
void xo_step(uint64_t s[4]) {
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = (s[3] << 45) | (s[3] >> 19);
}

// 256 x 256 matrix over GF(2) by columns: col[k] = image of state bit k
// (bit k of word k / 64)
struct XoMat {
  uint64_t col[256][4];
};

void mat_vec(const XoMat &m, const uint64_t s[4], uint64_t out[4]) {
  uint64_t o[4] = {0, 0, 0, 0};
  for (int k = 0; k < 256; ++k)
    if ((s[k >> 6] >> (k & 63)) & 1)
      for (int w = 0; w < 4; ++w) o[w] ^= m.col[k][w];
  for (int w = 0; w < 4; ++w) out[w] = o[w];
}

XoMat mat_mul(const XoMat &a, const XoMat &b) {   // a * b
  XoMat c;
  for (int k = 0; k < 256; ++k) mat_vec(a, b.col[k], c.col[k]);
  return c;
}

XoMat xo_step_matrix() {
  XoMat m;
  for (int k = 0; k < 256; ++k) {
    uint64_t s[4] = {0, 0, 0, 0};
    s[k >> 6] = 1ull << (k & 63);
    xo_step(s);
    for (int w = 0; w < 4; ++w) m.col[k][w] = s[w];
  }
  return m;
}

XoMat mat_pow(XoMat b, uint64_t m) {
  XoMat r;
  for (int k = 0; k < 256; ++k)
    for (int w = 0; w < 4; ++w) r.col[k][w] = (w == (k >> 6)) ? (1ull << (k & 63)) : 0;
  while (m) {
    if (m & 1) r = mat_mul(r, b);
    b = mat_mul(b, b);
    m >>= 1;
  }
  return r;
}

// --------------------------------------------------------------- kernels --

struct ChunkState {      // generator state at the start of a chunk, one lane per field
  uint64_t ring[kRing];  // x_{q0-61} .. x_{q0-1}
  uint64_t lcg;
  uint64_t xo[4];
};

struct JumpArgs {
  const msa_walker_stream *st;   // [nstream * R], state at the start of the refresh
  ChunkState *cs;                // [nchunk][nstream * R]
  const uint64_t *lf;            // [nchunk][55]: z^(j Lc - 6) mod P, row 0 unused
  const uint64_t *lcg;           // [nchunk][2]: a^(j Lc), (a^(j Lc) - 1) / (a - 1)
  const uint64_t *xo;            // [nbit][256][4]: columns of T^(Lc 2^b)
  uint32_t ns;                   // streams x replicas
};

// Block (chunk j, group of 256 streams); every thread one stream.  The
// coefficient vector is stepped (times z) by the whole block in LDS.
__global__ void __launch_bounds__(256) k_walk_jump(JumpArgs a) {
  __shared__ uint64_t cf[2][kLagN];
  const uint32_t j = blockIdx.x, sidx = blockIdx.y * blockDim.x + threadIdx.x;
  const bool live = sidx < a.ns;
  ChunkState *out = a.cs + (uint64_t)j * a.ns + sidx;
  const msa_walker_stream *in = a.st + (live ? sidx : 0);
  if (j == 0) {
    if (live) {
      for (int u = 0; u < kRing; ++u) out->ring[u] = in->ring[u];
      out->lcg = in->lcg;
      for (int w = 0; w < 4; ++w) out->xo[w] = in->xo[w];
    }
    return;
  }
  uint64_t b[kLagN];
#pragma unroll
  for (int i = 0; i < kLagN; ++i) b[i] = live ? in->ring[6 + i] : 0;   // x_{N-55+i}
  if (threadIdx.x < kLagN) cf[0][threadIdx.x] = a.lf[(uint64_t)j * kLagN + threadIdx.x];
  __syncthreads();
  for (int u = 0; u < kRing; ++u) {
    const uint64_t *c = cf[u & 1];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kLagN; ++i) acc += c[i] * b[i];
    if (live) out->ring[u] = acc;
    if (threadIdx.x < kLagN) {
      const int i = threadIdx.x;
      uint64_t v = i == 0 ? c[kLagN - 1] : c[i - 1];
      if (i == 31) v += c[kLagN - 1];
      cf[(u + 1) & 1][i] = v;
    }
    __syncthreads();
  }
  if (!live) return;
  out->lcg = a.lcg[2 * j] * in->lcg + a.lcg[2 * j + 1];
  uint64_t s[4] = {in->xo[0], in->xo[1], in->xo[2], in->xo[3]};
  for (uint32_t bit = 0; (j >> bit) != 0; ++bit) {
    if (!((j >> bit) & 1)) continue;
    const uint64_t *m = a.xo + (uint64_t)bit * 256 * 4;
    uint64_t o[4] = {0, 0, 0, 0};
    for (int k = 0; k < 256; ++k)
      if ((s[k >> 6] >> (k & 63)) & 1) {
        o[0] ^= m[4 * k];
        o[1] ^= m[4 * k + 1];
        o[2] ^= m[4 * k + 2];
        o[3] ^= m[4 * k + 3];
      }
    s[0] = o[0];
    s[1] = o[1];
    s[2] = o[2];
    s[3] = o[3];
  }
  for (int w = 0; w < 4; ++w) out->xo[w] = s[w];
}

struct GenArgs {
  const ChunkState *cs;          // [nchunk][nstream * R]
  msa_walker_stream *st_next;    // [nstream * R]: state after the refresh
  const int32_t *slot;           // [64]: stream of each lane, -1 none
  ulonglong2 *dm;                // demons of the current buffer
  msa::Geo g;
  uint32_t W, nchunk, ns, ML;
  uint64_t thr0, thr1, thr2;
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// One wavefront per (chunk j, replica r); lane t runs the stream of sample
// lane t.  Word iword of a replica's segment is colour iword / (HALF nc),
// layer (iword / HALF) % nc, half index iword % HALF (ini.c:750-775 with the
// uploads of refresh_walkers, main.cu:355-362).
__global__ void __launch_bounds__(64) k_walk_gen(GenArgs a) {
  const uint32_t t = threadIdx.x, j = blockIdx.x, r = blockIdx.y;
  const msa::Geo g = a.g;
  const int slot = a.slot[t];
  const bool live = slot >= 0;
  const uint32_t sidx = live ? (uint32_t)slot * g.R + r : 0;
  const ChunkState &c = a.cs[(uint64_t)j * a.ns + sidx];
  uint64_t x[64];
#pragma unroll
  for (int u = 0; u < 64; ++u) x[u] = 0;
  uint64_t lcg = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  if (live) {
#pragma unroll
    for (int u = 0; u < kRing; ++u) x[(u + 3) & 63] = c.ring[u];   // x_{q} at q & 63, q = -61..-1
    lcg = c.lcg;
    s0 = c.xo[0];
    s1 = c.xo[1];
    s2 = c.xo[2];
    s3 = c.xo[3];
  }
  const uint32_t ML = a.ML, MS = ML * ML;
  const uint64_t HALF = (uint64_t)MS * ML / 2;
  // this lane's site within every word: bit t = bx + 4 by + 16 bz
  const uint32_t bx = t & 3, by = (t >> 2) & 3, bz = t >> 4;
  for (uint32_t w = 0; w < a.W; ++w) {
    uint64_t d0 = 0, d1 = 0;
    if (live) {
#pragma unroll
      for (int bit = 0; bit < 64; ++bit) {
        const uint64_t xo_out = rotl64(s0 + s3, 23) + s0;
        const uint64_t tt = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= tt;
        s3 = rotl64(s3, 45);
        const uint64_t xn = x[(bit - 24) & 63] + x[(bit - 55) & 63];
        const uint64_t pr = xn ^ x[(bit - 61) & 63];
        x[bit] = xn;
        lcg = lcg * kLcgA + 1;
        const uint64_t e = lcg + pr + xo_out;
        const uint64_t dv = (uint64_t)(e < a.thr0) + (uint64_t)(e < a.thr1) + (uint64_t)(e < a.thr2);
        d0 |= (dv & 1) << bit;
        d1 |= (dv >> 1) << bit;
      }
    }
    // word for site t of the 64: bit s = lane s's bit t (all 64 lanes active)
    const uint64_t m0 = msa::wave_transpose(d0), m1 = msa::wave_transpose(d1);
    const uint64_t iw = (uint64_t)j * a.W + w;
    const uint32_t col = (uint32_t)(iw / (HALF * g.nc));
    const uint32_t layer = (uint32_t)((iw / HALF) % g.nc);
    const uint64_t half = iw % HALF;
    const uint32_t gz = (uint32_t)((2 * half) / MS), e_ = (uint32_t)(((2 * half) / ML) % ML);
    const uint32_t d = (col + e_ + gz) & 1u;
    const uint32_t aa = (uint32_t)((2 * half + d) % ML);
    const uint32_t xs = aa + ML * bx, ys = e_ + ML * by, zs = gz + ML * bz;
    const uint64_t h = (xs >> 1) + (uint64_t)g.Lh * (ys + (uint64_t)g.L * zs);
    a.dm[((uint64_t)(r * g.nc + layer) * 2 + col) * g.H + h] = make_ulonglong2(m0, m1);
  }
  if (live && j + 1 == a.nchunk) {
    msa_walker_stream &o = a.st_next[sidx];
#pragma unroll
    for (int u = 0; u < kRing; ++u) o.ring[u] = x[(u + 3) & 63];
    o.lcg = lcg;
    o.xo[0] = s0;
    o.xo[1] = s1;
    o.xo[2] = s2;
    o.xo[3] = s3;
  }
}

}  // namespace

// ------------------------------------------------------------------ host --

void walkers_free(msa_ctx *c) {
  (void)hipFree(c->w_st[0]);
  (void)hipFree(c->w_st[1]);
  (void)hipFree(c->w_cs);
  (void)hipFree(c->w_slot);
  (void)hipFree(c->w_lf);
  (void)hipFree(c->w_lcg);
  (void)hipFree(c->w_xo);
  c->w_st[0] = c->w_st[1] = nullptr;
  c->w_cs = nullptr;
  c->w_slot = nullptr;
  c->w_lf = c->w_lcg = c->w_xo = nullptr;
  c->w_n = 0;
  c->w_thr.clear();
}

extern "C" int msa_set_walker_streams(msa_ctx *c, int32_t n, const int32_t *lanes,
                                      const msa_walker_stream *st, int32_t nref,
                                      const uint64_t *thr) {
  if (!c) return msa_fail(MSA_ERR_ARG, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  walkers_free(c);
  if (n == 0) return MSA_OK;
  const msa::Geo &g = c->g;
  if (n < 0 || n > 64 || !lanes || !st || nref < 1 || !thr)
    return msa_fail(MSA_ERR_ARG, "walker streams: n=%d nref=%d", n, nref);
  if (g.L % 8) return msa_fail(MSA_ERR_ARG, "walker streams need L a multiple of 8 (L=%u)", g.L);
  int32_t slot[64];
  for (int b = 0; b < 64; ++b) slot[b] = -1;
  for (int i = 0; i < n; ++i) {
    if (lanes[i] < 0 || lanes[i] > 63 || slot[lanes[i]] >= 0)
      return msa_fail(MSA_ERR_ARG, "walker lane %d invalid or repeated", lanes[i]);
    slot[lanes[i]] = i;
  }
  // chunks of W words: the largest divisor of the words per segment up to 64
  const uint64_t words = g.V / 64 * g.nc;
  uint32_t W = 64;
  while (words % W) --W;
  const uint32_t nchunk = (uint32_t)(words / W);
  const uint64_t Lc = 64ull * W;
  uint32_t nbit = 0;
  while (nbit < 32 && (nchunk - 1) >> nbit) ++nbit;
  // tables
  std::vector<uint64_t> lf((size_t)nchunk * kLagN, 0), lcg((size_t)nchunk * 2, 0);
  if (nchunk > 1) {
    const Poly step = poly_zpow(Lc);
    Poly cur = poly_zpow(Lc - 6);
    for (uint32_t j = 1; j < nchunk; ++j) {
      for (int i = 0; i < kLagN; ++i) lf[(size_t)j * kLagN + i] = cur[i];
      cur = poly_mul(cur, step);
    }
  }
  {
    uint64_t A = 1, C = 0, Astep = 1, Cstep = 0;   // z -> A z + C
    for (uint64_t m = 0; m < Lc; ++m) {
      Astep = Astep * kLcgA;
      Cstep = Cstep * kLcgA + 1;
    }
    for (uint32_t j = 0; j < nchunk; ++j) {
      lcg[2 * j] = A;
      lcg[2 * j + 1] = C;
      const uint64_t An = Astep * A, Cn = Astep * C + Cstep;
      A = An;
      C = Cn;
    }
  }
  std::vector<uint64_t> xo((size_t)(nbit ? nbit : 1) * 256 * 4, 0);
  {
    XoMat m = mat_pow(xo_step_matrix(), Lc);
    for (uint32_t b = 0; b < nbit; ++b) {
      for (int k = 0; k < 256; ++k)
        for (int w = 0; w < 4; ++w) xo[((size_t)b * 256 + k) * 4 + w] = m.col[k][w];
      m = mat_mul(m, m);
    }
  }
  const uint32_t ns = (uint32_t)n * g.R;
  if (hipMalloc(&c->w_st[0], sizeof(msa_walker_stream) * ns) != hipSuccess ||
      hipMalloc(&c->w_st[1], sizeof(msa_walker_stream) * ns) != hipSuccess ||
      hipMalloc(&c->w_cs, sizeof(ChunkState) * ns * (size_t)nchunk) != hipSuccess ||
      hipMalloc(&c->w_slot, sizeof(int32_t) * 64) != hipSuccess ||
      hipMalloc(&c->w_lf, sizeof(uint64_t) * lf.size()) != hipSuccess ||
      hipMalloc(&c->w_lcg, sizeof(uint64_t) * lcg.size()) != hipSuccess ||
      hipMalloc(&c->w_xo, sizeof(uint64_t) * xo.size()) != hipSuccess) {
    walkers_free(c);
    return msa_fail(MSA_ERR_NOMEM, "walker stream buffers");
  }
  HIPCHK(hipMemcpy(c->w_st[0], st, sizeof(msa_walker_stream) * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->w_slot, slot, sizeof(slot), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->w_lf, lf.data(), sizeof(uint64_t) * lf.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->w_lcg, lcg.data(), sizeof(uint64_t) * lcg.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->w_xo, xo.data(), sizeof(uint64_t) * xo.size(), hipMemcpyHostToDevice));
  c->w_n = n;
  c->w_cur = 0;
  c->w_W = W;
  c->w_nchunk = nchunk;
  c->w_thr.assign(thr, thr + 3 * (size_t)nref);
  return MSA_OK;
}

extern "C" int msa_get_walker_streams(msa_ctx *c, msa_walker_stream *st) {
  if (!c || !st) return msa_fail(MSA_ERR_ARG, "NULL argument");
  if (c->w_n == 0) return msa_fail(MSA_ERR_STATE, "no walker streams set");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(st, c->w_st[c->w_cur], sizeof(msa_walker_stream) * c->w_n * c->g.R,
                   hipMemcpyDeviceToHost));
  return MSA_OK;
}

// Refresh k from the walker streams: every demon of the current buffer.
int walkers_refresh(msa_ctx *c, int64_t k) {
  if (k < 0 || (size_t)k * 3 >= c->w_thr.size())
    return msa_fail(MSA_ERR_RANGE, "walker refresh %lld but %zu thresholds are set", (long long)k,
                    c->w_thr.size() / 3);
  const uint32_t ns = (uint32_t)c->w_n * c->g.R;
  JumpArgs ja{};
  ja.st = c->w_st[c->w_cur];
  ja.cs = reinterpret_cast<ChunkState *>(c->w_cs);
  ja.lf = c->w_lf;
  ja.lcg = c->w_lcg;
  ja.xo = c->w_xo;
  ja.ns = ns;
  const uint32_t jb = ns < 64 ? 64 : (ns > 256 ? 256 : (ns + 63) / 64 * 64);
  hipLaunchKernelGGL(k_walk_jump, dim3(c->w_nchunk, (ns + jb - 1) / jb), dim3(jb), 0, c->stream, ja);
  GenArgs ga{};
  ga.cs = reinterpret_cast<const ChunkState *>(c->w_cs);
  ga.st_next = c->w_st[c->w_cur ^ 1];
  ga.slot = c->w_slot;
  ga.dm = c->dm[c->cur];
  ga.g = c->g;
  ga.W = c->w_W;
  ga.nchunk = c->w_nchunk;
  ga.ns = ns;
  ga.ML = c->g.L / 4;
  ga.thr0 = c->w_thr[3 * k];
  ga.thr1 = c->w_thr[3 * k + 1];
  ga.thr2 = c->w_thr[3 * k + 2];
  hipLaunchKernelGGL(k_walk_gen, dim3(c->w_nchunk, c->g.R), dim3(64), 0, c->stream, ga);
  c->w_cur ^= 1;
  HIPCHK(hipGetLastError());
  return MSA_OK;
}
