// msa_device.hpp -- device building blocks of the micSA path for gfx950.
//
// Integer/bitwise only: no MFMA.  One uint64 word = one site of 64 samples.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msa {

// ---------------------------------------------------------------- layout --
// Colour-split ("sublattice-ordered") storage: the sites of colour c = (x+y+z)&1
// are packed as h = i + Lh*(y + L*z), i = x>>1, Lh = L/2, so the sites one
// half-sweep writes, and the sites it reads, are each contiguous planes.
//   spins  sp[(r*2 + c)*H + h]
//   J      jp[(c*3 + d)*H + h]            (bond site -> site + e_d)
//   demons dm[((r*nc + l)*2 + c)*H + h]   as ulonglong2 {D0, D1}
struct Geo {
  uint32_t L, Lh, R, nc;
  uint64_t H, V;
};

__device__ __forceinline__ uint32_t wrapi(int a, uint32_t n) {
  return (uint32_t)(a < 0 ? a + (int)n : (a >= (int)n ? a - (int)n : a));
}

// Lattice direction e_l for demon layer l (main.cu:1125-1175): +x,+y,+z,-x,-y,-z.
__device__ __forceinline__ void layer_dir(int l, int &dx, int &dy, int &dz) {
  const int sgn = l < 3 ? 1 : -1;
  const int a = l % 3;
  dx = a == 0 ? sgn : 0;
  dy = a == 1 ? sgn : 0;
  dz = a == 2 ? sgn : 0;
}

// Split index of the site a demon of colour `col` at (i, y, z) moves to in the
// walk of its layer (d_Walk, main.cu:1084-1193): colour 0 by +2 e_l, colour 1
// by -2 e_l.  A shift by 2 keeps the colour, so only the index changes.
__device__ __forceinline__ uint64_t walk_dest(const Geo &g, int col, int dx, int dy, int dz,
                                              uint32_t i, uint32_t y, uint32_t z) {
  const int sg = col ? -1 : 1;
  const uint32_t wi = wrapi((int)i + sg * dx, g.Lh);
  const uint32_t wy = wrapi((int)y + 2 * sg * dy, g.L);
  const uint32_t wz = wrapi((int)z + 2 * sg * dz, g.L);
  return wi + (uint64_t)g.Lh * (wy + (uint64_t)g.L * wz);
}

// Access at a 32-bit byte offset from a wave-uniform base: the base stays in
// scalar registers and each access costs one offset VGPR (the saddr form of
// global_load / global_store).  Callers guarantee offset < 2^32.
template <class T>
__device__ __forceinline__ T ldo(const T *base, uint32_t boff) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
template <class T>
__device__ __forceinline__ void sto(T *base, uint32_t boff, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff) = v;
}

// Streaming forms for data a sweep touches once: non-temporal hint, so the
// lines do not displace reused data (spins and J shared with neighbouring
// tiles) from L2 and the last-level cache.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ ulonglong2 ldo_stream(const ulonglong2 *base, uint32_t boff) {
  const u64x2 v = __builtin_nontemporal_load(
      reinterpret_cast<const u64x2 *>(reinterpret_cast<const char *>(base) + boff));
  return make_ulonglong2(v.x, v.y);
}
__device__ __forceinline__ void sto_stream(uint64_t *base, uint32_t boff, uint64_t v) {
  __builtin_nontemporal_store(v, reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(base) + boff));
}
__device__ __forceinline__ void sto_stream(ulonglong2 *base, uint32_t boff, uint64_t x, uint64_t y) {
  u64x2 v;
  v.x = x;
  v.y = y;
  __builtin_nontemporal_store(v, reinterpret_cast<u64x2 *>(reinterpret_cast<char *>(base) + boff));
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt at their maxima, gfx9 encoding):
// every vector memory access this wave issued has completed.  Placed where
// the loads are needed anyway, it lets the compiler's wait counting treat
// them as done, so that loads issued after it for a later step are the only
// accesses in flight and later waits need not cover them.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---------------------------------------------------------------- Philox --
// Philox4x32-10 with the round and key schedule of main.cu:562-595
// (Random123's philox4x32 with 10 rounds).
__device__ __forceinline__ void philox10(uint32_t &v0, uint32_t &v1, uint32_t &v2, uint32_t &v3,
                                         uint32_t k0, uint32_t k1) {
  // 64-bit products: one v_mad_u64_u32 gives both halves (mulhi and mullo of
  // the reference's _update_state); the two xors fold into one v_bitop3.
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * v0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * v2;
    const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), v1, k0, 0x96);
    const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), v3, k1, 0x96);
    v0 = n0; v1 = (uint32_t)p1; v2 = n2; v3 = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// One heat-bath table as the device searches it: the 256 cumulative
// thresholds of s_lut_heat_bath (header.h:113-116) and a presearch over the
// top kPreBits bits of R: pre[b] = the byte whose interval holds the first R
// of bucket b, pre[2^kPreBits] = 255.  The byte of any R in bucket b lies in
// [pre[b], pre[b+1]] because the search answer is non-decreasing in R.  The
// reference's own presearch has 2^10 buckets of (min, max) pairs
// (NUMBITSPREBUSQUEDAS = 10); this one is built on the device from the same
// thresholds (k_fine_luts), finer so that almost every search ends in its
// bucket (a bracket of width 0, no threshold read, for all but ~1 % of R), and
// half the size per bucket so two tables (2 x 34 KiB) stage into LDS.
// The byte found is the same for every R: the unique n with
// umbral[n-1] <= R < umbral[n].
constexpr int kPreBits = 15;
struct FineLut {
  uint64_t umb[256];                       // umbral[n] = msb32 << 32 | lsb32
  uint8_t pre[(1u << kPreBits) + 16];      // + the sentinel 255, padded
};
static_assert(sizeof(FineLut) % 16 == 0, "FineLut is copied in 16-byte units");

struct LutView {
  const uint64_t *umb;
  const uint8_t *pre;
};
__device__ __forceinline__ LutView lut_view(const FineLut *t) { return LutView{t->umb, t->pre}; }

// Byte n with umbral[n-1] <= R < umbral[n], R = (rm:rl): the bisection of
// main.cu:615-630 started from the bracket [pre[b], pre[b+1]] of R's bucket
// b, which always contains the answer (the thresholds are non-decreasing,
// checked on upload).  Same byte for every R as the reference's search: the
// table construction and this search, restated on the host, agree with the
// reference bisection on every interval endpoint and bucket edge, and the
// device's refresh words equal the oracle's in every refresh parity test.
__device__ __forceinline__ uint32_t bisect(const LutView &t, uint32_t rm, uint32_t rl) {
  const uint64_t r = ((uint64_t)rm << 32) | rl;
  auto below = [&](uint32_t n) { return r < t.umb[n]; };
  const uint32_t b = rm >> (32 - kPreBits);
  uint32_t lo = t.pre[b], hi = t.pre[b + 1];
  if (lo == hi) return lo;                 // no threshold inside the bucket
  // at most 8 halvings of a bracket inside 0..255
  for (int it = 0; it < 9 && hi - lo > 1; ++it) {
    const uint32_t mid = (hi + lo) >> 1;
    if (below(mid)) hi = mid; else lo = mid;
  }
  return below(lo) ? lo : hi;
}

// Rotate left by r in 0..31, the last step of _get_rnd_word (main.cu:644;
// for r = 0 its right shift by 32 gives 0 on the reference's GPUs, so w is
// left as it is): one v_alignbit of (w:w) by 32 - r, whose shift count is
// taken mod 32, so r = 0 shifts by 0.
__device__ __forceinline__ uint32_t rotl32(uint32_t w, uint32_t r) {
  return __builtin_amdgcn_alignbit(w, w, 32u - r);
}

// _obten_aleatorio_contador + _get_rnd_word (main.cu:597-644): one Philox
// call -> four 64-bit uniforms -> four bytes of Bernoulli bits, rotated.
// The counter's x is incremented after the call (main.cu:606).
__device__ __forceinline__ uint32_t rnd_word(const LutView &t, uint32_t c[4], uint32_t k0,
                                             uint32_t k1) {
  uint32_t v0 = c[0], v1 = c[1], v2 = c[2], v3 = c[3];
  philox10(v0, v1, v2, v3, k0, k1);
  c[0] += 1u;
  const uint32_t a = v0 ^ v2, b = v1 ^ v3;
  const uint32_t l0 = ((v1 << 17) | (v0 >> 15)) ^ a ^ ((a << 21) | (b >> 11));
  const uint32_t l1 = ((v0 << 17) | (v1 >> 15)) ^ b ^ (b << 21);
  const uint32_t l2 = (a << 28) | (b >> 4);
  const uint32_t l3 = (b << 28) | (a >> 4);
  const uint32_t rot = (l0 + v1 + l2 + v3) >> 27;
  uint32_t w = bisect(t, v0, l0);
  w |= bisect(t, v1, l1) << 8;
  w |= bisect(t, v2, l2) << 16;
  w |= bisect(t, v3, l3) << 24;
  return rotl32(w, rot);
}

// The search of bisect() inside a bracket [lo, hi] known to hold the answer
// (hi > lo), on the 64-bit value R = (rm:rl) with one 8-byte threshold read
// per step.
__device__ __forceinline__ uint32_t bracket_search(const LutView &t, uint32_t rm, uint32_t rl,
                                                   uint32_t lo, uint32_t hi) {
  const uint64_t r = ((uint64_t)rm << 32) | rl;
  auto below = [&](uint32_t n) { return r < t.umb[n]; };
  for (int it = 0; it < 9 && hi - lo > 1; ++it) {
    const uint32_t mid = (hi + lo) >> 1;
    if (below(mid)) hi = mid; else lo = mid;
  }
  return below(lo) ? lo : hi;
}

// Two consecutive rnd_word() draws from one table (counter x, then x + 1),
// the same words rnd_word gives: both Philox calls and their mixing first,
// then the sixteen presearch reads issued together (one LDS wait for all of
// them rather than one per byte), then the bracket search of the few bytes
// whose bucket holds a threshold.
__device__ __forceinline__ void rnd_word_pair(const LutView &t, uint32_t cx, uint32_t cy,
                                              uint32_t cz, uint32_t cw, uint32_t k0, uint32_t k1,
                                              uint32_t &wa, uint32_t &wb) {
  uint32_t m[8], l[8], rot[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    uint32_t v0 = cx + (uint32_t)j, v1 = cy, v2 = cz, v3 = cw;
    philox10(v0, v1, v2, v3, k0, k1);
    const uint32_t a = v0 ^ v2, b = v1 ^ v3;
    const uint32_t l0 = ((v1 << 17) | (v0 >> 15)) ^ a ^ ((a << 21) | (b >> 11));
    const uint32_t l1 = ((v0 << 17) | (v1 >> 15)) ^ b ^ (b << 21);
    const uint32_t l2 = (a << 28) | (b >> 4);
    const uint32_t l3 = (b << 28) | (a >> 4);
    rot[j] = (l0 + v1 + l2 + v3) >> 27;
    m[4 * j] = v0; m[4 * j + 1] = v1; m[4 * j + 2] = v2; m[4 * j + 3] = v3;
    l[4 * j] = l0; l[4 * j + 1] = l1; l[4 * j + 2] = l2; l[4 * j + 3] = l3;
  }
  uint32_t lo[8], hi[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t bq = m[q] >> (32 - kPreBits);
    lo[q] = t.pre[bq];
    hi[q] = t.pre[bq + 1];
  }
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (lo[q] != hi[q]) lo[q] = bracket_search(t, m[q], l[q], lo[q], hi[q]);
  uint32_t w[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t x = lo[4 * j] | (lo[4 * j + 1] << 8) | (lo[4 * j + 2] << 16) | (lo[4 * j + 3] << 24);
    w[j] = rotl32(x, rot[j]);
  }
  wa = w[0];
  wb = w[1];
}

// Demon words of one (layer, site, replica) from the refresh stream:
// thread id tid = layer*V + site, counter ctr with .z += tid, key {tid, key},
// two words from the plane-0 table then two from the plane-1 table (rnd_word
// four times, the counter's x advancing after each).  refresh_plane(p) is
// plane p (D0 for p = 0 from the plane-0 table, D1 for p = 1 from the
// plane-1 table), so callers can place other work between the two.
__device__ __forceinline__ uint64_t refresh_plane(const LutView &t, uint4 ctr, uint32_t key,
                                                  uint32_t tid, int plane) {
  uint32_t wa, wb;
  rnd_word_pair(t, ctr.x + 2u * (uint32_t)plane, ctr.y, ctr.z + tid, ctr.w, tid, key, wa, wb);
  return ((uint64_t)wa << 32) | wb;
}

// ------------------------------------------------------------ the update --
// Bit-sliced local field (main.cu:898-944): the three + bonds give the 2-bit
// count (j1,j2), the three - bonds (j3,j4); f = s ^ J ^ s_nb = 1 marks an
// unsatisfied bond.
__device__ __forceinline__ void local_field(uint64_t s, uint64_t n0, uint64_t n1, uint64_t n2,
                                            uint64_t n3, uint64_t n4, uint64_t n5, uint64_t J0,
                                            uint64_t J1, uint64_t J2, uint64_t J3, uint64_t J4,
                                            uint64_t J5, uint64_t &j1, uint64_t &j2, uint64_t &j3,
                                            uint64_t &j4) {
  uint64_t f0 = s ^ J0 ^ n0, f1 = s ^ J1 ^ n1, f2 = s ^ J2 ^ n2;
  uint64_t k1 = f0 ^ f1;
  j1 = k1 ^ f2;
  j2 = (f0 & f1) ^ (k1 & f2);
  f0 = s ^ J3 ^ n3; f1 = s ^ J4 ^ n4; f2 = s ^ J5 ^ n5;
  k1 = f0 ^ f1;
  j3 = k1 ^ f2;
  j4 = (f0 & f1) ^ (k1 & f2);
}

// ISI demon algebra (main.cu:1037-1074): adds n_u - 3 to D = D0 + 2 D1 in
// bit-sliced two's complement; the carry decides acceptance.  Equivalent
// closed form: flip iff 0 <= D + n_u - 3 <= 3.
__device__ __forceinline__ uint64_t isi_update(uint64_t j1, uint64_t j2, uint64_t j3, uint64_t j4,
                                               uint64_t &D0, uint64_t &D1) {
  const uint64_t x13 = j1 ^ j3, x24 = j2 ^ j4;
  const uint64_t sub = (x24 & ~(j1 | j3)) | ~(j2 | j4);
  const uint64_t s0 = ~x13;
  uint64_t s1 = ~x24 & (x13 | (s0 & ~(j1 ^ j2)));
  s1 = (sub & (s1 ^ s0)) | (~sub & s1);
  const uint64_t n0 = D0 ^ s0;
  const uint64_t k1 = D0 & s0;
  const uint64_t k2 = D1 ^ s1;
  const uint64_t n1 = k1 ^ k2;
  const uint64_t carry = ((D1 & s1) | (k2 & k1)) ^ sub;
  const uint64_t flip = ~carry;
  D0 = (flip & n0) | (carry & D0);
  D1 = (flip & n1) | (carry & D1);
  return flip;
}

// ------------------------------------------------ gfx950 3-input logic --
// v_bitop3_b32 evaluates any 3-input boolean function given as an 8-entry
// truth table.  Only tables symmetric in their three operands are used here
// (xor3 0x96, xnor3 0x69, majority 0xE8, and3 0x80, "one or two set" 0x7E),
// so the result does not depend on the operand order of the table index.
template <uint32_t TT>
__device__ __forceinline__ uint64_t lop3(uint64_t a, uint64_t b, uint64_t c) {
  const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, (uint32_t)c, TT);
  const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32),
                                                  (uint32_t)(c >> 32), TT);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t xor3(uint64_t a, uint64_t b, uint64_t c) { return lop3<0x96>(a, b, c); }
__device__ __forceinline__ uint64_t xnor3(uint64_t a, uint64_t b, uint64_t c) { return lop3<0x69>(a, b, c); }
__device__ __forceinline__ uint64_t maj3(uint64_t a, uint64_t b, uint64_t c) { return lop3<0xE8>(a, b, c); }
__device__ __forceinline__ uint64_t and3(uint64_t a, uint64_t b, uint64_t c) { return lop3<0x80>(a, b, c); }
__device__ __forceinline__ uint64_t one_or_two3(uint64_t a, uint64_t b, uint64_t c) { return lop3<0x7E>(a, b, c); }

// One site update, 64 samples at once: the local field and the demon
// acceptance of main.cu:898-944 and main.cu:1037-1074, written as a smaller
// circuit with the same truth table (checked exhaustively against the
// reference algebra by the oracle tests).  With n+ = j1 + 2 j2 (+x,+y,+z
// bonds unsatisfied), n- = j3 + 2 j4 and D = D0 + 2 D1, a flip is accepted iff
// 3 <= D + n+ + n- <= 6 and then D' = D + n - 3:
//   c1 = [D0 + j1 + j3 in {1,2}], c2 = D0 & j1 & j3,
//   v0 = D1 ^ j2 ^ j4,           v1 = maj(D1, j2, j4),
//   flip = v1 ^ c2 ^ (c1 & v0),  D0' = ~(D0 ^ j1 ^ j3), D1' = v0 ^ c1.
// n0..n5 are the neighbour spins at +x,+y,+z,-x,-y,-z, J0..J5 the bonds.
__device__ __forceinline__ uint64_t site_update(uint64_t s, uint64_t n0, uint64_t n1, uint64_t n2,
                                                uint64_t n3, uint64_t n4, uint64_t n5,
                                                uint64_t J0, uint64_t J1, uint64_t J2,
                                                uint64_t J3, uint64_t J4, uint64_t J5,
                                                uint64_t &D0, uint64_t &D1) {
  const uint64_t f0 = xor3(s, J0, n0), f1 = xor3(s, J1, n1), f2 = xor3(s, J2, n2);
  const uint64_t j1 = xor3(f0, f1, f2), j2 = maj3(f0, f1, f2);
  const uint64_t f3 = xor3(s, J3, n3), f4 = xor3(s, J4, n4), f5 = xor3(s, J5, n5);
  const uint64_t j3 = xor3(f3, f4, f5), j4 = maj3(f3, f4, f5);
  const uint64_t c1 = one_or_two3(D0, j1, j3), c2 = and3(D0, j1, j3);
  const uint64_t v0 = xor3(D1, j2, j4), v1 = maj3(D1, j2, j4);
  const uint64_t flip = xor3(v1, c2, c1 & v0);
  const uint64_t d0n = xnor3(D0, j1, j3), d1n = v0 ^ c1;
  D0 = (flip & d0n) | (~flip & D0);
  D1 = (flip & d1n) | (~flip & D1);
  return s ^ flip;
}

// The Metropolis branch of the canonical thermalization (main.cu:945-958):
// (b0, b1, b2) are three Bernoulli(p) words, R = their leading ones
// (P(R >= k) = p^k), and a spin flips iff n_u + R >= 3, i.e. with
// probability min(1, exp(-beta dE)).  The flip algebra is the reference's,
// on the local field of site_update.
__device__ __forceinline__ uint64_t metropolis_update(uint64_t s, uint64_t n0, uint64_t n1,
                                                      uint64_t n2, uint64_t n3, uint64_t n4,
                                                      uint64_t n5, uint64_t J0, uint64_t J1,
                                                      uint64_t J2, uint64_t J3, uint64_t J4,
                                                      uint64_t J5, uint64_t b0, uint64_t b1,
                                                      uint64_t b2) {
  const uint64_t f0 = xor3(s, J0, n0), f1 = xor3(s, J1, n1), f2 = xor3(s, J2, n2);
  const uint64_t j1 = xor3(f0, f1, f2), j2 = maj3(f0, f1, f2);
  const uint64_t f3 = xor3(s, J3, n3), f4 = xor3(s, J4, n4), f5 = xor3(s, J5, n5);
  const uint64_t j3 = xor3(f3, f4, f5), j4 = maj3(f3, f4, f5);
  b1 &= b0;
  b2 &= b1;
  const uint64_t id2 = b1, id1 = (b0 ^ b1) | b2, j24 = j2 | j4;
  const uint64_t flip = (j1 & j3 & id1) | ((j24 | id2) & (j1 | j3 | id1)) | ((j2 & j4) | (j24 & id2));
  return s ^ flip;
}

// --------------------------------------------------- per-lane counting --
// Bit-sliced (vertical) counter: word j holds bit j of 64 independent lane
// counts.  add() is a ripple-carry add of a 1-bit-per-lane word at weight 2^pos.
// The value of lane t ^ s for s = 1, 2, 4, 8 (DPP moves: quad permutes,
// and row mirror / half-mirror, t ^ 15 and t ^ 7, composed with them),
// 16 (ds_swizzle, xor mask 16 within 32 lanes) and 32 (ds_bpermute).
__device__ __forceinline__ uint32_t lane_xor32(uint32_t v, int s) {
  constexpr int kXor1 = 0xB1, kXor2 = 0x4E, kXor3 = 0x1B, kMirror = 0x140, kHalfMirror = 0x141;
  switch (s) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, kXor1, 0xF, 0xF, false);
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, kXor2, 0xF, 0xF, false);
    case 4:
      return (uint32_t)__builtin_amdgcn_mov_dpp(
          __builtin_amdgcn_mov_dpp((int)v, kHalfMirror, 0xF, 0xF, false), kXor3, 0xF, 0xF, false);
    case 8:
      return (uint32_t)__builtin_amdgcn_mov_dpp(
          __builtin_amdgcn_mov_dpp((int)v, kMirror, 0xF, 0xF, false), kHalfMirror, 0xF, 0xF, false);
    case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
    default: return __shfl_xor(v, 32);
  }
}
__device__ __forceinline__ uint64_t lane_xor(uint64_t v, int s) {
  return ((uint64_t)lane_xor32((uint32_t)(v >> 32), s) << 32) | lane_xor32((uint32_t)v, s);
}
// Transpose the 64 x 64 bit matrix held one row per lane (bit b of lane t
// becomes bit t of lane b): six stages, each swapping the off-diagonal
// s x s blocks of every 2s x 2s block between lanes t and t ^ s.
__device__ __forceinline__ uint64_t wave_transpose(uint64_t x) {
  const uint32_t lane = __lane_id();
  constexpr uint64_t keep[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int s = 32 >> k;
    const uint64_t m = keep[k];
    const uint64_t y = lane_xor(x, s);
    x = (lane & (uint32_t)s) ? ((x & ~m) | ((y >> s) & m)) : ((x & m) | ((y & m) << s));
  }
  return x;
}

// Per bit lane, the number of set bits among six words as b0 + 2 b1 + 4 b2
// (two full adders, a half adder and a full adder over their carries).
__device__ __forceinline__ void sum6(const uint64_t w[6], uint64_t &b0, uint64_t &b1, uint64_t &b2) {
  const uint64_t s1 = w[0] ^ w[1] ^ w[2], c1 = (w[0] & w[1]) | (w[2] & (w[0] ^ w[1]));
  const uint64_t s2 = w[3] ^ w[4] ^ w[5], c2 = (w[3] & w[4]) | (w[5] & (w[3] ^ w[4]));
  const uint64_t c3 = s1 & s2;
  b0 = s1 ^ s2;
  b1 = c1 ^ c2 ^ c3;
  b2 = (c1 & c2) | (c3 & (c1 ^ c2));
}

template <int NB>
struct VCount {
  uint64_t c[NB];
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int j = 0; j < NB; ++j) c[j] = 0;
  }
  __device__ __forceinline__ void add(uint64_t w, int pos = 0) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j >= pos) {
        const uint64_t t = c[j] & w;
        c[j] ^= w;
        w = t;
      }
    }
  }
  // Add the bit-sliced number b0 + 2 b1 + 4 b2 (a ripple of full adders).
  __device__ __forceinline__ void add3(uint64_t b0, uint64_t b1, uint64_t b2) {
    static_assert(NB >= 3, "add3 needs three bit-planes");
    uint64_t carry = c[0] & b0;
    c[0] ^= b0;
    uint64_t s = c[1] ^ b1 ^ carry;
    carry = (c[1] & b1) | (carry & (c[1] ^ b1));
    c[1] = s;
    s = c[2] ^ b2 ^ carry;
    carry = (c[2] & b2) | (carry & (c[2] ^ b2));
    c[2] = s;
#pragma unroll
    for (int j = 3; j < NB; ++j) {
      const uint64_t t = c[j] & carry;
      c[j] ^= carry;
      carry = t;
    }
  }
  // Sum the counters of the 64 threads of the wavefront; lane b returns the
  // total for sample b.  Every lane of the wavefront must be active.
  __device__ __forceinline__ uint32_t wave_total() const {
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (__ballot(c[j] != 0) == 0) continue;
      acc += (uint32_t)__popcll(wave_transpose(c[j])) << j;
    }
    return acc;
  }
};

}  // namespace msa
