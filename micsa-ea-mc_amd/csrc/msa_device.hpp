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

// ---------------------------------------------------------------- Philox --
// Philox4x32-10 with the round and key schedule of main.cu:562-595
// (Random123's philox4x32 with 10 rounds).
__device__ __forceinline__ void philox10(uint32_t &v0, uint32_t &v1, uint32_t &v2, uint32_t &v3,
                                         uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint32_t lo0 = 0xD2511F53u * v0, hi0 = __umulhi(0xD2511F53u, v0);
    const uint32_t lo1 = 0xCD9E8D57u * v2, hi1 = __umulhi(0xCD9E8D57u, v2);
    const uint32_t n0 = hi1 ^ v1 ^ k0;
    const uint32_t n2 = hi0 ^ v3 ^ k1;
    v0 = n0; v1 = lo1; v2 = n2; v3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// One heat-bath table (s_lut_heat_bath) viewed as the reference kernel views
// it: 256 x uint2 thresholds and 1024 x uint16 presearch entries.
struct LutView {
  const uint2 *umb;
  const uint16_t *pre;
};

// Byte n with umbral[n-1] <= R < umbral[n], R = (rm:rl); the bisection of
// main.cu:615-630 seeded by the 10-bit presearch.
__device__ __forceinline__ uint32_t bisect(const LutView &t, uint32_t rm, uint32_t rl) {
  auto below = [&](uint32_t n) {
    const uint2 u = t.umb[n];
    return (rm == u.x) ? (rl < u.y) : (rm < u.x);
  };
  uint32_t lo = 255, hi = 255;
  if (below(254)) {
    const uint32_t e = t.pre[rm >> 22];
    lo = e & 255u;
    hi = e >> 8;
  }
  // A valid table (min <= max, checked on upload) needs at most 8 halvings;
  // the bound keeps a malformed table from looping.
  for (int it = 0; it < 9 && hi - lo > 1; ++it) {
    const uint32_t mid = (hi + lo) >> 1;
    if (below(mid)) hi = mid; else lo = mid;
  }
  return below(lo) ? lo : hi;
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
  return rot ? ((w << rot) | (w >> (32u - rot))) : w;
}

// Demon words of one (layer, site, replica) from the refresh stream:
// thread id tid = layer*V + site, counter ctr with .z += tid, key {tid, key},
// two words from the plane-0 table then two from the plane-1 table.
__device__ __forceinline__ void refresh_demon(const LutView &t0, const LutView &t1, uint4 ctr,
                                              uint32_t key, uint32_t tid, uint64_t &d0,
                                              uint64_t &d1) {
  uint32_t c[4] = {ctr.x, ctr.y, ctr.z + tid, ctr.w};
  const uint64_t w0 = rnd_word(t0, c, tid, key);
  const uint64_t w1 = rnd_word(t0, c, tid, key);
  const uint64_t w2 = rnd_word(t1, c, tid, key);
  const uint64_t w3 = rnd_word(t1, c, tid, key);
  d0 = (w0 << 32) | w1;
  d1 = (w2 << 32) | w3;
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

// --------------------------------------------------- per-lane counting --
// Bit-sliced (vertical) counter: word j holds bit j of 64 independent lane
// counts.  add() is a ripple-carry add of a 1-bit-per-lane word at weight 2^pos.
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
  // Sum the counters of the 64 threads of the wavefront; lane b returns the
  // total for sample b.  Each ballot transposes one bit-plane across the wave.
  __device__ __forceinline__ uint32_t wave_total() const {
    const uint32_t lane = __lane_id();
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (__ballot(c[j] != 0) == 0) continue;
      for (int b = 0; b < 64; ++b) {
        const uint64_t m = __ballot((int)((c[j] >> b) & 1ull));
        if (lane == (uint32_t)b) acc += (uint32_t)__popcll(m) << j;
      }
    }
    return acc;
  }
};

}  // namespace msa
