// photon.hpp -- device-side optical-photon physics for gfx950 (wave64).
//
// One lane owns one photon. The arithmetic of every operation follows the
// reference's physics (include/optphot/{philox,pmath,vec3,sampling,fresnel,
// geometry,transport}.hpp) in the same evaluation order, because only then do
// the correctly rounded f32 operations produce the same bits as the CPU path.
// The code is compiled with -ffp-contract=off and IEEE division / square root.
//
// What is organised differently from the reference, for wave64 SIMT:
//  * the step is mostly branch-free: every lane computes the travelled
//    distance d = min(absorb, scatter, wall) and advances with one update;
//    Fresnel (readout or polished face) is evaluated once for all lanes that
//    need it (and, in kernels that keep per-cell bounds of R, only for the
//    few decisions the bounds leave open), and one rotate_about() serves both
//    Rayleigh scattering and Lambertian reflection, so mixed waves do not
//    execute those bodies twice;
//  * specular reflection on an axis-aligned face is the exact negation of one
//    direction component (d - 2 (d.n) n rounds to exactly that);
//  * the Philox stream keeps an 8-word window (two blocks) that is topped up
//    once, at a wave-uniform point at the top of the step, so that the <= 5
//    words a step consumes never trigger a Philox call inside a divergent
//    branch. Each draw then reads the window at a compile-time offset.
// Draw order, decision inequalities, tie rules and fate codes are exactly the
// reference's (transport.hpp:55-145), so photon i's history is the same.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ophip {

constexpr float kInf = __builtin_huge_valf();
constexpr int kReadoutFace = 5;

enum : uint32_t { kPolished = 0, kBlack = 1, kSpecular = 2, kLambertian = 3 };
enum : uint32_t { kPoint = 0, kVolume = 1 };
enum : uint32_t {
  kDetected = 0,
  kAbsorbedBulk = 1,
  kAbsorbedSurface = 2,
  kEscaped = 3,
  kMaxSteps = 4
};

// Per-run constants in the form the step uses. Derived fields are computed on
// the host with the same single IEEE operation the reference performs inside
// the loop (inv_speed = n_scint / c, transport.hpp:74; eta = n1 / n2 and
// eta * eta, fresnel.hpp:37-38), so hoisting them changes no bit.
struct DevParams {
  float hx, hy, hz;
  float inv_speed;
  float abs_length, scat_length;  // +inf disables the process
  float n_scint, n_det, n_out;
  float eta2_det, eta2_out;       // (n_scint / n_det)^2, (n_scint / n_out)^2
  float reflectivity;
  float src_x, src_y, src_z;
  float tau;
  uint32_t det_matched, out_matched;  // n_scint == n_det, n_scint == n_out
  uint32_t surface, source, max_steps;
  // Operand ranges that make the short divisions (div_n) exact, set on the
  // host: bit 0 = every box half size in [2^-60, 2^40] (box exit and hit-map
  // bins), bit 1 = n_scint, n_det and n_out in [2^-20, 2^20] and eta^2 != 1
  // for each unmatched pair (Fresnel).
  uint32_t fast_div;
  // Rayleigh cosines of every 24-bit uniform (2^24 entries, entry k for
  // u = (k + 1) 2^-24, filled on the device by rayleigh_cos); nullptr: the
  // step computes them.
  const float* rayleigh;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC11; the reference's philox.hpp:37-76).
// Counter {c0, 0, c2, c3}; the second word is always 0 in this path.

struct PhiloxKey {
  uint32_t k0, k1;
};

__device__ __forceinline__ void philox10(uint32_t c0, uint32_t c2, uint32_t c3, PhiloxKey key,
                                         uint32_t& o0, uint32_t& o1, uint32_t& o2, uint32_t& o3) {
  uint32_t x0 = c0, x1 = 0u, x2 = c2, x3 = c3;
  uint32_t k0 = key.k0, k1 = key.k1;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(x0) * 0xD2511F53u;
    const uint64_t p1 = static_cast<uint64_t>(x2) * 0xCD9E8D57u;
    // hi ^ x ^ k as one gfx950 three-input bit operation (truth table 0x96 =
    // XOR of the three), instead of two v_xor_b32
    const uint32_t y0 = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(p1 >> 32), x1, k0, 0x96);
    const uint32_t y2 = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(p0 >> 32), x3, k1, 0x96);
    x1 = static_cast<uint32_t>(p1);
    x3 = static_cast<uint32_t>(p0);
    x0 = y0;
    x2 = y2;
    k0 += 0x9E3779B9u;  // Weyl key schedule
    k1 += 0xBB67AE85u;
    // Keep the schedule as two scalar adds per round: left visible, the
    // compiler hoists all 20 round keys out of the transport loop, and the
    // scalar registers they occupy spill into VGPR lanes (v_readlane in the
    // hot loop).
    asm volatile("" : "+s"(k0), "+s"(k1));
  }
  o0 = x0;
  o1 = x1;
  o2 = x2;
  o3 = x3;
}

// Top 24 bits -> k * 2^-24 in (0, 1] (philox.hpp:82-84).
__device__ __forceinline__ float to_unit(uint32_t bits) {
  // (k + 1) 2^-24 with k = bits >> 8 < 2^24, as one FMA: k 2^-24 and the sum
  // are exact in f32, so this is the reference's value bit for bit.
  return __builtin_fmaf(static_cast<float>(bits >> 8), 1.0f / 16777216.0f, 1.0f / 16777216.0f);
}

// 2 u - 1 for u = to_unit(bits), as the reference computes it (2.0f * u -
// 1.0f, transport.hpp:55-67): 2 u = (k + 1) 2^-23 is exact and so is
// 2 u - 1 = (k + 1 - 2^23) 2^-23 (an integer below 2^23 in magnitude times
// 2^-23). One FMA of k 2^-23 and 2^-23 - 1 (exact in f32) rounds that exact
// value, so it gives the same bits.
__device__ __forceinline__ float sym_unit(uint32_t bits) {
  return __builtin_fmaf(static_cast<float>(bits >> 8), 0x1p-23f, 0x1p-23f - 1.0f);
}

// A photon's stream: words [4*(blk-2), 4*blk) are held in w0..w7, `pos` is
// the next unread word inside that window. Refill keeps pos <= 3, so at least
// five words are readable without generating.
struct Stream {
  uint32_t w0, w1, w2, w3, w4, w5, w6, w7;
  uint32_t pos, blk;
  uint32_t id_lo, id_hi;

  __device__ __forceinline__ void start(uint64_t photon_id, PhiloxKey key, uint32_t draws) {
    id_lo = static_cast<uint32_t>(photon_id);
    id_hi = static_cast<uint32_t>(photon_id >> 32);
    const uint32_t b = draws >> 2;
    philox10(b, id_lo, id_hi, key, w0, w1, w2, w3);
    philox10(b + 1u, id_lo, id_hi, key, w4, w5, w6, w7);
    blk = b + 2u;
    pos = draws & 3u;
  }

  __device__ __forceinline__ void refill(PhiloxKey key) {
    while (pos >= 4u) {
      w0 = w4;
      w1 = w5;
      w2 = w6;
      w3 = w7;
      philox10(blk, id_lo, id_hi, key, w4, w5, w6, w7);
      ++blk;
      pos -= 4u;
    }
  }

  // Word at window offset pos + k, k in [0, 4] (pos in [0, 3]).
  template <int K>
  __device__ __forceinline__ uint32_t peek() const {
    const uint32_t i = pos + K;
    const uint32_t a = (i & 1u) ? w1 : w0;
    const uint32_t b = (i & 1u) ? w3 : w2;
    const uint32_t c = (i & 1u) ? w5 : w4;
    const uint32_t d = (i & 1u) ? w7 : w6;
    const uint32_t lo = (i & 2u) ? b : a;
    const uint32_t hi = (i & 2u) ? d : c;
    return (i & 4u) ? hi : lo;
  }

  __device__ __forceinline__ uint32_t draws() const { return 4u * (blk - 2u) + pos; }
};

// The same stream with its window in LDS: a ring of 16 words per lane,
// slot-major (slot j of lane l at byte j * 2^shift + 4 l of the ring region,
// 2^shift = 4 x the block size rounded up to a power of two), so a wave's 64
// accesses to one slot index fall in distinct banks. `pos` counts the words
// consumed and `wr` the words generated (4 per Philox block); words
// pos .. wr-1 are readable and wr - pos <= 16. Reading a word is one LDS load
// at a computed address instead of a select chain over registers, and a lane
// with room can take a block whenever the wave runs a generation pass, so
// passes are shared out over many steps.
//
// With kAligned the region's LDS byte address is a multiple of its size
// (16 << shift), so the slot bits of a lane's address are zero and the
// address of word w is ((w << shift) & slot mask) | lane: two instructions
// (v_add_lshl, v_and_or). Otherwise the lane's address is added (three).
typedef __attribute__((address_space(3))) uint32_t lds_u32;

constexpr uint32_t kRingSlots = 16u;

// log2 of the slot stride in bytes for the running block.
__device__ __forceinline__ uint32_t ring_shift() {
  const uint32_t b = blockDim.x;
  return 2u + (b > 1u ? 32u - static_cast<uint32_t>(__clz(b - 1u)) : 0u);
}

// 32-bit words of the ring region of the running block.
__device__ __forceinline__ uint32_t ring_region_words() { return kRingSlots << (ring_shift() - 2u); }

template <bool kAligned>
struct RingStreamT {
  uint32_t lane;   // LDS byte address of this lane's slot 0
  uint32_t shift;  // log2 of the slot stride in bytes (uniform)
  uint32_t mask;   // the slot bits of an address, (kSlots - 1) << shift (uniform)
  uint32_t pos, wr;
  uint32_t id_lo, id_hi;

  static constexpr uint32_t kSlots = kRingSlots;

  // `region`: kSlots << ring_shift() bytes of LDS (aligned as above when
  // kAligned); the stream starts empty.
  __device__ __forceinline__ void place(const uint32_t* region) { place(region, ring_shift()); }
  // The same with the slot stride given (it must equal ring_shift()): a
  // compile-time stride turns a block's four slot stores into two
  // ds_write2st64_b32 with immediate offsets.
  __device__ __forceinline__ void place(const uint32_t* region, uint32_t sh) {
    shift = sh;
    mask = (kSlots - 1u) << shift;
    // kept opaque: seen through, the compiler rewrites (w << shift) & mask as
    // (w & 15) << shift, one instruction more per access
    asm volatile("" : "+s"(mask));
    lane = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((const lds_u32*)region)) +
           4u * threadIdx.x;
    pos = wr = 0u;
    id_lo = id_hi = 0u;
  }
  __device__ __forceinline__ uint32_t addr(uint32_t word) const {
    const uint32_t slot = (word << shift) & mask;
    return kAligned ? (slot | lane) : slot + lane;
  }
  __device__ __forceinline__ static lds_u32* at(uint32_t a) {
    return reinterpret_cast<lds_u32*>(static_cast<uintptr_t>(a));
  }
  __device__ __forceinline__ void reset(uint64_t photon_id) {
    id_lo = static_cast<uint32_t>(photon_id);
    id_hi = static_cast<uint32_t>(photon_id >> 32);
    pos = 0u;
    wr = 0u;
  }
  __device__ __forceinline__ uint32_t level() const { return wr - pos; }
  __device__ __forceinline__ bool room() const { return wr - pos <= kSlots - 4u; }
  // Appends the next Philox block (counter wr / 4) to the ring; its four
  // slots are consecutive (wr is a multiple of 4).
  __device__ __forceinline__ void generate(PhiloxKey key) {
    uint32_t a, b, c, d;
    philox10(wr >> 2, id_lo, id_hi, key, a, b, c, d);
    const uint32_t p = addr(wr), s = 1u << shift;
    *at(p) = a;
    *at(p + s) = b;
    *at(p + 2u * s) = c;
    *at(p + 3u * s) = d;
    wr += 4u;
  }
  // Word pos + K (the caller keeps pos + K < wr).
  template <int K>
  __device__ __forceinline__ uint32_t peek() const {
    return *at(addr(pos + K));
  }
};
typedef RingStreamT<true> RingStream;

// Correctly rounded a / b and sqrt(x) for operands in a normal range, where
// the compiler's full IEEE sequences spend instructions on range scaling and
// special-case fix-up that cannot act:
//  * div_n: the standard refined-reciprocal sequence (rcp, one Newton step on
//    the reciprocal, two residual corrections of the quotient) without
//    v_div_scale / v_div_fixup. Those only rescale or replace the result when
//    an operand is 0, inf, NaN or denormal, the exponents differ by >= 96, or
//    the quotient or 1 / b is denormal; for b normal with 2^-100 < |b| < 2^100
//    and a == 0 or 2^-100 < |a| < 2^100 (quotient normal) both give the same
//    bits (a == 0 gives +-0 either way).
//  * sqrt_n: hardware estimate plus the one-ulp correction of both
//    neighbours, without the 2^32 pre-scaling of inputs below 2^-96 and the
//    +-0 / +inf pass-through: identical for x == +-0 and x >= 2^-96 finite.
// Every call site states why its operands are in range.
__device__ __forceinline__ float div_n(float a, float b) {
  float r = __builtin_amdgcn_rcpf(b);
  const float e = __builtin_fmaf(-b, r, 1.0f);
  r = __builtin_fmaf(e, r, r);
  float q = a * r;
  float t = __builtin_fmaf(-b, q, a);
  q = __builtin_fmaf(t, r, q);
  t = __builtin_fmaf(-b, q, a);
  return __builtin_fmaf(t, r, q);
}

__device__ __forceinline__ float sqrt_n(float x) {
  float y = __builtin_amdgcn_sqrtf(x);
  const float dn = __uint_as_float(__float_as_uint(y) - 1u);
  const float up = __uint_as_float(__float_as_uint(y) + 1u);
  const float r_dn = __builtin_fmaf(-dn, y, x);
  const float r_up = __builtin_fmaf(-up, y, x);
  y = (r_dn <= 0.0f) ? dn : y;
  y = (r_up > 0.0f) ? up : y;
  return y;
}

// ---------------------------------------------------------------------------
// Transcendentals from correctly rounded operations only (the reference's
// portable math, pmath.hpp:62-130): bit-identical to the CPU build with
// OPTPHOT_PORTABLE_MATH.

// x positive and normal. The reference splits x = m 2^e with m the mantissa
// in [1, 2), then halves m and increments e when m > 1.41421356f, i.e. when
// the mantissa field is at least 0x3504f4 (1.41421356f rounds to 0x3fb504f3).
// With d = bits(x) - 0x3f3504f4 (0x3f3504f4 = 0x3fb504f4 halved) the borrow
// out of the mantissa field makes e = d >> 23 (arithmetic) and
// bits(m) = (d & 0x7fffff) + 0x3f3504f4 exactly those: 4 integer operations
// instead of the extraction, compare and two selects.
__device__ __forceinline__ float p_log(float x) {
  const int d = static_cast<int>(__float_as_uint(x)) - 0x3f3504f4;
  const int e = d >> 23;
  const float m = __int_as_float((d & 0x007fffff) + 0x3f3504f4);
  // m in [0.707, 1.415): m - 1 is 0 or at least 2^-24 in magnitude, m + 1 in [1.7, 2.5]
  const float s = div_n(m - 1.0f, m + 1.0f);
  const float s2 = s * s;
  const float poly =
      1.0f + s2 * (1.0f / 3.0f +
                   s2 * (1.0f / 5.0f + s2 * (1.0f / 7.0f + s2 * (1.0f / 9.0f + s2 * (1.0f / 11.0f)))));
  const float fe = static_cast<float>(e);
  return fe * 6.9313812256e-01f + (2.0f * s * poly + fe * 9.0580006145e-06f);
}

// sin and cos of 2 pi u for the uniform u = (k + 1) 2^-24 drawn from word w
// (k = w >> 8, to_unit). The reference reduces v = 4 u: q = floor(v),
// r = v - q, and r > 1/2 becomes r - 1 with q + 1; then x = r (pi / 2). Every
// step before the product is exact: v = (k + 1) 2^-22, so q and r are the
// quotient and remainder of k + 1 by 2^22, rounded to nearest with ties down
// (r = 1/2 stays). With t = k + 2^21: q = t >> 22 and r = ri 2^-22 with the
// integer ri = (t & (2^22 - 1)) - (2^21 - 1) in (-2^21, 2^21], and
// x = ri (2^-22 pi/2) rounds as r (pi/2) does (a power-of-two scale commutes
// with rounding, and x is 0 or normal). The polynomials are unchanged.
__device__ __forceinline__ void p_sincos_word(uint32_t w, float& sn, float& cs) {
  const uint32_t t = (w >> 8) + (1u << 21);
  const uint32_t q = t >> 22;
  const int ri = static_cast<int>(t & 0x003fffffu) - ((1 << 21) - 1);
  const float x = static_cast<float>(ri) * (1.57079632679489662f * 0x1p-22f);
  const float x2 = x * x;
  const float sp =
      x * (1.0f + x2 * (-1.0f / 6.0f +
                        x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f)))));
  const float cp =
      1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f +
                                 x2 * (-1.0f / 720.0f +
                                       x2 * (1.0f / 40320.0f + x2 * (-1.0f / 3628800.0f)))));
  const int quadrant = static_cast<int>(q & 3u);
  // quadrant rotation as selects: (s, c) = (sp, cp), (cp, -sp), (-sp, -cp), (-cp, sp)
  const bool odd = quadrant & 1;
  const float a = odd ? cp : sp;
  const float b = odd ? sp : cp;
  sn = (quadrant & 2) ? -a : a;
  cs = (quadrant == 1 || quadrant == 2) ? -b : b;
}

// Called for x in [1, 4.3] only (Rayleigh sampling), where every Newton
// quotient is normal: y ~ x^(1/3) in [0.9, 1.7], 3 y^2 in [2.4, 8.7], and
// y^3 - x is 0 or at least 2^-24 in magnitude.
__device__ __forceinline__ float p_cbrt(float x) {
  float y = __uint_as_float(__float_as_uint(x) / 3u + 0x2a514067u);
#pragma unroll
  for (int i = 0; i < 4; ++i) y = y - div_n(y * y * y - x, 3.0f * y * y);
  return y;
}


// ---------------------------------------------------------------------------
// Vector helpers (vec3.hpp): only what the step needs, same operation order.

struct V3 {
  float x, y, z;
};

// 1 / sqrt(s) with both operations correctly rounded (the reference's
// normalisation, vec3.hpp:35-53), for |s| within 1024 ulps of 1, from integer
// operations on the bits of s. With s = 1 + i 2^-23 (i >= 0) the rounded
// root is 1 + floor(i / 2) 2^-23 and its rounded reciprocal 1 - 2 floor(i / 2)
// 2^-24; with s = 1 - |i| 2^-24 (i < 0) the root is 1 - ceil(|i| / 2) 2^-24
// and the reciprocal 1 + ceil(|i| / 4) 2^-23. The second-order terms of the
// root and reciprocal series are positive and below a quarter of an ulp, so
// they only decide on which side a tie falls. Checked against the IEEE
// operations for every |i| <= 2000. Any other s (a rotation leaves |d|^2 within
// a few ulps of 1) takes div_n(1, sqrt_n(s)).
__device__ __forceinline__ float inv_sqrt_near_one(float s) {
  const int i = static_cast<int>(__float_as_uint(s)) - 0x3f800000;
  if (static_cast<uint32_t>(i + 1024) <= 2048u)
    return __int_as_float(i >= 0 ? 0x3f800000 - (i & ~1) : 0x3f800000 + ((3 - i) >> 2));
  return div_n(1.0f, sqrt_n(s));  // s is normal here (|d|^2 of a rotated unit vector)
}

// rotate_about (vec3.hpp:35-53): Duff et al. basis around `n`, then
// normalize((st cp) t1 + (st sp) t2 + ct n) with normalize(a) = (1/|a|) a,
// the azimuth 2 pi to_unit(phi_word).
__device__ __forceinline__ V3 rotate_about(V3 n, float cos_theta, uint32_t phi_word) {
  const float sign = __builtin_copysignf(1.0f, n.z);
  // n is a unit vector and sign = copysign(1, n.z): sign + n.z in [1, 2] or [-2, -1]
  const float a = div_n(-1.0f, sign + n.z);
  const float b = n.x * n.y * a;
  const V3 t1{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  const V3 t2{b, sign + n.y * n.y * a, -n.y};
  // |cos| <= 1: 1 - cos^2 is 0 or at least 2^-24
  const float st = sqrt_n(__builtin_fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
  float sp, cp;
  p_sincos_word(phi_word, sp, cp);
  const float c1 = st * cp, c2 = st * sp;
  const V3 d{(c1 * t1.x + c2 * t2.x) + cos_theta * n.x, (c1 * t1.y + c2 * t2.y) + cos_theta * n.y,
             (c1 * t1.z + c2 * t2.z) + cos_theta * n.z};
  // |d|^2 = 1 up to rounding
  const float inv = inv_sqrt_near_one(d.x * d.x + d.y * d.y + d.z * d.z);
  return V3{inv * d.x, inv * d.y, inv * d.z};
}

// Cardano root of the Rayleigh CDF (sampling.hpp:39-44).
__device__ __forceinline__ float rayleigh_cos(float u) {
  const float q = 4.0f * u - 2.0f;
  const float s = sqrt_n(q * q + 1.0f);  // q in [-2, 2]: q^2 + 1 in [1, 5]
  // A = cbrt(q + s), or 1 / cbrt(s - q) for q < 0: one cube root on the
  // selected argument (a wave with lanes on both sides would otherwise run
  // two Newton iterations side by side), then the lane's own reciprocal.
  const bool pos = q >= 0.0f;
  const float c = p_cbrt(pos ? q + s : s - q);
  const float a = pos ? c : div_n(1.0f, c);  // c in [1, 1.62], so a in [0.61, 1.62]
  return __builtin_fminf(1.0f, __builtin_fmaxf(-1.0f, a - div_n(1.0f, a)));
}

// Rayleigh cosine of the uniform drawn from word w (to_unit keeps its top 24
// bits), from the table when the run has one. The table entries are this
// function's own values, so both give the same bits; the table replaces ~90
// instructions that a wave runs whenever any of its lanes scatters.
__device__ __forceinline__ float rayleigh_cos_word(const float* table, uint32_t w) {
  if (table) return table[w >> 8];
  return rayleigh_cos(to_unit(w));
}

// Tally binning (transport.hpp:170-181); `time_scale` = time_bins / time_max
// and `two_hx`, `two_hy` = 2 hx, 2 hy, computed once on the host.
__device__ __forceinline__ uint32_t time_bin(float t, float time_scale, uint32_t bins) {
  const float f = t * time_scale;
  return (f >= static_cast<float>(bins)) ? bins : static_cast<uint32_t>(f);
}

// With hx, hy in [2^-60, 2^40] (fast), x + hx is 0 or at least half an ulp
// of hx (x lies on the readout face, within [-hx, hx]) and 2 hx is normal, so
// div_n gives the IEEE quotients.
__device__ __forceinline__ uint32_t xy_bin(float x, float y, float hx, float hy, float two_hx,
                                           float two_hy, uint32_t nb_bins, bool fast) {
  const float nb = static_cast<float>(nb_bins);
  const float fx = fast ? div_n(x + hx, two_hx) : (x + hx) / two_hx;
  const float fy = fast ? div_n(y + hy, two_hy) : (y + hy) / two_hy;
  const uint32_t ix = static_cast<uint32_t>(__builtin_fminf(fx * nb, nb - 1.0f));
  const uint32_t iy = static_cast<uint32_t>(__builtin_fminf(fy * nb, nb - 1.0f));
  return iy * nb_bins + ix;
}

// ---------------------------------------------------------------------------
// Photon state and the transport step.

struct Photon {
  V3 p, d;
  float t, path;
  uint32_t nb, ns;
};

// Likelihood-ratio scores a photon accumulates along its history (adjoint
// kernel only): d log p(history) / d n_det and d n_scint from every Fresnel
// decision, and the number of reflector hits (for d / d reflectivity).
//
// Readout decisions carry a baseline: with b = w[bin(t)], the weight the
// photon would score if transmitted at this hit (known before the draw), the
// decision contributes (g - b) s instead of g s, g the photon's final weight.
// This is unbiased because E[s | state before the draw] = 0, and it makes the
// transmitted branch contribute exactly (b - b) s = 0: near total internal
// reflection that branch has s = -dR / (1 - R) ~ 1 / x with probability ~ x^(1/2)
// (x the distance to the critical angle), an infinite-variance term whose
// finite-sample means are visibly biased. Reflected readout decisions add s
// to the score and b s to `base`; transmitted ones add nothing.
struct AdjState {
  float s_ndet, s_nscint;
  uint32_t refl_hits;
  double base_ndet, base_nscint;
  const double* w;  // the set's weights, time_bins + 1 of them
  float time_scale;
  uint32_t time_bins;
};

// Score of one Fresnel decision: d log P(decision) / d n1 and / d n2 with
// P = R (reflected) or 1 - R (transmitted), R = (rs^2 + rp^2) / 2 and
//   cos_t = sqrt(1 - (n1/n2)^2 (1 - cos_i^2)),
//   rs = (n1 ci - n2 ct) / (n1 ci + n2 ct),  rp = (n2 ci - n1 ct) / (n2 ci + n1 ct).
// Evaluated in f64 from the f32 inputs of the history: close to total
// internal reflection cos_t is tiny, 1 - R and d cos_t / dn ~ 1 / cos_t lose
// most of their bits in f32, and those are the photons with the largest
// scores. 1 - R uses the cancellation-free form
//   1 - R = 2 n1 n2 ci ct (1 / (n1 ci + n2 ct)^2 + 1 / (n2 ci + n1 ct)^2).
// f64 reciprocal and inverse square root for the scores: an f32 hardware
// estimate refined by two Newton steps (each squares the relative error:
// ~2^-22 -> 2^-44 -> below f64 rounding). Arguments are positive, normal and
// within f32 range here.
__device__ __forceinline__ double rcp64(double x) {
  double r = static_cast<double>(__builtin_amdgcn_rcpf(static_cast<float>(x)));
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return r;
}
__device__ __forceinline__ double rsqrt64(double x) {
  double r = static_cast<double>(__builtin_amdgcn_rsqf(static_cast<float>(x)));
  r = r * __builtin_fma(-0.5 * x, r * r, 1.5);
  r = r * __builtin_fma(-0.5 * x, r * r, 1.5);
  return r;
}

__device__ __forceinline__ void fresnel_score(float n1f, float n2f, float cif, bool reflected,
                                              float& d1, float& d2) {
  const double n1 = n1f, n2 = n2f, ci = cif;
  const double s2i = 1.0 - ci * ci > 0.0 ? 1.0 - ci * ci : 0.0;
  const double rn2 = rcp64(n2);
  const double eta = n1 * rn2;
  const double v = 1.0 - eta * eta * s2i;  // cos_t^2
  d1 = d2 = 0.0f;
  if (!(v > 0.0)) return;
  const double rct = rsqrt64(v);  // 1 / cos_t
  const double ct = v * rct;
  const double A = n1 * ci - n2 * ct, B = n1 * ci + n2 * ct;
  const double C = n2 * ci - n1 * ct, D = n2 * ci + n1 * ct;
  const double rB = rcp64(B), rD = rcp64(D);
  const double rs = A * rB, rp = C * rD;
  const double dct1 = -n1 * s2i * rn2 * rn2 * rct;
  const double dct2 = n1 * n1 * s2i * rn2 * rn2 * rn2 * rct;
  const double drs1 = ((ci - n2 * dct1) * B - A * (ci + n2 * dct1)) * rB * rB;
  const double drs2 = ((-ct - n2 * dct2) * B - A * (ct + n2 * dct2)) * rB * rB;
  const double drp1 = ((-ct - n1 * dct1) * D - C * (ct + n1 * dct1)) * rD * rD;
  const double drp2 = ((ci - n1 * dct2) * D - C * (ci + n1 * dct2)) * rD * rD;
  const double dR1 = rs * drs1 + rp * drp1, dR2 = rs * drs2 + rp * drp2;
  const double inv = reflected ? 2.0 * rcp64(rs * rs + rp * rp)
                               : -rcp64(2.0 * n1 * n2 * ci * ct * (rB * rB + rD * rD));
  d1 = static_cast<float>(dR1 * inv);
  d2 = static_cast<float>(dR2 * inv);
}

// Per-step classification, exposed so kernels can count branch events.
struct StepInfo {
  bool absorb, scatter, readout, polished, black, refl, survived, lambert_reflect;
  bool abs_on, scat_on;  // the step sampled an absorption / scattering distance
};

// Emission (transport.hpp:55-67): [x, y, z if volume], cos, phi, time.
// The stream must hold the photon's first 6 words (Stream: pos == 0 after
// start(); RingStream: level() >= 6).
template <class RNG>
__device__ __forceinline__ void emit(const DevParams& P, RNG& rng, Photon& ph) {
  uint32_t o = 0;
  if (P.source == kVolume) {
    ph.p = V3{P.hx * sym_unit(rng.template peek<0>()), P.hy * sym_unit(rng.template peek<1>()),
              P.hz * sym_unit(rng.template peek<2>())};
    o = 3;
  } else {
    ph.p = V3{P.src_x, P.src_y, P.src_z};
  }
  // The window holds words 0..7 and o <= 3, so the offsets o..o+2 are read
  // through peek with pos advanced first.
  rng.pos += o;
  const float cos_t = sym_unit(rng.template peek<0>());
  const uint32_t wb = rng.template peek<1>();
  const float ut = to_unit(rng.template peek<2>());
  rng.pos += 3u;
  const float sin_t = sqrt_n(__builtin_fmaxf(0.0f, 1.0f - cos_t * cos_t));  // 0 or >= 2^-24
  float sp, cp;
  p_sincos_word(wb, sp, cp);
  ph.d = V3{sin_t * cp, sin_t * sp, cos_t};
  ph.t = P.tau > 0.0f ? -P.tau * p_log(ut) : 0.0f;
  ph.path = 0.0f;
  ph.nb = 0u;
  ph.ns = 0u;
}

// A genstep on the device: emission segment, its time span, and the global
// index of its first photon (photons of earlier gensteps come first).
struct DevGenstep {
  float x0, y0, z0, t0;
  float x1, y1, z1, t1;
  uint32_t n_photons;
  float cone_cos;  // 0: scintillation; (0, 1]: Cherenkov cone
  unsigned long long offset;
};

// Emission from a genstep (no counterpart in the reference, which lists
// generation from gensteps as future work, README.md:734-735). Draw order:
// [u along the segment, only for a non-degenerate segment or time span], cos,
// phi, emission delay; position x0 + u (x1 - x0) per axis, time
// t0 + u (t1 - t0) + delay. A degenerate genstep with t0 = 0 draws and computes
// exactly what the point source at (x0, y0, z0) does (transport.hpp:55-67).
// Cherenkov gensteps (cone_cos > 0; the host checks that the segment has a
// direction): draws u along the segment, then phi; position and time as
// above, no delay, and the direction rotate_about(axis, cone_cos, phi) with
// axis = (x1 - x0) / |x1 - x0| (IEEE square root and division).
// Needs 4 words in the stream.
template <class RNG>
__device__ __forceinline__ void emit_genstep(const DevParams& P, const DevGenstep& G, RNG& rng,
                                             Photon& ph) {
  if (G.cone_cos > 0.0f) {
    const float u = to_unit(rng.template peek<0>());
    const uint32_t wp = rng.template peek<1>();
    rng.pos += 2u;
    const float dx = G.x1 - G.x0, dy = G.y1 - G.y0, dz = G.z1 - G.z0;
    ph.p = V3{G.x0 + u * dx, G.y0 + u * dy, G.z0 + u * dz};
    const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    ph.d = rotate_about(V3{inv * dx, inv * dy, inv * dz}, G.cone_cos, wp);
    ph.t = G.t0 + u * (G.t1 - G.t0);
    ph.path = 0.0f;
    ph.nb = 0u;
    ph.ns = 0u;
    return;
  }
  ph.p = V3{G.x0, G.y0, G.z0};
  float t0 = G.t0;
  if (G.x0 != G.x1 || G.y0 != G.y1 || G.z0 != G.z1 || G.t0 != G.t1) {
    const float u = to_unit(rng.template peek<0>());
    ph.p = V3{G.x0 + u * (G.x1 - G.x0), G.y0 + u * (G.y1 - G.y0), G.z0 + u * (G.z1 - G.z0)};
    t0 = G.t0 + u * (G.t1 - G.t0);
    rng.pos += 1u;
  }
  const float cos_t = sym_unit(rng.template peek<0>());
  const uint32_t wb = rng.template peek<1>();
  const float ut = to_unit(rng.template peek<2>());
  rng.pos += 3u;
  const float sin_t = sqrt_n(__builtin_fmaxf(0.0f, 1.0f - cos_t * cos_t));  // 0 or >= 2^-24
  float sp, cp;
  p_sincos_word(wb, sp, cp);
  ph.d = V3{sin_t * cp, sin_t * sp, cos_t};
  ph.t = t0 + (P.tau > 0.0f ? -P.tau * p_log(ut) : 0.0f);
  ph.path = 0.0f;
  ph.nb = 0u;
  ph.ns = 0u;
}

// The phases of one transport step (transport.hpp:71-145), shared by the
// fused step() below and the per-interaction kernels, which run them in
// separate launches on compacted queues.

// (±h - p) / d for the box exit. With the box half sizes in [2^-60, 2^40]
// (P.fast_div bit 0) the numerator is 0 or between 2^-86 and 2^41 in
// magnitude (p lies in the box, so a nonzero ±h - p is at least half an ulp
// of h), and with |d| >= 2^-60 the quotient is normal: div_n then gives the
// IEEE quotient. A component below 2^-60, or a box outside that range, takes
// the full division.
__device__ __forceinline__ float box_div(const DevParams& P, float num, float d) {
  if ((P.fast_div & 1u) && __builtin_fabsf(d) >= 0x1p-60f) return div_n(num, d);
  return num / d;
}

// fminf(fmaxf(v, -h), h), the clamp of the face snap (geometry.hpp:57-63).
// With the box half sizes in [2^-60, 2^40] (P.fast_div bit 0) every position
// is finite (each advance is bounded by the distance to a wall) and h > 0, so
// the median of v, -h and h is the same value in one instruction (no NaN
// operand, and no tie between a zero and a signed zero).
__device__ __forceinline__ float clamp_half(const DevParams& P, float v, float h) {
  if (P.fast_div & 1u) return __builtin_amdgcn_fmed3f(v, -h, h);
  return __builtin_fminf(__builtin_fmaxf(v, -h), h);
}

// Where the step ends: absorbed in the bulk, scattered, or on face `face`
// (axis = face / 2, plus = odd face); cos_i = dot(d, outward normal) at a
// boundary.
struct Advance {
  bool absorb, scatter, boundary, plus;
  int face, axis;
  float cos_i;
};

// Some component of d is nonzero and below 2^-60 in magnitude (subnormals
// included): (bits << 1) - 2 drops the sign and sends +-0 to the top.
__device__ __forceinline__ bool tiny_component(float dx, float dy, float dz) {
  constexpr uint32_t kMin = (0x21800000u << 1) - 2u;  // 2^-60
  const uint32_t x = (__float_as_uint(dx) << 1) - 2u;
  const uint32_t y = (__float_as_uint(dy) << 1) - 2u;
  const uint32_t z = (__float_as_uint(dz) << 1) - 2u;
  return (x < y ? (x < z ? x : z) : (y < z ? y : z)) < kMin;
}


// Phase 1: the two exponential distances (draws 0 and 1, always drawn,
// transport.hpp:76-77), the distance to the wall approached (geometry.hpp:
// 25-46), which event wins, and one shared advance with the snap onto a face
// (geometry.hpp:57-63) and the boundary / scatter counts.
template <class RNG>
__device__ __forceinline__ Advance advance(const DevParams& P, const RNG& rng, Photon& ph) {
  const float u_abs = to_unit(rng.template peek<0>());
  const float u_sca = to_unit(rng.template peek<1>());
  float s_abs = kInf, s_sca = kInf;
  if (P.abs_length != kInf) s_abs = -P.abs_length * p_log(u_abs);
  if (P.scat_length != kInf) s_sca = -P.scat_length * p_log(u_sca);

  // --- distance to the wall approached, ties to the lower axis --------------
  float dist = kInf;
  int face = -1;
  // (evaluated ahead of the branch: inside its condition, hipcc 7.2 keeps a
  // stack copy of the direction and reads cos_i below from scratch memory)
  const bool tiny = tiny_component(ph.d.x, ph.d.y, ph.d.z);
  if ((P.fast_div & 1u) && !tiny) {
    // Every component is 0 or at least 2^-60 in magnitude. div_n gives the
    // IEEE quotient of a nonzero one (box_div), and NaN for a zero one (the
    // reciprocal estimate of 0 is infinite, and 0 * inf is NaN). NaN loses
    // every comparison, so a zero component leaves dist and face as the
    // skipped axis of the general path does, without a branch per axis.
    const float tx = div_n((ph.d.x > 0.0f ? P.hx : -P.hx) - ph.p.x, ph.d.x);
    const float ty = div_n((ph.d.y > 0.0f ? P.hy : -P.hy) - ph.p.y, ph.d.y);
    const float tz = div_n((ph.d.z > 0.0f ? P.hz : -P.hz) - ph.p.z, ph.d.z);
    if (tx < dist) {
      dist = tx;
      face = ph.d.x > 0.0f ? 1 : 0;
    }
    if (ty < dist) {
      dist = ty;
      face = ph.d.y > 0.0f ? 3 : 2;
    }
    if (tz < dist) {
      dist = tz;
      face = ph.d.z > 0.0f ? 5 : 4;
    }
  } else {  // one axis at a time, skipping zero components; the quotients of box_div
    if (ph.d.x != 0.0f) {
      const bool pos = ph.d.x > 0.0f;
      const float num = (pos ? P.hx : -P.hx) - ph.p.x;
      const float tt = box_div(P, num, ph.d.x);
      if (tt < dist) {
        dist = tt;
        face = pos ? 1 : 0;
      }
    }
    if (ph.d.y != 0.0f) {
      const bool pos = ph.d.y > 0.0f;
      const float num = (pos ? P.hy : -P.hy) - ph.p.y;
      const float tt = box_div(P, num, ph.d.y);
      if (tt < dist) {
        dist = tt;
        face = pos ? 3 : 2;
      }
    }
    if (ph.d.z != 0.0f) {
      const bool pos = ph.d.z > 0.0f;
      const float num = (pos ? P.hz : -P.hz) - ph.p.z;
      const float tt = box_div(P, num, ph.d.z);
      if (tt < dist) {
        dist = tt;
        face = pos ? 5 : 4;
      }
    }
  }
  dist = __builtin_fmaxf(dist, 0.0f);

  // --- which event wins, and one shared advance ------------------------------
  Advance a;
  a.absorb = s_abs < dist && s_abs <= s_sca;
  a.scatter = !a.absorb && s_sca < dist;
  a.boundary = !a.absorb && !a.scatter;
  const float go = a.absorb ? s_abs : (a.scatter ? s_sca : dist);
  ph.p = V3{ph.p.x + go * ph.d.x, ph.p.y + go * ph.d.y, ph.p.z + go * ph.d.z};
  ph.t += go * P.inv_speed;
  ph.path += go;

  a.face = face;
  a.axis = face >> 1;
  a.plus = face & 1;
  if (a.boundary) {  // snap onto the face plane, clamp the others
    const float s = a.plus ? 1.0f : -1.0f;
    ph.p.x = a.axis == 0 ? s * P.hx : clamp_half(P, ph.p.x, P.hx);
    ph.p.y = a.axis == 1 ? s * P.hy : clamp_half(P, ph.p.y, P.hy);
    ph.p.z = a.axis == 2 ? s * P.hz : clamp_half(P, ph.p.z, P.hz);
    ++ph.nb;
  }
  if (a.scatter) ++ph.ns;

  // cos_i = dot(d, outward normal) is exactly +-d[axis] (the zero components
  // contribute signed zeros that vanish in the sum).
  const float d_axis = a.axis == 0 ? ph.d.x : (a.axis == 1 ? ph.d.y : ph.d.z);
  a.cos_i = a.plus ? d_axis : -d_axis;
  return a;
}

// Phase 2 (readout and polished faces): the Fresnel reflection probability
// (fresnel.hpp:30-46) from the scintillator into the detector (readout) or
// the outside medium. `scored` is set when the decision carries an adjoint
// score: matched interfaces (R = 0 to second order) and total internal
// reflection (R = 1 on an open set of angles) carry none.
__device__ __forceinline__ float fresnel_prob(const DevParams& P, bool readout, float cos_i,
                                              bool& scored) {
  const float n1 = P.n_scint;
  const float n2 = readout ? P.n_det : P.n_out;
  const float eta2 = readout ? P.eta2_det : P.eta2_out;
  const bool matched = readout ? P.det_matched : P.out_matched;
  const float sin2_t = eta2 * __builtin_fmaxf(0.0f, 1.0f - cos_i * cos_i);
  // 1 - sin2_t is in (0, 1] and at least 2^-24 unless total internal
  // reflection (sin2_t >= 1), where R is not used
  const float cos_t = sqrt_n(1.0f - sin2_t);
  // With the indices in [2^-20, 2^20] and eta^2 != 1 unless matched
  // (P.fast_div bit 1), |1 - eta^2| >= 2^-24. When R is in use (no total
  // internal reflection, not matched), cos_i (eta^2 > 1) or cos_t
  // (eta^2 < 1) is then at least about 2^-12.5, so both denominators exceed
  // 2^-33 and the numerators are 0 or above 2^-58: div_n gives the IEEE
  // quotients. When R is not used its value does not matter.
  const float a_s = n1 * cos_i - n2 * cos_t, b_s = n1 * cos_i + n2 * cos_t;
  const float a_p = n2 * cos_i - n1 * cos_t, b_p = n2 * cos_i + n1 * cos_t;
  const bool fast = P.fast_div & 2u;
  const float rs = fast ? div_n(a_s, b_s) : a_s / b_s;
  const float rp = fast ? div_n(a_p, b_p) : a_p / b_p;
  const float R = 0.5f * (rs * rs + rp * rp);
  const bool tir = sin2_t >= 1.0f;
  // The score follows the exact geometry: a decision that float32 rounds into
  // total internal reflection (sin2_t within its rounding error of 1, so a
  // certain reflection here) is still scored when cos_t^2 > 0 in float64
  // (fresnel_score returns zero otherwise). Such grazing-critical hits carry
  // the 1/cos_t peak of dR/dn; leaving them out biases the gradient low
  // wherever photons revisit the readout. float32 sin2_t is within about six
  // rounding errors (< 2^-21) of its exact value, so a margin of 2^-18
  // keeps every such hit.
  scored = !matched && sin2_t < 1.0f + 0x1p-18f;
  return matched ? 0.0f : (tir ? 1.0f : R);
}

// Fresnel decisions from cell bounds. The decision of a readout or polished
// hit is u <= prob with u = (k + 1) 2^-24 (k the top 24 bits of the drawn
// word) and prob = fresnel_prob(cos_i). A kernel can keep, per interface, a
// table of kFresnelCells words: word j bounds prob over the cos_i in
// [j, j + 1] / kFresnelCells, packed as lo | hi << 16 with
//   lo 2^-16 <= prob <= (hi + 1) 2^-16   for every such cos_i.
// With k8 = word >> 16 (the top 16 bits of k): k8 < lo gives u <= lo 2^-16,
// so the photon is reflected; k8 > hi gives u > (hi + 1) 2^-16, so it is
// transmitted. Only when k8 lies in [lo, hi] is fresnel_prob evaluated, so
// every decision is the one fresnel_prob gives.
//
// The bounds are exact for the float32 computation, not for the
// mathematical R: every operation of fresnel_prob is a correctly rounded
// multiply, add, subtract, square root, maximum or division by a positive
// denominator, and each of these is monotone in each operand. Evaluating the
// same operations on the end points of each operand's range (all four
// corners of a quotient or difference) therefore bounds every value
// fresnel_prob can produce in the cell, with its own rounding; operands that
// depend on each other are only widened by this.
constexpr uint32_t kFresnelCells = 512u;

struct FInterval {
  float lo, hi;
};

__device__ __forceinline__ FInterval iv_quot(FInterval a, FInterval b) {  // b.lo > 0
  return FInterval{__builtin_fminf(a.lo / b.lo, a.lo / b.hi),
                   __builtin_fmaxf(a.hi / b.lo, a.hi / b.hi)};
}

__device__ __forceinline__ FInterval iv_square(FInterval a) {
  const float l = a.lo * a.lo, h = a.hi * a.hi;
  if (a.lo >= 0.0f) return FInterval{l, h};
  if (a.hi <= 0.0f) return FInterval{h, l};
  return FInterval{0.0f, __builtin_fmaxf(l, h)};
}

// Packed bounds of fresnel_prob(cos_i) for cos_i in [j, j + 1] / kFresnelCells,
// interface n1 -> n2 with eta2 = (n1 / n2)^2 as in DevParams. Uses IEEE
// division and square root, which give the bits of div_n and sqrt_n wherever
// fresnel_prob uses those.
__device__ __forceinline__ uint32_t fresnel_cell_bounds(float n1, float n2, float eta2,
                                                         bool matched, uint32_t j) {
  constexpr uint32_t kNever = 0xffff0000u;  // lo = 0, hi = 0xffff: always evaluated
  if (matched) return 0u;                    // prob = 0
  const float cl = static_cast<float>(j) * (1.0f / kFresnelCells);  // exact
  const float ch = static_cast<float>(j + 1u) * (1.0f / kFresnelCells);
  // sin2_t = eta2 max(0, 1 - c^2) does not increase with c
  const float s_hi = eta2 * __builtin_fmaxf(0.0f, 1.0f - cl * cl);
  const float s_lo = eta2 * __builtin_fmaxf(0.0f, 1.0f - ch * ch);
  if (s_lo >= 1.0f) return 0xffffffffu;  // total internal reflection throughout: prob = 1
  // cos_i of the cell outside total internal reflection: sin2_t <= 1 - 2^-24
  const float sn_hi = __builtin_fminf(s_hi, 0x1.fffffep-1f);
  const FInterval ct{__builtin_sqrtf(1.0f - sn_hi), __builtin_sqrtf(1.0f - s_lo)};
  const FInterval n1c{n1 * cl, n1 * ch}, n2c{n2 * cl, n2 * ch};
  const FInterval n1t{n1 * ct.lo, n1 * ct.hi}, n2t{n2 * ct.lo, n2 * ct.hi};
  const FInterval rs = iv_quot(FInterval{n1c.lo - n2t.hi, n1c.hi - n2t.lo},
                               FInterval{n1c.lo + n2t.lo, n1c.hi + n2t.hi});
  const FInterval rp = iv_quot(FInterval{n2c.lo - n1t.hi, n2c.hi - n1t.lo},
                               FInterval{n2c.lo + n1t.lo, n2c.hi + n1t.hi});
  const FInterval s2 = iv_square(rs), p2 = iv_square(rp);
  float lo = 0.5f * (s2.lo + p2.lo), hi = 0.5f * (s2.hi + p2.hi);
  if (s_hi >= 1.0f) {  // part of the cell is totally reflected (prob = 1)
    lo = __builtin_fminf(lo, 1.0f);
    hi = __builtin_fmaxf(hi, 1.0f);
  }
  if (!(lo == lo) || !(hi == hi)) return kNever;
  // lo16 = floor(lo 2^16), hi16 = ceil(hi 2^16) - 1, both clamped to [0, 0xffff]
  const float fl = __builtin_floorf(__builtin_fminf(__builtin_fmaxf(lo, 0.0f), 1.0f) * 65536.0f);
  const float fh = __builtin_ceilf(__builtin_fminf(__builtin_fmaxf(hi, 0.0f), 1.0f) * 65536.0f);
  const uint32_t l16 = fl >= 65535.0f ? 0xffffu : static_cast<uint32_t>(fl);
  const uint32_t h16 = fh < 1.0f ? 0u : (fh >= 65536.0f ? 0xffffu : static_cast<uint32_t>(fh) - 1u);
  return l16 | (h16 << 16);
}

// Specular reflection on an axis-aligned face: d - 2 (d.n) n rounds to
// exactly the negation of the approached component.
__device__ __forceinline__ void flip_axis(V3& d, int axis) {
  d.x = axis == 0 ? -d.x : d.x;
  d.y = axis == 1 ? -d.y : d.y;
  d.z = axis == 2 ? -d.z : d.z;
}

// Phase 3: the new direction after Rayleigh scattering (about the direction,
// cos from Cardano, sampling.hpp:39-49) or Lambertian reflection (about the
// inward normal -n of the face on `axis`, cos = sqrt(u), sampling.hpp:53-55),
// from the two words wc (cos) and wp (azimuth).
__device__ __forceinline__ V3 redirect(const DevParams& P, bool scatter, V3 d, int axis, bool plus,
                                       uint32_t wc, uint32_t wp) {
  float cos_theta;
  if (scatter)
    cos_theta = rayleigh_cos_word(P.rayleigh, wc);
  else
    cos_theta = sqrt_n(to_unit(wc));  // u in [2^-24, 1]
  const float s = plus ? 1.0f : -1.0f;
  // -n keeps the reference's signed zeros: -(+0) = -0 off-axis.
  const V3 inward{axis == 0 ? -s : -0.0f, axis == 1 ? -s : -0.0f, axis == 2 ? -s : -0.0f};
  return rotate_about(scatter ? d : inward, cos_theta, wp);
}

// One transport step (transport.hpp:71-145). Returns true while the photon
// is alive; otherwise *fate is set. The stream must hold the step's (up to 5)
// words: Stream after refill() (pos <= 3), RingStream with level() >= 5.
// kSurf >= 0 fixes the surface model at compile time (a launch whose events
// all share it); kSurf < 0 reads it from P. `fres` (forward steps only):
// Fresnel cell bounds, kFresnelCells words for the readout followed, when the
// side faces are polished, by kFresnelCells words for the outside medium; a
// decision they settle skips fresnel_prob.
template <int kSurf = -1, bool kAdj = false, class RNG>
__device__ __forceinline__ bool step(const DevParams& P, RNG& rng, Photon& ph, uint32_t& fate,
                                     StepInfo& info, AdjState* adj = nullptr,
                                     const uint32_t* fres = nullptr) {
  const uint32_t surface = kSurf >= 0 ? static_cast<uint32_t>(kSurf) : P.surface;
  const Advance a = advance(P, rng, ph);
  const bool absorb = a.absorb, scatter = a.scatter;
  const bool readout = a.boundary && a.face == kReadoutFace;
  const bool side = a.boundary && a.face != kReadoutFace;
  const bool polished = side && surface == kPolished;
  const bool black = side && surface == kBlack;
  const bool refl = side && surface >= kSpecular;
  const bool lambert = surface == kLambertian;

  // --- Fresnel, once for readout and polished lanes --------------------------
  float prob = P.reflectivity;
  bool scored = false;  // a Fresnel decision with a score (adjoint only)
  bool sure_r = false, sure_t = false;  // settled by the cell bounds
  if (!kAdj && fres != nullptr && (readout || polished)) {
    // cos_i in (0, 1]: cos_i * kFresnelCells is exact, its floor the cell
    const uint32_t c = static_cast<uint32_t>(a.cos_i * static_cast<float>(kFresnelCells));
    const uint32_t cell = (c < kFresnelCells ? c : kFresnelCells - 1u) + (readout ? 0u : kFresnelCells);
    const uint32_t b = fres[cell];
    const uint32_t k8 = rng.template peek<2>() >> 16;
    sure_r = k8 < (b & 0xffffu);
    sure_t = k8 > (b >> 16);
  }
  if ((readout || polished) && !(sure_r || sure_t))
    prob = fresnel_prob(P, readout, a.cos_i, scored);
  // --- the surface decision draw (u <= prob survives, transport.hpp:41) -----
  const bool decide = readout || polished || refl;
  const bool survived =
      decide && (sure_r || (!sure_t && to_unit(rng.template peek<2>()) <= prob));
  const bool lambert_reflect = refl && lambert && survived;
  if (kAdj && scored && (survived || !readout)) {
    // Bernoulli score: d log R (reflected) or d log (1 - R) (transmitted);
    // readout transmissions are skipped (their baseline term is zero).
    float d1, d2;
    fresnel_score(P.n_scint, readout ? P.n_det : P.n_out, a.cos_i, survived, d1, d2);
    adj->s_nscint += d1;
    if (readout) {
      adj->s_ndet += d2;
      const double b = adj->w[time_bin(ph.t, adj->time_scale, adj->time_bins)];
      adj->base_ndet += b * static_cast<double>(d2);
      adj->base_nscint += b * static_cast<double>(d1);
    }
  }
  if (kAdj && refl) ++adj->refl_hits;

  // --- new direction ----------------------------------------------------------
  if (survived && !lambert_reflect) flip_axis(ph.d, a.axis);
  if (scatter || lambert_reflect) {
    // Rayleigh: draws 2, 3. Lambertian: draws 3, 4.
    const uint32_t wc = scatter ? rng.template peek<2>() : rng.template peek<3>();
    const uint32_t wp = scatter ? rng.template peek<3>() : rng.template peek<4>();
    ph.d = redirect(P, scatter, ph.d, a.axis, a.plus, wc, wp);
  }
  rng.pos += 2u + (decide ? 1u : 0u) + ((scatter || lambert_reflect) ? 2u : 0u);

  // --- fate -----------------------------------------------------------------
  const bool alive = scatter || survived;
  if (!alive) {
    fate = absorb ? kAbsorbedBulk
                  : (readout ? kDetected : (polished ? kEscaped : kAbsorbedSurface));
  }
  info.absorb = absorb;
  info.scatter = scatter;
  info.readout = readout;
  info.polished = polished;
  info.black = black;
  info.refl = refl;
  info.survived = survived;
  info.lambert_reflect = lambert_reflect;
  info.abs_on = P.abs_length != kInf;
  info.scat_on = P.scat_length != kInf;
  return alive;
}

}  // namespace ophip
