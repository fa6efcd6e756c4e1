// mcl_device.h — device building blocks for the CDNA4 Monte Carlo kernels:
// Philox4x32-10 counter RNG, the library's own float32/float64 exp and expm1,
// the per-electron fire threshold, and wave64 reductions.
//
// Every float operation that feeds a recombination decision is written with an
// explicit rounding: the library is compiled with -ffp-contract=off, products
// that should fuse call __builtin_fma{,f}, and exp/expm1 are evaluated here
// (Cody-Waite reduction with ln2 split 12+24 bits in float32 / 21+53 bits in
// float64, Taylor polynomial in Horner form) rather than by the hardware
// v_exp_f32, whose last-bit behaviour no CPU reproduces.  A CPU evaluation of
// the same steps in the same order gives the same bits, which is what makes
// per-step replay bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcl {

constexpr int kWave = 64;
constexpr uint32_t kNoHole = 0xFFFFu;
constexpr uint32_t kDeadSlot = 0xFFFFFFFFu;
constexpr float kDeadHoleX = 3.0e38f;   // removed hole: x coordinate sentinel, d2 -> +inf

// ------------------------------------------------------------------ Philox --
constexpr uint32_t kPhM0 = 0xD2511F53u, kPhM1 = 0xCD9E8D57u;
constexpr uint32_t kPhW0 = 0x9E3779B9u, kPhW1 = 0xBB67AE85u;

struct u32x4 { uint32_t x, y, z, w; };

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The key is wave-uniform,
// so its bumps stay on the scalar unit; each round is two 32x32->64 multiplies
// (v_mad_u64_u32) and two three-way xors.
__device__ __forceinline__ u32x4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                          uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        if (i) { k0 += kPhW0; k1 += kPhW1; }
        const uint64_t p0 = (uint64_t)kPhM0 * c0;
        const uint64_t p1 = (uint64_t)kPhM1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    return u32x4{c0, c1, c2, c3};
}

__device__ __forceinline__ uint32_t pick_word(const u32x4& r, int word)
{
    return word == 0 ? r.x : word == 1 ? r.y : word == 2 ? r.z : r.w;
}

// ------------------------------------------------------- own exp / expm1 ----
constexpr float kFInvLn2 = 0x1.715476p+0f;
constexpr float kFLn2Hi = 0x1.62ep-1f;
constexpr float kFLn2Lo = 0x1.0bfbe8p-15f;

__device__ __forceinline__ float pow2i_f(int k)   // 2^k, -126 <= k <= 127
{
    return __int_as_float((k + 127) << 23);
}

// expm1(r) on |r| <= ln2/2: r + r^2 (1/2! + r/3! + ... + r^6/8!)
__device__ __forceinline__ float em1_core(float r)
{
    float p = 0x1.a01a02p-16f;
    p = __builtin_fmaf(p, r, 0x1.a01a02p-13f);
    p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
    p = __builtin_fmaf(p, r, 0x1.111112p-7f);
    p = __builtin_fmaf(p, r, 0x1.555556p-5f);
    p = __builtin_fmaf(p, r, 0x1.555556p-3f);
    p = __builtin_fmaf(p, r, 0x1p-1f);
    const float r2 = r * r;
    return __builtin_fmaf(r2, p, r);
}

__device__ __forceinline__ float expm1_own(float x)
{
    if (x != x) return x;
    if (x < -17.5f) return -1.0f;
    if (x > 88.0f) return __builtin_inff();
    const float kf = __builtin_rintf(x * kFInvLn2);
    float r = __builtin_fmaf(-kf, kFLn2Hi, x);
    r = __builtin_fmaf(-kf, kFLn2Lo, r);
    const float em1 = em1_core(r);
    if (kf == 0.0f) return em1;
    const float t = pow2i_f((int)kf);
    return __builtin_fmaf(t, em1, t - 1.0f);
}

__device__ __forceinline__ float exp_own(float x)
{
    if (x != x) return x;
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) return __builtin_inff();
    const float kf = __builtin_rintf(x * kFInvLn2);
    float r = __builtin_fmaf(-kf, kFLn2Hi, x);
    r = __builtin_fmaf(-kf, kFLn2Lo, r);
    const float e = 1.0f + em1_core(r);
    return e * pow2i_f((int)kf);
}

constexpr double kDInvLn2 = 0x1.71547652b82fep+0;
constexpr double kDLn2Hi = 0x1.62e42p-1;
constexpr double kDLn2Lo = 0x1.fdf473de6af28p-22;

__device__ __forceinline__ double pow2i_d(int k)  // 2^k, -1022 <= k <= 1023
{
    return __longlong_as_double((long long)(k + 1023) << 52);
}

__device__ __forceinline__ double exp_own(double x)
{
    if (x != x) return x;
    if (x < -708.0) return 0.0;
    if (x > 709.0) return __builtin_inf();
    const double kd = __builtin_rint(x * kDInvLn2);
    double r = __builtin_fma(-kd, kDLn2Hi, x);
    r = __builtin_fma(-kd, kDLn2Lo, r);
    double p = 0x1.6124613a86d09p-33;
    p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0x1p-1);
    p = __builtin_fma(p, r, 0x1p+0);
    const double e = __builtin_fma(p, r, 1.0);
    return e * pow2i_d((int)kd);
}

// Fire threshold for x = k dt: p = 1 - e^{-x}; an electron with draw u fires
// iff always || u < thr, thr = (uint32)(p * 2^32).
// float64 natural log for normal x > 0: x = 2^k m with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), by the odd series 2 s (1 + s^2/3 + ...
// + s^22/23) (|s| <= 0.1716), then k ln 2 with the two-part ln 2 of exp_own.
__device__ __forceinline__ double log_own(double x)
{
    if (x != x || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t b = (uint64_t)__double_as_longlong(x);
    int k = (int)((b >> 52) & 0x7ffu) - 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m = __longlong_as_double((long long)b);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m *= 0.5;
        k += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double q = s * s;
    double P = 1.0 / 23.0;
    P = __builtin_fma(P, q, 1.0 / 21.0);
    P = __builtin_fma(P, q, 1.0 / 19.0);
    P = __builtin_fma(P, q, 1.0 / 17.0);
    P = __builtin_fma(P, q, 1.0 / 15.0);
    P = __builtin_fma(P, q, 1.0 / 13.0);
    P = __builtin_fma(P, q, 1.0 / 11.0);
    P = __builtin_fma(P, q, 1.0 / 9.0);
    P = __builtin_fma(P, q, 1.0 / 7.0);
    P = __builtin_fma(P, q, 1.0 / 5.0);
    P = __builtin_fma(P, q, 1.0 / 3.0);
    const double t = 2.0 * s;
    const double lnm = __builtin_fma(t, q * P, t);
    const double kd = (double)k;
    return __builtin_fma(kd, kDLn2Hi, __builtin_fma(kd, kDLn2Lo, lnm));
}

struct FireThr { uint32_t thr; bool always; };

__device__ __forceinline__ FireThr fire_threshold(float x)
{
    const float p = -expm1_own(-x);
    FireThr t;
    t.always = p >= 1.0f;
    t.thr = t.always ? 0xFFFFFFFFu : (uint32_t)(p * 0x1p32f);
    return t;
}

// ---------------------------------------------------------- wave helpers ----
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// exclusive prefix count of set bits of mask below this lane
__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Value of lane l (l wave-uniform) in every lane: v_readlane, no LDS round trip.
__device__ __forceinline__ uint32_t lane_value(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), l));
}

// DPP controls (gfx9 encoding): quad permutations and the row mirrors.
constexpr int kDppQuad1032 = 0xB1;      // quad_perm [1, 0, 3, 2]
constexpr int kDppQuad2301 = 0x4E;      // quad_perm [2, 3, 0, 1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

template <int kCtrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, 0xF, 0xF, false);
}

// Minimum over the 64 lanes (all active), in every lane: four DPP steps give the
// minimum of each 16-lane row, four readlanes and scalar mins combine the rows.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    uint32_t u = dpp_u32<kDppQuad1032>(v);
    v = u < v ? u : v;
    u = dpp_u32<kDppQuad2301>(v);
    v = u < v ? u : v;
    u = dpp_u32<kDppRowHalfMirror>(v);
    v = u < v ? u : v;
    u = dpp_u32<kDppRowMirror>(v);
    v = u < v ? u : v;
    const uint32_t r0 = lane_value(v, 0), r1 = lane_value(v, 16);
    const uint32_t r2 = lane_value(v, 32), r3 = lane_value(v, 48);
    const uint32_t a = r0 < r1 ? r0 : r1, b = r2 < r3 ? r2 : r3;
    return a < b ? a : b;
}

// Minimum of a 64-bit key over the 64 lanes (all active): the high words first;
// the low word of the single lane holding that minimum, or the minimum low word
// among the tied lanes.
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
    const uint32_t hi = (uint32_t)(v >> 32), lo = (uint32_t)v;
    const uint32_t mh = wave_min_u32(hi);
    const uint64_t tied = __ballot(hi == mh);
    uint32_t ml;
    if ((tied & (tied - 1)) == 0) ml = lane_value(lo, (int)__builtin_ctzll(tied));
    else ml = wave_min_u32(hi == mh ? lo : 0xFFFFFFFFu);
    return ((uint64_t)mh << 32) | ml;
}

// Minimum of a 64-bit key over each 16-lane row (all lanes active), in every lane
// of the row: four DPP steps (pairs, quads, half rows, rows).
__device__ __forceinline__ uint64_t row_min_u64(uint64_t v)
{
    uint64_t u = ((uint64_t)dpp_u32<kDppQuad1032>((uint32_t)(v >> 32)) << 32) |
                 dpp_u32<kDppQuad1032>((uint32_t)v);
    v = u < v ? u : v;
    u = ((uint64_t)dpp_u32<kDppQuad2301>((uint32_t)(v >> 32)) << 32) | dpp_u32<kDppQuad2301>((uint32_t)v);
    v = u < v ? u : v;
    u = ((uint64_t)dpp_u32<kDppRowHalfMirror>((uint32_t)(v >> 32)) << 32) |
        dpp_u32<kDppRowHalfMirror>((uint32_t)v);
    v = u < v ? u : v;
    u = ((uint64_t)dpp_u32<kDppRowMirror>((uint32_t)(v >> 32)) << 32) | dpp_u32<kDppRowMirror>((uint32_t)v);
    return u < v ? u : v;
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), off);
        v |= ((uint64_t)hi << 32) | lo;
    }
    return v;
}

}  // namespace mcl
