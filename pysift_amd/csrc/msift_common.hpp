// Shared definitions for the pysift_amd HIP library (gfx950 / MI355X).
//
// Numerics contract (identical on host and device, compiled with
// -ffp-contract=off): every float/double operation is the IEEE operation as
// written; fmaf()/fma() appear only where an accumulation is defined as fused.
// exp/atan2/sincos/log2 are evaluated by the double-precision routines below
// (Cody-Waite reduction + Taylor / atanh series using + - * / floor only), so
// the GPU reproduces a CPU evaluation of the same expressions bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stddef.h>

#pragma clang fp contract(off)

#define MSIFT_HD __host__ __device__ __forceinline__

namespace msift {

constexpr int kLanes = 64;          // CDNA wavefront width
constexpr int kMaxOct = 16;
constexpr int kMaxR = 32;           // largest blur radius handled (sigma < 7.9)
constexpr int kUpsK = 30;           // spline prefilter half-length (|z|^31 < 2e-18)

// ---------------------------------------------------------------- math ----
MSIFT_HD double exp_d(double x) {
    if (x != x) return x;
    if (x > 709.0) return __builtin_huge_val();
    if (x < -708.0) return 0.0;
    const double INV_LN2 = 0x1.71547652b82fep+0;
    const double LN2_HI = 0x1.62e42fee00000p-1;
    const double LN2_LO = 0x1.a39ef35793c76p-33;
    double kd = floor(x * INV_LN2 + 0.5);
    int k = (int)kd;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    union { double d; uint64_t u; } v;
    v.u = (uint64_t)(k + 1023) << 52;
    return p * v.d;
}

MSIFT_HD double log2_d(double x) {
    union { double d; uint64_t u; } v;
    v.d = x;
    int e = (int)((v.u >> 52) & 0x7ff) - 1023;
    v.u = (v.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m = v.d;
    if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double t = 1.0 / 23.0;
    t = t * s2 + 1.0 / 21.0;
    t = t * s2 + 1.0 / 19.0;
    t = t * s2 + 1.0 / 17.0;
    t = t * s2 + 1.0 / 15.0;
    t = t * s2 + 1.0 / 13.0;
    t = t * s2 + 1.0 / 11.0;
    t = t * s2 + 1.0 / 9.0;
    t = t * s2 + 1.0 / 7.0;
    t = t * s2 + 1.0 / 5.0;
    t = t * s2 + 1.0 / 3.0;
    t = t * s2 + 1.0;
    double ln_m = 2.0 * (s * t);
    return (double)e + ln_m * 0x1.71547652b82fep+0;
}

MSIFT_HD double atan01_d(double t) {
    const double K8[9] = {
        0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2,
        0x1.dac670561bb4fp-2, 0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1,
        0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1 };
    int k = (int)floor(t * 8.0 + 0.5);
    double c = (double)k * 0.125;
    double u = (t - c) / (1.0 + t * c);
    double u2 = u * u;
    double q = -1.0 / 15.0;
    q = q * u2 + 1.0 / 13.0;
    q = q * u2 - 1.0 / 11.0;
    q = q * u2 + 1.0 / 9.0;
    q = q * u2 - 1.0 / 7.0;
    q = q * u2 + 1.0 / 5.0;
    q = q * u2 - 1.0 / 3.0;
    double at = u + u * (u2 * q);
    return K8[k] + at;
}

// atan2 for finite arguments (image gradients); signed zeros follow C atan2.
MSIFT_HD double atan2_d(double y, double x) {
    const double PI = 0x1.921fb54442d18p+1;
    const double PIO2 = 0x1.921fb54442d18p+0;
    bool ys = __builtin_signbit(y) != 0, xs = __builtin_signbit(x) != 0;
    double ax = fabs(x), ay = fabs(y);
    double a;
    if (ay == 0.0) {
        a = xs ? PI : 0.0;
    } else if (ax == 0.0) {
        a = PIO2;
    } else {
        if (ay > ax) a = PIO2 - atan01_d(ax / ay);
        else a = atan01_d(ay / ax);
        if (xs) a = PI - a;
    }
    return ys ? -a : a;
}

MSIFT_HD void sincos_d(double x, double* s, double* c) {
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double PIO2_1 = 0x1.921fb54400000p+0;
    const double PIO2_1T = 0x1.0b4611a626331p-34;
    double kd = floor(x * TWO_OVER_PI + 0.5);
    int q = ((int)kd) & 3;
    double r = (x - kd * PIO2_1) - kd * PIO2_1T;
    double r2 = r * r;
    double sp = -1.0 / 1307674368000.0;
    sp = sp * r2 + 1.0 / 6227020800.0;
    sp = sp * r2 - 1.0 / 39916800.0;
    sp = sp * r2 + 1.0 / 362880.0;
    sp = sp * r2 - 1.0 / 5040.0;
    sp = sp * r2 + 1.0 / 120.0;
    sp = sp * r2 - 1.0 / 6.0;
    double sn = r + r * (r2 * sp);
    double cp = 1.0 / 20922789888000.0;
    cp = cp * r2 - 1.0 / 87178291200.0;
    cp = cp * r2 + 1.0 / 479001600.0;
    cp = cp * r2 - 1.0 / 3628800.0;
    cp = cp * r2 + 1.0 / 40320.0;
    cp = cp * r2 - 1.0 / 720.0;
    cp = cp * r2 + 1.0 / 24.0;
    cp = cp * r2 - 0.5;
    double cs = 1.0 + r2 * cp;
    switch (q) {
    case 0: *s = sn;  *c = cs;  break;
    case 1: *s = cs;  *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

// Single-precision exp and atan2 of the orientation / descriptor weights.
// Plain IEEE float operations with explicit fmaf only, so the CPU oracle
// reproduces every bit.  exp: x = n ln2 + r (Cody-Waite, ln2 split 16+24
// bits, n ln2_hi exact), e^r by its degree-7 Taylor polynomial (Horner),
// times 2^n; <= 1 ulp on [-87, 88], 0 below -87.  atan2: t = min/max (IEEE
// division), atan(t) = t P(t^2) with a degree-9 polynomial fitted to
// atan(t)/t on [0, 1] (iteratively reweighted least squares on relative
// error, coefficients rounded to float), then the octant fix-ups; <= 2 ulp.
MSIFT_HD float exp_f(float x) {
    if (x != x) return x;
    if (x > 88.0f) return __builtin_huge_valf();
    if (x < -87.0f) return 0.0f;
    const float n = floorf(x * 0x1.715476p+0f + 0.5f);
    float r = fmaf(n, -0x1.62e4p-1f, x);
    r = fmaf(n, -0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;
    p = fmaf(p, r, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    union { float f; uint32_t u; } v;
    v.u = (uint32_t)((int)n + 127) << 23;
    return p * v.f;
}

MSIFT_HD float atan2_f(float y, float x) {
    const float PI = 0x1.921fb6p+1f, PIO2 = 0x1.921fb6p+0f;
    if (x != x || y != y) return x + y;
    const bool ys = signbit(y), xs = signbit(x);
    const float ax = fabsf(x), ay = fabsf(y);
    float a;
    if (ay == 0.0f) {
        a = xs ? PI : 0.0f;
    } else if (ax == 0.0f) {
        a = PIO2;
    } else if (isinf(ax) || isinf(ay)) {
        if (isinf(ax) && isinf(ay)) a = xs ? 0x1.2d97c8p+1f : 0x1.921fb6p-1f;
        else if (isinf(ax)) a = xs ? PI : 0.0f;
        else a = PIO2;
    } else {
        const float t = (ay > ax ? ax : ay) / (ay > ax ? ay : ax);
        const float s2 = t * t;
        float p = -0x1.cbd234p-10f;
        p = fmaf(p, s2, 0x1.5f85d2p-7f);
        p = fmaf(p, s2, -0x1.f8b94ep-6f);
        p = fmaf(p, s2, 0x1.d77a5cp-5f);
        p = fmaf(p, s2, -0x1.572366p-4f);
        p = fmaf(p, s2, 0x1.c02fd6p-4f);
        p = fmaf(p, s2, -0x1.2415a4p-3f);
        p = fmaf(p, s2, 0x1.999084p-3f);
        p = fmaf(p, s2, -0x1.555534p-2f);
        p = fmaf(p, s2, 1.0f);
        a = t * p;
        if (ay > ax) a = PIO2 - a;
        if (xs) a = PI - a;
    }
    return ys ? -a : a;
}

// The same operations as exp_f and atan2_f without their special-case
// branches, for the per-pixel loops of detection, whose inputs come from the
// level values of a u8 image (finite): exp_f_le88 for x <= 88 that is not
// NaN (the orientation and descriptor weights take x <= 0), where x < -87
// becomes a select of 0; atan2_f_finite for finite y, x, where the zero
// cases become selects.  Identical bits to the general forms on those
// inputs; the branch-free form keeps a wave's pixels in one basic block.
// exp_f for -87 <= x <= 88, where exp_f takes neither special case: the
// descriptor's Gaussian weight, whose argument -(xr^2 + yr^2) / (2 sd^2)
// stays above -3.2 over the patch (|xr|, |yr| <= ph sqrt 2, sd = 0.5657 ph)
__device__ __forceinline__ float exp_f_mid(float x) {
    const float n = floorf(x * 0x1.715476p+0f + 0.5f);
    float r = fmaf(n, -0x1.62e4p-1f, x);
    r = fmaf(n, -0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;
    p = fmaf(p, r, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    // p 2^n by the exponent add, the same value as the multiply by 2^n (normal here)
    return __builtin_ldexpf(p, (int)n);
}

__device__ __forceinline__ float exp_f_le88(float x) {
    const float n = floorf(x * 0x1.715476p+0f + 0.5f);
    float r = fmaf(n, -0x1.62e4p-1f, x);
    r = fmaf(n, -0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;
    p = fmaf(p, r, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    const float v = p * __builtin_bit_cast(float, (uint32_t)((int)n + 127) << 23);
    return x < -87.0f ? 0.0f : v;
}

// fmod360 for |x| < 720: at most one exact subtraction
__device__ __forceinline__ float fmod360_lt720(float x) {
    float a = fabsf(x);
    a = a >= 360.0f ? a - 360.0f : a;
    return copysignf(a, x);
}

__device__ __forceinline__ float atan2_f_finite(float y, float x) {
    const float PI = 0x1.921fb6p+1f, PIO2 = 0x1.921fb6p+0f;
    const bool ys = signbit(y), xs = signbit(x);
    const float ax = fabsf(x), ay = fabsf(y);
    const bool swap = ay > ax;
    const float t = (swap ? ax : ay) / (swap ? ay : ax);   // 0 / 0 only where a select below applies
    const float s2 = t * t;
    float p = -0x1.cbd234p-10f;
    p = fmaf(p, s2, 0x1.5f85d2p-7f);
    p = fmaf(p, s2, -0x1.f8b94ep-6f);
    p = fmaf(p, s2, 0x1.d77a5cp-5f);
    p = fmaf(p, s2, -0x1.572366p-4f);
    p = fmaf(p, s2, 0x1.c02fd6p-4f);
    p = fmaf(p, s2, -0x1.2415a4p-3f);
    p = fmaf(p, s2, 0x1.999084p-3f);
    p = fmaf(p, s2, -0x1.555534p-2f);
    p = fmaf(p, s2, 1.0f);
    float a = t * p;
    a = swap ? PIO2 - a : a;
    a = xs ? PI - a : a;
    // y != 0, x == 0 needs no select: t = 0 / |y| = 0, a = PIO2 - 0, and
    // PI - PIO2 == PIO2 in float for x == -0
    a = ay == 0.0f ? (xs ? PI : 0.0f) : a;        // y == 0 (takes precedence, as in atan2_f)
    return ys ? -a : a;
}

// C fmodf(x, 360) for |x| < 1440 (every subtraction is exact): the same
// subtractions as a loop while |x| >= 360, unrolled to three selects
// x rounded to binary16 (nearest, ties to even) and back: the value a
// float16 array stores (core.py:1244, 1308)
__device__ __forceinline__ float f16_round(float x) { return __half2float(__float2half_rn(x)); }

MSIFT_HD float fmod360(float x) {
    float a = fabsf(x);
    a = a >= 360.0f ? a - 360.0f : a;
    a = a >= 360.0f ? a - 360.0f : a;
    a = a >= 360.0f ? a - 360.0f : a;
    return copysignf(a, x);
}

// scipy 'reflect' (half-sample symmetric, periodic in 2N): core.py:477-483
MSIFT_HD int reflect_idx(int i, int N) {
    int p = 2 * N;
    i = i % p;
    if (i < 0) i += p;
    if (i >= N) i = 2 * N - 1 - i;
    return i;
}

// scipy 'mirror' (whole-sample symmetric, periodic in 2N-2)
MSIFT_HD int mirror_idx(int i, int N) {
    if (N == 1) return 0;
    int p = 2 * N - 2;
    if (i < 0) i = -i;
    i = i % p;
    return i >= N ? p - i : i;
}

// ------------------------------------------------------------ structs ----
struct Taps {
    float w[2 * kMaxR + 1];
    int r;
};

struct UpsTaps {
    double h[kUpsK + 1];
};

struct Kp {            // refined keypoint in octave coordinates
    float ri, ci, resp;
    int o_si;          // octave << 8 | scale index
};

struct OctInfo {
    int n_oct, S, border;
    int H[kMaxOct], W[kMaxOct];
    float* G[kMaxOct];   // (S+3) levels
    int blk0[kMaxOct + 1];  // first extrema block of each octave
    int nseg[kMaxOct];      // 256-column segments per row
};

// ---------------------------------------------------------- wave utils ----
__device__ __forceinline__ float wave_tree_f(float v) {
    // xor butterfly: every lane ends with the value the oracle's tree gives
    v = v + __shfl_xor(v, 32);
    v = v + __shfl_xor(v, 16);
    v = v + __shfl_xor(v, 8);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 1);
    return v;
}

__device__ __forceinline__ double wave_tree_d(double v) {
    v = v + __shfl_xor(v, 32);
    v = v + __shfl_xor(v, 16);
    v = v + __shfl_xor(v, 8);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 1);
    return v;
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() is a
// release/acquire fence on every address space, so it also waits for all of the
// wave's outstanding global loads and stores (vmcnt(0)) before the barrier; a
// loop that keeps global loads in flight across barriers (prefetch) or stores
// results it never re-reads uses this instead.  Global memory written before it
// is NOT visible to the other waves after it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace msift
