/*
 * CPU oracle (test infrastructure, never shipped as product code) for the
 * plain detect -> describe -> match -> RANSAC -> warp/feather path of
 * PySIFT v0.1.4.  Reference citations are pysift/core.py:<line>.
 *
 * Numerics pysift_amd defines where the reference leaves them open; the GPU
 * library must reproduce every one of them bit for bit:
 *   - all float arithmetic is IEEE single, no FMA contraction (-ffp-contract=off);
 *     explicit fmaf()/fma() where an accumulation is specified as fused;
 *   - sincos/log2 are this file's own double-precision routines (Cody-Waite
 *     reduction + Taylor/atanh series), rounded to float at the use; the
 *     per-pixel exp and atan2 of the orientation and descriptor histograms are
 *     its own single-precision routines (or_exp_f, or_atan2_f);
 *   - the orientation histogram accumulates in 64 lane-private partials (pixel
 *     p goes to partial p % 64, in increasing p), then each bin combines its
 *     64 partials with the 64-lane xor tree of wave_tree_f();
 *   - the descriptor histogram sums its contributions as integers: each,
 *     scaled by 2^30 (after clamping the pixel weight below 2^32), is
 *     truncated to an integer (desc_fx), the 64-bit sums are exact and so
 *     independent of order, and bin = (float)(sum * 2^-30);
 *   - 128-bin and per-point sums use the 64-lane xor tree of wave_tree_*();
 *   - keypoint refinement and every homography solve run in FP64.
 */
#include "sift_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LANES 64

/* ========================================================================
 * Elementary functions.  Only +,-,*,/, floor and exact bit manipulation.
 * ======================================================================== */
static double ldexp_exact(double x, int k)
{
    /* x * 2^k for k in [-1022, 1023]: multiply by an exactly built power of two */
    union { double d; uint64_t u; } v;
    v.u = (uint64_t)(k + 1023) << 52;
    return x * v.d;
}

double or_exp(double x)
{
    if (x != x) return x;
    if (x > 709.0) return HUGE_VAL;
    if (x < -708.0) return 0.0;                       /* flush: defined, not denormal */
    const double INV_LN2 = 0x1.71547652b82fep+0;
    const double LN2_HI = 0x1.62e42fee00000p-1;       /* 32 trailing zero bits */
    const double LN2_LO = 0x1.a39ef35793c76p-33;
    double kd = floor(x * INV_LN2 + 0.5);
    int k = (int)kd;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;       /* |r| <= 0.3466 */
    double p = 1.0 / 6227020800.0;                    /* 1/13! */
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
    return ldexp_exact(p, k);
}

double or_log2(double x)
{
    /* x > 0, finite, normal */
    union { double d; uint64_t u; } v;
    v.d = x;
    int e = (int)((v.u >> 52) & 0x7ff) - 1023;
    v.u = (v.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;   /* m in [1,2) */
    double m = v.d;
    if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }          /* m in [0.707,1.414] */
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

static const double ATAN_K8[9] = {
    0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2,
    0x1.dac670561bb4fp-2, 0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1,
    0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1 };

static double atan01(double t)
{
    /* t in [0,1]: atan(t) = atan(k/8) + atan((t - k/8) / (1 + t k/8)) */
    int k = (int)floor(t * 8.0 + 0.5);
    double c = (double)k * 0.125;
    double u = (t - c) / (1.0 + t * c);               /* |u| <= 0.0625 */
    double u2 = u * u;
    double q = -1.0 / 15.0;
    q = q * u2 + 1.0 / 13.0;
    q = q * u2 - 1.0 / 11.0;
    q = q * u2 + 1.0 / 9.0;
    q = q * u2 - 1.0 / 7.0;
    q = q * u2 + 1.0 / 5.0;
    q = q * u2 - 1.0 / 3.0;
    double at = u + u * (u2 * q);
    return ATAN_K8[k] + at;
}

double or_atan2(double y, double x)
{
    const double PI = 0x1.921fb54442d18p+1;
    const double PIO2 = 0x1.921fb54442d18p+0;
    if (x != x || y != y) return x + y;
    int ys = signbit(y) != 0, xs = signbit(x) != 0;
    double ax = fabs(x), ay = fabs(y);
    double a;
    if (ay == 0.0) {
        a = xs ? PI : 0.0;
    } else if (ax == 0.0) {
        a = PIO2;
    } else if (isinf(ax) || isinf(ay)) {
        if (isinf(ax) && isinf(ay)) a = xs ? 3.0 * 0x1.921fb54442d18p-1 : 0x1.921fb54442d18p-1;
        else if (isinf(ax)) a = xs ? PI : 0.0;
        else a = PIO2;
    } else {
        if (ay > ax) a = PIO2 - atan01(ax / ay);
        else a = atan01(ay / ax);
        if (xs) a = PI - a;
    }
    return ys ? -a : a;
}

void or_sincos(double x, double* s, double* c)
{
    /* |x| < 1e5; Cody-Waite pi/2 split with 33-bit high part */
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double PIO2_1 = 0x1.921fb54400000p+0;
    const double PIO2_1T = 0x1.0b4611a626331p-34;
    double kd = floor(x * TWO_OVER_PI + 0.5);
    int q = ((int)kd) & 3;
    double r = (x - kd * PIO2_1) - kd * PIO2_1T;     /* |r| <= pi/4 + eps */
    double r2 = r * r;
    double sp = -1.0 / 1307674368000.0;               /* -1/15! */
    sp = sp * r2 + 1.0 / 6227020800.0;
    sp = sp * r2 - 1.0 / 39916800.0;
    sp = sp * r2 + 1.0 / 362880.0;
    sp = sp * r2 - 1.0 / 5040.0;
    sp = sp * r2 + 1.0 / 120.0;
    sp = sp * r2 - 1.0 / 6.0;
    double sn = r + r * (r2 * sp);
    double cp = 1.0 / 20922789888000.0;               /* 1/16! */
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

/* Single-precision exp / atan2 for the orientation and descriptor weights,
 * restating atan2f / __expf / expf of the reference's CUDA kernels
 * (core.py:272-273 orientation, 758-760 descriptor) as a fixed sequence of IEEE float operations with explicit fmaf,
 * so that the device code can reproduce every bit.
 * exp: x = n ln2 + r, ln2 = 0x1.62e4p-1 + 0x1.7f7d1cp-20 (n ln2_hi exact for
 * |n| < 256), e^r = degree-7 Taylor polynomial, times 2^n (<= 1 ulp on
 * [-87, 88]; 0 below -87).
 * atan2: t = min(|x|,|y|)/max(|x|,|y|) (IEEE division), atan(t) = t P(t^2),
 * P of degree 9 fitted to atan(t)/t on [0, 1] by iteratively reweighted
 * least squares on the relative error (coefficients rounded to float); then
 * pi/2 - a, pi - a and the sign (<= 2 ulp).  Pinned by test_oracle against
 * numpy's float64 exp / arctan2 on dense sweeps. */
float or_exp_f(float x)
{
    if (x != x) return x;
    if (x > 88.0f) return HUGE_VALF;
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
    return ldexpf(p, (int)n);
}

float or_atan2_f(float y, float x)
{
    const float PI = 0x1.921fb6p+1f, PIO2 = 0x1.921fb6p+0f;
    if (x != x || y != y) return x + y;
    int ys = signbit(y) != 0, xs = signbit(x) != 0;
    float ax = fabsf(x), ay = fabsf(y), a;
    if (ay == 0.0f) {
        a = xs ? PI : 0.0f;
    } else if (ax == 0.0f) {
        a = PIO2;
    } else if (isinf(ax) || isinf(ay)) {
        if (isinf(ax) && isinf(ay)) a = xs ? 0x1.2d97c8p+1f : 0x1.921fb6p-1f;
        else if (isinf(ax)) a = xs ? PI : 0.0f;
        else a = PIO2;
    } else {
        float t = (ay > ax ? ax : ay) / (ay > ax ? ay : ax);
        float s2 = t * t;
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

/* x rounded to IEEE binary16 (round to nearest, ties to even; subnormals
 * kept; beyond the largest half, infinity) and back to float: the value a
 * float16 array stores (core.py:1244, 1308) */
float or_f16_round(float x)
{
    union { float f; uint32_t u; } v;
    v.f = x;
    const uint32_t sign = v.u & 0x80000000u, a = v.u & 0x7fffffffu;
    if (a >= 0x7f800000u) return x;
    union { uint32_t u; float f; } m;
    m.u = a;
    if (m.f >= 65520.0f) { v.u = sign | 0x7f800000u; return v.f; }
    if (m.f < 0x1p-14f) {
        /* half subnormals are the multiples of 2^-24 */
        m.f = rintf(m.f * 0x1p24f) * 0x1p-24f;
        v.u = sign | m.u;
        return v.f;
    }
    /* 23 -> 10 mantissa bits, ties to even; a carry moves into the exponent */
    const uint32_t r = (a + 0x0fffu + ((a >> 13) & 1u)) & ~0x1fffu;
    v.u = sign | r;
    return v.f;
}

void or_f16_round_arr(const float* x, float* y, int n)
{
    for (int i = 0; i < n; ++i) y[i] = or_f16_round(x[i]);
}

void or_exp_f_arr(const float* x, float* y, int n)
{
    for (int i = 0; i < n; ++i) y[i] = or_exp_f(x[i]);
}

void or_atan2_f_arr(const float* y, const float* x, float* a, int n)
{
    for (int i = 0; i < n; ++i) a[i] = or_atan2_f(y[i], x[i]);
}

static float f_exp(float x) { return or_exp_f(x); }
static float f_atan2(float y, float x) { return or_atan2_f(y, x); }

/* fmodf(x, 360) for |x| < 1440, exact (each subtraction is exact) */
static float fmod360(float x)
{
    float a = fabsf(x);
    while (a >= 360.0f) a = a - 360.0f;
    return copysignf(a, x);
}

/* 64-lane xor tree: partial[l] for l<64 -> total, combining l and l+off for
 * off = 32,16,8,4,2,1 (the lane-0 value of a __shfl_xor butterfly). */
static float wave_tree_f(float* p)
{
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; ++l) p[l] = p[l] + p[l + off];
    return p[0];
}
static double wave_tree_d(double* p)
{
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; ++l) p[l] = p[l] + p[l + off];
    return p[0];
}

/* ========================================================================
 * A2: separable Gaussian blur (core.py:475-608)
 * ======================================================================== */
int or_radius(double sigma) { return (int)(4.0 * sigma + 0.5); }

void or_taps(double sigma, int r, float* taps)
{
    /* core.py:564-567 (fp32 exp of -0.5(x/sigma)^2, normalised); computed in
     * double here and rounded once. */
    double t[129], sum = 0.0;
    for (int i = -r; i <= r; ++i) {
        double q = (double)i / sigma;
        t[i + r] = exp(-0.5 * (q * q));
    }
    for (int i = 0; i <= 2 * r; ++i) sum += t[i];
    for (int i = 0; i <= 2 * r; ++i) taps[i] = (float)(t[i] / sum);
}

static int reflect_idx(int i, int N)
{
    /* core.py:477-483: half-sample symmetric, periodic in 2N */
    int p = 2 * N;
    i = i % p;
    if (i < 0) i += p;
    if (i >= N) i = 2 * N - 1 - i;
    return i;
}

void or_blur(const float* in, float* out, int H, int W, double sigma)
{
    if (sigma < 0.01) { memcpy(out, in, sizeof(float) * (size_t)H * W); return; }
    int r = or_radius(sigma);
    float taps[129];
    or_taps(sigma, r, taps);
    float* tmp = (float*)malloc(sizeof(float) * (size_t)H * W);
    /* horizontal pass (core.py:486-509): sum = fma(tile[c-r+k], kern[k], sum), k ascending */
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        const float* row = in + (size_t)y * W;
        for (int x = 0; x < W; ++x) {
            float s = 0.0f;
            for (int k = 0; k <= 2 * r; ++k) s = fmaf(row[reflect_idx(x - r + k, W)], taps[k], s);
            tmp[(size_t)y * W + x] = s;
        }
    }
    /* vertical pass (core.py:519-547) */
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            float s = 0.0f;
            for (int k = 0; k <= 2 * r; ++k) s = fmaf(tmp[(size_t)reflect_idx(y - r + k, H) * W + x], taps[k], s);
            out[(size_t)y * W + x] = s;
        }
    }
    free(tmp);
}

/* ========================================================================
 * A3: cubic B-spline x2 zoom, scipy/cupyx zoom(order=3, mode='constant',
 * grid_mode=False) semantics (core.py:1274).  The prefilter is the mirror-
 * boundary inverse B-spline filter written as its symmetric impulse response
 * sqrt(3) z^|k| (z = sqrt(3)-2) truncated at |k| <= 30 (|z|^31 < 2e-18).
 * ======================================================================== */
#define UPS_K 30
static int mirror_idx(int i, int N)
{
    if (N == 1) return 0;
    int p = 2 * N - 2;
    if (i < 0) i = -i;
    i = i % p;
    return i >= N ? p - i : i;
}

static void ups_taps(double* h)
{
    h[0] = 0x1.bb67ae8584caap+0;                    /* sqrt(3) */
    for (int k = 1; k <= UPS_K; ++k) h[k] = h[k - 1] * -0x1.126145e9ecd58p-2;
}

static void bspline_w(double t, double* w)
{
    double s = 1.0 - t;
    double t2 = t * t, t3 = t2 * t, s2 = s * s, s3 = s2 * s;
    w[0] = s3 * (1.0 / 6.0);
    w[1] = (2.0 / 3.0 - t2) + 0.5 * t3;
    w[2] = (2.0 / 3.0 - s2) + 0.5 * s3;
    w[3] = t3 * (1.0 / 6.0);
}

void or_upsample2x(const float* in, int H, int W, float* out)
{
    double h[UPS_K + 1];
    ups_taps(h);
    double* t0 = (double*)malloc(sizeof(double) * (size_t)H * W);
    double* cf = (double*)malloc(sizeof(double) * (size_t)H * W);
    /* axis 0 (down columns) */
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double acc = 0.0;
            for (int k = -UPS_K; k <= UPS_K; ++k)
                acc = fma(h[k < 0 ? -k : k], (double)in[(size_t)mirror_idx(y + k, H) * W + x], acc);
            t0[(size_t)y * W + x] = acc;
        }
    /* axis 1 (along rows) */
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double acc = 0.0;
            for (int k = -UPS_K; k <= UPS_K; ++k)
                acc = fma(h[k < 0 ? -k : k], t0[(size_t)y * W + mirror_idx(x + k, W)], acc);
            cf[(size_t)y * W + x] = acc;
        }
    int Ho = 2 * H, Wo = 2 * W;
    double zy = (Ho > 1) ? (double)(H - 1) / (double)(Ho - 1) : 0.0;
    double zx = (Wo > 1) ? (double)(W - 1) / (double)(Wo - 1) : 0.0;
    #pragma omp parallel for schedule(static)
    for (int oy = 0; oy < Ho; ++oy) {
        double yy = (double)oy * zy;
        double fy = floor(yy);
        int iy = (int)fy;
        double wy[4];
        bspline_w(yy - fy, wy);
        for (int ox = 0; ox < Wo; ++ox) {
            double xx = (double)ox * zx;
            double fx = floor(xx);
            int ix = (int)fx;
            double wx[4];
            bspline_w(xx - fx, wx);
            double v = 0.0;
            for (int a = 0; a < 4; ++a) {
                const double* row = cf + (size_t)mirror_idx(iy - 1 + a, H) * W;
                double rv = 0.0;
                for (int b = 0; b < 4; ++b) rv = fma(wx[b], row[mirror_idx(ix - 1 + b, W)], rv);
                v = fma(wy[a], rv, v);
            }
            out[(size_t)oy * Wo + ox] = (float)v;
        }
    }
    free(t0);
    free(cf);
}

void or_u8_to_unit(const uint8_t* in, int H, int W, int pitch, float* out)
{
    /* core.py:1258: asarray(u8, float32) / float32(255) */
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) out[(size_t)y * W + x] = (float)in[(size_t)y * pitch + x] / 255.0f;
}

/* ========================================================================
 * A1: adaptive contrast threshold (core.py:1129-1148), entropy in FP64 with
 * bins summed in index order.
 * ======================================================================== */
double or_adaptive_thresh(const uint8_t* img, int H, int W, int pitch, double base)
{
    long long hist[256];
    memset(hist, 0, sizeof(hist));
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) hist[img[(size_t)y * pitch + x]]++;
    double total = (double)H * (double)W;
    /* lane l sums bins l, l+64, l+128, l+192; lanes combine by the xor tree */
    double part[LANES];
    for (int l = 0; l < LANES; ++l) {
        double acc = 0.0;
        for (int j = 0; j < 4; ++j) {
            int b = l + 64 * j;
            if (hist[b] == 0) continue;
            double p = (double)hist[b] / total;
            acc = acc - p * or_log2(p);
        }
        part[l] = acc;
    }
    double ent = wave_tree_d(part);
    double scale = 0.8 + ent / 40.0;
    if (scale > 1.0) scale = 1.0;
    return base * scale;
}

/* ========================================================================
 * Pyramid (A3/A4, core.py:1234-1365)
 * ======================================================================== */
typedef struct {
    int n_oct, S, doubled;
    int Ho[16], Wo[16];
    float* G[16];   /* (S+3) * Ho * Wo */
    float* D[16];   /* (S+2) * Ho * Wo */
} pyramid_t;

static void pyr_free(pyramid_t* P)
{
    for (int o = 0; o < P->n_oct; ++o) { free(P->G[o]); free(P->D[o]); }
}

static void sigma_schedule(const or_params* p, int doubled, double* sd0, double* inc)
{
    int S = p->n_scales;
    double k = pow(2.0, 1.0 / (double)S);
    double sig[64];
    for (int s = 0; s < S + 3; ++s) sig[s] = p->sigma0 * pow(k, (double)s);
    double b = doubled ? 1.0 : 0.5;
    double v = p->sigma0 * p->sigma0 - b * b;
    *sd0 = sqrt(v > 1e-6 ? v : 1e-6);
    for (int s = 1; s < S + 3; ++s) {
        double q = sig[s] * sig[s] - sig[s - 1] * sig[s - 1];
        inc[s - 1] = sqrt(q > 1e-6 ? q : 1e-6);
    }
}

static int build_pyramid(const uint8_t* img, int H, int W, int pitch, const or_params* p, pyramid_t* P)
{
    int S = p->n_scales;
    P->S = S;
    P->n_oct = p->n_octaves;
    P->doubled = p->double_image && ((long long)H * W < (long long)p->double_px_cap);
    float* unit = (float*)malloc(sizeof(float) * (size_t)H * W);
    or_u8_to_unit(img, H, W, pitch, unit);
    /* fp16 mode: the u8 / 255 base is a float16 array (core.py:1256), and the
     * zoomed base is cast back to float16 (core.py:1274) */
    if (p->fp16_pyramid)
        for (size_t i = 0; i < (size_t)H * W; ++i) unit[i] = or_f16_round(unit[i]);
    int bh = H, bw = W;
    float* base = unit;
    if (P->doubled) {
        base = (float*)malloc(sizeof(float) * (size_t)4 * H * W);
        or_upsample2x(unit, H, W, base);
        free(unit);
        bh = 2 * H; bw = 2 * W;
        if (p->fp16_pyramid)
            for (size_t i = 0; i < (size_t)bh * bw; ++i) base[i] = or_f16_round(base[i]);
    }
    double sd0, inc[64];
    sigma_schedule(p, P->doubled, &sd0, inc);
    float* b0 = (float*)malloc(sizeof(float) * (size_t)bh * bw);
    or_blur(base, b0, bh, bw, sd0);
    free(base);
    for (int o = 0; o < P->n_oct; ++o) {
        size_t n = (size_t)bh * bw;
        P->Ho[o] = bh; P->Wo[o] = bw;
        P->G[o] = (float*)malloc(sizeof(float) * n * (S + 3));
        P->D[o] = (float*)malloc(sizeof(float) * n * (S + 2));
        memcpy(P->G[o], b0, sizeof(float) * n);
        for (int s = 0; s < S + 2; ++s) {
            float* src = P->G[o] + n * s;
            float* dst = P->G[o] + n * (s + 1);
            or_blur(src, dst, bh, bw, inc[s]);
        }
        /* next base = level S decimated [::2, ::2] (core.py:1350), from the
         * float32 chain */
        int nh = (bh + 1) / 2, nw = (bw + 1) / 2;
        float* nb = (float*)malloc(sizeof(float) * (size_t)nh * nw);
        const float* gs = P->G[o] + n * S;
        for (int y = 0; y < nh; ++y)
            for (int x = 0; x < nw; ++x) nb[(size_t)y * nw + x] = gs[(size_t)(2 * y) * bw + 2 * x];
        /* fp16 mode: octaves >= 2 are stored as float16 (core.py:1306-1308);
         * the DoG, orientation and descriptor read the stored levels, the
         * blur chain does not (core.py:1313-1318) */
        if (p->fp16_pyramid && o >= 2)
            for (size_t i = 0; i < n * (S + 3); ++i) P->G[o][i] = or_f16_round(P->G[o][i]);
        for (int s = 0; s < S + 2; ++s) {
            const float* src = P->G[o] + n * s;
            const float* dst = P->G[o] + n * (s + 1);
            float* d = P->D[o] + n * s;
            for (size_t i = 0; i < n; ++i) d[i] = dst[i] - src[i];
        }
        free(b0);
        b0 = nb; bh = nh; bw = nw;
    }
    free(b0);
    return 0;
}

int or_pyramid_octave(const uint8_t* img, int H, int W, int pitch, const or_params* p,
                      int which, int* Ho, int* Wo, float* gauss, float* dog)
{
    pyramid_t P;
    memset(&P, 0, sizeof(P));
    build_pyramid(img, H, W, pitch, p, &P);
    if (which < 0 || which >= P.n_oct) { pyr_free(&P); return -1; }
    size_t n = (size_t)P.Ho[which] * P.Wo[which];
    *Ho = P.Ho[which]; *Wo = P.Wo[which];
    if (gauss) memcpy(gauss, P.G[which], sizeof(float) * n * (P.S + 3));
    if (dog) memcpy(dog, P.D[which], sizeof(float) * n * (P.S + 2));
    pyr_free(&P);
    return 0;
}

/* ========================================================================
 * A5: 3x3x3 extrema (core.py:176-217, 1367-1460), raster order (s,r,c) per
 * octave, first `cap` kept per octave.
 * ======================================================================== */
typedef struct { int o, s, r, c; } cand_t;
typedef struct { int o, si; float ri, ci, resp; } kpt_t;

static int find_extrema(const pyramid_t* P, float thr, int border, int cap, cand_t** out)
{
    int total = 0, capacity = 1024;
    cand_t* c = (cand_t*)malloc(sizeof(cand_t) * capacity);
    int S2 = P->S + 2;
    for (int o = 0; o < P->n_oct; ++o) {
        int H = P->Ho[o], W = P->Wo[o];
        if (H < 3 || W < 3) continue;
        size_t HW = (size_t)H * W;
        const float* D = P->D[o];
        int cnt = 0;
        for (int s = 1; s < S2 - 1; ++s)
            for (int r = border; r < H - border; ++r)
                for (int col = border; col < W - border; ++col) {
                    float v = D[s * HW + (size_t)r * W + col];
                    if (fabsf(v) <= thr) continue;
                    int is_max = 1, is_min = 1;
                    for (int ds = -1; ds <= 1 && (is_max || is_min); ++ds)
                        for (int dr = -1; dr <= 1 && (is_max || is_min); ++dr)
                            for (int dc = -1; dc <= 1; ++dc) {
                                if (!ds && !dr && !dc) continue;
                                float nb = D[(s + ds) * HW + (size_t)(r + dr) * W + col + dc];
                                is_max = is_max && (v >= nb);
                                is_min = is_min && (v <= nb);
                                if (!is_max && !is_min) break;
                            }
                    if (is_max || is_min) {
                        if (cnt < cap) {
                            if (total == capacity) { capacity *= 2; c = (cand_t*)realloc(c, sizeof(cand_t) * capacity); }
                            c[total].o = o; c[total].s = s; c[total].r = r; c[total].c = col;
                            total++;
                        }
                        cnt++;
                    }
                }
    }
    *out = c;
    return total;
}

/* ========================================================================
 * A6: sub-pixel refinement (core.py:343-463) in FP64.
 * ======================================================================== */
static int refine_one(const pyramid_t* P, cand_t cd, int border, double thr, double edge, kpt_t* kp)
{
    int o = cd.o, si = cd.s, row = cd.r, col = cd.c;
    int H = P->Ho[o], W = P->Wo[o], S = P->S;
    size_t HW = (size_t)H * W;
    const float* D = P->D[o];
#define DV(s_, r_, c_) ((double)D[(size_t)(s_) * HW + (size_t)(r_) * W + (c_)])
    double ox = 0.0, oy = 0.0, os = 0.0;
    int converged = 0;
    for (int it = 0; it < 5; ++it) {
        if (si < 1 || si > S || row < border || row >= H - border || col < border || col >= W - border) break;
        double cv = DV(si, row, col);
        double dx = (DV(si, row, col + 1) - DV(si, row, col - 1)) * 0.5;
        double dy = (DV(si, row + 1, col) - DV(si, row - 1, col)) * 0.5;
        double ds = (DV(si + 1, row, col) - DV(si - 1, row, col)) * 0.5;
        double dxx = (DV(si, row, col + 1) - 2.0 * cv) + DV(si, row, col - 1);
        double dyy = (DV(si, row + 1, col) - 2.0 * cv) + DV(si, row - 1, col);
        double dss = (DV(si + 1, row, col) - 2.0 * cv) + DV(si - 1, row, col);
        double dxy = (((DV(si, row + 1, col + 1) - DV(si, row + 1, col - 1)) - DV(si, row - 1, col + 1)) + DV(si, row - 1, col - 1)) * 0.25;
        double dxs = (((DV(si + 1, row, col + 1) - DV(si + 1, row, col - 1)) - DV(si - 1, row, col + 1)) + DV(si - 1, row, col - 1)) * 0.25;
        double dys = (((DV(si + 1, row + 1, col) - DV(si + 1, row - 1, col)) - DV(si - 1, row + 1, col)) + DV(si - 1, row - 1, col)) * 0.25;
        double det = (dxx * (dyy * dss - dys * dys) - dxy * (dxy * dss - dys * dxs)) + dxs * (dxy * dys - dyy * dxs);
        if (fabs(det) < 1e-12) break;
        double p = -dx, q = -dy, rv = -ds;
        ox = ((p * (dyy * dss - dys * dys) - dxy * (q * dss - dys * rv)) + dxs * (q * dys - dyy * rv)) / det;
        oy = ((dxx * (q * dss - dys * rv) - p * (dxy * dss - dys * dxs)) + dxs * (dxy * rv - q * dxs)) / det;
        os = ((dxx * (dyy * rv - q * dys) - dxy * (dxy * rv - q * dxs)) + p * (dxy * dys - dyy * dxs)) / det;
        if (fabs(ox) < 0.5 && fabs(oy) < 0.5 && fabs(os) < 0.5) { converged = 1; break; }
        col += (ox >= 0.0) ? (int)(ox + 0.5) : (int)(ox - 0.5);
        row += (oy >= 0.0) ? (int)(oy + 0.5) : (int)(oy - 0.5);
        si  += (os >= 0.0) ? (int)(os + 0.5) : (int)(os - 0.5);
    }
    if (!converged) return 0;
    if (si < 1 || si > S || row < border || row >= H - border || col < border || col >= W - border) return 0;
    double cv = DV(si, row, col);
    double dx = (DV(si, row, col + 1) - DV(si, row, col - 1)) * 0.5;
    double dy = (DV(si, row + 1, col) - DV(si, row - 1, col)) * 0.5;
    double ds = (DV(si + 1, row, col) - DV(si - 1, row, col)) * 0.5;
    double resp = fabs(cv + 0.5 * ((dx * ox + dy * oy) + ds * os));
    if (resp < thr) return 0;
    double dxx = (DV(si, row, col + 1) - 2.0 * cv) + DV(si, row, col - 1);
    double dyy = (DV(si, row + 1, col) - 2.0 * cv) + DV(si, row - 1, col);
    double dxy = (((DV(si, row + 1, col + 1) - DV(si, row + 1, col - 1)) - DV(si, row - 1, col + 1)) + DV(si, row - 1, col - 1)) * 0.25;
    double tr = dxx + dyy;
    double det2 = dxx * dyy - dxy * dxy;
    if (det2 <= 0.0 || (tr * tr) / det2 >= edge) return 0;
#undef DV
    kp->o = o; kp->si = si;
    kp->ri = (float)((double)row + oy);
    kp->ci = (float)((double)col + ox);
    kp->resp = (float)resp;
    return 1;
}

static int detect_keypoints(const pyramid_t* P, const or_params* p, double ct, kpt_t** out)
{
    float thr32 = (float)(ct / (double)p->n_scales);
    float edge32 = (float)((p->edge_thresh + 1.0) * (p->edge_thresh + 1.0) / p->edge_thresh);
    cand_t* c = NULL;
    int nc = find_extrema(P, thr32, p->border, p->max_cands_per_octave, &c);
    kpt_t* k = (kpt_t*)malloc(sizeof(kpt_t) * (nc > 0 ? nc : 1));
    int n = 0;
    for (int i = 0; i < nc; ++i)
        if (refine_one(P, c[i], p->border, (double)thr32, (double)edge32, &k[n])) n++;
    free(c);
    *out = k;
    return n;
}

/* ========================================================================
 * A8: orientation histogram (core.py:228-333)
 * ======================================================================== */
static int orientation_one(const float* G, int H, int W, float ri, float ci, float sigma, float* angles)
{
    int radius = (int)(4.5f * sigma + 0.5f);
    if (radius < 1) radius = 1;
    float sigma_w = 1.5f * sigma;
    float inv_2sw2 = 1.0f / (2.0f * sigma_w * sigma_w + 1e-12f);
    int r_min = (int)ri - radius; if (r_min < 1) r_min = 1;
    int r_max = (int)ri + radius; if (r_max > H - 2) r_max = H - 2;
    int c_min = (int)ci - radius; if (c_min < 1) c_min = 1;
    int c_max = (int)ci + radius; if (c_max > W - 2) c_max = W - 2;
    int n_cols = c_max - c_min + 1; if (n_cols < 1) n_cols = 1;
    int n_rows = r_max - r_min + 1; if (n_rows < 0) n_rows = 0;
    int total = n_rows * n_cols;
    float part[LANES][36];
    memset(part, 0, sizeof(part));
    for (int pidx = 0; pidx < total; ++pidx) {
        int lane = pidx % LANES;
        int r = r_min + pidx / n_cols;
        int c = c_min + pidx % n_cols;
        float dx = G[(size_t)r * W + c + 1] - G[(size_t)r * W + c - 1];
        float dy = G[(size_t)(r + 1) * W + c] - G[(size_t)(r - 1) * W + c];
        float mag = sqrtf(dx * dx + dy * dy);
        float ori = fmod360(f_atan2(dy, dx) * 57.29577951f + 360.0f);
        float dr = (float)r - ri, dc = (float)c - ci;
        float wt = f_exp(-(dr * dr + dc * dc) * inv_2sw2);
        float bin_f = ori * 0.1f;
        int bin_i = ((int)bin_f) % 36;
        int bin_r = (bin_i + 1) % 36;
        float frac = bin_f - floorf(bin_f);
        part[lane][bin_i] += mag * wt * (1.0f - frac);
        part[lane][bin_r] += mag * wt * frac;
    }
    float hist[36], smooth[36];
    for (int b = 0; b < 36; ++b) {
        float t[LANES];
        for (int l = 0; l < LANES; ++l) t[l] = part[l][b];
        hist[b] = wave_tree_f(t);
    }
    for (int it = 0; it < 2; ++it) {
        for (int b = 0; b < 36; ++b)
            smooth[b] = (0.25f * hist[(b + 35) % 36] + 0.5f * hist[b]) + 0.25f * hist[(b + 1) % 36];
        for (int b = 0; b < 36; ++b) hist[b] = smooth[b];
    }
    float peak = 0.0f;
    for (int b = 0; b < 36; ++b) if (hist[b] > peak) peak = hist[b];
    float thresh = 0.8f * peak;
    int n = 0;
    for (int b = 0; b < 36; ++b) {
        if (hist[b] >= thresh) {
            float l = hist[(b + 35) % 36], rv = hist[(b + 1) % 36];
            if (hist[b] >= l && hist[b] >= rv) {
                float denom = (l - 2.0f * hist[b]) + rv;
                float fine = (float)b + (fabsf(denom) > 1e-12f ? (0.5f * (l - rv)) / denom : 0.0f);
                float angle = fmod360((fine / 36.0f) * 360.0f + 360.0f);
                if (n < 4) angles[n++] = angle;
            }
        }
    }
    for (int k = n; k < 4; ++k) angles[k] = -1.0f;
    return n;
}

/* ========================================================================
 * A9: 4x4x8 descriptor (core.py:693-835) + RootSIFT (core.py:1870-1873)
 * ======================================================================== */
/* one trilinear contribution c in [0, 2^32), already scaled by 2^30, as the
 * integer floor(c) */
static uint64_t desc_fx(float c)
{
    return (uint64_t)(uint32_t)c;
}

/* the 4x4x8 histogram of one patch of half-size ph, added into hist
 * (core.py:693-792; one scale of the DSP loop, core.py:883-950) */
static void desc_hist_one(const float* G, int H, int W, float ri, float ci, float ang_rad, float cos_a,
                          float sin_a, int ph, uint64_t* hist)
{
    float scale_sp = 4.0f / (2.0f * (float)ph);
    float sigma_d = (float)ph * 0.5657f;
    float inv_2s2 = 1.0f / (2.0f * sigma_d * sigma_d + 1e-12f);
    int side = 2 * ph + 1;
    int total_px = side * side;
    int iri = (int)ri, ici = (int)ci;
    for (int flat = 0; flat < total_px; ++flat) {
        int r_off = flat / side - ph;
        int c_off = flat % side - ph;
        int rr = iri + r_off, cc = ici + c_off;
        if (rr < 1 || rr >= H - 1 || cc < 1 || cc >= W - 1) continue;
        float xr = cos_a * (float)c_off - sin_a * (float)r_off;
        float yr = sin_a * (float)c_off + cos_a * (float)r_off;
        float xb = xr * scale_sp + 1.5f;
        float yb = yr * scale_sp + 1.5f;
        if (xb < -0.5f || xb >= 3.5f || yb < -0.5f || yb >= 3.5f) continue;
        float dx = G[(size_t)rr * W + cc + 1] - G[(size_t)rr * W + cc - 1];
        float dy = G[(size_t)(rr + 1) * W + cc] - G[(size_t)(rr - 1) * W + cc];
        float mag = sqrtf(dx * dx + dy * dy);
        float ori_deg = fmod360((f_atan2(dy, dx) - ang_rad) * 57.29577951f + 360.0f);
        float gw = f_exp(-(xr * xr + yr * yr) * inv_2s2);
        /* the pixel weight bounds its 8 contributions: one clamp below 2^32 */
        float wms = fminf((mag * gw) * 0x1p30f, 0x1.fffffep31f);
        float ob = ori_deg * 0.02222222f;
        int oi0 = (int)floorf(ob);
        float df = ob - (float)oi0;
        int xi0 = (int)floorf(xb), yi0 = (int)floorf(yb);
        float dxb = xb - (float)xi0, dyb = yb - (float)yi0;
        int xi0c = xi0 < 0 ? 0 : (xi0 > 3 ? 3 : xi0);
        int yi0c = yi0 < 0 ? 0 : (yi0 > 3 ? 3 : yi0);
        int xi1c = xi0 + 1 < 0 ? 0 : (xi0 + 1 > 3 ? 3 : xi0 + 1);
        int yi1c = yi0 + 1 < 0 ? 0 : (yi0 + 1 > 3 ? 3 : yi0 + 1);
        int oi0w = ((oi0 % 8) + 8) % 8;
        int oi1w = ((oi0 + 1) % 8 + 8) % 8;
        float wy0 = 1.0f - dyb, wy1 = dyb, wx0 = 1.0f - dxb, wx1 = dxb, wo0 = 1.0f - df, wo1 = df;
        hist[yi0c * 32 + xi0c * 8 + oi0w] += desc_fx(wy0 * wx0 * wo0 * wms);
        hist[yi0c * 32 + xi0c * 8 + oi1w] += desc_fx(wy0 * wx0 * wo1 * wms);
        hist[yi0c * 32 + xi1c * 8 + oi0w] += desc_fx(wy0 * wx1 * wo0 * wms);
        hist[yi0c * 32 + xi1c * 8 + oi1w] += desc_fx(wy0 * wx1 * wo1 * wms);
        hist[yi1c * 32 + xi0c * 8 + oi0w] += desc_fx(wy1 * wx0 * wo0 * wms);
        hist[yi1c * 32 + xi0c * 8 + oi1w] += desc_fx(wy1 * wx0 * wo1 * wms);
        hist[yi1c * 32 + xi1c * 8 + oi0w] += desc_fx(wy1 * wx1 * wo0 * wms);
        hist[yi1c * 32 + xi1c * 8 + oi1w] += desc_fx(wy1 * wx1 * wo1 * wms);
    }
}

/* DSP-SIFT scale multipliers (core.py:1198-1202), float32 */
static int dsp_scales(int n, float* sm)
{
    const float r2 = (float)(1.0 / sqrt(2.0)), s2 = (float)sqrt(2.0);
    if (n == 3) { sm[0] = r2; sm[1] = 1.0f; sm[2] = s2; return 3; }
    if (n == 4) { sm[0] = 0.5f; sm[1] = r2; sm[2] = 1.0f; sm[3] = s2; return 4; }
    if (n == 5) { sm[0] = 0.5f; sm[1] = r2; sm[2] = 1.0f; sm[3] = s2; sm[4] = 2.0f; return 5; }
    sm[0] = 1.0f;
    return 0;
}

/* sift_descriptor_coop (core.py:693-835) or, with dsp_n_scales 3/4/5,
 * sift_descriptor_dsp_coop (core.py:845-997): one histogram over the patch
 * half-sizes max(4, (int)(ph * m + 0.5f)), divided by the number of scales;
 * then L2 -> clip 0.2 -> L2 and RootSIFT (core.py:1870-1873) */
static void descriptor_one(const float* G, int H, int W, float ri, float ci, float ang_rad, int ph,
                           int rootsift, int dsp_n_scales, float* out, uint8_t* out_u8)
{
    /* bins are exact integer sums of the rounded contributions: order-free */
    uint64_t hist[128];
    memset(hist, 0, sizeof(hist));
    double sd, cd;
    or_sincos((double)ang_rad, &sd, &cd);
    float cos_a = (float)cd, sin_a = (float)sd;
    float sm[5];
    int n_dsp = dsp_scales(dsp_n_scales, sm);
    if (n_dsp == 0) {
        desc_hist_one(G, H, W, ri, ci, ang_rad, cos_a, sin_a, ph, hist);
    } else {
        for (int d = 0; d < n_dsp; ++d) {
            int phd = (int)((float)ph * sm[d] + 0.5f);
            desc_hist_one(G, H, W, ri, ci, ang_rad, cos_a, sin_a, phd < 4 ? 4 : phd, hist);
        }
    }
    float v[128];
    for (int b = 0; b < 128; ++b) v[b] = (float)((double)hist[b] * 0x1p-30);
    if (n_dsp > 0) {
        const float inv_n = 1.0f / (float)n_dsp;
        for (int b = 0; b < 128; ++b) v[b] = v[b] * inv_n;
    }
    /* L2 -> clip 0.2 -> L2 (core.py:793-834); lane l holds bins 2l, 2l+1 */
    float pp[LANES];
    for (int l = 0; l < LANES; ++l) pp[l] = v[2 * l] * v[2 * l] + v[2 * l + 1] * v[2 * l + 1];
    float inv = 1.0f / sqrtf(wave_tree_f(pp) + 1e-14f);
    for (int b = 0; b < 128; ++b) { float t = v[b] * inv; v[b] = t > 0.2f ? 0.2f : t; }
    for (int l = 0; l < LANES; ++l) pp[l] = v[2 * l] * v[2 * l] + v[2 * l + 1] * v[2 * l + 1];
    inv = 1.0f / sqrtf(wave_tree_f(pp) + 1e-14f);
    for (int b = 0; b < 128; ++b) v[b] = v[b] * inv;
    if (rootsift) {
        for (int l = 0; l < LANES; ++l) pp[l] = fabsf(v[2 * l]) + fabsf(v[2 * l + 1]);
        float l1 = wave_tree_f(pp) + 1e-7f;
        for (int b = 0; b < 128; ++b) { float t = v[b] / l1; v[b] = sqrtf(t > 0.0f ? t : 0.0f); }
    }
    for (int b = 0; b < 128; ++b) {
        out[b] = v[b];
        float q = v[b] * 512.0f + 0.5f;
        int iq = (int)q;
        out_u8[b] = (uint8_t)(iq > 255 ? 255 : (iq < 0 ? 0 : iq));
    }
}

/* ========================================================================
 * A7/A10: full detectAndCompute (core.py:2021-2172)
 * ======================================================================== */
typedef struct { int kp; float angle; float x, y; } okp_t;

static int cmp_resp_desc(const void* a, const void* b)
{
    const kpt_t* ka = *(const kpt_t* const*)a;
    const kpt_t* kb = *(const kpt_t* const*)b;
    if (ka->resp > kb->resp) return -1;
    if (ka->resp < kb->resp) return 1;
    return (ka < kb) ? -1 : (ka > kb);   /* stable by original position */
}

static const okp_t* g_sort_base;
static int cmp_xy(const void* a, const void* b)
{
    int ia = *(const int*)a, ib = *(const int*)b;
    const okp_t* A = g_sort_base + ia;
    const okp_t* B = g_sort_base + ib;
    if (A->x < B->x) return -1;
    if (A->x > B->x) return 1;
    if (A->y < B->y) return -1;
    if (A->y > B->y) return 1;
    return (ia < ib) ? -1 : (ia > ib);
}

int or_detect_describe(const uint8_t* img, int H, int W, int pitch, const or_params* p,
                       int cap, float* kp_out, float* desc_out, uint8_t* desc_u8)
{
    if (p->n_octaves < 1 || p->n_octaves > 16 || p->n_scales < 1 || p->n_scales > 32) return -1;
    pyramid_t P;
    memset(&P, 0, sizeof(P));
    double ct = p->adaptive_thresh ? or_adaptive_thresh(img, H, W, pitch, p->contrast_thresh) : p->contrast_thresh;
    build_pyramid(img, H, W, pitch, p, &P);
    kpt_t* kps = NULL;
    int n = detect_keypoints(&P, p, ct, &kps);
    if (n < p->retry_below) {
        free(kps);
        n = detect_keypoints(&P, p, ct * 0.5, &kps);
    }
    /* top-max_keypoints by response, stable (core.py:2078-2080) */
    kpt_t* sel = kps;
    kpt_t* sorted = NULL;
    if (n > p->max_keypoints) {
        kpt_t** ptrs = (kpt_t**)malloc(sizeof(kpt_t*) * n);
        for (int i = 0; i < n; ++i) ptrs[i] = &kps[i];
        qsort(ptrs, n, sizeof(kpt_t*), cmp_resp_desc);
        sorted = (kpt_t*)malloc(sizeof(kpt_t) * p->max_keypoints);
        for (int i = 0; i < p->max_keypoints; ++i) sorted[i] = *ptrs[i];
        free(ptrs);
        sel = sorted;
        n = p->max_keypoints;
    }
    int S = p->n_scales;
    float kf = (float)pow(2.0, 1.0 / (double)S);
    double kd = pow(2.0, 1.0 / (double)S);
    double cs = P.doubled ? 0.5 : 1.0;
    /* orientation (core.py:1525-1603): sigma per si in float32 */
    float* ang = (float*)malloc(sizeof(float) * 4 * (n > 0 ? n : 1));
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        float pw = (float)pow((double)kf, (double)sel[i].si);
        float sig = (float)((float)p->sigma0 * pw);
        size_t HW = (size_t)P.Ho[sel[i].o] * P.Wo[sel[i].o];
        orientation_one(P.G[sel[i].o] + HW * sel[i].si, P.Ho[sel[i].o], P.Wo[sel[i].o],
                        sel[i].ri, sel[i].ci, sig, ang + 4 * i);
    }
    int m = 0;
    for (int i = 0; i < 4 * n; ++i) if (ang[i] >= 0.0f) m++;
    okp_t* ok = (okp_t*)malloc(sizeof(okp_t) * (m > 0 ? m : 1));
    int j = 0;
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 4; ++a)
            if (ang[4 * i + a] >= 0.0f) {
                double sf = ldexp(1.0, sel[i].o);
                ok[j].kp = i;
                ok[j].angle = ang[4 * i + a];
                ok[j].x = (float)(((double)sel[i].ci * sf) * cs);
                ok[j].y = (float)(((double)sel[i].ri * sf) * cs);
                j++;
            }
    int* order = (int*)malloc(sizeof(int) * (m > 0 ? m : 1));
    for (int i = 0; i < m; ++i) order[i] = i;
    g_sort_base = ok;
    qsort(order, m, sizeof(int), cmp_xy);
    int nout = m < cap ? m : cap;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < nout; ++t) {
        const okp_t* e = &ok[order[t]];
        const kpt_t* k = &sel[e->kp];
        double sf = ldexp(1.0, k->o);
        float* ko = kp_out + 6 * (size_t)t;
        ko[0] = e->x;
        ko[1] = e->y;
        ko[2] = (float)((((p->sigma0 * pow(kd, (double)k->si)) * sf) * cs) * 2.0);
        ko[3] = e->angle;
        ko[4] = k->resp;
        ko[5] = (float)k->o;
        /* descriptor geometry (core.py:1806-1810) in float32 */
        float pw = (float)pow((double)kf, (double)k->si);
        float t0 = (float)((float)(3.0 * p->sigma0) * pw);
        t0 = t0 * 1.41421354f;
        t0 = t0 * 2.5f;
        float phf = rintf(t0);
        int ph = phf < 4.0f ? 4 : (int)phf;
        float rad = e->angle * 0.0174532924f;
        size_t HW = (size_t)P.Ho[k->o] * P.Wo[k->o];
        float dtmp[128];
        uint8_t utmp[128];
        descriptor_one(P.G[k->o] + HW * k->si, P.Ho[k->o], P.Wo[k->o], k->ri, k->ci, rad, ph,
                       p->rootsift, p->dsp_n_scales, dtmp, utmp);
        if (desc_out) memcpy(desc_out + 128 * (size_t)t, dtmp, sizeof(dtmp));
        if (desc_u8) memcpy(desc_u8 + 128 * (size_t)t, utmp, sizeof(utmp));
    }
    free(order); free(ok); free(ang); free(sorted); free(kps);
    pyr_free(&P);
    return nout;
}

/* ========================================================================
 * A12: CLAHE (core.py:2562-2565: cv2.createCLAHE(clipLimit, (t, t)).apply on
 * the u8 gray image), restated from OpenCV's published algorithm:
 *   - when H or W is not a multiple of the tile count, the image is padded
 *     at the bottom by tiles_y - H % tiles_y and at the right by
 *     tiles_x - W % tiles_x (BORDER_REFLECT_101) before it is tiled;
 *   - per tile a 256-bin histogram, clipped at max(1, (int)(clip * area /
 *     256)); the clipped count is spread evenly, and the remainder adds one
 *     to every max(1, 256 / remainder)-th bin from bin 0;
 *   - LUT[v] = saturate(round_even((float)cumsum[v] * (255.0f / area)));
 *   - each pixel blends the LUTs of its 4 nearest tile centres bilinearly in
 *     float: ((l11 xa1 + l12 xa) ya1 + (l21 xa1 + l22 xa) ya), rounded.
 * ======================================================================== */
static int reflect101(int p, int len)
{
    if (len == 1) return 0;
    while (p < 0 || p >= len) p = p < 0 ? -p : 2 * len - 2 - p;
    return p;
}

static uint8_t sat_round_u8(float v)
{
    int r = (int)lrintf(v);
    return (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
}

int or_clahe(const uint8_t* src, int H, int W, int pitch, double clip, int tiles_x, int tiles_y, uint8_t* dst)
{
    if (H < 1 || W < 1 || tiles_x < 1 || tiles_y < 1) return -1;
    int He = H, We = W;
    if (H % tiles_y != 0 || W % tiles_x != 0) { He = H + tiles_y - H % tiles_y; We = W + tiles_x - W % tiles_x; }
    const int tw = We / tiles_x, th = He / tiles_y, area = tw * th;
    const float lut_scale = 255.0f / (float)area;
    int limit = 0;
    if (clip > 0.0) {
        const double l = clip * area / 256;
        limit = l >= (double)INT_MAX ? INT_MAX : (int)l;
        if (limit < 1) limit = 1;
    }
    uint8_t* lut = (uint8_t*)malloc((size_t)tiles_x * tiles_y * 256);
    for (int t = 0; t < tiles_x * tiles_y; ++t) {
        const int ty = t / tiles_x, tx = t % tiles_x;
        int hist[256];
        memset(hist, 0, sizeof(hist));
        for (int y = ty * th; y < (ty + 1) * th; ++y)
            for (int x = tx * tw; x < (tx + 1) * tw; ++x)
                hist[src[(size_t)reflect101(y, H) * pitch + reflect101(x, W)]]++;
        if (limit > 0) {
            int clipped = 0;
            for (int i = 0; i < 256; ++i)
                if (hist[i] > limit) { clipped += hist[i] - limit; hist[i] = limit; }
            const int batch = clipped / 256;
            int residual = clipped - batch * 256;
            for (int i = 0; i < 256; ++i) hist[i] += batch;
            if (residual != 0) {
                const int step = 256 / residual > 1 ? 256 / residual : 1;
                for (int i = 0; i < 256 && residual > 0; i += step, residual--) hist[i]++;
            }
        }
        int sum = 0;
        for (int i = 0; i < 256; ++i) {
            sum += hist[i];
            lut[(size_t)t * 256 + i] = sat_round_u8((float)sum * lut_scale);
        }
    }
    const float inv_tw = 1.0f / (float)tw, inv_th = 1.0f / (float)th;
    for (int y = 0; y < H; ++y) {
        const float tyf = (float)y * inv_th - 0.5f;
        int ty1 = (int)floorf(tyf), ty2 = ty1 + 1;
        const float ya = tyf - (float)ty1, ya1 = 1.0f - ya;
        if (ty1 < 0) ty1 = 0;
        if (ty2 > tiles_y - 1) ty2 = tiles_y - 1;
        for (int x = 0; x < W; ++x) {
            const float txf = (float)x * inv_tw - 0.5f;
            int tx1 = (int)floorf(txf), tx2 = tx1 + 1;
            const float xa = txf - (float)tx1, xa1 = 1.0f - xa;
            if (tx1 < 0) tx1 = 0;
            if (tx2 > tiles_x - 1) tx2 = tiles_x - 1;
            const int v = src[(size_t)y * pitch + x];
            const float l11 = lut[((size_t)ty1 * tiles_x + tx1) * 256 + v], l12 = lut[((size_t)ty1 * tiles_x + tx2) * 256 + v];
            const float l21 = lut[((size_t)ty2 * tiles_x + tx1) * 256 + v], l22 = lut[((size_t)ty2 * tiles_x + tx2) * 256 + v];
            const float res = (l11 * xa1 + l12 * xa) * ya1 + (l21 * xa1 + l22 * xa) * ya;
            dst[(size_t)y * W + x] = sat_round_u8(res);
        }
    }
    free(lut);
    return 0;
}

/* ========================================================================
 * Exact matcher (core.py:2840-2940 restated on uint8 descriptors):
 * d = sum (a-b)^2 in integers; top-2 with the smallest index kept on ties;
 * ratio test d1 < ratio^2 * d2 (double); mutual check.
 * ======================================================================== */
void or_match_u8(const uint8_t* a, int N, const uint8_t* b, int M, double ratio,
                 int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs)
{
    const int32_t BIG = 0x7fffffff;
    for (int j = 0; j < M; ++j) { bwd[3 * j] = -1; bwd[3 * j + 1] = BIG; bwd[3 * j + 2] = BIG; }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        int32_t d1 = BIG, d2 = BIG, j1 = -1;
        const uint8_t* ai = a + 128 * (size_t)i;
        for (int j = 0; j < M; ++j) {
            const uint8_t* bj = b + 128 * (size_t)j;
            int32_t d = 0;
            for (int k = 0; k < 128; ++k) { int t = (int)ai[k] - (int)bj[k]; d += t * t; }
            if (d < d1) { d2 = d1; d1 = d; j1 = j; }
            else if (d < d2) d2 = d;
        }
        fwd[3 * i] = j1; fwd[3 * i + 1] = d1; fwd[3 * i + 2] = d2;
    }
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < M; ++j) {
        int32_t d1 = BIG, d2 = BIG, i1 = -1;
        const uint8_t* bj = b + 128 * (size_t)j;
        for (int i = 0; i < N; ++i) {
            const uint8_t* ai = a + 128 * (size_t)i;
            int32_t d = 0;
            for (int k = 0; k < 128; ++k) { int t = (int)ai[k] - (int)bj[k]; d += t * t; }
            if (d < d1) { d2 = d1; d1 = d; i1 = i; }
            else if (d < d2) d2 = d;
        }
        bwd[3 * j] = i1; bwd[3 * j + 1] = d1; bwd[3 * j + 2] = d2;
    }
    double r2 = ratio * ratio;
    int np = 0;
    for (int i = 0; i < N; ++i) {
        int j = fwd[3 * i];
        if (j < 0 || M < 2) continue;
        if (!((double)fwd[3 * i + 1] < r2 * (double)fwd[3 * i + 2])) continue;
        if (bwd[3 * j] != i) continue;
        if (N < 2 || !((double)bwd[3 * j + 1] < r2 * (double)bwd[3 * j + 2])) continue;
        pairs[2 * np] = i; pairs[2 * np + 1] = j; np++;
    }
    *npairs = np;
}

/* ========================================================================
 * F4: the reference's float matcher (core.py:2840-2940 with int8_matching
 * off): rows L2-normalised with +1e-7 (core.py:2879-2880), PCA to `dims`
 * fitted on both sets (core.py:2703-2743), fp16-autocast similarity GEMM,
 * top-2 by similarity in both directions, ratio test on d = max(2 - 2 s, 0)
 * (core.py:2662-2690), mutual check (core.py:2933-2939).
 *
 * Arithmetic (defined here; the reference's CuPy/cuBLAS orders are not
 * reproducible):
 *   - norms: float sum of squares in index order, x / (sqrtf(s) + 1e-7f);
 *   - PCA: double mean and centred Gram, rows summed in index order within
 *     chunks of OR_PCA_CHUNK rows and the chunk partials summed in chunk
 *     order; covariance = Gram / n; eigenvectors by the cyclic parallel
 *     Jacobi method: each sweep runs 127 rounds of the circle-method pairing
 *     of the 128 indices (round r: (r, 127) and ((r + k) % 127, (r - k) %
 *     127) for k = 1..63), all 64 rotations of a round computed from the
 *     same matrix, then applied to rows, then to columns (then a_pq = a_qp =
 *     0), and to the eigenvector matrix; a pair rotates when |a_pq| >
 *     1e-15 sqrt(|a_pp a_qq|), sweeps stop after one without a rotation (at
 *     most OR_PCA_MAX_SWEEPS);
 *   - the `dims` largest eigenvalues (ties: lower index) give the columns of
 *     proj in ascending eigenvalue order (as eigh(cov)[:, -dims:]), each
 *     signed so that its largest-magnitude entry (first on ties) is positive;
 *   - projection: float c_j = x_j - mean_j, p_k = sum_j c_j proj_jk in j
 *     order (separate multiply and add), then p / (sqrtf(sum p_k^2) + 1e-7f);
 *   - similarity: inputs rounded to binary16, the exact sum of their
 *     products rounded to float, then to binary16 (the fp16 GEMM output).
 * ======================================================================== */
void or_l2n_rows(const float* x, int n, float* y)
{
    for (int i = 0; i < n; ++i) {
        const float* r = x + (size_t)i * 128;
        float s = 0.0f;
        for (int k = 0; k < 128; ++k) s = s + r[k] * r[k];
        const float den = sqrtf(s) + 1e-7f;
        for (int k = 0; k < 128; ++k) y[(size_t)i * 128 + k] = r[k] / den;
    }
}

static void jacobi_cs(double app, double aqq, double apq, double* c, double* s)
{
    const double theta = (aqq - app) / (2.0 * apq);
    double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
    if (theta < 0.0) t = -t;
    *c = 1.0 / sqrt(t * t + 1.0);
    *s = t * *c;
}

int or_pca_fit(const float* xa, int N, const float* xb, int M, int dims, float* mean32, float* proj,
               double* evals, int* sweeps_out)
{
    const int n = N + M;
    if (n < 1 || dims < 1 || dims > 128) return -1;
    const int nch = (n + OR_PCA_CHUNK - 1) / OR_PCA_CHUNK;
    double* part = (double*)calloc((size_t)nch * 128 * 128, sizeof(double));
    double mean[128];
#define ROW(r) ((r) < N ? xa + (size_t)(r) * 128 : xb + (size_t)((r) - N) * 128)
    for (int c = 0; c < nch; ++c)
        for (int r = c * OR_PCA_CHUNK; r < n && r < (c + 1) * OR_PCA_CHUNK; ++r)
            for (int j = 0; j < 128; ++j) part[(size_t)c * 128 + j] += (double)ROW(r)[j];
    for (int j = 0; j < 128; ++j) {
        double t = 0.0;
        for (int c = 0; c < nch; ++c) t += part[(size_t)c * 128 + j];
        mean[j] = t / (double)n;
    }
    memset(part, 0, (size_t)nch * 128 * 128 * sizeof(double));
    for (int c = 0; c < nch; ++c) {
        double* g = part + (size_t)c * 128 * 128;
        for (int r = c * OR_PCA_CHUNK; r < n && r < (c + 1) * OR_PCA_CHUNK; ++r) {
            double cr[128];
            for (int j = 0; j < 128; ++j) cr[j] = (double)ROW(r)[j] - mean[j];
            for (int j = 0; j < 128; ++j)
                for (int k = 0; k < 128; ++k) g[j * 128 + k] += cr[j] * cr[k];
        }
    }
#undef ROW
    double* A = (double*)malloc(128 * 128 * sizeof(double));
    double* V = (double*)malloc(128 * 128 * sizeof(double));
    for (int e = 0; e < 128 * 128; ++e) {
        double t = 0.0;
        for (int c = 0; c < nch; ++c) t += part[(size_t)c * 128 * 128 + e];
        A[e] = t / (double)n;
        V[e] = (e / 128 == e % 128) ? 1.0 : 0.0;
    }
    free(part);
    int sweeps = 0;
    for (; sweeps < OR_PCA_MAX_SWEEPS; ++sweeps) {
        int rotated = 0;
        for (int r = 0; r < 127; ++r) {
            int P[64], Q[64], rot[64];
            double C[64], S[64];
            for (int k = 0; k < 64; ++k) {
                int p, q;
                if (k == 0) { p = r; q = 127; }
                else { p = (r + k) % 127; q = (r - k + 127) % 127; }
                if (p > q) { int t = p; p = q; q = t; }
                P[k] = p; Q[k] = q;
                const double app = A[p * 128 + p], aqq = A[q * 128 + q], apq = A[p * 128 + q];
                rot[k] = fabs(apq) > 1e-15 * sqrt(fabs(app * aqq));
                if (rot[k]) jacobi_cs(app, aqq, apq, &C[k], &S[k]);
                rotated |= rot[k];
            }
            for (int k = 0; k < 64; ++k) {
                if (!rot[k]) continue;
                const int p = P[k], q = Q[k];
                for (int j = 0; j < 128; ++j) {
                    const double apj = A[p * 128 + j], aqj = A[q * 128 + j];
                    A[p * 128 + j] = C[k] * apj - S[k] * aqj;
                    A[q * 128 + j] = S[k] * apj + C[k] * aqj;
                }
            }
            for (int k = 0; k < 64; ++k) {
                if (!rot[k]) continue;
                const int p = P[k], q = Q[k];
                for (int i = 0; i < 128; ++i) {
                    const double aip = A[i * 128 + p], aiq = A[i * 128 + q];
                    A[i * 128 + p] = C[k] * aip - S[k] * aiq;
                    A[i * 128 + q] = S[k] * aip + C[k] * aiq;
                    const double vip = V[i * 128 + p], viq = V[i * 128 + q];
                    V[i * 128 + p] = C[k] * vip - S[k] * viq;
                    V[i * 128 + q] = S[k] * vip + C[k] * viq;
                }
            }
            for (int k = 0; k < 64; ++k)
                if (rot[k]) { A[P[k] * 128 + Q[k]] = 0.0; A[Q[k] * 128 + P[k]] = 0.0; }
        }
        if (!rotated) break;
    }
    /* eigenvalues in descending order (ties: lower index first) */
    int order[128];
    for (int j = 0; j < 128; ++j) order[j] = j;
    for (int a = 0; a < 128; ++a) {
        int best = a;
        for (int b = a + 1; b < 128; ++b) {
            const double vb = A[order[b] * 129], vbest = A[order[best] * 129];
            if (vb > vbest || (vb == vbest && order[b] < order[best])) best = b;
        }
        int t = order[a]; order[a] = order[best]; order[best] = t;
    }
    if (evals) for (int j = 0; j < 128; ++j) evals[j] = A[order[j] * 129];
    for (int k = 0; k < dims; ++k) {
        const int col = order[dims - 1 - k];
        int arg = 0;
        for (int j = 1; j < 128; ++j) if (fabs(V[j * 128 + col]) > fabs(V[arg * 128 + col])) arg = j;
        const double sg = V[arg * 128 + col] < 0.0 ? -1.0 : 1.0;
        for (int j = 0; j < 128; ++j) proj[j * dims + k] = (float)(sg * V[j * 128 + col]);
    }
    for (int j = 0; j < 128; ++j) mean32[j] = (float)mean[j];
    if (sweeps_out) *sweeps_out = sweeps;
    free(A); free(V);
    return 0;
}

void or_pca_project(const float* x, int n, const float* mean, const float* proj, int dims, float* y)
{
    for (int i = 0; i < n; ++i) {
        float c[128], p[128];
        for (int j = 0; j < 128; ++j) c[j] = x[(size_t)i * 128 + j] - mean[j];
        for (int k = 0; k < dims; ++k) {
            float acc = 0.0f;
            for (int j = 0; j < 128; ++j) acc = acc + c[j] * proj[j * dims + k];
            p[k] = acc;
        }
        float s = 0.0f;
        for (int k = 0; k < dims; ++k) s = s + p[k] * p[k];
        const float den = sqrtf(s) + 1e-7f;
        for (int k = 0; k < dims; ++k) y[(size_t)i * dims + k] = p[k] / den;
    }
}

void or_sim16(const float* a, int N, const float* b, int M, int D, float* s16)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        double ha[128];
        for (int k = 0; k < D; ++k) ha[k] = (double)or_f16_round(a[(size_t)i * D + k]);
        for (int j = 0; j < M; ++j) {
            double t = 0.0;   /* exact: products of binary16 values carry <= 22 bits */
            for (int k = 0; k < D; ++k) t += ha[k] * (double)or_f16_round(b[(size_t)j * D + k]);
            s16[(size_t)i * M + j] = or_f16_round((float)t);
        }
    }
}

static int32_t f_bits(float f) { int32_t u; memcpy(&u, &f, 4); return u; }

/* top-2 by similarity (ties: lower index), distances max(2 - 2 s, 0) */
static void sim_top2(const float* s16, int N, int M, int transposed, int32_t* out)
{
    const int rows = transposed ? M : N, cols = transposed ? N : M;
    for (int i = 0; i < rows; ++i) {
        float b1 = -INFINITY, b2 = -INFINITY;
        int j1 = -1;
        for (int j = 0; j < cols; ++j) {
            const float s = transposed ? s16[(size_t)j * M + i] : s16[(size_t)i * M + j];
            if (s > b1) { b2 = b1; b1 = s; j1 = j; }
            else if (s > b2) b2 = s;
        }
        out[3 * i] = j1;
        float d1 = 2.0f - 2.0f * b1, d2 = 2.0f - 2.0f * b2;
        d1 = d1 > 0.0f ? d1 : 0.0f;
        d2 = d2 > 0.0f ? d2 : 0.0f;
        out[3 * i + 1] = j1 < 0 ? 0x7fffffff : f_bits(d1);
        out[3 * i + 2] = cols < 2 ? 0x7fffffff : f_bits(d2);
    }
}

static int sim_pass(const int32_t* t, float r2)
{
    if (t[0] < 0 || t[2] == 0x7fffffff) return 0;
    float d1, d2;
    memcpy(&d1, &t[1], 4); memcpy(&d2, &t[2], 4);
    return d2 > 1e-10f && d1 < r2 * d2;
}

void or_match_sim16(const float* a, int N, const float* b, int M, int D, double ratio,
                    int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs)
{
    float* s16 = (float*)malloc((size_t)(N > 0 ? N : 1) * (M > 0 ? M : 1) * sizeof(float));
    or_sim16(a, N, b, M, D, s16);
    sim_top2(s16, N, M, 0, fwd);
    sim_top2(s16, N, M, 1, bwd);
    free(s16);
    const float r2 = (float)(ratio * ratio);
    int np = 0;
    for (int i = 0; i < N; ++i) {
        if (!sim_pass(fwd + 3 * i, r2)) continue;
        const int j = fwd[3 * i];
        if (bwd[3 * j] != i || !sim_pass(bwd + 3 * j, r2)) continue;
        pairs[2 * np] = i; pairs[2 * np + 1] = j; np++;
    }
    *npairs = np;
}

/* ========================================================================
 * F3: the reference's composition of a pair (core.py:3686-3714, 3783-3802):
 * grid_sample warp, gain compensation, DP seam, multiband blend.
 *
 * Arithmetic (defined here; the reference runs it in torch, NumPy, CuPy and
 * OpenCV):
 *   warp (core.py:3211-3239): X = u - (float)ox, source = hf (X, Y, 1) as
 *     ((h0 X + h1 Y) + h2), z clamped at 1e-10, g = 2 s / (n - 1) - 1, valid
 *     when -1 <= g <= 1; grid_sample bilinear (align_corners, zeros):
 *     i = ((g + 1) / 2)(n - 1), corner weights nw = (x_se - x)(y_se - y), ne =
 *     (x - x_sw)(y_sw - y), sw = (x_ne - x)(y - y_ne), se = (x - x_nw)(y - y_nw),
 *     out = ((nw v + ne v) + sw v) + se v over in-bounds corners (none when
 *     |i| >= 2^30), truncated;
 *   gain (core.py:3696-3697): exact integer sums over the overlap, float
 *     means m = (float)(sum / n), g = (m2 + 1e-3f) / (m1 + 1e-3f) in [0.5, 2],
 *     w1 = min(max(w1 g, 0), 255); or (color_match 'hist', core.py:3694-3695)
 *     w1 = lut[w1], the per-channel histogram-matching LUT of or_hist_lut;
 *   seam cost (core.py:3411-3417): sqrtf((d0^2 + d1^2) + d2^2) + 0.05f *
 *     sqrtf(gx^2 + gy^2) on the channel mean of (w1 + c2) * 0.5; Sobel as
 *     SciPy (derivative [-1, 0, 1] then smoothing [1, 2, 1], mode 'reflect',
 *     double accumulation); 1e9 outside the overlap;
 *   DP seam (core.py:3287-3384) as the reference, in double, first minimum
 *     on ties; orientation and side from the overlap and solo extents;
 *   seam mask blur: cv2.GaussianBlur(sigma 3): 25 float32 taps as
 *     getGaussianKernel, rows then columns, BORDER_REFLECT_101, float32 sums
 *     in tap order; written back only inside the overlap;
 *   pyramid (core.py:3475-3553): down = SciPy/CuPy gaussian_filter(sigma 1,
 *     'reflect', radius 4; float32 weights, fmaf in tap order, rows then
 *     columns) then [::2, ::2]; up = zoom(order 1) with align-corners
 *     coordinates in double, clamped to the input, double weights; blend
 *     g lp1 + (1 - g) lp2 and reconstruction in float32; output truncated.
 * ======================================================================== */
static void inv3(const double* a, double* o);

static int refl_half(int p, int n)
{
    if (n == 1) return 0;
    const int period = 2 * n;
    p %= period;
    if (p < 0) p += period;
    return p < n ? p : period - 1 - p;
}

static int refl_101(int p, int n)
{
    if (n == 1) return 0;
    const int period = 2 * n - 2;
    p %= period;
    if (p < 0) p += period;
    return p < n ? p : period - p;
}

void or_grid_warp(const uint8_t* img, int h, int w, const double* H, int cw, int ch, double ox, double oy,
                  uint8_t* out, uint8_t* mask)
{
    double Hi[9];
    inv3(H, Hi);
    float hf[9];
    for (int i = 0; i < 9; ++i) hf[i] = (float)Hi[i];
    const float oxf = (float)ox, oyf = (float)oy, wm = (float)(w - 1), hm = (float)(h - 1);
    #pragma omp parallel for schedule(static)
    for (int v = 0; v < ch; ++v)
        for (int u = 0; u < cw; ++u) {
            const float X = (float)u - oxf, Y = (float)v - oyf;
            const float sxn = (hf[0] * X + hf[1] * Y) + hf[2];
            const float syn = (hf[3] * X + hf[4] * Y) + hf[5];
            float szn = (hf[6] * X + hf[7] * Y) + hf[8];
            if (szn < 1e-10f) szn = 1e-10f;
            const float sx = sxn / szn, sy = syn / szn;
            const float gx = (2.0f * sx) / wm - 1.0f, gy = (2.0f * sy) / hm - 1.0f;
            const size_t o = (size_t)v * cw + u;
            mask[o] = gx >= -1.0f && gx <= 1.0f && gy >= -1.0f && gy <= 1.0f;
            const float ix = ((gx + 1.0f) / 2.0f) * wm, iy = ((gy + 1.0f) / 2.0f) * hm;
            const float fx = floorf(ix), fy = floorf(iy);
            const float ix_nw = fx, iy_nw = fy, ix_ne = fx + 1.0f, iy_ne = fy, ix_sw = fx, iy_sw = fy + 1.0f;
            const float ix_se = fx + 1.0f, iy_se = fy + 1.0f;
            const float nw = (ix_se - ix) * (iy_se - iy), ne = (ix - ix_sw) * (iy_sw - iy);
            const float sw = (ix_ne - ix) * (iy - iy_ne), se = (ix - ix_nw) * (iy - iy_nw);
            const int fin = fabsf(ix) < 1073741824.0f && fabsf(iy) < 1073741824.0f;
            const int x0 = fin ? (int)fx : -2, y0 = fin ? (int)fy : -2;
            const int in_nw = x0 >= 0 && x0 < w && y0 >= 0 && y0 < h, in_ne = x0 + 1 >= 0 && x0 + 1 < w && y0 >= 0 && y0 < h;
            const int in_sw = x0 >= 0 && x0 < w && y0 + 1 >= 0 && y0 + 1 < h;
            const int in_se = x0 + 1 >= 0 && x0 + 1 < w && y0 + 1 >= 0 && y0 + 1 < h;
            for (int c = 0; c < 3; ++c) {
                float acc = 0.0f;
                if (fin) {
                    if (in_nw) acc = acc + (float)img[((size_t)y0 * w + x0) * 3 + c] * nw;
                    if (in_ne) acc = acc + (float)img[((size_t)y0 * w + x0 + 1) * 3 + c] * ne;
                    if (in_sw) acc = acc + (float)img[((size_t)(y0 + 1) * w + x0) * 3 + c] * sw;
                    if (in_se) acc = acc + (float)img[((size_t)(y0 + 1) * w + x0 + 1) * 3 + c] * se;
                }
                acc = acc < 0.0f ? 0.0f : (acc > 255.0f ? 255.0f : acc);
                out[o * 3 + c] = (uint8_t)acc;
            }
        }
}

void or_gauss_weights_sp(float* w9)
{
    double phi[9];
    for (int k = 0; k < 9; ++k) { const double x = (double)(k - 4); phi[k] = exp(-0.5 * (x * x)); }
    /* numpy's pairwise sum of 9 terms */
    const double sum = (((phi[0] + phi[1]) + (phi[2] + phi[3])) + ((phi[4] + phi[5]) + (phi[6] + phi[7]))) + phi[8];
    for (int k = 0; k < 9; ++k) w9[k] = (float)(phi[k] / sum);
}

void or_gauss_weights_cv(float* w25)
{
    double sum = 0.0;
    for (int i = 0; i < 25; ++i) {
        const double x = (double)i - 12.0;
        w25[i] = (float)exp((-0.5 / 9.0) * x * x);
        sum += w25[i];
    }
    sum = 1.0 / sum;
    for (int i = 0; i < 25; ++i) w25[i] = (float)(w25[i] * sum);
}

void or_gauss_down(const float* in, int H, int W, const float* w9, float* out)
{
    const int Ho = (H + 1) / 2, Wo = (W + 1) / 2;
    float* t = (float*)malloc((size_t)Ho * W * sizeof(float));
    for (int yo = 0; yo < Ho; ++yo)
        for (int x = 0; x < W; ++x) {
            float acc = 0.0f;
            for (int k = 0; k < 9; ++k) acc = fmaf(in[(size_t)refl_half(2 * yo + k - 4, H) * W + x], w9[k], acc);
            t[(size_t)yo * W + x] = acc;
        }
    for (int yo = 0; yo < Ho; ++yo)
        for (int xo = 0; xo < Wo; ++xo) {
            float acc = 0.0f;
            for (int k = 0; k < 9; ++k) acc = fmaf(t[(size_t)yo * W + refl_half(2 * xo + k - 4, W)], w9[k], acc);
            out[(size_t)yo * Wo + xo] = acc;
        }
    free(t);
}

static void zoom_coord(int o, int n_in, int n_out, int* i0, int* i1, double* w0, double* w1)
{
    double c = n_out > 1 ? (double)o * ((double)(n_in - 1) / (double)(n_out - 1)) : 0.0;
    if (c > (double)(n_in - 1)) c = (double)(n_in - 1);
    const double f = floor(c);
    *i0 = (int)f;
    *i1 = *i0 + 1 < n_in ? *i0 + 1 : n_in - 1;
    *w1 = c - f;
    *w0 = 1.0 - *w1;
}

void or_zoom_up(const float* in, int h, int w, int th, int tw, float* out)
{
    for (int y = 0; y < th; ++y) {
        int y0, y1; double wy0, wy1;
        zoom_coord(y, h, th, &y0, &y1, &wy0, &wy1);
        for (int x = 0; x < tw; ++x) {
            int x0, x1; double wx0, wx1;
            zoom_coord(x, w, tw, &x0, &x1, &wx0, &wx1);
            double acc = (double)in[(size_t)y0 * w + x0] * (wy0 * wx0);
            acc = acc + (double)in[(size_t)y0 * w + x1] * (wy0 * wx1);
            acc = acc + (double)in[(size_t)y1 * w + x0] * (wy1 * wx0);
            acc = acc + (double)in[(size_t)y1 * w + x1] * (wy1 * wx1);
            out[(size_t)y * tw + x] = (float)acc;
        }
    }
}

/* SciPy sobel: derivative along `axis`, then [1, 2, 1] along the other,
   mode 'reflect', double accumulation in tap order, float32 between passes */
void or_sobel(const float* in, int H, int W, int axis, float* out)
{
    float* t = (float*)malloc((size_t)H * W * sizeof(float));
    const double dk[3] = { -1.0, 0.0, 1.0 }, sk[3] = { 1.0, 2.0, 1.0 };
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) {
                const float v = axis == 1 ? in[(size_t)y * W + refl_half(x + k - 1, W)] : in[(size_t)refl_half(y + k - 1, H) * W + x];
                acc += (double)v * dk[k];
            }
            t[(size_t)y * W + x] = (float)acc;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) {
                const float v = axis == 1 ? t[(size_t)refl_half(y + k - 1, H) * W + x] : t[(size_t)y * W + refl_half(x + k - 1, W)];
                acc += (double)v * sk[k];
            }
            out[(size_t)y * W + x] = (float)acc;
        }
    free(t);
}

void or_cv_blur25(const float* in, int H, int W, const float* w25, float* out)
{
    float* t = (float*)malloc((size_t)H * W * sizeof(float));
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float acc = 0.0f;
            for (int k = 0; k < 25; ++k) acc = acc + w25[k] * in[(size_t)y * W + refl_101(x + k - 12, W)];
            t[(size_t)y * W + x] = acc;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float acc = 0.0f;
            for (int k = 0; k < 25; ++k) acc = acc + w25[k] * t[(size_t)refl_101(y + k - 12, H) * W + x];
            out[(size_t)y * W + x] = acc;
        }
    free(t);
}

/* vertical: one column per row over the rows that hold overlap; horizontal:
   the transpose.  seam[H] (or seam[W]); rows without overlap keep the last
   value the reference leaves there (the final seam position). */
void or_dp_seam(const float* cost, const uint8_t* ov, int H, int W, int vertical, int32_t* seam)
{
    const int L = vertical ? H : W, N = vertical ? W : H;     /* L lines of N cells */
#define CST(l, n) (vertical ? cost[(size_t)(l) * W + (n)] : cost[(size_t)(n) * W + (l)])
#define OVM(l, n) (vertical ? ov[(size_t)(l) * W + (n)] : ov[(size_t)(n) * W + (l)])
    const double INF = 1e18;
    int* lines = (int*)malloc((size_t)L * sizeof(int));
    int nl = 0;
    for (int l = 0; l < L; ++l) {
        int any = 0;
        for (int n = 0; n < N && !any; ++n) any = OVM(l, n) != 0;
        if (any) lines[nl++] = l;
    }
    if (nl == 0) { for (int l = 0; l < L; ++l) seam[l] = 0; free(lines); return; }
    double* dp = (double*)malloc((size_t)N * sizeof(double));
    double* nd = (double*)malloc((size_t)N * sizeof(double));
    int32_t* bt = (int32_t*)calloc((size_t)L * N, sizeof(int32_t));
    for (int n = 0; n < N; ++n) dp[n] = OVM(lines[0], n) ? (double)CST(lines[0], n) : INF;
    for (int i = 1; i < nl; ++i) {
        const int l = lines[i];
        for (int n = 0; n < N; ++n) {
            const double pl = n > 0 ? dp[n - 1] : INF, pm = dp[n], pr = n < N - 1 ? dp[n + 1] : INF;
            int a = 0; double best = pl;
            if (pm < best) { best = pm; a = 1; }
            if (pr < best) { best = pr; a = 2; }
            nd[n] = OVM(l, n) ? (double)CST(l, n) + best : INF;
            int b = n + a - 1;
            bt[(size_t)l * N + n] = b < 0 ? 0 : (b > N - 1 ? N - 1 : b);
        }
        double* t = dp; dp = nd; nd = t;
    }
    const int last = lines[nl - 1];
    int sc = 0; double bc = INF; int first = 1;
    for (int n = 0; n < N; ++n) {
        const double c = OVM(last, n) ? dp[n] : INF;
        if (first || c < bc) { bc = c; sc = n; first = 0; }
    }
    for (int l = 0; l < L; ++l) seam[l] = sc;
    for (int i = nl - 1; i > 0; --i) seam[lines[i - 1]] = bt[(size_t)lines[i] * N + seam[lines[i]]];
#undef CST
#undef OVM
    free(lines); free(dp); free(nd); free(bt);
}

/* np.interp(x, xp, arange(n)) with NumPy's rules: fp[0] left of xp[0],
   fp[n-1] right of xp[n-1] and at it, fp[j] at an exact xp[j] (j the last
   index with xp[j] <= x), else slope * (x - xp[j]) + fp[j] with slope =
   (fp[j+1] - fp[j]) / (xp[j+1] - xp[j]) */
static double np_interp_index(double x, const double* xp, int n)
{
    if (x < xp[0]) return 0.0;
    if (x > xp[n - 1]) return (double)(n - 1);
    int lo = 0, hi = n;                 /* first index with xp[i] > x */
    while (lo < hi) { const int mid = (lo + hi) >> 1; if (x >= xp[mid]) lo = mid + 1; else hi = mid; }
    const int j = lo - 1;
    if (j == n - 1) return (double)(n - 1);
    if (xp[j] == x) return (double)j;
    const double slope = ((double)(j + 1) - (double)j) / (xp[j + 1] - xp[j]);
    return slope * (x - xp[j]) + (double)j;
}

/* _hist_match (core.py:3591-3626) on integer-valued images: per channel, the
   overlap histograms of src and ref, their CDFs cumsum / (total + 1e-8) in
   double, lut[v] = interp(s_cdf[v], r_cdf, 0..255) rounded to float */
void or_hist_lut(const uint64_t* hs /*3*256*/, const uint64_t* hr /*3*256*/, float* lut /*3*256*/)
{
    for (int c = 0; c < 3; ++c) {
        double sc[256], rc[256];
        uint64_t a = 0, b = 0;
        for (int v = 0; v < 256; ++v) { a += hs[c * 256 + v]; b += hr[c * 256 + v]; sc[v] = (double)a; rc[v] = (double)b; }
        const double st = sc[255] + 1e-8, rt = rc[255] + 1e-8;
        for (int v = 0; v < 256; ++v) { sc[v] = sc[v] / st; rc[v] = rc[v] / rt; }
        for (int v = 0; v < 256; ++v) lut[c * 256 + v] = (float)np_interp_index(sc[v], rc, 256);
    }
}

int or_composite(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2, const double* H,
                 int cw, int ch, double ox, double oy, int n_levels, int color_match,
                 uint8_t* pano, uint8_t* valid, float* seam_out, float* gain_out, double* rmse, int* vertical_out)
{
    if (h1 < 2 || w1 < 2 || h2 < 1 || w2 < 1 || cw < 1 || ch < 1 || n_levels < 1) return -1;
    const size_t P = (size_t)cw * ch;
    uint8_t* wu = (uint8_t*)malloc(P * 3);
    uint8_t* m1 = (uint8_t*)malloc(P);
    uint8_t* c2 = (uint8_t*)calloc(P * 3, 1);
    uint8_t* m2 = (uint8_t*)calloc(P, 1);
    or_grid_warp(img1, h1, w1, H, cw, ch, ox, oy, wu, m1);
    const int iox = (int)rint(ox), ioy = (int)rint(oy);
    for (int y = 0; y < h2; ++y)
        for (int x = 0; x < w2; ++x) {
            const int u = x + iox, v = y + ioy;
            if (u < 0 || u >= cw || v < 0 || v >= ch) continue;
            for (int c = 0; c < 3; ++c) c2[((size_t)v * cw + u) * 3 + c] = img2[((size_t)y * w2 + x) * 3 + c];
            m2[(size_t)v * cw + u] = 1;
        }
    /* gain */
    uint64_t s1[3] = { 0, 0, 0 }, s2[3] = { 0, 0, 0 }, nov = 0;
    for (size_t p = 0; p < P; ++p)
        if (m1[p] && m2[p]) {
            nov++;
            for (int c = 0; c < 3; ++c) { s1[c] += wu[p * 3 + c]; s2[c] += c2[p * 3 + c]; }
        }
    float g[3] = { 1.0f, 1.0f, 1.0f };
    float* hlut = NULL;
    if (nov && color_match == 1) {
        uint64_t* hs = (uint64_t*)calloc(3 * 256, sizeof(uint64_t));
        uint64_t* hr = (uint64_t*)calloc(3 * 256, sizeof(uint64_t));
        for (size_t p = 0; p < P; ++p)
            if (m1[p] && m2[p])
                for (int c = 0; c < 3; ++c) { hs[c * 256 + wu[p * 3 + c]]++; hr[c * 256 + c2[p * 3 + c]]++; }
        hlut = (float*)malloc(3 * 256 * sizeof(float));
        or_hist_lut(hs, hr, hlut);
        free(hs); free(hr);
    }
    if (nov && color_match != 1) for (int c = 0; c < 3; ++c) {
        const float a = (float)((double)s1[c] / (double)nov), b = (float)((double)s2[c] / (double)nov);
        float t = (b + 1e-3f) / (a + 1e-3f);
        g[c] = t < 0.5f ? 0.5f : (t > 2.0f ? 2.0f : t);
    }
    if (gain_out) for (int c = 0; c < 3; ++c) gain_out[c] = g[c];
    float* w1g = (float*)malloc(P * 3 * sizeof(float));
    for (size_t p = 0; p < P; ++p)
        for (int c = 0; c < 3; ++c) {
            float t = (float)wu[p * 3 + c];
            if (hlut) t = hlut[c * 256 + wu[p * 3 + c]];
            else if (nov) { t = t * g[c]; t = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t); }
            w1g[p * 3 + c] = t;
        }
    /* seam mask */
    float* seam = (float*)calloc(P, sizeof(float));
    uint8_t* ov = (uint8_t*)malloc(P);
    for (size_t p = 0; p < P; ++p) { ov[p] = m1[p] && m2[p]; seam[p] = (m1[p] && !m2[p]) ? 1.0f : 0.0f; }
    int vertical = 1;
    if (nov) {
        float* avg = (float*)malloc(P * sizeof(float));
        float* cost = (float*)malloc(P * sizeof(float));
        float* gx = (float*)malloc(P * sizeof(float));
        float* gy = (float*)malloc(P * sizeof(float));
        for (size_t p = 0; p < P; ++p) {
            float d2[3], hv[3];
            for (int c = 0; c < 3; ++c) {
                const float a = w1g[p * 3 + c], b = (float)c2[p * 3 + c];
                const float d = a - b;
                d2[c] = d * d;
                hv[c] = (a + b) * 0.5f;
            }
            cost[p] = sqrtf((d2[0] + d2[1]) + d2[2]);
            avg[p] = ((hv[0] + hv[1]) + hv[2]) / 3.0f;
        }
        or_sobel(avg, ch, cw, 1, gx);
        or_sobel(avg, ch, cw, 0, gy);
        for (size_t p = 0; p < P; ++p) {
            const float gr = sqrtf(gx[p] * gx[p] + gy[p] * gy[p]);
            cost[p] = ov[p] ? cost[p] + 0.05f * gr : 1e9f;
        }
        /* orientation and side */
        long long nr = 0, nc = 0, s1c = 0, n1c = 0, s2c = 0, n2c = 0, s1r = 0, n1r = 0, s2r = 0, n2r = 0;
        for (int y = 0; y < ch; ++y) {
            int a = 0, b = 0, d = 0;
            for (int x = 0; x < cw; ++x) {
                const size_t p = (size_t)y * cw + x;
                a |= ov[p]; b |= m1[p] && !m2[p]; d |= m2[p] && !m1[p];
            }
            nr += a; if (b) { s1r += y; n1r++; } if (d) { s2r += y; n2r++; }
        }
        for (int x = 0; x < cw; ++x) {
            int a = 0, b = 0, d = 0;
            for (int y = 0; y < ch; ++y) {
                const size_t p = (size_t)y * cw + x;
                a |= ov[p]; b |= m1[p] && !m2[p]; d |= m2[p] && !m1[p];
            }
            nc += a; if (b) { s1c += x; n1c++; } if (d) { s2c += x; n2c++; }
        }
        const int left = n1c == 0 || n2c == 0 || (double)s1c / (double)n1c <= (double)s2c / (double)n2c;
        const int top = n1r == 0 || n2r == 0 || (double)s1r / (double)n1r <= (double)s2r / (double)n2r;
        vertical = nr >= nc;
        const int L = vertical ? ch : cw;
        int32_t* sl = (int32_t*)malloc((size_t)L * sizeof(int32_t));
        or_dp_seam(cost, ov, ch, cw, vertical, sl);
        const int first = vertical ? left : top;
        for (int l = 0; l < L; ++l) {
            const int N = vertical ? cw : ch;
            int any = 0;
            for (int n = 0; n < N && !any; ++n) any = vertical ? ov[(size_t)l * cw + n] : ov[(size_t)n * cw + l];
            if (!any) continue;
            for (int n = 0; n < N; ++n) {
                const size_t p = vertical ? (size_t)l * cw + n : (size_t)n * cw + l;
                if (!ov[p]) continue;
                seam[p] = (n <= sl[l]) == first ? 1.0f : 0.0f;
            }
        }
        free(sl);
        float w25[25];
        or_gauss_weights_cv(w25);
        float* bl = (float*)malloc(P * sizeof(float));
        or_cv_blur25(seam, ch, cw, w25, bl);
        for (size_t p = 0; p < P; ++p) if (ov[p]) seam[p] = bl[p];
        free(bl); free(avg); free(cost); free(gx); free(gy);
    }
    if (seam_out) memcpy(seam_out, seam, P * sizeof(float));
    /* multiband: planes 0-2 image 1, 3-5 image 2, 6 mask */
    float w9[9];
    or_gauss_weights_sp(w9);
    int hs[16], ws[16];
    float* gp[7][16];
    hs[0] = ch; ws[0] = cw;
    for (int k = 1; k < n_levels; ++k) { hs[k] = (hs[k - 1] + 1) / 2; ws[k] = (ws[k - 1] + 1) / 2; }
    for (int q = 0; q < 7; ++q) {
        gp[q][0] = (float*)malloc(P * sizeof(float));
        for (size_t p = 0; p < P; ++p) {
            float v;
            if (q == 6) v = seam[p];
            else {
                const int c = q % 3;
                const float a = w1g[p * 3 + c], b = (float)c2[p * 3 + c];
                if (q < 3) v = m1[p] ? a : (m2[p] ? b : 0.0f);
                else v = m2[p] ? b : (m1[p] ? a : 0.0f);
            }
            gp[q][0][p] = v;
        }
        for (int k = 1; k < n_levels; ++k) {
            gp[q][k] = (float*)malloc((size_t)hs[k] * ws[k] * sizeof(float));
            or_gauss_down(gp[q][k - 1], hs[k - 1], ws[k - 1], w9, gp[q][k]);
        }
    }
    float* res[3];
    for (int c = 0; c < 3; ++c) {
        const int kL = n_levels - 1;
        const size_t nL = (size_t)hs[kL] * ws[kL];
        res[c] = (float*)malloc(nL * sizeof(float));
        for (size_t p = 0; p < nL; ++p) {
            const float m = gp[6][kL][p];
            res[c][p] = m * gp[c][kL][p] + (1.0f - m) * gp[c + 3][kL][p];
        }
        for (int k = n_levels - 2; k >= 0; --k) {
            const size_t n = (size_t)hs[k] * ws[k];
            float* u1 = (float*)malloc(n * sizeof(float));
            float* u2 = (float*)malloc(n * sizeof(float));
            float* ur = (float*)malloc(n * sizeof(float));
            or_zoom_up(gp[c][k + 1], hs[k + 1], ws[k + 1], hs[k], ws[k], u1);
            or_zoom_up(gp[c + 3][k + 1], hs[k + 1], ws[k + 1], hs[k], ws[k], u2);
            or_zoom_up(res[c], hs[k + 1], ws[k + 1], hs[k], ws[k], ur);
            for (size_t p = 0; p < n; ++p) {
                const float m = gp[6][k][p];
                const float l1 = gp[c][k][p] - u1[p], l2 = gp[c + 3][k][p] - u2[p];
                const float lb = m * l1 + (1.0f - m) * l2;
                ur[p] = ur[p] + lb;
            }
            free(u1); free(u2); free(res[c]);
            res[c] = ur;
        }
    }
    for (size_t p = 0; p < P; ++p) {
        valid[p] = m1[p] || m2[p];
        for (int c = 0; c < 3; ++c) {
            float v = res[c][p];
            v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
            pano[p * 3 + c] = valid[p] ? (uint8_t)v : 0;
        }
    }
    /* seam RMSE over the overlap of the (gained) uint8 images */
    if (rmse) {
        uint64_t sq = 0;
        for (size_t p = 0; p < P; ++p)
            if (ov[p]) for (int c = 0; c < 3; ++c) {
                const int d = (int)(uint8_t)w1g[p * 3 + c] - (int)c2[p * 3 + c];
                sq += (uint64_t)(d * d);
            }
        *rmse = nov ? sqrt((double)sq / (double)(3 * nov)) : NAN;
    }
    if (vertical_out) *vertical_out = nov ? vertical : -1;
    for (int q = 0; q < 7; ++q) for (int k = 0; k < n_levels; ++k) free(gp[q][k]);
    for (int c = 0; c < 3; ++c) free(res[c]);
    free(wu); free(m1); free(c2); free(m2); free(w1g); free(seam); free(ov); free(hlut);
    return 0;
}

/* ========================================================================
 * Philox4x32-10 (Salmon et al., SC'11)
 * ======================================================================== */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void or_ransac_sample(int B, int N, uint64_t seed, int32_t* idx)
{
    /* hypothesis b draws counter (b,0,0,0) under key (seed lo, seed hi);
     * index = (u * N) >> 32 (with replacement, like torch.randint) */
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    for (int b = 0; b < B; ++b) {
        uint32_t ctr[4] = { (uint32_t)b, 0u, 0u, 0u }, o[4];
        or_philox4x32_10(ctr, key, o);
        for (int k = 0; k < 4; ++k) idx[4 * b + k] = (int32_t)(((uint64_t)o[k] * (uint64_t)N) >> 32);
    }
}

/* ========================================================================
 * Homography (A14-A16, core.py:2946-2981, 2272-2313, 3011-3097)
 * ======================================================================== */
static double wave_sum_d(const double* v, int n)
{
    double p[LANES];
    for (int l = 0; l < LANES; ++l) p[l] = 0.0;
    for (int i = 0; i < n; ++i) p[i % LANES] = p[i % LANES] + v[i];
    return wave_tree_d(p);
}

/* Hartley normalisation (core.py:2946-2957): centroid, mean distance md,
 * s = sqrt(2)/(md + 1e-10).  Sums use the 64-lane order. */
static void hartley(const double* pts, int N, const uint8_t* sel, double* T)
{
    double* buf = (double*)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
    int m = 0;
    for (int i = 0; i < N; ++i) if (!sel || sel[i]) buf[m++] = pts[2 * i];
    double cx = wave_sum_d(buf, m) / (double)m;
    m = 0;
    for (int i = 0; i < N; ++i) if (!sel || sel[i]) buf[m++] = pts[2 * i + 1];
    double cy = wave_sum_d(buf, m) / (double)m;
    m = 0;
    for (int i = 0; i < N; ++i) if (!sel || sel[i]) {
        double dx = pts[2 * i] - cx, dy = pts[2 * i + 1] - cy;
        buf[m++] = sqrt(dx * dx + dy * dy);
    }
    double md = wave_sum_d(buf, m) / (double)m;
    double s = 0x1.6a09e667f3bcdp+0 / (md + 1e-10);
    T[0] = s; T[1] = 0.0; T[2] = -s * cx;
    T[3] = 0.0; T[4] = s; T[5] = -s * cy;
    T[6] = 0.0; T[7] = 0.0; T[8] = 1.0;
    free(buf);
}

/* H = inv(T2) * Hn * T1, then / H[2][2] when |H22| > 1e-10 */
static void denorm(const double* Hn, const double* T1, const double* T2, double* H)
{
    double s2 = T2[0], inv_s2 = 1.0 / s2;
    double cx2 = -T2[2] / s2, cy2 = -T2[5] / s2;
    double Ti[9] = { inv_s2, 0.0, cx2, 0.0, inv_s2, cy2, 0.0, 0.0, 1.0 };
    double A[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            A[3 * r + c] = (Ti[3 * r] * Hn[c] + Ti[3 * r + 1] * Hn[3 + c]) + Ti[3 * r + 2] * Hn[6 + c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            H[3 * r + c] = (A[3 * r] * T1[c] + A[3 * r + 1] * T1[3 + c]) + A[3 * r + 2] * T1[6 + c];
    double h22 = H[8];
    if (fabs(h22) > 1e-10) for (int i = 0; i < 9; ++i) H[i] = H[i] / h22;
}

/* Minimal 4-point solve: null vector of the 8x9 DLT matrix by Gaussian
 * elimination with partial pivoting (first max on ties), h8 = 1. */
static int solve8(double A[8][9], double* h)
{
    double amax = 0.0;
    for (int r = 0; r < 8; ++r) for (int c = 0; c < 9; ++c) if (fabs(A[r][c]) > amax) amax = fabs(A[r][c]);
    double tiny = amax * 1e-12;
    for (int c = 0; c < 8; ++c) {
        int p = c;
        double best = fabs(A[c][c]);
        for (int r = c + 1; r < 8; ++r) if (fabs(A[r][c]) > best) { best = fabs(A[r][c]); p = r; }
        if (!(best > tiny)) return 0;
        if (p != c) for (int k = 0; k < 9; ++k) { double t = A[c][k]; A[c][k] = A[p][k]; A[p][k] = t; }
        for (int r = c + 1; r < 8; ++r) {
            double f = A[r][c] / A[c][c];
            for (int k = c; k < 9; ++k) A[r][k] = A[r][k] - f * A[c][k];
        }
    }
    h[8] = 1.0;
    for (int c = 7; c >= 0; --c) {
        double s = A[c][8];
        for (int k = c + 1; k < 8; ++k) s = s + A[c][k] * h[k];
        h[c] = -s / A[c][c];
    }
    return 1;
}

static void dlt_rows(double x1, double y1, double x2, double y2, double* re, double* ro)
{
    re[0] = -x1; re[1] = -y1; re[2] = -1.0; re[3] = 0.0; re[4] = 0.0; re[5] = 0.0;
    re[6] = x2 * x1; re[7] = x2 * y1; re[8] = x2;
    ro[0] = 0.0; ro[1] = 0.0; ro[2] = 0.0; ro[3] = -x1; ro[4] = -y1; ro[5] = -1.0;
    ro[6] = y2 * x1; ro[7] = y2 * y1; ro[8] = y2;
}

/* smallest-eigenvalue eigenvector of a symmetric 9x9 by cyclic Jacobi */
static void jacobi9(double* M, double* vec)
{
    double V[81];
    for (int i = 0; i < 81; ++i) V[i] = (i % 10 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < 9; ++i) for (int j = 0; j < 9; ++j) {
            if (i != j) off = off + M[9 * i + j] * M[9 * i + j];
            else diag = diag + M[9 * i + j] * M[9 * i + j];
        }
        if (off <= diag * 1e-30) break;
        for (int p = 0; p < 8; ++p)
            for (int q = p + 1; q < 9; ++q) {
                double apq = M[9 * p + q];
                if (apq == 0.0) continue;
                double app = M[9 * p + p], aqq = M[9 * q + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) t = -t;
                double c = 1.0 / sqrt(t * t + 1.0);
                double s = t * c;
                for (int k = 0; k < 9; ++k) {
                    double mkp = M[9 * k + p], mkq = M[9 * k + q];
                    M[9 * k + p] = c * mkp - s * mkq;
                    M[9 * k + q] = s * mkp + c * mkq;
                }
                for (int k = 0; k < 9; ++k) {
                    double mpk = M[9 * p + k], mqk = M[9 * q + k];
                    M[9 * p + k] = c * mpk - s * mqk;
                    M[9 * q + k] = s * mpk + c * mqk;
                }
                for (int k = 0; k < 9; ++k) {
                    double vkp = V[9 * k + p], vkq = V[9 * k + q];
                    V[9 * k + p] = c * vkp - s * vkq;
                    V[9 * k + q] = s * vkp + c * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 9; ++i) if (M[9 * i + i] < M[9 * best + best]) best = i;
    for (int k = 0; k < 9; ++k) vec[k] = V[9 * k + best];
}

/* least-squares DLT on the selected points (core.py:2959-2981): null vector
 * of A^T A (accumulated in point order, 64-lane partials per entry). */
static int dlt_ls_sel(const double* p1, const double* p2, int N, const uint8_t* sel, double* H)
{
    int m = 0;
    for (int i = 0; i < N; ++i) if (!sel || sel[i]) m++;
    if (m < 4) return 0;
    double T1[9], T2[9];
    hartley(p1, N, sel, T1);
    hartley(p2, N, sel, T2);
    double part[LANES][45];
    memset(part, 0, sizeof(part));
    int k = 0;
    for (int i = 0; i < N; ++i) {
        if (sel && !sel[i]) continue;
        double x1 = T1[0] * p1[2 * i] + T1[2], y1 = T1[4] * p1[2 * i + 1] + T1[5];
        double x2 = T2[0] * p2[2 * i] + T2[2], y2 = T2[4] * p2[2 * i + 1] + T2[5];
        double re[9], ro[9];
        dlt_rows(x1, y1, x2, y2, re, ro);
        int lane = k % LANES, e = 0;
        for (int a = 0; a < 9; ++a)
            for (int b = a; b < 9; ++b) {
                part[lane][e] = part[lane][e] + (re[a] * re[b] + ro[a] * ro[b]);
                e++;
            }
        k++;
    }
    double M[81];
    int e = 0;
    for (int a = 0; a < 9; ++a)
        for (int b = a; b < 9; ++b) {
            double t[LANES];
            for (int l = 0; l < LANES; ++l) t[l] = part[l][e];
            double v = wave_tree_d(t);
            M[9 * a + b] = v; M[9 * b + a] = v;
            e++;
        }
    double h[9];
    jacobi9(M, h);
    denorm(h, T1, T2, H);
    return 1;
}

int or_dlt_ls(const double* p1, const double* p2, int N, double* H)
{
    return dlt_ls_sel(p1, p2, N, NULL, H) ? 0 : -1;
}

static double sym_err(const double* H, const double* Hi, double x1, double y1, double x2, double y2, int clampz)
{
    double fx = (H[0] * x1 + H[1] * y1) + H[2];
    double fy = (H[3] * x1 + H[4] * y1) + H[5];
    double fz = (H[6] * x1 + H[7] * y1) + H[8];
    if (clampz && fz < 1e-10) fz = 1e-10;
    double ex = fx / fz - x2, ey = fy / fz - y2;
    double bx = (Hi[0] * x2 + Hi[1] * y2) + Hi[2];
    double by = (Hi[3] * x2 + Hi[4] * y2) + Hi[5];
    double bz = (Hi[6] * x2 + Hi[7] * y2) + Hi[8];
    if (clampz && bz < 1e-10) bz = 1e-10;
    double gx = bx / bz - x1, gy = by / bz - y1;
    return (ex * ex + ey * ey) + (gx * gx + gy * gy);
}

/* adjugate inverse; det replaced by 1 when |det| <= 1e-12 (core.py:2298-2307) */
static void inv3(const double* a, double* o)
{
    double det = (a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6])) + a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (!(fabs(det) > 1e-12)) det = 1.0;
    o[0] = (a[4] * a[8] - a[5] * a[7]) / det; o[1] = (a[2] * a[7] - a[1] * a[8]) / det; o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    o[3] = (a[5] * a[6] - a[3] * a[8]) / det; o[4] = (a[0] * a[8] - a[2] * a[6]) / det; o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    o[6] = (a[3] * a[7] - a[4] * a[6]) / det; o[7] = (a[1] * a[6] - a[0] * a[7]) / det; o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

int or_ransac(const double* p1, const double* p2, int N, uint64_t seed, int B,
              double thr, double sigma_max, int method, int lo, int min_inliers,
              double* Hout, uint8_t* mask, double* scores, double* Hs)
{
    if (N < 4) return -1;
    double tq = thr * thr;
    double T1[9], T2[9];
    hartley(p1, N, NULL, T1);
    hartley(p2, N, NULL, T2);
    double* n1 = (double*)malloc(sizeof(double) * 2 * N);
    double* n2 = (double*)malloc(sizeof(double) * 2 * N);
    for (int i = 0; i < N; ++i) {
        n1[2 * i] = T1[0] * p1[2 * i] + T1[2]; n1[2 * i + 1] = T1[4] * p1[2 * i + 1] + T1[5];
        n2[2 * i] = T2[0] * p2[2 * i] + T2[2]; n2[2 * i + 1] = T2[4] * p2[2 * i + 1] + T2[5];
    }
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * 4 * B);
    or_ransac_sample(B, N, seed, idx);
    double* sc = (double*)malloc(sizeof(double) * B);
    double* HB = (double*)malloc(sizeof(double) * 9 * B);
    double den = 2.0 * (sigma_max * sigma_max);
    #pragma omp parallel for schedule(dynamic, 8)
    for (int b = 0; b < B; ++b) {
        double A[8][9], h[9], H[9], Hi[9];
        for (int k = 0; k < 4; ++k) {
            int i = idx[4 * b + k];
            dlt_rows(n1[2 * i], n1[2 * i + 1], n2[2 * i], n2[2 * i + 1], A[2 * k], A[2 * k + 1]);
        }
        if (!solve8(A, h)) {
            sc[b] = -1.0;
            for (int k = 0; k < 9; ++k) HB[9 * b + k] = (k % 4 == 0) ? 1.0 : 0.0;
            continue;
        }
        denorm(h, T1, T2, H);
        inv3(H, Hi);
        double part[LANES];
        for (int l = 0; l < LANES; ++l) part[l] = 0.0;
        for (int i = 0; i < N; ++i) {
            double e = sym_err(H, Hi, p1[2 * i], p1[2 * i + 1], p2[2 * i], p2[2 * i + 1], 1);
            double w = (method == 0) ? or_exp(-e / den) : (e < tq ? 1.0 : 0.0);
            part[i % LANES] = part[i % LANES] + w;
        }
        sc[b] = wave_tree_d(part);
        memcpy(HB + 9 * b, H, sizeof(H));
    }
    int best = 0;
    for (int b = 1; b < B; ++b) if (sc[b] > sc[best]) best = b;
    double Hb[9], Hbi[9];
    memcpy(Hb, HB + 9 * best, sizeof(Hb));
    inv3(Hb, Hbi);
    int nb = 0;
    for (int i = 0; i < N; ++i) {
        double e = sym_err(Hb, Hbi, p1[2 * i], p1[2 * i + 1], p2[2 * i], p2[2 * i + 1], 1);
        mask[i] = e < tq;
        nb += mask[i];
    }
    memcpy(Hout, Hb, sizeof(Hb));
    if (lo && nb >= min_inliers) {
        double H1[9], H1i[9];
        dlt_ls_sel(p1, p2, N, mask, H1);
        inv3(H1, H1i);
        uint8_t* m2 = (uint8_t*)malloc(N);
        int n2c = 0;
        for (int i = 0; i < N; ++i) {
            double e = sym_err(H1, H1i, p1[2 * i], p1[2 * i + 1], p2[2 * i], p2[2 * i + 1], 0);
            m2[i] = e < tq;
            n2c += m2[i];
        }
        if (n2c >= min_inliers) {
            dlt_ls_sel(p1, p2, N, m2, Hout);
            memcpy(mask, m2, N);
        } else {
            memcpy(Hout, H1, sizeof(H1));
        }
        free(m2);
    }
    if (scores) memcpy(scores, sc, sizeof(double) * B);
    if (Hs) memcpy(Hs, HB, sizeof(double) * 9 * B);
    free(n1); free(n2); free(idx); free(sc); free(HB);
    return best;
}

/* ========================================================================
 * A17 + feather blend.  Canvas (core.py:3171-3187).  Warp: inverse map with
 * H^-1 rounded to float32 (core.py:3225), bilinear with zero padding
 * (core.py:3233), valid where the source point lies in [0,w-1]x[0,h-1].
 * img2 is pasted at (round(ox), round(oy)) (core.py:3794-3800, Python round =
 * half to even).  Feather
 * weight of a valid pixel = min(sx+1, w-sx, sy+1, h-sy) in its own source
 * coordinates; output = trunc(clip(sum w_i I_i / sum w_i)).
 * ======================================================================== */
void or_canvas(int h1, int w1, int h2, int w2, const double* H,
               int* cw, int* ch, double* ox, double* oy)
{
    double cx[4] = { 0, (double)w1, (double)w1, 0 }, cy[4] = { 0, 0, (double)h1, (double)h1 };
    double ax[8], ay[8];
    for (int i = 0; i < 4; ++i) {
        double X = (H[0] * cx[i] + H[1] * cy[i]) + H[2];
        double Y = (H[3] * cx[i] + H[4] * cy[i]) + H[5];
        double Z = (H[6] * cx[i] + H[7] * cy[i]) + H[8];
        ax[i] = X / Z; ay[i] = Y / Z;
    }
    ax[4] = 0; ax[5] = w2; ax[6] = w2; ax[7] = 0;
    ay[4] = 0; ay[5] = 0; ay[6] = h2; ay[7] = h2;
    double mnx = ax[0], mxx = ax[0], mny = ay[0], mxy = ay[0];
    for (int i = 1; i < 8; ++i) {
        if (ax[i] < mnx) mnx = ax[i];
        if (ax[i] > mxx) mxx = ax[i];
        if (ay[i] < mny) mny = ay[i];
        if (ay[i] > mxy) mxy = ay[i];
    }
    *ox = -mnx > 0.0 ? -mnx : 0.0;
    *oy = -mny > 0.0 ? -mny : 0.0;
    *cw = (int)ceil(mxx + *ox) + 1;
    *ch = (int)ceil(mxy + *oy) + 1;
}

void or_warp_feather(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2,
                     const double* H, int cw, int ch, double ox, double oy,
                     uint8_t* pano, uint8_t* valid)
{
    double Hi[9];
    inv3(H, Hi);
    float hf[9];
    for (int i = 0; i < 9; ++i) hf[i] = (float)Hi[i];
    float oxf = (float)ox, oyf = (float)oy;
    int iox = (int)rint(ox), ioy = (int)rint(oy);   /* Python round(): half to even */
    #pragma omp parallel for schedule(static)
    for (int v = 0; v < ch; ++v)
        for (int u = 0; u < cw; ++u) {
            float X = (float)u - oxf, Y = (float)v - oyf;
            float sxn = (hf[0] * X + hf[1] * Y) + hf[2];
            float syn = (hf[3] * X + hf[4] * Y) + hf[5];
            float szn = (hf[6] * X + hf[7] * Y) + hf[8];
            if (szn < 1e-10f) szn = 1e-10f;
            float sx = sxn / szn, sy = syn / szn;
            float acc[3] = { 0, 0, 0 }, wsum = 0.0f;
            if (sx >= 0.0f && sx <= (float)(w1 - 1) && sy >= 0.0f && sy <= (float)(h1 - 1)) {
                float fx = floorf(sx), fy = floorf(sy);
                int x0 = (int)fx, y0 = (int)fy;
                float ax = sx - fx, ay = sy - fy;
                float w00 = (1.0f - ax) * (1.0f - ay), w01 = ax * (1.0f - ay);
                float w10 = (1.0f - ax) * ay, w11 = ax * ay;
                float wf = sx + 1.0f;
                if ((float)w1 - sx < wf) wf = (float)w1 - sx;
                if (sy + 1.0f < wf) wf = sy + 1.0f;
                if ((float)h1 - sy < wf) wf = (float)h1 - sy;
                for (int c = 0; c < 3; ++c) {
                    float s = 0.0f;
                    float v00 = img1[((size_t)y0 * w1 + x0) * 3 + c];
                    float v01 = (x0 + 1 < w1) ? img1[((size_t)y0 * w1 + x0 + 1) * 3 + c] : 0.0f;
                    float v10 = (y0 + 1 < h1) ? img1[((size_t)(y0 + 1) * w1 + x0) * 3 + c] : 0.0f;
                    float v11 = (x0 + 1 < w1 && y0 + 1 < h1) ? img1[((size_t)(y0 + 1) * w1 + x0 + 1) * 3 + c] : 0.0f;
                    s = ((w00 * v00 + w01 * v01) + w10 * v10) + w11 * v11;
                    acc[c] = wf * s;
                }
                wsum = wf;
            }
            int px = u - iox, py = v - ioy;
            if (px >= 0 && px < w2 && py >= 0 && py < h2) {
                float wf = (float)(px + 1);
                if ((float)(w2 - px) < wf) wf = (float)(w2 - px);
                if ((float)(py + 1) < wf) wf = (float)(py + 1);
                if ((float)(h2 - py) < wf) wf = (float)(h2 - py);
                for (int c = 0; c < 3; ++c) acc[c] = acc[c] + wf * (float)img2[((size_t)py * w2 + px) * 3 + c];
                wsum = wsum + wf;
            }
            size_t o = (size_t)v * cw + u;
            valid[o] = wsum > 0.0f;
            for (int c = 0; c < 3; ++c) {
                float val = wsum > 0.0f ? acc[c] / wsum : 0.0f;
                if (val < 0.0f) val = 0.0f;
                if (val > 255.0f) val = 255.0f;
                pano[o * 3 + c] = (uint8_t)(int)val;
            }
        }
}

void or_bgr_to_gray(const uint8_t* bgr, int H, int W, uint8_t* gray)
{
    /* fixed-point BT.601 (B*1868 + G*9617 + R*4899 + 2^13) >> 14 */
    for (size_t i = 0; i < (size_t)H * W; ++i) {
        int b = bgr[3 * i], g = bgr[3 * i + 1], r = bgr[3 * i + 2];
        gray[i] = (uint8_t)((b * 1868 + g * 9617 + r * 4899 + 8192) >> 14);
    }
}

int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ========================================================================
 * Single-stage hooks used by the oracle's own pinning tests.
 * ======================================================================== */
int or_extrema_octave(const float* D, int S, int H, int W, float thr, int border, int cap, int* out)
{
    pyramid_t P;
    memset(&P, 0, sizeof(P));
    P.n_oct = 1; P.S = S; P.Ho[0] = H; P.Wo[0] = W; P.D[0] = (float*)D;
    cand_t* c = NULL;
    int n = find_extrema(&P, thr, border, cap, &c);
    for (int i = 0; i < n; ++i) { out[3 * i] = c[i].s; out[3 * i + 1] = c[i].r; out[3 * i + 2] = c[i].c; }
    free(c);
    return n;
}

int or_refine_octave(const float* D, int S, int H, int W, int border, float thr, float edge,
                     const int* cands, int n, int* valid, float* out)
{
    pyramid_t P;
    memset(&P, 0, sizeof(P));
    P.n_oct = 1; P.S = S; P.Ho[0] = H; P.Wo[0] = W; P.D[0] = (float*)D;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        cand_t cd = { 0, cands[3 * i], cands[3 * i + 1], cands[3 * i + 2] };
        kpt_t k;
        valid[i] = refine_one(&P, cd, border, (double)thr, (double)edge, &k);
        if (valid[i]) {
            out[4 * i] = k.ri; out[4 * i + 1] = k.ci; out[4 * i + 2] = (float)k.si; out[4 * i + 3] = k.resp;
            m++;
        }
    }
    return m;
}

int or_orientation(const float* G, int H, int W, float ri, float ci, float sigma, float* angles)
{
    return orientation_one(G, H, W, ri, ci, sigma, angles);
}

void or_descriptor(const float* G, int H, int W, float ri, float ci, float ang_rad, int ph, int rootsift,
                   float* out, uint8_t* out_u8)
{
    descriptor_one(G, H, W, ri, ci, ang_rad, ph, rootsift, 0, out, out_u8);
}

void or_descriptor_dsp(const float* G, int H, int W, float ri, float ci, float ang_rad, int ph, int rootsift,
                       int dsp_n_scales, float* out, uint8_t* out_u8)
{
    descriptor_one(G, H, W, ri, ci, ang_rad, ph, rootsift, dsp_n_scales, out, out_u8);
}
