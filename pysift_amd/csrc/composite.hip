// The reference's composition of a stitched pair (core.py:3686-3714,
// 3783-3802): grid_sample warp of image 1 (core.py:3211-3239), paste of
// image 2, gain compensation over the overlap, a DP seam through a colour +
// gradient cost (core.py:3287-3472), a Gaussian-softened seam mask, and a
// Laplacian-pyramid (multiband) blend (core.py:3475-3553).  Every stage runs
// on the device; the decisions that depend on whole-image reductions (gain,
// seam orientation and side) are made by a one-block kernel and read by the
// later kernels, so the host never waits in the middle.
//
// Arithmetic (the CPU oracle reproduces it bit for bit):
//   warp: X = u - (float)ox; source = ((h0 X + h1 Y) + h2) / max(z, 1e-10);
//     g = 2 s / (n - 1) - 1, valid when -1 <= g <= 1; i = ((g + 1) / 2)(n - 1);
//     grid_sample corner weights nw = (x_se - x)(y_se - y), ne = (x - x_sw)
//     (y_sw - y), sw = (x_ne - x)(y - y_ne), se = (x - x_nw)(y - y_nw), summed
//     nw, ne, sw, se over in-bounds corners; clipped and truncated to u8;
//   gain: exact integer sums over the overlap, m = (float)(sum / n),
//     g = clamp((m2 + 1e-3f) / (m1 + 1e-3f), 0.5, 2), w1 = clamp(w1 g, 0, 255);
//     or (color_match 'hist') w1 = lut[w1] from the overlap CDFs (np.interp rules);
//   cost: sqrtf((d0^2 + d1^2) + d2^2) + 0.05f sqrtf(gx^2 + gy^2), Sobel of the
//     channel mean of (w1 + c2) * 0.5 with SciPy's passes (derivative, then
//     [1, 2, 1]; 'reflect'; double accumulation); 1e9 outside the overlap;
//   seam: dynamic programme in double over the lines holding overlap, first
//     minimum on ties; vertical when the overlap spans at least as many rows
//     as columns; the mask is 1 on image 1's side; then a 25-tap float32
//     Gaussian (sigma 3, BORDER_REFLECT_101, rows then columns) replaces it
//     inside the overlap;
//   pyramid: sigma-1 Gaussian (float32 taps, fmaf in tap order, 'reflect',
//     rows then columns) evaluated only at the kept [::2, ::2] samples;
//     bilinear up-sampling with align-corners coordinates and weights in
//     double; blend m l1 + (1 - m) l2 and reconstruction in float32;
//     output clipped and truncated inside the union of the two images.
#include "msift_internal.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#pragma clang fp contract(off)

namespace msift {

namespace {

constexpr int kCT = 256;
constexpr int kMaxLevels = 12;

// device-side accumulators and decisions; the layout of msift_composite_stats
typedef msift_composite_stats CmpStats;
static_assert(sizeof(CmpStats) == 104, "msift_composite_stats layout");

struct Hf9 { float h[9]; };
struct W9 { float w[9]; };
struct W25 { float w[25]; };

__device__ __forceinline__ int refl_half(int p, int n) {
    if ((unsigned)p < (unsigned)n) return p;
    if (n == 1) return 0;
    const int period = 2 * n;
    p %= period;
    if (p < 0) p += period;
    return p < n ? p : period - 1 - p;
}

__device__ __forceinline__ int refl_101(int p, int n) {
    if ((unsigned)p < (unsigned)n) return p;
    if (n == 1) return 0;
    const int period = 2 * n - 2;
    p %= period;
    if (p < 0) p += period;
    return p < n ? p : period - p;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long* sh) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
    for (int k = 0; k < kCT / 64; ++k) t += sh[k];
    __syncthreads();
    return t;
}

// grid (ceil(cw / 256), ch): warp image 1, paste image 2; each block adds
// its overlap sums to its row's partials (rowacc[v][0..6]: n, s1[3], s2[3])
// and ORs its flags into rowflags[v] (bit 0 overlap, bit 1 image 1 only,
// bit 2 image 2 only)
__global__ __launch_bounds__(kCT) void k_cmp_warp(const uint8_t* __restrict__ img1, int h1, int w1,
                                                  const uint8_t* __restrict__ img2, int h2, int w2, Hf9 hf, int cw,
                                                  float oxf, float oyf, int iox, int ioy, uint8_t* __restrict__ wu,
                                                  uint8_t* __restrict__ c2, uint8_t* __restrict__ m1,
                                                  uint8_t* __restrict__ m2, unsigned long long* __restrict__ rowacc,
                                                  uint32_t* __restrict__ rowflags) {
    __shared__ unsigned long long sh[kCT / 64];
    const int u = blockIdx.x * kCT + threadIdx.x, v = blockIdx.y;
    unsigned long long acc7[7] = {0, 0, 0, 0, 0, 0, 0};
    uint32_t bits = 0;
    if (u < cw) {
        const size_t o = (size_t)v * cw + u;
        const float wm = (float)(w1 - 1), hm = (float)(h1 - 1);
        const float X = (float)u - oxf, Y = (float)v - oyf;
        const float sxn = (hf.h[0] * X + hf.h[1] * Y) + hf.h[2];
        const float syn = (hf.h[3] * X + hf.h[4] * Y) + hf.h[5];
        float szn = (hf.h[6] * X + hf.h[7] * Y) + hf.h[8];
        if (szn < 1e-10f) szn = 1e-10f;
        const float sx = sxn / szn, sy = syn / szn;
        const float gx = (2.0f * sx) / wm - 1.0f, gy = (2.0f * sy) / hm - 1.0f;
        const bool in1 = gx >= -1.0f && gx <= 1.0f && gy >= -1.0f && gy <= 1.0f;
        m1[o] = in1;
        const float ix = ((gx + 1.0f) / 2.0f) * wm, iy = ((gy + 1.0f) / 2.0f) * hm;
        const float fx = floorf(ix), fy = floorf(iy);
        const float nw = ((fx + 1.0f) - ix) * ((fy + 1.0f) - iy), ne = (ix - fx) * ((fy + 1.0f) - iy);
        const float sw = ((fx + 1.0f) - ix) * (iy - fy), se = (ix - fx) * (iy - fy);
        const bool finite = fabsf(ix) < 1073741824.0f && fabsf(iy) < 1073741824.0f;
        const int x0 = finite ? (int)fx : -2, y0 = finite ? (int)fy : -2;
        const bool bx0 = x0 >= 0 && x0 < w1, bx1 = x0 + 1 >= 0 && x0 + 1 < w1;
        const bool by0 = y0 >= 0 && y0 < h1, by1 = y0 + 1 >= 0 && y0 + 1 < h1;
        const int px = u - iox, py = v - ioy;
        const bool in2 = px >= 0 && px < w2 && py >= 0 && py < h2;
        m2[o] = in2;
        bits = (in1 && in2 ? 1u : 0u) | (in1 && !in2 ? 2u : 0u) | (!in1 && in2 ? 4u : 0u);
        acc7[0] = in1 && in2;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float acc = 0.0f;
            if (finite) {
                if (bx0 && by0) acc = acc + (float)img1[((size_t)y0 * w1 + x0) * 3 + c] * nw;
                if (bx1 && by0) acc = acc + (float)img1[((size_t)y0 * w1 + x0 + 1) * 3 + c] * ne;
                if (bx0 && by1) acc = acc + (float)img1[((size_t)(y0 + 1) * w1 + x0) * 3 + c] * sw;
                if (bx1 && by1) acc = acc + (float)img1[((size_t)(y0 + 1) * w1 + x0 + 1) * 3 + c] * se;
            }
            acc = acc < 0.0f ? 0.0f : (acc > 255.0f ? 255.0f : acc);
            const uint8_t w8 = (uint8_t)acc;
            const uint8_t q8 = in2 ? img2[((size_t)py * w2 + px) * 3 + c] : 0;
            wu[o * 3 + c] = w8;
            c2[o * 3 + c] = q8;
            if (in1 && in2) { acc7[1 + c] = w8; acc7[4 + c] = q8; }
        }
    }
    const unsigned long long n = block_sum_u64(acc7[0], sh);
    if (n) {
        if (threadIdx.x == 0) atomicAdd(&rowacc[(size_t)v * 8], n);
        for (int k = 1; k < 7; ++k) {
            const unsigned long long t = block_sum_u64(acc7[k], sh);
            if (threadIdx.x == 0) atomicAdd(&rowacc[(size_t)v * 8 + k], t);
        }
    }
    uint32_t wb = bits;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wb |= __shfl_xor(wb, off, 64);
    if ((threadIdx.x & 63) == 0 && wb) atomicOr(&rowflags[v], wb);
}

// grid (ceil(cw / 256), ceil(ch / 64)): a thread ORs 64 rows of its column's
// flags into colflags[u]
__global__ __launch_bounds__(kCT) void k_cmp_colflags(const uint8_t* __restrict__ m1, const uint8_t* __restrict__ m2,
                                                      int cw, int ch, uint32_t* __restrict__ colflags) {
    const int u = blockIdx.x * kCT + threadIdx.x;
    if (u >= cw) return;
    const int v0 = blockIdx.y * 64, v1 = min(ch, v0 + 64);
    uint32_t bits = 0;
    for (int v = v0; v < v1; ++v) {
        const size_t o = (size_t)v * cw + u;
        const bool a = m1[o], b = m2[o];
        bits |= (a && b ? 1u : 0u) | (a && !b ? 2u : 0u) | (!a && b ? 4u : 0u);
    }
    if (bits) atomicOr(&colflags[u], bits);
}

// color_match 'hist': overlap histograms of the warped image 1 (hist[0..2])
// and of image 2 (hist[3..5]); per-block LDS counts, one atomic per bin
__global__ __launch_bounds__(kCT) void k_cmp_hist(const uint8_t* __restrict__ wu, const uint8_t* __restrict__ c2,
                                                  const uint8_t* __restrict__ m1, const uint8_t* __restrict__ m2,
                                                  size_t P, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[6 * 256];
    for (int i = threadIdx.x; i < 6 * 256; i += kCT) h[i] = 0;
    __syncthreads();
    for (size_t p = (size_t)blockIdx.x * kCT + threadIdx.x; p < P; p += (size_t)gridDim.x * kCT) {
        if (!(m1[p] && m2[p])) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            atomicAdd(&h[c * 256 + wu[p * 3 + c]], 1u);
            atomicAdd(&h[(3 + c) * 256 + c2[p * 3 + c]], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 6 * 256; i += kCT)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// np.interp(x, xp, 0..n-1) with NumPy's rules (see the oracle's restatement)
__device__ double np_interp_index(double x, const double* xp, int n) {
    if (x < xp[0]) return 0.0;
    if (x > xp[n - 1]) return (double)(n - 1);
    int lo = 0, hi = n;
    while (lo < hi) { const int mid = (lo + hi) >> 1; if (x >= xp[mid]) lo = mid + 1; else hi = mid; }
    const int j = lo - 1;
    if (j == n - 1) return (double)(n - 1);
    if (xp[j] == x) return (double)j;
    const double slope = ((double)(j + 1) - (double)j) / (xp[j + 1] - xp[j]);
    return slope * (x - xp[j]) + (double)j;
}

// one block: gain, seam orientation and side (core.py:3419-3431, 3696)
__global__ __launch_bounds__(kCT) void k_cmp_decide(const uint32_t* __restrict__ rowflags, int ch,
                                                    const uint32_t* __restrict__ colflags, int cw,
                                                    const unsigned long long* __restrict__ rowacc,
                                                    int color_match, const uint32_t* __restrict__ hist,
                                                    float* __restrict__ lut, CmpStats* __restrict__ st) {
    __shared__ unsigned long long sh[kCT / 64];
    __shared__ double sc[256], rc[256];
    {
        unsigned long long t7[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int v = threadIdx.x; v < ch; v += kCT)
            for (int k = 0; k < 7; ++k) t7[k] += rowacc[(size_t)v * 8 + k];
        for (int k = 0; k < 7; ++k) t7[k] = block_sum_u64(t7[k], sh);
        if (threadIdx.x == 0) {
            st->n_overlap = t7[0];
            for (int c = 0; c < 3; ++c) { st->sum1[c] = t7[1 + c]; st->sum2[c] = t7[4 + c]; }
        }
        __syncthreads();
    }
    unsigned long long r[5] = {0, 0, 0, 0, 0}, c[5] = {0, 0, 0, 0, 0};   // n_ov, sum1, n1, sum2, n2
    for (int v = threadIdx.x; v < ch; v += kCT) {
        const uint32_t f = rowflags[v];
        r[0] += f & 1u;
        if (f & 2u) { r[1] += v; r[2]++; }
        if (f & 4u) { r[3] += v; r[4]++; }
    }
    for (int u = threadIdx.x; u < cw; u += kCT) {
        const uint32_t f = colflags[u];
        c[0] += f & 1u;
        if (f & 2u) { c[1] += u; c[2]++; }
        if (f & 4u) { c[3] += u; c[4]++; }
    }
    for (int k = 0; k < 5; ++k) { r[k] = block_sum_u64(r[k], sh); c[k] = block_sum_u64(c[k], sh); }
    if (threadIdx.x == 0) {
        const unsigned long long n = st->n_overlap;
        st->has_overlap = n > 0;
        for (int k = 0; k < 3; ++k) {
            float g = 1.0f;   // 1 when no gain is applied (no overlap, or 'hist')
            if (n && color_match == 0) {
                const float a = (float)((double)st->sum1[k] / (double)n), b = (float)((double)st->sum2[k] / (double)n);
                g = (b + 1e-3f) / (a + 1e-3f);
                g = g < 0.5f ? 0.5f : (g > 2.0f ? 2.0f : g);
            }
            st->gain[k] = g;
        }
        const bool left = c[2] == 0 || c[4] == 0 || (double)c[1] / (double)c[2] <= (double)c[3] / (double)c[4];
        const bool top = r[2] == 0 || r[4] == 0 || (double)r[1] / (double)r[2] <= (double)r[3] / (double)r[4];
        st->vertical = r[0] >= c[0];
        st->img1_first = st->vertical ? left : top;
        st->sq_diff = 0;
    }
    // color_match 'hist' (core.py:3591-3626): per channel lut[v] =
    // interp(s_cdf[v], r_cdf, 0..255), CDFs cumsum / (total + 1e-8) in double
    if (color_match == 1) {
        __syncthreads();
        const bool any = st->n_overlap > 0;
        for (int c = 0; c < 3; ++c) {
            const int v = threadIdx.x;
            unsigned long long a = 0, b = 0;
            for (int u = 0; u <= v; ++u) { a += hist[c * 256 + u]; b += hist[(3 + c) * 256 + u]; }
            sc[v] = (double)a;
            rc[v] = (double)b;
            __syncthreads();
            const double stt = sc[255] + 1e-8, rtt = rc[255] + 1e-8;
            __syncthreads();
            sc[v] = sc[v] / stt;
            rc[v] = rc[v] / rtt;
            __syncthreads();
            lut[c * 256 + v] = any ? (float)np_interp_index(sc[v], rc, 256) : (float)v;
            __syncthreads();
        }
    }
}

// gained image 1, the level-0 planes of both images (each filled with the
// other outside its own region), the colour cost, the channel mean and the
// seam RMSE sums; planes: [I1 b, g, r, I2 b, g, r, mask] of P floats
__global__ __launch_bounds__(kCT) void k_cmp_prepare(const uint8_t* __restrict__ wu, const uint8_t* __restrict__ c2,
                                                     const uint8_t* __restrict__ m1, const uint8_t* __restrict__ m2,
                                                     size_t P, CmpStats* __restrict__ st, int color_match,
                                                     const float* __restrict__ lut, float* __restrict__ planes,
                                                     float* __restrict__ costc, float* __restrict__ avg,
                                                     float* __restrict__ seam0) {
    __shared__ unsigned long long sh[kCT / 64];
    const size_t p = (size_t)blockIdx.x * kCT + threadIdx.x;
    unsigned long long sq = 0;
    if (p < P) {
        const bool a = m1[p], b = m2[p], any = st->has_overlap;
        float d2[3], hv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float w = (float)wu[p * 3 + c];
            if (any) {
                if (color_match == 1) {
                    w = lut[c * 256 + wu[p * 3 + c]];
                } else {
                    w = w * st->gain[c];
                    w = w < 0.0f ? 0.0f : (w > 255.0f ? 255.0f : w);
                }
            }
            const float q = (float)c2[p * 3 + c];
            const float d = w - q;
            d2[c] = d * d;
            hv[c] = (w + q) * 0.5f;
            planes[(size_t)c * P + p] = a ? w : (b ? q : 0.0f);
            planes[(size_t)(c + 3) * P + p] = b ? q : (a ? w : 0.0f);
            if (a && b) { const int e = (int)(uint8_t)w - (int)c2[p * 3 + c]; sq += (unsigned long long)(e * e); }
        }
        costc[p] = sqrtf((d2[0] + d2[1]) + d2[2]);
        avg[p] = ((hv[0] + hv[1]) + hv[2]) / 3.0f;
        seam0[p] = (a && !b) ? 1.0f : 0.0f;
    }
    sq = block_sum_u64(sq, sh);
    if (threadIdx.x == 0 && sq) atomicAdd(&st->sq_diff, sq);
}

// Sobel first passes: dx = [-1, 0, 1] along x, dy = [-1, 0, 1] along y
__global__ __launch_bounds__(kCT) void k_cmp_sobel1(const float* __restrict__ avg, int cw, int ch,
                                                    float* __restrict__ dx, float* __restrict__ dy) {
    const int u = blockIdx.x * kCT + threadIdx.x, v = blockIdx.y;
    if (u >= cw) return;
    const double k3[3] = { -1.0, 0.0, 1.0 };
    double ax = 0.0, ay = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ax += (double)avg[(size_t)v * cw + refl_half(u + k - 1, cw)] * k3[k];
        ay += (double)avg[(size_t)refl_half(v + k - 1, ch) * cw + u] * k3[k];
    }
    dx[(size_t)v * cw + u] = (float)ax;
    dy[(size_t)v * cw + u] = (float)ay;
}

// Sobel second passes ([1, 2, 1] across), gradient magnitude, total cost;
// the cost is stored along the seam's lines (row-major for a vertical seam,
// column-major for a horizontal one) for the dynamic programme's reads
__global__ __launch_bounds__(kCT) void k_cmp_sobel2(const float* __restrict__ dx, const float* __restrict__ dy,
                                                    const float* __restrict__ costc, const uint8_t* __restrict__ m1,
                                                    const uint8_t* __restrict__ m2, int cw, int ch,
                                                    const CmpStats* __restrict__ st, float* __restrict__ cost) {
    const int u = blockIdx.x * kCT + threadIdx.x, v = blockIdx.y;
    if (u >= cw) return;
    const double k3[3] = { 1.0, 2.0, 1.0 };
    double ax = 0.0, ay = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ax += (double)dx[(size_t)refl_half(v + k - 1, ch) * cw + u] * k3[k];
        ay += (double)dy[(size_t)v * cw + refl_half(u + k - 1, cw)] * k3[k];
    }
    const float gx = (float)ax, gy = (float)ay;
    const size_t o = (size_t)v * cw + u;
    const float gr = sqrtf(gx * gx + gy * gy);
    const float c = (m1[o] && m2[o]) ? costc[o] + 0.05f * gr : 1e9f;
    cost[st->vertical ? o : (size_t)u * ch + v] = c;
}

// The seam's dynamic programme over the lines holding overlap (rows for a
// vertical seam, columns otherwise; a cell belongs to the overlap exactly when
// its cost is below 1e9), in bands of kDpK lines.  Line i's value at n needs
// line i - 1 at n - 1, n, n + 1 only, so the workgroup that owns cells
// [s0, s0 + kDpS) of a band can compute them from the band's input line over
// [s0 - kDpK, s0 + kDpS + kDpK), the valid range shrinking by one cell a side
// per line (not at the line's ends, where the neighbours are +inf).  Every
// cell is thus computed by the same operations as a line-by-line pass, partly
// twice; each workgroup writes the back-pointers and the band's output values
// of the cells it owns.  One launch per band; the backtrack runs in a last
// single workgroup.
//
// A workgroup is one wave: lane j holds cells a + j + 64 c (c < kDpC) in
// registers and takes its neighbours' values from the adjacent lanes by DPP
// wave rotations, so a line needs no LDS round trip and no barrier.
constexpr int kDpS = 64;                // cells a workgroup owns
constexpr int kDpK = 64;                // lines per band
constexpr int kDpR = kDpS + 2 * kDpK;   // cells a workgroup computes
constexpr int kDpC = kDpR / 64;         // cells per lane
constexpr int kDpPf = 8;                // lines of costs in flight
constexpr int kWin = 64;                // backtrack window (lines)
constexpr int kWw = 2 * kWin + 1;       // cells a window can reach
constexpr int kFill = (kWin * kWw + 1023) / 1024;   // back-pointers per thread and window
static_assert(kDpS == 64 && kDpK == 64, "k_cmp_dp_band: one owned cell and one band line per lane");

// back-pointer row stride: N rounded up to a whole number of workgroups' cells
__host__ __device__ __forceinline__ int dp_bt_stride(int N) { return (N + kDpS - 1) / kDpS * kDpS; }

struct DpGeom {
    int L, N;
    const uint32_t* lf;
};

__device__ __forceinline__ DpGeom dp_geom(const CmpStats* st, const uint32_t* rowflags, const uint32_t* colflags,
                                          int cw, int ch) {
    const bool vert = st->vertical;
    return DpGeom{vert ? ch : cw, vert ? cw : ch, vert ? rowflags : colflags};
}

// overlap lines in order into lines[0 .. nl), nl into *nlp, and the first
// line's values (its costs, +inf off the overlap) into dp0.  One block.
__global__ __launch_bounds__(1024) void k_cmp_dp_lines(const float* __restrict__ cost,
                                                       const uint32_t* __restrict__ rowflags,
                                                       const uint32_t* __restrict__ colflags, int cw, int ch,
                                                       const CmpStats* __restrict__ st, int32_t* __restrict__ lines,
                                                       int* __restrict__ nlp, double* __restrict__ dp0) {
    __shared__ int wsum[16];
    __shared__ int s_nl, s_first;
    if (!st->has_overlap) return;
    const DpGeom g = dp_geom(st, rowflags, colflags, cw, ch);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // thread tid takes lines [tid * per, (tid + 1) * per)
    const int per = (g.L + 1023) / 1024;
    const int l0 = min(g.L, tid * per), l1 = min(g.L, l0 + per);
    int cnt = 0;
    for (int l = l0; l < l1; ++l) cnt += (g.lf[l] & 1u) ? 1 : 0;
    int incl = cnt;   // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int k = 0; k < 16; ++k) { const int v = wsum[k]; wsum[k] = acc; acc += v; }
        s_nl = acc;
        *nlp = acc;
    }
    __syncthreads();
    int pos = wsum[w] + incl - cnt;
    for (int l = l0; l < l1; ++l)
        if (g.lf[l] & 1u) {
            if (pos == 0) s_first = l;
            lines[pos++] = l;
        }
    __syncthreads();
    if (s_nl == 0) return;
    const int first = s_first;
    for (int n = tid; n < g.N; n += 1024) {
        const float c = cost[(size_t)first * g.N + n];
        dp0[n] = c < 1e9f ? (double)c : 1e18;
    }
}

// 64-bit value of lane j - 1 (lane 0: lane 63) and of lane j + 1 (lane 63: lane 0)
__device__ __forceinline__ double dpp_ror1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x13C, 0xF, 0xF, false);          // wave_ror:1
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x13C, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double dpp_rol1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x134, 0xF, 0xF, false);          // wave_rol:1
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x134, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// band b: lines i0 = 1 + b kDpK .. min(nl, i0 + kDpK) - 1 from dp_in (line
// i0 - 1) to dp_out; blockIdx.x owns cells [kDpS x, +kDpS)
__global__ __launch_bounds__(64) void k_cmp_dp_band(const float* __restrict__ cost,
                                                    const uint32_t* __restrict__ rowflags,
                                                    const uint32_t* __restrict__ colflags, int cw, int ch,
                                                    const CmpStats* __restrict__ st,
                                                    const int32_t* __restrict__ lines, const int* __restrict__ nlp,
                                                    int band, const double* __restrict__ dp_in,
                                                    double* __restrict__ dp_out, uint8_t* __restrict__ bt) {
    if (!st->has_overlap) return;
    const DpGeom g = dp_geom(st, rowflags, colflags, cw, ch);
    const int N = g.N, nl = *nlp;
    const int i0 = 1 + band * kDpK;
    if (i0 >= nl) return;
    const int i1 = min(nl, i0 + kDpK);
    const int s0 = blockIdx.x * kDpS;
    if (s0 >= N) return;
    const int a = max(0, s0 - kDpK), e = min(N, s0 + kDpS + kDpK);   // computed cells [a, e)
    const int lane = threadIdx.x;
    const double INF = 1e18;
    // The owned cells are register cown = (s0 - a) / 64 of every lane (s0 - a
    // is 0 or kDpK); the owned cell s0 + lane may lie in the row's padding past
    // N, which bt rows of dp_bt_stride(N) bytes hold.  Cells outside the valid
    // range are computed too, from whatever their neighbours hold: no valid
    // cell ever reads them.
    const int np = dp_bt_stride(N);
    const int cown = (s0 - a) >> 6;
    // the band's line numbers, lane q holding line i0 + q (clamped to the band)
    const int lnv = lines[min(i0 + lane, i1 - 1)];
    double dv[kDpC];
    bool okl[kDpC], okr[kDpC];
#pragma unroll
    for (int c = 0; c < kDpC; ++c) {
        const int n = a + lane + 64 * c;
        dv[c] = n < e ? dp_in[n] : INF;
        okl[c] = n > 0;
        okr[c] = n < N - 1;
    }
    // Costs of the next kDpPf lines: slot k holds line i for i = k (mod kDpPf).
    // Every line issues the same two loads per lane (clamped addresses, cells
    // past e are never valid) and one back-pointer store, with no early exit
    // from the unrolled steps, so the compiler waits only for the load a step
    // needs.
    float cb[kDpPf][kDpC];
    auto load_line = [&](int i, float* dst) {
        const int q = (i < i1 ? i : i1 - 1) - i0;   // past the band: a line of the band, never used
        const float* row = cost + (size_t)__builtin_amdgcn_readlane(lnv, q) * N;
#pragma unroll
        for (int c = 0; c < kDpC; ++c) {
            const int n = a + lane + 64 * c;
            dst[c] = row[n < e ? n : e - 1];
        }
    };
#pragma unroll
    for (int k = 0; k < kDpPf; ++k) load_line(i0 + k, cb[k]);
    int own = 0;
    uint8_t* btl = bt + s0 + lane;
    // one line; a step past the band's last line (live false, only in the last
    // group) changes nothing and stores the last line's back-pointer again
    // (same lane, same value)
    auto step = [&](int k, int i, bool live) __attribute__((always_inline)) {
        const int l = __builtin_amdgcn_readlane(lnv, (live ? i : i1 - 1) - i0);
        double rr[kDpC], rl[kDpC], nv[kDpC];
#pragma unroll
        for (int c = 0; c < kDpC; ++c) {
            rr[c] = dpp_ror1(dv[c]);
            rl[c] = dpp_rol1(dv[c]);
        }
        int aown = 0;
#pragma unroll
        for (int c = 0; c < kDpC; ++c) {
            double pl = lane == 0 ? (c > 0 ? rr[c - 1] : INF) : rr[c];
            double pr = lane == 63 ? (c < kDpC - 1 ? rl[c + 1] : INF) : rl[c];
            pl = okl[c] ? pl : INF;
            pr = okr[c] ? pr : INF;
            const double pm = dv[c];
            int arg = 0;
            double best = pl;
            if (pm < best) { best = pm; arg = 1; }
            if (pr < best) { best = pr; arg = 2; }
            const float cst = cb[k][c];
            nv[c] = cst < 1e9f ? (double)cst + best : INF;
            aown = c == cown ? arg : aown;
        }
        if (live) {
#pragma unroll
            for (int c = 0; c < kDpC; ++c) dv[c] = nv[c];
            own = aown;
        }
        btl[(size_t)l * np] = (uint8_t)own;
        load_line(i + kDpPf, cb[k]);
    };
    // Whole groups of kDpPf lines, then the last partial group: the main loop
    // has no branch inside, so the compiler counts the loads in flight exactly.
    int ib = i0;
    for (; ib + kDpPf <= i1; ib += kDpPf) {
#pragma unroll
        for (int k = 0; k < kDpPf; ++k) step(k, ib + k, true);
    }
    if (ib < i1) {
#pragma unroll
        for (int k = 0; k < kDpPf; ++k) step(k, ib + k, ib + k < i1);
    }
    double vo = dv[0];
#pragma unroll
    for (int c = 1; c < kDpC; ++c) vo = c == cown ? dv[c] : vo;
    if (s0 + lane < N) dp_out[s0 + lane] = vo;
}

// the first minimum over the last line's overlap cells, then the backtrack in
// windows of kWin lines: the seam moves at most one cell per line, so the block
// loads those lines' back-pointers within kWin cells of the current position
// into LDS and resolves the window's cells by a scan of the back-pointer maps.
__global__ __launch_bounds__(1024) void k_cmp_dp_back(const float* __restrict__ cost,
                                                      const uint32_t* __restrict__ rowflags,
                                                      const uint32_t* __restrict__ colflags, int cw, int ch,
                                                      const CmpStats* __restrict__ st,
                                                      const int32_t* __restrict__ lines, const int* __restrict__ nlp,
                                                      const double* __restrict__ dpa, const double* __restrict__ dpb,
                                                      const uint8_t* __restrict__ bt, int32_t* __restrict__ seam) {
    __shared__ uint8_t F[2][kWin][kWw];
    __shared__ int s_ln[2][kWin], s_row[2][kWin];
    __shared__ double rv[16];
    __shared__ int ri[16];
    __shared__ int s_pos;
    if (!st->has_overlap) return;
    const DpGeom g = dp_geom(st, rowflags, colflags, cw, ch);
    const int L = g.L, N = g.N, nl = *nlp;
    const int np = dp_bt_stride(N);
    const int tid = threadIdx.x;
    if (nl == 0) {
        for (int l = tid; l < L; l += 1024) seam[l] = 0;
        return;
    }
    const double INF = 1e18;
    // bands 0 .. nb - 1 alternate dpa -> dpb -> dpa ...; the last line is in dp
    const int nb = (nl - 1 + kDpK - 1) / kDpK;
    const double* dp = (nb & 1) ? dpb : dpa;
    const int last = lines[nl - 1];
    double bv = INF;
    int bi = 0x7fffffff;
    for (int n = tid; n < N; n += 1024) {
        const float c = cost[(size_t)last * N + n];
        const double v = c < 1e9f ? dp[n] : INF;
        if (v < bv || (v == bv && n < bi)) { bv = v; bi = n; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { rv[tid >> 6] = bv; ri[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
        for (int k = 1; k < 16; ++k)
            if (rv[k] < rv[0] || (rv[k] == rv[0] && ri[k] < ri[0])) { rv[0] = rv[k]; ri[0] = ri[k]; }
        s_pos = ri[0] == 0x7fffffff ? 0 : ri[0];
    }
    __syncthreads();
    const int sc = s_pos;
    for (int l = tid; l < L; l += 1024) seam[l] = sc;
    if (tid < min(kWin, nl - 1)) {   // the first window's line numbers
        s_row[0][tid] = lines[nl - 1 - tid];
        s_ln[0][tid] = lines[nl - 2 - tid];
    }
    __syncthreads();
    // Walk back: entry i (from nl - 1 down to 1) sets seam[lines[i - 1]] to
    // clamp(pos + bt[lines[i]][pos] - 1, 0, N - 1), pos being the seam's cell
    // on line lines[i].  A window of kWin entries moves the seam at most kWin
    // cells, so in window coordinates x = pos - pos0 + kWin (pos0: the cell at
    // the window's first entry) each entry j is a map T_j of [0, 2 kWin] into
    // itself.  An inclusive scan of the maps under composition (log2 kWin
    // doubling rounds, F_j <- F_j o F_(j - d)) gives every entry's cell at
    // once; entries at x off the seam's path may hold clamped garbage, which no
    // evaluation on the path reads.
    int pos0 = sc, wb = 0;
    for (int i = nl - 1; i > 0; i -= kWin, wb ^= 1) {
        const int cnt = min(kWin, i);           // entries i, i - 1, ..., i - cnt + 1
        // All of a thread's back-pointer loads are in flight at once: the
        // indexes are clamped and nothing branches; a thread past the window's
        // last entry computes and writes that entry's own values again.
        const int tot = cnt * kWw;
        uint8_t v[kFill];
        int pk[kFill], jk[kFill], xk[kFill];
#pragma unroll
        for (int k = 0; k < kFill; ++k) {
            const int q = min(tid + 1024 * k, tot - 1);
            const int j = q / kWw, x = q - j * kWw;
            int p = pos0 + x - kWin;
            p = p < 0 ? 0 : (p > N - 1 ? N - 1 : p);
            pk[k] = p;
            jk[k] = j;
            xk[k] = x;
            v[k] = bt[(size_t)s_row[wb][j] * np + p];
        }
#pragma unroll
        for (int k = 0; k < kFill; ++k) {
            int b = pk[k] + (int)v[k] - 1;
            b = b < 0 ? 0 : (b > N - 1 ? N - 1 : b);
            const int t = b - pos0 + kWin;
            F[0][jk[k]][xk[k]] = (uint8_t)(t < 0 ? 0 : (t > kWw - 1 ? kWw - 1 : t));
        }
        lds_barrier();
        // the next window's line numbers load during this window's scan
        const int inx = i - kWin;
        if (inx > 0 && tid < min(kWin, inx)) {
            s_row[wb ^ 1][tid] = lines[inx - tid];
            s_ln[wb ^ 1][tid] = lines[inx - tid - 1];
        }
        int cur = 0;
        for (int d = 1; d < cnt; d <<= 1) {
            uint8_t nv[kFill];
#pragma unroll
            for (int k = 0; k < kFill; ++k) {
                const int j = jk[k], x = xk[k];
                const int y = j >= d ? (int)F[cur][j - d][x] : x;
                nv[k] = F[cur][j][y];
            }
#pragma unroll
            for (int k = 0; k < kFill; ++k) F[cur ^ 1][jk[k]][xk[k]] = nv[k];
            lds_barrier();
            cur ^= 1;
        }
        if (tid < cnt) seam[s_ln[wb][tid]] = pos0 + (int)F[cur][tid][kWin] - kWin;
        pos0 += (int)F[cur][cnt - 1][kWin] - kWin;
        lds_barrier();   // F and this window's line numbers are rewritten next
    }
}

// seam mask inside the overlap (1 on image 1's side of the seam)
__global__ __launch_bounds__(kCT) void k_cmp_seam_mask(const uint8_t* __restrict__ m1, const uint8_t* __restrict__ m2,
                                                       const int32_t* __restrict__ seam, int cw,
                                                       const CmpStats* __restrict__ st, float* __restrict__ seam0) {
    const int u = blockIdx.x * kCT + threadIdx.x, v = blockIdx.y;
    if (u >= cw || !st->has_overlap) return;
    const size_t o = (size_t)v * cw + u;
    if (!(m1[o] && m2[o])) return;
    const bool vert = st->vertical;
    const int l = vert ? v : u, n = vert ? u : v;
    seam0[o] = ((n <= seam[l]) == (st->img1_first != 0)) ? 1.0f : 0.0f;
}

// 25-tap Gaussian of the seam mask: rows into tmp, then columns into the
// mask plane (inside the overlap; the unblurred mask elsewhere).  The row pass
// stages the block's 256 + 24 source values in LDS once.
__global__ __launch_bounds__(kCT) void k_cmp_blur_rows(const float* __restrict__ in, int cw, W25 w,
                                                       const CmpStats* __restrict__ st, float* __restrict__ out) {
    __shared__ float t[kCT + 24];
    if (!st->has_overlap) return;
    const int u0 = blockIdx.x * kCT, v = blockIdx.y;
    const float* row = in + (size_t)v * cw;
    const int hi = min(u0 + kCT, cw) + 12;   // sources [u0 - 12, hi) feed the block's outputs
    for (int k = threadIdx.x; k < kCT + 24; k += kCT) {
        const int p = u0 - 12 + k;
        if (p < hi) t[k] = row[refl_101(p, cw)];
    }
    __syncthreads();
    const int u = u0 + threadIdx.x;
    if (u >= cw) return;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 25; ++k) acc = acc + w.w[k] * t[threadIdx.x + k];
    out[(size_t)v * cw + u] = acc;
}

__global__ __launch_bounds__(kCT) void k_cmp_blur_cols(const float* __restrict__ tmp, const float* __restrict__ seam0,
                                                       const uint8_t* __restrict__ m1, const uint8_t* __restrict__ m2,
                                                       int cw, int ch, W25 w, const CmpStats* __restrict__ st,
                                                       float* __restrict__ mask) {
    const int u = blockIdx.x * kCT + threadIdx.x, v = blockIdx.y;
    if (u >= cw) return;
    const size_t o = (size_t)v * cw + u;
    float r = seam0[o];
    if (st->has_overlap && m1[o] && m2[o]) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 25; ++k) acc = acc + w.w[k] * tmp[(size_t)refl_101(v + k - 12, ch) * cw + u];
        r = acc;
    }
    mask[o] = r;
}

// pyramid down, rows at the kept even rows: grid (ceil(W / 256),
// ceil(Ho / kPdR), planes).  A thread makes kPdR output rows of its column from
// 2 kPdR + 7 input rows held in registers (rows past the plane's end are
// reflected like any other and never used).
constexpr int kPdR = 8;
__global__ __launch_bounds__(kCT) void k_pyr_down_rows(const float* __restrict__ in, int H, int W, int Ho, W9 w,
                                                       float* __restrict__ tmp) {
    const int x = blockIdx.x * kCT + threadIdx.x, yo0 = blockIdx.y * kPdR, q = blockIdx.z;
    if (x >= W) return;
    const float* src = in + (size_t)q * H * W + x;
    float r[2 * kPdR + 7];
#pragma unroll
    for (int m = 0; m < 2 * kPdR + 7; ++m) r[m] = src[(size_t)refl_half(2 * yo0 + m - 4, H) * W];
    float* dst = tmp + (size_t)q * Ho * W + (size_t)yo0 * W + x;
#pragma unroll
    for (int i = 0; i < kPdR; ++i) {
        if (yo0 + i >= Ho) break;
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 9; ++k) acc = fmaf(r[2 * i + k], w.w[k], acc);
        dst[(size_t)i * W] = acc;
    }
}

// then columns: grid (ceil(Wo / 256), Ho, planes); the block's 2 x 256 + 8
// source values are staged in LDS once
__global__ __launch_bounds__(kCT) void k_pyr_down_cols(const float* __restrict__ tmp, int W, int Ho, int Wo, W9 w,
                                                       float* __restrict__ out) {
    __shared__ float t[2 * kCT + 8];
    const int xo0 = blockIdx.x * kCT, yo = blockIdx.y, q = blockIdx.z;
    const float* src = tmp + (size_t)q * Ho * W + (size_t)yo * W;
    const int hi = 2 * min(xo0 + kCT, Wo) + 3;   // sources [2 xo0 - 4, hi) feed the block's outputs
    for (int k = threadIdx.x; k < 2 * kCT + 8; k += kCT) {
        const int p = 2 * xo0 - 4 + k;
        if (p < hi) t[k] = src[refl_half(p, W)];
    }
    __syncthreads();
    const int xo = xo0 + threadIdx.x;
    if (xo >= Wo) return;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc = fmaf(t[2 * threadIdx.x + k], w.w[k], acc);
    out[(size_t)q * Ho * Wo + (size_t)yo * Wo + xo] = acc;
}

__device__ __forceinline__ void zoom_coord(int o, int n_in, int n_out, int& i0, int& i1, double& w0, double& w1) {
    double c = n_out > 1 ? (double)o * ((double)(n_in - 1) / (double)(n_out - 1)) : 0.0;
    if (c > (double)(n_in - 1)) c = (double)(n_in - 1);
    const double f = floor(c);
    i0 = (int)f;
    i1 = i0 + 1 < n_in ? i0 + 1 : n_in - 1;
    w1 = c - f;
    w0 = 1.0 - w1;
}

__device__ __forceinline__ float zoom_at(const float* __restrict__ in, int w, int y0, int y1, int x0, int x1,
                                         double wy0, double wy1, double wx0, double wx1) {
    double acc = (double)in[(size_t)y0 * w + x0] * (wy0 * wx0);
    acc = acc + (double)in[(size_t)y0 * w + x1] * (wy0 * wx1);
    acc = acc + (double)in[(size_t)y1 * w + x0] * (wy1 * wx0);
    acc = acc + (double)in[(size_t)y1 * w + x1] * (wy1 * wx1);
    return (float)acc;
}

// coarsest level: res = m g1 + (1 - m) g2, grid (ceil(W / 256), H, 3)
__global__ __launch_bounds__(kCT) void k_pyr_top(const float* __restrict__ lvl, int H, int W, float* __restrict__ res) {
    const int x = blockIdx.x * kCT + threadIdx.x, y = blockIdx.y, c = blockIdx.z;
    if (x >= W) return;
    const size_t n = (size_t)H * W, o = (size_t)y * W + x;
    const float m = lvl[6 * n + o];
    res[c * n + o] = m * lvl[c * n + o] + (1.0f - m) * lvl[(c + 3) * n + o];
}

// level k from level k + 1: res_k = up(res_k+1) + m (g1 - up g1') + (1 - m) (g2 - up g2'),
// the three channels per thread: grid (ceil(W / 256), H).  The block's two
// source rows of the nine upsampled planes (g1' and g2' of level k + 1, its
// result) are staged in LDS when the block's source columns fit kUpSpan.
constexpr int kUpSpan = 160;
static_assert(kUpSpan <= kCT, "k_pyr_up_blend stages one source column per thread");
__device__ __forceinline__ float zoom_row2(const float* __restrict__ r0, const float* __restrict__ r1, int x0, int x1,
                                           double wy0, double wy1, double wx0, double wx1) {
    double acc = (double)r0[x0] * (wy0 * wx0);
    acc = acc + (double)r0[x1] * (wy0 * wx1);
    acc = acc + (double)r1[x0] * (wy1 * wx0);
    acc = acc + (double)r1[x1] * (wy1 * wx1);
    return (float)acc;
}

__global__ __launch_bounds__(kCT) void k_pyr_up_blend(const float* __restrict__ lvk, int H, int W,
                                                      const float* __restrict__ lvn, int h, int w,
                                                      const float* __restrict__ resn, float* __restrict__ res) {
    __shared__ float sr[9][2][kUpSpan];
    const int xb = blockIdx.x * kCT, y = blockIdx.y, x = xb + threadIdx.x;
    int y0, y1;
    double wy0, wy1;
    zoom_coord(y, h, H, y0, y1, wy0, wy1);
    int xa, span;
    {
        int i0, i1;
        double a0, a1;
        zoom_coord(xb, w, W, i0, i1, a0, a1);
        xa = i0;
        zoom_coord(min(xb + kCT, W) - 1, w, W, i0, i1, a0, a1);
        span = i1 - xa + 1;
    }
    const size_t nn = (size_t)h * w;
    const bool staged = span <= kUpSpan;   // uniform
    if (staged) {
        const int cc = threadIdx.x;   // kUpSpan <= kCT: one column per thread
        if (cc < span) {
#pragma unroll
            for (int pl = 0; pl < 9; ++pl) {
                const float* base = (pl < 6 ? lvn + (size_t)pl * nn : resn + (size_t)(pl - 6) * nn) + xa + cc;
                sr[pl][0][cc] = base[(size_t)y0 * w];
                sr[pl][1][cc] = base[(size_t)y1 * w];
            }
        }
        __syncthreads();
    }
    if (x >= W) return;
    int x0, x1;
    double wx0, wx1;
    zoom_coord(x, w, W, x0, x1, wx0, wx1);
    const size_t n = (size_t)H * W, o = (size_t)y * W + x;
    const float m = lvk[6 * n + o];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float u1, u2, ur;
        if (staged) {
            u1 = zoom_row2(sr[c][0], sr[c][1], x0 - xa, x1 - xa, wy0, wy1, wx0, wx1);
            u2 = zoom_row2(sr[c + 3][0], sr[c + 3][1], x0 - xa, x1 - xa, wy0, wy1, wx0, wx1);
            ur = zoom_row2(sr[6 + c][0], sr[6 + c][1], x0 - xa, x1 - xa, wy0, wy1, wx0, wx1);
        } else {
            u1 = zoom_at(lvn + c * nn, w, y0, y1, x0, x1, wy0, wy1, wx0, wx1);
            u2 = zoom_at(lvn + (c + 3) * nn, w, y0, y1, x0, x1, wy0, wy1, wx0, wx1);
            ur = zoom_at(resn + c * nn, w, y0, y1, x0, x1, wy0, wy1, wx0, wx1);
        }
        const float l1 = lvk[c * n + o] - u1, l2 = lvk[(c + 3) * n + o] - u2;
        const float lb = m * l1 + (1.0f - m) * l2;
        res[c * n + o] = ur + lb;
    }
}

__global__ __launch_bounds__(kCT) void k_cmp_final(const float* __restrict__ res, const uint8_t* __restrict__ m1,
                                                   const uint8_t* __restrict__ m2, size_t P, uint8_t* __restrict__ pano,
                                                   uint8_t* __restrict__ valid, const float* __restrict__ mask,
                                                   float* __restrict__ seam_out, CmpStats* __restrict__ st) {
    const size_t p = (size_t)blockIdx.x * kCT + threadIdx.x;
    if (p == 0) st->rmse = st->n_overlap ? sqrt((double)st->sq_diff / (double)(3 * st->n_overlap)) : __builtin_nan("");
    if (p >= P) return;
    const bool vl = m1[p] || m2[p];
    valid[p] = vl;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = res[(size_t)c * P + p];
        v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
        pano[p * 3 + c] = vl ? (uint8_t)v : 0;
    }
    if (seam_out) seam_out[p] = mask[p];
}

__global__ void k_cmp_zero(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * kCT + threadIdx.x;
    if (i < n) p[i] = 0;
}

struct CmpPlan {
    int L, hs[kMaxLevels], ws[kMaxLevels];
    size_t P, o_wu, o_c2, o_m1, o_m2, o_rowf, o_colf, o_st, o_costc, o_avg, o_dx, o_dy, o_cost, o_seam0, o_tmp,
        o_lvl[kMaxLevels], o_res[kMaxLevels], o_dpa, o_dpb, o_bt, o_seam, o_lines, o_nl, o_rowacc, o_hist, o_lut,
        zero_bytes, total;
};

CmpPlan cmp_plan(int cw, int ch, int n_levels) {
    CmpPlan Pl;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    const size_t P = (size_t)cw * ch;
    Pl.P = P;
    Pl.L = n_levels;
    Pl.hs[0] = ch; Pl.ws[0] = cw;
    for (int k = 1; k < n_levels; ++k) { Pl.hs[k] = (Pl.hs[k - 1] + 1) / 2; Pl.ws[k] = (Pl.ws[k - 1] + 1) / 2; }
    Pl.o_wu = take(P * 3); Pl.o_c2 = take(P * 3); Pl.o_m1 = take(P); Pl.o_m2 = take(P);
    // zeroed at the start of every call: stats, row / column flags, row partials
    Pl.o_st = take(sizeof(CmpStats));
    Pl.o_rowf = take((size_t)ch * 4); Pl.o_colf = take((size_t)cw * 4);
    Pl.o_rowacc = take((size_t)ch * 8 * 8);
    Pl.o_hist = take(6 * 256 * 4);
    Pl.zero_bytes = off - Pl.o_st;
    Pl.o_lut = take(3 * 256 * 4);
    Pl.o_costc = take(P * 4); Pl.o_avg = take(P * 4); Pl.o_dx = take(P * 4); Pl.o_dy = take(P * 4);
    Pl.o_cost = take(P * 4); Pl.o_seam0 = take(P * 4);
    Pl.o_tmp = take(std::max(P, (size_t)7 * ((ch + 1) / 2) * cw) * 4);
    for (int k = 0; k < n_levels; ++k) Pl.o_lvl[k] = take((size_t)7 * Pl.hs[k] * Pl.ws[k] * 4);
    for (int k = 0; k < n_levels; ++k) Pl.o_res[k] = take((size_t)3 * Pl.hs[k] * Pl.ws[k] * 4);
    const int N = std::max(cw, ch);
    Pl.o_dpa = take((size_t)N * 8); Pl.o_dpb = take((size_t)N * 8);
    // back-pointers: L lines of dp_bt_stride(N) bytes, for either seam orientation
    Pl.o_bt = take(std::max((size_t)ch * dp_bt_stride(cw), (size_t)cw * dp_bt_stride(ch)));
    Pl.o_seam = take((size_t)N * 4);
    Pl.o_lines = take((size_t)N * 4);
    Pl.o_nl = take(4);
    Pl.total = off;
    return Pl;
}

void inv3_adj(const double* a, double* o) {
    double det = (a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6])) + a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (!(std::fabs(det) > 1e-12)) det = 1.0;
    o[0] = (a[4] * a[8] - a[5] * a[7]) / det; o[1] = (a[2] * a[7] - a[1] * a[8]) / det; o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    o[3] = (a[5] * a[6] - a[3] * a[8]) / det; o[4] = (a[0] * a[8] - a[2] * a[6]) / det; o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    o[6] = (a[3] * a[7] - a[4] * a[6]) / det; o[7] = (a[1] * a[6] - a[0] * a[7]) / det; o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

}  // namespace

size_t composite_workspace(int cw, int ch, int n_levels) { return cmp_plan(cw, ch, n_levels).total; }

int composite(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2, const double* H, int cw,
              int ch, double ox, double oy, int n_levels, int color_match, void* ws_, size_t ws_bytes, uint8_t* pano,
              uint8_t* valid, float* seam_out, void* stats_out, hipStream_t s) {
    if (!img1 || !img2 || !H || !pano || !valid || !stats_out || !ws_) {
        set_error("msift_composite_multiband: null buffer");
        return MSIFT_EINVAL;
    }
    if (h1 < 2 || w1 < 2 || h2 < 1 || w2 < 1 || cw < 1 || ch < 1 || ch > 65535 || n_levels < 1 ||
        n_levels > kMaxLevels || cw > 65535) {
        set_error("msift_composite_multiband: bad arguments (image 1 at least 2x2, canvas sides <= 65535, "
                  "1 <= n_levels <= 12)");
        return MSIFT_EINVAL;
    }
    const CmpPlan Pl = cmp_plan(cw, ch, n_levels);
    if (ws_bytes < Pl.total) { set_error("msift_composite_multiband: workspace too small"); return MSIFT_EWORKSPACE; }
    if (color_match != 0 && color_match != 1) {
        set_error("msift_composite_multiband: color_match must be 0 (gain) or 1 (hist)");
        return MSIFT_EINVAL;
    }
    if (((uintptr_t)ws_ & 255) != 0) { set_error("msift_composite_multiband: workspace must be 256-byte aligned"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    uint8_t* wu = (uint8_t*)(ws + Pl.o_wu);
    uint8_t* c2 = (uint8_t*)(ws + Pl.o_c2);
    uint8_t* m1 = (uint8_t*)(ws + Pl.o_m1);
    uint8_t* m2 = (uint8_t*)(ws + Pl.o_m2);
    uint32_t* rowf = (uint32_t*)(ws + Pl.o_rowf);
    uint32_t* colf = (uint32_t*)(ws + Pl.o_colf);
    CmpStats* st = (CmpStats*)(ws + Pl.o_st);
    float* costc = (float*)(ws + Pl.o_costc);
    float* avg = (float*)(ws + Pl.o_avg);
    float* dx = (float*)(ws + Pl.o_dx);
    float* dy = (float*)(ws + Pl.o_dy);
    float* cost = (float*)(ws + Pl.o_cost);
    float* seam0 = (float*)(ws + Pl.o_seam0);
    float* tmp = (float*)(ws + Pl.o_tmp);
    const size_t P = Pl.P;
    double Hi[9];
    inv3_adj(H, Hi);
    Hf9 hf;
    for (int i = 0; i < 9; ++i) hf.h[i] = (float)Hi[i];
    // sigma-1 taps (SciPy's kernel: exp(-x^2 / 2) over numpy's pairwise sum)
    W9 w9;
    {
        double phi[9];
        for (int k = 0; k < 9; ++k) { const double x = (double)(k - 4); phi[k] = std::exp(-0.5 * (x * x)); }
        const double sum = (((phi[0] + phi[1]) + (phi[2] + phi[3])) + ((phi[4] + phi[5]) + (phi[6] + phi[7]))) + phi[8];
        for (int k = 0; k < 9; ++k) w9.w[k] = (float)(phi[k] / sum);
    }
    // cv2.getGaussianKernel(25, 3) in float32
    W25 w25;
    {
        double sum = 0.0;
        for (int i = 0; i < 25; ++i) {
            const double x = (double)i - 12.0;
            w25.w[i] = (float)std::exp((-0.5 / 9.0) * x * x);
            sum += w25.w[i];
        }
        sum = 1.0 / sum;
        for (int i = 0; i < 25; ++i) w25.w[i] = (float)(w25.w[i] * sum);
    }
    const dim3 g2d((cw + kCT - 1) / kCT, ch);
    const unsigned gP = (unsigned)((P + kCT - 1) / kCT);
    const int nz = (int)(Pl.zero_bytes / 8);
    unsigned long long* rowacc = (unsigned long long*)(ws + Pl.o_rowacc);
    hipLaunchKernelGGL(k_cmp_zero, dim3((nz + kCT - 1) / kCT), dim3(kCT), 0, s, (unsigned long long*)st, nz);
    hipLaunchKernelGGL(k_cmp_warp, g2d, dim3(kCT), 0, s, img1, h1, w1, img2, h2, w2, hf, cw, (float)ox, (float)oy,
                       (int)std::rint(ox), (int)std::rint(oy), wu, c2, m1, m2, rowacc, rowf);
    hipLaunchKernelGGL(k_cmp_colflags, dim3((cw + kCT - 1) / kCT, (ch + 63) / 64), dim3(kCT), 0, s,
                       (const uint8_t*)m1, (const uint8_t*)m2, cw, ch, colf);
    uint32_t* hist = (uint32_t*)(ws + Pl.o_hist);
    float* lut = (float*)(ws + Pl.o_lut);
    if (color_match == 1)
        hipLaunchKernelGGL(k_cmp_hist, dim3((unsigned)std::min<size_t>((P + 4095) / 4096, 1024)), dim3(kCT), 0, s,
                           (const uint8_t*)wu, (const uint8_t*)c2, (const uint8_t*)m1, (const uint8_t*)m2, P, hist);
    hipLaunchKernelGGL(k_cmp_decide, dim3(1), dim3(kCT), 0, s, (const uint32_t*)rowf, ch, (const uint32_t*)colf, cw,
                       (const unsigned long long*)rowacc, color_match, (const uint32_t*)hist, lut, st);
    float* lvl0 = (float*)(ws + Pl.o_lvl[0]);
    hipLaunchKernelGGL(k_cmp_prepare, dim3(gP), dim3(kCT), 0, s, (const uint8_t*)wu, (const uint8_t*)c2,
                       (const uint8_t*)m1, (const uint8_t*)m2, P, st, color_match, (const float*)lut, lvl0, costc,
                       avg, seam0);
    hipLaunchKernelGGL(k_cmp_sobel1, g2d, dim3(kCT), 0, s, (const float*)avg, cw, ch, dx, dy);
    hipLaunchKernelGGL(k_cmp_sobel2, g2d, dim3(kCT), 0, s, (const float*)dx, (const float*)dy, (const float*)costc,
                       (const uint8_t*)m1, (const uint8_t*)m2, cw, ch, (const CmpStats*)st, cost);
    int32_t* seam = (int32_t*)(ws + Pl.o_seam);
    {
        // the seam: overlap lines, one launch per band of kDpK lines (launches past
        // the last band return at once), backtrack
        double* dpa = (double*)(ws + Pl.o_dpa);
        double* dpb = (double*)(ws + Pl.o_dpb);
        int32_t* lines = (int32_t*)(ws + Pl.o_lines);
        int* nlp = (int*)(ws + Pl.o_nl);
        uint8_t* bt = (uint8_t*)(ws + Pl.o_bt);
        const int Lmax = std::max(cw, ch), Nmax = std::max(cw, ch);
        hipLaunchKernelGGL(k_cmp_dp_lines, dim3(1), dim3(1024), 0, s, (const float*)cost, (const uint32_t*)rowf,
                           (const uint32_t*)colf, cw, ch, (const CmpStats*)st, lines, nlp, dpa);
        const int nbands = (Lmax - 1 + kDpK - 1) / kDpK;
        for (int b = 0; b < nbands; ++b)
            hipLaunchKernelGGL(k_cmp_dp_band, dim3((Nmax + kDpS - 1) / kDpS), dim3(64), 0, s, (const float*)cost,
                               (const uint32_t*)rowf, (const uint32_t*)colf, cw, ch, (const CmpStats*)st,
                               (const int32_t*)lines, (const int*)nlp, b, (const double*)((b & 1) ? dpb : dpa),
                               (b & 1) ? dpa : dpb, bt);
        hipLaunchKernelGGL(k_cmp_dp_back, dim3(1), dim3(1024), 0, s, (const float*)cost, (const uint32_t*)rowf,
                           (const uint32_t*)colf, cw, ch, (const CmpStats*)st, (const int32_t*)lines,
                           (const int*)nlp, (const double*)dpa, (const double*)dpb, (const uint8_t*)bt, seam);
    }
    hipLaunchKernelGGL(k_cmp_seam_mask, g2d, dim3(kCT), 0, s, (const uint8_t*)m1, (const uint8_t*)m2,
                       (const int32_t*)seam, cw, (const CmpStats*)st, seam0);
    hipLaunchKernelGGL(k_cmp_blur_rows, g2d, dim3(kCT), 0, s, (const float*)seam0, cw, w25, (const CmpStats*)st, tmp);
    hipLaunchKernelGGL(k_cmp_blur_cols, g2d, dim3(kCT), 0, s, (const float*)tmp, (const float*)seam0,
                       (const uint8_t*)m1, (const uint8_t*)m2, cw, ch, w25, (const CmpStats*)st, lvl0 + 6 * P);
    for (int k = 1; k < n_levels; ++k) {
        const int Hk = Pl.hs[k - 1], Wk = Pl.ws[k - 1], Ho = Pl.hs[k], Wo = Pl.ws[k];
        hipLaunchKernelGGL(k_pyr_down_rows, dim3((Wk + kCT - 1) / kCT, (Ho + kPdR - 1) / kPdR, 7), dim3(kCT), 0, s,
                           (const float*)(ws + Pl.o_lvl[k - 1]), Hk, Wk, Ho, w9, tmp);
        hipLaunchKernelGGL(k_pyr_down_cols, dim3((Wo + kCT - 1) / kCT, Ho, 7), dim3(kCT), 0, s, (const float*)tmp, Wk,
                           Ho, Wo, w9, (float*)(ws + Pl.o_lvl[k]));
    }
    const int kL = n_levels - 1;
    hipLaunchKernelGGL(k_pyr_top, dim3((Pl.ws[kL] + kCT - 1) / kCT, Pl.hs[kL], 3), dim3(kCT), 0, s,
                       (const float*)(ws + Pl.o_lvl[kL]), Pl.hs[kL], Pl.ws[kL], (float*)(ws + Pl.o_res[kL]));
    for (int k = n_levels - 2; k >= 0; --k)
        hipLaunchKernelGGL(k_pyr_up_blend, dim3((Pl.ws[k] + kCT - 1) / kCT, Pl.hs[k]), dim3(kCT), 0, s,
                           (const float*)(ws + Pl.o_lvl[k]), Pl.hs[k], Pl.ws[k], (const float*)(ws + Pl.o_lvl[k + 1]),
                           Pl.hs[k + 1], Pl.ws[k + 1], (const float*)(ws + Pl.o_res[k + 1]),
                           (float*)(ws + Pl.o_res[k]));
    hipLaunchKernelGGL(k_cmp_final, dim3(gP), dim3(kCT), 0, s, (const float*)(ws + Pl.o_res[0]), (const uint8_t*)m1,
                       (const uint8_t*)m2, P, pano, valid, (const float*)(lvl0 + 6 * P), seam_out, st);
    MSIFT_LAUNCH_CHECK();
    MSIFT_HIP(hipMemcpyAsync(stats_out, st, sizeof(CmpStats), hipMemcpyDeviceToDevice, s));
    return MSIFT_OK;
}

size_t composite_stats_bytes() { return sizeof(CmpStats); }

}  // namespace msift
