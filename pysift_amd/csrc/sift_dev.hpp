// Device bodies shared by the single-image and the batched detect kernels.
// Each function is the per-tile / per-candidate / per-keypoint work of one
// stage; the kernels around them only decide which image, octave and tile a
// block works on.  Keeping one body per stage keeps both paths bit-identical.
#pragma once
#include "msift_internal.hpp"

#pragma clang fp contract(off)

namespace msift {

struct Ent { int kp; float angle, x, y; };

// ---------------------------------------------------------------- spline --
MSIFT_HD void bspline_w(double t, double* w) {
    double s = 1.0 - t;
    double t2 = t * t, t3 = t2 * t, s2 = s * s, s3 = s2 * s;
    w[0] = s3 * (1.0 / 6.0);
    w[1] = (2.0 / 3.0 - t2) + 0.5 * t3;
    w[2] = (2.0 / 3.0 - s2) + 0.5 * s3;
    w[3] = t3 * (1.0 / 6.0);
}

// prefilter down a column of the u8 image at (y, x): sum_k h|k| u(mirror(y+k))
__device__ __forceinline__ double ups_pf_col(const uint8_t* __restrict__ img, int pitch, int H, int x, int y,
                                             const UpsTaps& h) {
    double acc = 0.0;
    if (y >= kUpsK && y + kUpsK < H) {
        const uint8_t* col = img + (size_t)(y - kUpsK) * pitch + x;
#pragma unroll
        for (int k = -kUpsK; k <= kUpsK; ++k) {
            float u = (float)col[(size_t)(k + kUpsK) * pitch] / 255.0f;
            acc = fma(h.h[k < 0 ? -k : k], (double)u, acc);
        }
    } else {
        for (int k = -kUpsK; k <= kUpsK; ++k) {
            float u = (float)img[(size_t)mirror_idx(y + k, H) * pitch + x] / 255.0f;
            acc = fma(h.h[k < 0 ? -k : k], (double)u, acc);
        }
    }
    return acc;
}

__device__ __forceinline__ double ups_pf_row(const double* __restrict__ row, int W, int x, const UpsTaps& h) {
    double acc = 0.0;
    if (x >= kUpsK && x + kUpsK < W) {
#pragma unroll
        for (int k = -kUpsK; k <= kUpsK; ++k) acc = fma(h.h[k < 0 ? -k : k], row[x + k], acc);
    } else {
        for (int k = -kUpsK; k <= kUpsK; ++k) acc = fma(h.h[k < 0 ? -k : k], row[mirror_idx(x + k, W)], acc);
    }
    return acc;
}

__device__ __forceinline__ float ups_interp(const double* __restrict__ cf, int H, int W, double zy, double zx,
                                            int oy, int ox) {
    double yy = (double)oy * zy;
    double fy = floor(yy);
    int iy = (int)fy;
    double wy[4];
    bspline_w(yy - fy, wy);
    double xx = (double)ox * zx;
    double fx = floor(xx);
    int ix = (int)fx;
    double wx[4];
    bspline_w(xx - fx, wx);
    int xi[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) xi[b] = mirror_idx(ix - 1 + b, W);
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double* row = cf + (size_t)mirror_idx(iy - 1 + a, H) * W;
        double rv = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) rv = fma(wx[b], row[xi[b]], rv);
        v = fma(wy[a], rv, v);
    }
    return (float)v;
}

// ------------------------------------------------------------------ blur --
// One 64 x TH output tile of G[s+1] = blur(G[s]) (H pass then V pass, fused
// multiply-add chains over the taps in ascending order), D[s] = G[s+1] - G[s]
// and the next octave's base G[S][::2, ::2].  The tile and its reflected
// halo are staged through LDS.
template <int R, int TH>
__device__ __forceinline__ void blur_tile(const float* __restrict__ in, float* __restrict__ gout,
                                          float* __restrict__ dout, float* __restrict__ dec, int H, int W, int decW,
                                          const Taps& taps, int bx, int by, float* sIn, float* sH) {
    constexpr int TW = 64;
    constexpr int IW = TW + 2 * R;
    constexpr int IWP = (IW + 4 + 3) & ~3;
    constexpr int IH = TH + 2 * R;
    constexpr int NV = 4 + 2 * R;
    constexpr int NV4 = (NV + 3) / 4;
    const int tid = threadIdx.x;
    const int x0 = bx * TW, y0 = by * TH;
    {
        // all loads issued before the LDS stores; one reflection suffices
        // unless the image is narrower than the halo
        constexpr int NL = (IH * IW + 255) / 256;
        const bool small = (H <= R) || (W <= R);
        float tv[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            int i = tid + 256 * j;
            if (i < IH * IW) {
                int ly = i / IW, lx = i - ly * IW;
                int gy = y0 - R + ly, gx = x0 - R + lx;
                if (small) {
                    gy = reflect_idx(gy, H);
                    gx = reflect_idx(gx, W);
                } else {
                    gy = gy < 0 ? -1 - gy : (gy >= H ? 2 * H - 1 - gy : gy);
                    gx = gx < 0 ? -1 - gx : (gx >= W ? 2 * W - 1 - gx : gx);
                }
                tv[j] = in[(size_t)gy * W + gx];
            }
        }
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            int i = tid + 256 * j;
            if (i < IH * IW) {
                int ly = i / IW, lx = i - ly * IW;
                sIn[ly * IWP + lx] = tv[j];
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < IH * (TW / 4); i += 256) {
        int ly = i / (TW / 4), q = i - ly * (TW / 4);
        const float* row = sIn + ly * IWP + 4 * q;
        float v[NV4 * 4];
#pragma unroll
        for (int k = 0; k < NV4; ++k) {
            float4 t = *reinterpret_cast<const float4*>(row + 4 * k);
            v[4 * k] = t.x; v[4 * k + 1] = t.y; v[4 * k + 2] = t.z; v[4 * k + 3] = t.w;
        }
        float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f, o3 = 0.0f;
#pragma unroll
        for (int k = 0; k <= 2 * R; ++k) {
            float w = taps.w[k];
            o0 = fmaf(v[k], w, o0);
            o1 = fmaf(v[k + 1], w, o1);
            o2 = fmaf(v[k + 2], w, o2);
            o3 = fmaf(v[k + 3], w, o3);
        }
        *reinterpret_cast<float4*>(sH + ly * TW + 4 * q) = make_float4(o0, o1, o2, o3);
    }
    __syncthreads();
    constexpr int RG = TH / 4;
    const int lx = tid & 63, g = tid >> 6;
    float v[RG + 2 * R];
#pragma unroll
    for (int k = 0; k < RG + 2 * R; ++k) v[k] = sH[(g * RG + k) * TW + lx];
    const int gx = x0 + lx;
#pragma unroll
    for (int j = 0; j < RG; ++j) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k <= 2 * R; ++k) s = fmaf(v[j + k], taps.w[k], s);
        int gy = y0 + g * RG + j;
        if (gy < H && gx < W) {
            size_t o = (size_t)gy * W + gx;
            if (gout) gout[o] = s;
            if (dout) dout[o] = s - sIn[(g * RG + j + R) * IWP + lx + R];
            if (dec && !(gy & 1) && !(gx & 1)) dec[(size_t)(gy >> 1) * decW + (gx >> 1)] = s;
        }
    }
}

template <int R, int TH>
struct BlurSmem {
    static constexpr int TW = 64;
    static constexpr int IWP = (TW + 2 * R + 4 + 3) & ~3;
    static constexpr int IH = TH + 2 * R;
    static constexpr int nIn = IH * IWP;
    static constexpr int nH = IH * TW;
};

// --------------------------------------------------------------- extrema --
__device__ __forceinline__ bool is_extremum(const float* __restrict__ D, int H, int W, int s, int r, int c,
                                            float thr) {
    size_t HW = (size_t)H * W;
    float v = D[(size_t)s * HW + (size_t)r * W + c];
    if (fabsf(v) <= thr) return false;
    bool mx = true, mn = true;
#pragma unroll
    for (int ds = -1; ds <= 1; ++ds)
#pragma unroll
        for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
            for (int dc = -1; dc <= 1; ++dc) {
                if (ds == 0 && dr == 0 && dc == 0) continue;
                float nb = D[(size_t)(s + ds) * HW + (size_t)(r + dr) * W + (c + dc)];
                mx = mx && (v >= nb);
                mn = mn && (v <= nb);
            }
    return mx || mn;
}

// block-wide exclusive scan of one int per thread (256 threads); returns the
// exclusive prefix and stores the block total in *tot
__device__ __forceinline__ int block_excl_scan256(int c, int* wsum, int* tot) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int pre = 0;
    for (int w = 0; w < wv; ++w) pre += wsum[w];
    *tot = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    __syncthreads();
    return pre + x - c;
}

// ---------------------------------------------------------------- refine --
// refine_keypoints (core.py:343-463) in FP64 on one octave's DoG stack;
// returns true and fills ri, ci, resp, si for an accepted keypoint
__device__ __forceinline__ bool refine_candidate(const float* __restrict__ D, int H, int W, int S, int bd, int si,
                                                 int row, int col, double thr, float edge32, float* ri, float* ci,
                                                 float* resp_out, int* si_out) {
    const size_t HW = (size_t)H * W;
#define DV(s_, r_, c_) ((double)D[(size_t)(s_) * HW + (size_t)(r_) * W + (c_)])
    double ox = 0.0, oy = 0.0, os = 0.0;
    bool conv = false;
    for (int it = 0; it < 5; ++it) {
        if (si < 1 || si > S || row < bd || row >= H - bd || col < bd || col >= W - bd) break;
        double cv = DV(si, row, col);
        double dx = (DV(si, row, col + 1) - DV(si, row, col - 1)) * 0.5;
        double dy = (DV(si, row + 1, col) - DV(si, row - 1, col)) * 0.5;
        double ds = (DV(si + 1, row, col) - DV(si - 1, row, col)) * 0.5;
        double dxx = (DV(si, row, col + 1) - 2.0 * cv) + DV(si, row, col - 1);
        double dyy = (DV(si, row + 1, col) - 2.0 * cv) + DV(si, row - 1, col);
        double dss = (DV(si + 1, row, col) - 2.0 * cv) + DV(si - 1, row, col);
        double dxy = (((DV(si, row + 1, col + 1) - DV(si, row + 1, col - 1)) - DV(si, row - 1, col + 1)) +
                      DV(si, row - 1, col - 1)) * 0.25;
        double dxs = (((DV(si + 1, row, col + 1) - DV(si + 1, row, col - 1)) - DV(si - 1, row, col + 1)) +
                      DV(si - 1, row, col - 1)) * 0.25;
        double dys = (((DV(si + 1, row + 1, col) - DV(si + 1, row - 1, col)) - DV(si - 1, row + 1, col)) +
                      DV(si - 1, row - 1, col)) * 0.25;
        double det = (dxx * (dyy * dss - dys * dys) - dxy * (dxy * dss - dys * dxs)) + dxs * (dxy * dys - dyy * dxs);
        if (fabs(det) < 1e-12) break;
        double pp = -dx, q = -dy, rv = -ds;
        ox = ((pp * (dyy * dss - dys * dys) - dxy * (q * dss - dys * rv)) + dxs * (q * dys - dyy * rv)) / det;
        oy = ((dxx * (q * dss - dys * rv) - pp * (dxy * dss - dys * dxs)) + dxs * (dxy * rv - q * dxs)) / det;
        os = ((dxx * (dyy * rv - q * dys) - dxy * (dxy * rv - q * dxs)) + pp * (dxy * dys - dyy * dxs)) / det;
        if (fabs(ox) < 0.5 && fabs(oy) < 0.5 && fabs(os) < 0.5) { conv = true; break; }
        col += (ox >= 0.0) ? (int)(ox + 0.5) : (int)(ox - 0.5);
        row += (oy >= 0.0) ? (int)(oy + 0.5) : (int)(oy - 0.5);
        si += (os >= 0.0) ? (int)(os + 0.5) : (int)(os - 0.5);
    }
    bool ok = false;
    if (conv && !(si < 1 || si > S || row < bd || row >= H - bd || col < bd || col >= W - bd)) {
        double cv = DV(si, row, col);
        double dx = (DV(si, row, col + 1) - DV(si, row, col - 1)) * 0.5;
        double dy = (DV(si, row + 1, col) - DV(si, row - 1, col)) * 0.5;
        double ds = (DV(si + 1, row, col) - DV(si - 1, row, col)) * 0.5;
        double resp = fabs(cv + 0.5 * ((dx * ox + dy * oy) + ds * os));
        if (resp >= thr) {
            double dxx = (DV(si, row, col + 1) - 2.0 * cv) + DV(si, row, col - 1);
            double dyy = (DV(si, row + 1, col) - 2.0 * cv) + DV(si, row - 1, col);
            double dxy = (((DV(si, row + 1, col + 1) - DV(si, row + 1, col - 1)) - DV(si, row - 1, col + 1)) +
                          DV(si, row - 1, col - 1)) * 0.25;
            double tr = dxx + dyy;
            double det2 = dxx * dyy - dxy * dxy;
            if (!(det2 <= 0.0 || (tr * tr) / det2 >= (double)edge32)) {
                *ri = (float)((double)row + oy);
                *ci = (float)((double)col + ox);
                *resp_out = (float)resp;
                *si_out = si;
                ok = true;
            }
        }
    }
#undef DV
    return ok;
}

// ----------------------------------------------------------- orientation --
// orientation_hist (core.py:228-333) for one keypoint on one wave: pixel p of
// the window goes to lane p % 64's private 36-bin column P[bin*64 + lane].
// Every lane returns; lane 0's out[]/n are the result.
__device__ __forceinline__ int orient_keypoint(const float* __restrict__ G, int H, int W, float ri, float ci,
                                               float sigma, float* P, int lane, float* out) {
#pragma unroll
    for (int b = 0; b < 36; ++b) P[b * 64 + lane] = 0.0f;
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
    for (int p = lane; p < total; p += 64) {
        int rq = p / n_cols;
        int r = r_min + rq;
        int c = c_min + (p - rq * n_cols);
        const float* g = G + (size_t)r * W + c;
        float dx = g[1] - g[-1];
        float dy = g[W] - g[-W];
        float mag = sqrtf(dx * dx + dy * dy);
        float ori = fmod360(atan2_f(dy, dx) * 57.29577951f + 360.0f);
        float dr = (float)r - ri, dc = (float)c - ci;
        float wt = exp_f(-(dr * dr + dc * dc) * inv_2sw2);
        float bin_f = ori * 0.1f;
        int bin_i = ((int)bin_f) % 36;
        int bin_r = (bin_i + 1) % 36;
        float frac = bin_f - floorf(bin_f);
        P[bin_i * 64 + lane] += mag * wt * (1.0f - frac);
        P[bin_r * 64 + lane] += mag * wt * frac;
    }
    float hist[36];
#pragma unroll
    for (int b = 0; b < 36; ++b) hist[b] = wave_tree_f(P[b * 64 + lane]);
    out[0] = out[1] = out[2] = out[3] = -1.0f;
    if (lane != 0) return 0;
    float smooth[36];
    for (int it = 0; it < 2; ++it) {
#pragma unroll
        for (int b = 0; b < 36; ++b)
            smooth[b] = (0.25f * hist[(b + 35) % 36] + 0.5f * hist[b]) + 0.25f * hist[(b + 1) % 36];
#pragma unroll
        for (int b = 0; b < 36; ++b) hist[b] = smooth[b];
    }
    float peak = 0.0f;
#pragma unroll
    for (int b = 0; b < 36; ++b) if (hist[b] > peak) peak = hist[b];
    float thresh = 0.8f * peak;
    int n = 0;
#pragma unroll
    for (int b = 0; b < 36; ++b) {
        if (hist[b] >= thresh) {
            float l = hist[(b + 35) % 36], rv = hist[(b + 1) % 36];
            if (hist[b] >= l && hist[b] >= rv) {
                float denom = (l - 2.0f * hist[b]) + rv;
                float fine = (float)b + (fabsf(denom) > 1e-12f ? (0.5f * (l - rv)) / denom : 0.0f);
                float angle = fmod360((fine / 36.0f) * 360.0f + 360.0f);
                if (n < 4) out[n++] = angle;
            }
        }
    }
    return n;
}

// ------------------------------------------------------------ descriptor --
// sift_descriptor_coop (core.py:693-835) + RootSIFT (core.py:1870-1873) + u8
// for one keypoint on one wave; lane l holds bins 2l, 2l+1 on return.
__device__ __forceinline__ void describe_keypoint(const float* __restrict__ G, int H, int W, float ri, float ci,
                                                  float angle_deg, int ph, int rootsift, float* P, int lane,
                                                  float* v0o, float* v1o) {
#pragma unroll
    for (int b = 0; b < 128; ++b) P[b * 64 + lane] = 0.0f;
    const float ang_rad = angle_deg * 0.0174532924f;
    double sd, cd;
    sincos_d((double)ang_rad, &sd, &cd);
    const float cos_a = (float)cd, sin_a = (float)sd;
    const float scale_sp = 4.0f / (2.0f * (float)ph);
    const float sigma_d = (float)ph * 0.5657f;
    const float inv_2s2 = 1.0f / (2.0f * sigma_d * sigma_d + 1e-12f);
    const int side = 2 * ph + 1;
    const int total_px = side * side;
    const int iri = (int)ri, ici = (int)ci;
    for (int flat = lane; flat < total_px; flat += 64) {
        int rq = flat / side;
        int r_off = rq - ph;
        int c_off = (flat - rq * side) - ph;
        int rr = iri + r_off, cc = ici + c_off;
        if (rr < 1 || rr >= H - 1 || cc < 1 || cc >= W - 1) continue;
        float xr = cos_a * (float)c_off - sin_a * (float)r_off;
        float yr = sin_a * (float)c_off + cos_a * (float)r_off;
        float xb = xr * scale_sp + 1.5f;
        float yb = yr * scale_sp + 1.5f;
        if (xb < -0.5f || xb >= 3.5f || yb < -0.5f || yb >= 3.5f) continue;
        const float* g = G + (size_t)rr * W + cc;
        float dx = g[1] - g[-1];
        float dy = g[W] - g[-W];
        float mag = sqrtf(dx * dx + dy * dy);
        float ori_deg = fmod360((atan2_f(dy, dx) - ang_rad) * 57.29577951f + 360.0f);
        float gw = exp_f(-(xr * xr + yr * yr) * inv_2s2);
        float wmag = mag * gw;
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
        float* h = P + lane;
        h[(yi0c * 32 + xi0c * 8 + oi0w) * 64] += wy0 * wx0 * wo0 * wmag;
        h[(yi0c * 32 + xi0c * 8 + oi1w) * 64] += wy0 * wx0 * wo1 * wmag;
        h[(yi0c * 32 + xi1c * 8 + oi0w) * 64] += wy0 * wx1 * wo0 * wmag;
        h[(yi0c * 32 + xi1c * 8 + oi1w) * 64] += wy0 * wx1 * wo1 * wmag;
        h[(yi1c * 32 + xi0c * 8 + oi0w) * 64] += wy1 * wx0 * wo0 * wmag;
        h[(yi1c * 32 + xi0c * 8 + oi1w) * 64] += wy1 * wx0 * wo1 * wmag;
        h[(yi1c * 32 + xi1c * 8 + oi0w) * 64] += wy1 * wx1 * wo0 * wmag;
        h[(yi1c * 32 + xi1c * 8 + oi1w) * 64] += wy1 * wx1 * wo1 * wmag;
    }
    float v0 = 0.0f, v1 = 0.0f;
#pragma unroll 8
    for (int b = 0; b < 128; ++b) {
        float t = wave_tree_f(P[b * 64 + lane]);
        if ((b >> 1) == lane) { if (b & 1) v1 = t; else v0 = t; }
    }
    float inv = 1.0f / sqrtf(wave_tree_f(v0 * v0 + v1 * v1) + 1e-14f);
    float t0 = v0 * inv, t1 = v1 * inv;
    v0 = t0 > 0.2f ? 0.2f : t0;
    v1 = t1 > 0.2f ? 0.2f : t1;
    inv = 1.0f / sqrtf(wave_tree_f(v0 * v0 + v1 * v1) + 1e-14f);
    v0 = v0 * inv;
    v1 = v1 * inv;
    if (rootsift) {
        float l1 = wave_tree_f(fabsf(v0) + fabsf(v1)) + 1e-7f;
        float q0 = v0 / l1, q1 = v1 / l1;
        v0 = sqrtf(q0 > 0.0f ? q0 : 0.0f);
        v1 = sqrtf(q1 > 0.0f ? q1 : 0.0f);
    }
    *v0o = v0;
    *v1o = v1;
}

__device__ __forceinline__ void store_descriptor(float v0, float v1, int lane, float* out_f, uint8_t* out_u8) {
    if (out_f) {
        out_f[2 * lane] = v0;
        out_f[2 * lane + 1] = v1;
    }
    if (out_u8) {
        int a0 = (int)(v0 * 512.0f + 0.5f), a1 = (int)(v1 * 512.0f + 0.5f);
        a0 = a0 > 255 ? 255 : (a0 < 0 ? 0 : a0);
        a1 = a1 > 255 ? 255 : (a1 < 0 ? 0 : a1);
        out_u8[2 * lane] = (uint8_t)a0;
        out_u8[2 * lane + 1] = (uint8_t)a1;
    }
}

}  // namespace msift
