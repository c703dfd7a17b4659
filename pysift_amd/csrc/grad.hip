// Differentiable SIFT descriptor: descriptors of float images at fixed
// keypoints, and the hand-written backward to the pixels.  No reference
// counterpart (the reference has no gradient path); the forward is the
// reference's pyramid (core.py:1234-1352, single-size base) and descriptor
// (core.py:693-835 + RootSIFT core.py:1870-1873) on a float image in [0, 1].
//
// Backward: the pyramid is recomputed (never stored between the calls); each
// keypoint's histogram is recomputed and the normalisation chain
// (L2 -> clip 0.2 -> L2 -> RootSIFT) differentiated on one wave; every patch
// pixel then sends d/d(dx), d/d(dy) to the four Gaussian-level pixels of its
// central differences.  Those scatters from all keypoints are accumulated as
// int64 fixed point (2^-32) with integer atomics, which are order-free, so
// the gradient is bitwise reproducible.  The level gradients then run through
// the adjoint of each blur (reflect boundary folded back) and of the [::2, ::2]
// decimation, octave by octave from the top, down to the image.  With a
// doubled base (double_image, below double_px_cap pixels) the image is first
// upsampled x2 by the cubic B-spline of the detector (mirror-boundary FIR
// prefilter down the columns and along the rows, then interpolation, FP64);
// its adjoint runs the interpolation and the two prefilters transposed, also
// in FP64, as gathers (every output written by one thread, no atomics).
#include "sift_dev.hpp"
#include <cstring>

#pragma clang fp contract(off)

namespace msift {

constexpr double kFix = 4294967296.0;   // fixed-point scale of the level gradients

struct GradPlan {
    int B, H, W, S, n_oct, doubled;
    int Ho[kMaxOct], Wo[kMaxOct];
    size_t oG[kMaxOct];    // G levels 0..S of octave o (floats), offsets within an image block
    size_t oGi[kMaxOct];   // int64 gradients of levels 1..S
    size_t oGf[kMaxOct];   // float gradients of levels 0..S
    size_t o_tmp;          // one level of scratch for the adjoint blur
    // doubled base: upsampled image (2H x 2W float, its gradient in backward),
    // two H x W double planes (prefilter passes / their adjoints) and the
    // 2H x W double plane of the separable interpolation adjoint
    size_t o_up, o_d0, o_d1, o_r;
    size_t per_img;        // bytes per image block (multiple of 256)
    size_t total;
};

static void grad_plan(const msift_handle* h, int B, int H, int W, GradPlan* P) {
    memset(P, 0, sizeof(*P));
    P->B = B; P->H = H; P->W = W; P->S = h->S; P->n_oct = h->p.n_octaves;
    P->doubled = h->p.double_image && ((long long)H * W < (long long)h->p.double_px_cap);
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    int bh = P->doubled ? 2 * H : H, bw = P->doubled ? 2 * W : W;
    for (int o = 0; o < P->n_oct; ++o) {
        P->Ho[o] = bh; P->Wo[o] = bw;
        bh = (bh + 1) / 2; bw = (bw + 1) / 2;
    }
    for (int o = 0; o < P->n_oct; ++o) P->oG[o] = take((size_t)P->Ho[o] * P->Wo[o] * 4 * (P->S + 1));
    for (int o = 0; o < P->n_oct; ++o) P->oGi[o] = take((size_t)P->Ho[o] * P->Wo[o] * 8 * P->S);
    for (int o = 0; o < P->n_oct; ++o) P->oGf[o] = take((size_t)P->Ho[o] * P->Wo[o] * 4 * (P->S + 1));
    P->o_tmp = take((size_t)P->Ho[0] * P->Wo[0] * 4);
    if (P->doubled) {
        P->o_up = take((size_t)P->Ho[0] * P->Wo[0] * 4);
        P->o_d0 = take((size_t)H * W * 8);
        P->o_d1 = take((size_t)H * W * 8);
        P->o_r = take((size_t)2 * H * W * 8);
    }
    P->per_img = off;
    P->total = off * (size_t)B;
}

struct GradGeo {           // kernel-side view of the plan
    int S, n_oct;
    int Ho[kMaxOct], Wo[kMaxOct];
    size_t oG[kMaxOct], oGi[kMaxOct];
    size_t per_img;
};

static GradGeo grad_geo(const GradPlan& P) {
    GradGeo g;
    memset(&g, 0, sizeof(g));
    g.S = P.S; g.n_oct = P.n_oct; g.per_img = P.per_img;
    for (int o = 0; o < P.n_oct; ++o) {
        g.Ho[o] = P.Ho[o]; g.Wo[o] = P.Wo[o]; g.oG[o] = P.oG[o]; g.oGi[o] = P.oGi[o];
    }
    return g;
}

// ------------------------------------------------------------- forward ----
// keypoints kps[B*K][5] = (octave, scale index si, row, col (octave pixels),
// angle in degrees); an octave outside [0, n_oct) or si outside [1, S] marks
// an empty slot (zero descriptor, no gradient)
__device__ __forceinline__ bool kp_valid(const GradGeo& g, const float* kp, int* o, int* si) {
    *o = (int)kp[0];
    *si = (int)kp[1];
    return kp[0] >= 0.0f && *o < g.n_oct && *si >= 1 && *si <= g.S;
}

__global__ __launch_bounds__(64) void k_gdesc_fwd(const char* __restrict__ ws, GradGeo g, const float* __restrict__ kps,
                                                  int B, int K, KpTables tab, int rootsift, float* __restrict__ desc) {
    __shared__ float P[128 * 64];
    const int lane = threadIdx.x;
    for (int e = blockIdx.x; e < B * K; e += gridDim.x) {
        const float* kp = kps + (size_t)e * 5;
        int o, si;
        float v0 = 0.0f, v1 = 0.0f;
        if (kp_valid(g, kp, &o, &si)) {
            const float* G = (const float*)(ws + (size_t)(e / K) * g.per_img + g.oG[o]) + (size_t)si * g.Ho[o] * g.Wo[o];
            describe_keypoint(G, g.Ho[o], g.Wo[o], kp[2], kp[3], kp[4], tab.ph[si], rootsift, P, lane, &v0, &v1);
        }
        desc[(size_t)e * 128 + 2 * lane] = v0;
        desc[(size_t)e * 128 + 2 * lane + 1] = v1;
    }
}

// ------------------------------------------------------------ backward ----
__device__ __forceinline__ void fix_add(unsigned long long* p, float v) {
    long long q = __double2ll_rn((double)v * kFix);
    if (q) atomicAdd(p, (unsigned long long)q);
}

__global__ __launch_bounds__(64) void k_gdesc_bwd(char* __restrict__ ws, GradGeo g, const float* __restrict__ kps,
                                                  int B, int K, KpTables tab, int rootsift,
                                                  const float* __restrict__ gdesc) {
    __shared__ float P[128 * 64];
    __shared__ float sdh[128];
    const int lane = threadIdx.x;
    for (int e = blockIdx.x; e < B * K; e += gridDim.x) {
        const float* kp = kps + (size_t)e * 5;
        int o, si;
        if (!kp_valid(g, kp, &o, &si)) continue;
        const int H = g.Ho[o], W = g.Wo[o];
        char* img = ws + (size_t)(e / K) * g.per_img;
        const float* G = (const float*)(img + g.oG[o]) + (size_t)si * H * W;
        unsigned long long* dG = (unsigned long long*)(img + g.oGi[o]) + (size_t)(si - 1) * H * W;
        const float ri = kp[2], ci = kp[3], angle_deg = kp[4];
        const int ph = tab.ph[si];
        // forward histogram and normalisation chain, intermediates kept
        float h0, h1;
        desc_hist(G, H, W, ri, ci, angle_deg, ph, P, lane, &h0, &h1);
        const float inv1 = 1.0f / sqrtf(wave_tree_f(h0 * h0 + h1 * h1) + 1e-14f);
        const float t0 = h0 * inv1, t1 = h1 * inv1;
        const float u0 = t0 > 0.2f ? 0.2f : t0, u1 = t1 > 0.2f ? 0.2f : t1;
        const float inv2 = 1.0f / sqrtf(wave_tree_f(u0 * u0 + u1 * u1) + 1e-14f);
        const float w0 = u0 * inv2, w1 = u1 * inv2;
        float g0 = gdesc[(size_t)e * 128 + 2 * lane], g1 = gdesc[(size_t)e * 128 + 2 * lane + 1];
        if (rootsift) {
            // out = sqrt(max(q, 0)), q = w / (sum|w| + 1e-7)
            const float l1 = wave_tree_f(fabsf(w0) + fabsf(w1)) + 1e-7f;
            const float q0 = w0 / l1, q1 = w1 / l1;
            const float gq0 = q0 > 0.0f ? g0 * 0.5f / sqrtf(q0) : 0.0f;
            const float gq1 = q1 > 0.0f ? g1 * 0.5f / sqrtf(q1) : 0.0f;
            const float sd = wave_tree_f(gq0 * w0 + gq1 * w1) / (l1 * l1);
            const float sg0 = w0 > 0.0f ? 1.0f : (w0 < 0.0f ? -1.0f : 0.0f);
            const float sg1 = w1 > 0.0f ? 1.0f : (w1 < 0.0f ? -1.0f : 0.0f);
            g0 = gq0 / l1 - sg0 * sd;
            g1 = gq1 / l1 - sg1 * sd;
        }
        // w = u / ||u||: g_u = g_w / n - u (g_w . u) / n^3
        const float s2 = wave_tree_f(g0 * u0 + g1 * u1) * (inv2 * inv2 * inv2);
        float gu0 = g0 * inv2 - u0 * s2, gu1 = g1 * inv2 - u1 * s2;
        // clip at 0.2 (the forward keeps t when t <= 0.2)
        const float gt0 = t0 > 0.2f ? 0.0f : gu0, gt1 = t1 > 0.2f ? 0.0f : gu1;
        const float s1 = wave_tree_f(gt0 * h0 + gt1 * h1) * (inv1 * inv1 * inv1);
        sdh[2 * lane] = gt0 * inv1 - h0 * s1;
        sdh[2 * lane + 1] = gt1 * inv1 - h1 * s1;
        __syncthreads();
        // pixel pass: the forward's geometry, then d/d(dx), d/d(dy)
        const float ang_rad = angle_deg * 0.0174532924f;
        double sdd, cdd;
        sincos_d((double)ang_rad, &sdd, &cdd);
        const float cos_a = (float)cdd, sin_a = (float)sdd;
        const float scale_sp = 4.0f / (2.0f * (float)ph);
        const float sigma_d = (float)ph * 0.5657f;
        const float inv_2s2 = 1.0f / (2.0f * sigma_d * sigma_d + 1e-12f);
        const int side = 2 * ph + 1;
        const int total_px = side * side;
        const int iri = (int)ri, ici = (int)ci;
        const float kAng = 57.29577951f * 0.02222222f;   // d(ob)/d(theta)
        for (int flat = lane; flat < total_px; flat += 64) {
            const int rq = flat / side;
            const int r_off = rq - ph, c_off = (flat - rq * side) - ph;
            const int rr = iri + r_off, cc = ici + c_off;
            if (rr < 1 || rr >= H - 1 || cc < 1 || cc >= W - 1) continue;
            const float xr = cos_a * (float)c_off - sin_a * (float)r_off;
            const float yr = sin_a * (float)c_off + cos_a * (float)r_off;
            const float xb = xr * scale_sp + 1.5f;
            const float yb = yr * scale_sp + 1.5f;
            if (xb < -0.5f || xb >= 3.5f || yb < -0.5f || yb >= 3.5f) continue;
            const float* gp = G + (size_t)rr * W + cc;
            const float dx = gp[1] - gp[-1];
            const float dy = gp[W] - gp[-W];
            const float mag = sqrtf(dx * dx + dy * dy);
            if (!(mag > 0.0f)) continue;   // gradient of |(dx, dy)| and atan2 taken as 0 at the origin
            const float ori_deg = fmod360((atan2_f(dy, dx) - ang_rad) * 57.29577951f + 360.0f);
            const float gw = exp_f(-(xr * xr + yr * yr) * inv_2s2);
            const float ob = ori_deg * 0.02222222f;
            const int oi0 = (int)floorf(ob);
            const float df = ob - (float)oi0;
            const int xi0 = (int)floorf(xb), yi0 = (int)floorf(yb);
            const float dxb = xb - (float)xi0, dyb = yb - (float)yi0;
            const int xi0c = xi0 < 0 ? 0 : (xi0 > 3 ? 3 : xi0);
            const int yi0c = yi0 < 0 ? 0 : (yi0 > 3 ? 3 : yi0);
            const int xi1c = xi0 + 1 < 0 ? 0 : (xi0 + 1 > 3 ? 3 : xi0 + 1);
            const int yi1c = yi0 + 1 < 0 ? 0 : (yi0 + 1 > 3 ? 3 : yi0 + 1);
            const int oi0w = ((oi0 % 8) + 8) % 8;
            const int oi1w = ((oi0 + 1) % 8 + 8) % 8;
            const float wy0 = 1.0f - dyb, wy1 = dyb, wx0 = 1.0f - dxb, wx1 = dxb;
            // spatial weights of the 4 cells and the histogram gradient there
            const float s00 = wy0 * wx0, s01 = wy0 * wx1, s10 = wy1 * wx0, s11 = wy1 * wx1;
            const float a00 = sdh[yi0c * 32 + xi0c * 8 + oi0w], b00 = sdh[yi0c * 32 + xi0c * 8 + oi1w];
            const float a01 = sdh[yi0c * 32 + xi1c * 8 + oi0w], b01 = sdh[yi0c * 32 + xi1c * 8 + oi1w];
            const float a10 = sdh[yi1c * 32 + xi0c * 8 + oi0w], b10 = sdh[yi1c * 32 + xi0c * 8 + oi1w];
            const float a11 = sdh[yi1c * 32 + xi1c * 8 + oi0w], b11 = sdh[yi1c * 32 + xi1c * 8 + oi1w];
            const float A = (s00 * a00 + s01 * a01) + (s10 * a10 + s11 * a11);   // sum over the o0 bins
            const float Bs = (s00 * b00 + s01 * b01) + (s10 * b10 + s11 * b11);  // sum over the o1 bins
            // contribution = s * wo * mag * gw, wo0 = 1 - df, wo1 = df
            const float g_mag = gw * ((1.0f - df) * A + df * Bs);
            const float g_th = gw * mag * (Bs - A) * kAng;
            const float inv_m = 1.0f / mag;
            const float gdx = g_mag * dx * inv_m - g_th * dy * inv_m * inv_m;
            const float gdy = g_mag * dy * inv_m + g_th * dx * inv_m * inv_m;
            unsigned long long* q = dG + (size_t)rr * W + cc;
            fix_add(q + 1, gdx);
            fix_add(q - 1, -gdx);
            fix_add(q + W, gdy);
            fix_add(q - W, -gdy);
        }
        __syncthreads();
    }
}

// int64 fixed-point gradients -> float: level l (1..S) of Gf; level 0 zeroed
__global__ __launch_bounds__(256) void k_gfix_to_float(char* __restrict__ ws, size_t per_img, size_t oGi, size_t oGf,
                                                       size_t n, int S) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const long long* gi = (const long long*)(img + oGi);
    float* gf = (float*)(img + oGf);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n * (size_t)(S + 1); i += (size_t)gridDim.x * 256)
        gf[i] = i < n ? 0.0f : (float)((double)gi[i - n] * (1.0 / kFix));
}

__global__ __launch_bounds__(256) void k_zero_u64(char* __restrict__ ws, size_t per_img, size_t off, size_t n) {
    unsigned long long* p = (unsigned long long*)(ws + (size_t)blockIdx.y * per_img + off);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0ull;
}

// adjoint of dec = G[::2, ::2]: G_grad[2i][2j] += dec_grad[i][j]
__global__ __launch_bounds__(256) void k_decim_adj(char* __restrict__ ws, size_t per_img, size_t o_src, int h2, int w2,
                                                   size_t o_dst, int W) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const float* src = (const float*)(img + o_src);
    float* dst = (float*)(img + o_dst);
    const size_t n = (size_t)h2 * w2;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / w2), x = (int)(i - (size_t)y * w2);
        dst[(size_t)(2 * y) * W + 2 * x] += src[i];
    }
}

// adjoint of the 1-D reflect-boundary correlation y[i] = sum_k w[k] x[refl(i+k-R)]:
// x_grad[j] = sum_k w[k] sum_{i : refl(i+k-R) = j} y_grad[i].  refl(m) = j for
// m = j + 2Nt or m = -1 - j + 2Nt; one reflection suffices when N > R.
__device__ __forceinline__ float reflect_adj(const float* __restrict__ y, size_t stride, int N, int j, const Taps& t) {
    const int R = t.r;
    float acc = 0.0f;
    if (N > R) {
        for (int k = 0; k <= 2 * R; ++k) {
            const float w = t.w[k];
            const int i0 = j - k + R;              // m = j
            const int i1 = -1 - j - k + R;         // m = -1 - j (low edge)
            const int i2 = 2 * N - 1 - j - k + R;  // m = 2N - 1 - j (high edge)
            if (i0 >= 0 && i0 < N) acc = fmaf(w, y[(size_t)i0 * stride], acc);
            if (i1 >= 0 && i1 < N) acc = fmaf(w, y[(size_t)i1 * stride], acc);
            if (i2 >= 0 && i2 < N) acc = fmaf(w, y[(size_t)i2 * stride], acc);
        }
    } else {
        const int P2 = 2 * N, T = (R + P2) / P2 + 1;
        for (int k = 0; k <= 2 * R; ++k) {
            const float w = t.w[k];
            for (int tt = -T; tt <= T; ++tt) {
                const int i0 = j - k + R + P2 * tt, i1 = -1 - j - k + R + P2 * tt;
                if (i0 >= 0 && i0 < N) acc = fmaf(w, y[(size_t)i0 * stride], acc);
                if (i1 >= 0 && i1 < N && i1 != i0) acc = fmaf(w, y[(size_t)i1 * stride], acc);
            }
        }
    }
    return acc;
}

// vertical adjoint: tmp = V^T(src)
__global__ __launch_bounds__(256) void k_adj_v(const char* __restrict__ ws, size_t per_img, size_t o_src, size_t o_tmp,
                                               int H, int W, Taps t) {
    const char* img = ws + (size_t)blockIdx.y * per_img;
    const float* src = (const float*)(img + o_src);
    float* tmp = (float*)(img + o_tmp);
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        tmp[i] = reflect_adj(src + x, (size_t)W, H, y, t);
    }
}

// horizontal adjoint: dst (+)= H^T(tmp); dst_base NULL = inside the image block
__global__ __launch_bounds__(256) void k_adj_h(char* __restrict__ ws, size_t per_img, size_t o_tmp, size_t o_dst,
                                               int H, int W, Taps t, float* __restrict__ dst_base, size_t dst_stride,
                                               int accumulate) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const float* tmp = (const float*)(img + o_tmp);
    float* dst = dst_base ? dst_base + (size_t)blockIdx.y * dst_stride : (float*)(img + o_dst);
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        const float v = reflect_adj(tmp + (size_t)y * W, 1, W, x, t);
        dst[i] = accumulate ? dst[i] + v : v;
    }
}

// ------------------------------------------------- x2 spline upsample ----
// Forward on a float image (the detector's spline, core.py:1274, on the
// values themselves instead of u8 / 255): prefilter down the columns, then
// along the rows, sum_k h|k| x(mirror(i + k)), |k| <= kUpsK, FP64; then the
// cubic B-spline interpolation at output o -> o (N - 1) / (2N - 1).
__global__ __launch_bounds__(256) void k_gups_pf_col(const float* __restrict__ imgs, char* __restrict__ ws,
                                                     size_t per_img, size_t o_dst, int H, int W, UpsTaps h) {
    const float* x = imgs + (size_t)blockIdx.y * H * W;
    double* d = (double*)(ws + (size_t)blockIdx.y * per_img + o_dst);
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / W), c = (int)(i - (size_t)y * W);
        double acc = 0.0;
        if (y >= kUpsK && y + kUpsK < H) {
            for (int k = -kUpsK; k <= kUpsK; ++k)
                acc = fma(h.h[k < 0 ? -k : k], (double)x[(size_t)(y + k) * W + c], acc);
        } else {
            for (int k = -kUpsK; k <= kUpsK; ++k)
                acc = fma(h.h[k < 0 ? -k : k], (double)x[(size_t)mirror_idx(y + k, H) * W + c], acc);
        }
        d[i] = acc;
    }
}

__global__ __launch_bounds__(256) void k_gups_pf_row(char* __restrict__ ws, size_t per_img, size_t o_src,
                                                     size_t o_dst, int H, int W, UpsTaps h) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const double* a = (const double*)(img + o_src);
    double* d = (double*)(img + o_dst);
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / W), c = (int)(i - (size_t)y * W);
        const double* row = a + (size_t)y * W;
        double acc = 0.0;
        if (c >= kUpsK && c + kUpsK < W) {
            for (int k = -kUpsK; k <= kUpsK; ++k) acc = fma(h.h[k < 0 ? -k : k], row[c + k], acc);
        } else {
            for (int k = -kUpsK; k <= kUpsK; ++k) acc = fma(h.h[k < 0 ? -k : k], row[mirror_idx(c + k, W)], acc);
        }
        d[i] = acc;
    }
}

__global__ __launch_bounds__(256) void k_gups_interp(char* __restrict__ ws, size_t per_img, size_t o_cf, size_t o_up,
                                                     int H, int W) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const double* cf = (const double*)(img + o_cf);
    float* up = (float*)(img + o_up);
    const int Ho = 2 * H, Wo = 2 * W;
    const double zy = (double)(H - 1) / (double)(Ho - 1), zx = (double)(W - 1) / (double)(Wo - 1);
    const size_t n = (size_t)Ho * Wo;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int oy = (int)(i / Wo), ox = (int)(i - (size_t)oy * Wo);
        const double yy = (double)oy * zy, fy = floor(yy);
        const double xx = (double)ox * zx, fx = floor(xx);
        const int iy = (int)fy, ix = (int)fx;
        double wy[4], wx[4];
        bspline_w(yy - fy, wy);
        bspline_w(xx - fx, wx);
        double v = 0.0;
        for (int a = 0; a < 4; ++a) {
            const double* row = cf + (size_t)mirror_idx(iy - 1 + a, H) * W;
            double rv = 0.0;
            for (int b = 0; b < 4; ++b) rv = fma(wx[b], row[mirror_idx(ix - 1 + b, W)], rv);
            v = fma(wy[a], rv, v);
        }
        up[i] = (float)v;
    }
}

// Weight of coefficient j in interpolated output o along an axis of N
// coefficients and 2N outputs: the sum of the B-spline weights of the taps of
// o that mirror to j (zero if none does).
__device__ __forceinline__ double interp_weight(int o, int j, int N, double z) {
    const double t = (double)o * z, f = floor(t);
    const int i0 = (int)f - 1;
    double w[4];
    bspline_w(t - f, w);
    double acc = 0.0;
    for (int b = 0; b < 4; ++b)
        if (mirror_idx(i0 + b, N) == j) acc += w[b];
    return acc;
}

// Outputs whose taps can reach coefficient j (N >= 3).  Output o has its taps
// at positions floor(o z) - 1 .. floor(o z) + 2, within [-1, N + 1]; position
// p reaches j for p = j, p = -j (j = 1) and p = 2(N - 1) - j (j >= N - 3).
// Returns a hull [lo, hi] of all of them (the weights are tested exactly).
__device__ __forceinline__ void interp_range(int j, int N, double z, int* lo, int* hi) {
    const int No = 2 * N;
    int a = (int)floor((double)(j - 2) / z) - 1, b = (int)ceil((double)(j + 2) / z) + 1;
    if (j <= 2) a = 0;
    if (j >= N - 3) b = No - 1;
    *lo = a < 0 ? 0 : a;
    *hi = b > No - 1 ? No - 1 : b;
}

// interpolation adjoint along x: r[oy][j] = sum_ox w(ox -> j) g_up[oy][ox]
__global__ __launch_bounds__(256) void k_gups_interp_adj_x(char* __restrict__ ws, size_t per_img, size_t o_up,
                                                           size_t o_r, int H, int W) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const float* g = (const float*)(img + o_up);
    double* r = (double*)(img + o_r);
    const int Wo = 2 * W;
    const double zx = (double)(W - 1) / (double)(Wo - 1);
    const size_t n = (size_t)2 * H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int oy = (int)(i / W), j = (int)(i - (size_t)oy * W);
        int lo, hi;
        interp_range(j, W, zx, &lo, &hi);
        double acc = 0.0;
        for (int ox = lo; ox <= hi; ++ox) {
            const double w = interp_weight(ox, j, W, zx);
            if (w != 0.0) acc = fma(w, (double)g[(size_t)oy * Wo + ox], acc);
        }
        r[i] = acc;
    }
}

// interpolation adjoint along y: gcf[i][x] = sum_oy w(oy -> i) r[oy][x]
__global__ __launch_bounds__(256) void k_gups_interp_adj_y(char* __restrict__ ws, size_t per_img, size_t o_r,
                                                           size_t o_dst, int H, int W) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const double* r = (const double*)(img + o_r);
    double* d = (double*)(img + o_dst);
    const double zy = (double)(H - 1) / (double)(2 * H - 1);
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        int lo, hi;
        interp_range(y, H, zy, &lo, &hi);
        double acc = 0.0;
        for (int oy = lo; oy <= hi; ++oy) {
            const double w = interp_weight(oy, y, H, zy);
            if (w != 0.0) acc = fma(w, r[(size_t)oy * W + x], acc);
        }
        d[i] = acc;
    }
}

// Adjoint of the mirror-boundary prefilter d[i] = sum_k h|k| x[mirror(i + k)]
// along an axis of N samples, at index j: sum over (i, k) with
// mirror(i + k) = j of h|k| g[i].
// * N > 2 kUpsK + 2: i + k stays within one mirror fold, so the pairs are
//   k = j - i, k = -j - i (j >= 1) and k = 2(N - 1) - j - i (j <= N - 2), each
//   kept when |k| <= kUpsK; away from the edges only the first exists.
// * smaller N: the window folds several times; every pair is tested.
__device__ __forceinline__ double pf_adj(const double* __restrict__ g, size_t stride, int N, int j, const UpsTaps& h) {
    double acc = 0.0;
    if (N > 2 * kUpsK + 2) {
        if (j > kUpsK && j < N - 1 - kUpsK) {
            for (int k = -kUpsK; k <= kUpsK; ++k) acc = fma(h.h[k < 0 ? -k : k], g[(size_t)(j - k) * stride], acc);
            return acc;
        }
        const int lo = j - 2 * kUpsK < 0 ? 0 : j - 2 * kUpsK;
        const int hi = j + 2 * kUpsK > N - 1 ? N - 1 : j + 2 * kUpsK;
        for (int i = lo; i <= hi; ++i) {
            const double gi = g[(size_t)i * stride];
            const int k1 = j - i, k2 = -j - i, k3 = 2 * (N - 1) - j - i;
            if (k1 >= -kUpsK && k1 <= kUpsK) acc = fma(h.h[k1 < 0 ? -k1 : k1], gi, acc);
            if (j >= 1 && k2 >= -kUpsK && k2 <= kUpsK) acc = fma(h.h[k2 < 0 ? -k2 : k2], gi, acc);
            if (j <= N - 2 && k3 >= -kUpsK && k3 <= kUpsK) acc = fma(h.h[k3 < 0 ? -k3 : k3], gi, acc);
        }
        return acc;
    }
    for (int i = 0; i < N; ++i)
        for (int k = -kUpsK; k <= kUpsK; ++k)
            if (mirror_idx(i + k, N) == j) acc = fma(h.h[k < 0 ? -k : k], g[(size_t)i * stride], acc);
    return acc;
}

__global__ __launch_bounds__(256) void k_gups_pf_adj_row(char* __restrict__ ws, size_t per_img, size_t o_src,
                                                         size_t o_dst, int H, int W, UpsTaps h) {
    char* img = ws + (size_t)blockIdx.y * per_img;
    const double* g = (const double*)(img + o_src);
    double* d = (double*)(img + o_dst);
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        d[i] = pf_adj(g + (size_t)y * W, 1, W, x, h);
    }
}

__global__ __launch_bounds__(256) void k_gups_pf_adj_col(const char* __restrict__ ws, size_t per_img, size_t o_src,
                                                         float* __restrict__ gimg, int H, int W, UpsTaps h) {
    const double* g = (const double*)(ws + (size_t)blockIdx.y * per_img + o_src);
    float* d = gimg + (size_t)blockIdx.y * H * W;
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        d[i] = (float)pf_adj(g + x, (size_t)W, H, y, h);
    }
}

static inline unsigned gcap(size_t n, unsigned cap) { return n < 1 ? 1u : (n < cap ? (unsigned)n : cap); }


static int check_args(msift_handle* h, int B, int H, int W, int K, const void* imgs, const void* kps, void* ws) {
    if (!h || B < 1 || H < 3 || W < 3 || K < 0 || !imgs || (K > 0 && !kps) || !ws) {
        set_error("descriptor gradient: bad arguments");
        return MSIFT_EINVAL;
    }
    if (h->p.dsp_n_scales || h->p.fp16_pyramid) {
        set_error("descriptor gradient: the plain float32 descriptor only (dsp_n_scales = 0, fp16_pyramid = 0)");
        return MSIFT_EINVAL;
    }
    if (B > 65535) { set_error("descriptor gradient: at most 65535 images"); return MSIFT_EINVAL; }
    return MSIFT_OK;
}

// pyramid levels 0..S of every octave for the B images (one launch per level)
static int grad_pyramid(msift_handle* h, const GradPlan& P, const float* imgs, char* ws, hipStream_t s) {
    const int S = P.S;
    const size_t zimg = P.per_img / 4;
    int rc;
    if (P.doubled) {
        const size_t n = (size_t)P.H * P.W;
        dim3 grid(gcap((n + 255) / 256, 2048), P.B), grid2(gcap((4 * n + 255) / 256, 2048), P.B);
        hipLaunchKernelGGL(k_gups_pf_col, grid, dim3(256), 0, s, imgs, ws, P.per_img, P.o_d0, P.H, P.W, h->ups);
        hipLaunchKernelGGL(k_gups_pf_row, grid, dim3(256), 0, s, ws, P.per_img, P.o_d0, P.o_d1, P.H, P.W, h->ups);
        hipLaunchKernelGGL(k_gups_interp, grid2, dim3(256), 0, s, ws, P.per_img, P.o_d1, P.o_up, P.H, P.W);
        rc = blur((const float*)(ws + P.o_up), (float*)(ws + P.oG[0]), nullptr, P.Ho[0], P.Wo[0], h->sd0[1],
                  h->taps_base[1], s, P.B, zimg, zimg);
    } else {
        rc = blur(imgs, (float*)(ws + P.oG[0]), nullptr, P.H, P.W, h->sd0[0], h->taps_base[0], s, P.B,
                  (size_t)P.H * P.W, zimg);
    }
    if (rc) return rc;
    for (int o = 0; o < P.n_oct; ++o) {
        const size_t n = (size_t)P.Ho[o] * P.Wo[o];
        float* G = (float*)(ws + P.oG[o]);
        for (int lv = 0; lv < S; ++lv) {
            float* dec = (lv + 1 == S && o + 1 < P.n_oct) ? (float*)(ws + P.oG[o + 1]) : nullptr;
            if ((rc = blur(G + n * lv, G + n * (lv + 1), dec, P.Ho[o], P.Wo[o], h->inc[lv], h->taps_inc[lv], s,
                           P.B, zimg, zimg)))
                return rc;
        }
    }
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

// dst (+)= blur^T(src) for one level size; sigma < 0.01 is the identity
static void blur_adjoint(const GradPlan& P, char* ws, size_t o_src, size_t o_dst, int H, int W, double sigma,
                         const Taps& t, float* dst_base, size_t dst_stride, int accumulate, hipStream_t s) {
    const size_t n = (size_t)H * W;
    dim3 grid(gcap((n + 255) / 256, 2048), P.B);
    if (sigma < 0.01) {
        Taps id;
        memset(&id, 0, sizeof(id));
        id.r = 0; id.w[0] = 1.0f;
        hipLaunchKernelGGL(k_adj_v, grid, dim3(256), 0, s, (const char*)ws, P.per_img, o_src, P.o_tmp, H, W, id);
        hipLaunchKernelGGL(k_adj_h, grid, dim3(256), 0, s, ws, P.per_img, P.o_tmp, o_dst, H, W, id, dst_base,
                           dst_stride, accumulate);
        return;
    }
    hipLaunchKernelGGL(k_adj_v, grid, dim3(256), 0, s, (const char*)ws, P.per_img, o_src, P.o_tmp, H, W, t);
    hipLaunchKernelGGL(k_adj_h, grid, dim3(256), 0, s, ws, P.per_img, P.o_tmp, o_dst, H, W, t, dst_base, dst_stride,
                       accumulate);
}

int grad_workspace(msift_handle* h, int B, int H, int W, size_t* bytes) {
    if (!h || !bytes || B < 1 || H < 3 || W < 3) { set_error("grad workspace: bad arguments"); return MSIFT_EINVAL; }
    GradPlan P;
    grad_plan(h, B, H, W, &P);
    *bytes = P.total;
    return MSIFT_OK;
}

int grad_level_view(msift_handle* h, int B, int H, int W, void* ws, int b, int o, int* Ho, int* Wo, float** G,
                    long long** dG) {
    if (!h || !ws || b < 0 || b >= B || H < 3 || W < 3) { set_error("grad view: bad arguments"); return MSIFT_EINVAL; }
    GradPlan P;
    grad_plan(h, B, H, W, &P);
    if (o < 0 || o >= P.n_oct) { set_error("grad view: octave out of range"); return MSIFT_EINVAL; }
    char* img = (char*)ws + (size_t)b * P.per_img;
    *Ho = P.Ho[o];
    *Wo = P.Wo[o];
    *G = (float*)(img + P.oG[o]);
    *dG = (long long*)(img + P.oGi[o]);
    return MSIFT_OK;
}

int desc_forward(msift_handle* h, const float* imgs, int B, int H, int W, const float* kps, int K, void* ws_,
                 size_t ws_bytes, float* desc, hipStream_t s) {
    int rc = check_args(h, B, H, W, K, imgs, kps, ws_);
    if (rc) return rc;
    GradPlan P;
    grad_plan(h, B, H, W, &P);
    if (ws_bytes < P.total) { set_error("descriptor forward: workspace too small"); return MSIFT_EWORKSPACE; }
    if (K > 0 && !desc) { set_error("descriptor forward: null output"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    if ((rc = grad_pyramid(h, P, imgs, ws, s))) return rc;
    if (K > 0)
        hipLaunchKernelGGL(k_gdesc_fwd, dim3(gcap((size_t)B * K, 16384)), dim3(64), 0, s, (const char*)ws, grad_geo(P),
                           kps, B, K, h->tab, h->p.rootsift, desc);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

int desc_backward(msift_handle* h, const float* imgs, int B, int H, int W, const float* kps, int K,
                  const float* gdesc, void* ws_, size_t ws_bytes, float* gimg, hipStream_t s) {
    int rc = check_args(h, B, H, W, K, imgs, kps, ws_);
    if (rc) return rc;
    GradPlan P;
    grad_plan(h, B, H, W, &P);
    if (ws_bytes < P.total) { set_error("descriptor backward: workspace too small"); return MSIFT_EWORKSPACE; }
    if (!gimg || (K > 0 && !gdesc)) { set_error("descriptor backward: null buffer"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    const int S = P.S;
    if ((rc = grad_pyramid(h, P, imgs, ws, s))) return rc;
    for (int o = 0; o < P.n_oct; ++o) {
        const size_t n = (size_t)P.Ho[o] * P.Wo[o] * S;
        hipLaunchKernelGGL(k_zero_u64, dim3(gcap((n + 255) / 256, 2048), B), dim3(256), 0, s, ws, P.per_img, P.oGi[o], n);
    }
    if (K > 0)
        hipLaunchKernelGGL(k_gdesc_bwd, dim3(gcap((size_t)B * K, 16384)), dim3(64), 0, s, ws, grad_geo(P), kps, B, K,
                           h->tab, h->p.rootsift, gdesc);
    // octaves from the top: level gradients -> blur adjoints -> decimation adjoint
    for (int o = P.n_oct - 1; o >= 0; --o) {
        const int Ho = P.Ho[o], Wo = P.Wo[o];
        const size_t n = (size_t)Ho * Wo;
        hipLaunchKernelGGL(k_gfix_to_float, dim3(gcap((n * (S + 1) + 255) / 256, 2048), B), dim3(256), 0, s, ws,
                           P.per_img, P.oGi[o], P.oGf[o], n, S);
        if (o + 1 < P.n_oct)
            hipLaunchKernelGGL(k_decim_adj, dim3(gcap(((size_t)P.Ho[o + 1] * P.Wo[o + 1] + 255) / 256, 2048), B),
                               dim3(256), 0, s, ws, P.per_img, P.oGf[o + 1], P.Ho[o + 1], P.Wo[o + 1],
                               P.oGf[o] + n * 4 * S, Wo);
        for (int l = S; l >= 1; --l)
            blur_adjoint(P, ws, P.oGf[o] + n * 4 * l, P.oGf[o] + n * 4 * (l - 1), Ho, Wo, h->inc[l - 1],
                         h->taps_inc[l - 1], nullptr, 0, 1, s);
    }
    if (P.doubled) {
        // base blur adjoint into the upsampled image's gradient, then the
        // spline adjoint: interpolation along x and y, row and column prefilters
        blur_adjoint(P, ws, P.oGf[0], P.o_up, P.Ho[0], P.Wo[0], h->sd0[1], h->taps_base[1], nullptr, 0, 0, s);
        const size_t n = (size_t)H * W;
        dim3 grid(gcap((n + 255) / 256, 2048), B), grid2(gcap((2 * n + 255) / 256, 2048), B);
        hipLaunchKernelGGL(k_gups_interp_adj_x, grid2, dim3(256), 0, s, ws, P.per_img, P.o_up, P.o_r, H, W);
        hipLaunchKernelGGL(k_gups_interp_adj_y, grid, dim3(256), 0, s, ws, P.per_img, P.o_r, P.o_d1, H, W);
        hipLaunchKernelGGL(k_gups_pf_adj_row, grid, dim3(256), 0, s, ws, P.per_img, P.o_d1, P.o_d0, H, W, h->ups);
        hipLaunchKernelGGL(k_gups_pf_adj_col, grid, dim3(256), 0, s, (const char*)ws, P.per_img, P.o_d0, gimg, H, W,
                           h->ups);
    } else {
        blur_adjoint(P, ws, P.oGf[0], 0, H, W, h->sd0[0], h->taps_base[0], gimg, (size_t)H * W, 0, s);
    }
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift

using namespace msift;

extern "C" {

int msift_grad_workspace_bytes(msift_handle* h, int B, int H, int W, size_t* bytes) {
    return grad_workspace(h, B, H, W, bytes);
}

int msift_grad_level_view(msift_handle* h, int B, int H, int W, void* d_ws, int b, int o, int* Ho, int* Wo,
                          float** d_gauss, long long** d_grad_fixed) {
    return grad_level_view(h, B, H, W, d_ws, b, o, Ho, Wo, d_gauss, d_grad_fixed);
}

int msift_desc_forward(msift_handle* h, const float* d_imgs, int B, int H, int W, const float* d_kps, int K,
                       void* d_ws, size_t ws_bytes, float* d_desc, void* stream) {
    return desc_forward(h, d_imgs, B, H, W, d_kps, K, d_ws, ws_bytes, d_desc, (hipStream_t)stream);
}

int msift_desc_backward(msift_handle* h, const float* d_imgs, int B, int H, int W, const float* d_kps, int K,
                        const float* d_gdesc, void* d_ws, size_t ws_bytes, float* d_gimg, void* stream) {
    return desc_backward(h, d_imgs, B, H, W, d_kps, K, d_gdesc, d_ws, ws_bytes, d_gimg, (hipStream_t)stream);
}

}  // extern "C"
