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
// decimation, octave by octave from the top, down to the image.
#include "sift_dev.hpp"
#include <cstring>

#pragma clang fp contract(off)

namespace msift {

constexpr double kFix = 4294967296.0;   // fixed-point scale of the level gradients

struct GradPlan {
    int B, H, W, S, n_oct;
    int Ho[kMaxOct], Wo[kMaxOct];
    size_t oG[kMaxOct];    // G levels 0..S of octave o (floats), offsets within an image block
    size_t oGi[kMaxOct];   // int64 gradients of levels 1..S
    size_t oGf[kMaxOct];   // float gradients of levels 0..S
    size_t o_tmp;          // one level of scratch for the adjoint blur
    size_t per_img;        // bytes per image block (multiple of 256)
    size_t total;
};

static void grad_plan(const msift_handle* h, int B, int H, int W, GradPlan* P) {
    memset(P, 0, sizeof(*P));
    P->B = B; P->H = H; P->W = W; P->S = h->S; P->n_oct = h->p.n_octaves;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    int bh = H, bw = W;
    for (int o = 0; o < P->n_oct; ++o) {
        P->Ho[o] = bh; P->Wo[o] = bw;
        bh = (bh + 1) / 2; bw = (bw + 1) / 2;
    }
    for (int o = 0; o < P->n_oct; ++o) P->oG[o] = take((size_t)P->Ho[o] * P->Wo[o] * 4 * (P->S + 1));
    for (int o = 0; o < P->n_oct; ++o) P->oGi[o] = take((size_t)P->Ho[o] * P->Wo[o] * 8 * P->S);
    for (int o = 0; o < P->n_oct; ++o) P->oGf[o] = take((size_t)P->Ho[o] * P->Wo[o] * 4 * (P->S + 1));
    P->o_tmp = take((size_t)H * W * 4);
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

static inline unsigned gcap(size_t n, unsigned cap) { return n < 1 ? 1u : (n < cap ? (unsigned)n : cap); }

static int check_args(msift_handle* h, int B, int H, int W, int K, const void* imgs, const void* kps, void* ws) {
    if (!h || B < 1 || H < 3 || W < 3 || K < 0 || !imgs || (K > 0 && !kps) || !ws) {
        set_error("descriptor gradient: bad arguments");
        return MSIFT_EINVAL;
    }
    if (h->p.double_image) {
        set_error("descriptor gradient: needs double_image = 0 (the spline-upsample adjoint is not built)");
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
    int rc = blur(imgs, (float*)(ws + P.oG[0]), nullptr, nullptr, P.H, P.W, h->sd0[0], h->taps_base[0], s, P.B,
                  (size_t)P.H * P.W, zimg);
    if (rc) return rc;
    for (int o = 0; o < P.n_oct; ++o) {
        const size_t n = (size_t)P.Ho[o] * P.Wo[o];
        float* G = (float*)(ws + P.oG[o]);
        for (int lv = 0; lv < S; ++lv) {
            float* dec = (lv + 1 == S && o + 1 < P.n_oct) ? (float*)(ws + P.oG[o + 1]) : nullptr;
            if ((rc = blur(G + n * lv, G + n * (lv + 1), nullptr, dec, P.Ho[o], P.Wo[o], h->inc[lv], h->taps_inc[lv], s,
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
    blur_adjoint(P, ws, P.oGf[0], 0, H, W, h->sd0[0], h->taps_base[0], gimg, (size_t)H * W, 0, s);
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
