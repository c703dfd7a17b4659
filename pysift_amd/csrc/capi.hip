// C-ABI of libpysift_amd.so: parameter validation, per-handle precomputed
// tables, and dispatch to the kernel translation units.  No exception or
// torch type crosses this boundary.
#include "msift_internal.hpp"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#pragma clang fp contract(off)

namespace msift {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

// taps exp(-0.5 (i/sigma)^2) / sum, computed in double and rounded once
// (core.py:561-567: radius int(4 sigma + 0.5), normalised kernel)
void make_taps(double sigma, Taps* t) {
    int r = (int)(4.0 * sigma + 0.5);
    t->r = r;
    if (r > kMaxR) return;
    double v[2 * kMaxR + 1], sum = 0.0;
    for (int i = -r; i <= r; ++i) {
        double q = (double)i / sigma;
        v[i + r] = exp(-0.5 * (q * q));
    }
    for (int i = 0; i <= 2 * r; ++i) sum += v[i];
    for (int i = 0; i <= 2 * r; ++i) t->w[i] = (float)(v[i] / sum);
    for (int i = 2 * r + 1; i < 2 * kMaxR + 1; ++i) t->w[i] = 0.0f;
}

}  // namespace msift

using namespace msift;

extern "C" {

int msift_abi_version(void) { return MSIFT_ABI_VERSION; }

int msift_last_error(char* buf, size_t n) {
    if (!buf || n == 0) return MSIFT_EINVAL;
    snprintf(buf, n, "%s", g_err.c_str());
    return MSIFT_OK;
}

int msift_create(const msift_params* p, msift_handle** out) {
    if (!p || !out) { set_error("msift_create: null argument"); return MSIFT_EINVAL; }
    if (p->n_octaves < 1 || p->n_octaves > kMaxOct) { set_error("n_octaves must be in [1, 16]"); return MSIFT_EINVAL; }
    if (p->n_scales < 1 || p->n_scales > 28) { set_error("n_scales must be in [1, 28]"); return MSIFT_EINVAL; }
    if (!(p->sigma0 > 0.0) || !(p->edge_thresh > 0.0) || !(p->contrast_thresh >= 0.0)) {
        set_error("sigma0 and edge_thresh must be > 0, contrast_thresh >= 0");
        return MSIFT_EINVAL;
    }
    if (p->max_cands_per_octave < 1 || p->max_keypoints < 1 || p->border < 1) {
        set_error("caps and border must be positive");
        return MSIFT_EINVAL;
    }
    if (!(p->dsp_n_scales == 0 || p->dsp_n_scales == 3 || p->dsp_n_scales == 4 || p->dsp_n_scales == 5)) {
        set_error("dsp_n_scales must be 3, 4, or 5 (or 0 for the plain descriptor)");
        return MSIFT_EINVAL;
    }
    if (p->fp16_pyramid != 0 && p->fp16_pyramid != 1) {
        set_error("fp16_pyramid must be 0 or 1");
        return MSIFT_EINVAL;
    }
    msift_handle* h = new (std::nothrow) msift_handle();
    if (!h) { set_error("out of host memory"); return MSIFT_EINVAL; }
    memset(h, 0, sizeof(*h));
    h->p = *p;
    const int S = p->n_scales;
    h->S = S;
    // sigma schedule (core.py:1240-1242, 1279-1284)
    const double k = pow(2.0, 1.0 / (double)S);
    double sig[64];
    for (int s = 0; s < S + 3; ++s) sig[s] = p->sigma0 * pow(k, (double)s);
    for (int d = 0; d < 2; ++d) {
        double b = d ? 1.0 : 0.5;
        double v = p->sigma0 * p->sigma0 - b * b;
        h->sd0[d] = sqrt(v > 1e-6 ? v : 1e-6);
        make_taps(h->sd0[d], &h->taps_base[d]);
        if (h->taps_base[d].r > kMaxR) { set_error("base sigma too large"); delete h; return MSIFT_EINVAL; }
    }
    for (int s = 1; s < S + 3; ++s) {
        double q = sig[s] * sig[s] - sig[s - 1] * sig[s - 1];
        h->inc[s - 1] = sqrt(q > 1e-6 ? q : 1e-6);
        make_taps(h->inc[s - 1], &h->taps_inc[s - 1]);
        if (h->taps_inc[s - 1].r > kMaxR) { set_error("level sigma too large (radius > 32)"); delete h; return MSIFT_EINVAL; }
    }
    // spline prefilter impulse response sqrt(3) z^k, z = sqrt(3) - 2
    h->ups.h[0] = 0x1.bb67ae8584caap+0;
    for (int i = 1; i <= kUpsK; ++i) h->ups.h[i] = h->ups.h[i - 1] * -0x1.126145e9ecd58p-2;
    // per-scale tables (core.py:1552-1554, 1806-1810, 1518-1519)
    const float kf = (float)k;
    for (int si = 0; si <= S && si < 32; ++si) {
        float pw = (float)pow((double)kf, (double)si);
        h->tab.sig_ori[si] = (float)((float)p->sigma0 * pw);
        float t0 = (float)((float)(3.0 * p->sigma0) * pw);
        t0 = t0 * 1.41421354f;
        t0 = t0 * 2.5f;
        float phf = rintf(t0);
        h->tab.ph[si] = phf < 4.0f ? 4 : (int)phf;
        if (h->tab.ph[si] > 200) { set_error("descriptor patch too large"); delete h; return MSIFT_EINVAL; }
        h->tab.base_size[si] = p->sigma0 * pow(k, (double)si);
    }
    h->edge32 = (float)((p->edge_thresh + 1.0) * (p->edge_thresh + 1.0) / p->edge_thresh);
    // DSP-SIFT presets (core.py:1198-1202) as float32
    {
        const float r2 = (float)(1.0 / sqrt(2.0)), s2 = (float)sqrt(2.0);
        const float p3[3] = {r2, 1.0f, s2}, p4[4] = {0.5f, r2, 1.0f, s2}, p5[5] = {0.5f, r2, 1.0f, s2, 2.0f};
        const float* src = p->dsp_n_scales == 3 ? p3 : (p->dsp_n_scales == 4 ? p4 : p5);
        h->tab.n_dsp = p->dsp_n_scales;
        for (int i = 0; i < 5; ++i) h->tab.dsp[i] = i < p->dsp_n_scales ? src[i] : 0.0f;
    }
    h->tab.q16_oct = p->fp16_pyramid ? 2 : (1 << 30);
    *out = h;
    return MSIFT_OK;
}

int msift_destroy(msift_handle* h) {
    if (!h) return MSIFT_OK;
    if (h->ev_init)
        for (int i = 0; i < 10; ++i) (void)hipEventDestroy(h->ev[i]);
    msift::release_side_stream(h);
    delete h;
    return MSIFT_OK;
}

int msift_set_blur_tile(msift_handle* h, int tile_h) {
    if (!h || !(tile_h == 0 || tile_h == 8 || tile_h == 16 || tile_h == 32 || tile_h == 64)) {
        set_error("msift_set_blur_tile: tile_h must be 0, 8, 16, 32 or 64");
        return MSIFT_EINVAL;
    }
    h->blur_th = tile_h;
    return MSIFT_OK;
}

int msift_workspace_bytes(msift_handle* h, int H, int W, size_t* bytes) {
    if (!h || !bytes || H < 1 || W < 1) { set_error("msift_workspace_bytes: bad arguments"); return MSIFT_EINVAL; }
    return detect_workspace(h, H, W, bytes);
}

int msift_detect_describe(msift_handle* h, const uint8_t* d_img, int H, int W, int pitch, void* d_ws,
                          size_t ws_bytes, int cap, float* d_kp, float* d_desc_f32, uint8_t* d_desc_u8,
                          int* d_count, float* phase_ms, void* stream) {
    if (!h) { set_error("null handle"); return MSIFT_EINVAL; }
    return detect_describe(h, d_img, H, W, pitch, d_ws, ws_bytes, cap, d_kp, d_desc_f32, d_desc_u8, d_count,
                           phase_ms, (hipStream_t)stream);
}

int msift_pyramid_view(msift_handle* h, int H, int W, void* d_ws, int o, int* Ho, int* Wo, float** d_gauss,
                       float** d_dog) {
    if (!h || !d_ws || !Ho || !Wo || !d_gauss || !d_dog) { set_error("msift_pyramid_view: bad arguments"); return MSIFT_EINVAL; }
    return pyramid_view(h, H, W, d_ws, o, Ho, Wo, d_gauss, d_dog);
}

int msift_match_workspace_bytes(int N, int M, size_t* bytes) {
    if (N < 0 || M < 0 || !bytes) { set_error("msift_match_workspace_bytes: bad arguments"); return MSIFT_EINVAL; }
    *bytes = match_workspace(N, M);
    return MSIFT_OK;
}

int msift_match_u8(const uint8_t* d_a, int N, const uint8_t* d_b, int M, double ratio, int method, void* d_ws,
                   size_t ws_bytes, int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs, int* d_npairs,
                   void* stream) {
    if (N < 0 || M < 0 || !(ratio > 0.0) || (method < 0 || method > 2) || !d_npairs) {
        set_error("msift_match_u8: bad arguments");
        return MSIFT_EINVAL;
    }
    return match_u8(d_a, N, d_b, M, ratio, method, d_ws, ws_bytes, d_fwd, d_bwd, d_pairs, d_npairs,
                    (hipStream_t)stream);
}

int msift_match_u8_noalloc(const uint8_t* d_a, int N, const uint8_t* d_b, int M, double ratio, int method,
                           int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs, int* d_npairs, void* stream) {
    if (N < 0 || M < 0 || !(ratio > 0.0) || (method < 0 || method > 2) || !d_npairs) {
        set_error("msift_match_u8_noalloc: bad arguments");
        return MSIFT_EINVAL;
    }
    return match_u8_noalloc(d_a, N, d_b, M, ratio, method, d_fwd, d_bwd, d_pairs, d_npairs, (hipStream_t)stream);
}

int msift_match_set_tile(int shape, int chunks_per_block) { return match_set_tile(shape, chunks_per_block); }

int msift_match_f16(const float* d_a, int N, const float* d_b, int M, double ratio, void* d_ws, size_t ws_bytes,
                    int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs, int* d_npairs, void* stream) {
    if (N < 0 || M < 0 || !(ratio > 0.0) || !d_npairs) { set_error("msift_match_f16: bad arguments"); return MSIFT_EINVAL; }
    return match_f16(d_a, N, d_b, M, ratio, d_ws, ws_bytes, d_fwd, d_bwd, d_pairs, d_npairs, (hipStream_t)stream);
}

int msift_ransac_workspace_bytes(int N, int n_hyp, size_t* bytes) {
    if (N < 0 || n_hyp < 1 || !bytes) { set_error("msift_ransac_workspace_bytes: bad arguments"); return MSIFT_EINVAL; }
    *bytes = ransac_workspace(N, n_hyp);
    return MSIFT_OK;
}

int msift_ransac_homography(const double* d_p1, const double* d_p2, int N, uint64_t seed, int n_hyp,
                            double thr_px, double sigma_max, int method, int lo, int min_inliers, void* d_ws,
                            size_t ws_bytes, double* d_H, uint8_t* d_mask, int* d_ninl, int* d_best,
                            void* stream) {
    if (N < 4) { set_error("Need >= 4 correspondences"); return MSIFT_EINVAL; }
    if (n_hyp < 1 || !(thr_px > 0.0) || !(sigma_max > 0.0) || (method != 0 && method != 1)) {
        set_error("msift_ransac_homography: bad arguments");
        return MSIFT_EINVAL;
    }
    if (lo && min_inliers < 4) {
        /* the LO refit solves a DLT on the inliers, which needs 4 points
           (_compute_homography asserts it, core.py:2959-2961) */
        set_error("msift_ransac_homography: min_inliers must be >= 4 when lo is set");
        return MSIFT_EINVAL;
    }
    return ransac(d_p1, d_p2, N, seed, n_hyp, thr_px, sigma_max, method, lo, min_inliers, d_ws, ws_bytes, d_H,
                  d_mask, d_ninl, d_best, (hipStream_t)stream);
}

int msift_warp_feather(const uint8_t* d_img1, int h1, int w1, const uint8_t* d_img2, int h2, int w2,
                       const double* H, int cw, int ch, double ox, double oy, uint8_t* d_pano, uint8_t* d_valid,
                       void* stream) {
    if (h1 < 1 || w1 < 1 || h2 < 1 || w2 < 1 || cw < 1 || ch < 1 || !H) {
        set_error("msift_warp_feather: bad arguments");
        return MSIFT_EINVAL;
    }
    return warp_feather(d_img1, h1, w1, d_img2, h2, w2, H, cw, ch, ox, oy, d_pano, d_valid, (hipStream_t)stream);
}

int msift_bgr_to_gray(const uint8_t* d_bgr, int H, int W, uint8_t* d_gray, void* stream) {
    if (H < 1 || W < 1) { set_error("msift_bgr_to_gray: bad arguments"); return MSIFT_EINVAL; }
    return bgr_to_gray(d_bgr, H, W, d_gray, (hipStream_t)stream);
}

int msift_composite_workspace_bytes(int cw, int ch, int n_levels, size_t* bytes) {
    if (cw < 1 || ch < 1 || n_levels < 1 || n_levels > 12 || !bytes) {
        set_error("msift_composite_workspace_bytes: bad arguments (1 <= n_levels <= 12)");
        return MSIFT_EINVAL;
    }
    *bytes = composite_workspace(cw, ch, n_levels);
    return MSIFT_OK;
}

int msift_composite_multiband(const uint8_t* d_img1, int h1, int w1, const uint8_t* d_img2, int h2, int w2,
                              const double* H, int cw, int ch, double ox, double oy, int n_levels,
                              int color_match, void* d_ws, size_t ws_bytes, uint8_t* d_pano, uint8_t* d_valid,
                              float* d_seam_mask, msift_composite_stats* d_stats, void* stream) {
    return composite(d_img1, h1, w1, d_img2, h2, w2, H, cw, ch, ox, oy, n_levels, color_match, d_ws, ws_bytes, d_pano,
                     d_valid, d_seam_mask, d_stats, (hipStream_t)stream);
}

int msift_l2_normalize(const float* d_x, int n, float* d_y, void* stream) {
    return l2_normalize_rows(d_x, n, d_y, (hipStream_t)stream);
}

int msift_pca_workspace_bytes(int N, int M, size_t* bytes) {
    if (N < 0 || M < 0 || !bytes) { set_error("msift_pca_workspace_bytes: bad arguments"); return MSIFT_EINVAL; }
    *bytes = pca_workspace(N, M);
    return MSIFT_OK;
}

int msift_pca_fit(const float* d_a, int N, const float* d_b, int M, int dims, void* d_ws, size_t ws_bytes,
                  float* d_mean, float* d_proj, double* d_evals, int* d_sweeps, void* stream) {
    return pca_fit(d_a, N, d_b, M, dims, d_ws, ws_bytes, d_mean, d_proj, d_evals, d_sweeps, (hipStream_t)stream);
}

int msift_pca_project(const float* d_x, int n, const float* d_mean, const float* d_proj, int dims, float* d_y,
                      void* stream) {
    return pca_project(d_x, n, d_mean, d_proj, dims, d_y, (hipStream_t)stream);
}

int msift_match_sim_f16(const float* d_a, int N, const float* d_b, int M, int D, double ratio, void* d_ws,
                        size_t ws_bytes, int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs, int* d_npairs,
                        void* stream) {
    return match_sim16(d_a, N, d_b, M, D, ratio, d_ws, ws_bytes, d_fwd, d_bwd, d_pairs, d_npairs, (hipStream_t)stream);
}

int msift_clahe_workspace_bytes(int tiles_x, int tiles_y, size_t* bytes) {
    if (tiles_x < 1 || tiles_y < 1 || tiles_x > 128 || tiles_y > 128 || !bytes) {
        set_error("msift_clahe_workspace_bytes: bad arguments (1 <= tiles <= 128)");
        return MSIFT_EINVAL;
    }
    *bytes = clahe_workspace(tiles_x, tiles_y);
    return MSIFT_OK;
}

int msift_clahe(const uint8_t* d_src, int H, int W, int pitch, double clip_limit, int tiles_x, int tiles_y,
                void* d_ws, size_t ws_bytes, uint8_t* d_dst, void* stream) {
    return clahe(d_src, H, W, pitch, clip_limit, tiles_x, tiles_y, d_ws, ws_bytes, d_dst, (hipStream_t)stream);
}

}  // extern "C"
