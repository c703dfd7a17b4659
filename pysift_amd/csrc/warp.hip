// Canvas warp + feather blend (restates _warp_image / compositing of
// _stitch_pair, pysift/core.py:3210-3239, 3786-3802) and BGR->gray
// (cv2.cvtColor fixed-point BT.601, core.py:2548).
//
// For canvas pixel (u,v): source of img1 = H^-1 (u-ox, v-oy, 1) with H^-1
// rounded to float32 (core.py:3225), z clamped at 1e-10; bilinear sample with
// zero padding, valid when the source point lies in [0,w-1]x[0,h-1]
// (core.py:3233-3237).  img2 sits at (round(ox), round(oy)).  Each valid
// contribution is weighted by its feather weight min(x+1, w-x, y+1, h-y) in
// its own pixel coordinates; output = trunc(clip(sum w I / sum w)).
#include "msift_internal.hpp"
#include <cmath>

#pragma clang fp contract(off)

namespace msift {

struct Hf { float h[9]; };

__global__ __launch_bounds__(256) void k_warp_feather(const uint8_t* __restrict__ img1, int h1, int w1,
                                                       const uint8_t* __restrict__ img2, int h2, int w2, Hf hf,
                                                       int cw, int ch, float oxf, float oyf, int iox, int ioy,
                                                       uint8_t* __restrict__ pano, uint8_t* __restrict__ valid) {
    int u = blockIdx.x * 256 + threadIdx.x;
    int v = blockIdx.y;
    if (u >= cw) return;
    float X = (float)u - oxf, Y = (float)v - oyf;
    float sxn = (hf.h[0] * X + hf.h[1] * Y) + hf.h[2];
    float syn = (hf.h[3] * X + hf.h[4] * Y) + hf.h[5];
    float szn = (hf.h[6] * X + hf.h[7] * Y) + hf.h[8];
    if (szn < 1e-10f) szn = 1e-10f;
    float sx = sxn / szn, sy = syn / szn;
    float acc[3] = { 0.0f, 0.0f, 0.0f };
    float wsum = 0.0f;
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
        bool xr = x0 + 1 < w1, yb = y0 + 1 < h1;
        const uint8_t* r0 = img1 + ((size_t)y0 * w1 + x0) * 3;
        const uint8_t* r1 = img1 + ((size_t)(yb ? y0 + 1 : y0) * w1 + x0) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float v00 = r0[c];
            float v01 = xr ? (float)r0[3 + c] : 0.0f;
            float v10 = yb ? (float)r1[c] : 0.0f;
            float v11 = (xr && yb) ? (float)r1[3 + c] : 0.0f;
            float sv = ((w00 * v00 + w01 * v01) + w10 * v10) + w11 * v11;
            acc[c] = wf * sv;
        }
        wsum = wf;
    }
    int px = u - iox, py = v - ioy;
    if (px >= 0 && px < w2 && py >= 0 && py < h2) {
        float wf = (float)(px + 1);
        if ((float)(w2 - px) < wf) wf = (float)(w2 - px);
        if ((float)(py + 1) < wf) wf = (float)(py + 1);
        if ((float)(h2 - py) < wf) wf = (float)(h2 - py);
        const uint8_t* q = img2 + ((size_t)py * w2 + px) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + wf * (float)q[c];
        wsum = wsum + wf;
    }
    size_t o = (size_t)v * cw + u;
    valid[o] = wsum > 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float val = wsum > 0.0f ? acc[c] / wsum : 0.0f;
        if (val < 0.0f) val = 0.0f;
        if (val > 255.0f) val = 255.0f;
        pano[o * 3 + c] = (uint8_t)(int)val;
    }
}

__global__ __launch_bounds__(256) void k_bgr_to_gray(const uint8_t* __restrict__ bgr, int n, uint8_t* __restrict__ gray) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int b = bgr[3 * i], g = bgr[3 * i + 1], r = bgr[3 * i + 2];
    gray[i] = (uint8_t)((b * 1868 + g * 9617 + r * 4899 + 8192) >> 14);
}

// adjugate inverse with the determinant replaced by 1 when |det| <= 1e-12
static void inv3_host(const double* a, double* o) {
    double det = (a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6])) + a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (!(std::fabs(det) > 1e-12)) det = 1.0;
    o[0] = (a[4] * a[8] - a[5] * a[7]) / det; o[1] = (a[2] * a[7] - a[1] * a[8]) / det; o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    o[3] = (a[5] * a[6] - a[3] * a[8]) / det; o[4] = (a[0] * a[8] - a[2] * a[6]) / det; o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    o[6] = (a[3] * a[7] - a[4] * a[6]) / det; o[7] = (a[1] * a[6] - a[0] * a[7]) / det; o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

int warp_feather(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2, const double* H, int cw,
                 int ch, double ox, double oy, uint8_t* pano, uint8_t* valid, hipStream_t s) {
    if (!img1 || !img2 || !pano || !valid) { set_error("warp_feather: null buffer"); return MSIFT_EINVAL; }
    double Hi[9];
    inv3_host(H, Hi);
    Hf hf;
    for (int i = 0; i < 9; ++i) hf.h[i] = (float)Hi[i];
    int iox = (int)std::rint(ox), ioy = (int)std::rint(oy);
    hipLaunchKernelGGL(k_warp_feather, dim3((cw + 255) / 256, ch), dim3(256), 0, s, img1, h1, w1, img2, h2, w2, hf,
                       cw, ch, (float)ox, (float)oy, iox, ioy, pano, valid);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

int bgr_to_gray(const uint8_t* bgr, int H, int W, uint8_t* gray, hipStream_t s) {
    if (!bgr || !gray) { set_error("bgr_to_gray: null buffer"); return MSIFT_EINVAL; }
    int n = H * W;
    hipLaunchKernelGGL(k_bgr_to_gray, dim3((n + 255) / 256), dim3(256), 0, s, bgr, n, gray);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift
