// CLAHE on a u8 gray image in HBM (core.py:2562-2565: cv2.createCLAHE(clip,
// (t, t)).apply(gray) ahead of detection), with OpenCV's 8-bit semantics:
//   - if H or W is not a multiple of its tile count, the image is tiled as if
//     padded at the bottom by tiles_y - H % tiles_y rows and at the right by
//     tiles_x - W % tiles_x columns with BORDER_REFLECT_101 (both pads apply
//     as soon as either axis is ragged); tiles are tw x th of that extent;
//   - per tile: 256-bin histogram, clipped at `limit`, the clipped count
//     spread as clipped / 256 to every bin plus one to each
//     max(1, 256 / residual)-th bin from 0 while the residual lasts;
//     LUT[v] = round_even((float)cumsum[v] * (255.0f / area));
//   - each pixel blends the LUTs of its four nearest tile centres with
//     float weights, ((l11 xa1 + l12 xa) ya1 + (l21 xa1 + l22 xa) ya),
//     rounded to nearest even and saturated.
//
// Four launches: k_clahe_zero clears the per-tile histograms, k_clahe_hist
// builds them (blocks over (tile, band of rows), LDS per-wave histograms
// merged with one global atomic per bin), k_clahe_lut turns each into its
// LUT (one block per tile, a lane per bin), k_clahe_apply writes the output
// (4 pixels per lane, the LUTs its rectangle needs staged in LDS; src == dst is
// allowed, each lane reads its pixels before it writes them).  HBM traffic: the image is read twice and written once.
#include "msift_internal.hpp"
#include <climits>
#include <cmath>

#pragma clang fp contract(off)

namespace msift {

namespace {

constexpr int kClaheT = 256;          // threads per block; one histogram bin per lane in k_clahe_lut
constexpr int kClaheBandPx = 4096;    // pixels of a tile per k_clahe_hist block
constexpr int kClaheApplyPx = 4;      // pixels per lane in k_clahe_apply
constexpr int kClaheMaxTiles = 128;   // per axis
constexpr int kClaheLdsLuts = 64;     // LUTs a k_clahe_apply block stages in LDS (16 KiB)
constexpr int kClaheApplyBlocks = 512;  // k_clahe_apply blocks aimed for (rows per block adapt)

__device__ __forceinline__ int reflect101(int p, int len) {
    if (len == 1) return 0;
    while (p < 0 || p >= len) p = p < 0 ? -p : 2 * len - 2 - p;
    return p;
}

__device__ __forceinline__ uint8_t sat_round_u8(float v) {
    int r = __float2int_rn(v);
    return (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
}

__global__ __launch_bounds__(kClaheT) void k_clahe_zero(int* __restrict__ hist, int n) {
    int i = blockIdx.x * kClaheT + threadIdx.x;
    if (i < n) hist[i] = 0;
}

// grid (tiles, bands): block (t, b) counts rows [b*rows, (b+1)*rows) of tile
// t; rows past H (and columns past W) read their BORDER_REFLECT_101 source.
__global__ __launch_bounds__(kClaheT) void k_clahe_hist(const uint8_t* __restrict__ src, int H, int W, int pitch,
                                                        int tiles_x, int tw, int th, int rows,
                                                        int* __restrict__ hist) {
    __shared__ int h[kClaheT / 64][256];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int i = t; i < (kClaheT / 64) * 256; i += kClaheT) (&h[0][0])[i] = 0;
    __syncthreads();
    const int tile = blockIdx.x, ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int r0 = blockIdx.y * rows, r1 = min(r0 + rows, th);
    const int x0 = tx * tw;
    int* hw = h[wv];
    const bool x_inside = x0 + tw <= W;
    const bool vec4 = x_inside && ((x0 | tw | pitch) & 3) == 0 && (((uintptr_t)src) & 3) == 0;
    for (int r = r0 + wv; r < r1; r += kClaheT / 64) {
        int y = ty * th + r;
        if (y >= H) y = reflect101(y, H);
        const uint8_t* row = src + (size_t)y * pitch;
        if (vec4) {
            for (int c = lane * 4; c < tw; c += 256) {
                const uint32_t v = *reinterpret_cast<const uint32_t*>(row + x0 + c);
                atomicAdd(&hw[v & 255u], 1);
                atomicAdd(&hw[(v >> 8) & 255u], 1);
                atomicAdd(&hw[(v >> 16) & 255u], 1);
                atomicAdd(&hw[v >> 24], 1);
            }
        } else if (x_inside) {
            for (int c = lane; c < tw; c += 64) atomicAdd(&hw[row[x0 + c]], 1);
        } else {
            for (int c = lane; c < tw; c += 64) {
                int x = x0 + c;
                if (x >= W) x = reflect101(x, W);
                atomicAdd(&hw[row[x]], 1);
            }
        }
    }
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int k = 0; k < kClaheT / 64; ++k) s += h[k][t];
    if (s) atomicAdd(&hist[tile * 256 + t], s);
}

// one block per tile, lane b owns bin b
__global__ __launch_bounds__(kClaheT) void k_clahe_lut(const int* __restrict__ hist, int limit, float lut_scale,
                                                       uint8_t* __restrict__ lut) {
    __shared__ int part[kClaheT / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int h = hist[blockIdx.x * 256 + t];
    if (limit > 0) {
        int ex = h > limit ? h - limit : 0;
        h -= ex;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) ex += __shfl_xor(ex, off, 64);
        if (lane == 0) part[wv] = ex;
        __syncthreads();
        int clipped = 0;
#pragma unroll
        for (int k = 0; k < kClaheT / 64; ++k) clipped += part[k];
        const int batch = clipped / 256, residual = clipped - batch * 256;
        h += batch;
        if (residual != 0) {
            const int step = max(256 / residual, 1);
            if (t % step == 0 && t / step < residual) h += 1;
        }
        __syncthreads();
    }
    int v = h;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    if (lane == 63) part[wv] = v;
    __syncthreads();
    for (int k = 0; k < wv; ++k) v += part[k];
    lut[blockIdx.x * 256 + t] = sat_round_u8((float)v * lut_scale);
}

__device__ __forceinline__ int tile_lo(int p, float inv_t) { return (int)floorf((float)p * inv_t - 0.5f); }

// One output pixel from the LUT rows of tile rows ty1 / ty2 (l1 / l2, each
// indexed [tx * stride + v]): the bilinear blend of the four nearest tile
// LUTs with float weights, rounded to nearest even.
template <typename LutPtr>
__device__ __forceinline__ uint8_t clahe_px(int x, int v, float inv_tw, int tx_base, int tx_max, LutPtr l1, LutPtr l2,
                                            float ya, float ya1) {
    const float txf = (float)x * inv_tw - 0.5f;
    int tx1 = (int)floorf(txf), tx2 = tx1 + 1;
    const float xa = txf - (float)tx1, xa1 = 1.0f - xa;
    tx1 = max(tx1, 0) - tx_base;
    tx2 = min(tx2, tx_max) - tx_base;
    const float a = (float)l1[tx1 * 256 + v], b = (float)l1[tx2 * 256 + v];
    const float c = (float)l2[tx1 * 256 + v], d = (float)l2[tx2 * 256 + v];
    const float res = (a * xa1 + b * xa) * ya1 + (c * xa1 + d * xa) * ya;
    return sat_round_u8(res);
}

// grid (ceil(W / 1024), ceil(H / rows)): block (bx, by) writes columns
// [1024 bx, 1024 bx + 1024) of rows [rows by, rows by + rows), 4 pixels per
// lane per row.  The LUTs of the tiles whose centres bound that rectangle
// (tile columns tx_lo..tx_hi by rows ty_lo..ty_hi, found with the per-pixel
// formula at its corners, which is monotone in x and y) are staged in LDS
// when they fit kClaheLdsLuts; otherwise the block reads them from global
// memory.  Both paths compute the same expression.
__global__ __launch_bounds__(kClaheT) void k_clahe_apply(const uint8_t* src, int H, int W, int pitch, int tiles_x,
                                                         int tiles_y, float inv_tw, float inv_th, int rows,
                                                         const uint8_t* __restrict__ lut, uint8_t* dst) {
    __shared__ uint32_t s_lut[kClaheLdsLuts * 64];
    const int xb = blockIdx.x * kClaheT * kClaheApplyPx;
    const int y0 = blockIdx.y * rows, y1 = min(y0 + rows, H);
    const int xe = min(xb + kClaheT * kClaheApplyPx, W) - 1;
    const int tx_lo = max(tile_lo(xb, inv_tw), 0), tx_hi = min(tile_lo(xe, inv_tw) + 1, tiles_x - 1);
    const int ty_lo = max(tile_lo(y0, inv_th), 0), ty_hi = min(tile_lo(y1 - 1, inv_th) + 1, tiles_y - 1);
    const int ntx = tx_hi - tx_lo + 1, nty = ty_hi - ty_lo + 1;
    const bool staged = ntx * nty <= kClaheLdsLuts;
    if (staged) {
        for (int i = threadIdx.x; i < ntx * nty * 64; i += kClaheT) {
            const int k = i >> 6, w = i & 63, ty = k / ntx, tx = k - ty * ntx;
            s_lut[i] = reinterpret_cast<const uint32_t*>(lut + ((size_t)(ty_lo + ty) * tiles_x + tx_lo + tx) * 256)[w];
        }
        __syncthreads();
    }
    const int x = xb + threadIdx.x * kClaheApplyPx;
    if (x >= W) return;
    const uint8_t* s8 = reinterpret_cast<const uint8_t*>(s_lut);
    for (int y = y0; y < y1; ++y) {
        const float tyf = (float)y * inv_th - 0.5f;
        int ty1 = (int)floorf(tyf), ty2 = ty1 + 1;
        const float ya = tyf - (float)ty1, ya1 = 1.0f - ya;
        ty1 = max(ty1, 0);
        ty2 = min(ty2, tiles_y - 1);
        const uint8_t* srow = src + (size_t)y * pitch;
        uint8_t* drow = dst + (size_t)y * W;
        const bool vec4 = x + kClaheApplyPx <= W && (((uintptr_t)(srow + x) | (uintptr_t)(drow + x)) & 3) == 0;
        uint8_t in[kClaheApplyPx], out[kClaheApplyPx];
        const int n = min(kClaheApplyPx, W - x);
        if (vec4) {
            const uint32_t w4 = *reinterpret_cast<const uint32_t*>(srow + x);
#pragma unroll
            for (int k = 0; k < kClaheApplyPx; ++k) in[k] = (uint8_t)(w4 >> (8 * k));
        } else {
            for (int k = 0; k < n; ++k) in[k] = srow[x + k];
        }
        if (staged) {
            const uint8_t* l1 = s8 + (ty1 - ty_lo) * ntx * 256;
            const uint8_t* l2 = s8 + (ty2 - ty_lo) * ntx * 256;
#pragma unroll
            for (int k = 0; k < kClaheApplyPx; ++k)
                out[k] = clahe_px(x + k, in[k], inv_tw, tx_lo, tiles_x - 1, l1, l2, ya, ya1);
        } else {
            const uint8_t* l1 = lut + (size_t)ty1 * tiles_x * 256;
            const uint8_t* l2 = lut + (size_t)ty2 * tiles_x * 256;
#pragma unroll
            for (int k = 0; k < kClaheApplyPx; ++k)
                out[k] = clahe_px(x + k, in[k], inv_tw, 0, tiles_x - 1, l1, l2, ya, ya1);
        }
        if (vec4) {
            uint32_t w4 = 0;
#pragma unroll
            for (int k = 0; k < kClaheApplyPx; ++k) w4 |= (uint32_t)out[k] << (8 * k);
            *reinterpret_cast<uint32_t*>(drow + x) = w4;
        } else {
            for (int k = 0; k < n; ++k) drow[x + k] = out[k];
        }
    }
}

}  // namespace

size_t clahe_workspace(int tiles_x, int tiles_y) {
    return (((size_t)tiles_x * tiles_y * 256 * (sizeof(int) + 1)) + 255) & ~(size_t)255;
}

int clahe(const uint8_t* src, int H, int W, int pitch, double clip, int tiles_x, int tiles_y, void* ws,
          size_t ws_bytes, uint8_t* dst, hipStream_t s) {
    if (!src || !dst || !ws) { set_error("msift_clahe: null buffer"); return MSIFT_EINVAL; }
    if (H < 1 || W < 1 || pitch < W || tiles_x < 1 || tiles_y < 1 || tiles_x > kClaheMaxTiles ||
        tiles_y > kClaheMaxTiles || !(clip >= 0.0)) {
        set_error("msift_clahe: bad arguments (need H, W >= 1, pitch >= W, 1 <= tiles <= 128, clip >= 0)");
        return MSIFT_EINVAL;
    }
    if (((uintptr_t)ws & 3) != 0) { set_error("msift_clahe: workspace must be 4-byte aligned"); return MSIFT_EINVAL; }
    if (ws_bytes < clahe_workspace(tiles_x, tiles_y)) { set_error("msift_clahe: workspace too small"); return MSIFT_EWORKSPACE; }
    long long He = H, We = W;
    if (H % tiles_y != 0 || W % tiles_x != 0) { He = H + tiles_y - H % tiles_y; We = W + tiles_x - W % tiles_x; }
    const long long tw = We / tiles_x, th = He / tiles_y;
    if (tw * th > INT_MAX) { set_error("msift_clahe: tile area exceeds 2^31 - 1 pixels"); return MSIFT_EINVAL; }
    const int area = (int)(tw * th);
    int limit = 0;
    if (clip > 0.0) {
        const double l = clip * area / 256;
        limit = l >= (double)INT_MAX ? INT_MAX : (int)l;
        if (limit < 1) limit = 1;
    }
    const int n_tiles = tiles_x * tiles_y;
    int* hist = static_cast<int*>(ws);
    uint8_t* lut = reinterpret_cast<uint8_t*>(hist + (size_t)n_tiles * 256);
    hipLaunchKernelGGL(k_clahe_zero, dim3((n_tiles * 256 + kClaheT - 1) / kClaheT), dim3(kClaheT), 0, s, hist,
                       n_tiles * 256);
    MSIFT_LAUNCH_CHECK();
    const int rows = (int)std::max<long long>(kClaheT / 64, std::min<long long>(th, (kClaheBandPx + tw - 1) / tw));
    const int bands = (int)((th + rows - 1) / rows);
    hipLaunchKernelGGL(k_clahe_hist, dim3(n_tiles, bands), dim3(kClaheT), 0, s, src, H, W, pitch, tiles_x, (int)tw,
                       (int)th, rows, hist);
    MSIFT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_clahe_lut, dim3(n_tiles), dim3(kClaheT), 0, s, hist, limit, 255.0f / (float)area, lut);
    MSIFT_LAUNCH_CHECK();
    const int per_block = kClaheT * kClaheApplyPx;
    const int bx = (W + per_block - 1) / per_block;
    const long long want_rows = ((long long)H * bx + kClaheApplyBlocks - 1) / kClaheApplyBlocks;
    const int arows = (int)std::min<long long>(std::max<long long>(want_rows, 1), 16);
    const long long by = ((long long)H + arows - 1) / arows;
    const int rows_a = by > 65535 ? (H + 65534) / 65535 : arows;
    hipLaunchKernelGGL(k_clahe_apply, dim3(bx, (H + rows_a - 1) / rows_a), dim3(kClaheT), 0, s, src, H, W, pitch,
                       tiles_x, tiles_y, 1.0f / (float)tw, 1.0f / (float)th, rows_a, lut, dst);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift
