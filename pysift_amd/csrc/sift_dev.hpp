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
    int xi[4], yi[4];
    const bool inner = ix >= 1 && ix + 2 < W && iy >= 1 && iy + 2 < H;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        xi[b] = inner ? ix - 1 + b : mirror_idx(ix - 1 + b, W);
        yi[b] = inner ? iy - 1 + b : mirror_idx(iy - 1 + b, H);
    }
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double* row = cf + (size_t)yi[a] * W;
        double rv = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) rv = fma(wx[b], row[xi[b]], rv);
        v = fma(wy[a], rv, v);
    }
    return (float)v;
}

// tiled forms of the three spline kernels (whole block, 256 threads); every
// output is the same fma chain as ups_pf_col / ups_pf_row / ups_interp
constexpr int kUpsColRB = 16;             // column prefilter: rows per thread
constexpr int kUpsColTW = 64, kUpsColTH = 4 * kUpsColRB;
constexpr int kUpsRowPT = 8;              // row prefilter: outputs per thread
constexpr int kUpsRowTH = 32, kUpsRowTW = 8 * kUpsRowPT;   // its 32 x 64 tile
constexpr int kUpsRowSP = kUpsRowTW + 2 * kUpsK + 1;       // odd LDS row stride
constexpr int kUpsRowS = kUpsRowTH * kUpsRowSP;            // its LDS doubles
constexpr int kUpsInTW = 64, kUpsInTH = 16;   // interpolation output tile
constexpr int kUpsInSW = 36, kUpsInSH = 12;   // its input tile (see ups_interp_tile)
constexpr int kUpsInS = kUpsInSH * (kUpsInSW + kUpsInTW) + 5 * kUpsInTH;  // its LDS doubles

// (float)v / 255 for a byte v, without a division: the double product
// v * (1 / 255) is within 2^-52 relative of v / 255, while v / 255 (a fraction
// with denominator 255) is never within 2^-33 relative of a float rounding
// boundary, so rounding the product to float gives the correctly rounded
// quotient for every v in [0, 255] (checked exhaustively on the CPU).
__device__ __forceinline__ float u8_unit(int v) { return (float)((double)v * (1.0 / 255.0)); }

// column prefilter of a 64 x 32 tile: each thread walks kUpsColRB rows of one
// column with the 8 + 60 input values (u8 / 255 as float, rounded to binary16
// when q16, then double) held in registers
__device__ __forceinline__ void ups_cols_tile(const uint8_t* __restrict__ img, int pitch, int H, int W,
                                              const UpsTaps& h, double* __restrict__ out, int tx, int ty,
                                              bool q16) {
    const int x = tx * kUpsColTW + (int)(threadIdx.x & 63);
    const int y0 = ty * kUpsColTH + (int)(threadIdx.x >> 6) * kUpsColRB;
    if (x >= W || y0 >= H) return;
    double u[kUpsColRB + 2 * kUpsK];
    const bool inner = y0 >= kUpsK && y0 + kUpsColRB - 1 + kUpsK < H;
#pragma unroll
    for (int i = 0; i < kUpsColRB + 2 * kUpsK; ++i) {
        int yy = y0 - kUpsK + i;
        if (!inner) yy = mirror_idx(yy, H);
        const float v = u8_unit(img[(size_t)yy * pitch + x]);
        u[i] = (double)(q16 ? f16_round(v) : v);    // fp16 mode: the u8 / 255 base is float16
    }
#pragma unroll
    for (int j = 0; j < kUpsColRB; ++j) {
        if (y0 + j < H) {
            double acc = 0.0;
#pragma unroll
            for (int k = -kUpsK; k <= kUpsK; ++k) acc = fma(h.h[k < 0 ? -k : k], u[j + k + kUpsK], acc);
            out[(size_t)(y0 + j) * W + x] = acc;
        }
    }
}

// row prefilter of a 32-row x 64-column tile.  The mirrored input rows are
// staged in LDS with an odd row stride, so the lanes of a wave that read one
// column of different rows hit different banks.  Thread (r, g) = (tid % 32,
// tid / 32) walks the 8 + 2 kUpsK inputs of row r from column 8 g once,
// advancing the 8 outputs' fma chains, each in ascending tap order (the
// chain of ups_pf_row), with the taps held in registers.  s holds kUpsRowS
// doubles; ends with a barrier.
__device__ __forceinline__ void ups_rows_tile(const double* __restrict__ in, int H, int W, const UpsTaps& h,
                                              double* __restrict__ out, int tx, int ty, double* s) {
    const int x0 = tx * kUpsRowTW, y0 = ty * kUpsRowTH;
    constexpr int IW = kUpsRowTW + 2 * kUpsK;
    static_assert(IW <= 256, "one staged column per thread");
    if ((int)threadIdx.x < IW) {
        // column c of all 32 rows: every load issued before the LDS stores
        const int c = (int)threadIdx.x;
        const double* src = in + (size_t)y0 * W + mirror_idx(x0 - kUpsK + c, W);
        const int nr = H - y0 < kUpsRowTH ? H - y0 : kUpsRowTH;
        double v[kUpsRowTH];
#pragma unroll
        for (int r = 0; r < kUpsRowTH; ++r) v[r] = r < nr ? src[(size_t)r * W] : 0.0;
#pragma unroll
        for (int r = 0; r < kUpsRowTH; ++r) s[r * kUpsRowSP + c] = v[r];
    }
    __syncthreads();
    const int r = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int y = y0 + r;
    if (y < H) {
        const double* sr = s + r * kUpsRowSP + g * kUpsRowPT;
        double hv[kUpsK + 1];
#pragma unroll
        for (int k = 0; k <= kUpsK; ++k) hv[k] = h.h[k];
        double acc[kUpsRowPT];
#pragma unroll
        for (int u = 0; u < kUpsRowPT; ++u) acc[u] = 0.0;
#pragma unroll
        for (int m = 0; m < kUpsRowPT + 2 * kUpsK; ++m) {
            const double v = sr[m];
#pragma unroll
            for (int u = 0; u < kUpsRowPT; ++u) {
                const int k = m - u - kUpsK;             // tap of input m for output u
                if (k >= -kUpsK && k <= kUpsK) acc[u] = fma(hv[k < 0 ? -k : k], v, acc[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < kUpsRowPT; ++u) {
            const int x = x0 + g * kUpsRowPT + u;
            if (x < W) out[(size_t)y * W + x] = acc[u];
        }
    }
    __syncthreads();
}

// interpolation of a 64 x 16 output tile.  Its outputs read input rows
// floor(oy zy) - 1 .. + 2 and columns floor(ox zx) - 1 .. + 2; with zy, zx
// < 1/2 these span at most 12 rows and 36 columns from (cy0, cx0), staged
// (mirrored) in LDS.  The interpolation is separable: an output's inner chain
// over its 4 columns depends only on (input row, output column), so it is
// formed once per pair in tmp[kUpsInSH][64], and each output then runs its
// 4-term chain over rows -- the same fma chains as ups_interp.  The row
// weights and first input row of each output row are formed once, in wyt.
// q16 rounds the output to binary16.  s holds kUpsInS doubles; ends with a
// barrier.
__device__ __forceinline__ void ups_interp_tile(const double* __restrict__ cf, int H, int W, double zy, double zx,
                                                float* __restrict__ out, int tx, int ty, double* s, bool q16) {
    double* tmp = s + kUpsInSH * kUpsInSW;
    double* wyt = tmp + kUpsInSH * kUpsInTW;        // per output row: wy[4], row offset
    const int Wo = 2 * W, Ho = 2 * H;
    const int ox0 = tx * kUpsInTW, oy0 = ty * kUpsInTH;
    const int cy0 = (int)floor((double)oy0 * zy) - 1, cx0 = (int)floor((double)ox0 * zx) - 1;
    for (int i = threadIdx.x; i < kUpsInSH * kUpsInSW; i += 256) {
        const int r = i / kUpsInSW, c = i - r * kUpsInSW;
        s[i] = cf[(size_t)mirror_idx(cy0 + r, H) * W + mirror_idx(cx0 + c, W)];
    }
    if (threadIdx.x < kUpsInTH) {
        const int oy = oy0 + (int)threadIdx.x;
        double yy = (double)oy * zy;
        double fy = floor(yy);
        bspline_w(yy - fy, wyt + 5 * threadIdx.x);
        wyt[5 * threadIdx.x + 4] = (double)(((int)fy - 1 - cy0) * kUpsInTW);
    }
    __syncthreads();
    const int oxl = (int)(threadIdx.x & 63);
    const int ox = ox0 + oxl;
    {
        double xx = (double)ox * zx;
        double fx = floor(xx);
        const int ix = (int)fx;
        double wx[4];
        bspline_w(xx - fx, wx);
        for (int r = (int)(threadIdx.x >> 6); r < kUpsInSH; r += 4) {
            const double* srow = s + r * kUpsInSW + (ix - 1 - cx0);
            double rv = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) rv = fma(wx[b], srow[b], rv);
            tmp[r * kUpsInTW + oxl] = rv;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kUpsInTH / 4; ++j) {
        const int ly = (int)(threadIdx.x >> 6) * (kUpsInTH / 4) + j;
        const int oy = oy0 + ly;
        if (ox < Wo && oy < Ho) {
            const double* w = wyt + 5 * ly;
            const double* t = tmp + (int)w[4] + oxl;
            double v = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) v = fma(w[a], t[a * kUpsInTW], v);
            // fp16 mode: the zoomed base is cast to float16 (core.py:1274)
            out[(size_t)oy * Wo + ox] = q16 ? f16_round((float)v) : (float)v;
        }
    }
    __syncthreads();
}

// ----------------------------------------------------------- histogram --
// 256-bin histogram of rows blockIdx.x, + gridDim.x, ... of a u8 image, added
// into hist[256] with integer atomics (order-free, exact).  Lane l counts into
// LDS copy l % 32 (row stride 257, so lanes that see equal values hit
// different banks); the block sums the copies per bin at the end.
constexpr int kHistCopies = 32, kHistStride = 257;
__device__ __forceinline__ void hist256_body(const uint8_t* __restrict__ img, int H, int W, size_t pitch,
                                             int* __restrict__ hist, int* sh /* kHistCopies * kHistStride */) {
    const int tid = (int)threadIdx.x;
    for (int i = tid; i < kHistCopies * kHistStride; i += 256) sh[i] = 0;
    __syncthreads();
    int* mine = sh + (tid & (kHistCopies - 1)) * kHistStride;
    for (int y = blockIdx.x; y < H; y += gridDim.x) {
        const uint8_t* row = img + (size_t)y * pitch;
        int x = tid;
        for (; x + 768 < W; x += 1024) {
            const int v0 = row[x], v1 = row[x + 256], v2 = row[x + 512], v3 = row[x + 768];
            atomicAdd(mine + v0, 1); atomicAdd(mine + v1, 1); atomicAdd(mine + v2, 1); atomicAdd(mine + v3, 1);
        }
        for (; x < W; x += 256) atomicAdd(mine + row[x], 1);
    }
    __syncthreads();
    int t = 0;
    for (int c = 0; c < kHistCopies; ++c) t += sh[c * kHistStride + tid];
    if (t) atomicAdd(&hist[tid], t);
}

// ------------------------------------------------------- table lookups --
// Octave tables hold ExtOct entries in increasing job0 / row0.  For a key v
// and a start entry `from` (whose key is <= v) the entry that owns v is the
// last one with key <= v.  A short table in kernel arguments is scanned
// linearly; a long one in device memory is searched by the whole wave, 64
// entries per step, with a ballot.
template <class Tab>
__device__ __forceinline__ int tab_scan_job(const Tab& tab, int v, int from, int n) {
    while (from + 1 < n && tab.get(from + 1).job0 <= v) ++from;
    return from;
}
template <class Tab>
__device__ __forceinline__ int tab_scan_row(const Tab& tab, int v, int from, int n) {
    while (from + 1 < n && tab.get(from + 1).row0 <= v) ++from;
    return from;
}
template <class KeyAt>
__device__ __forceinline__ int wave_find_last_le(KeyAt key_at, int v, int from, int n) {
    const int lane = (int)(threadIdx.x & 63);
    int last = from;
    for (int b0 = from; b0 < n; b0 += 64) {
        const int k = b0 + lane;
        const unsigned long long m = __ballot(k < n && key_at(k) <= v);
        if (m) last = b0 + 63 - __clzll(m);
        if (m != ~0ull) break;
    }
    return last;
}

// ------------------------------------------------------------ placement --
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2:
// linear block b runs on XCD b % 8.  xcd_remap hands every XCD a contiguous
// range of logical tiles (a bijection of [0, n)), so the halo rows that
// neighbouring tiles share are fetched into one L2 instead of eight.
// Stable lexicographic rank of entries by (x, y, index): the rank of entry i
// is #{j : (x_j, y_j, j) < (x_i, y_i, i)}, added into rank[] (zero on entry)
// by integer atomics.  Grid x strides over 256-entry i-blocks, grid y over
// kRankJ-entry j-chunks staged in LDS; the chunk is read as float4 and the
// loop is fully unrolled so the 64 comparisons overlap their LDS latency.
constexpr int kRankJ = 64;
__device__ __forceinline__ void rank_xy_body(const Ent* __restrict__ E, int m, int* __restrict__ rank) {
    __shared__ float4 sxy[kRankJ / 2];                 // (x_j, y_j, x_j+1, y_j+1)
    const int nbi = (m + 255) / 256, nbj = (m + kRankJ - 1) / kRankJ;
    for (int jb = blockIdx.y; jb < nbj; jb += gridDim.y) {
        const int t0 = jb * kRankJ;
        __syncthreads();
        if (threadIdx.x < kRankJ) {
            const int j = t0 + threadIdx.x;
            float x = __builtin_inff(), y = __builtin_inff();
            if (j < m) { x = E[j].x; y = E[j].y; }
            float* f = reinterpret_cast<float*>(sxy);
            f[2 * threadIdx.x] = x;
            f[2 * threadIdx.x + 1] = y;
        }
        __syncthreads();
        for (int b = blockIdx.x; b < nbi; b += gridDim.x) {
            const int i = b * 256 + threadIdx.x;
            if (i >= m) continue;
            const float xi = E[i].x, yi = E[i].y;
            const int li = i - t0;                      // j < i  <=>  local j < li
            int r = 0;
#pragma unroll
            for (int q = 0; q < kRankJ / 2; ++q) {
                const float4 v = sxy[q];
                const int j = 2 * q;
                r += (v.x < xi) || (v.x == xi && (v.y < yi || (v.y == yi && j < li)));
                r += (v.z < xi) || (v.z == xi && (v.w < yi || (v.w == yi && j + 1 < li)));
            }
            if (r) atomicAdd(&rank[i], r);
        }
    }
}

__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int x = b & 7, i = b >> 3;
    const int q = n >> 3, r = n & 7;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// ------------------------------------------------------------------ blur --
// One 64 x TH output tile of G[s+1] = blur(G[s]) (H pass then V pass, fused
// multiply-add chains over the taps in ascending order) and the next
// octave's base G[S][::2, ::2].  The tile and its reflected halo are staged
// through LDS.  The DoG is not written here: its readers form it from the
// stored levels.
template <int R, int TH>
__device__ __forceinline__ void blur_tile(const float* __restrict__ in, float* __restrict__ gout,
                                          float* __restrict__ dec, int H, int W, int decW,
                                          const Taps& taps, int bx, int by, float* sIn) {
    constexpr int TW = 64;
    constexpr int IW = TW + 2 * R;
    constexpr int IWP = (IW + 4 + 3) & ~3;
    constexpr int IH = TH + 2 * R;
    constexpr int NV = 4 + 2 * R;
    constexpr int NV4 = (NV + 3) / 4;
    const int tid = threadIdx.x;
    const int x0 = bx * TW, y0 = by * TH;
    {
        // all loads issued before the LDS stores.  Element i = tid + 256 j of
        // the IH x IW input tile sits at (i / IW, i % IW); it advances by 256 =
        // DY rows + DX columns per j, tracked as running offsets.
        constexpr int NL = (IH * IW + 255) / 256;
        constexpr int DY = 256 / IW, DX = 256 - DY * IW;
        float tv[NL];
        const int ly0 = tid / IW, lx0 = tid - ly0 * IW;
        if (y0 >= R && y0 + TH + R <= H && x0 >= R && x0 + TW + R <= W) {
            // interior tile: no reflection
            const float* base = in + (size_t)(y0 - R) * W + (x0 - R);
            int lx = lx0;
            size_t off = (size_t)ly0 * W + lx0;
#pragma unroll
            for (int j = 0; j < NL; ++j) {
                if (j + 1 < NL || tid + 256 * j < IH * IW) tv[j] = base[off];
                off += (size_t)DY * W + DX;
                lx += DX;
                if (lx >= IW) { lx -= IW; off += W - IW; }
            }
        } else {
            // one reflection suffices unless the image is narrower than the halo or
            // the tile reaches past 2H (2W), as a tile taller than what is left of
            // the level can (the last 8-row tile of a 13-row level at R = 12 would
            // otherwise read row -2); the periodic form gives the same index wherever
            // one reflection suffices, so stored outputs are unchanged
            const bool small = (H <= R) || (W <= R) || (y0 + TH + R > 2 * H) || (x0 + TW + R > 2 * W);
            int ly = ly0, lx = lx0;
#pragma unroll
            for (int j = 0; j < NL; ++j) {
                if (j + 1 < NL || tid + 256 * j < IH * IW) {
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
                ly += DY; lx += DX;
                if (lx >= IW) { lx -= IW; ly += 1; }
            }
        }
        int soff = ly0 * IWP + lx0, lx = lx0;
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            if (j + 1 < NL || tid + 256 * j < IH * IW) sIn[soff] = tv[j];
            soff += DY * IWP + DX;
            lx += DX;
            if (lx >= IW) { lx -= IW; soff += IWP - IW; }
        }
    }
    __syncthreads();
    // H pass, in place: row ly's 64 outputs overwrite the first 64 floats of
    // row ly of the input tile.  A row's 16 float4 items fall in one wave and
    // one iteration, whose LDS reads all precede its writes, so a row is read
    // whole before it is overwritten, and no other wave touches it.
    constexpr int RG = TH / 4;
    const int lx = tid & 63, g = tid >> 6;
    for (int i = tid; i < IH * (TW / 4); i += 256) {
        int ly = i / (TW / 4), q = i - ly * (TW / 4);
        float* row = sIn + ly * IWP;
        float v[NV4 * 4];
#pragma unroll
        for (int k = 0; k < NV4; ++k) {
            float4 t = *reinterpret_cast<const float4*>(row + 4 * q + 4 * k);
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
        *reinterpret_cast<float4*>(row + 4 * q) = make_float4(o0, o1, o2, o3);
    }
    __syncthreads();
    float v[RG + 2 * R];
#pragma unroll
    for (int k = 0; k < RG + 2 * R; ++k) v[k] = sIn[(g * RG + k) * IWP + lx];
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
            if (dec && !(gy & 1) && !(gx & 1)) dec[(size_t)(gy >> 1) * decW + (gx >> 1)] = s;
        }
    }
}

template <int R, int TH>
struct BlurSmem {
    static constexpr int TW = 64;
    static constexpr int IWP = (TW + 2 * R + 4 + 3) & ~3;
    static constexpr int IH = TH + 2 * R;
    static constexpr int nIn = IH * IWP;   // the input tile; the H pass overwrites it row by row
};

// --------------------------------------------------------------- extrema --
// 3x3x3 extrema (core.py:260-340): D[s][r][c] with |v| > thr that is >= (or
// <=) all 26 neighbours at scales 1..S, rows/cols in [border, dim-border).
// The DoG is not stored: D[s] = G[s+1] - G[s] is formed where it is read,
// the subtraction of the pyramid (core.py:146), between the binary16 values
// in the fp16 mode's octaves (core.py:1306-1308, 1362-1364).
// One lane per column: a wave covers 62 interior columns (lanes 1..62; lanes
// 0 and 63 load the halo columns) and walks kExtRB rows of up to 3 scales,
// keeping each level's row maxima / minima over (c-1, c, c+1) for the rows
// r-1, r, r+1 in registers, so every Gaussian value is loaded once per strip.
// v >= every neighbour  <=>  v >= the max over the 3x3x3 cube including v.
// Output per (scale, row, 62-column segment): a 64-bit lane mask, and the
// per-(scale, row) candidate count (integer atomics, order-free).  The write
// pass enumerates the rows in (s, r) order and each row's bits in column
// order, which is the raster order of the reference's nonzero().
constexpr int kExtSeg = 62;
constexpr int kExtRB = 16;

struct ExtOct {          // extrema geometry of one (image, octave)
    size_t G_off;        // Gaussian stack (S+3 levels, H x W) in the workspace
    int q16;             // levels read as binary16 (fp16 mode)
    int H, W, Hi, Wi;    // size, interior size (minus 2*border)
    int nseg, nstrip;    // 62-column segments, kExtRB-row strips (0 = no interior)
    int img;             // image index (threshold and retry gate)
    int job0, row0;      // first wave job; first (scale, row) slot
    long long unit0;     // first (scale, row, segment) mask unit
};

// whole-wave DPP shifts: lane i receives lane i-1 (wave_shr:1) or lane i+1
// (wave_shl:1); the lane with no source receives 0
__device__ __forceinline__ float dpp_from_left(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float dpp_from_right(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ float max3f(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }
__device__ __forceinline__ float min3f(float a, float b, float c) { return fminf(a, fminf(b, c)); }

// one wave job: rows [strip*kExtRB, +kExtRB) x segment `seg` of an octave,
// all scales; must be called by the whole wave (uniform control flow)
template <bool Q16>
__device__ __forceinline__ float dog_value(float g1, float g0) {
    return Q16 ? f16_round(g1) - f16_round(g0) : g1 - g0;
}

template <bool Q16>
__device__ __forceinline__ void ext_strip(const float* __restrict__ G, const ExtOct& O, int S, int border,
                                          int strip, int seg, float thr, unsigned long long* __restrict__ masks,
                                          int* __restrict__ rowcnt) {
    const int lane = threadIdx.x & 63;
    const size_t HW = (size_t)O.H * O.W;
    const int ri0 = strip * kExtRB;
    const int nr = O.Hi - ri0 < kExtRB ? O.Hi - ri0 : kExtRB;
    const int cx = seg * kExtSeg + lane - 1;            // interior column of this lane
    const bool out_lane = lane >= 1 && lane <= kExtSeg && cx < O.Wi;
    int col = border + cx;
    if (col > O.W - 1) col = O.W - 1;                   // halo lane past the edge: unused value
    for (int s0 = 1; s0 <= S; s0 += 3) {
        // scales s0 .. s0+gs-1 read DoG levels s0-1 .. s0+gs, i.e. Gaussian
        // levels s0-1 .. s0+gs+1; slots past them repeat the top DoG level
        const int gs = S - s0 + 1 < 3 ? S - s0 + 1 : 3;
        // level bases are wave-uniform (scalar registers); a lane adds its
        // element offset r W + col (below H W, 32 bits)
        const float* lg[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            int L = s0 - 1 + (m < gs + 2 ? m : gs + 2);
            lg[m] = G + (size_t)L * HW;
        }
        float a0[5], a1[5], a2[5], b0[5], b1[5], b2[5], c1[5], c2[5], x[6];
#pragma unroll
        for (int l = 0; l < 5; ++l) { a1[l] = a2[l] = b1[l] = b2[l] = c2[l] = 0.0f; }
        auto load = [&](int r, float* g) {
            const unsigned off = (unsigned)r * (unsigned)O.W + (unsigned)col;
#pragma unroll
            for (int m = 0; m < 6; ++m) g[m] = lg[m][off];
        };
        auto push = [&](const float* g) {
            float v[5];
#pragma unroll
            for (int l = 0; l < 5; ++l) v[l] = dog_value<Q16>(g[l + 1], g[l]);
            if (gs < 3) {
                v[4] = gs == 2 ? v[3] : v[2];
                if (gs == 1) v[3] = v[2];
            }
#pragma unroll
            for (int l = 0; l < 5; ++l) {
                const float xl = dpp_from_left(v[l]), xr = dpp_from_right(v[l]);
                a0[l] = a1[l]; a1[l] = a2[l]; a2[l] = max3f(xl, v[l], xr);
                b0[l] = b1[l]; b1[l] = b2[l]; b2[l] = min3f(xl, v[l], xr);
                c1[l] = c2[l]; c2[l] = v[l];
            }
        };
        const int rA = border + ri0 - 1;
        load(rA, x); push(x);
        load(rA + 1, x); push(x);
        load(rA + 2, x); push(x);
        for (int i = 0; i < nr; ++i) {
            const bool more = i + 1 < nr;
            if (more) load(rA + 3 + i, x);
            float M[5], m[5];
#pragma unroll
            for (int l = 0; l < 5; ++l) { M[l] = max3f(a0[l], a1[l], a2[l]); m[l] = min3f(b0[l], b1[l], b2[l]); }
            const int ri = ri0 + i;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k < gs) {
                    const float v = c1[k + 1];
                    const bool pass = out_lane && fabsf(v) > thr &&
                                      (v >= max3f(M[k], M[k + 1], M[k + 2]) || v <= min3f(m[k], m[k + 1], m[k + 2]));
                    const unsigned long long bits = __ballot(pass);
                    if (lane == 0) {
                        const int rs = (s0 + k - 1) * O.Hi + ri;
                        masks[O.unit0 + (long long)rs * O.nseg + seg] = bits;
                        if (bits) atomicAdd(&rowcnt[O.row0 + rs], __popcll(bits));
                    }
                }
            }
            if (more) push(x);
        }
    }
}

// candidates of (scale, row) slot `rs` of an octave, in column order, at
// dst[q] for q = first + 0, 1, ... while q < capc; whole wave
__device__ __forceinline__ void ext_write_row(const ExtOct& O, int io, int rs, int border,
                                              const unsigned long long* __restrict__ masks, int first, int capc,
                                              int4* __restrict__ dst) {
    const int lane = threadIdx.x & 63;
    const int s = 1 + rs / O.Hi, ri = rs - (s - 1) * O.Hi;
    const long long ub = O.unit0 + (long long)rs * O.nseg;
    int carry = first;
    for (int j0 = 0; j0 < O.nseg && carry < capc; j0 += 64) {
        const int j = j0 + lane;
        unsigned long long m = j < O.nseg ? masks[ub + j] : 0ull;
        const int c = __popcll(m);
        int x = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        int q = carry + x - c;
        while (m && q < capc) {
            const int b = __builtin_ctzll(m);
            m &= m - 1;
            dst[q] = make_int4(s, border + ri, border + j * kExtSeg + b - 1, io);
            ++q;
        }
        carry += __shfl(x, 63);
    }
}

// host-side geometry of one octave's extrema jobs (counters advanced)
inline ExtOct ext_geometry(size_t G_off, int q16, int H, int W, int border, int S, int img, int* jobs, int* rows,
                           long long* units) {
    ExtOct e;
    e.G_off = G_off; e.q16 = q16; e.H = H; e.W = W; e.img = img;
    e.Hi = H - 2 * border; e.Wi = W - 2 * border;
    if (H < 3 || W < 3 || e.Hi <= 0 || e.Wi <= 0) { e.Hi = 0; e.Wi = 0; e.nseg = 0; e.nstrip = 0; }
    else { e.nseg = (e.Wi + kExtSeg - 1) / kExtSeg; e.nstrip = (e.Hi + kExtRB - 1) / kExtRB; }
    e.job0 = *jobs; e.row0 = *rows; e.unit0 = *units;
    *jobs += e.nseg * e.nstrip;
    *rows += S * e.Hi;
    *units += (long long)S * e.Hi * e.nseg;
    return e;
}

// count pass over a table of octaves: Tab::get(i) -> ExtOct; wave jobs are
// strided over the grid and each wave walks the table forward
template <class Tab>
__device__ __forceinline__ void ext_count_waves(const Tab& tab, int n_io, int njobs, const char* ws, int S,
                                                int border, const float* __restrict__ thr, int thr_idx,
                                                const int* __restrict__ gate, int below,
                                                unsigned long long* __restrict__ masks, int* __restrict__ rowcnt) {
    const int nw = gridDim.x * (blockDim.x >> 6);
    int io = 0;
    // the job index is wave-uniform: in a scalar register, with all it selects
    for (int j = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); j < njobs;
         j += nw) {
        io = tab.find_job(j, io, n_io);
        const ExtOct O = tab.get(io);
        if (gate && gate[O.img] >= below) continue;
        const int jj = j - O.job0;
        const int strip = jj / O.nseg, seg = jj - strip * O.nseg;
        if (O.q16)
            ext_strip<true>((const float*)(ws + O.G_off), O, S, border, strip, seg, thr[2 * O.img + thr_idx], masks,
                            rowcnt);
        else
            ext_strip<false>((const float*)(ws + O.G_off), O, S, border, strip, seg, thr[2 * O.img + thr_idx], masks,
                             rowcnt);
    }
}

// write pass: one wave per (octave, scale, row) slot with candidates.
// rowoff = exclusive scan of rowcnt; cbase (NULL: io * capc) = output base
template <class Tab>
__device__ __forceinline__ void ext_write_waves(const Tab& tab, int n_io, int nrows, int border,
                                                const unsigned long long* __restrict__ masks,
                                                const int* __restrict__ rowcnt, const int* __restrict__ rowoff,
                                                const int* __restrict__ cbase, int capc, int4* __restrict__ cands) {
    const int nw = gridDim.x * (blockDim.x >> 6);
    int io = 0;
    for (int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < nrows; g += nw) {
        if (rowcnt[g] == 0) continue;
        io = tab.find_row(g, io, n_io);
        const ExtOct O = tab.get(io);
        const int first = rowoff[g] - rowoff[O.row0];
        if (first >= capc) continue;
        int4* dst = cands + (cbase ? (size_t)cbase[io] : (size_t)io * capc);
        ext_write_row(O, io, g - O.row0, border, masks, first, capc, dst);
    }
}

// ---------------------------------------------------------------- refine --
// refine_keypoints (core.py:343-463) in FP64 on one octave's DoG, formed
// from the Gaussian stack G as the extrema form it; returns true and fills
// ri, ci, resp, si for an accepted keypoint
template <bool Q16>
__device__ __forceinline__ bool refine_candidate_t(const float* __restrict__ G, int H, int W, int S, int bd,
                                                   int si, int row, int col, double thr, float edge32, float* ri,
                                                   float* ci, float* resp_out, int* si_out) {
    const size_t HW = (size_t)H * W;
#define DV(s_, r_, c_)                                                                                   \
    ((double)dog_value<Q16>(G[(size_t)((s_) + 1) * HW + (size_t)(r_) * W + (c_)],                         \
                            G[(size_t)(s_) * HW + (size_t)(r_) * W + (c_)]))
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

__device__ __forceinline__ bool refine_candidate(const float* __restrict__ G, int q16, int H, int W, int S, int bd,
                                                 int si, int row, int col, double thr, float edge32, float* ri,
                                                 float* ci, float* resp_out, int* si_out) {
    return q16 ? refine_candidate_t<true>(G, H, W, S, bd, si, row, col, thr, edge32, ri, ci, resp_out, si_out)
               : refine_candidate_t<false>(G, H, W, S, bd, si, row, col, thr, edge32, ri, ci, resp_out, si_out);
}

// ----------------------------------------------------------- orientation --
// The window's pixels into lane-private 36-bin columns (orientation_hist,
// core.py:228-333): pixel p of the n_rows x n_cols window at (r_min, c_min)
// goes to lane p % 64's column P[bin*64 + ((lane + bin) & 63)].
template <bool Q16>
__device__ __forceinline__ void orient_accum(const float* __restrict__ G, int H, int W, float ri, float ci,
                                             float inv_2sw2, int r_min, int c_min, int n_rows, int n_cols,
                                             float* P, int lane) {
    const int total = n_rows * n_cols;
    // chunks of kPf pixels per lane: all gradient taps of a chunk are loaded
    // before any is used (memory-level parallelism); accumulation order per
    // lane stays p = lane, lane + 64, ...
    constexpr int kPf = 8;
    const int q64 = 64 / n_cols, r64 = 64 - q64 * n_cols;
    int rq = lane / n_cols, cq = lane - rq * n_cols;
    for (int p0 = lane; p0 < total; p0 += 64 * kPf) {
        float gl[kPf], gr[kPf], gu[kPf], gd[kPf];
        int rk[kPf], ck[kPf];
#pragma unroll
        for (int k = 0; k < kPf; ++k) {
            const int p = p0 + 64 * k;
            rk[k] = r_min + rq; ck[k] = c_min + cq;
            cq += r64; rq += q64;
            if (cq >= n_cols) { cq -= n_cols; rq += 1; }
            gl[k] = gr[k] = gu[k] = gd[k] = 0.0f;
            if (p < total) {
                const float* g = G + (size_t)rk[k] * W + ck[k];
                gl[k] = g[-1]; gr[k] = g[1]; gu[k] = g[-W]; gd[k] = g[W];
            }
            if (Q16) { gl[k] = f16_round(gl[k]); gr[k] = f16_round(gr[k]); gu[k] = f16_round(gu[k]); gd[k] = f16_round(gd[k]); }
        }
        // branch-free: pixels past the end add +0.0f (bins are sums of
        // non-negative terms, x + 0 == x bitwise)
#pragma unroll
        for (int k = 0; k < kPf; ++k) {
            const int p = p0 + 64 * k;
            int r = rk[k];
            int c = ck[k];
            float dx = gr[k] - gl[k];
            float dy = gd[k] - gu[k];
            float mag = sqrtf(dx * dx + dy * dy);
            // atan2 in [-pi, pi]: the argument lies in [179.99, 540.01]
            float ori = fmod360_lt720(atan2_f_finite(dy, dx) * 57.29577951f + 360.0f);
            float dr = (float)r - ri, dc = (float)c - ci;
            float wt = exp_f_le88(-(dr * dr + dc * dc) * inv_2sw2);
            // ori in [0, 360) gives bin_f in [0, 36): 359.99997f * 0.1f rounds
            // to 35.9999962f, so (int)bin_f % 36 == (int)bin_f
            float bin_f = ori * 0.1f;
            int bin_i = (int)bin_f;
            int bin_r = bin_i == 35 ? 0 : bin_i + 1;
            float frac = bin_f - floorf(bin_f);
            // two distinct bins: both reads issued before the stores.  The
            // partial of (bin, lane) sits at bin * 64 + ((lane + bin) & 63).
            float w0 = mag * wt * (1.0f - frac), w1 = mag * wt * frac;
            if (p >= total) { w0 = 0.0f; w1 = 0.0f; }
            float* s0 = P + bin_i * 64 + ((lane + bin_i) & 63);
            float* s1 = P + bin_r * 64 + ((lane + bin_r) & 63);
            const float q0 = *s0, q1 = *s1;
            *s0 = q0 + w0;
            *s1 = q1 + w1;
        }
    }
}

// orientation_hist (core.py:228-333) for one keypoint on one wave: pixel p of
// the window goes to lane p % 64's private 36-bin column P[bin*64 + lane].
// Every lane returns; lane 0's out[]/n are the result.  Q16: samples are read
// as the binary16 values of the fp16 mode's stored octaves.
template <bool Q16>
__device__ __forceinline__ int orient_keypoint(const float* __restrict__ G, int H, int W, float ri, float ci,
                                               float sigma, float* P, int lane, float* out) {
    __builtin_amdgcn_wave_barrier();     // the previous keypoint's reads of P are done
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
    orient_accum<Q16>(G, H, W, ri, ci, inv_2sw2, r_min, c_min, n_rows, n_cols, P, lane);
    // lane b < 36 owns bin b: it combines the 64 partials in the pairing of
    // the xor-butterfly tree (t[l] += t[l + off], off = 32 .. 1; wave_tree_f
    // gives every lane that value); the lane rotation spreads these reads
    // over the banks
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float h = 0.0f;
    if (lane < 36) {
        const float* row = P + lane * 64;
        float t[32];
#pragma unroll
        for (int l = 0; l < 32; ++l) t[l] = row[(l + lane) & 63] + row[(l + 32 + lane) & 63];
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1)
#pragma unroll
            for (int l = 0; l < off; ++l) t[l] = t[l] + t[l + off];
        h = t[0];
    }
    // two circular [1/4, 1/2, 1/4] passes over bins (lane b = bin b)
    const int bl = lane < 36 ? (lane + 35) % 36 : lane, br = lane < 36 ? (lane + 1) % 36 : lane;
    for (int it = 0; it < 2; ++it) {
        const float hl = __shfl(h, bl), hr = __shfl(h, br);
        h = (0.25f * hl + 0.5f * h) + 0.25f * hr;
    }
    if (lane >= 36) h = 0.0f;
    float peak = h;                      // max over bins and 0 (bins are >= 0)
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) peak = fmaxf(peak, __shfl_xor(peak, off));
    const float thresh = 0.8f * peak;
    const float l = __shfl(h, bl), rv = __shfl(h, br);
    const bool is_peak = lane < 36 && h >= thresh && h >= l && h >= rv;
    float angle = -1.0f;
    if (is_peak) {
        const float denom = (l - 2.0f * h) + rv;
        const float fine = (float)lane + (fabsf(denom) > 1e-12f ? (0.5f * (l - rv)) / denom : 0.0f);
        angle = fmod360((fine / 36.0f) * 360.0f + 360.0f);
    }
    // the first four peaks in bin order
    unsigned long long m = __ballot(is_peak);
    int n = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        out[k] = -1.0f;
        if (m) {
            const int b = __builtin_ctzll(m);
            m &= m - 1;
            out[k] = __shfl(angle, b);
            ++n;
        }
    }
    return n;
}

// ------------------------------------------------------------ descriptor --
// raw 4x4x8 histogram of sift_descriptor_coop (core.py:693-792) for one
// keypoint on one wave; lane l returns bins 2l, 2l+1
__device__ __forceinline__ void desc_hist(const float* __restrict__ G, int H, int W, float ri, float ci,
                                          float angle_deg, int ph, float* P, int lane, float* h0o, float* h1o) {
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
    // chunks of kPf pixels per lane, taps loaded before use (see orientation)
    constexpr int kPf = 8;
    // (row, col) of flat = lane + 64 i in the side x side patch, stepped by
    // 64 = q64 rows + r64 columns without a division per pixel
    const int q64 = 64 / side, r64 = 64 - q64 * side;
    int rq = lane / side, cq = lane - rq * side;
    for (int f0 = lane; f0 < total_px; f0 += 64 * kPf) {
        float gl[kPf], gr[kPf], gu[kPf], gd[kPf], xrs[kPf], yrs[kPf];
        bool okk[kPf];
#pragma unroll
        for (int k = 0; k < kPf; ++k) {
            const int flat = f0 + 64 * k;
            int r_off = rq - ph;
            int c_off = cq - ph;
            cq += r64; rq += q64;
            if (cq >= side) { cq -= side; rq += 1; }
            int rr = iri + r_off, cc = ici + c_off;
            float xr = cos_a * (float)c_off - sin_a * (float)r_off;
            float yr = sin_a * (float)c_off + cos_a * (float)r_off;
            float xb = xr * scale_sp + 1.5f;
            float yb = yr * scale_sp + 1.5f;
            bool ok = flat < total_px && !(rr < 1 || rr >= H - 1 || cc < 1 || cc >= W - 1) &&
                      !(xb < -0.5f || xb >= 3.5f || yb < -0.5f || yb >= 3.5f);
            okk[k] = ok;
            xrs[k] = xr; yrs[k] = yr;
            gl[k] = gr[k] = gu[k] = gd[k] = 0.0f;
            if (ok) {
                const float* g = G + (size_t)rr * W + cc;
                gl[k] = g[-1]; gr[k] = g[1]; gu[k] = g[-W]; gd[k] = g[W];
            }
        }
        // branch-free: a pixel outside the patch / window adds +0.0f to
        // in-range bins (every bin holds a sum of products of non-negative
        // factors, so x + 0 == x bitwise), which keeps the 8 pixels of a
        // chunk in one basic block for the scheduler
#pragma unroll
        for (int k = 0; k < kPf; ++k) {
            const float xr = xrs[k], yr = yrs[k];
            float xb = xr * scale_sp + 1.5f;
            float yb = yr * scale_sp + 1.5f;
            float dx = gr[k] - gl[k];
            float dy = gd[k] - gu[k];
            float mag = sqrtf(dx * dx + dy * dy);
            float ori_deg = fmod360((atan2_f_finite(dy, dx) - ang_rad) * 57.29577951f + 360.0f);
            float gw = exp_f_le88(-(xr * xr + yr * yr) * inv_2s2);
            float wmag = okk[k] ? mag * gw : 0.0f;
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
            // the 8 trilinear updates in the reference's order.  Targets
            // coincide only through the clamps (x pair: xi0c == xi1c, y pair:
            // yi0c == yi1c); all 8 reads are issued together and coinciding
            // targets continue from the running sum, so every bin receives
            // exactly the sequential chain of float additions
            float* h = P + lane;
            const int t0 = (yi0c * 32 + xi0c * 8 + oi0w) * 64, t1 = (yi0c * 32 + xi0c * 8 + oi1w) * 64;
            const int t2 = (yi0c * 32 + xi1c * 8 + oi0w) * 64, t3 = (yi0c * 32 + xi1c * 8 + oi1w) * 64;
            const int t4 = (yi1c * 32 + xi0c * 8 + oi0w) * 64, t5 = (yi1c * 32 + xi0c * 8 + oi1w) * 64;
            const int t6 = (yi1c * 32 + xi1c * 8 + oi0w) * 64, t7 = (yi1c * 32 + xi1c * 8 + oi1w) * 64;
            const bool xd = xi0c == xi1c, yd = yi0c == yi1c;
            const float p0 = h[t0], p1 = h[t1], p2 = h[t2], p3 = h[t3];
            const float p4 = h[t4], p5 = h[t5], p6 = h[t6], p7 = h[t7];
            const float r0 = p0 + wy0 * wx0 * wo0 * wmag;
            const float r1 = p1 + wy0 * wx0 * wo1 * wmag;
            const float r2 = (xd ? r0 : p2) + wy0 * wx1 * wo0 * wmag;
            const float r3 = (xd ? r1 : p3) + wy0 * wx1 * wo1 * wmag;
            const float r4 = (yd ? (xd ? r2 : r0) : p4) + wy1 * wx0 * wo0 * wmag;
            const float r5 = (yd ? (xd ? r3 : r1) : p5) + wy1 * wx0 * wo1 * wmag;
            const float r6 = (xd ? r4 : (yd ? r2 : p6)) + wy1 * wx1 * wo0 * wmag;
            const float r7 = (xd ? r5 : (yd ? r3 : p7)) + wy1 * wx1 * wo1 * wmag;
            // in-order stores: a coinciding target ends with its last value
            h[t0] = r0; h[t1] = r1; h[t2] = r2; h[t3] = r3;
            h[t4] = r4; h[t5] = r5; h[t6] = r6; h[t7] = r7;
        }
    }
    float v0 = 0.0f, v1 = 0.0f;
#pragma unroll 8
    for (int b = 0; b < 128; ++b) {
        float t = wave_tree_f(P[b * 64 + lane]);
        if ((b >> 1) == lane) { if (b & 1) v1 = t; else v0 = t; }
    }
    *h0o = v0;
    *h1o = v1;
}

// L2 -> clip 0.2 -> L2 (core.py:793-834) and RootSIFT (core.py:1870-1873)
// of a histogram held as (bins 2l, 2l+1) in lane l
__device__ __forceinline__ void desc_normalize(float v0, float v1, int rootsift, float* v0o, float* v1o) {
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

// sift_descriptor_coop (core.py:693-835) + RootSIFT for one keypoint on one
// wave; lane l holds bins 2l, 2l+1 on return
__device__ __forceinline__ void describe_keypoint(const float* __restrict__ G, int H, int W, float ri, float ci,
                                                  float angle_deg, int ph, int rootsift, float* P, int lane,
                                                  float* v0o, float* v1o) {
    float h0, h1;
    desc_hist(G, H, W, ri, ci, angle_deg, ph, P, lane, &h0, &h1);
    desc_normalize(h0, h1, rootsift, v0o, v1o);
}

// ----------------------------------------- descriptor, block per keypoint --
// The 4x4x8 histogram of sift_descriptor_coop (core.py:693-792) summed
// order-free.  A pixel's weight w = min(mag * gw * 2^30, 2^32 - 256) bounds
// each of its 8 trilinear contributions c = wy * wx * wo * w (the other
// factors lie in [0, 1]); c becomes the integer floor(c) and the bins are
// exact 64-bit sums, so every assignment of pixels to threads and every
// atomic order give the same bins; bin = (float)(sum 2^-30).  (The reference
// adds float contributions with atomics in no fixed order; the integer sum is
// within 2^-30 per contribution of the exact one.)  mag * gw stays below 2
// for images in [0, 1], so the clamp never acts on the detector's levels.
constexpr int kDescT = 256;            // threads per keypoint
constexpr int kDescCopies = kDescT / 64;   // one histogram copy per wave
constexpr float kDescFx = 0x1p30f;
constexpr float kDescFxMax = 0x1.fffffep31f;
// A wave's copy holds padded cells (y, x) for y, x in -1..4 at ((y + 1) 6 +
// (x + 1)) 9, each with orientation slots 0..8, slot 8 standing for bin 0
// reached as bin 7 + 1.  A pixel's 8 trilinear targets then sit at fixed
// offsets 0, 1, 9, 10, 54, 55, 63, 64 from one address.  The reference clamps
// the spatial targets to 0..3 and wraps the orientation; here the clamped and
// wrapped slots are added into their bins after the pixel loop, which gives
// the same integer sums.
constexpr int kDescCell = 9;
constexpr int kDescRow = 6 * kDescCell;
constexpr int kDescSlots = 6 * kDescRow;   // 324 per copy
constexpr int kDescHist = kDescCopies * kDescSlots;

// Pixel p of the side x side patch is visited as p = (f K) mod side^2 for
// f = tid + kDescT i, K = 97 (101 when 97 divides side): a bijection that
// sends the 64 lanes of a wave to pixels spread over the patch, so their
// atomics fall on different bins.  p advances by (kDescT K) mod side^2 as a
// (row, col) step.  hw = this wave's bins.
template <bool kInterior, bool Q16>
__device__ __forceinline__ void desc_hist_fx(const float* __restrict__ G, int H, int W, float ri, float ci,
                                             float angle_deg, int ph, unsigned long long* hw, int tid) {
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
    constexpr int kPf = 4;
    const int K = (side % 97) ? 97 : 101;
    const int step = (kDescT * K) % total_px;
    const int srow = step / side, scol = step - srow * side;
    const int p0 = (tid * K) % total_px;
    int rq = p0 / side, cq = p0 - rq * side;
    for (int f0 = tid; f0 < total_px; f0 += kDescT * kPf) {
        float gl[kPf], gr[kPf], gu[kPf], gd[kPf], xrs[kPf], yrs[kPf];
        bool okk[kPf];
#pragma unroll
        for (int k = 0; k < kPf; ++k) {
            const int flat = f0 + kDescT * k;
            const int r_off = rq - ph;
            const int c_off = cq - ph;
            cq += scol; rq += srow;
            if (cq >= side) { cq -= side; rq += 1; }
            if (rq >= side) rq -= side;
            const int rr = iri + r_off, cc = ici + c_off;
            const float xr = cos_a * (float)c_off - sin_a * (float)r_off;
            const float yr = sin_a * (float)c_off + cos_a * (float)r_off;
            const float xb = xr * scale_sp + 1.5f;
            const float yb = yr * scale_sp + 1.5f;
            bool ok = flat < total_px && !(xb < -0.5f || xb >= 3.5f || yb < -0.5f || yb >= 3.5f);
            if (!kInterior) ok = ok && !(rr < 1 || rr >= H - 1 || cc < 1 || cc >= W - 1);
            okk[k] = ok;
            xrs[k] = xr; yrs[k] = yr;
            if (kInterior) {
                // every (rr, cc) of the patch has its 4 neighbours in the level:
                // the taps are read for all pixels, the atomics skip the others
                const float* g = G + (size_t)rr * W + cc;
                gl[k] = g[-1]; gr[k] = g[1]; gu[k] = g[-W]; gd[k] = g[W];
            } else {
                gl[k] = gr[k] = gu[k] = gd[k] = 0.0f;
                if (ok) {
                    const float* g = G + (size_t)rr * W + cc;
                    gl[k] = g[-1]; gr[k] = g[1]; gu[k] = g[-W]; gd[k] = g[W];
                }
            }
            if (Q16) { gl[k] = f16_round(gl[k]); gr[k] = f16_round(gr[k]); gu[k] = f16_round(gu[k]); gd[k] = f16_round(gd[k]); }
        }
#pragma unroll
        for (int k = 0; k < kPf; ++k) {
            const float xr = xrs[k], yr = yrs[k];
            const float xb = xr * scale_sp + 1.5f;
            const float yb = yr * scale_sp + 1.5f;
            const float dx = gr[k] - gl[k];
            const float dy = gd[k] - gu[k];
            const float mag = sqrtf(dx * dx + dy * dy);
            // atan2 in [-pi, pi], ang_rad in [0, 2 pi): the argument lies in (-180.01, 540.01]
            const float ori_deg = fmod360_lt720((atan2_f_finite(dy, dx) - ang_rad) * 57.29577951f + 360.0f);
            const float gw = exp_f_mid(-(xr * xr + yr * yr) * inv_2s2);
            const float wms = fminf(mag * gw * kDescFx, kDescFxMax);
            const float ob = ori_deg * 0.02222222f;
            // (float)(int)floorf(v) == floorf(v) for the |v| < 2^24 here
            const float fo = floorf(ob), fx = floorf(xb), fy = floorf(yb);
            const int oi0 = (int)fo;
            const float df = ob - fo;
            const int xi0 = (int)fx, yi0 = (int)fy;
            const float dxb = xb - fx, dyb = yb - fy;
            // inside the window xb, yb lie in [-0.5, 3.5), so xi0, yi0 lie in
            // [-1, 3].  A pixel outside it adds nothing (its lanes skip the
            // atomics: sent to one slot they would serialise on it).  & 7 is
            // the residue mod 8 in [0, 8), for negative oi0 too.
            if (!okk[k]) continue;
            const int slot = __mul24(yi0, kDescRow) + __mul24(xi0, kDescCell) + (oi0 & 7) + kDescRow + kDescCell;
            const float wy0 = 1.0f - dyb, wy1 = dyb, wx0 = 1.0f - dxb, wx1 = dxb, wo0 = 1.0f - df, wo1 = df;
            unsigned long long* h = hw + slot;
            atomicAdd(h, (unsigned long long)(unsigned)(wy0 * wx0 * wo0 * wms));
            atomicAdd(h + 1, (unsigned long long)(unsigned)(wy0 * wx0 * wo1 * wms));
            atomicAdd(h + kDescCell, (unsigned long long)(unsigned)(wy0 * wx1 * wo0 * wms));
            atomicAdd(h + kDescCell + 1, (unsigned long long)(unsigned)(wy0 * wx1 * wo1 * wms));
            atomicAdd(h + kDescRow, (unsigned long long)(unsigned)(wy1 * wx0 * wo0 * wms));
            atomicAdd(h + kDescRow + 1, (unsigned long long)(unsigned)(wy1 * wx0 * wo1 * wms));
            atomicAdd(h + kDescRow + kDescCell, (unsigned long long)(unsigned)(wy1 * wx1 * wo0 * wms));
            atomicAdd(h + kDescRow + kDescCell + 1, (unsigned long long)(unsigned)(wy1 * wx1 * wo1 * wms));
        }
    }
}

// bin b = 32 y + 8 x + o of the reference's histogram from the padded copy:
// the padded cells that its clamp sends to (y, x) (row -1 to 0, row 4 to 3,
// likewise for columns), and slot 8 into orientation 0
__device__ __forceinline__ unsigned long long desc_bin_sum(const unsigned long long* h, int b) {
    const int y = b >> 5, x = (b >> 3) & 3, o = b & 7;
    const int ye = y == 0 ? 0 : (y == 3 ? 5 : -1);
    const int xe = x == 0 ? 0 : (x == 3 ? 5 : -1);
    unsigned long long s = 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int ys = i ? ye : y + 1;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int xs = j ? xe : x + 1;
            if (ys >= 0 && xs >= 0) {
                const unsigned long long* c = h + ys * kDescRow + xs * kDescCell;
                s += c[o];
                if (o == 0) s += c[8];
            }
        }
    }
    return s;
}

// sift_descriptor_coop (core.py:693-835), or with tab.n_dsp scales
// sift_descriptor_dsp_coop (core.py:845-997: one histogram over the patch
// half-sizes max(4, (int)(ph m + 0.5f)), divided by the number of scales),
// + RootSIFT, for one keypoint on a kDescT-thread block (every thread calls
// it): integer histogram in hist (kDescCopies padded copies, zeroed here), then wave
// 0 converts and normalises; thread l < 64 returns bins 2l, 2l+1.  q16: the
// level is read as binary16 (fp16 mode, octaves >= 2).
template <bool Q16>
__device__ __forceinline__ void desc_hist_scale(const float* __restrict__ G, int H, int W, float ri, float ci,
                                                float angle_deg, int ph, unsigned long long* hw, int tid) {
    const int iri = (int)ri, ici = (int)ci;
    if (iri - ph >= 2 && iri + ph < H - 2 && ici - ph >= 2 && ici + ph < W - 2)
        desc_hist_fx<true, Q16>(G, H, W, ri, ci, angle_deg, ph, hw, tid);
    else
        desc_hist_fx<false, Q16>(G, H, W, ri, ci, angle_deg, ph, hw, tid);
}

__device__ __forceinline__ void describe_keypoint_block(const float* __restrict__ G, int H, int W, float ri,
                                                        float ci, float angle_deg, int ph, int rootsift,
                                                        const KpTables& tab, bool q16,
                                                        unsigned long long* hist, float* v0o, float* v1o) {
    const int tid = threadIdx.x;
    // the block's barriers order LDS only: the previous keypoint's descriptor
    // stores need not drain before the next keypoint's histogram starts
    for (int i = tid; i < kDescHist; i += kDescT) hist[i] = 0ull;
    lds_barrier();
    unsigned long long* hw = hist + (tid >> 6) * kDescSlots;
    const int nd = tab.n_dsp > 0 ? tab.n_dsp : 1;
    for (int d = 0; d < nd; ++d) {
        int phd = ph;
        if (tab.n_dsp > 0) {
            phd = (int)((float)ph * tab.dsp[d] + 0.5f);
            phd = phd < 4 ? 4 : phd;
        }
        if (q16) desc_hist_scale<true>(G, H, W, ri, ci, angle_deg, phd, hw, tid);
        else desc_hist_scale<false>(G, H, W, ri, ci, angle_deg, phd, hw, tid);
    }
    lds_barrier();
    // the copies summed into copy 0, slot by slot
    for (int i = tid; i < kDescSlots; i += kDescT) {
        unsigned long long s = hist[i];
#pragma unroll
        for (int w = 1; w < kDescCopies; ++w) s += hist[w * kDescSlots + i];
        hist[i] = s;
    }
    lds_barrier();
    float v0 = 0.0f, v1 = 0.0f;
    if (tid < 64) {
        const unsigned long long s0 = desc_bin_sum(hist, 2 * tid), s1 = desc_bin_sum(hist, 2 * tid + 1);
        float h0 = (float)((double)s0 * 0x1p-30), h1 = (float)((double)s1 * 0x1p-30);
        if (tab.n_dsp > 0) {
            const float inv_n = 1.0f / (float)tab.n_dsp;
            h0 = h0 * inv_n;
            h1 = h1 * inv_n;
        }
        desc_normalize(h0, h1, rootsift, &v0, &v1);
    }
    *v0o = v0;
    *v1o = v1;
    lds_barrier();                         // hist is reused by the caller's next keypoint
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
