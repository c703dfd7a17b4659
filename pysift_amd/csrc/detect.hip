// Detect + describe on one image: Gaussian pyramid (the DoG formed by its
// readers), 3x3x3 extrema with deterministic raster-order compaction, FP64
// sub-pixel refinement, orientation histograms in wave64 lane-private
// partials and 4x4x8 descriptor histograms as exact integer sums.  Restates
// PySIFT.detectAndCompute (pysift/core.py:2021-2172).
#include "sift_dev.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

#pragma clang fp contract(off)

namespace msift {

// ======================================================================
// misc small kernels
// ======================================================================
__global__ void k_zero_i32(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// 256-bin histogram of the u8 image (integer atomics: order-free, exact);
// per-wave LDS sub-histograms, rows strided over the grid
__global__ __launch_bounds__(256) void k_hist256(const uint8_t* __restrict__ img, int H, int W,
                                                  int pitch, int* __restrict__ hist) {
    __shared__ int sh[kHistCopies * kHistStride];
    hist256_body(img, H, W, (size_t)pitch, hist, sh);
}

// _adaptive_contrast_thresh (core.py:1129-1148): entropy in FP64; lane l
// sums bins l, l+64, l+128, l+192, lanes combine by the xor tree; writes
// thr[0] = f32(ct'/S), thr[1] = f32(ct'*0.5/S)
__global__ __launch_bounds__(64) void k_thresh(const int* __restrict__ hist, double total, double base, int S,
                                                int adaptive, float* __restrict__ thr) {
    const int l = threadIdx.x;
    double ct = base;
    if (adaptive) {
        double acc = 0.0;
        for (int j = 0; j < 4; ++j) {
            int c = hist[l + 64 * j];
            if (c == 0) continue;
            double p = (double)c / total;
            acc = acc - p * log2_d(p);
        }
        double ent = wave_tree_d(acc);
        double scale = 0.8 + ent / 40.0;
        if (scale > 1.0) scale = 1.0;
        ct = base * scale;
    }
    if (l == 0) {
        thr[0] = (float)(ct / (double)S);
        thr[1] = (float)((ct * 0.5) / (double)S);
    }
}

// ======================================================================
// A3: cubic B-spline x2 zoom (core.py:1274), mirror-boundary prefilter as a
// 61-tap symmetric FIR in FP64, then 4x4 tensor-product interpolation.
// ======================================================================
__global__ __launch_bounds__(256) void k_ups_pf_cols(const uint8_t* __restrict__ img, int pitch, int H, int W,
                                                      UpsTaps h, double* __restrict__ out, int q16) {
    const int ntx = (W + kUpsColTW - 1) / kUpsColTW, nt = ntx * ((H + kUpsColTH - 1) / kUpsColTH);
    for (int t = blockIdx.x; t < nt; t += gridDim.x)
        ups_cols_tile(img, pitch, H, W, h, out, t % ntx, t / ntx, q16 != 0);
}

__global__ __launch_bounds__(256) void k_ups_pf_rows(const double* __restrict__ in, int H, int W, UpsTaps h,
                                                      double* __restrict__ out) {
    __shared__ double s[kUpsRowS];
    const int ntx = (W + kUpsRowTW - 1) / kUpsRowTW, nt = ntx * ((H + kUpsRowTH - 1) / kUpsRowTH);
    for (int t = blockIdx.x; t < nt; t += gridDim.x) ups_rows_tile(in, H, W, h, out, t % ntx, t / ntx, s);
}

__global__ __launch_bounds__(256) void k_ups_interp(const double* __restrict__ cf, int H, int W, double zy, double zx,
                                                     float* __restrict__ out, int q16) {
    __shared__ double s[kUpsInS];
    const int ntx = (2 * W + kUpsInTW - 1) / kUpsInTW, nt = ntx * ((2 * H + kUpsInTH - 1) / kUpsInTH);
    for (int t = blockIdx.x; t < nt; t += gridDim.x)
        ups_interp_tile(cf, H, W, zy, zx, out, t % ntx, t / ntx, s, q16 != 0);
}

// u8 / 255 (core.py:1256); fp16 mode: a float16 array
__global__ __launch_bounds__(256) void k_u8_unit(const uint8_t* __restrict__ img, int pitch, int H,
                                                  int W, float* __restrict__ out, int q16) {
    int x = blockIdx.x * 256 + threadIdx.x;
    int y = blockIdx.y;
    if (x < W) {
        const float v = (float)img[(size_t)y * pitch + x] / 255.0f;
        out[(size_t)y * W + x] = q16 ? f16_round(v) : v;
    }
}

// a level whose sigma is below 0.01 equals its input (core.py:559-560)
__global__ __launch_bounds__(256) void k_copy_level(const float* __restrict__ in, float* __restrict__ gout,
                                                     float* __restrict__ dec, int H, int W,
                                                     int decW, size_t zin, size_t zout) {
    const size_t z = blockIdx.z;
    in += z * zin;
    if (gout) gout += z * zout;
    if (dec) dec += z * zout;
    size_t n = (size_t)H * W;
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float v = in[i];
    if (gout) gout[i] = v;
    int y = (int)(i / W), x = (int)(i - (size_t)y * W);
    if (dec && !(y & 1) && !(x & 1)) dec[(size_t)(y >> 1) * decW + (x >> 1)] = v;
}

// ======================================================================
// A2/A4: one pyramid level.  G[s+1] = blur(G[s]) (H pass then V pass, each
// output a fused-multiply-add chain over taps in ascending order, the
// reference's gauss_h/gauss_v, core.py:475-547), and optionally the next
// octave's base G[S][::2, ::2].  A 64 x TH output tile is staged through LDS
// with its reflected halo.  D[s] = G[s+1] - G[s] (core.py:146) is formed by
// the extrema and refinement kernels from the stored levels.
// ======================================================================
// gridDim.z images of one size: pointers advance by zin (input) and zout
// (outputs) floats per image
template <int R, int TH>
__global__ __launch_bounds__(256) void k_blur(const float* __restrict__ in, float* __restrict__ gout,
                                               float* __restrict__ dec,
                                               int H, int W, int decW, Taps taps, size_t zin, size_t zout) {
    __shared__ __attribute__((aligned(16))) float sIn[BlurSmem<R, TH>::nIn];
    const size_t z = blockIdx.z;
    const int t = xcd_remap((int)(blockIdx.x + blockIdx.y * gridDim.x), (int)(gridDim.x * gridDim.y));
    blur_tile<R, TH>(in + z * zin, gout ? gout + z * zout : nullptr, dec ? dec + z * zout : nullptr, H, W, decW,
                     taps, t % (int)gridDim.x, t / (int)gridDim.x, sIn);
}

// Tile height for a level: the tallest tile (least halo re-read) whose tile
// count still gives every CU four workgroups; small levels use short tiles
// (down to 8 rows below 256 tiles) so their latency is spread over more CUs.
// LDS bounds it by the radius (64 rows up to R=12, 32 up to R=24, else 16).
constexpr int kBlurMinTiles = 4 * 256;
static int blur_tile_h(int R, int H, int W, int nb, int th) {
    const int cap = R <= 12 ? 64 : (R <= 24 ? 32 : 16);
    if (th) return th < cap ? th : cap;
    const long long cols = (long long)((W + 63) / 64) * nb;
    for (int t = cap; t > 16; t >>= 1)
        if (cols * ((H + t - 1) / t) >= kBlurMinTiles) return t;
    return cols * ((H + 15) / 16) >= kBlurMinTiles / 4 ? 16 : 8;
}

template <int R, int TH>
static void blur_launch_th(const float* in, float* gout, float* dec, int H, int W, int decW,
                           const Taps& t, hipStream_t s, int nb, size_t zin, size_t zout) {
    dim3 grid((W + 63) / 64, (H + TH - 1) / TH, nb);
    hipLaunchKernelGGL((k_blur<R, TH>), grid, dim3(256), 0, s, in, gout, dec, H, W, decW, t, zin, zout);
}

template <int R>
static void blur_launch(const float* in, float* gout, float* dec, int H, int W, int decW,
                        const Taps& t, hipStream_t s, int nb, size_t zin, size_t zout, int th) {
    const int TH = blur_tile_h(R, H, W, nb, th);
    if constexpr (R <= 12) {
        if (TH == 64) { blur_launch_th<R, 64>(in, gout, dec, H, W, decW, t, s, nb, zin, zout); return; }
    }
    if constexpr (R <= 24) {
        if (TH == 32) { blur_launch_th<R, 32>(in, gout, dec, H, W, decW, t, s, nb, zin, zout); return; }
    }
    if (TH == 16) { blur_launch_th<R, 16>(in, gout, dec, H, W, decW, t, s, nb, zin, zout); return; }
    blur_launch_th<R, 8>(in, gout, dec, H, W, decW, t, s, nb, zin, zout);
}

int blur(const float* in, float* gout, float* dec, int H, int W, double sigma, const Taps& t,
         hipStream_t s, int nb, size_t zin, size_t zout, int th) {
    int decW = (W + 1) / 2;
    if (sigma < 0.01) {
        // _fast_gaussian returns its input unchanged (core.py:559-560)
        size_t n = (size_t)H * W;
        hipLaunchKernelGGL(k_copy_level, dim3((unsigned)((n + 255) / 256), 1, nb), dim3(256), 0, s, in, gout,
                           dec, H, W, decW, zin, zout);
        return MSIFT_OK;
    }
    switch (t.r) {
#define MSIFT_BLUR_CASE(RR) case RR: blur_launch<RR>(in, gout, dec, H, W, decW, t, s, nb, zin, zout, th); break;
        MSIFT_BLUR_CASE(0) MSIFT_BLUR_CASE(1) MSIFT_BLUR_CASE(2) MSIFT_BLUR_CASE(3) MSIFT_BLUR_CASE(4)
        MSIFT_BLUR_CASE(5) MSIFT_BLUR_CASE(6) MSIFT_BLUR_CASE(7) MSIFT_BLUR_CASE(8) MSIFT_BLUR_CASE(9)
        MSIFT_BLUR_CASE(10) MSIFT_BLUR_CASE(11) MSIFT_BLUR_CASE(12) MSIFT_BLUR_CASE(13) MSIFT_BLUR_CASE(14)
        MSIFT_BLUR_CASE(15) MSIFT_BLUR_CASE(16) MSIFT_BLUR_CASE(17) MSIFT_BLUR_CASE(18) MSIFT_BLUR_CASE(19)
        MSIFT_BLUR_CASE(20) MSIFT_BLUR_CASE(21) MSIFT_BLUR_CASE(22) MSIFT_BLUR_CASE(23) MSIFT_BLUR_CASE(24)
        MSIFT_BLUR_CASE(25) MSIFT_BLUR_CASE(26) MSIFT_BLUR_CASE(27) MSIFT_BLUR_CASE(28) MSIFT_BLUR_CASE(29)
        MSIFT_BLUR_CASE(30) MSIFT_BLUR_CASE(31) MSIFT_BLUR_CASE(32)
#undef MSIFT_BLUR_CASE
    default:
        set_error("blur radius above 32 (sigma too large)");
        return MSIFT_EINVAL;
    }
    return MSIFT_OK;
}

// ======================================================================
// A5: 3x3x3 extrema (core.py:176-217).  Wave jobs (octave, 16-row strip,
// 62-column segment) mark extrema in per-(scale, row, segment) lane masks
// and count them per (scale, row); a scan of the row counts gives each row's
// raster-order position; the write pass keeps the first `capc` per octave.
// ======================================================================
struct ExtTabArg {
    ExtOct o[kMaxOct];
    __device__ __forceinline__ const ExtOct& get(int i) const { return o[i]; }
    __device__ __forceinline__ int find_job(int v, int from, int n) const { return tab_scan_job(*this, v, from, n); }
    __device__ __forceinline__ int find_row(int v, int from, int n) const { return tab_scan_row(*this, v, from, n); }
};

__global__ __launch_bounds__(256) void k_ext_count(ExtTabArg tab, int n_oct, int njobs, const char* ws, int S,
                                                    int border, const float* __restrict__ thr, int thr_idx,
                                                    const int* __restrict__ gate, int gate_below,
                                                    unsigned long long* __restrict__ masks,
                                                    int* __restrict__ rowcnt) {
    ext_count_waves(tab, n_oct, njobs, ws, S, border, thr, thr_idx, gate, gate_below, masks, rowcnt);
}

__global__ __launch_bounds__(256) void k_ext_write(ExtTabArg tab, int n_oct, int nrows, int border,
                                                    const unsigned long long* __restrict__ masks,
                                                    const int* __restrict__ rowcnt, const int* __restrict__ rowoff,
                                                    int capc, int4* __restrict__ cands,
                                                    const int* __restrict__ gate, int gate_below) {
    if (gate && *gate >= gate_below) return;
    ext_write_waves(tab, n_oct, nrows, border, masks, rowcnt, rowoff, nullptr, capc, cands);
}

// ======================================================================
// A6: sub-pixel refinement (core.py:343-463) in FP64
// ======================================================================
__global__ __launch_bounds__(256) void k_refine(OctInfo oi, int q16_oct, const int4* __restrict__ cands,
                                                 const int* __restrict__ offs, int capc,
                                                 const float* __restrict__ thr, int thr_idx, float edge32,
                                                 uint8_t* __restrict__ valid, Kp* __restrict__ kslot,
                                                 const int* __restrict__ gate, int gate_below) {
    if (gate && *gate >= gate_below) return;
    const int n_slots = oi.n_oct * capc;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < n_slots; t += gridDim.x * 256) {
    int o = t / capc, i = t - o * capc;
    int cnt = offs[oi.blk0[o + 1]] - offs[oi.blk0[o]];
    uint8_t ok = 0;
    if (i < cnt && i < capc) {
        int4 cd = cands[t];
        Kp k;
        int si;
        if (refine_candidate(oi.G[o], o >= q16_oct, oi.H[o], oi.W[o], oi.S, oi.border, cd.x, cd.y, cd.z, (double)thr[thr_idx],
                             edge32, &k.ri, &k.ci, &k.resp, &si)) {
            k.o_si = (o << 8) | si;
            kslot[t] = k;
            ok = 1;
        }
    }
    valid[t] = ok;
    }
}

// compaction of refine slots (slot order = octave, raster), logical blocks of 256
__global__ __launch_bounds__(256) void k_flag_count(const uint8_t* __restrict__ flags, int n, int nblk,
                                                     int* __restrict__ counts, const int* __restrict__ gate,
                                                     int gate_below) {
    if (gate && *gate >= gate_below) return;
    __shared__ int wc[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
        int t = b * 256 + threadIdx.x;
        bool f = t < n && flags[t];
        unsigned long long m = __ballot(f);
        if (lane == 0) wc[wave] = __popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) counts[b] = (wc[0] + wc[1]) + (wc[2] + wc[3]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_flag_write(const uint8_t* __restrict__ flags, const Kp* __restrict__ src,
                                                     const int* __restrict__ offs, int n, int nblk,
                                                     Kp* __restrict__ dst, const int* __restrict__ gate,
                                                     int gate_below) {
    if (gate && *gate >= gate_below) return;
    __shared__ int wc[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
        int t = b * 256 + threadIdx.x;
        bool f = t < n && flags[t];
        unsigned long long m = __ballot(f);
        if (lane == 0) wc[wave] = __popcll(m);
        __syncthreads();
        if (f) {
            int pre = 0;
            for (int w = 0; w < wave; ++w) pre += wc[w];
            dst[offs[b] + pre + __popcll(m & lanemask_lt())] = src[t];
        }
        __syncthreads();
    }
}

// ======================================================================
// A7: keep the top max_keypoints by response (stable, descending;
// core.py:2078-2080).  Rank by counting over LDS tiles.
// ======================================================================
__global__ __launch_bounds__(256) void k_select(const Kp* __restrict__ k1, const int* __restrict__ kcount,
                                                 int maxkp, Kp* __restrict__ k2, int* __restrict__ n2) {
    __shared__ float sr[256];
    const int n = *kcount;
    if (blockIdx.x == 0 && threadIdx.x == 0) *n2 = n < maxkp ? n : maxkp;
    const int nblk = (n + 255) / 256;
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int i = b * 256 + threadIdx.x;
        if (n <= maxkp) {
            if (i < n) k2[i] = k1[i];
            continue;
        }
        float ri = i < n ? k1[i].resp : 0.0f;
        int rank = 0;
        for (int t0 = 0; t0 < n; t0 += 256) {
            __syncthreads();
            if (t0 + threadIdx.x < n) sr[threadIdx.x] = k1[t0 + threadIdx.x].resp;
            __syncthreads();
            int lim = min(256, n - t0);
            for (int j = 0; j < lim; ++j) {
                float rj = sr[j];
                rank += (rj > ri) || (rj == ri && t0 + j < i);
            }
        }
        if (i < n && rank < maxkp) k2[rank] = k1[i];
    }
}

// ======================================================================
// A8: orientation histogram (core.py:228-333).  One wave per keypoint;
// pixel p of the window goes to lane p % 64's private 36-bin column in LDS.
// ======================================================================
__global__ __launch_bounds__(256) void k_orient(OctInfo oi, const Kp* __restrict__ k2,
                                                 const int* __restrict__ n2p, KpTables tab,
                                                 float4* __restrict__ ang, int* __restrict__ nang, int maxkp) {
    __shared__ float part[4][36 * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n2 = *n2p;
    // angle counts beyond n2 read as zero for the scan
    for (int j = n2 + blockIdx.x * 256 + threadIdx.x; j < maxkp; j += gridDim.x * 256) nang[j] = 0;
    for (int i = blockIdx.x * 4 + wave; i < n2; i += gridDim.x * 4) {
        Kp k = k2[i];
        int o = k.o_si >> 8, si = k.o_si & 255;
        int H = oi.H[o], W = oi.W[o];
        float out[4];
        const float* Gl = oi.G[o] + (size_t)si * H * W;
        int n = o >= tab.q16_oct ? orient_keypoint<true>(Gl, H, W, k.ri, k.ci, tab.sig_ori[si], part[wave], lane, out)
                                 : orient_keypoint<false>(Gl, H, W, k.ri, k.ci, tab.sig_ori[si], part[wave], lane, out);
        if (lane == 0) {
            ang[i] = make_float4(out[0], out[1], out[2], out[3]);
            nang[i] = n;
        }
    }
}


// one entry per (keypoint, angle) in keypoint-major order (core.py:1594-1598)
__global__ __launch_bounds__(256) void k_expand(const Kp* __restrict__ k2, const int* __restrict__ n2p,
                                                 const float4* __restrict__ ang, const int* __restrict__ nang,
                                                 const int* __restrict__ aoff, double cs, Ent* __restrict__ E) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= *n2p) return;
    Kp k = k2[i];
    int o = k.o_si >> 8;
    double sf = (double)(1 << o);
    float x = (float)(((double)k.ci * sf) * cs);
    float y = (float)(((double)k.ri * sf) * cs);
    float4 a = ang[i];
    float av[4] = { a.x, a.y, a.z, a.w };
    int base = aoff[i];
    for (int j = 0; j < nang[i]; ++j) {
        Ent e;
        e.kp = i; e.angle = av[j]; e.x = x; e.y = y;
        E[base + j] = e;
    }
}

// canonical order: np.lexsort((ys, xs)) (core.py:2130), stable.  rank[i] =
// #{j : (x_j, y_j, j) < (x_i, y_i, i)}.  Block (bx, by) counts the 256
// entries i of i-block bx against the kRankJ entries j of chunk by (chunks
// past m are padded with +inf, which no finite coordinate exceeds) and adds
// its partial with an integer atomic (order-free); rank must be zero on entry
__global__ __launch_bounds__(256) void k_rank_xy(const Ent* __restrict__ E, const int* __restrict__ mp,
                                                  int* __restrict__ rank) {
    const int m = *mp;
    rank_xy_body(E, m, rank);
}

// ======================================================================
// A9: 4x4x8 descriptor (core.py:693-835), L2 / clip 0.2 / L2, RootSIFT
// (core.py:1870-1873), uint8 quantisation.  One wave per entry; pixel p of
// the patch goes to lane p % 64's private 128-bin column in LDS.
// ======================================================================
__global__ __launch_bounds__(kDescT) void k_descriptor(OctInfo oi, const Kp* __restrict__ k2,
                                                    const Ent* __restrict__ E, const int* __restrict__ mp,
                                                    const int* __restrict__ rank, int cap, KpTables tab,
                                                    double cs, int rootsift, float* __restrict__ out_kp,
                                                    float* __restrict__ out_f, uint8_t* __restrict__ out_u8,
                                                    int* __restrict__ count) {
    __shared__ unsigned long long hist[kDescHist];
    const int m = *mp;
    const int lane = threadIdx.x;
    if (blockIdx.x == 0 && lane == 0) *count = m;
    for (int e = blockIdx.x; e < m; e += gridDim.x) {
        const int rk = rank[e];
        if (rk >= cap) continue;
        Ent en = E[e];
        Kp k = k2[en.kp];
        const int o = k.o_si >> 8, si = k.o_si & 255;
        const int H = oi.H[o], W = oi.W[o];
        float v0, v1;
        describe_keypoint_block(oi.G[o] + (size_t)si * H * W, H, W, k.ri, k.ci, en.angle, tab.ph[si], rootsift, tab,
                                o >= tab.q16_oct, hist, &v0, &v1);
        if (lane < 64)
            store_descriptor(v0, v1, lane, out_f ? out_f + (size_t)rk * 128 : nullptr,
                             out_u8 ? out_u8 + (size_t)rk * 128 : nullptr);
        if (lane == 0 && out_kp) {
            double sf = (double)(1 << o);
            float* ko = out_kp + (size_t)rk * 6;
            ko[0] = en.x;
            ko[1] = en.y;
            ko[2] = (float)(((tab.base_size[si] * sf) * cs) * 2.0);
            ko[3] = en.angle;
            ko[4] = k.resp;
            ko[5] = (float)o;
        }
    }
}

// ======================================================================
// host side: workspace plan and launch sequence
// ======================================================================
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    }
};

struct Plan {
    int H, W, doubled, n_oct, S, capc, maxkp, cap_e, n_slots;
    int Ho[kMaxOct], Wo[kMaxOct];
    size_t oG[kMaxOct];
    size_t zero_ints;   // ints from o_hist on that the call's first kernel zeroes
    size_t o_unit, o_pf0, o_pf1, o_hist, o_thr, o_ints, o_masks, o_ecnt, o_ecnt2, o_eoff, o_cands, o_ccount,
        o_valid, o_kslot, o_ccnt, o_coff, o_k1, o_k2, o_ang, o_nang, o_aoff, o_E, o_rank, o_stmp;
    int n_ext_jobs, n_ext_rows, n_cmp_blk;
    ExtOct ext[kMaxOct];
    size_t total;
};

static void make_plan(const msift_handle* h, int H, int W, Plan* P) {
    const msift_params& p = h->p;
    memset(P, 0, sizeof(*P));
    P->H = H; P->W = W; P->S = h->S;
    P->doubled = p.double_image && ((long long)H * W < (long long)p.double_px_cap);
    P->n_oct = p.n_octaves;
    P->capc = p.max_cands_per_octave;
    P->maxkp = p.max_keypoints;
    P->cap_e = 4 * p.max_keypoints;
    int bh = P->doubled ? 2 * H : H, bw = P->doubled ? 2 * W : W;
    Carve c;
    for (int o = 0; o < P->n_oct; ++o) {
        P->Ho[o] = bh; P->Wo[o] = bw;
        size_t n = (size_t)bh * bw;
        P->oG[o] = c.take(n * sizeof(float) * (P->S + 3));
        bh = (bh + 1) / 2; bw = (bw + 1) / 2;
    }
    // scratch for the base image: aliases octave 0's levels 1..S+2, which
    // are written only after the base level exists
    size_t n_in = (size_t)H * W;
    size_t scratch = P->doubled ? n_in * (sizeof(double) * 2 + sizeof(float) * 4) : n_in * sizeof(float);
    const size_t lv0 = (size_t)P->Ho[0] * P->Wo[0] * sizeof(float);
    size_t sbase;
    if (scratch <= lv0 * (P->S + 2)) sbase = P->oG[0] + lv0;
    else sbase = c.take(scratch);
    if (P->doubled) {
        P->o_pf0 = sbase;
        P->o_pf1 = sbase + n_in * sizeof(double);
        P->o_unit = sbase + 2 * n_in * sizeof(double);   // upsampled float image (4 n_in)
    } else {
        P->o_unit = sbase;
        P->o_pf0 = P->o_pf1 = 0;
    }
    // extrema: wave jobs, (scale, row) slots, lane-mask units
    int nj = 0, nr = 0;
    long long nu = 0;
    for (int o = 0; o < P->n_oct; ++o)
        P->ext[o] = ext_geometry(P->oG[o], o >= h->tab.q16_oct, P->Ho[o], P->Wo[o], p.border, P->S, 0, &nj, &nr, &nu);
    P->n_ext_jobs = nj;
    P->n_ext_rows = nr;
    // zeroed together by the call's first kernel: histogram (256) + counters
    // (16), the row counts of both keypoint passes, the rank counts
    P->o_hist = c.take((256 + 16) * sizeof(int));
    P->o_ints = P->o_hist + 256 * sizeof(int);
    P->o_ecnt = c.take((size_t)(nr + 1) * sizeof(int));
    P->o_ecnt2 = c.take((size_t)(nr + 1) * sizeof(int));
    P->o_rank = c.take((size_t)P->cap_e * sizeof(int));
    P->zero_ints = (c.off - P->o_hist) / sizeof(int);
    P->o_thr = c.take(4 * sizeof(float));
    P->o_masks = c.take((size_t)(nu > 0 ? nu : 1) * 8);
    P->o_eoff = c.take((size_t)(nr + 1) * sizeof(int));
    P->n_slots = P->n_oct * P->capc;
    P->o_cands = c.take((size_t)P->n_slots * sizeof(int4));
    P->o_ccount = c.take(kMaxOct * sizeof(int));
    P->o_valid = c.take((size_t)P->n_slots);
    P->o_kslot = c.take((size_t)P->n_slots * sizeof(Kp));
    P->n_cmp_blk = (P->n_slots + 255) / 256;
    P->o_ccnt = c.take((size_t)(P->n_cmp_blk + 1) * sizeof(int));
    P->o_coff = c.take((size_t)(P->n_cmp_blk + 1) * sizeof(int));
    P->o_k1 = c.take((size_t)P->n_slots * sizeof(Kp));
    P->o_k2 = c.take((size_t)P->maxkp * sizeof(Kp));
    P->o_ang = c.take((size_t)P->maxkp * sizeof(float4));
    P->o_nang = c.take((size_t)(P->maxkp + 1) * sizeof(int));
    P->o_aoff = c.take((size_t)(P->maxkp + 1) * sizeof(int));
    P->o_E = c.take((size_t)P->cap_e * sizeof(Ent));
    P->o_stmp = c.take((size_t)scan_tmp_ints(std::max(P->n_ext_rows, P->maxkp + 1)) * sizeof(int));
    P->total = c.off;
}

static OctInfo oct_info(const Plan& P, char* ws, int border) {
    OctInfo oi;
    memset(&oi, 0, sizeof(oi));
    oi.n_oct = P.n_oct; oi.S = P.S; oi.border = border;
    for (int o = 0; o < P.n_oct; ++o) {
        oi.H[o] = P.Ho[o]; oi.W[o] = P.Wo[o];
        oi.G[o] = (float*)(ws + P.oG[o]);
        oi.blk0[o] = P.ext[o].row0;
    }
    oi.blk0[P.n_oct] = P.n_ext_rows;
    return oi;
}

int detect_workspace(msift_handle* h, int H, int W, size_t* bytes) {
    Plan P;
    make_plan(h, H, W, &P);
    *bytes = P.total;
    return MSIFT_OK;
}

int pyramid_view(msift_handle* h, int H, int W, void* ws, int o, int* Ho, int* Wo, float** g, float** d) {
    Plan P;
    make_plan(h, H, W, &P);
    if (o < 0 || o >= P.n_oct) { set_error("octave out of range"); return MSIFT_EINVAL; }
    *Ho = P.Ho[o]; *Wo = P.Wo[o];
    *g = (float*)((char*)ws + P.oG[o]);
    *d = nullptr;   // the DoG is formed by its readers, not stored
    return MSIFT_OK;
}

static inline int grid_cap(int n, int cap) { return n < 1 ? 1 : (n < cap ? n : cap); }

// extrema -> refine -> compaction, once (gate NULL) or as the ct/2 retry
static int keypoint_pass(const Plan& P, char* ws, const OctInfo& oi, msift_handle* h, int thr_idx,
                         const int* gate, int* total_a, int* total_b, hipStream_t s) {
    const int gb = h->p.retry_below;
    float* thr = (float*)(ws + P.o_thr);
    ExtTabArg tab;
    memcpy(tab.o, P.ext, sizeof(P.ext));
    // each pass counts into its own row counts, zeroed at the call's start
    int* rowcnt = (int*)(ws + (gate ? P.o_ecnt2 : P.o_ecnt));
    int* rowoff = (int*)(ws + P.o_eoff);
    if (P.n_ext_rows > 0) {
        hipLaunchKernelGGL(k_ext_count, dim3(grid_cap((P.n_ext_jobs + 3) / 4, 4096)), dim3(256), 0, s, tab, P.n_oct,
                           P.n_ext_jobs, (const char*)ws, P.S, h->p.border, thr, thr_idx, gate, gb,
                           (unsigned long long*)(ws + P.o_masks), rowcnt);
    }
    launch_scan(rowcnt, rowoff, P.n_ext_rows, nullptr, nullptr, gate, gb, s, (int*)(ws + P.o_stmp));
    if (P.n_ext_rows > 0) {
        hipLaunchKernelGGL(k_ext_write, dim3(grid_cap((P.n_ext_rows + 3) / 4, 4096)), dim3(256), 0, s, tab, P.n_oct,
                           P.n_ext_rows, h->p.border, (const unsigned long long*)(ws + P.o_masks),
                           (const int*)rowcnt, (const int*)rowoff, P.capc, (int4*)(ws + P.o_cands), gate, gb);
    }
    hipLaunchKernelGGL(k_refine, dim3(grid_cap((P.n_slots + 255) / 256, 1024)), dim3(256), 0, s, oi,
                       h->tab.q16_oct, (const int4*)(ws + P.o_cands), (const int*)(ws + P.o_eoff), P.capc, thr, thr_idx,
                       h->edge32, (uint8_t*)(ws + P.o_valid), (Kp*)(ws + P.o_kslot), gate, gb);
    hipLaunchKernelGGL(k_flag_count, dim3(grid_cap(P.n_cmp_blk, 1024)), dim3(256), 0, s,
                       (const uint8_t*)(ws + P.o_valid), P.n_slots, P.n_cmp_blk, (int*)(ws + P.o_ccnt), gate, gb);
    launch_scan((int*)(ws + P.o_ccnt), (int*)(ws + P.o_coff), P.n_cmp_blk, total_a, total_b, gate, gb, s);
    hipLaunchKernelGGL(k_flag_write, dim3(grid_cap(P.n_cmp_blk, 1024)), dim3(256), 0, s,
                       (const uint8_t*)(ws + P.o_valid), (const Kp*)(ws + P.o_kslot), (const int*)(ws + P.o_coff),
                       P.n_slots, P.n_cmp_blk, (Kp*)(ws + P.o_k1), gate, gb);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

void release_side_stream(msift_handle* h) {
    if (!h->side_init) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(h->side_dev);
    (void)hipStreamDestroy(h->side);
    for (int i = 0; i <= kMaxOct + 1; ++i) (void)hipEventDestroy(h->fork[i]);
    (void)hipSetDevice(cur);
    h->side_init = false;
}

int side_stream(msift_handle* h, hipStream_t s, hipStream_t* out) {
    int dev = 0;
    MSIFT_HIP(hipStreamGetDevice(s, &dev));
    if (h->side_init && h->side_dev == dev) { *out = h->side; return MSIFT_OK; }
    release_side_stream(h);
    int cur = 0;
    MSIFT_HIP(hipGetDevice(&cur));
    MSIFT_HIP(hipSetDevice(dev));
    hipError_t e = hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking);
    int ne = 0;
    for (; e == hipSuccess && ne <= kMaxOct + 1; ++ne) e = hipEventCreateWithFlags(&h->fork[ne], hipEventDisableTiming);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) {
        set_error(std::string("side stream: ") + hipGetErrorString(e));
        return MSIFT_EHIP;
    }
    h->side_dev = dev;
    h->side_init = true;
    *out = h->side;
    return MSIFT_OK;
}

int detect_describe(msift_handle* h, const uint8_t* img, int H, int W, int pitch, void* ws_, size_t ws_bytes,
                    int cap, float* kp, float* df, uint8_t* du, int* count, float* phase_ms, hipStream_t s) {
    if (H < 1 || W < 1 || pitch < W || !img || !ws_ || !count || cap < 0) {
        set_error("detect_describe: bad arguments");
        return MSIFT_EINVAL;
    }
    Plan P;
    make_plan(h, H, W, &P);
    if (ws_bytes < P.total) {
        set_error("detect_describe: workspace too small");
        return MSIFT_EWORKSPACE;
    }
    char* ws = (char*)ws_;
    const msift_params& p = h->p;
    const int S = P.S;
    OctInfo oi = oct_info(P, ws, p.border);
    int* ints = (int*)(ws + P.o_ints);
    float* thr = (float*)(ws + P.o_thr);
    bool prof = phase_ms != nullptr;
    if (prof && !h->ev_init) {
        for (int i = 0; i < 10; ++i) MSIFT_HIP(hipEventCreate(&h->ev[i]));
        h->ev_init = true;
    }
    if (prof) MSIFT_HIP(hipEventRecord(h->ev[0], s));

    int rc;
    // the second stream of the handle: the threshold (needed first by the
    // extrema) and the pyramid levels off the octave-to-octave chain run there
    hipStream_t side = nullptr;
    if (P.n_oct > 1 && (rc = side_stream(h, s, &side))) return rc;
    // A1: adaptive threshold, off the pyramid's dependent chain
    hipStream_t ts = s;
    if (side) {
        MSIFT_HIP(hipEventRecord(h->fork[kMaxOct + 1], s));
        MSIFT_HIP(hipStreamWaitEvent(side, h->fork[kMaxOct + 1], 0));
        ts = side;
    }
    // histogram, counters, both passes' row counts and the rank counts; their
    // readers on the launch stream follow the join of the second stream
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((P.zero_ints + 255) / 256)), dim3(256), 0, ts, (int*)(ws + P.o_hist),
                       (int)P.zero_ints);
    {
        hipLaunchKernelGGL(k_hist256, dim3(grid_cap(H, 128)), dim3(256), 0, ts, img, H, W, pitch,
                           (int*)(ws + P.o_hist));
        hipLaunchKernelGGL(k_thresh, dim3(1), dim3(64), 0, ts, (const int*)(ws + P.o_hist), (double)H * (double)W,
                           p.contrast_thresh, S, p.adaptive_thresh, thr);
    }
    // A3: base image
    float* G0 = oi.G[0];
    if (P.doubled) {
        double* pf0 = (double*)(ws + P.o_pf0);
        double* pf1 = (double*)(ws + P.o_pf1);
        float* up = (float*)(ws + P.o_unit);
        const int nt_c = ((W + kUpsColTW - 1) / kUpsColTW) * ((H + kUpsColTH - 1) / kUpsColTH);
        const int nt_r = ((W + kUpsRowTW - 1) / kUpsRowTW) * ((H + kUpsRowTH - 1) / kUpsRowTH);
        hipLaunchKernelGGL(k_ups_pf_cols, dim3(grid_cap(nt_c, 8192)), dim3(256), 0, s, img, pitch, H, W, h->ups, pf0,
                           p.fp16_pyramid);
        hipLaunchKernelGGL(k_ups_pf_rows, dim3(grid_cap(nt_r, 8192)), dim3(256), 0, s, pf0, H, W, h->ups, pf1);
        double zy = (2 * H > 1) ? (double)(H - 1) / (double)(2 * H - 1) : 0.0;
        double zx = (2 * W > 1) ? (double)(W - 1) / (double)(2 * W - 1) : 0.0;
        const int nt_i = ((2 * W + kUpsInTW - 1) / kUpsInTW) * ((2 * H + kUpsInTH - 1) / kUpsInTH);
        hipLaunchKernelGGL(k_ups_interp, dim3(grid_cap(nt_i, 8192)), dim3(256), 0, s, pf1, H, W, zy, zx, up,
                           p.fp16_pyramid);
        if ((rc = blur(up, G0, nullptr, 2 * H, 2 * W, h->sd0[1], h->taps_base[1], s, 1, 0, 0, h->blur_th)))
            return rc;
    } else {
        float* unit = (float*)(ws + P.o_unit);
        hipLaunchKernelGGL(k_u8_unit, dim3((W + 255) / 256, H), dim3(256), 0, s, img, pitch, H, W, unit,
                           p.fp16_pyramid);
        if ((rc = blur(unit, G0, nullptr, H, W, h->sd0[0], h->taps_base[0], s, 1, 0, 0, h->blur_th)))
            return rc;
    }
    MSIFT_LAUNCH_CHECK();
    // A3/A4: octaves.  The next octave's base is G[S][::2, ::2], so levels
    // S and S+1 of octave o are off the octave-to-octave chain: they run on
    // the side stream (after a fork event recorded once G[S] exists) while the
    // launch stream goes on to octave o+1; one join event brings them back
    // before the extrema.
    for (int o = 0; o < P.n_oct; ++o) {
        int Ho = P.Ho[o], Wo = P.Wo[o];
        size_t n = (size_t)Ho * Wo;
        const bool fork = o + 1 < P.n_oct;
        for (int lv = 0; lv < S + 2; ++lv) {
            float* dec = (lv + 1 == S && fork) ? oi.G[o + 1] : nullptr;
            float* gnext = oi.G[o] + n * (lv + 1);
            hipStream_t st = s;
            if (fork && lv >= S) {
                if (lv == S) {
                    MSIFT_HIP(hipEventRecord(h->fork[o], s));
                    MSIFT_HIP(hipStreamWaitEvent(side, h->fork[o], 0));
                }
                st = side;
            }
            if ((rc = blur(oi.G[o] + n * lv, gnext, dec, Ho, Wo, h->inc[lv],
                           h->taps_inc[lv], st, 1, 0, 0, h->blur_th)))
                return rc;
        }
    }
    MSIFT_LAUNCH_CHECK();
    if (side) {
        MSIFT_HIP(hipEventRecord(h->fork[kMaxOct], side));
        MSIFT_HIP(hipStreamWaitEvent(s, h->fork[kMaxOct], 0));
    }
    if (prof) MSIFT_HIP(hipEventRecord(h->ev[1], s));
    if (prof) MSIFT_HIP(hipEventRecord(h->ev[2], s));
    // A5/A6 (+ retry at ct/2 when fewer than retry_below keypoints, core.py:2064-2071)
    if ((rc = keypoint_pass(P, ws, oi, h, 0, nullptr, &ints[0], &ints[1], s))) return rc;
    if ((rc = keypoint_pass(P, ws, oi, h, 1, &ints[0], &ints[1], nullptr, s))) return rc;
    if (prof) MSIFT_HIP(hipEventRecord(h->ev[3], s));
    // A7: cap
    hipLaunchKernelGGL(k_select, dim3(grid_cap((P.n_slots + 255) / 256, 512)), dim3(256), 0, s, (const Kp*)(ws + P.o_k1),
                       (const int*)&ints[1], P.maxkp, (Kp*)(ws + P.o_k2), &ints[2]);
    if (prof) MSIFT_HIP(hipEventRecord(h->ev[4], s));
    // A8: orientation
    hipLaunchKernelGGL(k_orient, dim3(grid_cap((P.maxkp + 3) / 4, 1024)), dim3(256), 0, s, oi,
                       (const Kp*)(ws + P.o_k2), (const int*)&ints[2], h->tab, (float4*)(ws + P.o_ang),
                       (int*)(ws + P.o_nang), P.maxkp);
    MSIFT_LAUNCH_CHECK();
    if (prof) MSIFT_HIP(hipEventRecord(h->ev[5], s));
    // expansion + canonical order
    {
        launch_scan((int*)(ws + P.o_nang), (int*)(ws + P.o_aoff), P.maxkp, &ints[3], nullptr, nullptr, 0, s);
        double cs = P.doubled ? 0.5 : 1.0;
        hipLaunchKernelGGL(k_expand, dim3((P.maxkp + 255) / 256), dim3(256), 0, s, (const Kp*)(ws + P.o_k2),
                           (const int*)&ints[2], (const float4*)(ws + P.o_ang), (const int*)(ws + P.o_nang),
                           (const int*)(ws + P.o_aoff), cs, (Ent*)(ws + P.o_E));
        {
            hipLaunchKernelGGL(k_rank_xy, dim3(grid_cap((P.cap_e + 255) / 256, 8), grid_cap((P.cap_e + kRankJ - 1) / kRankJ, 128)),
                               dim3(256), 0, s, (const Ent*)(ws + P.o_E), (const int*)&ints[3],
                               (int*)(ws + P.o_rank));
        }
        // A9: descriptors written at their sorted position
        hipLaunchKernelGGL(k_descriptor, dim3(grid_cap(P.cap_e, 2048)), dim3(kDescT), 0, s, oi, (const Kp*)(ws + P.o_k2),
                           (const Ent*)(ws + P.o_E), (const int*)&ints[3], (const int*)(ws + P.o_rank), cap, h->tab,
                           cs, p.rootsift, kp, df, du, count);
    }
    MSIFT_LAUNCH_CHECK();
    if (prof) {
        MSIFT_HIP(hipEventRecord(h->ev[6], s));
        MSIFT_HIP(hipEventSynchronize(h->ev[6]));
        float t[7];
        for (int i = 0; i < 6; ++i) MSIFT_HIP(hipEventElapsedTime(&t[i], h->ev[i], h->ev[i + 1]));
        MSIFT_HIP(hipEventElapsedTime(&t[6], h->ev[0], h->ev[6]));
        // gauss_pyr, dog (formed by the extrema: 0), extrema (+refine), stack/cap,
        // orient, descriptor (+expand/sort), cv_kpt (0 on device), total
        phase_ms[0] = t[0]; phase_ms[1] = t[1]; phase_ms[2] = t[2]; phase_ms[3] = t[3];
        phase_ms[4] = t[4]; phase_ms[5] = t[5]; phase_ms[6] = 0.0f; phase_ms[7] = t[6];
    }
    return MSIFT_OK;
}

}  // namespace msift
