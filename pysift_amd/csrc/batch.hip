// Batched detect + describe over images of uneven size: every stage is ONE
// launch over all images, driven by per-image / per-octave tables in device
// memory (blocks find their image and tile by binary search), with no host
// synchronisation, so a caller can capture the whole call in a HIP graph.
// The per-tile and per-keypoint work is the single-image path's own device
// bodies (sift_dev.hpp), so every image's result is bit-identical to
// msift_detect_describe on that image alone.
#include "sift_dev.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#pragma clang fp contract(off)

namespace msift {

struct BImg {
    int H, W, doubled, oct0;
    size_t in_off;      // byte offset in the packed input (rows of W bytes)
    size_t pf0_off, pf1_off, base_off;  // spline scratch / base image (workspace offsets)
    double cs;
};

struct BOct {
    int img, o, H, W;
    int q16;            // levels read as binary16 (fp16 mode, octaves >= 2)
    size_t G_off;
};

struct BJob {           // one image's share of one pyramid step
    size_t in_off, g_off, dec_off;   // SIZE_MAX = not written
    int H, W, decW, tiles_x, tile0;
};

struct PixJob { int img; int px0; };          // per-image pixel ranges for the spline kernels

constexpr size_t kNone = ~(size_t)0;

}  // namespace msift

struct msift_batch {
    msift_handle* h;
    int n, n_oct, S, capc, maxkp;
    std::vector<msift::BImg> img;
    std::vector<msift::BOct> oct;
    // pyramid steps: each step = one launch = a contiguous range of jobs
    struct Step { int job0, njobs, tiles, R; double sigma; int th; const msift::Taps* taps; };
    std::vector<Step> steps;
    std::vector<msift::BJob> jobs;
    std::vector<msift::PixJob> ups_cols, ups_interp, unit;   // pixel-range tables
    std::vector<msift::ExtOct> ext;                          // extrema geometry per (image, octave)
    int ups_cols_px, ups_interp_px, unit_px;
    int n_ext_jobs, n_ext_rows;
    long long n_ext_units;
    size_t in_bytes;
    // workspace carve
    size_t o_hist, o_thr, o_cnt, o_masks, o_ecnt, o_eoff, o_cbase, o_cands, o_valid, o_kslot, o_fcnt, o_foff,
        o_k1a, o_k1b, o_kstart, o_k1, o_k2, o_ang, o_nang, o_aoff, o_E, o_rank, o_stmp, total;
    int n_cand_cap, n_cmp_blk;
    // device copies of the tables
    void* d_tables;
    msift::BImg* d_img;
    msift::BOct* d_oct;
    msift::BJob* d_jobs;
    msift::PixJob *d_ups_cols, *d_ups_interp, *d_unit;
    msift::ExtOct* d_ext;
    int blur_th;        // tile height (autotuned): 0 = default per radius
};

namespace msift {

// --------------------------------------------------------------- helpers --
template <typename T, typename F>
__device__ __forceinline__ int bsearch_last(const T* a, int n, int key, F get) {
    // largest i with get(a[i]) <= key
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (get(a[mid]) <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_bzero(int* p, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

// ------------------------------------------------------------- threshold --
__global__ __launch_bounds__(256) void k_bhist(const BImg* __restrict__ im, const uint8_t* __restrict__ in,
                                                int* __restrict__ hist) {
    __shared__ int sh[kHistCopies * kHistStride];
    const BImg I = im[blockIdx.y];
    hist256_body(in + I.in_off, I.H, I.W, (size_t)I.W, hist + blockIdx.y * 256, sh);
}

__global__ __launch_bounds__(64) void k_bthresh(const BImg* __restrict__ im, const int* __restrict__ hist, double base,
                                                 int S, int adaptive, float* __restrict__ thr) {
    const int l = threadIdx.x, i = blockIdx.x;
    const BImg I = im[i];
    const double total = (double)I.H * (double)I.W;
    double ct = base;
    if (adaptive) {
        double acc = 0.0;
        for (int j = 0; j < 4; ++j) {
            int c = hist[i * 256 + l + 64 * j];
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
        thr[2 * i] = (float)(ct / (double)S);
        thr[2 * i + 1] = (float)((ct * 0.5) / (double)S);
    }
}

// ---------------------------------------------------------------- spline --
// spline / unit kernels: blockIdx.y = image job, x strides over its tiles
__global__ __launch_bounds__(256) void k_bups_cols(const BImg* __restrict__ im, const PixJob* __restrict__ jobs,
                                                    const uint8_t* __restrict__ in, char* ws, UpsTaps h, int q16) {
    const BImg I = im[jobs[blockIdx.y].img];
    double* out = (double*)(ws + I.pf0_off);
    const int ntx = (I.W + kUpsColTW - 1) / kUpsColTW, nt = ntx * ((I.H + kUpsColTH - 1) / kUpsColTH);
    for (int t = blockIdx.x; t < nt; t += gridDim.x)
        ups_cols_tile(in + I.in_off, I.W, I.H, I.W, h, out, t % ntx, t / ntx, q16 != 0);
}

__global__ __launch_bounds__(256) void k_bups_rows(const BImg* __restrict__ im, const PixJob* __restrict__ jobs, char* ws,
                                                    UpsTaps h) {
    __shared__ double s[kUpsRowS];
    const BImg I = im[jobs[blockIdx.y].img];
    const double* src = (const double*)(ws + I.pf0_off);
    double* out = (double*)(ws + I.pf1_off);
    const int ntx = (I.W + kUpsRowTW - 1) / kUpsRowTW, nt = ntx * ((I.H + kUpsRowTH - 1) / kUpsRowTH);
    for (int t = blockIdx.x; t < nt; t += gridDim.x) ups_rows_tile(src, I.H, I.W, h, out, t % ntx, t / ntx, s);
}

__global__ __launch_bounds__(256) void k_bups_interp(const BImg* __restrict__ im, const PixJob* __restrict__ jobs, char* ws,
                                                      int q16) {
    __shared__ double s[kUpsInS];
    const BImg I = im[jobs[blockIdx.y].img];
    const double zy = (2 * I.H > 1) ? (double)(I.H - 1) / (double)(2 * I.H - 1) : 0.0;
    const double zx = (2 * I.W > 1) ? (double)(I.W - 1) / (double)(2 * I.W - 1) : 0.0;
    const double* cf = (const double*)(ws + I.pf1_off);
    float* out = (float*)(ws + I.base_off);
    const int ntx = (2 * I.W + kUpsInTW - 1) / kUpsInTW, nt = ntx * ((2 * I.H + kUpsInTH - 1) / kUpsInTH);
    for (int t = blockIdx.x; t < nt; t += gridDim.x)
        ups_interp_tile(cf, I.H, I.W, zy, zx, out, t % ntx, t / ntx, s, q16 != 0);
}

__global__ __launch_bounds__(256) void k_bunit(const BImg* __restrict__ im, const PixJob* __restrict__ jobs,
                                                const uint8_t* __restrict__ in, char* ws, int q16) {
    const BImg I = im[jobs[blockIdx.y].img];
    float* out = (float*)(ws + I.base_off);
    const int n = I.H * I.W;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const float v = (float)in[I.in_off + p] / 255.0f;
        out[p] = q16 ? f16_round(v) : v;
    }
}

// ------------------------------------------------------------------ blur --
template <int R, int TH>
__global__ __launch_bounds__(256) void k_bblur(const BJob* __restrict__ jobs, int njobs, char* ws, Taps taps) {
    __shared__ __attribute__((aligned(16))) float sIn[BlurSmem<R, TH>::nIn];
    __shared__ int s_j;
    const int tile = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (threadIdx.x < 64) {
        // the wave scans the job starts 64 at a time (no dependent loads)
        int last = -1;
        for (int b0 = 0; b0 < njobs; b0 += 64) {
            int k = b0 + (int)threadIdx.x;
            bool le = k < njobs && jobs[k].tile0 <= tile;
            unsigned long long m = __ballot(le);
            if (m) last = b0 + 63 - __clzll(m);
            if (m != ~0ull) break;
        }
        if (threadIdx.x == 0) s_j = last;
    }
    __syncthreads();
    const int j = s_j;
    const BJob J = jobs[j];
    const int t = tile - J.tile0;
    const int ty = t / J.tiles_x, tx = t - ty * J.tiles_x;
    blur_tile<R, TH>((const float*)(ws + J.in_off), J.g_off == kNone ? nullptr : (float*)(ws + J.g_off),
                     J.dec_off == kNone ? nullptr : (float*)(ws + J.dec_off), J.H, J.W, J.decW, taps, tx, ty, sIn);
}

__global__ __launch_bounds__(256) void k_bcopy_level(const BJob* __restrict__ jobs, int njobs, char* ws) {
    // sigma below 0.01: the level equals its input (core.py:559-560)
    const BJob J = jobs[blockIdx.y];
    const float* in = (const float*)(ws + J.in_off);
    size_t n = (size_t)J.H * J.W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float v = in[i];
        if (J.g_off != kNone) ((float*)(ws + J.g_off))[i] = v;
        int y = (int)(i / J.W), x = (int)(i - (size_t)y * J.W);
        if (J.dec_off != kNone && !(y & 1) && !(x & 1)) ((float*)(ws + J.dec_off))[(size_t)(y >> 1) * J.decW + (x >> 1)] = v;
    }
}

// --------------------------------------------------------------- extrema --
struct ExtTabPtr {
    const ExtOct* p;
    __device__ __forceinline__ ExtOct get(int i) const { return p[i]; }
    __device__ __forceinline__ int find_job(int v, int from, int n) const {
        return wave_find_last_le([&](int k) { return p[k].job0; }, v, from, n);
    }
    __device__ __forceinline__ int find_row(int v, int from, int n) const {
        return wave_find_last_le([&](int k) { return p[k].row0; }, v, from, n);
    }
};

// pass 1 (gate == NULL) or the ct/2 retry (only images with gate[i] < below)
__global__ __launch_bounds__(256) void k_bext_count(const ExtOct* __restrict__ ext, int n_io, int njobs, int S,
                                                     int border, const char* ws, const float* __restrict__ thr,
                                                     int thr_idx, const int* __restrict__ gate, int below,
                                                     unsigned long long* __restrict__ masks, int* __restrict__ rowcnt,
                                                     const int* __restrict__ skip) {
    if (skip && *skip) return;
    ext_count_waves(ExtTabPtr{ext}, n_io, njobs, ws, S, border, thr, thr_idx, gate, below, masks, rowcnt);
}

// truncated per-(image, octave) candidate counts and their exclusive scan
__global__ __launch_bounds__(1024) void k_bcand_base(const ExtOct* __restrict__ ext, int n_io, int nrows,
                                                      const int* __restrict__ rowoff, int capc, int* __restrict__ cbase,
                                                      const int* __restrict__ skip) {
    __shared__ int sm[1024];
    if (skip && *skip) return;
    int carry = 0;
    for (int base = 0; base < n_io; base += 1024) {
        int io = base + threadIdx.x;
        int c = 0;
        if (io < n_io) {
            int r0 = ext[io].row0;
            int r1 = io + 1 < n_io ? ext[io + 1].row0 : nrows;
            int t = rowoff[r1] - rowoff[r0];
            c = t < capc ? t : capc;
        }
        sm[threadIdx.x] = c;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int v = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
            __syncthreads();
            sm[threadIdx.x] += v;
            __syncthreads();
        }
        if (io < n_io) cbase[io] = carry + sm[threadIdx.x] - c;
        carry += sm[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) cbase[n_io] = carry;
}

__global__ __launch_bounds__(256) void k_bext_write(const ExtOct* __restrict__ ext, int n_io, int nrows, int border,
                                                     const unsigned long long* __restrict__ masks,
                                                     const int* __restrict__ rowcnt, const int* __restrict__ rowoff,
                                                     const int* __restrict__ cbase, int capc, int4* __restrict__ cands,
                                                     const int* __restrict__ skip) {
    if (skip && *skip) return;
    ext_write_waves(ExtTabPtr{ext}, n_io, nrows, border, masks, rowcnt, rowoff, cbase, capc, cands);
}

__global__ __launch_bounds__(256) void k_brefine(const BOct* __restrict__ oct, int n_io, int S, int border, char* ws,
                                                  const int4* __restrict__ cands, const int* __restrict__ cbase,
                                                  const float* __restrict__ thr, int thr_idx, float edge32,
                                                  uint8_t* __restrict__ valid, Kp* __restrict__ kslot,
                                                  int* __restrict__ img_count, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int n = cbase[n_io];
    for (int t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) {
        int4 cd = cands[t];
        const BOct O = oct[cd.w];
        Kp k;
        int si;
        uint8_t ok = 0;
        if (refine_candidate((const float*)(ws + O.G_off), O.q16, O.H, O.W, S, border, cd.x, cd.y, cd.z,
                             (double)thr[2 * O.img + thr_idx], edge32, &k.ri, &k.ci, &k.resp, &si)) {
            k.o_si = (O.img << 16) | (O.o << 8) | si;
            kslot[t] = k;
            ok = 1;
            atomicAdd(&img_count[O.img], 1);
        }
        valid[t] = ok;
    }
}

__global__ __launch_bounds__(256) void k_bflag_count(const uint8_t* __restrict__ flags, const int* __restrict__ np,
                                                      int nblk_cap, int* __restrict__ counts,
                                                      const int* __restrict__ skip) {
    __shared__ int wc[4];
    if (skip && *skip) return;
    const int n = *np;
    const int nblk = (n + 255) / 256;
    // blocks beyond this call's candidates count zero (the scan covers the capacity)
    for (int b = nblk + blockIdx.x * 256 + threadIdx.x; b <= nblk_cap; b += gridDim.x * 256) counts[b] = 0;
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

__global__ __launch_bounds__(256) void k_bflag_write(const uint8_t* __restrict__ flags, const Kp* __restrict__ src,
                                                      const int* __restrict__ np, const int* __restrict__ offs,
                                                      Kp* __restrict__ dst, const int* __restrict__ skip) {
    __shared__ int wc[4];
    if (skip && *skip) return;
    const int n = *np;
    const int nblk = (n + 255) / 256;
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

// skip = 1 when no image's pass-1 count is below the retry gate (the whole
// retry pass is then a no-op and every kernel of it returns at once)
__global__ __launch_bounds__(256) void k_bneed(int n, const int* __restrict__ cnt1, int below, int* __restrict__ skip) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256)
        if (cnt1[i] < below) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) *skip = any ? 0 : 1;
}

// per-image keypoint list: pass 1, or the retry for images below the gate
__global__ __launch_bounds__(1024) void k_bmerge_counts(int n, const int* __restrict__ cnt1, const int* __restrict__ cnt2,
                                                         int below, int* __restrict__ start1, int* __restrict__ start2,
                                                         int* __restrict__ fstart, int* __restrict__ fcount) {
    // sequential prefix over at most a few thousand images (one thread)
    if (threadIdx.x != 0) return;
    int a = 0, b = 0, f = 0;
    for (int i = 0; i < n; ++i) {
        start1[i] = a; a += cnt1[i];
        int c2 = cnt1[i] < below ? cnt2[i] : 0;
        start2[i] = b; b += c2;
        int c = cnt1[i] < below ? cnt2[i] : cnt1[i];
        fstart[i] = f; fcount[i] = c; f += c;
    }
    fstart[n] = f;
}

__global__ __launch_bounds__(256) void k_bmerge(int n, const int* __restrict__ cnt1, int below,
                                                 const int* __restrict__ start1, const int* __restrict__ start2,
                                                 const int* __restrict__ fstart, const Kp* __restrict__ k1a,
                                                 const Kp* __restrict__ k1b, Kp* __restrict__ k1) {
    const int i = blockIdx.y;
    const int c = fstart[i + 1] - fstart[i];
    const bool retry = cnt1[i] < below;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < c; j += gridDim.x * 256)
        k1[fstart[i] + j] = retry ? k1b[start2[i] + j] : k1a[start1[i] + j];
}

// top max_keypoints by response per image (stable, descending)
__global__ __launch_bounds__(256) void k_bselect(const int* __restrict__ fstart, const Kp* __restrict__ k1, int maxkp,
                                                  Kp* __restrict__ k2, int* __restrict__ n2) {
    __shared__ float sr[256];
    const int i = blockIdx.y;
    const int s0 = fstart[i], n = fstart[i + 1] - s0;
    if (blockIdx.x == 0 && threadIdx.x == 0) n2[i] = n < maxkp ? n : maxkp;
    const int nblk = (n + 255) / 256;
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int l = b * 256 + threadIdx.x;
        if (n <= maxkp) {
            if (l < n) k2[(size_t)i * maxkp + l] = k1[s0 + l];
            continue;
        }
        float rl = l < n ? k1[s0 + l].resp : 0.0f;
        int rank = 0;
        for (int t0 = 0; t0 < n; t0 += 256) {
            __syncthreads();
            if (t0 + threadIdx.x < n) sr[threadIdx.x] = k1[s0 + t0 + threadIdx.x].resp;
            __syncthreads();
            int lim = min(256, n - t0);
            for (int j = 0; j < lim; ++j) {
                float rj = sr[j];
                rank += (rj > rl) || (rj == rl && t0 + j < l);
            }
        }
        if (l < n && rank < maxkp) k2[(size_t)i * maxkp + rank] = k1[s0 + l];
    }
}

__global__ __launch_bounds__(256) void k_borient(const BImg* __restrict__ im, const BOct* __restrict__ oct, int n,
                                                  int maxkp, char* ws, const Kp* __restrict__ k2,
                                                  const int* __restrict__ n2, KpTables tab, float4* __restrict__ ang,
                                                  int* __restrict__ nang) {
    __shared__ float part[4][36 * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nslots = n * maxkp;
    for (int slot = blockIdx.x * 4 + wave; slot < nslots; slot += gridDim.x * 4) {
        const int i = slot / maxkp, l = slot - i * maxkp;
        if (l >= n2[i]) {
            if (lane == 0) nang[slot] = 0;
            continue;
        }
        Kp k = k2[slot];
        const int o = (k.o_si >> 8) & 255, si = k.o_si & 255;
        const BOct O = oct[im[i].oct0 + o];
        float out[4];
        const float* Gl = (const float*)(ws + O.G_off) + (size_t)si * O.H * O.W;
        int c = o >= tab.q16_oct ? orient_keypoint<true>(Gl, O.H, O.W, k.ri, k.ci, tab.sig_ori[si], part[wave], lane, out)
                                 : orient_keypoint<false>(Gl, O.H, O.W, k.ri, k.ci, tab.sig_ori[si], part[wave], lane, out);
        if (lane == 0) {
            ang[slot] = make_float4(out[0], out[1], out[2], out[3]);
            nang[slot] = c;
        }
    }
}

__global__ __launch_bounds__(256) void k_bexpand(const BImg* __restrict__ im, int n, int maxkp,
                                                  const Kp* __restrict__ k2, const int* __restrict__ n2,
                                                  const float4* __restrict__ ang, const int* __restrict__ nang,
                                                  const int* __restrict__ aoff, Ent* __restrict__ E) {
    const int nslots = n * maxkp;
    for (int slot = blockIdx.x * 256 + threadIdx.x; slot < nslots; slot += gridDim.x * 256) {
        const int i = slot / maxkp;
        if (slot - i * maxkp >= n2[i]) continue;
        Kp k = k2[slot];
        const int o = (k.o_si >> 8) & 255;
        const double sf = (double)(1 << o), cs = im[i].cs;
        float x = (float)(((double)k.ci * sf) * cs);
        float y = (float)(((double)k.ri * sf) * cs);
        float4 a = ang[slot];
        float av[4] = { a.x, a.y, a.z, a.w };
        for (int q = 0; q < nang[slot]; ++q) {
            Ent e;
            e.kp = slot; e.angle = av[q]; e.x = x; e.y = y;
            E[aoff[slot] + q] = e;
        }
    }
}

// per-image stable (x, y) rank (core.py:2130)
// stable (x, y, index) rank within image z (blockIdx.z), added into rank[]
// (zero on entry) at the image's entry offset
__global__ __launch_bounds__(256) void k_brank_xy(const Ent* __restrict__ E, const int* __restrict__ aoff, int maxkp,
                                                   int* __restrict__ rank) {
    const int i = blockIdx.z;
    const int e0 = aoff[(size_t)i * maxkp], m = aoff[(size_t)(i + 1) * maxkp] - e0;
    rank_xy_body(E + e0, m, rank + e0);
}

__global__ __launch_bounds__(kDescT) void k_bdescriptor(const BImg* __restrict__ im, const BOct* __restrict__ oct,
                                                     int n, int maxkp, char* ws, const Kp* __restrict__ k2,
                                                     const Ent* __restrict__ E, const int* __restrict__ aoff,
                                                     const int* __restrict__ rank, int cap, KpTables tab, int rootsift,
                                                     float* __restrict__ out_kp, float* __restrict__ out_f,
                                                     uint8_t* __restrict__ out_u8, int* __restrict__ counts,
                                                     int* __restrict__ offsets) {
    __shared__ unsigned long long hist[kDescHist];
    const int lane = threadIdx.x;
    const int m = aoff[(size_t)n * maxkp];
    if (blockIdx.x == 0)
        for (int i = lane; i <= n; i += kDescT) {
            if (offsets) offsets[i] = aoff[(size_t)i * maxkp];
            if (counts && i < n) counts[i] = aoff[(size_t)(i + 1) * maxkp] - aoff[(size_t)i * maxkp];
        }
    for (int e = blockIdx.x; e < m; e += gridDim.x) {
        Ent en = E[e];
        const int i = en.kp / maxkp;
        const int row = aoff[(size_t)i * maxkp] + rank[e];
        if (row >= cap) continue;
        Kp k = k2[en.kp];
        const int o = (k.o_si >> 8) & 255, si = k.o_si & 255;
        const BOct O = oct[im[i].oct0 + o];
        float v0, v1;
        describe_keypoint_block((const float*)(ws + O.G_off) + (size_t)si * O.H * O.W, O.H, O.W, k.ri, k.ci,
                                en.angle, tab.ph[si], rootsift, tab, o >= tab.q16_oct, hist, &v0, &v1);
        if (lane < 64)
            store_descriptor(v0, v1, lane, out_f ? out_f + (size_t)row * 128 : nullptr,
                             out_u8 ? out_u8 + (size_t)row * 128 : nullptr);
        if (lane == 0 && out_kp) {
            double sf = (double)(1 << o);
            float* ko = out_kp + (size_t)row * 6;
            ko[0] = en.x;
            ko[1] = en.y;
            ko[2] = (float)(((tab.base_size[si] * sf) * im[i].cs) * 2.0);
            ko[3] = en.angle;
            ko[4] = k.resp;
            ko[5] = (float)o;
        }
    }
}

// ---------------------------------------------------------------- host ----
static inline int gcap(long long n, int cap) { return n < 1 ? 1 : (n < cap ? (int)n : cap); }

template <int R>
static void bblur_launch(const msift_batch* B, const msift_batch::Step& st, char* ws, const Taps& t, hipStream_t s) {
    const BJob* jobs = B->d_jobs + st.job0;
    if (st.th == 16 || R > 24)
        hipLaunchKernelGGL((k_bblur<R, 16>), dim3(st.tiles), dim3(256), 0, s, jobs, st.njobs, ws, t);
    else if (st.th == 32 || R > 12)
        hipLaunchKernelGGL((k_bblur<R, (R <= 24 ? 32 : 16)>), dim3(st.tiles), dim3(256), 0, s, jobs, st.njobs, ws, t);
    else
        hipLaunchKernelGGL((k_bblur<R, (R <= 12 ? 64 : 16)>), dim3(st.tiles), dim3(256), 0, s, jobs, st.njobs, ws, t);
}

static int bblur(const msift_batch* B, const msift_batch::Step& st, char* ws, hipStream_t s) {
    const Taps& t = *st.taps;
    if (st.sigma < 0.01) {
        hipLaunchKernelGGL(k_bcopy_level, dim3(256, st.njobs), dim3(256), 0, s, B->d_jobs + st.job0, st.njobs, ws);
        return MSIFT_OK;
    }
    switch (t.r) {
#define MSIFT_BB(RR) case RR: bblur_launch<RR>(B, st, ws, t, s); break;
        MSIFT_BB(0) MSIFT_BB(1) MSIFT_BB(2) MSIFT_BB(3) MSIFT_BB(4) MSIFT_BB(5) MSIFT_BB(6) MSIFT_BB(7)
        MSIFT_BB(8) MSIFT_BB(9) MSIFT_BB(10) MSIFT_BB(11) MSIFT_BB(12) MSIFT_BB(13) MSIFT_BB(14) MSIFT_BB(15)
        MSIFT_BB(16) MSIFT_BB(17) MSIFT_BB(18) MSIFT_BB(19) MSIFT_BB(20) MSIFT_BB(21) MSIFT_BB(22) MSIFT_BB(23)
        MSIFT_BB(24) MSIFT_BB(25) MSIFT_BB(26) MSIFT_BB(27) MSIFT_BB(28) MSIFT_BB(29) MSIFT_BB(30) MSIFT_BB(31)
        MSIFT_BB(32)
#undef MSIFT_BB
    default:
        set_error("blur radius above 32");
        return MSIFT_EINVAL;
    }
    return MSIFT_OK;
}

static int tile_h(int R, int th) {
    if (R > 24) return 16;
    if (R > 12) return th == 16 ? 16 : 32;
    return th ? th : 64;
}

// fills the blur job list of every step for tile height choice `th`
static void plan_steps(msift_batch* B) {
    msift_handle* h = B->h;
    B->steps.clear();
    B->jobs.clear();
    const int S = B->S, n = B->n;
    auto add_step = [&](std::vector<BJob> js, const Taps* taps, double sigma) {
        msift_batch::Step st;
        st.job0 = (int)B->jobs.size();
        st.njobs = (int)js.size();
        st.taps = taps;
        st.sigma = sigma;
        st.R = taps->r;
        st.th = tile_h(taps->r, B->blur_th);
        int tiles = 0;
        for (auto& j : js) {
            j.tile0 = tiles;
            j.tiles_x = (j.W + 63) / 64;
            tiles += j.tiles_x * ((j.H + st.th - 1) / st.th);
        }
        st.tiles = tiles;
        for (auto& j : js) B->jobs.push_back(j);
        B->steps.push_back(st);
    };
    // base level: doubled and single-size images use different sigmas
    for (int d = 1; d >= 0; --d) {
        std::vector<BJob> js;
        for (int i = 0; i < n; ++i) {
            const BImg& I = B->img[i];
            if (I.doubled != d) continue;
            const BOct& O = B->oct[I.oct0];
            BJob j;
            j.in_off = I.base_off; j.g_off = O.G_off; j.dec_off = kNone;
            j.H = O.H; j.W = O.W; j.decW = (O.W + 1) / 2;
            js.push_back(j);
        }
        if (!js.empty()) add_step(js, &h->taps_base[d], h->sd0[d]);
    }
    for (int o = 0; o < B->n_oct; ++o)
        for (int lv = 0; lv < S + 2; ++lv) {
            std::vector<BJob> js;
            for (int i = 0; i < n; ++i) {
                const BOct& O = B->oct[B->img[i].oct0 + o];
                size_t HW4 = (size_t)O.H * O.W * 4;
                BJob j;
                j.in_off = O.G_off + HW4 * lv;
                j.g_off = O.G_off + HW4 * (lv + 1);
                j.dec_off = (lv + 1 == S && o + 1 < B->n_oct) ? B->oct[B->img[i].oct0 + o + 1].G_off : kNone;
                j.H = O.H; j.W = O.W; j.decW = (O.W + 1) / 2;
                js.push_back(j);
            }
            add_step(js, &h->taps_inc[lv], h->inc[lv]);
        }
}

static int upload_tables(msift_batch* B) {
    auto al = [](size_t x) { return ((x + 255) & ~(size_t)255) + 256; };
    const void* src[7] = { B->img.data(), B->oct.data(), B->jobs.data(), B->ups_cols.data(), B->ups_interp.data(),
                           B->unit.data(), B->ext.data() };
    size_t sz[7] = { B->img.size() * sizeof(BImg), B->oct.size() * sizeof(BOct), B->jobs.size() * sizeof(BJob),
                     B->ups_cols.size() * sizeof(PixJob), B->ups_interp.size() * sizeof(PixJob),
                     B->unit.size() * sizeof(PixJob), B->ext.size() * sizeof(ExtOct) };
    size_t off[7], total = 0;
    for (int k = 0; k < 7; ++k) { off[k] = total; total += al(sz[k]); }
    std::vector<char> host(total, 0);
    for (int k = 0; k < 7; ++k) if (sz[k]) memcpy(host.data() + off[k], src[k], sz[k]);
    if (B->d_tables) { (void)hipFree(B->d_tables); B->d_tables = nullptr; }
    MSIFT_HIP(hipMalloc(&B->d_tables, total));
    MSIFT_HIP(hipMemcpy(B->d_tables, host.data(), total, hipMemcpyHostToDevice));
    char* base = (char*)B->d_tables;
    B->d_img = (BImg*)(base + off[0]);
    B->d_oct = (BOct*)(base + off[1]);
    B->d_jobs = (BJob*)(base + off[2]);
    B->d_ups_cols = (PixJob*)(base + off[3]);
    B->d_ups_interp = (PixJob*)(base + off[4]);
    B->d_unit = (PixJob*)(base + off[5]);
    B->d_ext = (ExtOct*)(base + off[6]);
    return MSIFT_OK;
}

int batch_create(msift_handle* h, int n, const int* Hs, const int* Ws, msift_batch** out) {
    if (n < 1 || n > 65535 || !Hs || !Ws || !out) { set_error("batch_create: bad arguments"); return MSIFT_EINVAL; }
    const msift_params& p = h->p;
    msift_batch* B = new (std::nothrow) msift_batch();
    if (!B) { set_error("out of host memory"); return MSIFT_EINVAL; }
    B->h = h; B->n = n; B->n_oct = p.n_octaves; B->S = h->S; B->capc = p.max_cands_per_octave;
    B->maxkp = p.max_keypoints; B->d_tables = nullptr; B->blur_th = 0;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    // packed input layout
    size_t in_off = 0;
    int nj = 0, nr = 0;
    long long nu = 0;
    for (int i = 0; i < n; ++i) {
        if (Hs[i] < 1 || Ws[i] < 1) { set_error("batch_create: image sizes must be positive"); delete B; return MSIFT_EINVAL; }
        BImg I;
        memset(&I, 0, sizeof(I));
        I.H = Hs[i]; I.W = Ws[i];
        I.doubled = p.double_image && ((long long)Hs[i] * Ws[i] < (long long)p.double_px_cap);
        I.cs = I.doubled ? 0.5 : 1.0;
        I.in_off = in_off;
        in_off += (size_t)Hs[i] * Ws[i];
        I.oct0 = (int)B->oct.size();
        int bh = I.doubled ? 2 * I.H : I.H, bw = I.doubled ? 2 * I.W : I.W;
        for (int o = 0; o < B->n_oct; ++o) {
            BOct O;
            memset(&O, 0, sizeof(O));
            O.img = i; O.o = o; O.H = bh; O.W = bw;
            O.q16 = o >= h->tab.q16_oct;
            size_t hw = (size_t)bh * bw;
            O.G_off = take(hw * 4 * (B->S + 3));
            B->ext.push_back(ext_geometry(O.G_off, O.q16, bh, bw, p.border, B->S, i, &nj, &nr, &nu));
            B->oct.push_back(O);
            bh = (bh + 1) / 2; bw = (bw + 1) / 2;
        }
        // spline scratch and base image alias octave 0's levels 1..S+2 (written
        // only after the base level exists) when they fit
        size_t n_in = (size_t)I.H * I.W;
        size_t need = I.doubled ? n_in * (16 + 16) : n_in * 4;
        const size_t lv0 = (size_t)B->oct[I.oct0].H * B->oct[I.oct0].W * 4;
        size_t sb = need <= lv0 * (B->S + 2) ? B->oct[I.oct0].G_off + lv0 : take(need);
        if (I.doubled) { I.pf0_off = sb; I.pf1_off = sb + n_in * 8; I.base_off = sb + n_in * 16; }
        else { I.base_off = sb; I.pf0_off = I.pf1_off = 0; }
        B->img.push_back(I);
    }
    B->in_bytes = in_off;
    B->n_ext_jobs = nj;
    B->n_ext_rows = nr;
    B->n_ext_units = nu;
    // pixel-range tables for the spline / unit kernels
    int pc = 0, pi = 0, pu = 0;
    for (int i = 0; i < n; ++i) {
        const BImg& I = B->img[i];
        if (I.doubled) {
            B->ups_cols.push_back({ i, pc }); pc += I.H * I.W;
            B->ups_interp.push_back({ i, pi }); pi += 4 * I.H * I.W;
        } else {
            B->unit.push_back({ i, pu }); pu += I.H * I.W;
        }
    }
    B->ups_cols_px = pc; B->ups_interp_px = pi; B->unit_px = pu;
    // per-call state
    const int n_io = n * B->n_oct;
    B->o_hist = take((size_t)n * 256 * 4 + (size_t)n * 8 * 4);   // hist + 8 counters per image
    B->o_cnt = B->o_hist + (size_t)n * 256 * 4;
    B->o_thr = take((size_t)n * 2 * 4);
    B->o_masks = take((size_t)(nu > 0 ? nu : 1) * 8);
    B->o_ecnt = take((size_t)(nr + 1) * 4);
    B->o_eoff = take((size_t)(nr + 1) * 4);
    B->o_cbase = take((size_t)(n_io + 1) * 4);
    B->n_cand_cap = n_io * B->capc;
    B->o_cands = take((size_t)B->n_cand_cap * 16);
    B->o_valid = take((size_t)B->n_cand_cap);
    B->o_kslot = take((size_t)B->n_cand_cap * sizeof(Kp));
    B->n_cmp_blk = (B->n_cand_cap + 255) / 256;
    B->o_fcnt = take((size_t)(B->n_cmp_blk + 1) * 4);
    B->o_foff = take((size_t)(B->n_cmp_blk + 1) * 4);
    B->o_k1a = take((size_t)B->n_cand_cap * sizeof(Kp));
    B->o_k1b = take((size_t)B->n_cand_cap * sizeof(Kp));
    B->o_kstart = take((size_t)(5 * n + 2) * 4);
    B->o_k1 = take((size_t)B->n_cand_cap * sizeof(Kp));
    B->o_k2 = take((size_t)n * B->maxkp * sizeof(Kp));
    B->o_ang = take((size_t)n * B->maxkp * 16);
    B->o_nang = take((size_t)n * B->maxkp * 4);
    B->o_aoff = take(((size_t)n * B->maxkp + 1) * 4);
    B->o_E = take((size_t)n * 4 * B->maxkp * sizeof(Ent));
    B->o_rank = take((size_t)n * 4 * B->maxkp * 4);
    {
        long long big = std::max<long long>((long long)nr, (long long)n * B->maxkp);
        big = std::max<long long>(big, (long long)B->n_cmp_blk);
        B->o_stmp = take((size_t)scan_tmp_ints((int)big) * 4);
    }
    B->total = off;
    plan_steps(B);
    int rc = upload_tables(B);
    if (rc) { delete B; return rc; }
    *out = B;
    return MSIFT_OK;
}

int batch_set_blur_tile(msift_batch* B, int th) {
    if (th != 0 && th != 16 && th != 32 && th != 64) { set_error("blur tile height must be 0, 16, 32 or 64"); return MSIFT_EINVAL; }
    B->blur_th = th;
    plan_steps(B);
    return upload_tables(B);
}

int batch_destroy(msift_batch* B) {
    if (!B) return MSIFT_OK;
    if (B->d_tables) (void)hipFree(B->d_tables);
    delete B;
    return MSIFT_OK;
}

static int bkeypoint_pass(msift_batch* B, char* ws, const uint8_t* in, int thr_idx, const int* gate, int* img_count,
                          Kp* dst, const int* skip, hipStream_t s) {
    (void)in;
    msift_handle* h = B->h;
    const msift_params& p = h->p;
    const int n_io = B->n * B->n_oct;
    int* ecnt = (int*)(ws + B->o_ecnt);
    int* eoff = (int*)(ws + B->o_eoff);
    int* cbase = (int*)(ws + B->o_cbase);
    float* thr = (float*)(ws + B->o_thr);
    int* stmp = (int*)(ws + B->o_stmp);
    const int nr = B->n_ext_rows;
    if (nr > 0) {
        hipLaunchKernelGGL(k_bzero, dim3(gcap((nr + 255) / 256, 1024)), dim3(256), 0, s, ecnt, nr);
        hipLaunchKernelGGL(k_bext_count, dim3(gcap((B->n_ext_jobs + 3) / 4, 16384)), dim3(256), 0, s, B->d_ext, n_io,
                           B->n_ext_jobs, B->S, p.border, (const char*)ws, thr, thr_idx, gate, p.retry_below,
                           (unsigned long long*)(ws + B->o_masks), ecnt, skip);
    }
    launch_scan(ecnt, eoff, nr, nullptr, nullptr, skip, 1, s, stmp);
    hipLaunchKernelGGL(k_bcand_base, dim3(1), dim3(1024), 0, s, B->d_ext, n_io, nr, eoff, B->capc, cbase, skip);
    if (nr > 0)
        hipLaunchKernelGGL(k_bext_write, dim3(gcap((nr + 3) / 4, 16384)), dim3(256), 0, s, B->d_ext, n_io, nr, p.border,
                           (const unsigned long long*)(ws + B->o_masks), (const int*)ecnt, (const int*)eoff, cbase,
                           B->capc, (int4*)(ws + B->o_cands), skip);
    hipLaunchKernelGGL(k_brefine, dim3(gcap(B->n_cmp_blk, 2048)), dim3(256), 0, s, B->d_oct, n_io, B->S, p.border, ws,
                       (const int4*)(ws + B->o_cands), cbase, thr, thr_idx, h->edge32, (uint8_t*)(ws + B->o_valid),
                       (Kp*)(ws + B->o_kslot), img_count, skip);
    hipLaunchKernelGGL(k_bflag_count, dim3(gcap(B->n_cmp_blk, 2048)), dim3(256), 0, s, (const uint8_t*)(ws + B->o_valid),
                       (const int*)(cbase + n_io), B->n_cmp_blk, (int*)(ws + B->o_fcnt), skip);
    launch_scan((int*)(ws + B->o_fcnt), (int*)(ws + B->o_foff), B->n_cmp_blk, nullptr, nullptr, skip, 1, s, stmp);
    hipLaunchKernelGGL(k_bflag_write, dim3(gcap(B->n_cmp_blk, 2048)), dim3(256), 0, s, (const uint8_t*)(ws + B->o_valid),
                       (const Kp*)(ws + B->o_kslot), (const int*)(cbase + n_io), (const int*)(ws + B->o_foff), dst,
                       skip);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

int batch_run(msift_batch* B, const uint8_t* in, void* ws_, size_t ws_bytes, int cap, float* kp, float* df,
              uint8_t* du, int* counts, int* offsets, hipStream_t s) {
    if (ws_bytes < B->total) { set_error("detect_describe_batched: workspace too small"); return MSIFT_EWORKSPACE; }
    if (!in || !ws_) { set_error("detect_describe_batched: null buffer"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    msift_handle* h = B->h;
    const msift_params& p = h->p;
    const int n = B->n, S = B->S;
    int* hist = (int*)(ws + B->o_hist);
    int* cnt = (int*)(ws + B->o_cnt);           // [n] pass-1 counts, [n] retry counts
    float* thr = (float*)(ws + B->o_thr);
    // A1: per-image histograms and thresholds
    hipLaunchKernelGGL(k_bzero, dim3(gcap((n * 264 + 255) / 256, 1024)), dim3(256), 0, s, hist, n * 256 + n * 8);
    hipLaunchKernelGGL(k_bhist, dim3(32, n), dim3(256), 0, s, B->d_img, in, hist);
    hipLaunchKernelGGL(k_bthresh, dim3(n), dim3(64), 0, s, B->d_img, (const int*)hist, p.contrast_thresh, S,
                       p.adaptive_thresh, thr);
    // A3: base images
    if (B->ups_cols_px > 0) {
        dim3 g(256, (unsigned)B->ups_cols.size());
        hipLaunchKernelGGL(k_bups_cols, g, dim3(256), 0, s, B->d_img, B->d_ups_cols, in, ws, h->ups, p.fp16_pyramid);
        hipLaunchKernelGGL(k_bups_rows, g, dim3(256), 0, s, B->d_img, B->d_ups_cols, ws, h->ups);
        hipLaunchKernelGGL(k_bups_interp, dim3(512, (unsigned)B->ups_cols.size()), dim3(256), 0, s, B->d_img,
                           B->d_ups_cols, ws, p.fp16_pyramid);
    }
    if (B->unit_px > 0)
        hipLaunchKernelGGL(k_bunit, dim3(256, (unsigned)B->unit.size()), dim3(256), 0, s, B->d_img, B->d_unit, in, ws,
                           p.fp16_pyramid);
    MSIFT_LAUNCH_CHECK();
    // A2-A4: every pyramid step is one launch over all images
    for (const auto& st : B->steps) {
        int rc = bblur(B, st, ws, s);
        if (rc) return rc;
    }
    MSIFT_LAUNCH_CHECK();
    // A5-A7
    int* k_start = (int*)(ws + B->o_kstart);    // start1[n], start2[n], fstart[n+1], fcount[n]
    int* start1 = k_start, *start2 = k_start + n, *fstart = k_start + 2 * n, *fcount = k_start + 3 * n + 1;
    int* skip2 = cnt + 3 * n;
    int rc = bkeypoint_pass(B, ws, in, 0, nullptr, cnt, (Kp*)(ws + B->o_k1a), nullptr, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bneed, dim3(1), dim3(256), 0, s, n, (const int*)cnt, p.retry_below, skip2);
    rc = bkeypoint_pass(B, ws, in, 1, cnt, cnt + n, (Kp*)(ws + B->o_k1b), skip2, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bmerge_counts, dim3(1), dim3(64), 0, s, n, (const int*)cnt, (const int*)(cnt + n),
                       p.retry_below, start1, start2, fstart, fcount);
    hipLaunchKernelGGL(k_bmerge, dim3(64, n), dim3(256), 0, s, n, (const int*)cnt, p.retry_below, (const int*)start1,
                       (const int*)start2, (const int*)fstart, (const Kp*)(ws + B->o_k1a), (const Kp*)(ws + B->o_k1b),
                       (Kp*)(ws + B->o_k1));
    int* n2 = cnt + 2 * n;
    hipLaunchKernelGGL(k_bselect, dim3(64, n), dim3(256), 0, s, (const int*)fstart, (const Kp*)(ws + B->o_k1), B->maxkp,
                       (Kp*)(ws + B->o_k2), n2);
    // A8
    hipLaunchKernelGGL(k_borient, dim3(gcap(((long long)n * B->maxkp + 3) / 4, 4096)), dim3(256), 0, s, B->d_img,
                       B->d_oct, n, B->maxkp, ws, (const Kp*)(ws + B->o_k2), (const int*)n2, h->tab,
                       (float4*)(ws + B->o_ang), (int*)(ws + B->o_nang));
    launch_scan((int*)(ws + B->o_nang), (int*)(ws + B->o_aoff), n * B->maxkp, nullptr, nullptr, nullptr, 0, s,
                (int*)(ws + B->o_stmp));
    hipLaunchKernelGGL(k_bexpand, dim3(gcap(((long long)n * B->maxkp + 255) / 256, 4096)), dim3(256), 0, s, B->d_img, n,
                       B->maxkp, (const Kp*)(ws + B->o_k2), (const int*)n2, (const float4*)(ws + B->o_ang),
                       (const int*)(ws + B->o_nang), (const int*)(ws + B->o_aoff), (Ent*)(ws + B->o_E));
    hipLaunchKernelGGL(k_bzero, dim3(gcap(((long long)n * 4 * B->maxkp + 255) / 256, 2048)), dim3(256), 0, s,
                       (int*)(ws + B->o_rank), n * 4 * B->maxkp);
    hipLaunchKernelGGL(k_brank_xy, dim3(8, 32, n), dim3(256), 0, s, (const Ent*)(ws + B->o_E), (const int*)(ws + B->o_aoff),
                       B->maxkp, (int*)(ws + B->o_rank));
    // A9 + outputs (images concatenated, each sorted by (x, y))
    hipLaunchKernelGGL(k_bdescriptor, dim3(gcap((long long)n * B->maxkp * 4, 8192)), dim3(kDescT), 0, s, B->d_img, B->d_oct,
                       n, B->maxkp, ws, (const Kp*)(ws + B->o_k2), (const Ent*)(ws + B->o_E),
                       (const int*)(ws + B->o_aoff), (const int*)(ws + B->o_rank), cap, h->tab, p.rootsift, kp, df, du,
                       counts, offsets);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift

using namespace msift;

extern "C" {

int msift_batch_create(msift_handle* h, int n, const int* H, const int* W, msift_batch** plan) {
    if (!h) { set_error("null handle"); return MSIFT_EINVAL; }
    return batch_create(h, n, H, W, plan);
}

int msift_batch_destroy(msift_batch* plan) { return batch_destroy(plan); }

int msift_batch_sizes(msift_batch* B, size_t* ws_bytes, size_t* input_bytes, size_t* input_offsets) {
    if (!B) { set_error("null plan"); return MSIFT_EINVAL; }
    if (ws_bytes) *ws_bytes = B->total;
    if (input_bytes) *input_bytes = B->in_bytes;
    if (input_offsets)
        for (int i = 0; i < B->n; ++i) input_offsets[i] = B->img[i].in_off;
    return MSIFT_OK;
}

int msift_batch_set_blur_tile(msift_batch* B, int tile_h) {
    if (!B) { set_error("null plan"); return MSIFT_EINVAL; }
    return batch_set_blur_tile(B, tile_h);
}

int msift_detect_describe_batched(msift_batch* B, const uint8_t* d_packed, void* d_ws, size_t ws_bytes, int cap,
                                  float* d_kp, float* d_desc_f32, uint8_t* d_desc_u8, int* d_counts, int* d_offsets,
                                  void* stream) {
    if (!B) { set_error("null plan"); return MSIFT_EINVAL; }
    return batch_run(B, d_packed, d_ws, ws_bytes, cap, d_kp, d_desc_f32, d_desc_u8, d_counts, d_offsets,
                     (hipStream_t)stream);
}

}  // extern "C"
