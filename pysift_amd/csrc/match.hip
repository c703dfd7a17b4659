// Exact brute-force ratio-test matcher on uint8 descriptors (restates
// _match_features / _match_symmetric, pysift/core.py:2840-2940, 2652-2701).
//
// d(i,j) = |a_i|^2 + |b_j|^2 - 2 a_i.b_j with a.b on INT8 MFMA
// (v_mfma_i32_32x32x32_i8, INT32 accumulation), exact:
//   method 0 (high/low split): a = 128 h + l, h in {0,1}, l in [0,127];
//     a.b = l.l + 128 (h_a.l_b + l_a.h_b + 2 h_a.(64 h_b))  (two accumulators,
//     every operand byte in [0,127])
//   method 1 (shift): a' = a - 128 is int8 and d = |a'|^2 + |b'|^2 - 2 a'.b'.
// A workgroup (8 waves) owns 256 rows x a 512-column chunk; each wave
// computes 32x64 tiles.  Top-2 per row and per column is tracked with 32-bit keys
// (d << bits | local index), which are unique, so the result does not depend
// on evaluation order; small reduce kernels merge the per-chunk partials.
#include "msift_internal.hpp"

#pragma clang fp contract(off)

namespace msift {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

constexpr int kRowBlk = 256;   // rows per workgroup
constexpr int kColChunk = 512; // columns per workgroup (9-bit local index)

__device__ __forceinline__ void top2_insert(uint32_t& k1, uint32_t& k2, uint32_t k) {
    uint32_t t = max(k1, k);
    k1 = min(k1, k);
    k2 = min(k2, t);
}

__device__ __forceinline__ void top2_merge(uint32_t& k1, uint32_t& k2, uint32_t o1, uint32_t o2) {
    uint32_t n2 = min(max(k1, o1), min(k2, o2));
    k1 = min(k1, o1);
    k2 = n2;
}

__device__ __forceinline__ void top2_insert64(uint64_t& k1, uint64_t& k2, uint64_t k) {
    uint64_t t = k1 > k ? k1 : k;
    k1 = k1 < k ? k1 : k;
    k2 = k2 < t ? k2 : t;
}

// squared norms (method 0: of a, method 1: of a - 128), exact int32
__global__ __launch_bounds__(256) void k_norms(const uint8_t* __restrict__ a, int n, int method,
                                                int* __restrict__ out) {
    int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (i >= n) return;
    int v0 = a[(size_t)i * 128 + 2 * lane], v1 = a[(size_t)i * 128 + 2 * lane + 1];
    if (method == 1) { v0 -= 128; v1 -= 128; }
    int s = v0 * v0 + v1 * v1;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if (lane == 0) out[i] = s;
}

template <int METHOD>
__global__ __launch_bounds__(512) void k_match_tiles(const uint8_t* __restrict__ A, int N, const uint8_t* __restrict__ B,
                                                      int M, const int* __restrict__ na, const int* __restrict__ nb,
                                                      uint32_t* __restrict__ rowpart, uint32_t* __restrict__ colpart) {
    __shared__ uint32_t colk[8][64][2];
    __shared__ int s_na[kRowBlk];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const int rb = blockIdx.x, cc = blockIdx.y;
    const int row0 = rb * kRowBlk + w * 32;     // this wave's 32 rows
    if (threadIdx.x < kRowBlk) {
        int i = rb * kRowBlk + threadIdx.x;
        s_na[threadIdx.x] = i < N ? na[i] : 0;
    }
    // A fragment: lane holds row (row0 + l32), bytes [32 ks + 16 hi, +16) for ks = 0..3
    v4i al[4], ah[4];
    {
        int i = row0 + l32;
        int ic = i < N ? i : N - 1;
        const v4i* src = reinterpret_cast<const v4i*>(A + (size_t)ic * 128);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            v4i v = src[2 * ks + hi];
            if (METHOD == 1) {
                al[ks] = v ^ (v4i)(int)0x80808080;
            } else {
                al[ks] = v & (v4i)0x7f7f7f7f;
                ah[ks] = (v >> 7) & (v4i)0x01010101;
            }
        }
    }
    uint32_t rk1[16], rk2[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) { rk1[r] = 0xffffffffu; rk2[r] = 0xffffffffu; }
    __syncthreads();

    for (int ct = 0; ct < kColChunk / 64; ++ct) {
        const int col0 = cc * kColChunk + ct * 64;
        if (col0 >= M) break;
        v16i acc[2], accx[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            int j = col0 + 32 * b + l32;
            int jc = j < M ? j : M - 1;
            const v4i* src = reinterpret_cast<const v4i*>(B + (size_t)jc * 128);
            acc[b] = (v16i)0;
            if (METHOD == 0) accx[b] = (v16i)0;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                v4i v = src[2 * ks + hi];
                if (METHOD == 1) {
                    acc[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(al[ks], v ^ (v4i)(int)0x80808080, acc[b], 0, 0, 0);
                } else {
                    v4i bl = v & (v4i)0x7f7f7f7f;
                    v4i bh = (v >> 7) & (v4i)0x01010101;
                    v4i bh64 = bh << 6;
                    acc[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(al[ks], bl, acc[b], 0, 0, 0);
                    // cross term h_a.l_b + l_a.h_b + 2 h_a.(64 h_b) = h_a.l_b + l_a.h_b + 128 h_a.h_b
                    accx[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[ks], bl, accx[b], 0, 0, 0);
                    accx[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(al[ks], bh, accx[b], 0, 0, 0);
                    accx[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[ks], bh64, accx[b], 0, 0, 0);
                    accx[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[ks], bh64, accx[b], 0, 0, 0);
                }
            }
        }
        // epilogue: lane holds column (col0 + 32b + l32), rows row0 + 8(r/4) + 4 hi + r%4
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = col0 + 32 * b + l32;
            const bool jv = j < M;
            const int nbj = jv ? nb[j] : 0;
            const uint32_t lj = (uint32_t)(ct * 64 + 32 * b + l32);
            uint32_t c1 = 0xffffffffu, c2 = 0xffffffffu;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int il = 32 * w + 8 * (r >> 2) + 4 * hi + (r & 3);   // row within the block
                const bool v = jv && (rb * kRowBlk + il < N);
                int dot = acc[b][r];
                if (METHOD == 0) dot = dot + 128 * accx[b][r];
                const int d = (s_na[il] + nbj) - 2 * dot;
                const uint32_t kc = v ? (((uint32_t)d << 8) | (uint32_t)il) : 0xffffffffu;
                top2_insert(c1, c2, kc);
                const uint32_t kr = v ? (((uint32_t)d << 9) | lj) : 0xffffffffu;
                top2_insert(rk1[r], rk2[r], kr);
            }
            uint32_t o1 = __shfl_xor(c1, 32), o2 = __shfl_xor(c2, 32);
            top2_merge(c1, c2, o1, o2);
            if (hi == 0) {
                colk[w][32 * b + l32][0] = c1;
                colk[w][32 * b + l32][1] = c2;
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const int c = threadIdx.x, j = col0 + c;
            uint32_t k1 = colk[0][c][0], k2 = colk[0][c][1];
#pragma unroll
            for (int ww = 1; ww < 8; ++ww) top2_merge(k1, k2, colk[ww][c][0], colk[ww][c][1]);
            if (j < M) {
                colpart[((size_t)rb * M + j) * 2] = k1;
                colpart[((size_t)rb * M + j) * 2 + 1] = k2;
            }
        }
        __syncthreads();
    }
    // rows: merge across the 32 lanes that share each row slot
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        uint32_t k1 = rk1[r], k2 = rk2[r];
#pragma unroll
        for (int m = 1; m < 32; m <<= 1) {
            uint32_t o1 = __shfl_xor(k1, m), o2 = __shfl_xor(k2, m);
            top2_merge(k1, k2, o1, o2);
        }
        const int i = row0 + 8 * (r >> 2) + 4 * hi + (r & 3);
        if (l32 == 0 && i < N) {
            rowpart[((size_t)cc * N + i) * 2] = k1;
            rowpart[((size_t)cc * N + i) * 2 + 1] = k2;
        }
    }
}

// merge per-chunk partials into (best index, d1, d2); key 0xffffffff = none
__global__ __launch_bounds__(256) void k_reduce_top2(const uint32_t* __restrict__ part, int n, int nparts,
                                                      int shift, int span, int32_t* __restrict__ out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint64_t k1 = ~0ull, k2 = ~0ull;
    for (int p = 0; p < nparts; ++p) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            uint32_t k = part[((size_t)p * n + i) * 2 + q];
            if (k == 0xffffffffu) continue;
            uint64_t d = k >> shift;
            uint64_t idx = (uint64_t)p * span + (k & ((1u << shift) - 1));
            top2_insert64(k1, k2, (d << 32) | idx);
        }
    }
    out[3 * i] = k1 == ~0ull ? -1 : (int32_t)(k1 & 0xffffffffu);
    out[3 * i + 1] = k1 == ~0ull ? 0x7fffffff : (int32_t)(k1 >> 32);
    out[3 * i + 2] = k2 == ~0ull ? 0x7fffffff : (int32_t)(k2 >> 32);
}

__global__ void k_fill_top2_none(int32_t* __restrict__ out, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[3 * i] = -1; out[3 * i + 1] = 0x7fffffff; out[3 * i + 2] = 0x7fffffff;
}

// mutual ratio test (core.py:2935-2939), then raster-order compaction
__device__ __forceinline__ bool mutual_ok(const int32_t* fwd, const int32_t* bwd, int N, int M, double r2, int i) {
    if (N < 2 || M < 2) return false;
    int j = fwd[3 * i];
    if (j < 0) return false;
    if (!((double)fwd[3 * i + 1] < r2 * (double)fwd[3 * i + 2])) return false;
    if (bwd[3 * j] != i) return false;
    return (double)bwd[3 * j + 1] < r2 * (double)bwd[3 * j + 2];
}

__global__ __launch_bounds__(256) void k_mutual_count(const int32_t* __restrict__ fwd, const int32_t* __restrict__ bwd,
                                                       int N, int M, double r2, int* __restrict__ counts) {
    __shared__ int wc[4];
    int i = blockIdx.x * 256 + threadIdx.x;
    bool ok = i < N && mutual_ok(fwd, bwd, N, M, r2, i);
    unsigned long long m = __ballot(ok);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = (wc[0] + wc[1]) + (wc[2] + wc[3]);
}

__global__ __launch_bounds__(256) void k_mutual_write(const int32_t* __restrict__ fwd, const int32_t* __restrict__ bwd,
                                                       int N, int M, double r2, const int* __restrict__ offs,
                                                       int32_t* __restrict__ pairs) {
    __shared__ int wc[4];
    int i = blockIdx.x * 256 + threadIdx.x;
    bool ok = i < N && mutual_ok(fwd, bwd, N, M, r2, i);
    unsigned long long m = __ballot(ok);
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wc[wave] = __popcll(m);
    __syncthreads();
    if (!ok) return;
    int pre = 0;
    for (int w = 0; w < wave; ++w) pre += wc[w];
    int pos = offs[blockIdx.x] + pre + __popcll(m & lanemask_lt());
    pairs[2 * pos] = i;
    pairs[2 * pos + 1] = fwd[3 * i];
}

// ---------------------------------------------------------------- host ----
struct MatchPlan {
    int nrb_a, ncc_a;   // row blocks of a, column chunks of b
    size_t o_na, o_nb, o_rowpart, o_colpart, o_cnt, o_off, total;
};

static MatchPlan match_plan(int N, int M) {
    MatchPlan P;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    P.nrb_a = (N + kRowBlk - 1) / kRowBlk;
    P.ncc_a = (M + kColChunk - 1) / kColChunk;
    P.o_na = take((size_t)(N > 0 ? N : 1) * 4);
    P.o_nb = take((size_t)(M > 0 ? M : 1) * 4);
    P.o_rowpart = take((size_t)P.ncc_a * (N > 0 ? N : 1) * 8);
    P.o_colpart = take((size_t)P.nrb_a * (M > 0 ? M : 1) * 8);
    int nblk = (N + 255) / 256;
    P.o_cnt = take((size_t)(nblk + 1) * 4);
    P.o_off = take((size_t)(nblk + 1) * 4);
    P.total = off;
    return P;
}

size_t match_workspace(int N, int M) { return match_plan(N, M).total; }

int match_u8(const uint8_t* a, int N, const uint8_t* b, int M, double ratio, int method, void* ws_, size_t ws_bytes,
             int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s) {
    MatchPlan P = match_plan(N, M);
    if (ws_bytes < P.total) { set_error("match_u8: workspace too small"); return MSIFT_EWORKSPACE; }
    if ((N > 0 && (!a || !fwd)) || (M > 0 && (!b || !bwd)) || (!pairs && N > 0 && M > 0)) {
        set_error("match_u8: null buffer");
        return MSIFT_EINVAL;
    }
    if (((uintptr_t)a & 15) || ((uintptr_t)b & 15)) { set_error("match_u8: descriptors must be 16-byte aligned"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    int* na = (int*)(ws + P.o_na);
    int* nb = (int*)(ws + P.o_nb);
    uint32_t* rowpart = (uint32_t*)(ws + P.o_rowpart);
    uint32_t* colpart = (uint32_t*)(ws + P.o_colpart);
    if (N > 0 && M > 0) {
        hipLaunchKernelGGL(k_norms, dim3((N + 3) / 4), dim3(256), 0, s, a, N, method, na);
        hipLaunchKernelGGL(k_norms, dim3((M + 3) / 4), dim3(256), 0, s, b, M, method, nb);
        dim3 grid(P.nrb_a, P.ncc_a);
        if (method == 0)
            hipLaunchKernelGGL(k_match_tiles<0>, grid, dim3(512), 0, s, a, N, b, M, na, nb, rowpart, colpart);
        else
            hipLaunchKernelGGL(k_match_tiles<1>, grid, dim3(512), 0, s, a, N, b, M, na, nb, rowpart, colpart);
        hipLaunchKernelGGL(k_reduce_top2, dim3((N + 255) / 256), dim3(256), 0, s, rowpart, N, P.ncc_a, 9, kColChunk, fwd);
        hipLaunchKernelGGL(k_reduce_top2, dim3((M + 255) / 256), dim3(256), 0, s, colpart, M, P.nrb_a, 8, kRowBlk, bwd);
    } else {
        if (N > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((N + 255) / 256), dim3(256), 0, s, fwd, N);
        if (M > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((M + 255) / 256), dim3(256), 0, s, bwd, M);
    }
    int nblk = (N + 255) / 256;
    double r2 = ratio * ratio;
    int* cnt = (int*)(ws + P.o_cnt);
    int* off = (int*)(ws + P.o_off);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_count, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, cnt);
    launch_scan(cnt, off, nblk, npairs, nullptr, nullptr, 0, s);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_write, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, off, pairs);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

int match_f16(const float*, int, const float*, int, double, void*, size_t, int32_t*, int32_t*, int32_t*, int*,
              hipStream_t) {
    set_error("match_f16: float descriptor path is not built");
    return MSIFT_ESTATE;
}

}  // namespace msift
