// Exact brute-force ratio-test matcher on uint8 descriptors (restates
// _match_features / _match_symmetric, pysift/core.py:2840-2940, 2652-2701).
//
// d(i,j) = |a_i|^2 + |b_j|^2 - 2 a_i.b_j with a.b on INT8 MFMA
// (v_mfma_i32_32x32x32_i8, INT32 accumulation), exact:
//   method 0 (high/low split): a = 128 h + l, h in {0,1}, l in [0,127];
//     a.b = l.l + 128 (h_a.l_b + l_a.h_b + 2 h_a.(64 h_b))  (two accumulators,
//     every operand byte in [0,127])
//   method 1 (shift): a' = a - 128 is int8 and d = |a'|^2 + |b'|^2 - 2 a'.b';
//     the MFMA takes ~a' = -a' - 1 (also int8) and b', returning
//     dn = -a'.b' - sum b', so d = |a'|^2 + (|b'|^2 + 2 sum b') + 2 dn and
//     each key is two integer instructions away from the accumulator.
// A workgroup (8 waves) owns 256 rows x a 512-column chunk; each wave
// computes 32x64 tiles.  Top-2 per row and per column is tracked with 32-bit keys
// (d << bits | local index), which are unique, so the result does not depend
// on evaluation order; small reduce kernels merge the per-chunk partials.
#include "msift_internal.hpp"
#include <algorithm>
#include <cstring>
#include <vector>
#include <mutex>
#include <atomic>

#pragma clang fp contract(off)

namespace msift {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

constexpr int kRowBlk = 256;   // rows per workgroup
constexpr uint32_t kKeyOff = 1u << 21;   // offset of the shift method's keys (see match_tile)
constexpr int kColChunk = 512; // columns per workgroup (9-bit local index)

// top-2 of unique 32-bit keys, k1 <= k2 always: the new second is the median
// of (k1, k2, k) and the new first min(k1, k).  The median is written as the
// instruction: from a chain that starts at (~0, ~0) the compiler folds the
// min/max form into three operations instead of one v_med3_u32.
__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ void top2_insert(uint32_t& k1, uint32_t& k2, uint32_t k) {
    const uint32_t m = med3_u32(k1, k2, k);
    k1 = min(k1, k);
    k2 = m;
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

// squared norms, exact int32: method 0 |a|^2, method 1 |a - 128|^2, mode 2
// |b'|^2 + 2 sum b' with b' = b - 128 (the column term of the shift
// method's keys, whose A operand is fed negated)
__global__ __launch_bounds__(256) void k_norms(const uint8_t* __restrict__ a, int n, int method,
                                                int* __restrict__ out) {
    int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (i >= n) return;
    int v0 = a[(size_t)i * 128 + 2 * lane], v1 = a[(size_t)i * 128 + 2 * lane + 1];
    if (method != 0) { v0 -= 128; v1 -= 128; }
    int s = v0 * v0 + v1 * v1;
    if (method == 2) s += 2 * (v0 + v1);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if (lane == 0) out[i] = s;
}

// one pair of a batched match
struct MPair {
    int a_off, N, b_off, M;   // rows of a / b in the packed sets
    int nrb, ncc, tile0;      // row blocks, column chunks, first tile
    int blk0, nblk;           // first 256-row block of the mutual pass, its count
    int tp0, xb1, cg1, xb2, cg2;   // two-pass: first tile; x-blocks and chunk groups of pass 1 and 2
    size_t o_rowpart, o_colpart;   // workspace offsets of the partial top-2 keys
};

// two-pass prologue over the rows of a and b (8 threads a row, 16 bytes each):
// the y' = y - 128 copy, |y'|^2 (the reduction's adj) and the finished key term
// ((|y'|^2 + 2 sum y' + 2^21) << 9) + (row mod 512) that a row contributes as y
// (batched: rows of many pairs packed back to back, pairs != nullptr; the row
// index in the key term is then counted from the row's own pair start)
__global__ __launch_bounds__(256) void k_tp_prep(const uint8_t* __restrict__ a, int N, const uint8_t* __restrict__ b,
                                                  int M, uint8_t* __restrict__ ac, uint8_t* __restrict__ bc,
                                                  int* __restrict__ na, int* __restrict__ nb,
                                                  uint32_t* __restrict__ ta, uint32_t* __restrict__ tb,
                                                  const MPair* __restrict__ pairs, int np) {
    const int r = blockIdx.x * 32 + (threadIdx.x >> 3), q = threadIdx.x & 7;
    if (r >= N + M) return;   // whole groups of 8 leave together
    const bool isb = r >= N;
    const int i = isb ? r - N : r;
    int off = 0;
    if (pairs) {   // last pair whose set starts at or before row i (pair 0 starts at 0)
        int lo = 0, hi = np - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((isb ? pairs[mid].b_off : pairs[mid].a_off) <= i) lo = mid; else hi = mid - 1;
        }
        off = isb ? pairs[lo].b_off : pairs[lo].a_off;
    }
    const uint8_t* src = (isb ? b : a) + (size_t)i * 128 + 16 * q;
    const v4i v = *reinterpret_cast<const v4i*>(src) ^ (v4i)(int)0x80808080;
    *reinterpret_cast<v4i*>((isb ? bc : ac) + (size_t)i * 128 + 16 * q) = v;
    int s2 = 0, s1 = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = (int)(int8_t)(v[e] >> (8 * k));
            s2 += x * x;
            s1 += x;
        }
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        s2 += __shfl_xor(s2, m);
        s1 += __shfl_xor(s1, m);
    }
    if (q == 0) {
        (isb ? nb : na)[i] = s2;
        (isb ? tb : ta)[i] = (((uint32_t)(s2 + 2 * s1) + kKeyOff) << 9) + (uint32_t)((i - off) & 511);
    }
}

// one 256-row x 512-column tile (row block rb, column chunk cc) of a
// workgroup of 8 waves.  Each 64-column step of b is staged once in LDS
// (rows padded to 144 B: conflict-free 16-byte reads), double-buffered: the
// next step's global load is issued before the current step's MFMA and top-2
// work.  Per-wave column partials stay in LDS for the whole chunk and are
// merged once at the end.  Results do not depend on this schedule (unique
// keys).
constexpr int kBStride = 144;                     // bytes per staged column
constexpr int kBTile = 64 * kBStride;             // one 64-column step

struct MatchSmem {
    uint8_t bt[2][kBTile];                        // staged b columns, double buffer
    uint32_t colk[8][kColChunk][2];               // per-wave column top-2 for the chunk
    int s_na[kRowBlk];                            // method 0: |a|^2
    uint32_t s_nac[kRowBlk];                      // method 1: ((|a'|^2 + 2^21) << 8) | row (column keys)
    int s_nb[kColChunk];                          // the chunk's column terms (no global load in
                                                  // the epilogue, which would wait on the prefetch)
};

template <int METHOD>
__device__ __forceinline__ void match_tile(const uint8_t* __restrict__ A, int N, const uint8_t* __restrict__ B, int M,
                                           const int* __restrict__ na, const int* __restrict__ nb,
                                           uint32_t* __restrict__ rowpart, uint32_t* __restrict__ colpart, int rb,
                                           int cc, MatchSmem& sm) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const int row0 = rb * kRowBlk + w * 32;     // this wave's 32 rows
    {
        const int j = cc * kColChunk + threadIdx.x;
        // method 1 stages its row keys' whole column term ((nbx + 2^21) << 9) + local column
        const int v = j < M ? nb[j] : 0;
        sm.s_nb[threadIdx.x] = METHOD == 1 ? (int)((((uint32_t)v + kKeyOff) << 9) + threadIdx.x) : v;
    }
    if (threadIdx.x < kRowBlk) {
        int i = rb * kRowBlk + threadIdx.x;
        const int v = i < N ? na[i] : 0;
        if (METHOD == 1) {
            sm.s_nac[threadIdx.x] = (((uint32_t)v + kKeyOff) << 8) | (uint32_t)threadIdx.x;
        } else {
            sm.s_na[threadIdx.x] = v;
        }
    }
    // A fragment: lane holds row (row0 + l32), bytes [32 ks + 16 hi, +16) for ks = 0..3;
    // method 1 feeds ~(a - 128) = -(a - 128) - 1, so the MFMA returns -a'.b' - sum b'
    v4i al[4], ah[4];
    {
        int i = row0 + l32;
        int ic = i < N ? i : N - 1;
        const v4i* src = reinterpret_cast<const v4i*>(A + (size_t)ic * 128);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            v4i v = src[2 * ks + hi];
            if (METHOD == 1) {
                al[ks] = v ^ (v4i)(int)0x7f7f7f7f;
            } else {
                al[ks] = v & (v4i)0x7f7f7f7f;
                ah[ks] = (v >> 7) & (v4i)0x01010101;
            }
        }
    }
    uint32_t rk1[16], rk2[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) { rk1[r] = 0xffffffffu; rk2[r] = 0xffffffffu; }
    // staging of step ct: thread t moves bytes [16 (t % 8), +16) of column t / 8
    const int n_ct = min(kColChunk / 64, (M - cc * kColChunk + 63) / 64);
    const int sc = threadIdx.x >> 3, sq = threadIdx.x & 7;
    auto gload = [&](int ct) -> v4i {
        int j = cc * kColChunk + ct * 64 + sc;
        j = j < M ? j : M - 1;
        return *reinterpret_cast<const v4i*>(B + (size_t)j * 128 + 16 * sq);
    };
    auto lstore = [&](int buf, v4i v) {
        if (METHOD == 1) v = v ^ (v4i)(int)0x80808080;   // staged as b' = b - 128
        *reinterpret_cast<v4i*>(&sm.bt[buf][sc * kBStride + 16 * sq]) = v;
    };
    lstore(0, gload(0));
    __syncthreads();
    // the lane's 16 rows keep their norm terms in registers for the whole chunk
    uint32_t rn[16], rc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int il = 32 * w + 8 * (r >> 2) + 4 * hi + (r & 3);
        rn[r] = METHOD == 1 ? 0u : (uint32_t)sm.s_na[il];
        rc[r] = METHOD == 1 ? sm.s_nac[il] : 0u;
    }
    auto mfma_step = [&](int step, v16i* acc, v16i* accx) {
        const uint8_t* bt = sm.bt[step & 1];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const uint8_t* src = bt + (32 * b + l32) * kBStride + 16 * hi;
            acc[b] = (v16i)0;
            if (METHOD == 0) accx[b] = (v16i)0;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                v4i v = *reinterpret_cast<const v4i*>(src + 32 * ks);
                if (METHOD == 1) {
                    acc[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(al[ks], v, acc[b], 0, 0, 0);
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
    };
    for (int ct = 0; ct < n_ct; ++ct) {
        const int col0 = cc * kColChunk + ct * 64;
        const bool more = ct + 1 < n_ct;
        v4i nxt;
        if (more) nxt = gload(ct + 1);
        v16i acc[2], accx[2];
        mfma_step(ct, acc, accx);
        // epilogue: lane holds column (col0 + 32b + l32), rows row0 + 8(r/4) + 4 hi + r%4.
        // Keys: row (d << 9) | local column, column (d << 8) | local row.  Method 1
        // drops from each key the term that is constant along it: with dn the MFMA
        // result -a'.b' - sum b', nbx = |b'|^2 + 2 sum b' and d = |a'|^2 + nbx + 2 dn,
        // the row key is ((d - |a'|^2 + 2^21) << 9) | lj = (dn << 10) + ((nbx + 2^21) << 9) + lj
        // and the column key ((d - nbx + 2^21) << 8) | il = (dn << 9) + ((|a'|^2 + 2^21) << 8) + il,
        // one shift-add each, exact mod 2^32 (both offset values lie in [0, 2^23) for every
        // pair of u8 vectors).  The reduction adds the dropped term back.  A step whose
        // rows and columns are all in range skips the validity selects (the condition is
        // uniform over the block).
        const bool full = (rb + 1) * kRowBlk <= N && col0 + 64 <= M;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = col0 + 32 * b + l32;
            const bool jv = j < M;
            const uint32_t lj = (uint32_t)(ct * 64 + 32 * b + l32);
            const int nbj = sm.s_nb[lj];
            uint32_t c1 = 0xffffffffu, c2 = 0xffffffffu;
            const uint32_t nbr = METHOD == 1 ? (uint32_t)nbj : ((uint32_t)nbj << 9) + lj;
            if (full) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int il = 32 * w + 8 * (r >> 2) + 4 * hi + (r & 3);
                    uint32_t kc, kr;
                    if (METHOD == 1) {
                        const uint32_t dn = (uint32_t)acc[b][r];
                        kr = nbr + (dn << 10);
                        kc = rc[r] + (dn << 9);
                    } else {
                        const int d = ((int)rn[r] + nbj) - 2 * (acc[b][r] + 128 * accx[b][r]);
                        kc = ((uint32_t)d << 8) | (uint32_t)il;
                        kr = ((uint32_t)d << 9) | lj;
                    }
                    top2_insert(c1, c2, kc);
                    top2_insert(rk1[r], rk2[r], kr);
                }
            } else {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int il = 32 * w + 8 * (r >> 2) + 4 * hi + (r & 3);   // row within the block
                    const bool v = jv && (rb * kRowBlk + il < N);
                    uint32_t kc, kr;
                    if (METHOD == 1) {
                        const uint32_t dn = (uint32_t)acc[b][r];
                        kr = nbr + (dn << 10);
                        kc = rc[r] + (dn << 9);
                    } else {
                        const int d = ((int)rn[r] + nbj) - 2 * (acc[b][r] + 128 * accx[b][r]);
                        kc = ((uint32_t)d << 8) | (uint32_t)il;
                        kr = ((uint32_t)d << 9) | lj;
                    }
                    top2_insert(c1, c2, v ? kc : 0xffffffffu);
                    top2_insert(rk1[r], rk2[r], v ? kr : 0xffffffffu);
                }
            }
            uint32_t o1 = __shfl_xor(c1, 32), o2 = __shfl_xor(c2, 32);
            top2_merge(c1, c2, o1, o2);
            if (hi == 0) {
                sm.colk[w][ct * 64 + 32 * b + l32][0] = c1;
                sm.colk[w][ct * 64 + 32 * b + l32][1] = c2;
            }
        }
        if (more) lstore((ct + 1) & 1, nxt);
        __syncthreads();
    }
    // columns: one merge over the 8 waves per column of the chunk
    {
        const int c = threadIdx.x, j = cc * kColChunk + c;
        if (c < n_ct * 64 && j < M) {
            uint32_t k1 = sm.colk[0][c][0], k2 = sm.colk[0][c][1];
#pragma unroll
            for (int ww = 1; ww < 8; ++ww) top2_merge(k1, k2, sm.colk[ww][c][0], sm.colk[ww][c][1]);
            colpart[((size_t)rb * M + j) * 2] = k1;
            colpart[((size_t)rb * M + j) * 2 + 1] = k2;
        }
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

template <int METHOD>
__global__ __launch_bounds__(512) void k_match_tiles(const uint8_t* __restrict__ A, int N, const uint8_t* __restrict__ B,
                                                      int M, const int* __restrict__ na, const int* __restrict__ nb,
                                                      uint32_t* __restrict__ rowpart, uint32_t* __restrict__ colpart) {
    __shared__ MatchSmem sm;
    match_tile<METHOD>(A, N, B, M, na, nb, rowpart, colpart, blockIdx.x, blockIdx.y, sm);
}

// ------------------------------------------------------------ two-pass ----
// method 2 (two passes): the shift method's arithmetic, run as two passes of
// one kernel, each the top-2 of every descriptor x of X over all of Y (pass
// 1: X = a, Y = b; pass 2: X = b, Y = a).  The MFMA takes the streamed
// y' = y - 128 as its A operand and the fixed ~x' = -(x - 128) - 1 as B, so
// lane l's accumulator holds the distances of ONE x (column l % 32 of the
// 32x32 output) to 16 y (rows 8 (r / 4) + 4 (l / 32) + r % 4): the top-2 is a
// register pair per x, one shift-add and one median insert per distance, and
// the two half-waves merge once per 512-y chunk.  It does twice the fused
// kernel's MFMA work and none of its per-step column merges, per-chunk row
// butterflies or per-step barriers: the fused kernel is bound by that VALU
// and synchronisation work, not by the MFMA.
//   Keys, with dn = -x'.y' - sum y' (the accumulator) and ty = |y'|^2 + 2 sum y':
//   ((d - |x'|^2 + 2^21) << 9) | lj = (dn << 10) + (((ty + 2^21) << 9) + lj),
//   lj = y mod 512: the shift method's row keys, so k_reduce_top2 with adj
//   |x'|^2 finishes either pass.  k_tp_prep writes the y' copies and the
//   bracketed term per row, so a stage is a plain copy (global -> LDS direct,
//   no registers or VALU).
constexpr int kTpStage = 256;   // y per LDS stage, two stages per 512-y chunk
constexpr int kTpWaves = 4;     // waves per workgroup

// block swizzle of the staged rows: distinct over any 8 aligned consecutive rows
// and over the rows of one parity among 16, so the 16-byte A-fragment reads
// of 8 (32 banks) or 16 (64 banks) consecutive lanes hit disjoint banks
__device__ __forceinline__ int tp_swz(int c) { return (c & 7) ^ ((c >> 3) & 1); }

struct TpSmem {
    alignas(16) uint8_t y[kTpStage * 128];   // y' rows: 16-byte block q of row c at block position q ^ tp_swz(c)
    alignas(16) uint32_t t[kTpStage];        // key terms ((ty + 2^21) << 9) + lj
};

// NT x-tiles of 32 per wave; PIPE: in a stage whose 256 y are all in range,
// tile t + 1's MFMAs are issued before tile t's inserts (two accumulator sets)
// x-block xb (kTpWaves * 32 * NT descriptors of X) against 512-y chunks
// cg * cpw .. +cpw of Y; part[(chunk * NX + x) * 2 + {0, 1}] = that chunk's top-2 keys
template <int NT, bool PIPE>
__device__ __forceinline__ void match_side_tile(const uint8_t* __restrict__ X, int NX, const uint8_t* __restrict__ Yc,
                                                int NY, const uint32_t* __restrict__ ty, uint32_t* __restrict__ part,
                                                int xb, int cg, int cpw, TpSmem& sm) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const int xw = xb * (kTpWaves * 32 * NT) + w * 32 * NT;   // this wave's first x
    // B fragments: lane holds x = xw + 32 n + l32, bytes [32 ks + 16 hi, +16), fed as ~x'
    v4i xf[NT][4];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int x = min(xw + 32 * n + l32, NX - 1);
        const v4i* src = reinterpret_cast<const v4i*>(X + (size_t)x * 128);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) xf[n][ks] = src[2 * ks + hi] ^ (v4i)(int)0x7f7f7f7f;
    }
    // LDS offsets of the lane's A fragment blocks (row l32 of a 32-y tile; tp_swz(32 t + l32) = tp_swz(l32))
    int yoff[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) yoff[ks] = l32 * 128 + 16 * ((2 * ks + hi) ^ tp_swz(l32));
    auto fetch = [&](int t, v4i* yf, v4i* tv) {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) yf[ks] = *reinterpret_cast<const v4i*>(&sm.y[t * 4096 + yoff[ks]]);
#pragma unroll
        for (int q = 0; q < 4; ++q) tv[q] = *reinterpret_cast<const v4i*>(&sm.t[32 * t + 8 * q + 4 * hi]);
    };
    auto mma = [&](const v4i* yf, v16i* acc) {
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            acc[n] = (v16i)0;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                acc[n] = __builtin_amdgcn_mfma_i32_32x32x32_i8(yf[ks], xf[n][ks], acc[n], 0, 0, 0);
        }
    };
    const int nch = (NY + 511) / 512;
    const int ch0 = cg * cpw, ch1 = min(nch, ch0 + cpw);
    for (int ch = ch0; ch < ch1; ++ch) {
        uint32_t k1[NT], k2[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) { k1[n] = 0xffffffffu; k2[n] = 0xffffffffu; }
        auto insert_full = [&](const v16i* acc, const v4i* tv) {
#pragma unroll
            for (int n = 0; n < NT; ++n)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    top2_insert(k1[n], k2[n], ((uint32_t)acc[n][r] << 10) + (uint32_t)tv[r >> 2][r & 3]);
        };
        const int yend = min(NY, ch * 512 + 512);
        for (int ybase = ch * 512; ybase < yend; ybase += kTpStage) {
            __syncthreads();   // the previous stage's reads are done
            // Instruction i of wave w fills stage rows 8 (8 w + i) .. +8 (1 KB): lane L's 16
            // bytes land at byte 16 L, i.e. row 8 (8 w + i) + L / 8, block position L % 8,
            // which holds the row's block (L % 8) ^ tp_swz(row).  Rows past NY repeat row NY - 1
            // (their keys are never inserted).
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int c = (8 * w + i) * 8 + (lane >> 3), q = (lane & 7) ^ tp_swz(c);
                const int y = min(ybase + c, NY - 1);
                __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(Yc + (size_t)y * 128 + 16 * q),
                                                 (__attribute__((address_space(3))) void*)&sm.y[(8 * w + i) * 1024], 16,
                                                 0, 0);
            }
            {
                const int y = min(ybase + 64 * w + lane, NY - 1);
                __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(ty + y),
                                                 (__attribute__((address_space(3))) void*)&sm.t[64 * w], 4, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(0);   // the copies have landed in LDS
            __syncthreads();
            if (PIPE && ybase + kTpStage <= NY) {
                v16i accP[NT], accQ[NT];
                v4i yf[4], tvP[4], tvQ[4];
                fetch(0, yf, tvP);
                mma(yf, accP);
#pragma unroll
                for (int t = 0; t < kTpStage / 32; t += 2) {
                    fetch(t + 1, yf, tvQ);
                    mma(yf, accQ);
                    insert_full(accP, tvP);
                    if (t + 2 < kTpStage / 32) {
                        fetch(t + 2, yf, tvP);
                        mma(yf, accP);
                    }
                    insert_full(accQ, tvQ);
                }
                continue;
            }
            const int nt = min(kTpStage / 32, (NY - ybase + 31) / 32);
            for (int t = 0; t < nt; ++t) {
                v4i yf[4], tv[4];
                fetch(t, yf, tv);
                v16i acc[NT];
                mma(yf, acc);
                if (ybase + 32 * t + 32 <= NY) {   // uniform over the block
                    insert_full(acc, tv);
                } else {
#pragma unroll
                    for (int n = 0; n < NT; ++n)
#pragma unroll
                        for (int r = 0; r < 16; ++r) {
                            const bool v = ybase + 32 * t + 8 * (r >> 2) + 4 * hi + (r & 3) < NY;
                            const uint32_t k = ((uint32_t)acc[n][r] << 10) + (uint32_t)tv[r >> 2][r & 3];
                            top2_insert(k1[n], k2[n], v ? k : 0xffffffffu);
                        }
                }
            }
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const uint32_t o1 = __shfl_xor(k1[n], 32), o2 = __shfl_xor(k2[n], 32);
            top2_merge(k1[n], k2[n], o1, o2);
            const int x = xw + 32 * n + l32;
            if (hi == 0 && x < NX) {
                part[((size_t)ch * NX + x) * 2] = k1[n];
                part[((size_t)ch * NX + x) * 2 + 1] = k2[n];
            }
        }
    }
}

// both passes in one launch: blockIdx.z = pass; blocks past a pass's extent leave
struct TpPass {
    const uint8_t* X;
    const uint8_t* Yc;
    const uint32_t* ty;
    uint32_t* part;
    int NX, NY;
};

template <int NT, bool PIPE>
__global__ __launch_bounds__(256) void k_match_side(TpPass p0, TpPass p1, int cpw) {
    __shared__ TpSmem sm;
    const TpPass P = blockIdx.z ? p1 : p0;
    if ((int)blockIdx.x * (kTpWaves * 32 * NT) >= P.NX || (int)blockIdx.y * cpw * 512 >= P.NY) return;
    match_side_tile<NT, PIPE>(P.X, P.NX, P.Yc, P.NY, P.ty, P.part, blockIdx.x, blockIdx.y, cpw, sm);
}

// tile shape (x per wave, pipelining) and 512-y chunks per workgroup of the
// two-pass kernel; results do not depend on them (set by the autotuner)
static std::atomic<int> g_tp_shape{0}, g_tp_cpw{4};

static void tp_launch(const TpPass& p0, const TpPass& p1, hipStream_t s) {
    const int shape = g_tp_shape.load(std::memory_order_relaxed), cpw = g_tp_cpw.load(std::memory_order_relaxed);
    static const int nt_of[5] = {2, 2, 1, 3, 4};
    const int xblk = kTpWaves * 32 * nt_of[shape];
    auto groups = [&](int ny) { return ((ny + 511) / 512 + cpw - 1) / cpw; };
    const dim3 grid((std::max(p0.NX, p1.NX) + xblk - 1) / xblk, std::max(groups(p0.NY), groups(p1.NY)), 2);
    switch (shape) {
    case 0: hipLaunchKernelGGL((k_match_side<2, true>), grid, dim3(256), 0, s, p0, p1, cpw); break;
    case 1: hipLaunchKernelGGL((k_match_side<2, false>), grid, dim3(256), 0, s, p0, p1, cpw); break;
    case 2: hipLaunchKernelGGL((k_match_side<1, false>), grid, dim3(256), 0, s, p0, p1, cpw); break;
    case 3: hipLaunchKernelGGL((k_match_side<3, false>), grid, dim3(256), 0, s, p0, p1, cpw); break;
    default: hipLaunchKernelGGL((k_match_side<4, false>), grid, dim3(256), 0, s, p0, p1, cpw); break;
    }
}

int match_set_tile(int shape, int cpw) {
    if (shape < 0 || shape > 4 || cpw < 1 || cpw > 64) {
        set_error("match_set_tile: shape must be in [0, 4] and chunks per workgroup in [1, 64]");
        return MSIFT_EINVAL;
    }
    g_tp_shape.store(shape);
    g_tp_cpw.store(cpw);
    return MSIFT_OK;
}

// ---------------------------------------------------------------- float ----
// Float descriptor path: inputs rounded to fp16, a.b on v_mfma_f32_32x32x16_f16
// with fp32 accumulation, norms in fp32 from the rounded values.  For unit
// descriptors d = |a|^2 + |b|^2 - 2 a.b lies in [0, 4]; it is clamped there and
// quantised to q = round(d 2^21) < 2^23, so the exact path's 32-bit top-2 keys
// (q << 9 | j), reduce and mutual kernels apply unchanged (the quantum 4.8e-7
// is far below the fp16 rounding of the product).
constexpr float kF16Q = 2097152.0f;   // 2^21

typedef float v16f __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void k_to_f16(const float* __restrict__ in, size_t n, _Float16* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        out[i] = (_Float16)in[i];
}

// squared norms of fp16 rows (128 halves), fp32, lane pairs + xor tree
__global__ __launch_bounds__(256) void k_norms_f16(const _Float16* __restrict__ a, int n, float* __restrict__ out) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const float v0 = (float)a[(size_t)i * 128 + 2 * lane], v1 = (float)a[(size_t)i * 128 + 2 * lane + 1];
    float s = v0 * v0 + v1 * v1;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s = s + __shfl_xor(s, m);
    if (lane == 0) out[i] = s;
}

__device__ __forceinline__ uint32_t f16_q(float d) {
    d = d < 0.0f ? 0.0f : (d > 4.0f ? 4.0f : d);
    uint32_t q = (uint32_t)(d * kF16Q + 0.5f);
    return q > 0x7fffffu ? 0x7fffffu : q;
}

// Similarity keys of the reference's fp16 path (core.py:2668-2677): s is
// rounded to binary16 (the GEMM's output type) and ordered largest first,
// ties by index; -0 counts as +0.  key = 0xffff - (order-preserving map of
// the binary16 bits), so a smaller key is a larger similarity.  With h the
// bits after a binary16 add of +0 (which turns -0 into +0 and keeps every
// other value), that is h for negative h and 0x7fff - h = h ^ 0x7fff for
// the others (equal to the map for all 65536 patterns).
__device__ __forceinline__ uint32_t sim16_key(float s) {
    const _Float16 f = (_Float16)s + (_Float16)0.0f;
    const uint32_t h = (uint32_t)__builtin_bit_cast(uint16_t, f);
    const uint32_t m = (uint32_t)((int32_t)(h << 16) >> 31);   // all ones for negative h
    return h ^ (~m & 0x7fffu);
}

__device__ __forceinline__ float sim16_of_key(uint32_t key) {
    const uint32_t u = 0xffffu - key;
    const uint32_t h = (u & 0x8000u) ? (u & 0x7fffu) : (~u & 0xffffu);
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
}

// KS = 16-element MFMA steps per row (rows are 16 KS halves); SIM selects the
// epilogue: false = quantised squared distance from the norms (128-D rows),
// true = the binary16 similarity key above.
template <int KS, bool SIM>
__global__ __launch_bounds__(512) void k_match_tiles_f16(const _Float16* __restrict__ A, int N,
                                                          const _Float16* __restrict__ B, int M,
                                                          const float* __restrict__ na, const float* __restrict__ nb,
                                                          uint32_t* __restrict__ rowpart, uint32_t* __restrict__ colpart) {
    constexpr int D = 16 * KS;
    __shared__ uint32_t colk[8][64][2];
    __shared__ float s_na[kRowBlk];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const int rb = blockIdx.x, cc = blockIdx.y;
    const int row0 = rb * kRowBlk + w * 32;
    if (!SIM && threadIdx.x < kRowBlk) {
        int i = rb * kRowBlk + threadIdx.x;
        s_na[threadIdx.x] = i < N ? na[i] : 0.0f;
    }
    // A fragment: lane holds row (row0 + l32), halves [16 ks + 8 hi, +8) for ks < KS
    v8h af[KS];
    {
        int i = row0 + l32;
        int ic = i < N ? i : N - 1;
        const v8h* src = reinterpret_cast<const v8h*>(A + (size_t)ic * D);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) af[ks] = src[2 * ks + hi];
    }
    uint32_t rk1[16], rk2[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) { rk1[r] = 0xffffffffu; rk2[r] = 0xffffffffu; }
    __syncthreads();
    for (int ct = 0; ct < kColChunk / 64; ++ct) {
        const int col0 = cc * kColChunk + ct * 64;
        if (col0 >= M) break;
        // rows and columns all in range: no validity selects (uniform over the block)
        const bool full = (rb + 1) * kRowBlk <= N && col0 + 64 <= M;
        v16f acc[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            int j = col0 + 32 * b + l32;
            int jc = j < M ? j : M - 1;
            const v8h* src = reinterpret_cast<const v8h*>(B + (size_t)jc * D);
            acc[b] = (v16f)0.0f;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[ks], src[2 * ks + hi], acc[b], 0, 0, 0);
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = col0 + 32 * b + l32;
            const bool jv = j < M;
            const float nbj = (!SIM && jv) ? nb[j] : 0.0f;
            uint32_t lj = (uint32_t)(ct * 64 + 32 * b + l32);
            asm("" : "+v"(lj));   // one register: keeps (q << 9) | lj a single shift-or
            uint32_t c1 = 0xffffffffu, c2 = 0xffffffffu;
            if (full) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int il = 32 * w + 8 * (r >> 2) + 4 * hi + (r & 3);
                    const uint32_t q = SIM ? sim16_key(acc[b][r]) : f16_q((s_na[il] + nbj) - 2.0f * acc[b][r]);
                    top2_insert(c1, c2, (q << 8) | (uint32_t)il);
                    top2_insert(rk1[r], rk2[r], (q << 9) | lj);
                }
            } else {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int il = 32 * w + 8 * (r >> 2) + 4 * hi + (r & 3);
                    const bool v = jv && (rb * kRowBlk + il < N);
                    const uint32_t q = SIM ? sim16_key(acc[b][r]) : f16_q((s_na[il] + nbj) - 2.0f * acc[b][r]);
                    const uint32_t kc = v ? ((q << 8) | (uint32_t)il) : 0xffffffffu;
                    top2_insert(c1, c2, kc);
                    const uint32_t kr = v ? ((q << 9) | lj) : 0xffffffffu;
                    top2_insert(rk1[r], rk2[r], kr);
                }
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

// distance slots of the top-2 triples: quantised q -> float bits q / 2^21
// ("none", 0x7fffffff, -> +inf)
__global__ __launch_bounds__(256) void k_q_to_float(int32_t* __restrict__ t, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 1; k <= 2; ++k) {
        const int32_t q = t[3 * i + k];
        const float d = q == 0x7fffffff ? __builtin_huge_valf() : (float)q * (1.0f / kF16Q);
        t[3 * i + k] = __builtin_bit_cast(int32_t, d);
    }
}

// similarity path: key slots -> float bits of d = max(2 - 2 s, 0)
// (core.py:2674-2675); "none" stays 0x7fffffff
__global__ __launch_bounds__(256) void k_sim_to_dist(int32_t* __restrict__ t, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 1; k <= 2; ++k) {
        const int32_t key = t[3 * i + k];
        if (key == 0x7fffffff) continue;
        float d = 2.0f - 2.0f * sim16_of_key((uint32_t)key);
        d = d > 0.0f ? d : 0.0f;
        t[3 * i + k] = __builtin_bit_cast(int32_t, d);
    }
}

// binary16 copies of n rows of D floats, zero-padded to Dp halves
__global__ __launch_bounds__(256) void k_to_f16_pad(const float* __restrict__ in, int n, int D, int Dp,
                                                    _Float16* __restrict__ out) {
    const size_t tot = (size_t)n * Dp;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (size_t)gridDim.x * 256) {
        const size_t i = e / Dp;
        const int k = (int)(e - i * Dp);
        out[e] = k < D ? (_Float16)in[i * D + k] : (_Float16)0.0f;
    }
}

// ------------------------------------------------------------- batched ----
// Pairs (a_p, b_p) of descriptor sets packed back to back; one launch covers
// every tile of every pair, a block finds its pair in the table by a ballot
// over the pairs' first tile.

template <bool TWOPASS = false>
__device__ __forceinline__ int pair_of_tile(const MPair* __restrict__ pairs, int np, int t) {
    int last = 0;
    for (int b0 = 0; b0 < np; b0 += 64) {
        const int k = b0 + (int)(threadIdx.x & 63);
        const bool le = k < np && (TWOPASS ? pairs[k].tp0 : pairs[k].tile0) <= t;
        const unsigned long long m = __ballot(le);
        if (m) last = b0 + 63 - __clzll(m);
        if (m != ~0ull) break;
    }
    return last;
}

template <int METHOD>
__global__ __launch_bounds__(512) void k_match_tiles_b(const MPair* __restrict__ pairs, int np,
                                                        const uint8_t* __restrict__ A, const uint8_t* __restrict__ B,
                                                        const int* __restrict__ na, const int* __restrict__ nb,
                                                        char* __restrict__ ws) {
    __shared__ MatchSmem sm;
    __shared__ int s_p;
    if (threadIdx.x < 64) {
        const int p = pair_of_tile(pairs, np, (int)blockIdx.x);
        if (threadIdx.x == 0) s_p = p;
    }
    __syncthreads();
    const MPair P = pairs[s_p];
    const int t = (int)blockIdx.x - P.tile0;
    match_tile<METHOD>(A + (size_t)P.a_off * 128, P.N, B + (size_t)P.b_off * 128, P.M, na + P.a_off, nb + P.b_off,
                       (uint32_t*)(ws + P.o_rowpart), (uint32_t*)(ws + P.o_colpart), t % P.nrb, t / P.nrb, sm);
}

// two-pass tiles of every pair (pass 1 tiles, then pass 2), fixed shape: 64
// x per wave with the next tile's MFMAs issued ahead, one 512-y chunk a workgroup
constexpr int kTpBatchXBlk = kTpWaves * 64;
__global__ __launch_bounds__(256) void k_match_side_b(const MPair* __restrict__ pairs, int np,
                                                       const uint8_t* __restrict__ A, const uint8_t* __restrict__ B,
                                                       const uint8_t* __restrict__ Ac, const uint8_t* __restrict__ Bc,
                                                       const uint32_t* __restrict__ ta, const uint32_t* __restrict__ tb,
                                                       char* __restrict__ ws) {
    __shared__ TpSmem sm;
    __shared__ int s_p;
    if (threadIdx.x < 64) {
        const int p = pair_of_tile<true>(pairs, np, (int)blockIdx.x);
        if (threadIdx.x == 0) s_p = p;
    }
    __syncthreads();
    const MPair P = pairs[s_p];
    int t = (int)blockIdx.x - P.tp0;
    if (t < P.xb1 * P.cg1)
        match_side_tile<2, true>(A + (size_t)P.a_off * 128, P.N, Bc + (size_t)P.b_off * 128, P.M, tb + P.b_off,
                                 (uint32_t*)(ws + P.o_rowpart), t % P.xb1, t / P.xb1, 1, sm);
    else {
        t -= P.xb1 * P.cg1;
        match_side_tile<2, true>(B + (size_t)P.b_off * 128, P.M, Ac + (size_t)P.a_off * 128, P.N, ta + P.a_off,
                                 (uint32_t*)(ws + P.o_colpart), t % P.xb2, t / P.xb2, 1, sm);
    }
}

// merge per-chunk partials into (best index, d1, d2); key 0xffffffff = none.
// Partials p0, p0 + pstep, ... of entry i fold into the 64-bit top-2 (k1, k2)
// of (d << 32) | global index: unique keys, so any split of the partials gives
// the same pair.  adj (the shift method): the keys hold d - adj[i] + 2^21,
// adj being |a'|^2 for rows and nbx for columns; the distances written are d.
__device__ __forceinline__ void reduce_top2_scan(const uint32_t* __restrict__ part, int n, int p0, int pstep,
                                                 int nparts, int shift, int span, int i, uint64_t& k1,
                                                 uint64_t& k2) {
    // 8-byte loads, unrolled so several are in flight
    const uint2* p2 = reinterpret_cast<const uint2*>(part);
#pragma unroll 8
    for (int p = p0; p < nparts; p += pstep) {
        const uint2 kk = p2[(size_t)p * n + i];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t k = q ? kk.y : kk.x;
            if (k == 0xffffffffu) continue;
            uint64_t d = k >> shift;
            uint64_t idx = (uint64_t)p * span + (k & ((1u << shift) - 1));
            top2_insert64(k1, k2, (d << 32) | idx);
        }
    }
}

__device__ __forceinline__ void reduce_top2_write(int32_t* __restrict__ out, int i, uint64_t k1, uint64_t k2,
                                                  const int* __restrict__ adj) {
    const int32_t da = adj ? adj[i] - (int32_t)kKeyOff : 0;
    out[3 * i] = k1 == ~0ull ? -1 : (int32_t)(k1 & 0xffffffffu);
    out[3 * i + 1] = k1 == ~0ull ? 0x7fffffff : (int32_t)(k1 >> 32) + da;
    out[3 * i + 2] = k2 == ~0ull ? 0x7fffffff : (int32_t)(k2 >> 32) + da;
}

__device__ __forceinline__ void reduce_top2_row(const uint32_t* __restrict__ part, int n, int nparts, int shift,
                                                int span, int32_t* __restrict__ out, int i,
                                                const int* __restrict__ adj) {
    uint64_t k1 = ~0ull, k2 = ~0ull;
    reduce_top2_scan(part, n, 0, 1, nparts, shift, span, i, k1, k2);
    reduce_top2_write(out, i, k1, k2, adj);
}

// 64 entries per block; wave w scans partials w, w + 4, ... and wave 0 merges
// the four pairs (n / 64 blocks keep enough loads in flight to stream the
// partials at HBM rate)
__global__ __launch_bounds__(256) void k_reduce_top2(const uint32_t* __restrict__ part, int n, int nparts,
                                                      int shift, int span, int32_t* __restrict__ out,
                                                      const int* __restrict__ adj) {
    __shared__ uint64_t s_k[3][64][2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    uint64_t k1 = ~0ull, k2 = ~0ull;
    if (i < n) reduce_top2_scan(part, n, w, 4, nparts, shift, span, i, k1, k2);
    if (w > 0) {
        s_k[w - 1][lane][0] = k1;
        s_k[w - 1][lane][1] = k2;
    }
    __syncthreads();
    if (w == 0 && i < n) {
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            top2_insert64(k1, k2, s_k[o][lane][0]);
            top2_insert64(k1, k2, s_k[o][lane][1]);
        }
        reduce_top2_write(out, i, k1, k2, adj);
    }
}

// the two-pass method's two reductions in one launch (blockIdx.y = side),
// both with 9-bit local indices over 512-y chunks
__global__ __launch_bounds__(256) void k_reduce_top2_2(const uint32_t* __restrict__ part0, int n0, int nparts0,
                                                        int32_t* __restrict__ out0, const int* __restrict__ adj0,
                                                        const uint32_t* __restrict__ part1, int n1, int nparts1,
                                                        int32_t* __restrict__ out1, const int* __restrict__ adj1) {
    __shared__ uint64_t s_k[3][64][2];
    const bool side = blockIdx.y != 0;
    const uint32_t* part = side ? part1 : part0;
    const int n = side ? n1 : n0, nparts = side ? nparts1 : nparts0;
    if ((int)blockIdx.x * 64 >= n) return;   // whole block
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    uint64_t k1 = ~0ull, k2 = ~0ull;
    if (i < n) reduce_top2_scan(part, n, w, 4, nparts, 9, 512, i, k1, k2);
    if (w > 0) {
        s_k[w - 1][lane][0] = k1;
        s_k[w - 1][lane][1] = k2;
    }
    __syncthreads();
    if (w == 0 && i < n) {
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            top2_insert64(k1, k2, s_k[o][lane][0]);
            top2_insert64(k1, k2, s_k[o][lane][1]);
        }
        reduce_top2_write(side ? out1 : out0, i, k1, k2, side ? adj1 : adj0);
    }
}

// rows (cols == 0: top-2 over b, into fwd) or columns (top-2 over a, into
// bwd) of pair blockIdx.y; no partials (an empty side) gives "none"
__global__ __launch_bounds__(256) void k_reduce_top2_b(const MPair* __restrict__ pairs, const char* __restrict__ ws,
                                                        int cols, int32_t* __restrict__ fwd, int32_t* __restrict__ bwd,
                                                        const int* __restrict__ na, const int* __restrict__ nb,
                                                        int method) {
    const MPair P = pairs[blockIdx.y];
    const int n = cols ? P.M : P.N;
    // method 2: both sides hold 9-bit indices over 512-chunks (rows: of b, columns: of a)
    const bool c8 = cols && method != 2;
    const int nparts = (P.N > 0 && P.M > 0) ? (cols ? (method == 2 ? (P.N + 511) / 512 : P.nrb) : P.ncc) : 0;
    const uint32_t* part = (const uint32_t*)(ws + (cols ? P.o_colpart : P.o_rowpart));
    int32_t* out = cols ? bwd + 3 * (size_t)P.b_off : fwd + 3 * (size_t)P.a_off;
    const int* adj = method != 0 ? (cols ? nb + P.b_off : na + P.a_off) : nullptr;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        reduce_top2_row(part, n, nparts, c8 ? 8 : 9, c8 ? kRowBlk : kColChunk, out, i, adj);
}

__global__ void k_fill_top2_none(int32_t* __restrict__ out, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[3 * i] = -1; out[3 * i + 1] = 0x7fffffff; out[3 * i + 2] = 0x7fffffff;
}

// mutual ratio test (core.py:2935-2939), then raster-order compaction
// fmode 0: integer distances, d1 < r2 d2 in double; fmode 1: float bits of the
// similarity path, d2 > 1e-10 and d1 < (float)r2 d2 in float (core.py:2676)
__device__ __forceinline__ bool ratio_ok(const int32_t* t, double r2, int fmode) {
    if (!fmode) return (double)t[1] < r2 * (double)t[2];
    if (t[2] == 0x7fffffff) return false;
    const float d1 = __builtin_bit_cast(float, t[1]), d2 = __builtin_bit_cast(float, t[2]);
    return d2 > 1e-10f && d1 < (float)r2 * d2;
}

__device__ __forceinline__ bool mutual_ok(const int32_t* fwd, const int32_t* bwd, int N, int M, double r2, int i,
                                          int fmode = 0) {
    if (N < 2 || M < 2) return false;
    int j = fwd[3 * i];
    if (j < 0) return false;
    if (!ratio_ok(fwd + 3 * i, r2, fmode)) return false;
    if (bwd[3 * j] != i) return false;
    return ratio_ok(bwd + 3 * j, r2, fmode);
}

// block b (256 rows): number of mutual matches; whole block
__device__ __forceinline__ int mutual_block_count(const int32_t* __restrict__ fwd, const int32_t* __restrict__ bwd,
                                                  int N, int M, double r2, int b, int* wc, int fmode = 0) {
    const int i = b * 256 + (int)threadIdx.x;
    const bool ok = i < N && mutual_ok(fwd, bwd, N, M, r2, i, fmode);
    const unsigned long long m = __ballot(ok);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    const int c = (wc[0] + wc[1]) + (wc[2] + wc[3]);
    __syncthreads();
    return c;
}

// block b: write its mutual pairs (i, fwd j) from position `pos0` on; whole block
__device__ __forceinline__ void mutual_block_write(const int32_t* __restrict__ fwd, const int32_t* __restrict__ bwd,
                                                   int N, int M, double r2, int b, int pos0, int32_t* __restrict__ pairs,
                                                   int* wc, int fmode = 0) {
    const int i = b * 256 + (int)threadIdx.x;
    const bool ok = i < N && mutual_ok(fwd, bwd, N, M, r2, i, fmode);
    const unsigned long long m = __ballot(ok);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wc[wave] = __popcll(m);
    __syncthreads();
    if (ok) {
        int pre = 0;
        for (int w = 0; w < wave; ++w) pre += wc[w];
        const int pos = pos0 + pre + __popcll(m & lanemask_lt());
        pairs[2 * pos] = i;
        pairs[2 * pos + 1] = fwd[3 * i];
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_mutual_count(const int32_t* __restrict__ fwd, const int32_t* __restrict__ bwd,
                                                       int N, int M, double r2, int* __restrict__ counts,
                                                       int fmode) {
    __shared__ int wc[4];
    const int c = mutual_block_count(fwd, bwd, N, M, r2, blockIdx.x, wc, fmode);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_mutual_write(const int32_t* __restrict__ fwd, const int32_t* __restrict__ bwd,
                                                       int N, int M, double r2, const int* __restrict__ offs,
                                                       int32_t* __restrict__ pairs, int fmode) {
    __shared__ int wc[4];
    mutual_block_write(fwd, bwd, N, M, r2, blockIdx.x, offs[blockIdx.x], pairs, wc, fmode);
}

__global__ __launch_bounds__(256) void k_mutual_count_b(const MPair* __restrict__ pairs, const int32_t* __restrict__ fwd,
                                                         const int32_t* __restrict__ bwd, double r2,
                                                         int* __restrict__ counts) {
    __shared__ int wc[4];
    const MPair P = pairs[blockIdx.y];
    for (int b = blockIdx.x; b < P.nblk; b += gridDim.x) {
        const int c = mutual_block_count(fwd + 3 * (size_t)P.a_off, bwd + 3 * (size_t)P.b_off, P.N, P.M, r2, b, wc);
        if (threadIdx.x == 0) counts[P.blk0 + b] = c;
    }
}

// pair p's mutual pairs go to out[2 * a_off ...] (at most N of them)
__global__ __launch_bounds__(256) void k_mutual_write_b(const MPair* __restrict__ pairs, const int32_t* __restrict__ fwd,
                                                         const int32_t* __restrict__ bwd, double r2,
                                                         const int* __restrict__ offs, int32_t* __restrict__ out) {
    __shared__ int wc[4];
    const MPair P = pairs[blockIdx.y];
    for (int b = blockIdx.x; b < P.nblk; b += gridDim.x)
        mutual_block_write(fwd + 3 * (size_t)P.a_off, bwd + 3 * (size_t)P.b_off, P.N, P.M, r2, b,
                           offs[P.blk0 + b] - offs[P.blk0], out + 2 * (size_t)P.a_off, wc);
}

__global__ __launch_bounds__(64) void k_npairs_b(const MPair* __restrict__ pairs, int np, const int* __restrict__ offs,
                                                  int* __restrict__ npairs) {
    for (int p = threadIdx.x; p < np; p += 64) npairs[p] = offs[pairs[p].blk0 + pairs[p].nblk] - offs[pairs[p].blk0];
}

// matched point coordinates of every pair: p1[off + k] = (x, y) of a's row,
// p2[off + k] = (x, y) of b's row, for k < npairs[p]; off = a_off (room for N)
__global__ __launch_bounds__(256) void k_gather_points_b(const MPair* __restrict__ pairs,
                                                          const float* __restrict__ kpa, int sa,
                                                          const float* __restrict__ kpb, int sb,
                                                          const int32_t* __restrict__ mp, const int* __restrict__ npairs,
                                                          double* __restrict__ p1, double* __restrict__ p2) {
    const MPair P = pairs[blockIdx.y];
    const int n = npairs[blockIdx.y];
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        const int i = mp[2 * ((size_t)P.a_off + k)], j = mp[2 * ((size_t)P.a_off + k) + 1];
        const float* a = kpa + (size_t)(P.a_off + i) * sa;
        const float* b = kpb + (size_t)(P.b_off + j) * sb;
        p1[2 * ((size_t)P.a_off + k)] = (double)a[0];
        p1[2 * ((size_t)P.a_off + k) + 1] = (double)a[1];
        p2[2 * ((size_t)P.a_off + k)] = (double)b[0];
        p2[2 * ((size_t)P.a_off + k) + 1] = (double)b[1];
    }
}

// ---------------------------------------------------------------- host ----
struct MatchPlan {
    int nrb_a, ncc_a;   // row blocks of a, column chunks of b
    size_t o_na, o_nb, o_na2, o_nb2, o_ac, o_bc, o_rowpart, o_colpart, o_cnt, o_off, o_ha, o_hb, total;
};

static MatchPlan match_plan(int N, int M) {
    MatchPlan P;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    P.nrb_a = (N + kRowBlk - 1) / kRowBlk;
    P.ncc_a = (M + kColChunk - 1) / kColChunk;
    P.o_na = take((size_t)(N > 0 ? N : 1) * 4);
    P.o_nb = take((size_t)(M > 0 ? M : 1) * 4);
    P.o_na2 = take((size_t)(N > 0 ? N : 1) * 4);   // method 2's second pass
    P.o_nb2 = take((size_t)(M > 0 ? M : 1) * 4);
    P.o_ac = take((size_t)(N > 0 ? N : 1) * 128);   // method 2's y' copies
    P.o_bc = take((size_t)(M > 0 ? M : 1) * 128);
    P.o_rowpart = take((size_t)P.ncc_a * (N > 0 ? N : 1) * 8);
    P.o_colpart = take((size_t)P.nrb_a * (M > 0 ? M : 1) * 8);
    int nblk = (N + 255) / 256;
    P.o_cnt = take((size_t)(nblk + 1) * 4);
    P.o_off = take((size_t)(nblk + 1) * 4);
    // fp16 copies for the float path (the integer path leaves them unused)
    P.o_ha = take((size_t)(N > 0 ? N : 1) * 256);
    P.o_hb = take((size_t)(M > 0 ? M : 1) * 256);
    P.total = off;
    return P;
}

size_t match_workspace(int N, int M) { return match_plan(N, M).total; }

// The integer path's share of the workspace: everything before the fp16
// copies, which only the float path uses.
static size_t match_int_bytes(const MatchPlan& P) { return P.o_ha; }

static int match_u8_run(const uint8_t* a, int N, const uint8_t* b, int M, double ratio, int method, char* ws,
                        const MatchPlan& P, int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s);

int match_u8(const uint8_t* a, int N, const uint8_t* b, int M, double ratio, int method, void* ws_, size_t ws_bytes,
             int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s) {
    MatchPlan P = match_plan(N, M);
    if (ws_bytes < P.total) { set_error("match_u8: workspace too small"); return MSIFT_EWORKSPACE; }
    return match_u8_run(a, N, b, M, ratio, method, (char*)ws_, P, fwd, bwd, pairs, npairs, s);
}

// ------------------------------------------------ allocation-free variant ----
// Scratch of the integer matcher for N, M <= kNoAllocMaxRows, in a file-scope
// device array: every GPU the library runs on holds its own copy until the
// process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).
constexpr int kNoAllocMaxRows = 16384;
constexpr size_t rup256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNoAllocScratchBytes =
    rup256((size_t)kNoAllocMaxRows * 4) * 4 + rup256((size_t)kNoAllocMaxRows * 128) * 2 +
    rup256((size_t)((kNoAllocMaxRows + kColChunk - 1) / kColChunk) * kNoAllocMaxRows * 8) +
    rup256((size_t)((kNoAllocMaxRows + kRowBlk - 1) / kRowBlk) * kNoAllocMaxRows * 8) +
    rup256((size_t)((kNoAllocMaxRows + 255) / 256 + 1) * 4) * 2;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device, looked up once per device.
static int scratch_on_current_device(char** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("match_u8_noalloc: no current HIP device");
        return MSIFT_EHIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("match_u8_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return MSIFT_EHIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = (char*)g_scratch_ptr[dev];
    return MSIFT_OK;
}

int match_u8_noalloc(const uint8_t* a, int N, const uint8_t* b, int M, double ratio, int method, int32_t* fwd,
                     int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s) {
    if (N > kNoAllocMaxRows || M > kNoAllocMaxRows) {
        set_error("match_u8_noalloc: N and M must be <= " + std::to_string(kNoAllocMaxRows) + " (got " +
                  std::to_string(N) + " x " + std::to_string(M) + "); use msift_match_u8 with a workspace");
        return MSIFT_EINVAL;
    }
    const MatchPlan P = match_plan(N, M);
    if (match_int_bytes(P) > kNoAllocScratchBytes) {   // cannot happen within the row limit
        set_error("match_u8_noalloc: scratch too small");
        return MSIFT_EWORKSPACE;
    }
    char* ws = nullptr;
    const int rc = scratch_on_current_device(&ws);
    if (rc != MSIFT_OK) return rc;
    return match_u8_run(a, N, b, M, ratio, method, ws, P, fwd, bwd, pairs, npairs, s);
}

static int match_u8_run(const uint8_t* a, int N, const uint8_t* b, int M, double ratio, int method, char* ws,
                        const MatchPlan& P, int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s) {
    if ((N > 0 && (!a || !fwd)) || (M > 0 && (!b || !bwd)) || (!pairs && N > 0 && M > 0)) {
        set_error("match_u8: null buffer");
        return MSIFT_EINVAL;
    }
    if (((uintptr_t)a & 15) || ((uintptr_t)b & 15)) { set_error("match_u8: descriptors must be 16-byte aligned"); return MSIFT_EINVAL; }
    int* na = (int*)(ws + P.o_na);
    int* nb = (int*)(ws + P.o_nb);
    uint32_t* rowpart = (uint32_t*)(ws + P.o_rowpart);
    uint32_t* colpart = (uint32_t*)(ws + P.o_colpart);
    if (N > 0 && M > 0) {
        if (method == 2) {
            // na: |a'|^2, nb2: |b'|^2 (the reductions' adj); nb / na2: b's / a's key terms
            int* na2 = (int*)(ws + P.o_na2);
            int* nb2 = (int*)(ws + P.o_nb2);
            uint8_t* ac = (uint8_t*)(ws + P.o_ac);
            uint8_t* bc = (uint8_t*)(ws + P.o_bc);
            const int cha = (M + 511) / 512, chb = (N + 511) / 512;
            hipLaunchKernelGGL(k_tp_prep, dim3((N + M + 31) / 32), dim3(256), 0, s, a, N, b, M, ac, bc, na, nb2,
                               (uint32_t*)na2, (uint32_t*)nb, (const MPair*)nullptr, 0);
            tp_launch(TpPass{a, bc, (const uint32_t*)nb, rowpart, N, M}, TpPass{b, ac, (const uint32_t*)na2, colpart, M, N},
                      s);
            hipLaunchKernelGGL(k_reduce_top2_2, dim3((std::max(N, M) + 63) / 64, 2), dim3(256), 0, s, rowpart, N, cha,
                               fwd, (const int*)na, colpart, M, chb, bwd, (const int*)nb2);
        } else {
            hipLaunchKernelGGL(k_norms, dim3((N + 3) / 4), dim3(256), 0, s, a, N, method, na);
            hipLaunchKernelGGL(k_norms, dim3((M + 3) / 4), dim3(256), 0, s, b, M, method == 1 ? 2 : method, nb);
            dim3 grid(P.nrb_a, P.ncc_a);
            if (method == 0)
                hipLaunchKernelGGL(k_match_tiles<0>, grid, dim3(512), 0, s, a, N, b, M, na, nb, rowpart, colpart);
            else
                hipLaunchKernelGGL(k_match_tiles<1>, grid, dim3(512), 0, s, a, N, b, M, na, nb, rowpart, colpart);
            hipLaunchKernelGGL(k_reduce_top2, dim3((N + 63) / 64), dim3(256), 0, s, rowpart, N, P.ncc_a, 9, kColChunk,
                               fwd, method == 1 ? (const int*)na : nullptr);
            hipLaunchKernelGGL(k_reduce_top2, dim3((M + 63) / 64), dim3(256), 0, s, colpart, M, P.nrb_a, 8, kRowBlk,
                               bwd, method == 1 ? (const int*)nb : nullptr);
        }
    } else {
        if (N > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((N + 255) / 256), dim3(256), 0, s, fwd, N);
        if (M > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((M + 255) / 256), dim3(256), 0, s, bwd, M);
    }
    int nblk = (N + 255) / 256;
    double r2 = ratio * ratio;
    int* cnt = (int*)(ws + P.o_cnt);
    int* off = (int*)(ws + P.o_off);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_count, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, cnt, 0);
    launch_scan(cnt, off, nblk, npairs, nullptr, nullptr, 0, s);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_write, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, off, pairs, 0);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift

// plan of a batched match: pair table in device memory, workspace layout
struct msift_match_batch {
    int np;
    std::vector<msift::MPair> pairs;
    int sumN, sumM, tiles, tp_tiles, blks, max_rows, max_cols, max_blk;
    size_t o_na, o_nb, o_ac, o_bc, o_ta, o_tb, o_cnt, o_off, o_stmp, total;
    msift::MPair* d_pairs;
};

namespace msift {

static int match_batch_create(int np, const int* N, const int* M, msift_match_batch** out) {
    if (np < 1 || np > 65535 || !N || !M || !out) { set_error("match batch: bad arguments"); return MSIFT_EINVAL; }
    msift_match_batch* B = new (std::nothrow) msift_match_batch();
    if (!B) { set_error("out of host memory"); return MSIFT_EINVAL; }
    B->np = np;
    B->d_pairs = nullptr;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    long long sa = 0, sb = 0;
    int tiles = 0, tp_tiles = 0, blks = 0;
    B->max_rows = B->max_cols = B->max_blk = 0;
    for (int p = 0; p < np; ++p) {
        if (N[p] < 0 || M[p] < 0) { set_error("match batch: negative count"); delete B; return MSIFT_EINVAL; }
        MPair P;
        memset(&P, 0, sizeof(P));
        P.a_off = (int)sa; P.N = N[p]; P.b_off = (int)sb; P.M = M[p];
        const bool both = N[p] > 0 && M[p] > 0;
        P.nrb = both ? (N[p] + kRowBlk - 1) / kRowBlk : 0;
        P.ncc = both ? (M[p] + kColChunk - 1) / kColChunk : 0;
        P.tile0 = tiles;
        tiles += P.nrb * P.ncc;
        P.tp0 = tp_tiles;
        P.xb1 = both ? (N[p] + kTpBatchXBlk - 1) / kTpBatchXBlk : 0;
        P.cg1 = both ? (M[p] + 511) / 512 : 0;
        P.xb2 = both ? (M[p] + kTpBatchXBlk - 1) / kTpBatchXBlk : 0;
        P.cg2 = both ? (N[p] + 511) / 512 : 0;
        tp_tiles += P.xb1 * P.cg1 + P.xb2 * P.cg2;
        P.nblk = (N[p] + 255) / 256;
        P.blk0 = blks;
        blks += P.nblk;
        P.o_rowpart = take((size_t)P.ncc * (N[p] > 0 ? N[p] : 1) * 8);
        P.o_colpart = take((size_t)P.nrb * (M[p] > 0 ? M[p] : 1) * 8);
        B->pairs.push_back(P);
        sa += N[p]; sb += M[p];
        B->max_rows = std::max(B->max_rows, N[p]);
        B->max_cols = std::max(B->max_cols, M[p]);
        B->max_blk = std::max(B->max_blk, P.nblk);
    }
    if (sa > (1LL << 30) || sb > (1LL << 30)) { set_error("match batch: too many descriptors"); delete B; return MSIFT_EINVAL; }
    B->sumN = (int)sa; B->sumM = (int)sb; B->tiles = tiles; B->tp_tiles = tp_tiles; B->blks = blks;
    B->o_na = take((size_t)(sa > 0 ? sa : 1) * 4);
    B->o_nb = take((size_t)(sb > 0 ? sb : 1) * 4);
    B->o_ac = take((size_t)(sa > 0 ? sa : 1) * 128);   // method 2: y' copies and key terms
    B->o_bc = take((size_t)(sb > 0 ? sb : 1) * 128);
    B->o_ta = take((size_t)(sa > 0 ? sa : 1) * 4);
    B->o_tb = take((size_t)(sb > 0 ? sb : 1) * 4);
    B->o_cnt = take((size_t)(blks + 1) * 4);
    B->o_off = take((size_t)(blks + 1) * 4);
    B->o_stmp = take((size_t)scan_tmp_ints(blks) * 4);
    B->total = off;
    if (hipMalloc(&B->d_pairs, sizeof(MPair) * np) != hipSuccess ||
        hipMemcpy(B->d_pairs, B->pairs.data(), sizeof(MPair) * np, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("match batch: device table allocation failed");
        if (B->d_pairs) (void)hipFree(B->d_pairs);
        delete B;
        return MSIFT_EHIP;
    }
    *out = B;
    return MSIFT_OK;
}

static inline unsigned gcap_u(long long n, unsigned cap) { return n < 1 ? 1u : (n < cap ? (unsigned)n : cap); }

static int match_batch_run(msift_match_batch* B, const uint8_t* a, const uint8_t* b, double ratio, int method,
                           void* ws_, size_t ws_bytes, int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs,
                           hipStream_t s) {
    if (ws_bytes < B->total) { set_error("match_u8_batched: workspace too small"); return MSIFT_EWORKSPACE; }
    if ((B->sumN > 0 && (!a || !fwd || !pairs)) || (B->sumM > 0 && (!b || !bwd)) || !npairs || !ws_) {
        set_error("match_u8_batched: null buffer");
        return MSIFT_EINVAL;
    }
    if (((uintptr_t)a & 15) || ((uintptr_t)b & 15)) { set_error("match_u8_batched: descriptors must be 16-byte aligned"); return MSIFT_EINVAL; }
    if (method < 0 || method > 2) { set_error("match_u8_batched: method must be 0, 1 or 2"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    int* na = (int*)(ws + B->o_na);
    int* nb = (int*)(ws + B->o_nb);
    const int np = B->np;
    if (method == 2) {
        // na, nb: |a'|^2, |b'|^2 (the reductions' adj)
        uint8_t* ac = (uint8_t*)(ws + B->o_ac);
        uint8_t* bc = (uint8_t*)(ws + B->o_bc);
        uint32_t* ta = (uint32_t*)(ws + B->o_ta);
        uint32_t* tb = (uint32_t*)(ws + B->o_tb);
        if (B->sumN + B->sumM > 0)
            hipLaunchKernelGGL(k_tp_prep, dim3((B->sumN + B->sumM + 31) / 32), dim3(256), 0, s, a, B->sumN, b, B->sumM,
                               ac, bc, na, nb, ta, tb, (const MPair*)B->d_pairs, np);
        if (B->tp_tiles > 0)
            hipLaunchKernelGGL(k_match_side_b, dim3(B->tp_tiles), dim3(256), 0, s, (const MPair*)B->d_pairs, np, a, b,
                               (const uint8_t*)ac, (const uint8_t*)bc, (const uint32_t*)ta, (const uint32_t*)tb, ws);
    } else {
        if (B->sumN > 0) hipLaunchKernelGGL(k_norms, dim3((B->sumN + 3) / 4), dim3(256), 0, s, a, B->sumN, method, na);
        if (B->sumM > 0)
            hipLaunchKernelGGL(k_norms, dim3((B->sumM + 3) / 4), dim3(256), 0, s, b, B->sumM, method == 1 ? 2 : method, nb);
        if (B->tiles > 0) {
            if (method == 0)
                hipLaunchKernelGGL(k_match_tiles_b<0>, dim3(B->tiles), dim3(512), 0, s, (const MPair*)B->d_pairs, np, a,
                                   b, (const int*)na, (const int*)nb, ws);
            else
                hipLaunchKernelGGL(k_match_tiles_b<1>, dim3(B->tiles), dim3(512), 0, s, (const MPair*)B->d_pairs, np, a,
                                   b, (const int*)na, (const int*)nb, ws);
        }
    }
    if (B->max_rows > 0)
        hipLaunchKernelGGL(k_reduce_top2_b, dim3(gcap_u((B->max_rows + 255) / 256, 1024), np), dim3(256), 0, s,
                           (const MPair*)B->d_pairs, (const char*)ws, 0, fwd, bwd, (const int*)na, (const int*)nb, method);
    if (B->max_cols > 0)
        hipLaunchKernelGGL(k_reduce_top2_b, dim3(gcap_u((B->max_cols + 255) / 256, 1024), np), dim3(256), 0, s,
                           (const MPair*)B->d_pairs, (const char*)ws, 1, fwd, bwd, (const int*)na, (const int*)nb, method);
    const double r2 = ratio * ratio;
    int* cnt = (int*)(ws + B->o_cnt);
    int* offs = (int*)(ws + B->o_off);
    if (B->max_blk > 0)
        hipLaunchKernelGGL(k_mutual_count_b, dim3(gcap_u(B->max_blk, 1024), np), dim3(256), 0, s,
                           (const MPair*)B->d_pairs, (const int32_t*)fwd, (const int32_t*)bwd, r2, cnt);
    launch_scan(cnt, offs, B->blks, nullptr, nullptr, nullptr, 0, s, (int*)(ws + B->o_stmp));
    if (B->max_blk > 0)
        hipLaunchKernelGGL(k_mutual_write_b, dim3(gcap_u(B->max_blk, 1024), np), dim3(256), 0, s,
                           (const MPair*)B->d_pairs, (const int32_t*)fwd, (const int32_t*)bwd, r2, (const int*)offs,
                           pairs);
    hipLaunchKernelGGL(k_npairs_b, dim3(1), dim3(64), 0, s, (const MPair*)B->d_pairs, np, (const int*)offs, npairs);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

static int match_batch_gather(msift_match_batch* B, const float* kpa, int sa, const float* kpb, int sb,
                              const int32_t* mp, const int* npairs, double* p1, double* p2, hipStream_t s) {
    if (!kpa || !kpb || !mp || !npairs || !p1 || !p2 || sa < 2 || sb < 2) {
        set_error("match batch gather: bad arguments");
        return MSIFT_EINVAL;
    }
    if (B->max_rows > 0)
        hipLaunchKernelGGL(k_gather_points_b, dim3(gcap_u((B->max_rows + 255) / 256, 256), B->np), dim3(256), 0, s,
                           (const MPair*)B->d_pairs, kpa, sa, kpb, sb, mp, npairs, p1, p2);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

int match_f16(const float* a, int N, const float* b, int M, double ratio, void* ws_, size_t ws_bytes, int32_t* fwd,
              int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s) {
    MatchPlan P = match_plan(N, M);
    if (ws_bytes < P.total) { set_error("match_f16: workspace too small"); return MSIFT_EWORKSPACE; }
    if ((N > 0 && (!a || !fwd)) || (M > 0 && (!b || !bwd)) || (!pairs && N > 0 && M > 0)) {
        set_error("match_f16: null buffer");
        return MSIFT_EINVAL;
    }
    char* ws = (char*)ws_;
    float* na = (float*)(ws + P.o_na);
    float* nb = (float*)(ws + P.o_nb);
    _Float16* ha = (_Float16*)(ws + P.o_ha);
    _Float16* hb = (_Float16*)(ws + P.o_hb);
    uint32_t* rowpart = (uint32_t*)(ws + P.o_rowpart);
    uint32_t* colpart = (uint32_t*)(ws + P.o_colpart);
    if (N > 0 && M > 0) {
        const size_t ea = (size_t)N * 128, eb = (size_t)M * 128;
        hipLaunchKernelGGL(k_to_f16, dim3((unsigned)std::min<size_t>((ea + 255) / 256, 16384)), dim3(256), 0, s, a, ea, ha);
        hipLaunchKernelGGL(k_to_f16, dim3((unsigned)std::min<size_t>((eb + 255) / 256, 16384)), dim3(256), 0, s, b, eb, hb);
        hipLaunchKernelGGL(k_norms_f16, dim3((N + 3) / 4), dim3(256), 0, s, (const _Float16*)ha, N, na);
        hipLaunchKernelGGL(k_norms_f16, dim3((M + 3) / 4), dim3(256), 0, s, (const _Float16*)hb, M, nb);
        hipLaunchKernelGGL((k_match_tiles_f16<8, false>), dim3(P.nrb_a, P.ncc_a), dim3(512), 0, s, (const _Float16*)ha, N,
                           (const _Float16*)hb, M, (const float*)na, (const float*)nb, rowpart, colpart);
        hipLaunchKernelGGL(k_reduce_top2, dim3((N + 63) / 64), dim3(256), 0, s, rowpart, N, P.ncc_a, 9, kColChunk, fwd,
                           nullptr);
        hipLaunchKernelGGL(k_reduce_top2, dim3((M + 63) / 64), dim3(256), 0, s, colpart, M, P.nrb_a, 8, kRowBlk, bwd,
                           nullptr);
    } else {
        if (N > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((N + 255) / 256), dim3(256), 0, s, fwd, N);
        if (M > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((M + 255) / 256), dim3(256), 0, s, bwd, M);
    }
    // ratio and mutual tests on the quantised distances, then float bits
    const int nblk = (N + 255) / 256;
    const double r2 = ratio * ratio;
    int* cnt = (int*)(ws + P.o_cnt);
    int* off = (int*)(ws + P.o_off);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_count, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, cnt, 0);
    launch_scan(cnt, off, nblk, npairs, nullptr, nullptr, 0, s);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_write, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, off, pairs, 0);
    if (N > 0) hipLaunchKernelGGL(k_q_to_float, dim3((N + 255) / 256), dim3(256), 0, s, fwd, N);
    if (M > 0) hipLaunchKernelGGL(k_q_to_float, dim3((M + 255) / 256), dim3(256), 0, s, bwd, M);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift

using namespace msift;

extern "C" {

int msift_match_batch_create(int n_pairs, const int* N, const int* M, msift_match_batch** plan) {
    return match_batch_create(n_pairs, N, M, plan);
}

int msift_match_batch_destroy(msift_match_batch* plan) {
    if (!plan) return MSIFT_OK;
    if (plan->d_pairs) (void)hipFree(plan->d_pairs);
    delete plan;
    return MSIFT_OK;
}

int msift_match_batch_sizes(msift_match_batch* plan, size_t* ws_bytes, int* a_offsets, int* b_offsets) {
    if (!plan) { set_error("null plan"); return MSIFT_EINVAL; }
    if (ws_bytes) *ws_bytes = plan->total;
    for (int p = 0; p < plan->np; ++p) {
        if (a_offsets) a_offsets[p] = plan->pairs[p].a_off;
        if (b_offsets) b_offsets[p] = plan->pairs[p].b_off;
    }
    return MSIFT_OK;
}

int msift_match_u8_batched(msift_match_batch* plan, const uint8_t* d_a, const uint8_t* d_b, double ratio, int method,
                           void* d_ws, size_t ws_bytes, int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs,
                           int* d_npairs, void* stream) {
    if (!plan) { set_error("null plan"); return MSIFT_EINVAL; }
    return match_batch_run(plan, d_a, d_b, ratio, method, d_ws, ws_bytes, d_fwd, d_bwd, d_pairs, d_npairs,
                           (hipStream_t)stream);
}

int msift_match_batch_gather_points(msift_match_batch* plan, const float* d_kpa, int stride_a, const float* d_kpb,
                                    int stride_b, const int32_t* d_pairs, const int* d_npairs, double* d_p1,
                                    double* d_p2, void* stream) {
    if (!plan) { set_error("null plan"); return MSIFT_EINVAL; }
    return match_batch_gather(plan, d_kpa, stride_a, d_kpb, stride_b, d_pairs, d_npairs, d_p1, d_p2,
                              (hipStream_t)stream);
}

}  // extern "C"

namespace msift {

// the reference's similarity matcher (core.py:2652-2701) on unit rows of D
// floats: binary16 inputs, a.b on FP16 MFMA with fp32 accumulation, the
// result rounded to binary16, top-2 by similarity in both directions, ratio
// test on d = max(2 - 2 s, 0), mutual pairs
int match_sim16(const float* a, int N, const float* b, int M, int D, double ratio, void* ws_, size_t ws_bytes,
                int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s) {
    if (N < 0 || M < 0 || D < 1 || D > 128) { set_error("msift_match_sim_f16: bad arguments (1 <= D <= 128)"); return MSIFT_EINVAL; }
    MatchPlan P = match_plan(N, M);
    if (ws_bytes < P.total) { set_error("msift_match_sim_f16: workspace too small"); return MSIFT_EWORKSPACE; }
    if ((N > 0 && (!a || !fwd)) || (M > 0 && (!b || !bwd)) || !npairs || (!pairs && N > 0 && M > 0)) {
        set_error("msift_match_sim_f16: null buffer");
        return MSIFT_EINVAL;
    }
    char* ws = (char*)ws_;
    _Float16* ha = (_Float16*)(ws + P.o_ha);
    _Float16* hb = (_Float16*)(ws + P.o_hb);
    uint32_t* rowpart = (uint32_t*)(ws + P.o_rowpart);
    uint32_t* colpart = (uint32_t*)(ws + P.o_colpart);
    const int KS = (D + 15) / 16, Dp = 16 * KS;
    if (N > 0 && M > 0) {
        const size_t ea = (size_t)N * Dp, eb = (size_t)M * Dp;
        hipLaunchKernelGGL(k_to_f16_pad, dim3((unsigned)std::min<size_t>((ea + 255) / 256, 16384)), dim3(256), 0, s, a,
                           N, D, Dp, ha);
        hipLaunchKernelGGL(k_to_f16_pad, dim3((unsigned)std::min<size_t>((eb + 255) / 256, 16384)), dim3(256), 0, s, b,
                           M, D, Dp, hb);
        const dim3 g(P.nrb_a, P.ncc_a);
        const _Float16 *cha = ha, *chb = hb;
        switch (KS) {
        case 1: hipLaunchKernelGGL((k_match_tiles_f16<1, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        case 2: hipLaunchKernelGGL((k_match_tiles_f16<2, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        case 3: hipLaunchKernelGGL((k_match_tiles_f16<3, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        case 4: hipLaunchKernelGGL((k_match_tiles_f16<4, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        case 5: hipLaunchKernelGGL((k_match_tiles_f16<5, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        case 6: hipLaunchKernelGGL((k_match_tiles_f16<6, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        case 7: hipLaunchKernelGGL((k_match_tiles_f16<7, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        default: hipLaunchKernelGGL((k_match_tiles_f16<8, true>), g, dim3(512), 0, s, cha, N, chb, M, nullptr, nullptr, rowpart, colpart); break;
        }
        hipLaunchKernelGGL(k_reduce_top2, dim3((N + 63) / 64), dim3(256), 0, s, rowpart, N, P.ncc_a, 9, kColChunk, fwd,
                           nullptr);
        hipLaunchKernelGGL(k_reduce_top2, dim3((M + 63) / 64), dim3(256), 0, s, colpart, M, P.nrb_a, 8, kRowBlk, bwd,
                           nullptr);
        hipLaunchKernelGGL(k_sim_to_dist, dim3((N + 255) / 256), dim3(256), 0, s, fwd, N);
        hipLaunchKernelGGL(k_sim_to_dist, dim3((M + 255) / 256), dim3(256), 0, s, bwd, M);
    } else {
        if (N > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((N + 255) / 256), dim3(256), 0, s, fwd, N);
        if (M > 0) hipLaunchKernelGGL(k_fill_top2_none, dim3((M + 255) / 256), dim3(256), 0, s, bwd, M);
    }
    const int nblk = (N + 255) / 256;
    const double r2 = ratio * ratio;
    int* cnt = (int*)(ws + P.o_cnt);
    int* off = (int*)(ws + P.o_off);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_count, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, cnt, 1);
    launch_scan(cnt, off, nblk, npairs, nullptr, nullptr, 0, s);
    if (nblk > 0) hipLaunchKernelGGL(k_mutual_write, dim3(nblk), dim3(256), 0, s, fwd, bwd, N, M, r2, off, pairs, 1);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift
