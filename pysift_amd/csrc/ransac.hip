// GPU-batched RANSAC with LO refit (restates _ransac / _ransac_batch_eval /
// _compute_homography / _normalize_points, pysift/core.py:3011-3097,
// 2272-2313, 2946-2981), all in FP64.
//
// Hypothesis b draws Philox4x32-10(counter = (b,0,0,0), key = seed) and maps
// each 32-bit word u to the point index (u * N) >> 32 (with replacement).
// The minimal solve is Gaussian elimination with partial pivoting (first
// maximum on ties) of the 8x9 DLT system with h8 = 1; the least-squares refit
// is the smallest-eigenvalue eigenvector of A^T A by cyclic Jacobi.  Sums over
// points use 64 lane partials (point k -> lane k % 64) and the xor tree, so
// every result is reproduced bit for bit by the CPU oracle.
#include "msift_internal.hpp"
#include <vector>

#pragma clang fp contract(off)

namespace msift {

constexpr int kHypPerBlock = 64;
constexpr int kScoreWaves = 16;    // scoring: one hypothesis per wave, 16 waves per block

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ void dlt_rows(double x1, double y1, double x2, double y2, double* re, double* ro) {
    re[0] = -x1; re[1] = -y1; re[2] = -1.0; re[3] = 0.0; re[4] = 0.0; re[5] = 0.0;
    re[6] = x2 * x1; re[7] = x2 * y1; re[8] = x2;
    ro[0] = 0.0; ro[1] = 0.0; ro[2] = 0.0; ro[3] = -x1; ro[4] = -y1; ro[5] = -1.0;
    ro[6] = y2 * x1; ro[7] = y2 * y1; ro[8] = y2;
}

// null vector of the 8x9 system, h8 = 1; static indices only (registers)
__device__ __forceinline__ bool solve8(double A[8][9], double* h) {
    double amax = 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 9; ++c) amax = fabs(A[r][c]) > amax ? fabs(A[r][c]) : amax;
    const double tiny = amax * 1e-12;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int p = c;
        double best = fabs(A[c][c]);
#pragma unroll
        for (int r = c + 1; r < 8; ++r)
            if (fabs(A[r][c]) > best) { best = fabs(A[r][c]); p = r; }
        if (!(best > tiny)) ok = false;
#pragma unroll
        for (int r = c + 1; r < 8; ++r) {
            if (p == r) {
#pragma unroll
                for (int k = 0; k < 9; ++k) { double t = A[c][k]; A[c][k] = A[r][k]; A[r][k] = t; }
            }
        }
#pragma unroll
        for (int r = c + 1; r < 8; ++r) {
            double f = A[r][c] / A[c][c];
#pragma unroll
            for (int k = c; k < 9; ++k) A[r][k] = A[r][k] - f * A[c][k];
        }
    }
    h[8] = 1.0;
#pragma unroll
    for (int c = 7; c >= 0; --c) {
        double s = A[c][8];
#pragma unroll
        for (int k = c + 1; k < 8; ++k) s = s + A[c][k] * h[k];
        h[c] = -s / A[c][c];
    }
    return ok;
}

__device__ __forceinline__ void denorm(const double* Hn, const double* T1, const double* T2, double* H) {
    double s2 = T2[0], inv_s2 = 1.0 / s2;
    double cx2 = -T2[2] / s2, cy2 = -T2[5] / s2;
    double Ti[9] = { inv_s2, 0.0, cx2, 0.0, inv_s2, cy2, 0.0, 0.0, 1.0 };
    double A[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            A[3 * r + c] = (Ti[3 * r] * Hn[c] + Ti[3 * r + 1] * Hn[3 + c]) + Ti[3 * r + 2] * Hn[6 + c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            H[3 * r + c] = (A[3 * r] * T1[c] + A[3 * r + 1] * T1[3 + c]) + A[3 * r + 2] * T1[6 + c];
    double h22 = H[8];
    if (fabs(h22) > 1e-10) {
#pragma unroll
        for (int i = 0; i < 9; ++i) H[i] = H[i] / h22;
    }
}

__device__ __forceinline__ void inv3(const double* a, double* o) {
    double det = (a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6])) + a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (!(fabs(det) > 1e-12)) det = 1.0;
    o[0] = (a[4] * a[8] - a[5] * a[7]) / det; o[1] = (a[2] * a[7] - a[1] * a[8]) / det; o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    o[3] = (a[5] * a[6] - a[3] * a[8]) / det; o[4] = (a[0] * a[8] - a[2] * a[6]) / det; o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    o[6] = (a[3] * a[7] - a[4] * a[6]) / det; o[7] = (a[1] * a[6] - a[0] * a[7]) / det; o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

__device__ __forceinline__ double sym_err(const double* H, const double* Hi, double x1, double y1, double x2, double y2,
                                          bool clampz) {
    double fx = (H[0] * x1 + H[1] * y1) + H[2];
    double fy = (H[3] * x1 + H[4] * y1) + H[5];
    double fz = (H[6] * x1 + H[7] * y1) + H[8];
    if (clampz && fz < 1e-10) fz = 1e-10;
    double ex = fx / fz - x2, ey = fy / fz - y2;
    double bx = (Hi[0] * x2 + Hi[1] * y2) + Hi[2];
    double by = (Hi[3] * x2 + Hi[4] * y2) + Hi[5];
    double bz = (Hi[6] * x2 + Hi[7] * y2) + Hi[8];
    if (clampz && bz < 1e-10) bz = 1e-10;
    double gx = bx / bz - x1, gy = by / bz - y1;
    return (ex * ex + ey * ey) + (gx * gx + gy * gy);
}

// Hartley normalisation over the points listed in idx[0..m) (core.py:2946-2957)
// The lane's points are fetched kGather at a time (indices, then
// coordinates, all loads in flight together) and summed in the same order
// k = lane, lane + 64, ... as one at a time.
constexpr int kGather = 8;
__device__ void hartley_wave(const double* __restrict__ pts, const int* __restrict__ idx, int m, double* T) {
    const int lane = threadIdx.x & 63;
    double sx = 0.0, sy = 0.0;
    for (int k0 = lane; k0 < m; k0 += 64 * kGather) {
        int ii[kGather];
        double xv[kGather], yv[kGather];
#pragma unroll
        for (int u = 0; u < kGather; ++u) {
            const int k = k0 + 64 * u;
            ii[u] = k < m ? (idx ? idx[k] : k) : -1;
        }
#pragma unroll
        for (int u = 0; u < kGather; ++u)
            if (ii[u] >= 0) { xv[u] = pts[2 * ii[u]]; yv[u] = pts[2 * ii[u] + 1]; }
#pragma unroll
        for (int u = 0; u < kGather; ++u)
            if (ii[u] >= 0) { sx = sx + xv[u]; sy = sy + yv[u]; }
    }
    double cx = wave_tree_d(sx) / (double)m;
    double cy = wave_tree_d(sy) / (double)m;
    double sd = 0.0;
    for (int k0 = lane; k0 < m; k0 += 64 * kGather) {
        int ii[kGather];
        double xv[kGather], yv[kGather];
#pragma unroll
        for (int u = 0; u < kGather; ++u) {
            const int k = k0 + 64 * u;
            ii[u] = k < m ? (idx ? idx[k] : k) : -1;
        }
#pragma unroll
        for (int u = 0; u < kGather; ++u)
            if (ii[u] >= 0) { xv[u] = pts[2 * ii[u]]; yv[u] = pts[2 * ii[u] + 1]; }
#pragma unroll
        for (int u = 0; u < kGather; ++u)
            if (ii[u] >= 0) {
                double dx = xv[u] - cx, dy = yv[u] - cy;
                sd = sd + sqrt(dx * dx + dy * dy);
            }
    }
    double md = wave_tree_d(sd) / (double)m;
    double s = 0x1.6a09e667f3bcdp+0 / (md + 1e-10);
    T[0] = s; T[1] = 0.0; T[2] = -s * cx;
    T[3] = 0.0; T[4] = s; T[5] = -s * cy;
    T[6] = 0.0; T[7] = 0.0; T[8] = 1.0;
}

// ws layout: scores[B] | Hs[B*9] | T[18] | best(int) ninl(int) | idx[N] | mask2[N]
struct RansacWs {
    double* scores;
    double* Hs;
    double* T;
    int* ints;
    int* idx;
    uint8_t* mask2;
};

__device__ __forceinline__ void normalize_body(const double* __restrict__ p1, const double* __restrict__ p2, int N,
                                               double* __restrict__ T) {
    double T1[9], T2[9];
    hartley_wave(p1, nullptr, N, T1);
    hartley_wave(p2, nullptr, N, T2);
    if (threadIdx.x == 0)
        for (int k = 0; k < 9; ++k) { T[k] = T1[k]; T[9 + k] = T2[k]; }
}

__global__ __launch_bounds__(64) void k_normalize(const double* __restrict__ p1, const double* __restrict__ p2, int N,
                                                   double* __restrict__ T) {
    normalize_body(p1, p2, N, T);
}

// hypotheses blk*64 .. blk*64+63, one per lane: the minimal solve.  Hs[b]
// gets H (the identity when the solve fails) and scores[b] -1 for a failed
// solve, 0 otherwise (the scoring kernel then fills in the score).
__device__ __forceinline__ void hyp_solve_body(const double* __restrict__ p1, const double* __restrict__ p2, int N,
                                               int B, uint32_t k0, uint32_t k1, const double* __restrict__ T,
                                               double* __restrict__ scores, double* __restrict__ Hs, int blk) {
    const int lane = threadIdx.x & 63;
    const int b = blk * kHypPerBlock + lane;
    double T1[9], T2[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) { T1[k] = T[k]; T2[k] = T[9 + k]; }
    uint32_t c[4] = { (uint32_t)b, 0u, 0u, 0u };
    philox4x32_10(c, k0, k1);
    double A[8][9];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int i = (int)(((uint64_t)c[q] * (uint64_t)N) >> 32);
        double x1 = T1[0] * p1[2 * i] + T1[2], y1 = T1[4] * p1[2 * i + 1] + T1[5];
        double x2 = T2[0] * p2[2 * i] + T2[2], y2 = T2[4] * p2[2 * i + 1] + T2[5];
        dlt_rows(x1, y1, x2, y2, A[2 * q], A[2 * q + 1]);
    }
    double h[9], H[9];
    const bool ok = solve8(A, h) && b < B;
    if (ok) {
        denorm(h, T1, T2, H);
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) H[k] = (k % 4 == 0) ? 1.0 : 0.0;
    }
    if (b < B) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Hs[(size_t)b * 9 + k] = H[k];
        scores[b] = ok ? 0.0 : -1.0;
    }
}

// hypothesis blk * kScoreWaves + w scored by wave w, lanes over the points
// (the same sum whichever wave computes it).  A kernel of its own with one
// hypothesis per wave: its few registers and many waves keep the FP64 units
// busy.
__device__ __forceinline__ void hyp_score_body(const double* __restrict__ p1, const double* __restrict__ p2, int N,
                                               int B, int method, double den, double tq,
                                               double* __restrict__ scores, const double* __restrict__ Hs, int blk) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    {
        const int bb = blk * kScoreWaves + wv;
        if (bb >= B || scores[bb] < 0.0) return;   // failed solve: stays -1
        double H[9], Hi[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) H[k] = Hs[(size_t)bb * 9 + k];
        inv3(H, Hi);
        double part = 0.0;
        for (int i = lane; i < N; i += 64) {
            double e = sym_err(H, Hi, p1[2 * i], p1[2 * i + 1], p2[2 * i], p2[2 * i + 1], true);
            double w = (method == 0) ? exp_d(-e / den) : (e < tq ? 1.0 : 0.0);
            part = part + w;
        }
        double sc = wave_tree_d(part);
        if (lane == 0) scores[bb] = sc;
    }
}

__global__ __launch_bounds__(64) void k_hyp_solve(const double* __restrict__ p1, const double* __restrict__ p2, int N,
                                                   int B, uint32_t k0, uint32_t k1, const double* __restrict__ T,
                                                   double* __restrict__ scores, double* __restrict__ Hs) {
    hyp_solve_body(p1, p2, N, B, k0, k1, T, scores, Hs, blockIdx.x);
}

__global__ __launch_bounds__(64 * kScoreWaves) void k_hyp_score(const double* __restrict__ p1,
                                                               const double* __restrict__ p2, int N, int B, int method,
                                                               double den, double tq, double* __restrict__ scores,
                                                               const double* __restrict__ Hs) {
    hyp_score_body(p1, p2, N, B, method, den, tq, scores, Hs, blockIdx.x);
}

// argmax with the smallest index on ties (a total order: any reduction order)
__device__ __forceinline__ void argmax_body(const double* __restrict__ scores, int B, int* __restrict__ best,
                                            double* ss, int* si) {
    double bs = -__builtin_huge_val();
    int bi = 0x7fffffff;
    for (int b = threadIdx.x; b < B; b += 1024) {
        double v = scores[b];
        if (v > bs || (v == bs && b < bi)) { bs = v; bi = b; }
    }
    ss[threadIdx.x] = bs; si[threadIdx.x] = bi;
    __syncthreads();
    for (int off = 512; off >= 1; off >>= 1) {
        if ((int)threadIdx.x < off) {
            double v = ss[threadIdx.x + off];
            int i = si[threadIdx.x + off];
            if (v > ss[threadIdx.x] || (v == ss[threadIdx.x] && i < si[threadIdx.x])) { ss[threadIdx.x] = v; si[threadIdx.x] = i; }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *best = si[0];
}

__global__ __launch_bounds__(1024) void k_argmax(const double* __restrict__ scores, int B, int* __restrict__ best) {
    __shared__ double ss[1024];
    __shared__ int si[1024];
    argmax_body(scores, B, best, ss, si);
}

// compacts mask (in index order) into idx, returns the count (all lanes)
// (the flags of 8 chunks of 64 are loaded together, then compacted chunk by
// chunk in order)
__device__ int compact_wave(const uint8_t* __restrict__ mask, int N, int* __restrict__ idx) {
    const int lane = threadIdx.x & 63;
    int base = 0;
    for (int c0 = 0; c0 < N; c0 += 64 * 8) {
        bool fl[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = c0 + 64 * u + lane;
            fl[u] = i < N && mask[i];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = c0 + 64 * u + lane;
            const unsigned long long m = __ballot(fl[u]);
            if (fl[u]) idx[base + __popcll(m & lanemask_lt())] = i;
            base += __popcll(m);
        }
    }
    return base;
}

__device__ void jacobi9(double* M, double* vec) {
    double V[81];
#pragma unroll
    for (int i = 0; i < 81; ++i) V[i] = (i % 10 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, diag = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                if (i != j) off = off + M[9 * i + j] * M[9 * i + j];
                else diag = diag + M[9 * i + j] * M[9 * i + j];
            }
        if (off <= diag * 1e-30) break;
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int q = p + 1; q < 9; ++q) {
                double apq = M[9 * p + q];
                if (apq != 0.0) {
                    double app = M[9 * p + p], aqq = M[9 * q + q];
                    double theta = (aqq - app) / (2.0 * apq);
                    double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                    if (theta < 0.0) t = -t;
                    double c = 1.0 / sqrt(t * t + 1.0);
                    double s = t * c;
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        double mkp = M[9 * k + p], mkq = M[9 * k + q];
                        M[9 * k + p] = c * mkp - s * mkq;
                        M[9 * k + q] = s * mkp + c * mkq;
                    }
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        double mpk = M[9 * p + k], mqk = M[9 * q + k];
                        M[9 * p + k] = c * mpk - s * mqk;
                        M[9 * q + k] = s * mpk + c * mqk;
                    }
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        double vkp = V[9 * k + p], vkq = V[9 * k + q];
                        V[9 * k + p] = c * vkp - s * vkq;
                        V[9 * k + q] = s * vkp + c * vkq;
                    }
                }
            }
    }
    int best = 0;
    double bv = M[0];
#pragma unroll
    for (int i = 1; i < 9; ++i)
        if (M[9 * i + i] < bv) { bv = M[9 * i + i]; best = i; }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double v = V[9 * k];
#pragma unroll
        for (int i = 1; i < 9; ++i) v = (best == i) ? V[9 * k + i] : v;
        vec[k] = v;
    }
}

// least-squares DLT over the points in idx[0..m) (core.py:2959-2981)
__device__ void dlt_ls_wave(const double* __restrict__ p1, const double* __restrict__ p2, const int* __restrict__ idx,
                            int m, double* __restrict__ Hout, double* __restrict__ sM) {
    const int lane = threadIdx.x & 63;
    double T1[9], T2[9];
    hartley_wave(p1, idx, m, T1);
    hartley_wave(p2, idx, m, T2);
    double part[45];
#pragma unroll
    for (int e = 0; e < 45; ++e) part[e] = 0.0;
    // points fetched 4 at a time (as in hartley_wave), accumulated in order
    for (int k0 = lane; k0 < m; k0 += 64 * 4) {
        int ii[4];
        double q[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) ii[u] = k0 + 64 * u < m ? idx[k0 + 64 * u] : -1;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (ii[u] >= 0) {
                q[u][0] = p1[2 * ii[u]]; q[u][1] = p1[2 * ii[u] + 1];
                q[u][2] = p2[2 * ii[u]]; q[u][3] = p2[2 * ii[u] + 1];
            }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (ii[u] < 0) continue;
            double x1 = T1[0] * q[u][0] + T1[2], y1 = T1[4] * q[u][1] + T1[5];
            double x2 = T2[0] * q[u][2] + T2[2], y2 = T2[4] * q[u][3] + T2[5];
            double re[9], ro[9];
            dlt_rows(x1, y1, x2, y2, re, ro);
            int e = 0;
#pragma unroll
            for (int a = 0; a < 9; ++a)
#pragma unroll
                for (int b = a; b < 9; ++b) {
                    part[e] = part[e] + (re[a] * re[b] + ro[a] * ro[b]);
                    e++;
                }
        }
    }
    int e = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
        for (int b = a; b < 9; ++b) {
            double v = wave_tree_d(part[e]);
            if (lane == 0) { sM[9 * a + b] = v; sM[9 * b + a] = v; }
            e++;
        }
    __syncthreads();
    if (lane == 0) {
        double M[81], h[9], H[9];
#pragma unroll
        for (int k = 0; k < 81; ++k) M[k] = sM[k];
        jacobi9(M, h);
        denorm(h, T1, T2, H);
#pragma unroll
        for (int k = 0; k < 9; ++k) Hout[k] = H[k];
    }
    __syncthreads();
}

// inlier flags e < tq of H over all N points into f (order-free: a byte per
// point and an integer count), the coordinates of 4 points per lane fetched
// together; returns the wave's count
__device__ __forceinline__ int flag_inliers(const double* H, const double* Hi, const double* __restrict__ p1,
                                            const double* __restrict__ p2, int N, double tq, bool clampz,
                                            uint8_t* __restrict__ f) {
    const int lane = threadIdx.x & 63;
    int n = 0;
    for (int i0 = lane; i0 < N; i0 += 64 * 4) {
        double q[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + 64 * u;
            if (i < N) { q[u][0] = p1[2 * i]; q[u][1] = p1[2 * i + 1]; q[u][2] = p2[2 * i]; q[u][3] = p2[2 * i + 1]; }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + 64 * u;
            if (i < N) {
                const uint8_t fl = sym_err(H, Hi, q[u][0], q[u][1], q[u][2], q[u][3], clampz) < tq;
                f[i] = fl;
                n += fl;
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) n += __shfl_xor(n, m);
    return n;
}

__device__ __forceinline__ void finalize_body(const double* __restrict__ p1, const double* __restrict__ p2, int N,
                                              double tq, int lo, int min_inl, const double* __restrict__ Hs,
                                              const int* __restrict__ bestp, int* __restrict__ idx,
                                              uint8_t* __restrict__ mask2, double* __restrict__ Hout,
                                              uint8_t* __restrict__ mask, int* __restrict__ ninl,
                                              int* __restrict__ best_out, double* sM, double* sH1, double* sH2) {
    const int lane = threadIdx.x;
    const int best = *bestp;
    double Hb[9], Hbi[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Hb[k] = Hs[(size_t)best * 9 + k];
    inv3(Hb, Hbi);
    const int nb = flag_inliers(Hb, Hbi, p1, p2, N, tq, true, mask);
    if (lane == 0) {
        for (int k = 0; k < 9; ++k) Hout[k] = Hb[k];
        *best_out = best;
    }
    int nfinal = nb;
    if (lo && nb >= min_inl) {
        __syncthreads();
        int m = compact_wave(mask, N, idx);
        __syncthreads();
        dlt_ls_wave(p1, p2, idx, m, sH1, sM);
        double H1[9], H1i[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) H1[k] = sH1[k];
        inv3(H1, H1i);
        const int n2 = flag_inliers(H1, H1i, p1, p2, N, tq, false, mask2);
        __syncthreads();
        if (n2 >= min_inl) {
            int m2 = compact_wave(mask2, N, idx);
            __syncthreads();
            dlt_ls_wave(p1, p2, idx, m2, sH2, sM);
            for (int i = lane; i < N; i += 64) mask[i] = mask2[i];
            if (lane == 0)
                for (int k = 0; k < 9; ++k) Hout[k] = sH2[k];
            nfinal = n2;
        } else {
            if (lane == 0)
                for (int k = 0; k < 9; ++k) Hout[k] = H1[k];
        }
    }
    if (lane == 0) *ninl = nfinal;
}

__global__ __launch_bounds__(64) void k_finalize(const double* __restrict__ p1, const double* __restrict__ p2, int N,
                                                  double tq, int lo, int min_inl, const double* __restrict__ Hs,
                                                  const int* __restrict__ bestp, int* __restrict__ idx,
                                                  uint8_t* __restrict__ mask2, double* __restrict__ Hout,
                                                  uint8_t* __restrict__ mask, int* __restrict__ ninl,
                                                  int* __restrict__ best_out) {
    __shared__ double sM[81];
    __shared__ double sH1[9], sH2[9];
    finalize_body(p1, p2, N, tq, lo, min_inl, Hs, bestp, idx, mask2, Hout, mask, ninl, best_out, sM, sH1, sH2);
}

// ------------------------------------------------------------- batched ----
// Pairs of point sets packed back to back (pair p at point offset poff, room
// for cap points); the point count of each pair is read from device memory,
// so a caller can chain the batched matcher without a host round trip.
// Pair p uses key = seed + p: its result equals a single call with that seed.
struct RPair {
    size_t poff, cap;
    size_t o_scores, o_Hs, o_T, o_ints, o_idx, o_mask2;
};

__device__ __forceinline__ int pair_npts(const RPair& P, const int* npts) {
    int n = npts[blockIdx.y];
    return n < 0 ? 0 : (n > (int)P.cap ? (int)P.cap : n);
}

__global__ __launch_bounds__(64) void k_bnormalize(const RPair* __restrict__ rp, const double* __restrict__ p1,
                                                    const double* __restrict__ p2, const int* __restrict__ npts,
                                                    char* __restrict__ ws) {
    const RPair P = rp[blockIdx.y];
    const int N = pair_npts(P, npts);
    if (N < 4) return;
    normalize_body(p1 + 2 * P.poff, p2 + 2 * P.poff, N, (double*)(ws + P.o_T));
}

__global__ __launch_bounds__(64) void k_bhyp_solve(const RPair* __restrict__ rp, const double* __restrict__ p1,
                                                    const double* __restrict__ p2, const int* __restrict__ npts, int B,
                                                    uint64_t seed, char* __restrict__ ws) {
    const RPair P = rp[blockIdx.y];
    const int N = pair_npts(P, npts);
    if (N < 4) return;
    const uint64_t key = seed + (uint64_t)blockIdx.y;
    hyp_solve_body(p1 + 2 * P.poff, p2 + 2 * P.poff, N, B, (uint32_t)key, (uint32_t)(key >> 32),
                   (const double*)(ws + P.o_T), (double*)(ws + P.o_scores), (double*)(ws + P.o_Hs), blockIdx.x);
}

__global__ __launch_bounds__(64 * kScoreWaves) void k_bhyp_score(const RPair* __restrict__ rp,
                                                                const double* __restrict__ p1,
                                                                const double* __restrict__ p2,
                                                                const int* __restrict__ npts, int B, int method,
                                                                double den, double tq, char* __restrict__ ws) {
    const RPair P = rp[blockIdx.y];
    const int N = pair_npts(P, npts);
    if (N < 4) return;
    hyp_score_body(p1 + 2 * P.poff, p2 + 2 * P.poff, N, B, method, den, tq, (double*)(ws + P.o_scores),
                   (const double*)(ws + P.o_Hs), blockIdx.x);
}

__global__ __launch_bounds__(1024) void k_bargmax(const RPair* __restrict__ rp, const int* __restrict__ npts, int B,
                                                   char* __restrict__ ws) {
    __shared__ double ss[1024];
    __shared__ int si[1024];
    const RPair P = rp[blockIdx.y];
    if (pair_npts(P, npts) < 4) return;
    argmax_body((const double*)(ws + P.o_scores), B, (int*)(ws + P.o_ints), ss, si);
}

// pairs with fewer than 4 points: H = 0, no inliers, best = -1
__global__ __launch_bounds__(64) void k_bfinalize(const RPair* __restrict__ rp, const double* __restrict__ p1,
                                                   const double* __restrict__ p2, const int* __restrict__ npts, double tq,
                                                   int lo, int min_inl, char* __restrict__ ws, double* __restrict__ Hout,
                                                   uint8_t* __restrict__ mask, int* __restrict__ ninl,
                                                   int* __restrict__ best) {
    __shared__ double sM[81];
    __shared__ double sH1[9], sH2[9];
    const RPair P = rp[blockIdx.y];
    const int N = pair_npts(P, npts);
    const int p = blockIdx.y;
    if (N < 4) {
        if (threadIdx.x < 9) Hout[9 * p + threadIdx.x] = 0.0;
        for (int i = threadIdx.x; i < N; i += 64) mask[P.poff + i] = 0;
        if (threadIdx.x == 0) { ninl[p] = 0; if (best) best[p] = -1; }
        return;
    }
    int* ints = (int*)(ws + P.o_ints);
    finalize_body(p1 + 2 * P.poff, p2 + 2 * P.poff, N, tq, lo, min_inl, (const double*)(ws + P.o_Hs), ints,
                  (int*)(ws + P.o_idx), (uint8_t*)(ws + P.o_mask2), Hout + 9 * p, mask + P.poff, ninl + p,
                  best ? best + p : ints + 1, sM, sH1, sH2);
}

// ---------------------------------------------------------------- host ----
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t ransac_workspace(int N, int B) {
    return align256((size_t)B * 8) + align256((size_t)B * 72) + align256(18 * 8) + align256(16) +
           align256((size_t)(N > 0 ? N : 1) * 4) + align256((size_t)(N > 0 ? N : 1));
}

int ransac(const double* p1, const double* p2, int N, uint64_t seed, int B, double thr, double sigma_max, int method,
           int lo, int min_inl, void* ws_, size_t ws_bytes, double* H, uint8_t* mask, int* ninl, int* best,
           hipStream_t s) {
    if (ws_bytes < ransac_workspace(N, B)) { set_error("ransac: workspace too small"); return MSIFT_EWORKSPACE; }
    if (!p1 || !p2 || !H || !mask || !ninl) { set_error("ransac: null buffer"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    size_t off = 0;
    double* scores = (double*)(ws + off); off += align256((size_t)B * 8);
    double* Hs = (double*)(ws + off); off += align256((size_t)B * 72);
    double* T = (double*)(ws + off); off += align256(18 * 8);
    int* ints = (int*)(ws + off); off += align256(16);
    int* idx = (int*)(ws + off); off += align256((size_t)N * 4);
    uint8_t* mask2 = (uint8_t*)(ws + off);
    const double tq = thr * thr;
    const double den = 2.0 * (sigma_max * sigma_max);
    hipLaunchKernelGGL(k_normalize, dim3(1), dim3(64), 0, s, p1, p2, N, T);
    const unsigned nhb = (unsigned)((B + kHypPerBlock - 1) / kHypPerBlock);
    hipLaunchKernelGGL(k_hyp_solve, dim3(nhb), dim3(64), 0, s, p1, p2, N, B, (uint32_t)seed, (uint32_t)(seed >> 32),
                       (const double*)T, scores, Hs);
    hipLaunchKernelGGL(k_hyp_score, dim3((unsigned)((B + kScoreWaves - 1) / kScoreWaves)), dim3(64 * kScoreWaves), 0, s,
                       p1, p2, N, B, method, den, tq, scores, (const double*)Hs);
    hipLaunchKernelGGL(k_argmax, dim3(1), dim3(1024), 0, s, (const double*)scores, B, &ints[0]);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, s, p1, p2, N, tq, lo, min_inl, (const double*)Hs,
                       (const int*)&ints[0], idx, mask2, H, mask, ninl, best ? best : &ints[1]);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift

struct msift_ransac_batch {
    int np, B;
    std::vector<msift::RPair> pairs;
    size_t total;
    msift::RPair* d_pairs;
};

namespace msift {

static int ransac_batch_create(int np, const int* cap, int B, msift_ransac_batch** out) {
    if (np < 1 || np > 65535 || !cap || B < 1 || !out) { set_error("ransac batch: bad arguments"); return MSIFT_EINVAL; }
    msift_ransac_batch* R = new (std::nothrow) msift_ransac_batch();
    if (!R) { set_error("out of host memory"); return MSIFT_EINVAL; }
    R->np = np; R->B = B; R->d_pairs = nullptr;
    size_t off = 0, poff = 0;
    auto take = [&](size_t b) { size_t o = off; off += align256(b); return o; };
    for (int p = 0; p < np; ++p) {
        if (cap[p] < 0) { set_error("ransac batch: negative capacity"); delete R; return MSIFT_EINVAL; }
        RPair P;
        P.poff = poff; P.cap = (size_t)cap[p];
        poff += (size_t)cap[p];
        P.o_scores = take((size_t)B * 8);
        P.o_Hs = take((size_t)B * 72);
        P.o_T = take(18 * 8);
        P.o_ints = take(16);
        P.o_idx = take((size_t)(cap[p] > 0 ? cap[p] : 1) * 4);
        P.o_mask2 = take((size_t)(cap[p] > 0 ? cap[p] : 1));
        R->pairs.push_back(P);
    }
    R->total = off;
    if (hipMalloc(&R->d_pairs, sizeof(RPair) * np) != hipSuccess ||
        hipMemcpy(R->d_pairs, R->pairs.data(), sizeof(RPair) * np, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("ransac batch: device table allocation failed");
        if (R->d_pairs) (void)hipFree(R->d_pairs);
        delete R;
        return MSIFT_EHIP;
    }
    *out = R;
    return MSIFT_OK;
}

static int ransac_batch_run(msift_ransac_batch* R, const double* p1, const double* p2, const int* npts, uint64_t seed,
                            double thr, double sigma_max, int method, int lo, int min_inl, void* ws_, size_t ws_bytes,
                            double* H, uint8_t* mask, int* ninl, int* best, hipStream_t s) {
    if (ws_bytes < R->total) { set_error("ransac_homography_batched: workspace too small"); return MSIFT_EWORKSPACE; }
    if (!p1 || !p2 || !npts || !H || !mask || !ninl || !ws_) {
        set_error("ransac_homography_batched: null buffer");
        return MSIFT_EINVAL;
    }
    if (!(thr > 0.0) || !(sigma_max > 0.0) || (method != 0 && method != 1)) {
        set_error("ransac_homography_batched: bad arguments");
        return MSIFT_EINVAL;
    }
    char* ws = (char*)ws_;
    const int np = R->np, B = R->B;
    const double tq = thr * thr;
    const double den = 2.0 * (sigma_max * sigma_max);
    const RPair* rp = R->d_pairs;
    hipLaunchKernelGGL(k_bnormalize, dim3(1, np), dim3(64), 0, s, rp, p1, p2, npts, ws);
    const unsigned nhb = (unsigned)((B + kHypPerBlock - 1) / kHypPerBlock);
    hipLaunchKernelGGL(k_bhyp_solve, dim3(nhb, np), dim3(64), 0, s, rp, p1, p2, npts, B, seed, ws);
    hipLaunchKernelGGL(k_bhyp_score, dim3((unsigned)((B + kScoreWaves - 1) / kScoreWaves), np), dim3(64 * kScoreWaves), 0,
                       s, rp, p1, p2, npts, B, method, den, tq, ws);
    hipLaunchKernelGGL(k_bargmax, dim3(1, np), dim3(1024), 0, s, rp, npts, B, ws);
    hipLaunchKernelGGL(k_bfinalize, dim3(1, np), dim3(64), 0, s, rp, p1, p2, npts, tq, lo, min_inl, ws, H, mask, ninl,
                       best);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift

using namespace msift;

extern "C" {

int msift_ransac_batch_create(int n_pairs, const int* cap, int n_hyp, msift_ransac_batch** plan) {
    return ransac_batch_create(n_pairs, cap, n_hyp, plan);
}

int msift_ransac_batch_destroy(msift_ransac_batch* plan) {
    if (!plan) return MSIFT_OK;
    if (plan->d_pairs) (void)hipFree(plan->d_pairs);
    delete plan;
    return MSIFT_OK;
}

int msift_ransac_batch_sizes(msift_ransac_batch* plan, size_t* ws_bytes, long long* point_offsets) {
    if (!plan) { set_error("null plan"); return MSIFT_EINVAL; }
    if (ws_bytes) *ws_bytes = plan->total;
    if (point_offsets)
        for (int p = 0; p < plan->np; ++p) point_offsets[p] = (long long)plan->pairs[p].poff;
    return MSIFT_OK;
}

int msift_ransac_homography_batched(msift_ransac_batch* plan, const double* d_p1, const double* d_p2,
                                    const int* d_npts, uint64_t seed, double thr_px, double sigma_max, int method,
                                    int lo, int min_inliers, void* d_ws, size_t ws_bytes, double* d_H,
                                    uint8_t* d_mask, int* d_ninl, int* d_best, void* stream) {
    if (!plan) { set_error("null plan"); return MSIFT_EINVAL; }
    if (lo && min_inliers < 4) {
        set_error("msift_ransac_homography_batched: min_inliers must be >= 4 when lo is set");
        return MSIFT_EINVAL;
    }
    return ransac_batch_run(plan, d_p1, d_p2, d_npts, seed, thr_px, sigma_max, method, lo, min_inliers, d_ws,
                            ws_bytes, d_H, d_mask, d_ninl, d_best, (hipStream_t)stream);
}

}  // extern "C"
