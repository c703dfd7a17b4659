// The reference's float matcher front end (core.py:2875-2886, 2703-2743):
// rows L2-normalised with +1e-7, a PCA fitted on both descriptor sets
// (cached by the caller across pairs, as the reference caches
// _pca_matrix), projection to `dims` and re-normalisation.
//
// Arithmetic (fixed, reproduced by the CPU oracle bit for bit):
//   - norms: float sum of squares in index order, x / (sqrtf(s) + 1e-7f);
//   - mean and centred Gram in double: rows summed in index order within
//     chunks of kPcaChunk rows, chunk partials summed in chunk order;
//     covariance = Gram / n;
//   - eigenvectors by the cyclic parallel Jacobi method in one workgroup (the
//     matrix in LDS, row stride 129 doubles so column walks hit distinct
//     banks; the rotations recorded and applied to the transposed eigenvector
//     matrix afterwards, one wave per column): each
//     sweep is 127 rounds of the circle-method pairing of the 128 indices
//     (round r: (r, 127) and ((r + k) % 127, (r - k) % 127), k = 1..63); the
//     64 rotations of a round are computed from the same matrix, then applied
//     to rows, then to columns and the eigenvectors, then a_pq = a_qp = 0; a
//     pair rotates when |a_pq| > 1e-15 sqrt(|a_pp a_qq|); sweeps stop after
//     one without a rotation (at most kPcaMaxSweeps);
//   - the `dims` largest eigenvalues (ties: lower index) give proj's columns
//     in ascending eigenvalue order (eigh(cov)[:, -dims:]), each signed so its
//     largest-magnitude entry (first on ties) is positive;
//   - projection: float c_j = x_j - mean_j, p_k = sum_j c_j proj_jk in j order
//     (separate multiply and add), then p / (sqrtf(sum_k p_k^2) + 1e-7f).
#include "msift_internal.hpp"
#include <algorithm>

#pragma clang fp contract(off)

namespace msift {

namespace {

constexpr int kPcaChunk = 128;     // rows per partial sum
constexpr int kPcaStage = 32;      // rows of a chunk staged in LDS at a time
constexpr int kPcaMaxSweeps = 40;
constexpr int kAStride = 129;      // LDS row stride of the Jacobi matrix (doubles)

// a block per 64 rows: loads into LDS (rows padded to 129 floats; float4 when
// x and y are 16-byte aligned), the first wave's lanes sum their row's
// squares in index order, then every thread writes normalised values
__global__ __launch_bounds__(256) void k_l2n_rows(const float* __restrict__ x, int n, float* __restrict__ y, int vec4) {
    __shared__ float t[64][129];
    __shared__ float den[64];
    const int r0 = blockIdx.x * 64, tid = threadIdx.x;
    const int nr = min(64, n - r0);
    if (vec4) {
        const float4* x4 = reinterpret_cast<const float4*>(x + (size_t)r0 * 128);
        for (int e = tid; e < nr * 32; e += 256) {
            const float4 v = x4[e];
            float* d = &t[e >> 5][(e & 31) * 4];
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
    } else {
        for (int e = tid; e < nr * 128; e += 256) t[e >> 7][e & 127] = x[(size_t)r0 * 128 + e];
    }
    __syncthreads();
    if (tid < nr) {
        float s = 0.0f;
        for (int k = 0; k < 128; ++k) s = s + t[tid][k] * t[tid][k];
        den[tid] = sqrtf(s) + 1e-7f;
    }
    __syncthreads();
    if (vec4) {
        float4* y4 = reinterpret_cast<float4*>(y + (size_t)r0 * 128);
        for (int e = tid; e < nr * 32; e += 256) {
            const float* d = &t[e >> 5][(e & 31) * 4];
            const float dn = den[e >> 5];
            y4[e] = make_float4(d[0] / dn, d[1] / dn, d[2] / dn, d[3] / dn);
        }
    } else {
        for (int e = tid; e < nr * 128; e += 256) y[(size_t)r0 * 128 + e] = t[e >> 7][e & 127] / den[e >> 7];
    }
}

inline int aligned16(const void* a, const void* b) { return ((((uintptr_t)a) | ((uintptr_t)b)) & 15) == 0; }

__global__ __launch_bounds__(128) void k_pca_mean_part(const float* __restrict__ X, int n, double* __restrict__ part) {
    const int c = blockIdx.x, j = threadIdx.x;
    const int r1 = min(n, (c + 1) * kPcaChunk);
    double s = 0.0;
    for (int r = c * kPcaChunk; r < r1; ++r) s += (double)X[(size_t)r * 128 + j];
    part[(size_t)c * 128 + j] = s;
}

__global__ __launch_bounds__(128) void k_pca_mean(const double* __restrict__ part, int nch, int n,
                                                  double* __restrict__ mean64, float* __restrict__ mean32) {
    const int j = threadIdx.x;
    double t = 0.0;
    for (int c = 0; c < nch; ++c) t += part[(size_t)c * 128 + j];
    const double m = t / (double)n;
    mean64[j] = m;
    mean32[j] = (float)m;
}

// chunk c's centred Gram: thread (tj, tk) owns the 8x8 block of rows
// 8 tj.. and columns 8 tk.., accumulated over the chunk's rows in order
__global__ __launch_bounds__(256) void k_pca_gram_part(const float* __restrict__ X, int n,
                                                       const double* __restrict__ mean64, double* __restrict__ part) {
    __shared__ double cs[kPcaStage][128];
    const int c = blockIdx.x, t = threadIdx.x, tj = t >> 4, tk = t & 15;
    double acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0.0;
    const int r0 = c * kPcaChunk, r1 = min(n, r0 + kPcaChunk);
    for (int s0 = r0; s0 < r1; s0 += kPcaStage) {
        const int ns = min(kPcaStage, r1 - s0);
        for (int e = t; e < ns * 128; e += 256)
            cs[e >> 7][e & 127] = (double)X[(size_t)s0 * 128 + e] - mean64[e & 127];
        __syncthreads();
        for (int r = 0; r < ns; ++r) {
            double va[8], vb[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) { va[q] = cs[r][8 * tj + q]; vb[q] = cs[r][8 * tk + q]; }
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) acc[a][b] += va[a] * vb[b];
        }
        __syncthreads();
    }
    double* g = part + (size_t)c * 16384;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) g[(8 * tj + a) * 128 + 8 * tk + b] = acc[a][b];
}

__global__ __launch_bounds__(256) void k_pca_cov(const double* __restrict__ part, int nch, int n, double* __restrict__ cov) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    double t = 0.0;
    for (int c = 0; c < nch; ++c) t += part[(size_t)c * 16384 + e];
    cov[e] = t / (double)n;
}

__device__ __forceinline__ void jacobi_cs(double app, double aqq, double apq, double& c, double& s) {
    const double theta = (aqq - app) / (2.0 * apq);
    double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
    if (theta < 0.0) t = -t;
    c = 1.0 / sqrt(t * t + 1.0);
    s = t * c;
}

// one workgroup: Jacobi sweeps on cov.  The eigenvector updates are not done
// here: every round that rotates records its 64 pairs (p | q << 8 | rot << 16)
// and (c, s) in rpq / rcs, k_pca_vt applies them, and the diagonal (the
// eigenvalues) goes to diag.
__global__ __launch_bounds__(1024) void k_pca_jacobi(const double* __restrict__ cov, int* __restrict__ rpq,
                                                     double* __restrict__ rcs, int* __restrict__ nrounds,
                                                     double* __restrict__ diag, int* __restrict__ sweeps_out) {
    __shared__ double A[128 * kAStride];
    __shared__ double Cs[64], Ss[64];
    __shared__ int Ps[64], Qs[64], Rs[64];
    const int tid = threadIdx.x;
    for (int e = tid; e < 16384; e += 1024) A[(e >> 7) * kAStride + (e & 127)] = cov[e];
    __syncthreads();
    int sweeps = 0, g = 0;
    for (; sweeps < kPcaMaxSweeps; ++sweeps) {
        int rotated = 0;
        for (int r = 0; r < 127; ++r) {
            int rot = 0, p = 0, q = 0;
            double c = 1.0, sn = 0.0;
            if (tid < 64) {
                if (tid == 0) { p = r; q = 127; }
                else { p = (r + tid) % 127; q = (r - tid + 127) % 127; }
                if (p > q) { const int t = p; p = q; q = t; }
                const double app = A[p * kAStride + p], aqq = A[q * kAStride + q], apq = A[p * kAStride + q];
                rot = fabs(apq) > 1e-15 * sqrt(fabs(app * aqq));
                if (rot) jacobi_cs(app, aqq, apq, c, sn);
                Ps[tid] = p; Qs[tid] = q; Rs[tid] = rot; Cs[tid] = c; Ss[tid] = sn;
            }
            if (!__syncthreads_or(rot)) continue;
            rotated = 1;
            if (tid < 64) {
                rpq[(size_t)g * 64 + tid] = p | (q << 8) | (rot << 16);
                rcs[((size_t)g * 64 + tid) * 2] = c;
                rcs[((size_t)g * 64 + tid) * 2 + 1] = sn;
            }
            ++g;
            // The row phase (rows p_k, q_k rotated by pair k) and then the column
            // phase (columns p_m, q_m by pair m) touch the 2x2 block {p_k, q_k} x
            // {p_m, q_m} only through its own four values, so one thread applies
            // both phases to its blocks in that order (the same operations as two
            // passes over the matrix) and zeroes a_pq of a rotated pair; 64 x 64
            // blocks over 1024 threads, every value owned by one thread.
            // pair m = tid % 64 is the same in all four blocks, pair k = tid / 64 + 16 u
            // is uniform over the wave; both are read before any A store
            const int m = tid & 63;
            const int pm = Ps[m], qm = Qs[m], rm = Rs[m];
            const double cm = Cs[m], sm = Ss[m];
            int pkk[4], qkk[4], rkk[4];
            double ckk[4], skk[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = (tid >> 6) + 16 * u;
                pkk[u] = Ps[k]; qkk[u] = Qs[k]; rkk[u] = Rs[k]; ckk[u] = Cs[k]; skk[u] = Ss[k];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = (tid >> 6) + 16 * u;
                const int pk = pkk[u], qk = qkk[u];
                double a00 = A[pk * kAStride + pm], a01 = A[pk * kAStride + qm];
                double a10 = A[qk * kAStride + pm], a11 = A[qk * kAStride + qm];
                if (rkk[u]) {
                    const double cc = ckk[u], ss = skk[u];
                    const double n00 = cc * a00 - ss * a10, n10 = ss * a00 + cc * a10;
                    const double n01 = cc * a01 - ss * a11, n11 = ss * a01 + cc * a11;
                    a00 = n00; a10 = n10; a01 = n01; a11 = n11;
                }
                if (rm) {
                    const double cc = cm, ss = sm;
                    const double n00 = cc * a00 - ss * a01, n01 = ss * a00 + cc * a01;
                    const double n10 = cc * a10 - ss * a11, n11 = ss * a10 + cc * a11;
                    a00 = n00; a01 = n01; a10 = n10; a11 = n11;
                }
                if (k == m && rkk[u]) { a01 = 0.0; a10 = 0.0; }
                A[pk * kAStride + pm] = a00; A[pk * kAStride + qm] = a01;
                A[qk * kAStride + pm] = a10; A[qk * kAStride + qm] = a11;
            }
            __syncthreads();
        }
        if (!rotated) break;
    }
    if (tid < 128) diag[tid] = A[tid * (kAStride + 1)];
    if (tid == 0) {
        *sweeps_out = sweeps;
        *nrounds = g;
    }
}

// the recorded rotations applied to the identity: wave w of block b carries
// column i = 4 b + w of Vt (Vt[p][i], Vt[q][i] per pair, the operations of a
// row update of Vt), lane k pair k of every round, the column in LDS.  The
// next kPcaPf rounds' records are in flight while a round is applied.
constexpr int kPcaPf = 4;
__global__ __launch_bounds__(256) void k_pca_vt(const int* __restrict__ rpq, const double* __restrict__ rcs,
                                                const int* __restrict__ nrounds, double* __restrict__ Vt) {
    __shared__ double col[4][128];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + w;
    double* v = col[w];
    v[lane] = lane == i ? 1.0 : 0.0;
    v[lane + 64] = lane + 64 == i ? 1.0 : 0.0;
    const int G = *nrounds;
    int pq[kPcaPf];
    double c[kPcaPf], sn[kPcaPf];
#pragma unroll
    for (int d = 0; d < kPcaPf; ++d) {
        const size_t e = (size_t)d * 64 + lane;
        pq[d] = d < G ? rpq[e] : 0;
        c[d] = d < G ? rcs[2 * e] : 1.0;
        sn[d] = d < G ? rcs[2 * e + 1] : 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int g0 = 0; g0 < G; g0 += kPcaPf) {
#pragma unroll
        for (int d = 0; d < kPcaPf; ++d) {
            const int g = g0 + d;
            if (g >= G) break;   // uniform
            const int p = pq[d] & 255, q = (pq[d] >> 8) & 255;
            if (pq[d] >> 16) {
                const double v0 = v[p], v1 = v[q];
                v[p] = c[d] * v0 - sn[d] * v1;
                v[q] = sn[d] * v0 + c[d] * v1;
            }
            // the round's 64 pairs are disjoint; the next round reads the
            // updated column (one wave: LDS operations complete in order)
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const int gn = g + kPcaPf;
            if (gn < G) {
                const size_t e = (size_t)gn * 64 + lane;
                pq[d] = rpq[e];
                c[d] = rcs[2 * e];
                sn[d] = rcs[2 * e + 1];
            }
        }
    }
    Vt[(size_t)lane * 128 + i] = v[lane];
    Vt[(size_t)(lane + 64) * 128 + i] = v[lane + 64];
}

// eigenvalue order (descending, ties: lower index), evals, proj.  Thread t
// ranks eigenvalue t by counting those ordered before it (a strict order, so
// the ranks are a permutation: the order a selection sort by the same
// comparison produces).
__global__ __launch_bounds__(128) void k_pca_finish(const double* __restrict__ Vt, const double* __restrict__ diag,
                                                    int dims, float* __restrict__ proj, double* __restrict__ evals) {
    __shared__ int order[128];
    __shared__ double dv[128];
    const int tid = threadIdx.x;
    dv[tid] = diag[tid];
    __syncthreads();
    {
        const double mine = dv[tid];
        int rank = 0;
        for (int b = 0; b < 128; ++b) {
            const double vb = dv[b];
            rank += (vb > mine || (vb == mine && b < tid)) ? 1 : 0;
        }
        order[rank] = tid;
    }
    __syncthreads();
    evals[tid] = dv[order[tid]];
    if (tid < dims) {
        const double* v = Vt + (size_t)order[dims - 1 - tid] * 128;
        // first largest magnitude: the loads do not depend on the running argmax
        int arg = 0;
        double best = fabs(v[0]);
#pragma unroll 8
        for (int j = 1; j < 128; ++j) {
            const double a = fabs(v[j]);
            if (a > best) { best = a; arg = j; }
        }
        const double sg = v[arg] < 0.0 ? -1.0 : 1.0;
#pragma unroll 8
        for (int j = 0; j < 128; ++j) proj[j * dims + tid] = (float)(sg * v[j]);
    }
}

// a block per 8 rows: centre, project (j order), re-normalise (k order)
__global__ __launch_bounds__(256) void k_pca_project(const float* __restrict__ x, int n, const float* __restrict__ mean,
                                                     const float* __restrict__ proj, int dims, float* __restrict__ y) {
    __shared__ float c[8][128];
    __shared__ float p[8][129];
    __shared__ float den[8];
    const int r0 = blockIdx.x * 8, t = threadIdx.x;
    const int nr = min(8, n - r0);
    for (int e = t; e < nr * 128; e += 256) c[e >> 7][e & 127] = x[(size_t)r0 * 128 + e] - mean[e & 127];
    __syncthreads();
    for (int e = t; e < nr * dims; e += 256) {
        const int r = e / dims, k = e - r * dims;
        float acc = 0.0f;
        for (int j = 0; j < 128; ++j) acc = acc + c[r][j] * proj[j * dims + k];
        p[r][k] = acc;
    }
    __syncthreads();
    if (t < nr) {
        float s = 0.0f;
        for (int k = 0; k < dims; ++k) s = s + p[t][k] * p[t][k];
        den[t] = sqrtf(s) + 1e-7f;
    }
    __syncthreads();
    for (int e = t; e < nr * dims; e += 256) {
        const int r = e / dims, k = e - r * dims;
        y[(size_t)(r0 + r) * dims + k] = p[r][k] / den[r];
    }
}

struct PcaPlan {
    int n, nch;
    size_t o_x, o_part, o_mean64, o_cov, o_vt, o_rpq, o_rcs, o_diag, o_nr, total;
};

PcaPlan pca_plan(int N, int M) {
    PcaPlan P;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~(size_t)255; return o; };
    P.n = N + M;
    P.nch = (P.n + kPcaChunk - 1) / kPcaChunk;
    P.o_x = take((size_t)std::max(P.n, 1) * 128 * sizeof(float));
    P.o_part = take((size_t)std::max(P.nch, 1) * 16384 * sizeof(double));
    P.o_mean64 = take(128 * sizeof(double));
    P.o_cov = take(16384 * sizeof(double));
    P.o_vt = take(16384 * sizeof(double));
    // the Jacobi rotations, at most kPcaMaxSweeps x 127 rounds of 64 pairs
    P.o_rpq = take((size_t)kPcaMaxSweeps * 127 * 64 * sizeof(int));
    P.o_rcs = take((size_t)kPcaMaxSweeps * 127 * 64 * 2 * sizeof(double));
    P.o_diag = take(128 * sizeof(double));
    P.o_nr = take(sizeof(int));
    P.total = off;
    return P;
}

}  // namespace

size_t pca_workspace(int N, int M) { return pca_plan(N, M).total; }

int l2_normalize_rows(const float* x, int n, float* y, hipStream_t s) {
    if (n < 0 || (n > 0 && (!x || !y))) { set_error("msift_l2_normalize: bad arguments"); return MSIFT_EINVAL; }
    if (n > 0) hipLaunchKernelGGL(k_l2n_rows, dim3((n + 63) / 64), dim3(256), 0, s, x, n, y, aligned16(x, y));
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

int pca_fit(const float* a, int N, const float* b, int M, int dims, void* ws_, size_t ws_bytes, float* mean,
            float* proj, double* evals, int* sweeps, hipStream_t s) {
    if (N < 0 || M < 0 || N + M < 1 || dims < 1 || dims > 128 || (N > 0 && !a) || (M > 0 && !b) || !ws_ || !mean ||
        !proj || !evals || !sweeps) {
        set_error("msift_pca_fit: bad arguments (need N + M >= 1, 1 <= dims <= 128, non-null outputs)");
        return MSIFT_EINVAL;
    }
    const PcaPlan P = pca_plan(N, M);
    if (ws_bytes < P.total) { set_error("msift_pca_fit: workspace too small"); return MSIFT_EWORKSPACE; }
    if (((uintptr_t)ws_ & 7) != 0) { set_error("msift_pca_fit: workspace must be 8-byte aligned"); return MSIFT_EINVAL; }
    char* ws = (char*)ws_;
    float* X = (float*)(ws + P.o_x);
    double* part = (double*)(ws + P.o_part);
    double* mean64 = (double*)(ws + P.o_mean64);
    double* cov = (double*)(ws + P.o_cov);
    double* vt = (double*)(ws + P.o_vt);
    if (N > 0) hipLaunchKernelGGL(k_l2n_rows, dim3((N + 63) / 64), dim3(256), 0, s, a, N, X, aligned16(a, X));
    if (M > 0) hipLaunchKernelGGL(k_l2n_rows, dim3((M + 63) / 64), dim3(256), 0, s, b, M, X + (size_t)N * 128,
                                  aligned16(b, X + (size_t)N * 128));
    hipLaunchKernelGGL(k_pca_mean_part, dim3(P.nch), dim3(128), 0, s, (const float*)X, P.n, part);
    hipLaunchKernelGGL(k_pca_mean, dim3(1), dim3(128), 0, s, (const double*)part, P.nch, P.n, mean64, mean);
    hipLaunchKernelGGL(k_pca_gram_part, dim3(P.nch), dim3(256), 0, s, (const float*)X, P.n, (const double*)mean64, part);
    hipLaunchKernelGGL(k_pca_cov, dim3(64), dim3(256), 0, s, (const double*)part, P.nch, P.n, cov);
    int* rpq = (int*)(ws + P.o_rpq);
    double* rcs = (double*)(ws + P.o_rcs);
    double* diag = (double*)(ws + P.o_diag);
    int* nr = (int*)(ws + P.o_nr);
    hipLaunchKernelGGL(k_pca_jacobi, dim3(1), dim3(1024), 0, s, (const double*)cov, rpq, rcs, nr, diag, sweeps);
    hipLaunchKernelGGL(k_pca_vt, dim3(32), dim3(256), 0, s, (const int*)rpq, (const double*)rcs, (const int*)nr, vt);
    hipLaunchKernelGGL(k_pca_finish, dim3(1), dim3(128), 0, s, (const double*)vt, (const double*)diag, dims, proj,
                       evals);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

int pca_project(const float* x, int n, const float* mean, const float* proj, int dims, float* y, hipStream_t s) {
    if (n < 0 || dims < 1 || dims > 128 || (n > 0 && (!x || !mean || !proj || !y))) {
        set_error("msift_pca_project: bad arguments (1 <= dims <= 128)");
        return MSIFT_EINVAL;
    }
    if (n > 0) hipLaunchKernelGGL(k_pca_project, dim3((n + 7) / 8), dim3(256), 0, s, x, n, mean, proj, dims, y);
    MSIFT_LAUNCH_CHECK();
    return MSIFT_OK;
}

}  // namespace msift
