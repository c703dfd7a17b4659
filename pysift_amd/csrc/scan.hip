// Device-wide exclusive scan of int32 counts in one workgroup (counts are
// small: extrema blocks, compaction blocks, per-keypoint angle counts).
#include "msift_internal.hpp"

namespace msift {

// out[i] = sum(in[0..i)), out[n] = total; optional copies of the total.
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ in, int* __restrict__ out, int n,
                                                int* __restrict__ total_a, int* __restrict__ total_b,
                                                const int* __restrict__ gate, int gate_below) {
    if (gate && *gate >= gate_below) return;
    __shared__ int sm[1024];
    const int tid = threadIdx.x;
    // blocked layout: 1024 contiguous chunks, read with a strided (coalesced) sweep
    const int chunk = (n + 1023) / 1024;
    const int beg = tid * chunk;
    const int end = min(n, beg + chunk);
    int s = 0;
    for (int i = beg; i < end; ++i) s += in[i];
    sm[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int t = tid >= off ? sm[tid - off] : 0;
        __syncthreads();
        sm[tid] += t;
        __syncthreads();
    }
    int run = sm[tid] - s;
    for (int i = beg; i < end; ++i) {
        int v = in[i];
        out[i] = run;
        run += v;
    }
    if (tid == 1023) {
        out[n] = sm[1023];
        if (total_a) *total_a = sm[1023];
        if (total_b) *total_b = sm[1023];
    }
}

void launch_scan(const int* in, int* out, int n, int* total_a, int* total_b, const int* gate, int gate_below,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, in, out, n, total_a, total_b, gate, gate_below);
}

}  // namespace msift
