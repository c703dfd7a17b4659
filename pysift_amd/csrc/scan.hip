// Device-wide exclusive scan of int32 counts in one workgroup: tiles of
// 4096 (4 per thread), wave scans by shuffles, a carry across tiles.
#include "msift_internal.hpp"

namespace msift {

// out[i] = sum(in[0..i)), out[n] = total; optional copies of the total.
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ in, int* __restrict__ out, int n,
                                                int* __restrict__ total_a, int* __restrict__ total_b,
                                                const int* __restrict__ gate, int gate_below) {
    if (gate && *gate >= gate_below) return;
    __shared__ int wsum[16];
    __shared__ int tile_tot;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int carry = 0;
    for (int base = 0; base < n; base += 4096) {
        const int i0 = base + 4 * tid;
        int v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (i0 + k < n) ? in[i0 + k] : 0;
        const int c = (v[0] + v[1]) + (v[2] + v[3]);
        int x = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        if (wv == 0) {
            int t = lane < 16 ? wsum[lane] : 0;
            int xs = t;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                int y = __shfl_up(xs, off);
                if (lane >= off) xs += y;
            }
            if (lane < 16) wsum[lane] = xs - t;
            if (lane == 15) tile_tot = xs;
        }
        __syncthreads();
        int e = carry + wsum[wv] + (x - c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < n) out[i0 + k] = e;
            e += v[k];
        }
        carry += tile_tot;
        __syncthreads();
    }
    if (tid == 0) {
        out[n] = carry;
        if (total_a) *total_a = carry;
        if (total_b) *total_b = carry;
    }
}

void launch_scan(const int* in, int* out, int n, int* total_a, int* total_b, const int* gate, int gate_below,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, in, out, n, total_a, total_b, gate, gate_below);
}

}  // namespace msift
