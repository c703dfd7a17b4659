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

// multi-workgroup path: per-tile sums, one-workgroup scan of the sums, then
// each tile rescans itself from its carry-in (tiles of kScanTile ints)
constexpr int kScanTile = 4096;

__global__ __launch_bounds__(256) void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ tsum,
                                                      const int* __restrict__ gate, int gate_below) {
    if (gate && *gate >= gate_below) return;
    __shared__ int ws4[4];
    const int base = blockIdx.x * kScanTile;
    int acc = 0;
#pragma unroll 4
    for (int k = 0; k < kScanTile / 256; ++k) {
        int i = base + k * 256 + (int)threadIdx.x;
        if (i < n) acc += in[i];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) ws4[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) tsum[blockIdx.x] = (ws4[0] + ws4[1]) + (ws4[2] + ws4[3]);
}

__global__ __launch_bounds__(256) void k_scan_apply(const int* __restrict__ in, int* __restrict__ out, int n,
                                                     const int* __restrict__ toff, int ntiles,
                                                     const int* __restrict__ gate, int gate_below) {
    if (gate && *gate >= gate_below) return;
    __shared__ int ws4[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int base = blockIdx.x * kScanTile;
    int carry = toff[blockIdx.x];
    for (int k = 0; k < kScanTile / 256; ++k) {
        const int i = base + k * 256 + (int)threadIdx.x;
        const int v = i < n ? in[i] : 0;
        int x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) ws4[wv] = x;
        __syncthreads();
        int pre = 0;
        for (int w = 0; w < wv; ++w) pre += ws4[w];
        const int tot = (ws4[0] + ws4[1]) + (ws4[2] + ws4[3]);
        if (i < n) out[i] = carry + pre + x - v;
        carry += tot;
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = toff[ntiles];
}

int scan_tmp_ints(int n) { return 2 * ((n + kScanTile - 1) / kScanTile + 1); }

void launch_scan(const int* in, int* out, int n, int* total_a, int* total_b, const int* gate, int gate_below,
                 hipStream_t s, int* tmp) {
    if (!tmp || n <= 4 * kScanTile) {
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, in, out, n, total_a, total_b, gate, gate_below);
        return;
    }
    const int nt = (n + kScanTile - 1) / kScanTile;
    int* tsum = tmp;
    int* toff = tmp + nt + 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(256), 0, s, in, n, tsum, gate, gate_below);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const int*)tsum, toff, nt, total_a, total_b, gate,
                       gate_below);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(256), 0, s, in, out, n, (const int*)toff, nt, gate, gate_below);
}

}  // namespace msift
