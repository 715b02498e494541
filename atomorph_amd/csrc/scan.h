// Exclusive prefix sum of int counts on the device: per-block sums, one
// block scanning the block sums, then per-block scans with the block offset.
// Used to turn per-pixel and per-cell counts into segment starts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amh {
namespace scan {

constexpr int kBlock = 1024;  // elements per scan block (256 threads)

inline int64_t blocks(int64_t n) { return (n + kBlock - 1) / kBlock; }

namespace {

__global__ __launch_bounds__(256) void k_blk_sum(const int *__restrict__ in, int64_t n, int *__restrict__ sums) {
    __shared__ int part[4];
    const int64_t base = int64_t(blockIdx.x) * kBlock;
    int s = 0;
    for (int l = threadIdx.x; l < kBlock; l += 256) {
        const int64_t i = base + l;
        if (i < n) s += in[i];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// offs[b] = sums[0] + ... + sums[b-1]; offs[nb] = total
__global__ __launch_bounds__(1024) void k_scan_sums(const int *__restrict__ in, int nb, int *__restrict__ out) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? in[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) out[i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[nb] = carry;
}

// exclusive scan inside each block plus the block's offset; out[n] = total
__global__ __launch_bounds__(256) void k_blk_apply(const int *__restrict__ in, int64_t n,
                                                   const int *__restrict__ offs, int nb, int *__restrict__ out) {
    __shared__ int wsum[4];
    const int64_t base = int64_t(blockIdx.x) * kBlock;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int running = offs[blockIdx.x];
    for (int chunk = 0; chunk < kBlock / 256; ++chunk) {
        const int64_t i = base + chunk * 256 + threadIdx.x;
        const int v = i < n ? in[i] : 0;
        int incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int before = running;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (i < n) out[i] = before + incl - v;
        running += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = offs[nb];
}

// k_blk_apply for few blocks, without the scan of the block sums: each block
// adds up the sums of the blocks before it itself (at most kFoldBlocks
// values, 256 threads); the last block writes out[n] = total.
constexpr int kFoldBlocks = 2048;

__global__ __launch_bounds__(256) void k_blk_apply_fold(const int *__restrict__ in, int64_t n,
                                                        const int *__restrict__ sums, int nb, int *__restrict__ out) {
    __shared__ int wsum[4];
    const int64_t base = int64_t(blockIdx.x) * kBlock;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int o = 0;
    for (int j = threadIdx.x; j < int(blockIdx.x); j += 256) o += sums[j];
    for (int off = 32; off > 0; off >>= 1) o += __shfl_down(o, off, 64);
    if (lane == 0) wsum[wave] = o;
    __syncthreads();
    int running = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();  // wsum is reused below
    for (int chunk = 0; chunk < kBlock / 256; ++chunk) {
        const int64_t i = base + chunk * 256 + threadIdx.x;
        const int v = i < n ? in[i] : 0;
        int incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int before = running;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (i < n) out[i] = before + incl - v;
        running += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (int(blockIdx.x) == nb - 1 && threadIdx.x == 0) out[n] = running;
}

}  // namespace

// out[0..n] = exclusive scan of in[0..n) with out[n] = total.  Workspace:
// sums[blocks(n)], offs[blocks(n) + 1].  n > 0.  Up to kFoldBlocks blocks
// the block sums are not scanned by a launch of their own.
inline void exclusive(const int *in, int64_t n, int *sums, int *offs, int *out, hipStream_t s) {
    const int64_t nb = blocks(n);
    hipLaunchKernelGGL(k_blk_sum, dim3(uint32_t(nb)), dim3(256), 0, s, in, n, sums);
    if (nb <= kFoldBlocks) {
        hipLaunchKernelGGL(k_blk_apply_fold, dim3(uint32_t(nb)), dim3(256), 0, s, in, n, sums, int(nb), out);
        return;
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, int(nb), offs);
    hipLaunchKernelGGL(k_blk_apply, dim3(uint32_t(nb)), dim3(256), 0, s, in, n, offs, int(nb), out);
}

}  // namespace scan
}  // namespace amh
