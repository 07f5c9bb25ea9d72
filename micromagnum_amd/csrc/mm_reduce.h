// mm_reduce.h -- deterministic FP64 block reductions (wave64 shuffles, then
// one LDS step across the block's waves) and the single-block final passes.
#ifndef MM_REDUCE_H
#define MM_REDUCE_H

#include <hip/hip_runtime.h>

#define MM_RED_THREADS 256
#define MM_RED_BLOCKS 1024

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// block of MM_RED_THREADS threads; result written to out[blockIdx.x] by thread 0
__device__ __forceinline__ void block_sum_store(double v, double *out)
{
    __shared__ double sred[MM_RED_THREADS / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sred[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += sred[q];
        out[blockIdx.x] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ void block_max_store(double v, double *out)
{
    __shared__ double sred[MM_RED_THREADS / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sred[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = sred[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) s = fmax(s, sred[q]);
        out[blockIdx.x] = s;
    }
    __syncthreads();
}

static __global__ void k_reduce_final_sum(const double *__restrict__ part, int n, double *__restrict__ out)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    __shared__ double sred[MM_RED_THREADS / 64];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sred[q];
        out[0] = t;
    }
}

static __global__ void k_reduce_final_max(const double *__restrict__ part, int n, double *__restrict__ out)
{
    double s = -1.0e308;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s = fmax(s, part[i]);
    __shared__ double sred[MM_RED_THREADS / 64];
    s = wave_max(s);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sred[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) t = fmax(t, sred[q]);
        out[0] = t;
    }
}

#endif
