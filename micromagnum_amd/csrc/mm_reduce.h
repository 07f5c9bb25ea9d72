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

// ---- compensated sums ----------------------------------------------------
// A sum carried as an unevaluated pair s + e.  dd_add adds a term with Knuth's
// TwoSum, whose rounding error is exact, into the pair (the compensated
// summation of the reference's accurate_sum, src/magneto/math/accurate_sum.h:
// 26-50, without its descending sort: the error term holds what the sort
// protects against); dd_merge adds two pairs.  The result s + e carries the
// sum to about twice the double precision before its final rounding, so the
// energies and <M> do not depend on the reduction's order or on cancellation
// between cells.
struct Dd {
    double s, e;
};

__host__ __device__ __forceinline__ void two_sum(double a, double b, double &x, double &y)
{
#pragma clang fp contract(off)
    x = a + b;
    const double z = x - a;
    y = (a - (x - z)) + (b - z);
}

__host__ __device__ __forceinline__ void dd_add(Dd &d, double v)
{
#pragma clang fp contract(off)
    double x, y;
    two_sum(d.s, v, x, y);
    d.s = x;
    d.e += y;
}

__host__ __device__ __forceinline__ Dd dd_merge(Dd a, Dd b)
{
#pragma clang fp contract(off)
    double x, y;
    two_sum(a.s, b.s, x, y);
    Dd r;
    r.s = x;
    r.e = (a.e + b.e) + y;
    return r;
}

__host__ __device__ __forceinline__ double dd_value(Dd d)
{
#pragma clang fp contract(off)
    return d.s + d.e;
}

__device__ __forceinline__ Dd wave_sum_dd(Dd v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        Dd o;
        o.s = __shfl_down(v.s, off, 64);
        o.e = __shfl_down(v.e, off, 64);
        v = dd_merge(v, o);
    }
    return v;
}

// block of MM_RED_THREADS threads; the block's pair goes to out[blockIdx.x]
// (s) and out[stride + blockIdx.x] (e)
__device__ __forceinline__ void block_sum_dd_store(Dd v, double *out, size_t stride)
{
    __shared__ double sred[2][MM_RED_THREADS / 64];
    v = wave_sum_dd(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sred[0][w] = v.s;
        sred[1][w] = v.e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Dd t = {sred[0][0], sred[1][0]};
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            const Dd o = {sred[0][q], sred[1][q]};
            t = dd_merge(t, o);
        }
        out[blockIdx.x] = t.s;
        out[stride + blockIdx.x] = t.e;
    }
    __syncthreads();
}

// the pairs (part[b], part[stride + b]), b < n, merged by one block of
// MM_RED_THREADS threads in a fixed order: out[0] = their rounded sum
static __global__ void k_reduce_final_sum_dd(const double *__restrict__ part, size_t stride, int n,
                                             double *__restrict__ out)
{
    Dd t = {0.0, 0.0};
    for (int b = threadIdx.x; b < n; b += blockDim.x) {
        const Dd o = {part[b], part[stride + b]};
        t = dd_merge(t, o);
    }
    __shared__ double sred[2][MM_RED_THREADS / 64];
    t = wave_sum_dd(t);
    if ((threadIdx.x & 63) == 0) {
        sred[0][threadIdx.x >> 6] = t.s;
        sred[1][threadIdx.x >> 6] = t.e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Dd r = {sred[0][0], sred[1][0]};
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            const Dd o = {sred[0][q], sred[1][q]};
            r = dd_merge(r, o);
        }
        out[0] = dd_value(r);
    }
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
