// mm_local.hip -- standalone local field terms, LLG, Zhang-Li, reductions,
// normalisation and Runge-Kutta stage algebra (the reference's per-call
// magneto functions), one thread per cell, SoA arrays in HBM.
//
// Reductions are deterministic: each block reduces its grid-stride slice in
// FP64 and writes one partial; a single-block pass combines the partials in
// a fixed order.  Scalars are returned through a small per-process device
// scratch guarded by a mutex (calls that return a scalar synchronise).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "mm_cell.h"
#include "mm_conv.h"
#include "mm_reduce.h"

// ---------------------------------------------------------------------------
// reduction scratch
// ---------------------------------------------------------------------------
static std::mutex g_red_mtx;
static double *g_red_partial = nullptr; // MM_RED_BLOCKS * 6: up to three compensated pairs
static double *g_red_out = nullptr;     // 4

static void red_scratch()
{
    if (!g_red_partial) {
        MM_HIP(hipMalloc(&g_red_partial, sizeof(double) * MM_RED_BLOCKS * 6));
        MM_HIP(hipMalloc(&g_red_out, sizeof(double) * 4));
    }
}

static int red_grid(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>(MM_RED_BLOCKS, (n + 255) / 256)); }

// ---------------------------------------------------------------------------
// scratch of the allocation-free entry point (mm_scaled_abs_max_noalloc):
// the per-block partial maxima, one per block of at most MM_RED_BLOCKS
// blocks, in a file-scope device array that every GPU the library runs on
// holds until the process ends (8 KiB).  Its name is a host-side stand-in in
// host code: kernels get the address hipGetSymbolAddress returns on the
// device, looked up once per device (noalloc_scratch) with that device
// current.  Calls on one GPU share the array: calls that overlap on one GPU
// race.
// ---------------------------------------------------------------------------
#define MM_NOALLOC_PARTIALS MM_RED_BLOCKS
#define MM_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[MM_NOALLOC_PARTIALS];
static std::mutex g_scratch_mtx;
static double *g_scratch_ptr[MM_NOALLOC_MAX_DEVICES];

// the current device's address of `scratch` (looked up on the first call
// per device)
static double *noalloc_scratch()
{
    int dev = -1;
    MM_HIP(hipGetDevice(&dev));
    MM_REQUIRE(dev >= 0 && dev < MM_NOALLOC_MAX_DEVICES, "device index beyond the scratch address table");
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        MM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        MM_REQUIRE(ptr != nullptr, "no device address for the scratch array");
        g_scratch_ptr[dev] = (double *)ptr;
    }
    return g_scratch_ptr[dev];
}

// set-up from mm_init with the device current
void mm_noalloc_setup() { (void)noalloc_scratch(); }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_exchange(MeshDev m, T wx, T wy, T wz, T coef, const T *__restrict__ Ms, const T *__restrict__ A,
                           const T *__restrict__ M, T *__restrict__ H, double *__restrict__ epart)
{
    const size_t n = (size_t)m.nx * m.ny * m.nz;
    Dd e = {0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % m.nx);
        const int y = (int)((i / m.nx) % m.ny);
        const int z = (int)(i / ((size_t)m.nx * m.ny));
        const T ms = Ms[i];
        if (ms == (T)0) {
            store3<T>(H, n, i, v3<T>(0, 0, 0));
            continue;
        }
        const V3<T> h = exchange_cell<T>(m, x, y, z, i, n, wx, wy, wz, coef, Ms, A, M, ms);
        store3<T>(H, n, i, h);
        if (epart)
            dd_add(e, (double)(M[i] / ms) * h.x + (double)(M[n + i] / ms) * h.y + (double)(M[2 * n + i] / ms) * h.z);
    }
    if (epart) block_sum_dd_store(e, epart, MM_RED_BLOCKS);
}

template <typename T>
__global__ void k_uniaxial(size_t n, const T *__restrict__ axis, const T *__restrict__ k, const T *__restrict__ Ms,
                           const T *__restrict__ M, T *__restrict__ H, T mu0, double *__restrict__ epart)
{
    Dd e = {0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        T ei;
        const V3<T> h = uniaxial_cell<T>(load3(M, n, i), load3(axis, n, i), k[i], Ms[i], mu0, &ei);
        store3<T>(H, n, i, h);
        dd_add(e, (double)ei);
    }
    if (epart) block_sum_dd_store(e, epart, MM_RED_BLOCKS);
}

template <typename T>
__global__ void k_cubic(size_t n, const T *__restrict__ a1, const T *__restrict__ a2, const T *__restrict__ k,
                        const T *__restrict__ Ms, const T *__restrict__ M, T *__restrict__ H, T cm2mu0,
                        double *__restrict__ epart)
{
    Dd e = {0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        T ei;
        const V3<T> h = cubic_cell<T>(load3(M, n, i), load3(a1, n, i), load3(a2, n, i), k[i], Ms[i], cm2mu0, &ei);
        store3<T>(H, n, i, h);
        dd_add(e, (double)ei);
    }
    if (epart) block_sum_dd_store(e, epart, MM_RED_BLOCKS);
}

template <typename T>
__global__ void k_llge(size_t n, const T *__restrict__ f1, const T *__restrict__ f2, const T *__restrict__ M,
                       const T *__restrict__ H, T *__restrict__ dM)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        store3<T>(dM, n, i, llge_cell<T>(load3(M, n, i), load3(H, n, i), f1[i], f2[i]));
}

template <typename T>
__global__ void k_zhangli(MeshDev m, T dx, T dy, T dz, int do_precess, const T *__restrict__ P,
                          const T *__restrict__ xi, const T *__restrict__ Ms, const T *__restrict__ alpha,
                          const T *__restrict__ j, const T *__restrict__ M, T *__restrict__ dM, T mu_b, T e_charge)
{
    const size_t n = (size_t)m.nx * m.ny * m.nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % m.nx);
        const int y = (int)((i / m.nx) % m.ny);
        const int z = (int)(i / ((size_t)m.nx * m.ny));
        store3<T>(dM, n, i,
                  zhangli_cell<T>(m, x, y, z, i, n, dx, dy, dz, do_precess != 0, P[i], xi[i], Ms[i], alpha[i],
                                  load3(j, n, i), M, mu_b, e_charge));
    }
}

// max over cells of |M|^2 (absmax3) or |M|^2/s^2 for s != 0 (scaled)
template <typename T>
__global__ void k_maxsq(size_t n, const T *__restrict__ M, const T *__restrict__ s, double *__restrict__ part)
{
    double mx = s ? 0.0 : -1.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double a = M[i], b = M[n + i], c = M[2 * n + i];
        double q = a * a + b * b + c * c;
        if (s) {
            const double si = s[i];
            if (si == 0.0) continue;
            q = q / (si * si);
        }
        mx = fmax(mx, q);
    }
    block_max_store(mx, part);
}

template <typename T>
__global__ void k_dot(size_t n, const T *__restrict__ A, const T *__restrict__ B, double *__restrict__ part)
{
    Dd acc = {0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        dd_add(acc, (double)A[i] * B[i]);
        dd_add(acc, (double)A[n + i] * B[n + i]);
        dd_add(acc, (double)A[2 * n + i] * B[2 * n + i]);
    }
    block_sum_dd_store(acc, part, MM_RED_BLOCKS);
}

template <typename T> __global__ void k_sum3(size_t n, const T *__restrict__ A, double *__restrict__ part)
{
    Dd s0 = {0.0, 0.0}, s1 = {0.0, 0.0}, s2 = {0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        dd_add(s0, (double)A[i]);
        dd_add(s1, (double)A[n + i]);
        dd_add(s2, (double)A[2 * n + i]);
    }
    // component c's pair at part[2c MM_RED_BLOCKS + b], part[(2c+1) MM_RED_BLOCKS + b]
    block_sum_dd_store(s0, part, MM_RED_BLOCKS);
    block_sum_dd_store(s1, part + 2 * MM_RED_BLOCKS, MM_RED_BLOCKS);
    block_sum_dd_store(s2, part + 4 * MM_RED_BLOCKS, MM_RED_BLOCKS);
}

// CPUDevice.cpp:244-271 (mode 0) or FP64 norm (mode 1)
template <typename T>
__global__ void k_normalize(size_t n, T *__restrict__ M, const T *__restrict__ len, int mode)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double norm;
        if (mode == 0) {
            norm = float_cast_norm((double)M[i], (double)M[n + i], (double)M[2 * n + i]);
        } else {
            norm = fp64_norm((double)M[i], (double)M[n + i], (double)M[2 * n + i]);
        }
        if (norm == 0.0) continue;
        const double sc = (double)len[i] / norm;
        M[i] = (T)(M[i] * sc);
        M[n + i] = (T)(M[n + i] * sc);
        M[2 * n + i] = (T)(M[2 * n + i] * sc);
    }
}

struct RKTerms {
    const void *k[8];
    double f[8];
    int nk;
};

// out = y + h * sum_j f_j k_j   (runge_kutta_cpu.cpp:28-107 arithmetic)
template <typename T>
__global__ void k_rk_axpy_sum(size_t m, const T *__restrict__ y, T *__restrict__ out, T h, RKTerms t)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
        T s = (T)t.f[0] * ((const T *)t.k[0])[i];
        for (int q = 1; q < t.nk; ++q) s += (T)t.f[q] * ((const T *)t.k[q])[i];
        out[i] = y[i] + h * s;
    }
}

// out = y; out += k_j * f_j sequentially (runge_kutta.py:107-111 arithmetic)
template <typename T>
__global__ void k_rk_seq(size_t m, const T *__restrict__ y, T *__restrict__ out, RKTerms t)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
        T v = y[i];
        for (int q = 0; q < t.nk; ++q) v += ((const T *)t.k[q])[i] * (T)t.f[q];
        out[i] = v;
    }
}

// y = y + h*sum c k ; err = h*sum ec k   (runge_kutta_cpu.cpp:139-194)
template <typename T>
__global__ void k_rk_combine6(size_t m, T *__restrict__ y, T *__restrict__ err, T h, RKTerms c, RKTerms ec)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
        T s = (T)c.f[0] * ((const T *)c.k[0])[i];
        T e = (T)ec.f[0] * ((const T *)ec.k[0])[i];
        for (int q = 1; q < c.nk; ++q) {
            s += (T)c.f[q] * ((const T *)c.k[q])[i];
            e += (T)ec.f[q] * ((const T *)ec.k[q])[i];
        }
        y[i] = y[i] + h * s;
        err[i] = h * e;
    }
}

// err = 0; y += k_j*(h c_j); err += k_j*(h ec_j)   (runge_kutta.py:125-129)
template <typename T>
__global__ void k_rk_combine_seq(size_t m, T *__restrict__ y, T *__restrict__ err, RKTerms c, RKTerms ec)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
        T v = y[i], e = (T)0;
        for (int q = 0; q < c.nk; ++q) {
            const T kq = ((const T *)c.k[q])[i];
            v += kq * (T)c.f[q];
            e += kq * (T)ec.f[q];
        }
        y[i] = v;
        err[i] = e;
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static int grid_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 16384)); }

// part: compensated pairs (part[b], part[MM_RED_BLOCKS + b]) of nblocks blocks
static double finish_sum(hipStream_t s, double *part, int nblocks)
{
    hipLaunchKernelGGL(k_reduce_final_sum_dd, dim3(1), dim3(MM_RED_THREADS), 0, s, part, (size_t)MM_RED_BLOCKS,
                       nblocks, g_red_out);
    MM_HIP(hipGetLastError());
    double v = 0.0;
    MM_HIP(hipMemcpyAsync(&v, g_red_out, sizeof(double), hipMemcpyDeviceToHost, s));
    MM_HIP(hipStreamSynchronize(s));
    return v;
}

static double finish_max(hipStream_t s, double *part, int nblocks)
{
    hipLaunchKernelGGL(k_reduce_final_max, dim3(1), dim3(MM_RED_THREADS), 0, s, part, nblocks, g_red_out);
    MM_HIP(hipGetLastError());
    double v = 0.0;
    MM_HIP(hipMemcpyAsync(&v, g_red_out, sizeof(double), hipMemcpyDeviceToHost, s));
    MM_HIP(hipStreamSynchronize(s));
    return v;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int mm_exchange(int nx, int ny, int nz, double dx, double dy, double dz, int px, int py, int pz,
                           const double *Ms, const double *A, const double *M, double *H, double *energy,
                           void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad mesh size");
    MM_REQUIRE(Ms && A && M && H, "null array");
    const MeshDev m = {nx, ny, nz, px != 0, py != 0, pz != 0};
    const size_t n = (size_t)nx * ny * nz;
    hipStream_t s = (hipStream_t)stream;
    const double mu0 = MM_MU0_VALUE;
    std::unique_lock<std::mutex> lk(g_red_mtx, std::defer_lock);
    double *part = nullptr;
    int grid = grid_for(n);
    if (energy) {
        lk.lock();
        red_scratch();
        part = g_red_partial;
        grid = red_grid(n);
    }
    hipLaunchKernelGGL(k_exchange<double>, dim3(grid), dim3(256), 0, s, m, 1.0 / (dx * dx), 1.0 / (dy * dy),
                       1.0 / (dz * dz), 2.0 / mu0, Ms, A, M, H, part);
    MM_HIP(hipGetLastError());
    if (energy) *energy = finish_sum(s, part, grid) * (-mu0 / 2.0 * dx * dy * dz);
    MM_CAPI_END
}

extern "C" int mm_exchange_f32(int nx, int ny, int nz, double dx, double dy, double dz, int px, int py, int pz,
                               const float *Ms, const float *A, const float *M, float *H, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad mesh size");
    MM_REQUIRE(Ms && A && M && H, "null array");
    const MeshDev m = {nx, ny, nz, px != 0, py != 0, pz != 0};
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_exchange<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, m,
                       (float)(1.0 / (dx * dx)), (float)(1.0 / (dy * dy)), (float)(1.0 / (dz * dz)),
                       (float)(2.0 / MM_MU0_VALUE), Ms, A, M, H, (double *)nullptr);
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_uniaxial_anisotropy(int n, const double *axis, const double *k, const double *Ms,
                                      const double *M, double *H, double *energy, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && axis && k && Ms && M && H, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    std::unique_lock<std::mutex> lk(g_red_mtx, std::defer_lock);
    double *part = nullptr;
    int grid = grid_for(n);
    if (energy) {
        lk.lock();
        red_scratch();
        part = g_red_partial;
        grid = red_grid(n);
    }
    hipLaunchKernelGGL(k_uniaxial<double>, dim3(grid), dim3(256), 0, s, (size_t)n, axis, k, Ms, M, H,
                       MM_MU0_VALUE, part);
    MM_HIP(hipGetLastError());
    if (energy) *energy = finish_sum(s, part, grid);
    MM_CAPI_END
}

extern "C" int mm_cubic_anisotropy(int n, const double *axis1, const double *axis2, const double *k,
                                   const double *Ms, const double *M, double *H, double *energy, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && axis1 && axis2 && k && Ms && M && H, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    std::unique_lock<std::mutex> lk(g_red_mtx, std::defer_lock);
    double *part = nullptr;
    int grid = grid_for(n);
    if (energy) {
        lk.lock();
        red_scratch();
        part = g_red_partial;
        grid = red_grid(n);
    }
    hipLaunchKernelGGL(k_cubic<double>, dim3(grid), dim3(256), 0, s, (size_t)n, axis1, axis2, k, Ms, M, H,
                       -2.0 / MM_MU0_VALUE, part);
    MM_HIP(hipGetLastError());
    if (energy) *energy = finish_sum(s, part, grid);
    MM_CAPI_END
}

extern "C" int mm_llge(int n, const double *f1, const double *f2, const double *M, const double *H, double *dM,
                       void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && f1 && f2 && M && H && dM, "bad argument");
    hipLaunchKernelGGL(k_llge<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (size_t)n, f1, f2, M,
                       H, dM);
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_fdm_zhangli(int nx, int ny, int nz, double dx, double dy, double dz, int do_precess,
                              const double *P, const double *xi, const double *Ms, const double *alpha,
                              const double *j, const double *M, double *dM, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad mesh size");
    MM_REQUIRE(P && xi && Ms && alpha && j && M && dM, "null array");
    const MeshDev m = {nx, ny, nz, 0, 0, 0};
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_zhangli<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, m, dx, dy, dz,
                       do_precess, P, xi, Ms, alpha, j, M, dM, MM_MU_BOHR, MM_E_CHARGE);
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_scaled_abs_max(int n, const double *M, const double *scale, int scale_uniform, double *out,
                                 void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && M && scale && out, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_red_mtx);
    red_scratch();
    const int grid = red_grid(n);
    if (scale_uniform) {
        double s0 = 0.0;
        MM_HIP(hipMemcpyAsync(&s0, scale, sizeof(double), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_maxsq<double>, dim3(grid), dim3(MM_RED_THREADS), 0, s, (size_t)n, M,
                           (const double *)nullptr, g_red_partial);
        MM_HIP(hipGetLastError());
        const double mx = finish_max(s, g_red_partial, grid);
        *out = std::sqrt(mx) / s0;
    } else {
        hipLaunchKernelGGL(k_maxsq<double>, dim3(grid), dim3(MM_RED_THREADS), 0, s, (size_t)n, M, scale,
                           g_red_partial);
        MM_HIP(hipGetLastError());
        *out = std::sqrt(finish_max(s, g_red_partial, grid));
    }
    MM_CAPI_END
}

// out[0] = sqrt(max of the partials) (/ scale[0] for a uniform scale): the
// final step of mm_scaled_abs_max on the device
static __global__ void k_scaled_abs_max_final(const double *__restrict__ part, int nb,
                                              const double *__restrict__ scale, int uniform,
                                              double *__restrict__ out)
{
    double s = -1.0e308;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s = fmax(s, part[i]);
    __shared__ double sred[MM_RED_THREADS / 64];
    s = wave_max(s);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sred[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) t = fmax(t, sred[q]);
        t = sqrt(t);
        out[0] = uniform ? t / scale[0] : t;
    }
}

extern "C" int mm_scaled_abs_max_noalloc(int n, const double *M, const double *scale, int scale_uniform,
                                         double *out_dev, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && M && scale && out_dev, "bad argument");
    const int grid = red_grid(n);
    MM_REQUIRE(grid <= MM_NOALLOC_PARTIALS, "more blocks than the scratch array holds partial maxima");
    hipStream_t s = (hipStream_t)stream;
    double *part = noalloc_scratch();
    hipLaunchKernelGGL(k_maxsq<double>, dim3(grid), dim3(MM_RED_THREADS), 0, s, (size_t)n, M,
                       scale_uniform ? (const double *)nullptr : scale, part);
    MM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scaled_abs_max_final, dim3(1), dim3(MM_RED_THREADS), 0, s, (const double *)part, grid,
                       scale, scale_uniform, out_dev);
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_abs_max(int n, const double *M, double *out, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && M && out, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_red_mtx);
    red_scratch();
    const int grid = red_grid(n);
    hipLaunchKernelGGL(k_maxsq<double>, dim3(grid), dim3(MM_RED_THREADS), 0, s, (size_t)n, M,
                       (const double *)nullptr, g_red_partial);
    MM_HIP(hipGetLastError());
    *out = std::sqrt(finish_max(s, g_red_partial, grid));
    MM_CAPI_END
}

extern "C" int mm_dot_sum(int n, const double *A, const double *B, double *out, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && A && B && out, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_red_mtx);
    red_scratch();
    const int grid = red_grid(n);
    hipLaunchKernelGGL(k_dot<double>, dim3(grid), dim3(MM_RED_THREADS), 0, s, (size_t)n, A, B, g_red_partial);
    MM_HIP(hipGetLastError());
    *out = finish_sum(s, g_red_partial, grid);
    MM_CAPI_END
}

extern "C" int mm_average(int n, const double *M, double *out3, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && M && out3, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_red_mtx);
    red_scratch();
    const int grid = red_grid(n);
    hipLaunchKernelGGL(k_sum3<double>, dim3(grid), dim3(MM_RED_THREADS), 0, s, (size_t)n, M, g_red_partial);
    MM_HIP(hipGetLastError());
    for (int c = 0; c < 3; ++c) out3[c] = finish_sum(s, g_red_partial + 2 * c * MM_RED_BLOCKS, grid) / n;
    MM_CAPI_END
}

extern "C" int mm_normalize(int n, double *M, const double *len, int mode, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(n > 0 && M && len, "bad argument");
    MM_REQUIRE(mode == 0 || mode == 1, "mode must be 0 or 1");
    hipLaunchKernelGGL(k_normalize<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (size_t)n, M, len,
                       mode);
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_rk_prepare_step(int step, double h, const mm_tableau_t *tab, int n, const double *const *k,
                                  const double *y, double *ytmp, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(tab && k && y && ytmp && n > 0, "bad argument");
    MM_REQUIRE(step >= 1 && step < tab->num_steps, "step out of range");
    RKTerms t;
    t.nk = step;
    for (int q = 0; q < step; ++q) {
        MM_REQUIRE(k[q] != nullptr, "missing stage vector");
        t.k[q] = k[q];
    }
    const size_t m = 3 * (size_t)n;
    if (tab->num_steps == 6) {
        for (int q = 0; q < step; ++q) t.f[q] = tab->b[step][q];
        hipLaunchKernelGGL(k_rk_axpy_sum<double>, dim3(grid_for(m)), dim3(256), 0, (hipStream_t)stream, m, y, ytmp,
                           h, t);
    } else {
        for (int q = 0; q < step; ++q) t.f[q] = h * tab->b[step][q];
        hipLaunchKernelGGL(k_rk_seq<double>, dim3(grid_for(m)), dim3(256), 0, (hipStream_t)stream, m, y, ytmp, t);
    }
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_rk_combine_result(double h, const mm_tableau_t *tab, int n, const double *const *k, double *y,
                                    double *y_err, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(tab && k && y && y_err && n > 0, "bad argument");
    const size_t m = 3 * (size_t)n;
    RKTerms c, ec;
    c.nk = ec.nk = 0;
    if (tab->num_steps == 6) {
        const bool skip1 = tab->c[1] == 0.0 && tab->ec[1] == 0.0;
        for (int q = 0; q < 6; ++q) {
            if (q == 1 && skip1) continue;
            MM_REQUIRE(k[q] != nullptr, "missing stage vector");
            c.k[c.nk] = k[q];
            c.f[c.nk++] = tab->c[q];
            ec.k[ec.nk] = k[q];
            ec.f[ec.nk++] = tab->ec[q];
        }
        hipLaunchKernelGGL(k_rk_combine6<double>, dim3(grid_for(m)), dim3(256), 0, (hipStream_t)stream, m, y, y_err,
                           h, c, ec);
    } else {
        for (int q = 0; q < tab->num_steps; ++q) {
            MM_REQUIRE(k[q] != nullptr, "missing stage vector");
            c.k[c.nk] = k[q];
            c.f[c.nk++] = h * tab->c[q];
            ec.k[ec.nk] = k[q];
            ec.f[ec.nk++] = h * tab->ec[q];
        }
        hipLaunchKernelGGL(k_rk_combine_seq<double>, dim3(grid_for(m)), dim3(256), 0, (hipStream_t)stream, m, y,
                           y_err, c, ec);
    }
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}
