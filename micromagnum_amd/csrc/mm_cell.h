// mm_cell.h -- per-cell physics shared by the standalone kernels (mm_local.hip)
// and the fused stage kernels (mm_stepper.hip).
#ifndef MM_CELL_H
#define MM_CELL_H

#include <hip/hip_runtime.h>

// A pointer read from device memory (a parameter record) is generic to the
// compiler, which then emits flat loads and stores (they count against the
// LDS counter too, so waits drain both).  gfield reads such a record field as
// a global-address-space pointer, so address-space inference turns every
// access through it into a global load or store.  (A cast of the loaded
// generic pointer to global and back is folded away and does not do this.)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wincompatible-pointer-types-discards-qualifiers"
template <typename T> __device__ __forceinline__ T *gfield(T *const &field)
{
    typedef __attribute__((address_space(1))) T *gptr;
    return (T *)*reinterpret_cast<const gptr *>(&field);
}
#pragma clang diagnostic pop

template <typename T> struct V3 {
    T x, y, z;
};

template <typename T> __device__ __forceinline__ V3<T> v3(T a, T b, T c)
{
    V3<T> r;
    r.x = a;
    r.y = b;
    r.z = c;
    return r;
}

template <typename T> __device__ __forceinline__ V3<T> cross3(const V3<T> &a, const V3<T> &b)
{
    return v3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

template <typename T> __device__ __forceinline__ T dot3(const V3<T> &a, const V3<T> &b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

template <typename T> __device__ __forceinline__ V3<T> load3(const T *__restrict__ a, size_t n, size_t i)
{
    return v3<T>(a[i], a[n + i], a[2 * n + i]);
}

template <typename T> __device__ __forceinline__ void store3(T *__restrict__ a, size_t n, size_t i, const V3<T> &v)
{
    a[i] = v.x;
    a[n + i] = v.y;
    a[2 * n + i] = v.z;
}

// |v| as CPUDevice::normalize3 computes it
// (src/magneto/matrix/device/cpu/CPUDevice.cpp:255-259): the components cast
// to float, every product and sum rounded to float with no contraction, and
// std::sqrt of that float expression, i.e. the float overload, so the norm
// itself is a float.  The square root is taken in double and rounded to
// float: double rounding is harmless for sqrt when 53 >= 2*24 + 2 bits, so
// this equals the correctly rounded float square root whatever the
// compiler's single-precision sqrt options are.
__device__ __forceinline__ double float_cast_norm(double a, double b, double c)
{
#pragma clang fp contract(off)
    const float e0 = (float)a, e1 = (float)b, e2 = (float)c;
    const float s = e0 * e0 + e1 * e1 + e2 * e2;
    return (double)(float)sqrt((double)s);
}

// |v| in FP64, products and sums rounded separately (no contraction), as
// the CPU path's sqrt(a*a + b*b + c*c)
__device__ __forceinline__ double fp64_norm(double a, double b, double c)
{
#pragma clang fp contract(off)
    return sqrt(a * a + b * b + c * c);
}

struct MeshDev {
    int nx, ny, nz;
    int px, py, pz;
};

// Exchange field at cell (x,y,z): exchange_cpu.cpp:60-235.  Neighbours with
// Ms == 0 or outside an open boundary are skipped; periodic axes wrap.
// coef = 2/mu0; w = 1/delta^2 per axis.
template <typename T>
__device__ __forceinline__ V3<T> exchange_cell(const MeshDev &m, int x, int y, int z, size_t i, size_t n,
                                               T wx, T wy, T wz, T coef, const T *__restrict__ Ms,
                                               const T *__restrict__ A, const T *__restrict__ M, T ms_i)
{
    const V3<T> mi = v3<T>(M[i] / ms_i, M[n + i] / ms_i, M[2 * n + i] / ms_i);
    V3<T> s = v3<T>(0, 0, 0);
    const long sx = 1, sy = m.nx, sz = (long)m.nx * m.ny;
    auto add = [&](long j, T w) {
        const T msj = Ms[j];
        if (msj != (T)0) {
            s.x += (M[j] / msj - mi.x) * w;
            s.y += (M[n + j] / msj - mi.y) * w;
            s.z += (M[2 * n + j] / msj - mi.z) * w;
        }
    };
    const long li = (long)i;
    if (x > 0) add(li - sx, wx);
    else if (m.px) add(li - sx + sx * m.nx, wx);
    if (x < m.nx - 1) add(li + sx, wx);
    else if (m.px) add(li + sx - sx * m.nx, wx);
    if (y > 0) add(li - sy, wy);
    else if (m.py) add(li - sy + sy * m.ny, wy);
    if (y < m.ny - 1) add(li + sy, wy);
    else if (m.py) add(li + sy - sy * m.ny, wy);
    if (z > 0) add(li - sz, wz);
    else if (m.pz) add(li - sz + sz * m.nz, wz);
    if (z < m.nz - 1) add(li + sz, wz);
    else if (m.pz) add(li + sz - sz * m.nz, wz);
    const T f = coef * A[i];
    return v3<T>(f * s.x / ms_i, f * s.y / ms_i, f * s.z / ms_i);
}

// uniaxial anisotropy: anisotropy_cpu.cpp:29-73
template <typename T>
__device__ __forceinline__ V3<T> uniaxial_cell(const V3<T> &Mv, const V3<T> &ax, T k, T ms, T mu0, T *energy)
{
    if (ms == (T)0 || k == (T)0) {
        *energy = 0;
        return v3<T>(0, 0, 0);
    }
    const V3<T> sp = v3<T>(Mv.x / ms, Mv.y / ms, Mv.z / ms);
    const T d = dot3(sp, ax);
    const T f = (T)2 * k * d / ms / mu0;
    const V3<T> c = cross3(ax, sp);
    *energy = k * dot3(c, c);
    return v3<T>(f * ax.x, f * ax.y, f * ax.z);
}

// cubic anisotropy: anisotropy_cpu.cpp:75-119 (m = M/|M|, OOMMF form);
// cm2mu0 = -2/mu0
template <typename T>
__device__ __forceinline__ V3<T> cubic_cell(const V3<T> &Mv, const V3<T> &u1, const V3<T> &u2, T k, T ms,
                                            T cm2mu0, T *energy)
{
    if (ms == (T)0) {
        *energy = 0;
        return v3<T>(0, 0, 0);
    }
    const T len = sqrt(dot3(Mv, Mv));
    const T sc = len == (T)0 ? (T)1 : (T)1 / len;
    const V3<T> mm = v3<T>(Mv.x * sc, Mv.y * sc, Mv.z * sc);
    const V3<T> u3 = cross3(u1, u2);
    const T a1 = dot3(u1, mm), a2 = dot3(u2, mm), a3 = dot3(u3, mm);
    const T a1s = a1 * a1, a2s = a2 * a2, a3s = a3 * a3;
    const T f = cm2mu0 * k / ms;
    const T c1 = (a2s + a3s) * a1, c2 = (a1s + a3s) * a2, c3 = (a1s + a2s) * a3;
    *energy = k * (a1s * a2s + a1s * a3s + a2s * a3s);
    return v3<T>(f * (c1 * u1.x + c2 * u2.x + c3 * u3.x), f * (c1 * u1.y + c2 * u2.y + c3 * u3.y),
                 f * (c1 * u1.z + c2 * u2.z + c3 * u3.z));
}

// LLG right-hand side: llge_cpu.cpp:25-45
template <typename T> __device__ __forceinline__ V3<T> llge_cell(const V3<T> &Mv, const V3<T> &H, T f1, T f2)
{
    const V3<T> a = cross3(Mv, H);
    const V3<T> b = cross3(Mv, a);
    return v3<T>(f1 * a.x + f2 * b.x, f1 * a.y + f2 * b.y, f1 * a.z + f2 * b.z);
}

// Zhang-Li torque: fdm_zhangli_cpu.cpp:95-200.  Gradients: central inside,
// one-sided at the faces, zero along axes of one cell.
template <typename T>
__device__ __forceinline__ V3<T> zhangli_cell(const MeshDev &m, int x, int y, int z, size_t i, size_t n,
                                              T dx, T dy, T dz, bool do_precess, T P, T xi, T ms, T alpha,
                                              const V3<T> &j, const T *__restrict__ M, T mu_b, T e_charge)
{
    if (ms == (T)0) return v3<T>(0, 0, 0);
    const T bj = P * mu_b / (e_charge * ms * ((T)1 + xi * xi));
    const T bjp = bj / ((T)1 + alpha * alpha);
    T f1, f2;
    if (do_precess) {
        f1 = -bjp / (ms * ms) * ((T)1 + alpha * xi);
        f2 = -bjp / ms * (xi - alpha);
    } else {
        f1 = -bjp / (ms * ms) * (alpha * xi);
        f2 = -bjp / ms * (-alpha);
    }
    const long li = (long)i;
    auto grad = [&](int pos, int dim, long st, T del) -> V3<T> {
        if (dim <= 1) return v3<T>(0, 0, 0);
        if (pos == 0) {
            const long b = li + st;
            return v3<T>((M[b] - M[li]) / del, (M[n + b] - M[n + li]) / del, (M[2 * n + b] - M[2 * n + li]) / del);
        }
        if (pos == dim - 1) {
            const long a = li - st;
            return v3<T>((M[li] - M[a]) / del, (M[n + li] - M[n + a]) / del, (M[2 * n + li] - M[2 * n + a]) / del);
        }
        const long a = li - st, b = li + st;
        const T d2 = (T)2 * del;
        return v3<T>((M[b] - M[a]) / d2, (M[n + b] - M[n + a]) / d2, (M[2 * n + b] - M[2 * n + a]) / d2);
    };
    const V3<T> gx = grad(x, m.nx, 1, dx);
    const V3<T> gy = grad(y, m.ny, m.nx, dy);
    const V3<T> gz = grad(z, m.nz, (long)m.nx * m.ny, dz);
    const V3<T> jg = v3<T>(j.x * gx.x + j.y * gy.x + j.z * gz.x, j.x * gx.y + j.y * gy.y + j.z * gz.y,
                           j.x * gx.z + j.y * gy.z + j.z * gz.z);
    const V3<T> Mv = v3<T>(M[li], M[n + li], M[2 * n + li]);
    const V3<T> B = cross3(Mv, jg);
    const V3<T> Av = cross3(Mv, B);
    return v3<T>(f1 * Av.x + f2 * B.x, f1 * Av.y + f2 * B.y, f1 * Av.z + f2 * B.z);
}

#endif
