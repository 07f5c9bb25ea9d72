// mm_tensor.hip -- demagnetisation tensor for open boundaries on the GPU.
//
// The same algorithm as the host generator (demag_tensor.cpp, restating
// src/magneto/mmm/demag/tensor.cpp:35-191 and tensor_integrals.h:313-585),
// evaluated in FP64 with one thread per entry (i,j,k) of the positive
// octant: axis sort, the large-x asymptotic sum inside the narrow cone,
// Taylor series of the second difference along axes whose edge is small
// against the distance (radius 30 edges), and the 27-point (3 s^2 - 2)
// second difference of Newell's f/g summands in the near field.  The
// generalized Newell functions come from newell_table_dev.inc, the FP64 form
// of the host's newell_table.inc.
//
// The series and asymptotic branches evaluate smooth functions and lose no
// digits; the finite differences cancel: an entry at distance r (in cells)
// is a difference of terms ~r^6 times larger, so FP64 keeps about
// 16 - 6 log10(r) digits (1e-10 relative at r = 10, 1e-8 at the expansion
// radius 30).  mm_generate_demag_tensor_gpu therefore lets the kernel flag
// every entry that takes a finite-difference branch -- and every entry whose
// branch choice is within 1e-12 of a threshold, where FP64 and long double
// could decide differently -- and evaluates those on the host in long double
// (unless near_on_gpu), exactly as the host generator does.  Each entry's
// mirror images are accumulated in the host's loop order, so entries
// computed on the host are bit-identical to the host generator.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "mm_internal.h"

namespace {

#include "newell_table_dev.inc"

#define MM_HIPCHK(x)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw mm_error(MM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// pi as the reference declares it (src/magneto/mmm/constants.h:26), a double
__device__ const double PI_D = 3.141592653589793238462643383279502884197169399;
__device__ const double EXPANSION_RADIUS = 30.0;
__device__ const double CONE_RATIO = 6.0;
// relative distance to a branch threshold below which the host decides
__device__ const double DECIDE_MARGIN = 1e-12;

template <typename T> __device__ inline void swp(T &a, T &b)
{
    T t = a;
    a = b;
    b = t;
}

// f(a,b,c) at any derivative order: the axes are swapped into a <= b <= c in
// the order demag_tensor.cpp's newell_f swaps them (first pair, then second)
__device__ __noinline__ double newell_f(int a, int b, int c, double lx, double ly, double lz, long i, long j, long k,
                                        int *bad)
{
    for (;;) {
        if (a > b) {
            swp(a, b);
            swp(lx, ly);
            swp(i, j);
            continue;
        }
        if (b > c) {
            swp(b, c);
            swp(ly, lz);
            swp(j, k);
            continue;
        }
        break;
    }
    double out = 0.0;
    if (!newell_f_sorted(a, b, c, (double)i * lx, (double)j * ly, (double)k * lz, &out)) *bad = 1;
    return out;
}

// near-field summands of Newell's f (0,0,2) and g (0,1,1); a summand whose
// polynomial prefactor vanishes is its limit 0 (demag_tensor.cpp f002_term,
// f011_term)
__device__ double f002_term(int n, double x, double y, double z)
{
    const double R = sqrt(x * x + y * y + z * z);
    switch (n) {
    case 0: return -R * x * x / 6.0;
    case 1: return -R * y * y / 6.0;
    case 2: return R * z * z / 3.0;
    case 3: return (x == 0 || y == 0 || z == 0) ? 0.0 : -x * y * z * atan((x * y) / (R * z));
    case 4: return (x == 0 || y == 0) ? 0.0 : x * y * y * atanh(x / R) / 2.0;
    case 5: return (x == 0 || z == 0) ? 0.0 : -x * z * z * atanh(x / R) / 2.0;
    case 6: return (x == 0 || y == 0) ? 0.0 : x * x * y * atanh(y / R) / 2.0;
    case 7: return (y == 0 || z == 0) ? 0.0 : -y * z * z * atanh(y / R) / 2.0;
    default: return 0.0;
    }
}

__device__ double f011_term(int n, double x, double y, double z)
{
    const double R = sqrt(x * x + y * y + z * z);
    switch (n) {
    case 0: return -R * y * z / 3.0;
    case 1: return (x == 0 || y == 0 || z == 0) ? 0.0 : -x * z * z * atan((x * y) / (R * z)) / 2.0;
    case 2: return (x == 0 || y == 0 || z == 0) ? 0.0 : -x * y * y * atan((x * z) / (R * y)) / 2.0;
    case 3: return (x == 0 || y == 0 || z == 0) ? 0.0 : -x * x * x * atan((y * z) / (R * x)) / 6.0;
    case 4: return (x == 0 || y == 0 || z == 0) ? 0.0 : x * y * z * atanh(x / R);
    case 5: return (x == 0 || z == 0) ? 0.0 : x * x * z * atanh(y / R) / 2.0;
    case 6: return (z == 0) ? 0.0 : -z * z * z * atanh(y / R) / 6.0;
    case 7: return (x == 0 || y == 0) ? 0.0 : x * x * y * atanh(z / R) / 2.0;
    case 8: return (y == 0) ? 0.0 : -y * y * y * atanh(z / R) / 6.0;
    default: return 0.0;
    }
}

__device__ __noinline__ double base_term(int a, int b, int c, double lx, double ly, double lz, long i, long j, long k,
                                         int n, int *bad)
{
    for (;;) {
        if (a > b) {
            swp(a, b);
            swp(lx, ly);
            swp(i, j);
            continue;
        }
        if (b > c) {
            swp(b, c);
            swp(ly, lz);
            swp(j, k);
            continue;
        }
        break;
    }
    const double x = (double)i * lx, y = (double)j * ly, z = (double)k * lz;
    if (a == 0 && b == 0 && c == 2) return f002_term(n, x, y, z);
    if (a == 0 && b == 1 && c == 1) return f011_term(n, x, y, z);
    *bad = 1;
    return 0.0;
}

__device__ inline double w3(int s) { return (double)(3 * s * s - 2); }

__device__ inline double factorial(int n)
{
    double r = 1.0;
    for (int q = 2; q <= n; ++q) r *= q;
    return r;
}

enum { AT_CENTER = 0, AT_BORDER = 1 };

// Taylor expansion of the second difference along x to order O (P, Q:
// expansion orders along y, z, or 0 for the exact second difference there)
__device__ __noinline__ double fseries(int o, int p, int q, int O, int P, int Q, double lx, double ly, double lz,
                                       long i, long j, long k, int where, int *bad)
{
    double xp = 1.0;
    if (where == AT_CENTER) {
        xp *= 2.0 * ipow(lx, O) / factorial(O);
    } else {
        xp *= (double)((1 << O) - 2) * ipow(-lx, O) / factorial(O);
        i += 1;
    }
    if (Q != 0) {
        return xp * 2.0 * ipow(ly, P) / factorial(P) * 2.0 * ipow(lz, Q) / factorial(Q) *
               newell_f(o + O, p + P, q + Q, lx, ly, lz, i, j, k, bad);
    }
    if (P != 0) {
        double s1 = 0.0;
        for (int s = -1; s <= 1; ++s) s1 += w3(s) * newell_f(o + O, p + P, q, lx, ly, lz, i, j, k + s, bad);
        return xp * 2.0 * ipow(ly, P) / factorial(P) * s1;
    }
    double s1 = 0.0;
    for (int r = -1; r <= 1; ++r) {
        double s2 = 0.0;
        for (int s = -1; s <= 1; ++s) s2 += w3(s) * newell_f(o + O, p, q, lx, ly, lz, i, j + r, k + s, bad);
        s1 += w3(r) * s2;
    }
    return xp * s1;
}

__device__ __noinline__ double large_x(int o, int p, int q, double lx, double ly, double lz, long i, long j, long k,
                                       int *bad)
{
    if (p > q) {
        swp(p, q);
        swp(ly, lz);
        swp(j, k);
    }
    const double x = (double)i * lx, y = (double)j * ly, z = (double)k * lz;
    if (o == 0 && p == 0 && q == 2) return nlarge_002(lx, ly, lz, x, y, z);
    if (o == 0 && p == 1 && q == 1) return nlarge_011(lx, ly, lz, x, y, z);
    *bad = 1;
    return 0.0;
}

__device__ inline long min_sq(long i)
{
    const long a = (i + 1) * (i + 1), b = (i - 1) * (i - 1), c = i * i;
    return a < b ? (a < c ? a : c) : (b < c ? b : c);
}

__device__ inline bool near_threshold(double a, double b) { return fabs(a - b) <= DECIDE_MARGIN * fabs(b); }

// Second difference of the base function for cell edges (1, ly, lz), the
// branch structure of demag_tensor.cpp second_difference.  *host: the
// branch cancels (finite differences) or its choice is marginal.
__device__ double second_difference(int o, int p, int q, double ly, double lz, long i, long j, long k, int *host,
                                     int *bad)
{
    const double lx = 1.0;
    const double er = EXPANSION_RADIUS;
    const double Rsq = (double)min_sq(i) * lx * lx + (double)min_sq(j) * ly * ly + (double)min_sq(k) * lz * lz;
    const double lim = Rsq / (er * er);
    int expand = 0;
    if (lx * lx < lim) ++expand;
    if (ly * ly < lim) ++expand;
    if (lz * lz < lim) ++expand;
    if (near_threshold(lx * lx, lim) || near_threshold(ly * ly, lim) || near_threshold(lz * lz, lim)) *host = 1;

    const double cone = Rsq / (CONE_RATIO * CONE_RATIO);
    const long ai = (i < 0 ? -i : i) + 1, aj = (j < 0 ? -j : j) + 1, ak = (k < 0 ? -k : k) + 1;
    const double cx = (double)(ai * ai) * lx * lx, cy = (double)(aj * aj) * ly * ly, cz = (double)(ak * ak) * lz * lz;
    const bool nx0 = cx < cone, ny0 = cy < cone, nz0 = cz < cone;
    if (near_threshold(cx, cone) || near_threshold(cy, cone) || near_threshold(cz, cone)) *host = 1;
    if (ny0 && nz0 && ((o == 0 && p == 2 && q == 0) || (o == 0 && p == 0 && q == 2) || (o == 0 && p == 1 && q == 1)))
        return large_x(o, p, q, lx, ly, lz, i, j, k, bad);
    if (nx0 && nz0 && ((o == 2 && p == 0 && q == 0) || (o == 0 && p == 0 && q == 2) || (o == 1 && p == 0 && q == 1)))
        return large_x(p, o, q, ly, lx, lz, j, i, k, bad);
    if (nx0 && ny0 && ((o == 2 && p == 0 && q == 0) || (o == 0 && p == 2 && q == 0) || (o == 1 && p == 1 && q == 0)))
        return large_x(q, o, p, lz, lx, ly, k, i, j, bad);

    if (expand != 3) *host = 1;
    double sum = 0.0;
    if (expand == 3) {
        sum += fseries(o, p, q, 2, 2, 2, lx, ly, lz, i, j, k, AT_CENTER, bad);
        sum += fseries(o, p, q, 4, 2, 2, lx, ly, lz, i, j, k, AT_CENTER, bad);
        sum += fseries(o, p, q, 2, 4, 2, lx, ly, lz, i, j, k, AT_CENTER, bad);
        sum += fseries(o, p, q, 2, 2, 4, lx, ly, lz, i, j, k, AT_CENTER, bad);
    } else if (expand == 2) {
        if (i == 0 && j == 0) {
            for (int O = 2; O <= 6; ++O) sum += fseries(o, p, q, O, 2, 0, lx, ly, lz, i, j, k, AT_BORDER, bad);
            sum += fseries(o, p, q, 2, 4, 0, lx, ly, lz, i, j, k, AT_BORDER, bad);
        } else {
            sum += fseries(o, p, q, 2, 2, 0, lx, ly, lz, i, j, k, AT_CENTER, bad);
            sum += fseries(o, p, q, 4, 2, 0, lx, ly, lz, i, j, k, AT_CENTER, bad);
            sum += fseries(o, p, q, 2, 4, 0, lx, ly, lz, i, j, k, AT_CENTER, bad);
        }
    } else if (expand == 1) {
        if (i == 0) {
            for (int O = 2; O <= 7; ++O) sum += fseries(o, p, q, O, 0, 0, lx, ly, lz, i, j, k, AT_BORDER, bad);
        } else {
            for (int O = 2; O <= 6; O += 2) sum += fseries(o, p, q, O, 0, 0, lx, ly, lz, i, j, k, AT_CENTER, bad);
        }
    } else {
        for (int n = 0; n < 13; ++n) {
            double s1 = 0.0;
            for (int c = -1; c <= 1; ++c) {
                double s2 = 0.0;
                for (int b = -1; b <= 1; ++b) {
                    double s3 = 0.0;
                    for (int a = -1; a <= 1; ++a) s3 += w3(a) * base_term(o, p, q, lx, ly, lz, i + a, j + b, k + c, n, bad);
                    s2 += w3(b) * s3;
                }
                s1 += w3(c) * s2;
            }
            sum += s1;
        }
    }
    return sum;
}

// averaged-cell interaction integral of component (o,p,q) at (i,j,k):
// demag_tensor.cpp cell_integral (axes sorted so that lx <= ly <= lz)
__device__ __noinline__ double cell_integral(int o, int p, int q, double lx, double ly, double lz, long i, long j,
                                             long k, int *host, int *bad)
{
    for (;;) {
        if (lx > ly) {
            swp(o, p);
            swp(lx, ly);
            swp(i, j);
            continue;
        }
        if (ly > lz) {
            swp(p, q);
            swp(ly, lz);
            swp(j, k);
            continue;
        }
        break;
    }
    if ((o & 1) && i == 0) return 0.0;
    if ((p & 1) && j == 0) return 0.0;
    if ((q & 1) && k == 0) return 0.0;
    const double pref = ipow(lx, 5 - o - p - q) / (4.0 * PI_D * lx * ly * lz);
    return pref * second_difference(o, p, q, ly / lx, lz / lx, i, j, k, host, bad);
}

// One thread per octant entry: the six components, mirrored into the padded
// grid in the host's loop order (mm_demag_fold_open) unless the entry goes
// to the host (flag[entry] = 1; near_on_gpu keeps finite differences here).
// flag[entry] = 2: an order outside the Newell table (cannot happen for the
// six components; reported as an error).
__global__ void k_demag_tensor(int nx, int ny, int nz, double dx, double dy, double dz, int ex, int ey, int ez,
                               int near_on_gpu, double *__restrict__ N, unsigned char *__restrict__ flag)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nx * ny * nz) return;
    const int i = (int)(idx % nx);
    const int j = (int)((idx / nx) % ny);
    const int k = (int)(idx / ((long)nx * ny));
    const int CA[6] = {0, 0, 0, 1, 1, 2}, CB[6] = {0, 1, 2, 1, 2, 2};
    int mcs = 1;
    if (i == 0) mcs *= 2;
    if (j == 0) mcs *= 2;
    if (k == 0) mcs *= 2;
    double v[6];
    int host = 0, bad = 0;
    for (int e = 0; e < 6; ++e) {
        int o[3] = {0, 0, 0};
        ++o[CA[e]];
        ++o[CB[e]];
        v[e] = 0.0 - cell_integral(o[0], o[1], o[2], dx, dy, dz, i, j, k, &host, &bad) / (double)mcs;
    }
    flag[idx] = (unsigned char)(bad ? 2 : (host && !near_on_gpu));
    if (bad || (host && !near_on_gpu)) return;
    for (int o = -1; o <= 1; o += 2)
        for (int p = -1; p <= 1; p += 2)
            for (int q = -1; q <= 1; q += 2) {
                const int I = (o * i + ex) % ex, J = (p * j + ey) % ey, K = (q * k + ez) % ez;
                double *d = N + 6 * ((size_t)I + (size_t)ex * ((size_t)J + (size_t)ey * K));
                d[0] += v[0];
                d[1] += o * p * v[1];
                d[2] += o * q * v[2];
                d[3] += v[3];
                d[4] += p * q * v[4];
                d[5] += v[5];
            }
}

} // namespace

extern "C" int mm_generate_demag_tensor_gpu(int nx, int ny, int nz, double dx, double dy, double dz, int padding,
                                            int near_on_gpu, int nthreads, double *N_host, int *exp_out,
                                            double *info4)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad mesh size");
    MM_REQUIRE(dx > 0 && dy > 0 && dz > 0, "cell sizes must be positive");
    const int ex = mm_round_tensor_dimension_impl(nx, false, padding);
    const int ey = mm_round_tensor_dimension_impl(ny, false, padding);
    const int ez = mm_round_tensor_dimension_impl(nz, false, padding);
    if (exp_out) {
        exp_out[0] = ex;
        exp_out[1] = ey;
        exp_out[2] = ez;
    }
    if (!N_host) return MM_OK;
    const size_t n = (size_t)nx * ny * nz;
    const size_t count = (size_t)6 * ex * ey * ez;
    const auto t0 = std::chrono::steady_clock::now();
    double *dN = nullptr;
    unsigned char *dflag = nullptr;
    std::vector<unsigned char> flag(n);
    try {
        MM_HIPCHK(hipMalloc(&dN, sizeof(double) * count));
        MM_HIPCHK(hipMalloc(&dflag, n));
        MM_HIPCHK(hipMemset(dN, 0, sizeof(double) * count));
        const int block = 64;
        const size_t grid = (n + block - 1) / block;
        MM_REQUIRE(grid <= 0x7fffffffu, "mesh too large");
        hipLaunchKernelGGL(k_demag_tensor, dim3((unsigned)grid), dim3(block), 0, 0, nx, ny, nz, dx, dy, dz, ex, ey, ez,
                           near_on_gpu, dN, dflag);
        MM_HIPCHK(hipGetLastError());
        MM_HIPCHK(hipMemcpy(N_host, dN, sizeof(double) * count, hipMemcpyDeviceToHost));
        MM_HIPCHK(hipMemcpy(flag.data(), dflag, n, hipMemcpyDeviceToHost));
    } catch (...) {
        if (dN) (void)hipFree(dN);
        if (dflag) (void)hipFree(dflag);
        throw;
    }
    (void)hipFree(dN);
    (void)hipFree(dflag);
    for (size_t q = 0; q < n; ++q)
        if (flag[q] == 2) throw mm_error(MM_ERR_RUNTIME, "demag tensor: generalized Newell function not tabulated");
    const auto t1 = std::chrono::steady_clock::now();

    // entries flagged for the host: long double, as the host generator
    std::vector<size_t> todo;
    for (size_t q = 0; q < n; ++q)
        if (flag[q]) todo.push_back(q);
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    nthreads = std::max(1, std::min<int>(nthreads, (int)std::max<size_t>(1, todo.size() / 16)));
    std::vector<std::vector<double>> vals(nthreads);
    auto work = [&](int t) {
        vals[t].resize(6 * ((todo.size() + nthreads - 1) / nthreads));
        size_t m = 0;
        for (size_t r = t; r < todo.size(); r += nthreads, ++m) {
            const size_t q = todo[r];
            mm_demag_entry_open_ld((int)(q % nx), (int)((q / nx) % ny), (int)(q / ((size_t)nx * ny)), dx, dy, dz,
                                   &vals[t][6 * m]);
        }
    };
    if (nthreads == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; ++t) pool.emplace_back(work, t);
        for (auto &th : pool) th.join();
    }
    // each entry owns its mirror locations (ex >= 2 nx - 1, ...): fold them in
    for (int t = 0; t < nthreads; ++t) {
        size_t m = 0;
        for (size_t r = t; r < todo.size(); r += nthreads, ++m) {
            const size_t q = todo[r];
            mm_demag_fold_open((int)(q % nx), (int)((q / nx) % ny), (int)(q / ((size_t)nx * ny)), ex, ey, ez,
                               &vals[t][6 * m], N_host);
        }
    }
    for (size_t q = 0; q < count; ++q) N_host[q] = -N_host[q];
    const auto t2 = std::chrono::steady_clock::now();
    if (info4) {
        info4[0] = (double)(n - todo.size());
        info4[1] = (double)todo.size();
        info4[2] = std::chrono::duration<double>(t1 - t0).count();
        info4[3] = std::chrono::duration<double>(t2 - t1).count();
    }
    MM_CAPI_END
}
