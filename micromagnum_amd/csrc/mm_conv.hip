// mm_conv.hip -- FFT demag convolution on gfx950.
//
// Replaces SymmetricMatrixVectorConvolution_FFT / MatrixVectorConvolution_FFT
// (src/magneto/math/conv/MatrixVectorConvolution_FFT.cpp:110-375) with five
// LDS-staged passes per evaluation instead of thirteen:
//
//   x fwd  : real rows of M (x < nx) zero-extended to Px, complex FFT,
//            keep kx < Hx = Px/2+1                    M  -> S1[c][z][y][kx]
//   y fwd  : columns of S1 (y < ny) zero-extended to Py  S1 -> S2[c][z][ky][kx]
//   z      : lines of S2 (z < nz) zero-extended to Pz, FFT, 3x3 multiply with
//            the real octant-stored Fourier kernel, inverse FFT, keep z < nz
//   y inv  : inverse FFT of S2 columns, keep y < ny     S2 -> S1
//   x inv  : Hermitian extension, inverse FFT, keep x < nx  S1 -> H
//
// 2-D meshes (nz == 1) fuse y fwd + multiply + y inv into one pass on S1.
// No transpose passes: strided passes load tiles of adjacent kx columns so
// every global access is a contiguous row segment, and the zero padding is
// never stored (only nonzero input rows are read, only kept outputs written).
//
// Fourier kernel: the padded tensor is even/odd in every axis, so its
// transform is real; only kx < Hx, ky <= Py/2, kz <= Pz/2 are stored (6 real
// planes, sign flips for the odd components on the other octants).  It is
// computed once on the device in FP64 with the same passes and cast to the
// working precision.  The reference keeps the full complex kernel; the
// dropped imaginary parts are rounding noise (~1e-17 relative).
#include "mm_conv.h"
#include "mm_cell.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

// ---------------------------------------------------------------------------
// FFT plan
// ---------------------------------------------------------------------------
void FftPlanHost::build(int len)
{
    MM_REQUIRE(len >= 1, "FFT length must be positive");
    L = len;
    rad.clear();
    int n = len;
    int twos = 0;
    while (n % 2 == 0) { n /= 2; ++twos; }
    while (twos >= 3) { rad.push_back(8); twos -= 3; }
    if (twos == 2) rad.push_back(4);
    if (twos == 1) rad.push_back(2);
    while (n % 5 == 0) { rad.push_back(5); n /= 5; }
    while (n % 3 == 0) { rad.push_back(3); n /= 3; }
    if (n != 1) throw mm_error(MM_ERR_UNSUPPORTED, "FFT length " + std::to_string(len) + " has a prime factor > 5");
    MM_REQUIRE(rad.size() <= MM_FFT_MAX_PASSES, "FFT length needs too many passes");
    tline = std::max(1, (L + MM_FFT_EPT - 1) / MM_FFT_EPT);
    if (tline > 1024) throw mm_error(MM_ERR_UNSUPPORTED, "FFT length " + std::to_string(len) + " above 8192");
    std::vector<float2> t32(L);
    std::vector<double2> t64(L);
    for (int m = 0; m < L; ++m) {
        const long double a = -2.0L * MM_PI_L * (long double)m / (long double)L;
        const long double c = cosl(a), s = sinl(a);
        t64[m].x = (double)c;
        t64[m].y = (double)s;
        t32[m].x = (float)c;
        t32[m].y = (float)s;
    }
    MM_HIP(hipMalloc(&tw32, sizeof(float2) * L));
    MM_HIP(hipMalloc(&tw64, sizeof(double2) * L));
    MM_HIP(hipMemcpy(tw32, t32.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
    MM_HIP(hipMemcpy(tw64, t64.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
    // per-pass layout: pass p (radix R, Ns > 1) holds W_{Ns R}^{r k} at
    // (r-1) Ns + k, copied from the plain table at index r k L / (Ns R)
    std::vector<float2> p32;
    std::vector<double2> p64;
    int Ns = 1;
    for (int R : rad) {
        if (Ns > 1)
            for (int r = 1; r < R; ++r)
                for (int k = 0; k < Ns; ++k) {
                    const int m = r * k * (L / (Ns * R));
                    p32.push_back(t32[m]);
                    p64.push_back(t64[m]);
                }
        Ns *= R;
    }
    ntw = (int)p64.size();
    MM_REQUIRE(ntw < L || L == 1, "per-pass twiddle table larger than the line");
    const size_t nalloc = std::max<size_t>(1, p64.size());
    MM_HIP(hipMalloc(&twp32, sizeof(float2) * nalloc));
    MM_HIP(hipMalloc(&twp64, sizeof(double2) * nalloc));
    if (ntw > 0) {
        MM_HIP(hipMemcpy(twp32, p32.data(), sizeof(float2) * ntw, hipMemcpyHostToDevice));
        MM_HIP(hipMemcpy(twp64, p64.data(), sizeof(double2) * ntw, hipMemcpyHostToDevice));
    }
}

void FftPlanHost::release()
{
    if (tw32) (void)hipFree(tw32);
    if (tw64) (void)hipFree(tw64);
    if (twp32) (void)hipFree(twp32);
    if (twp64) (void)hipFree(twp64);
    tw32 = tw64 = twp32 = twp64 = nullptr;
}

bool mm_is_smooth(int n)
{
    if (n < 1) return false;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

// Any padded size P >= 2n-1 gives the same linear convolution for open
// boundaries (the tensor is zero between the +/- offsets), so the given size
// is kept if it is radix-friendly and even, else the next such size is used.
// A periodic axis (given == n) is a circular convolution of period n and
// keeps its size, which must then be radix-friendly itself.
int mm_fft_size_for(int n, int given)
{
    if (n == 1) return 1;
    if (given < 2 * n - 1) {
        MM_REQUIRE(given == n, "tensor size along an axis must be the mesh size (periodic) or >= 2n-1 (open)");
        MM_REQUIRE(mm_is_smooth(n), "a periodic axis needs a mesh size of the form 2^a 3^b 5^c");
        return n;
    }
    if (given % 2 == 0 && mm_is_smooth(given)) return given;
    int p = 2 * n - 1;
    while (!(p % 2 == 0 && mm_is_smooth(p))) ++p;
    return p;
}

static int g_max_lds = 0;
static std::once_flag g_lds_once;

int mm_max_lds_bytes()
{
    std::call_once(g_lds_once, []() {
        int dev = 0;
        (void)hipGetDevice(&dev);
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        g_max_lds = v;
    });
    return g_max_lds;
}

// Dynamic LDS above 64 KB needs no opt-in on gfx950 (the device reports its
// per-workgroup maximum); launches are error-checked.  Kept as the single
// place that validates a request against that maximum.
template <typename K> static void set_lds(K, size_t bytes)
{
    MM_REQUIRE(bytes <= (size_t)mm_max_lds_bytes(), "LDS request exceeds the device maximum");
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// x forward: rows (c, z, y); input rows of nx_in reals with row stride nx_in;
// workgroup blk takes lines blk lpw .. (blk + 1) lpw - 1 (threads beyond the
// lpw lines stay idle)
template <typename T, typename TIN, int LC = 0>
__device__ __forceinline__ void xfwd_body(const TIN *__restrict__ in, size_t in_cstride, C2<T> *__restrict__ out,
                                          size_t out_cstride, int Hx, int Hp, int ny, int nz, int nrows, int nx_in,
                                          const FftPlanDev &p, int lpw, int blk)
{
    constexpr int SK = MM_SKEW_X;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int L = PlanC<LC>::L(p), LP = lds_len<SK>(L), TL = PlanC<LC>::tline(p);
    // two real rows per complex line: line l of the block holds rows
    // 2 g and 2 g + 1 (g = blockIdx.x lpw + l) as real and imaginary part
    const int l = threadIdx.x / TL, t = threadIdx.x % TL;
    const int g0 = 2 * (blk * lpw + l);
    const bool active = l < lpw && g0 < nrows;
    const bool has1 = g0 + 1 < nrows;
    C2<T> *line = sm + (size_t)l * LP;
    C2<T> *twl = sm + (size_t)lpw * LP;
    const TwFetch<T> twf = TwFetch<T>::fetch(p);
    const int rows_c = ny * nz;
    auto row_src = [&](int g) { return in + (g / rows_c) * in_cstride + (size_t)(g % rows_c) * nx_in; };
    const TIN *src0 = row_src(active ? g0 : 0);
    const TIN *src1 = row_src(active && has1 ? g0 + 1 : 0);
    const LineMap m = {t, active};
    fft_line<T, false, true, false, LC, SK>(
        p, twl, line, m, line, m,
        [&](int x) {
            const bool in_row = x < nx_in;
            return cmk<T>(in_row ? (T)src0[x] : (T)0, in_row && has1 ? (T)src1[x] : (T)0);
        },
        [](int, C2<T>) {}, [&] { twf.store(twl); });
    // X0[k] = (Z[k] + conj Z[L-k]) / 2, X1[k] = (Z[k] - conj Z[L-k]) / 2i
    for (int l2 = 0; l2 < lpw; ++l2) {
        const int h0 = 2 * (blk * lpw + l2);
        if (h0 >= nrows) break;
        const C2<T> *ln = sm + (size_t)l2 * LP;
        C2<T> *d0 = out + (h0 / rows_c) * out_cstride + (size_t)(h0 % rows_c) * Hp;
        C2<T> *d1 = h0 + 1 < nrows ? out + ((h0 + 1) / rows_c) * out_cstride + (size_t)((h0 + 1) % rows_c) * Hp : nullptr;
        for (int k = threadIdx.x; k < Hx; k += blockDim.x) {
            const C2<T> zk = ln[sk<SK>(k)], zm = ln[sk<SK>(k == 0 ? 0 : L - k)];
            d0[k] = cmk<T>((zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5);
            if (d1) d1[k] = cmk<T>((zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5);
        }
    }
}

// sel (may be null): device index choosing in (0) or in_alt (1)
template <typename T, typename TIN, int LC = 0>
__global__ void k_xfwd(const TIN *__restrict__ in, size_t in_cstride, C2<T> *__restrict__ out,
                       size_t out_cstride, int Hx, int Hp, int ny, int nz, int nrows, int nx_in,
                       FftPlanDev p, int lpw, const TIN *__restrict__ in_alt = nullptr,
                       const int *__restrict__ sel = nullptr)
{
    if (sel && *sel) in = in_alt;
    xfwd_body<T, TIN, LC>(in, in_cstride, out, out_cstride, Hx, Hp, ny, nz, nrows, nx_in, p, lpw, (int)blockIdx.x);
}

// n / d for the tile indices of a launch without an integer division (the
// GPU has none; the compiler emits some twenty instructions for one):
// m = floor(2^32 / d) + 1 gives floor(n / d) = umulhi(n, m) exactly when
// n * d < 2^32.  mdiv returns 0 where that does not hold, and the division
// is then done the long way.
__device__ __forceinline__ unsigned udiv(unsigned n, unsigned d, unsigned m)
{
    return d == 1 ? n : (m ? __umulhi(n, m) : n / d);
}

// y forward: columns kx of planes (c, z); rows y < ny_in read, Py written.
// O32: both arrays under 4 GiB and a row under 16 MiB: 32-bit byte offsets
// from the kernel-argument bases, rows stepped by 24-bit multiplies
// tile: (kx tile, z, c) with the kx tile fastest, ntx kx tiles per plane
template <typename T, int LC = 0, bool O32 = false>
__device__ __forceinline__ void yfwd_body(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int Hx, int Hp,
                                          int ny_in, int nzl, const FftPlanDev &p, int tk, int tile, int ntx,
                                          unsigned mtx, unsigned mzl)
{
    constexpr int SK = MM_SKEW_Y;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int L = PlanC<LC>::L(p), LP = lds_len<SK>(L);
    // tk is a power of two
    const int ltk = __ffs(tk) - 1;
    const int q1 = (int)udiv((unsigned)tile, (unsigned)ntx, mtx), c = (int)udiv((unsigned)q1, (unsigned)nzl, mzl);
    const int kx0 = (tile - q1 * ntx) * tk;
    const int z = q1 - c * nzl;
    const C2<T> *src = in + ((size_t)c * nzl + z) * (size_t)ny_in * Hp;
    C2<T> *dst = out + ((size_t)c * nzl + z) * (size_t)L * Hp;
    C2<T> *twl = sm + (size_t)tk * LP;
    const TwFetch<T> twf = TwFetch<T>::fetch(p);
    // global passes: columns fastest across the threads (row segments of tk
    // values per load); LDS passes: one line per group of tline threads
    const int lg = threadIdx.x & (tk - 1), ls = threadIdx.x / PlanC<LC>::tline(p);
    const int tg = (int)(threadIdx.x >> ltk);
    const LineMap mg = {tg, tg < PlanC<LC>::tline(p)}, ms = {(int)(threadIdx.x % PlanC<LC>::tline(p)), ls < tk};
    const int kx = kx0 + lg;
    constexpr unsigned CS = sizeof(C2<T>);
    const unsigned rs = (unsigned)Hp * CS, pl = ((unsigned)c * nzl + z);
    const unsigned sb = O32 ? (pl * (unsigned)ny_in * Hp + kx) * CS : 0u, db = O32 ? (pl * (unsigned)L * Hp + kx) * CS : 0u;
    fft_line<T, false, true, true, LC, SK>(
        p, twl, sm + (size_t)lg * LP, mg, sm + (size_t)ls * LP, ms,
        [&](int y) {
            if (!(y < ny_in && kx < Hx)) return cmk<T>((T)0, (T)0);
            if constexpr (O32) return *(const C2<T> *)((const char *)in + (sb + __umul24((unsigned)y, rs)));
            else return src[(size_t)y * Hp + kx];
        },
        [&](int y, C2<T> v) {
            if (kx < Hx) {
                if constexpr (O32) *(C2<T> *)((char *)out + (db + __umul24((unsigned)y, rs))) = v;
                else dst[(size_t)y * Hp + kx] = v;
            }
        }, [&] { twf.store(twl); });
}

template <typename T, int LC = 0, bool O32 = false>
__global__ void k_yfwd(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int Hx, int Hp, int ny_in,
                       int nzl, FftPlanDev p, int tk, int remap, unsigned mtx, unsigned mzl)
{
    yfwd_body<T, LC, O32>(in, out, Hx, Hp, ny_in, nzl, p, tk, block_tile(remap), (int)gridDim.x, mtx, mzl);
}

// y inverse: columns of S2 planes (c, z), all Py rows read, y < ny_out
// written; O32 as k_yfwd
template <typename T, int LC = 0, bool O32 = false>
__device__ __forceinline__ void yinv_body(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int Hx, int Hp,
                                          int ny_out, int nzl, const FftPlanDev &p, int tk, int tile, int ntx,
                                          unsigned mtx, unsigned mzl)
{
    constexpr int SK = MM_SKEW_Y;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int L = PlanC<LC>::L(p), LP = lds_len<SK>(L);
    const int ltk = __ffs(tk) - 1;
    const int q1 = (int)udiv((unsigned)tile, (unsigned)ntx, mtx), c = (int)udiv((unsigned)q1, (unsigned)nzl, mzl);
    const int kx0 = (tile - q1 * ntx) * tk;
    const int z = q1 - c * nzl;
    const C2<T> *src = in + ((size_t)c * nzl + z) * (size_t)L * Hp;
    C2<T> *dst = out + ((size_t)c * nzl + z) * (size_t)ny_out * Hp;
    C2<T> *twl = sm + (size_t)tk * LP;
    const TwFetch<T> twf = TwFetch<T>::fetch(p);
    const int lg = threadIdx.x & (tk - 1), ls = threadIdx.x / PlanC<LC>::tline(p);
    const int tg = (int)(threadIdx.x >> ltk);
    const LineMap mg = {tg, tg < PlanC<LC>::tline(p)}, ms = {(int)(threadIdx.x % PlanC<LC>::tline(p)), ls < tk};
    const int kx = kx0 + lg;
    constexpr unsigned CS = sizeof(C2<T>);
    const unsigned rs = (unsigned)Hp * CS, pl = ((unsigned)c * nzl + z);
    const unsigned sb = O32 ? (pl * (unsigned)L * Hp + kx) * CS : 0u, db = O32 ? (pl * (unsigned)ny_out * Hp + kx) * CS : 0u;
    fft_line<T, true, true, true, LC, SK>(
        p, twl, sm + (size_t)lg * LP, mg, sm + (size_t)ls * LP, ms,
        [&](int y) {
            if (!(kx < Hx)) return cmk<T>((T)0, (T)0);
            if constexpr (O32) return *(const C2<T> *)((const char *)in + (sb + __umul24((unsigned)y, rs)));
            else return src[(size_t)y * Hp + kx];
        },
        [&](int y, C2<T> v) {
            if (y < ny_out && kx < Hx) {
                if constexpr (O32) *(C2<T> *)((char *)out + (db + __umul24((unsigned)y, rs))) = v;
                else dst[(size_t)y * Hp + kx] = v;
            }
        }, [&] { twf.store(twl); });
}

template <typename T, int LC = 0, bool O32 = false>
__global__ void k_yinv(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int Hx, int Hp, int ny_out,
                       int nzl, FftPlanDev p, int tk, int remap, unsigned mtx, unsigned mzl)
{
    yinv_body<T, LC, O32>(in, out, Hx, Hp, ny_out, nzl, p, tk, block_tile(remap), (int)gridDim.x, mtx, mzl);
}

// Pruned y transforms for zero-padded lines (P = 2h even, data only in
// y < ny <= h).  Forward: X[2m] = DFT_h(x)[m], X[2m+1] = DFT_h(x W_P^n)[m];
// inverse: x[n] = IDFT_h(X_even)[n] + W_P^-n IDFT_h(X_odd)[n] for n < h.
// Two half-length transforms per line halve the LDS per line, which lets a
// workgroup take 128-byte column tiles.  twP: the length-P twiddle table.
// Each thread keeps the elements e = tid + q*blockDim (col = e % tk,
// n = e / tk) of its tile in registers between the two halves; blockDim =
// tk * tline with tline >= h/8 gives at most 8 per thread.
template <typename T>
__global__ void k_yfwd_pr(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int Hx, int Hp, int ny_in,
                          int nzl, FftPlanDev ph, const C2<T> *__restrict__ twP, int tk)
{
    constexpr int SK = MM_SKEW_Y;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int h = ph.L, LP = lds_len<SK>(h);
    const int kx0 = blockIdx.x * tk;
    const int z = blockIdx.y, c = blockIdx.z;
    const C2<T> *src = in + ((size_t)c * nzl + z) * (size_t)ny_in * Hp;
    C2<T> *dst = out + ((size_t)c * nzl + z) * (size_t)(2 * h) * Hp;
    C2<T> *twl = sm + (size_t)tk * LP;
    load_twiddles<T>(twl, ph);
    constexpr int MAXE = 8;
    C2<T> xr[MAXE];
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
        const int e = threadIdx.x + q * blockDim.x;
        const int col = e % tk, n = e / tk, kx = kx0 + col;
        xr[q] = cmk<T>((T)0, (T)0);
        if (e < tk * h) {
            if (n < ny_in && kx < Hx) xr[q] = src[(size_t)n * Hp + kx];
            sm[col * LP + sk<SK>(n)] = xr[q];
        }
    }
    __syncthreads();
    const int l = threadIdx.x / ph.tline, t = threadIdx.x % ph.tline;
    lds_fft<T, false, SK>(sm + (size_t)l * LP, ph, t, l < tk, twl);
    for (int e = threadIdx.x; e < tk * h; e += blockDim.x) {
        const int col = e % tk, m = e / tk, kx = kx0 + col;
        if (kx < Hx) dst[(size_t)(2 * m) * Hp + kx] = sm[col * LP + sk<SK>(m)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
        const int e = threadIdx.x + q * blockDim.x;
        const int col = e % tk, n = e / tk;
        if (e < tk * h) sm[col * LP + sk<SK>(n)] = cmul<T>(xr[q], twP[n]);
    }
    __syncthreads();
    lds_fft<T, false, SK>(sm + (size_t)l * LP, ph, t, l < tk, twl);
    for (int e = threadIdx.x; e < tk * h; e += blockDim.x) {
        const int col = e % tk, m = e / tk, kx = kx0 + col;
        if (kx < Hx) dst[(size_t)(2 * m + 1) * Hp + kx] = sm[col * LP + sk<SK>(m)];
    }
}

template <typename T>
__global__ void k_yinv_pr(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int Hx, int Hp, int ny_out,
                          int nzl, FftPlanDev ph, const C2<T> *__restrict__ twP, int tk)
{
    constexpr int SK = MM_SKEW_Y;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int h = ph.L, LP = lds_len<SK>(h);
    const int kx0 = blockIdx.x * tk;
    const int z = blockIdx.y, c = blockIdx.z;
    const C2<T> *src = in + ((size_t)c * nzl + z) * (size_t)(2 * h) * Hp;
    C2<T> *dst = out + ((size_t)c * nzl + z) * (size_t)ny_out * Hp;
    const int l = threadIdx.x / ph.tline, t = threadIdx.x % ph.tline;
    C2<T> *twl = sm + (size_t)tk * LP;
    load_twiddles<T>(twl, ph);
    constexpr int MAXE = 8;
    C2<T> ar[MAXE];
    for (int e = threadIdx.x; e < tk * h; e += blockDim.x) {
        const int col = e % tk, m = e / tk, kx = kx0 + col;
        sm[col * LP + sk<SK>(m)] = (kx < Hx) ? src[(size_t)(2 * m) * Hp + kx] : cmk<T>((T)0, (T)0);
    }
    __syncthreads();
    lds_fft<T, true, SK>(sm + (size_t)l * LP, ph, t, l < tk, twl);
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
        const int e = threadIdx.x + q * blockDim.x;
        const int col = e % tk, n = e / tk;
        ar[q] = (e < tk * h) ? sm[col * LP + sk<SK>(n)] : cmk<T>((T)0, (T)0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < tk * h; e += blockDim.x) {
        const int col = e % tk, m = e / tk, kx = kx0 + col;
        sm[col * LP + sk<SK>(m)] = (kx < Hx) ? src[(size_t)(2 * m + 1) * Hp + kx] : cmk<T>((T)0, (T)0);
    }
    __syncthreads();
    lds_fft<T, true, SK>(sm + (size_t)l * LP, ph, t, l < tk, twl);
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
        const int e = threadIdx.x + q * blockDim.x;
        const int col = e % tk, n = e / tk, kx = kx0 + col;
        if (e < tk * h && n < ny_out && kx < Hx)
            dst[(size_t)n * Hp + kx] = cadd<T>(ar[q], cmul<T>(sm[col * LP + sk<SK>(n)], cconj<T>(twP[n])));
    }
}

// real octant-stored kernel entry e at (kx, ky, kz) of the full spectrum
template <typename T>
__device__ __forceinline__ void kernel_at(const T *__restrict__ K, size_t plane, int Hp, int Hy2,
                                          int Py, int Pz, int kx, int ky, int kz, T k[6])
{
    const bool fy = ky > Py / 2, fz = kz > Pz / 2;
    const int kyf = fy ? Py - ky : ky;
    const int kzf = fz ? Pz - kz : kz;
    const size_t o = ((size_t)kzf * Hy2 + kyf) * Hp + kx;
    const T sy = fy ? (T)-1 : (T)1, sz = fz ? (T)-1 : (T)1;
    k[0] = K[o];
    k[1] = sy * K[plane + o];
    k[2] = sz * K[2 * plane + o];
    k[3] = K[3 * plane + o];
    k[4] = sy * sz * K[4 * plane + o];
    k[5] = K[5 * plane + o];
}

template <typename T>
__device__ __forceinline__ void mult3(C2<T> &mx, C2<T> &my, C2<T> &mz, const T k[6])
{
    const C2<T> x = mx, y = my, z = mz;
    mx = cmk<T>(k[0] * x.x + k[1] * y.x + k[2] * z.x, k[0] * x.y + k[1] * y.y + k[2] * z.y);
    my = cmk<T>(k[1] * x.x + k[3] * y.x + k[4] * z.x, k[1] * x.y + k[3] * y.y + k[4] * z.y);
    mz = cmk<T>(k[2] * x.x + k[4] * y.x + k[5] * z.x, k[2] * x.y + k[4] * y.y + k[5] * z.y);
}

// kernel items (column, kz pair) per thread of the z pass
#define ZM_ITEMS 2

// z pass (3-D): forward over z (nz rows, zero-extended), multiply, inverse,
// keep z < nz.  Lines: 3 components x tk columns for one ky.
// O32: S2 and K are under 4 GiB and a z step of S2 under 16 MiB (zsb bytes),
// so every access is a 32-bit byte offset from the kernel-argument base
// (global_load with an SGPR base: no 64-bit address arithmetic per element;
// the offsets' z steps are 24-bit multiplies); pb: bytes of one K plane.
template <typename T, int LC = 0, bool O32 = false>
__device__ __forceinline__ void zmi_body(C2<T> *__restrict__ S2, const T *__restrict__ K, size_t plane, int Hx,
                                         int Hp, int Hy2, int Py, int nz, const FftPlanDev &p, int tk, int remap,
                                         int tile, int ntx, unsigned mtx, unsigned zsb, unsigned pb)
{
    constexpr int SK = MM_SKEW_Z;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int L = PlanC<LC>::L(p), LP = lds_len<SK>(L);
    // tiles (kx tile, m), kx tile fastest.  With remap, rows come in the order
    // 0, 1, Py-1, 2, Py-2, ...: ky and Py-ky read the same kernel row
    // tk is a power of two
    const int ltk = __ffs(tk) - 1;
    const int m = (int)udiv((unsigned)tile, (unsigned)ntx, mtx), hm = (m + 1) >> 1;
    const int kx0 = (tile - m * ntx) * tk;
    const int ky = !remap || m == 0 ? m : ((m & 1) ? hm : Py - hm);
    const size_t cstride = (size_t)nz * Py * Hp;
    const size_t zstride = (size_t)Py * Hp;
    C2<T> *base = S2 + (size_t)ky * Hp;
    C2<T> *twl = sm + (size_t)3 * tk * LP;
    const TwFetch<T> twf = TwFetch<T>::fetch(p);
    // global passes: columns fastest, then components; LDS passes: one line
    // (component c, column) per group of tline threads
    const int colg = threadIdx.x & (tk - 1), cg = (int)(threadIdx.x >> ltk) % 3;
    const int tg = (int)(threadIdx.x >> ltk) / 3;
    const LineMap mg = {tg, tg < PlanC<LC>::tline(p)};
    C2<T> *lineG = sm + (size_t)(cg * tk + colg) * LP;
    const int l = threadIdx.x / PlanC<LC>::tline(p), t = threadIdx.x % PlanC<LC>::tline(p);
    const LineMap ms = {t, l < 3 * tk};
    C2<T> *line = sm + (size_t)l * LP;
    const int kxg = kx0 + colg;
    C2<T> *colp = base + cg * cstride + kxg;
    const unsigned cb = O32 ? (unsigned)(((size_t)cg * cstride + (size_t)ky * Hp + kxg) * sizeof(C2<T>)) : 0u;
    auto col_at = [&](int z) -> C2<T> & {
        if constexpr (O32) return *(C2<T> *)((char *)S2 + (cb + __umul24((unsigned)z, zsb)));
        else return colp[z * zstride];
    };
    // The kernel entries this thread multiplies load before the forward
    // transform, so their latency hides behind it.  Items (col, kzf) with
    // kzf <= L/2 each serve kz = kzf and the mirror kz = L - kzf, whose xz
    // and yz entries change sign (real octant storage, as kernel_at); at most
    // ZM_ITEMS per thread (the host checks).
    const int nitems = tk * (L / 2 + 1);
    const bool fy = ky > Py / 2;
    const int kyf = fy ? Py - ky : ky;
    const T sy = fy ? (T)-1 : (T)1;
    T kk[ZM_ITEMS][6];
#pragma unroll
    for (int j = 0; j < ZM_ITEMS; ++j) {
        const int e = threadIdx.x + j * blockDim.x;
        const int col = e & (tk - 1), kzf = e >> ltk, kx = kx0 + col;
        if (e < nitems && kx < Hx) {
            const size_t o = ((size_t)kzf * Hy2 + kyf) * Hp + kx;
            if constexpr (O32) {
                const unsigned ob = (unsigned)o * (unsigned)sizeof(T);
#pragma unroll
                for (int q = 0; q < 6; ++q) kk[j][q] = *(const T *)((const char *)K + (ob + (unsigned)q * pb));
            } else {
#pragma unroll
                for (int q = 0; q < 6; ++q) kk[j][q] = K[q * plane + o];
            }
        }
    }
    fft_line<T, false, true, false, LC, SK>(
        p, twl, lineG, mg, line, ms,
        [&](int z) { return (z < nz && kxg < Hx) ? col_at(z) : cmk<T>((T)0, (T)0); },
        [](int, C2<T>) {}, [&] { twf.store(twl); });
#pragma unroll
    for (int j = 0; j < ZM_ITEMS; ++j) {
        const int e = threadIdx.x + j * blockDim.x;
        const int col = e & (tk - 1), kzf = e >> ltk, kx = kx0 + col;
        if (e < nitems && kx < Hx) {
            T k[6] = {kk[j][0], sy * kk[j][1], kk[j][2], kk[j][3], sy * kk[j][4], kk[j][5]};
            C2<T> *l0 = sm + (size_t)(0 * tk + col) * LP, *l1 = sm + (size_t)(1 * tk + col) * LP,
                  *l2 = sm + (size_t)(2 * tk + col) * LP;
            mult3<T>(l0[sk<SK>(kzf)], l1[sk<SK>(kzf)], l2[sk<SK>(kzf)], k);
            const int kz2 = L - kzf;
            if (kzf > 0 && kz2 != kzf) {
                k[2] = -k[2];
                k[4] = -k[4];
                mult3<T>(l0[sk<SK>(kz2)], l1[sk<SK>(kz2)], l2[sk<SK>(kz2)], k);
            }
        }
    }
    __syncthreads();
    fft_line<T, true, false, true, LC, SK>(
        p, twl, lineG, mg, line, ms, [](int) { return cmk<T>((T)0, (T)0); },
        [&](int z, C2<T> v) {
            if (z < nz && kxg < Hx) col_at(z) = v;
        });
}

template <typename T, int LC = 0, bool O32 = false>
__global__ void k_zmi(C2<T> *__restrict__ S2, const T *__restrict__ K, size_t plane, int Hx, int Hp, int Hy2,
                      int Py, int nz, FftPlanDev p, int tk, int remap, unsigned mtx, unsigned zsb, unsigned pb)
{
    zmi_body<T, LC, O32>(S2, K, plane, Hx, Hp, Hy2, Py, nz, p, tk, remap, block_tile(remap), (int)gridDim.x, mtx,
                         zsb, pb);
}

// 2-D: y forward (ny rows, zero-extended) + multiply (kz = 0) + y inverse,
// keep y < ny, in place on S1 planes [c][y][kx].
template <typename T, int LC = 0>
__device__ __forceinline__ void ymi2d_body(C2<T> *__restrict__ S1, const T *__restrict__ K, size_t plane, int Hx,
                                           int Hp, int Hy2, int ny, const FftPlanDev &p, int tk, int blk)
{
    constexpr int SK = MM_SKEW_Y;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int L = PlanC<LC>::L(p), LP = lds_len<SK>(L);
    const int kx0 = blk * tk;
    const size_t cstride = (size_t)ny * Hp;
    C2<T> *twl = sm + (size_t)3 * tk * LP;
    const TwFetch<T> twf = TwFetch<T>::fetch(p);
    const int colg = threadIdx.x % tk, cg = (threadIdx.x / tk) % 3;
    const int tg = (int)(threadIdx.x / (3 * tk));
    const LineMap mg = {tg, tg < PlanC<LC>::tline(p)};
    C2<T> *lineG = sm + (size_t)(cg * tk + colg) * LP;
    const int l = threadIdx.x / PlanC<LC>::tline(p), t = threadIdx.x % PlanC<LC>::tline(p);
    const LineMap ms = {t, l < 3 * tk};
    C2<T> *line = sm + (size_t)l * LP;
    const int kxg = kx0 + colg;
    C2<T> *colp = S1 + cg * cstride + kxg;
    fft_line<T, false, true, false, LC, SK>(
        p, twl, lineG, mg, line, ms,
        [&](int y) { return (y < ny && kxg < Hx) ? colp[(size_t)y * Hp] : cmk<T>((T)0, (T)0); },
        [](int, C2<T>) {}, [&] { twf.store(twl); });
    for (int e = threadIdx.x; e < tk * L; e += blockDim.x) {
        const int col = e % tk, ky = e / tk, kx = kx0 + col;
        if (kx < Hx) {
            T k[6];
            kernel_at<T>(K, plane, Hp, Hy2, L, 1, kx, ky, 0, k);
            mult3<T>(sm[(0 * tk + col) * LP + sk<SK>(ky)], sm[(1 * tk + col) * LP + sk<SK>(ky)], sm[(2 * tk + col) * LP + sk<SK>(ky)], k);
        }
    }
    __syncthreads();
    fft_line<T, true, false, true, LC, SK>(
        p, twl, lineG, mg, line, ms, [](int) { return cmk<T>((T)0, (T)0); },
        [&](int y, C2<T> v) {
            if (y < ny && kxg < Hx) colp[(size_t)y * Hp] = v;
        });
}

template <typename T, int LC = 0>
__global__ void k_ymi2d(C2<T> *__restrict__ S1, const T *__restrict__ K, size_t plane, int Hx, int Hp, int Hy2,
                        int ny, FftPlanDev p, int tk)
{
    ymi2d_body<T, LC>(S1, K, plane, Hx, Hp, Hy2, ny, p, tk, (int)blockIdx.x);
}

// x inverse (C2R): rows (c, z, y) of S1, Hermitian extension, keep x < nx
template <typename T, typename TOUT, int LC = 0>
__global__ void k_xinv(const C2<T> *__restrict__ in, TOUT *__restrict__ out, size_t out_cstride, int Hx,
                       int Hp, int ny, int nz, int nrows, int nx, FftPlanDev p, int lpw)
{
    constexpr int SK = MM_SKEW_X;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int L = PlanC<LC>::L(p);
    const int l = threadIdx.x / PlanC<LC>::tline(p), t = threadIdx.x % PlanC<LC>::tline(p);
    const int g = blockIdx.x * lpw + l;
    const bool active = g < nrows;
    C2<T> *line = sm + (size_t)l * lds_len<SK>(L);
    C2<T> *twl = sm + (size_t)lpw * lds_len<SK>(L);
    const TwFetch<T> twf = TwFetch<T>::fetch(p);
    const int c = active ? g / (ny * nz) : 0, r = active ? g % (ny * nz) : 0;
    const C2<T> *src = in + (size_t)(active ? g : 0) * Hp;
    TOUT *dst = out + c * out_cstride + (size_t)r * nx;
    const LineMap m = {t, active};
    fft_line<T, true, true, true, LC, SK>(
        p, twl, line, m, line, m,
        [&](int k) {
            // one unconditional load per element (the conjugate by an exact
            // sign): a load in each arm of a branch made the compiler wait
            // for every element's load before issuing the next
            const bool lo = k < Hx;
            const C2<T> v = src[lo ? k : L - k];
            return cmk<T>(v.x, (lo ? (T)1 : (T)-1) * v.y);
        },
        [&](int x, C2<T> v) {
            if (x < nx) dst[x] = (TOUT)v.x;
        }, [&] { twf.store(twl); });
}

// setup: forward z over full lines (no multiply), all Pz outputs stored
__global__ void k_zfwd_setup(double2 *__restrict__ S, int Hx, int Hp, int Py, FftPlanDev p, int tk)
{
    constexpr int SK = MM_SKEW_Z;
    extern __shared__ char smem_raw[];
    double2 *sm = reinterpret_cast<double2 *>(smem_raw);
    const int L = p.L, LP = lds_len<SK>(L);
    const int kx0 = blockIdx.x * tk, ky = blockIdx.y;
    const size_t zstride = (size_t)Py * Hp;
    double2 *base = S + (size_t)ky * Hp;
    double2 *twl = sm + (size_t)tk * LP;
    load_twiddles<double>(twl, p);
    for (int e = threadIdx.x; e < tk * L; e += blockDim.x) {
        const int col = e % tk, z = e / tk, kx = kx0 + col;
        sm[col * LP + sk<SK>(z)] = kx < Hx ? base[z * zstride + kx] : cmk<double>(0.0, 0.0);
    }
    __syncthreads();
    const int l = threadIdx.x / p.tline, t = threadIdx.x % p.tline;
    lds_fft<double, false, SK>(sm + (size_t)l * LP, p, t, l < tk, twl);
    for (int e = threadIdx.x; e < tk * L; e += blockDim.x) {
        const int col = e % tk, z = e / tk, kx = kx0 + col;
        if (kx < Hx) base[z * zstride + kx] = sm[col * LP + sk<SK>(z)];
    }
}

// setup: real part of the octant ky <= Py/2, kz <= Pz/2, scaled
template <typename T>
__global__ void k_extract_kernel(const double2 *__restrict__ S, T *__restrict__ K, int Hx, int Hp, int Py,
                                 int Hy2, int Hz2, double scale)
{
    const size_t n = (size_t)Hx * Hy2 * Hz2;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(q % Hx);
        const size_t r = q / Hx;
        const int ky = (int)(r % Hy2), kz = (int)(r / Hy2);
        K[((size_t)kz * Hy2 + ky) * Hp + kx] = (T)(scale * S[((size_t)kz * Py + ky) * Hp + kx].x);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int grid_1d(size_t n, int bs)
{
    size_t b = (n + bs - 1) / bs;
    return (int)std::min<size_t>(b, 65535 * 64);
}

// multiplier for udiv: n / d for every n < nmax (0: not exact, divide;
// wide: always divide)
static unsigned mdiv(unsigned d, size_t nmax, bool wide = false)
{
    MM_REQUIRE(d >= 1, "zero divisor");
    if (d == 1 || (unsigned long long)nmax * d >= (1ull << 32) || wide) return 0u;
    return (unsigned)((1ull << 32) / d + 1);
}

static int floor_pow2(int v) { return v < 1 ? 1 : 1 << (31 - __builtin_clz((unsigned)v)); }

static int env_int(const char *name, int dflt)
{
    const char *v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

// Launch shapes: x passes take several rows per workgroup; the strided y and
// z passes take tiles of adjacent kx columns (row segments of tile*16 or
// tile*8 bytes).  Tiles are sized for several resident workgroups per CU so
// that global loads of one overlap the FFT of another.  MM_LPW_X, MM_TK_Y and
// MM_TK_Z override the choice (launch-shape sweeps).
// LDS bytes of `nlines` transform lines of length L plus the plan's twiddle table
template <int SK> static size_t lines_lds(int nlines, int L, size_t csize)
{
    return ((size_t)nlines * lds_len<SK>(L) + (size_t)L) * csize;
}

// ---------------------------------------------------------------------------
// launch-shape table ("wisdom")
// ---------------------------------------------------------------------------
static std::mutex g_wis_mu;
static std::map<std::string, MmShape> g_wisdom;
static thread_local MmShape g_candidate;
static thread_local bool g_have_candidate = false;

// "<arch>/<CUs>" of the current device, e.g. gfx950/256
static std::string device_tag()
{
    static std::mutex mu;
    static std::map<int, std::string> tags;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    auto it = tags.find(dev);
    if (it != tags.end()) return it->second;
    hipDeviceProp_t p;
    std::string tag = "unknown";
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
        std::string arch = p.gcnArchName;
        tag = arch.substr(0, arch.find(':')) + "/" + std::to_string(p.multiProcessorCount);
    }
    tags[dev] = tag;
    return tag;
}

static std::string shape_key(int nx, int ny, int nz, size_t real_bytes)
{
    return std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) +
           (real_bytes == 4 ? " fp32 " : " fp64 ") + device_tag();
}

MmShape mm_shape_for(int nx, int ny, int nz, size_t real_bytes)
{
    if (g_have_candidate) return g_candidate;
    const std::string key = shape_key(nx, ny, nz, real_bytes);
    std::lock_guard<std::mutex> lk(g_wis_mu);
    auto it = g_wisdom.find(key);
    return it != g_wisdom.end() ? it->second : MmShape{0, 0, 0, 0, 0};
}

void mm_shape_set_candidate(const MmShape *s)
{
    g_have_candidate = s != nullptr;
    if (s) g_candidate = *s;
}

void mm_wisdom_put(int nx, int ny, int nz, size_t real_bytes, const MmShape &s)
{
    const std::string key = shape_key(nx, ny, nz, real_bytes);
    std::lock_guard<std::mutex> lk(g_wis_mu);
    g_wisdom[key] = s;
}

// File format: '#' comment lines, then one entry per line,
// "<nx>x<ny>x<nz> <fp32|fp64> <arch>/<CUs> lpw_x tk_y tk_z rpw stage_threads".
// An import reads the whole file first and takes nothing from an invalid one.
extern "C" int mm_wisdom_import(const char *path, int *n_entries)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(path, "null argument");
    std::ifstream f(path);
    MM_REQUIRE((bool)f, std::string("cannot open wisdom file ") + path);
    std::map<std::string, MmShape> got;
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        const size_t b = line.find_first_not_of(" \t\r");
        if (b == std::string::npos || line[b] == '#') continue;
        std::istringstream ls(line);
        std::string mesh, prec, dev, extra;
        MmShape v;
        const bool ok = (bool)(ls >> mesh >> prec >> dev >> v.lpw_x >> v.tk_y >> v.tk_z >> v.rpw >> v.stage_threads) &&
                        !(ls >> extra);
        int nx = 0, ny = 0, nz = 0;
        char x1 = 0, x2 = 0;
        std::istringstream ms(mesh);
        const bool mesh_ok = (bool)(ms >> nx >> x1 >> ny >> x2 >> nz) && x1 == 'x' && x2 == 'x' && nx > 0 && ny > 0 &&
                             nz > 0 && !(ms >> extra);
        const bool vals_ok = v.lpw_x >= 0 && v.tk_y >= 0 && v.tk_z >= 0 && v.rpw >= 0 && v.stage_threads >= 0 &&
                             v.lpw_x <= 1024 && v.tk_y <= 1024 && v.tk_z <= 1024 && v.rpw <= 1024 &&
                             v.stage_threads <= 1024;
        MM_REQUIRE(ok && mesh_ok && vals_ok && (prec == "fp32" || prec == "fp64"),
                   std::string("invalid wisdom file ") + path + " at line " + std::to_string(lineno));
        got[mesh + " " + prec + " " + dev] = v;
    }
    {
        std::lock_guard<std::mutex> lk(g_wis_mu);
        for (const auto &kv : got) g_wisdom[kv.first] = kv.second;
    }
    if (n_entries) *n_entries = (int)got.size();
    MM_CAPI_END
}

extern "C" int mm_wisdom_export(const char *path)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(path, "null argument");
    std::ostringstream os;
    os << "# libmmhip launch-shape wisdom: mesh precision device lpw_x tk_y tk_z rpw stage_threads\n";
    {
        std::lock_guard<std::mutex> lk(g_wis_mu);
        for (const auto &kv : g_wisdom)
            os << kv.first << " " << kv.second.lpw_x << " " << kv.second.tk_y << " " << kv.second.tk_z << " "
               << kv.second.rpw << " " << kv.second.stage_threads << "\n";
    }
    std::ofstream f(path);
    MM_REQUIRE((bool)f, std::string("cannot write wisdom file ") + path);
    f << os.str();
    MM_REQUIRE((bool)f, std::string("cannot write wisdom file ") + path);
    MM_CAPI_END
}

extern "C" int mm_wisdom_forget(void)
{
    MM_CAPI_BEGIN
    std::lock_guard<std::mutex> lk(g_wis_mu);
    g_wisdom.clear();
    MM_CAPI_END
}

static ConvLaunch choose_launch(const ConvGeom &g, const FftPlanHost &px, const FftPlanHost &py,
                                const FftPlanHost &pz, size_t csize)
{
    const int lds = mm_max_lds_bytes();
    const MmShape w = mm_shape_for(g.nx, g.ny, g.nz, csize / 2);
    ConvLaunch ln;
    ln.lpw_x = env_int("MM_LPW_X", w.lpw_x > 0 ? w.lpw_x : std::max(1, 256 / px.tline));
    while (ln.lpw_x > 1 && (lines_lds<MM_SKEW_X>(ln.lpw_x, px.L, csize) > (size_t)lds || ln.lpw_x * px.tline > 1024)) --ln.lpw_x;
    const int nl_y = g.is2d ? 3 : 1;
    // default tile widths by line length, measured on MI355X (FP32 steps/s,
    // same-run A/B): y 8 columns, 2 for lines of 2048 or more (1024x1024x2
    // cells: +3 %; 256x256x16: 8 best; 512x512x64: 8 better than 4 by 4 %);
    // z 512 / Pz columns within [16, 64] (256x256x16: 16 best of 8..64;
    // 1024x1024x2, Pz 4: 64 +4 % against 16), halved while a workgroup
    // needs more than 64 KiB of LDS (512x512x64 FP64, Pz 128: 8 columns,
    // 56 KiB, +11 % against 16 at 111 KiB, one workgroup per CU; FP32 keeps
    // 16 at 55 KiB, 8 and 4 slower)
    const int tk_y0 = g.Py >= 2048 ? 2 : 8;
    int tk_z0 = std::max(16, std::min(64, 512 / std::max(1, g.Pz)));
    while (tk_z0 > 1 && lines_lds<MM_SKEW_Z>(3 * tk_z0, pz.L, csize) > 65536) tk_z0 /= 2;
    int tk = floor_pow2(env_int("MM_TK_Y", w.tk_y > 0 ? w.tk_y : tk_y0));
    while (tk > 1 && ((size_t)nl_y * tk * py.tline > 1024 || lines_lds<MM_SKEW_Y>(nl_y * tk, py.L, csize) > (size_t)lds)) tk /= 2;
    ln.tk_y = tk;
    MM_REQUIRE((size_t)nl_y * tk * py.tline <= 1024 && lines_lds<MM_SKEW_Y>(nl_y * tk, py.L, csize) <= (size_t)lds,
               "y transform length too large for one workgroup");
    int tz = floor_pow2(env_int("MM_TK_Z", w.tk_z > 0 ? w.tk_z : tk_z0));
    while (tz > 1 && ((size_t)3 * tz * pz.tline > 1024 || lines_lds<MM_SKEW_Z>(3 * tz, pz.L, csize) > (size_t)lds)) tz /= 2;
    ln.tk_z = tz;
    return ln;
}

template <typename T>
void ConvImpl<T>::init(const double *N_host, int nx, int ny, int nz, int ex, int ey, int ez)
{
    MM_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad mesh size");
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.Px = mm_fft_size_for(nx, ex);
    g.Py = mm_fft_size_for(ny, ey);
    g.Pz = mm_fft_size_for(nz, ez);
    g.Hx = g.Px / 2 + 1;
    // spectral row pitch: a multiple of 16 complex values, so that rows and
    // the column tiles of the y/z passes start on 128-byte boundaries
    g.Hp = (g.Hx + 15) / 16 * 16;
    g.Hy2 = g.Py / 2 + 1;
    g.Hz2 = g.Pz / 2 + 1;
    g.is2d = (nz == 1 && g.Pz == 1);
    px.build(g.Px);
    py.build(g.Py);
    pz.build(g.Pz);
    ln = choose_launch(g, px, py, pz, sizeof(C2<T>));
    // MM_FFT_GENERIC=1: runtime FFT plans for every length (tests compare
    // them with the compile-time plans)
    {
        const char *e = std::getenv("MM_FFT_GENERIC");
        fft_generic = e && std::atoi(e) != 0;
    }
    // MM_WIDE_INDEX=1: the general forms of the index arithmetic (integer
    // division, 64-bit element addressing) that grids beyond the 32-bit
    // ranges take, for any grid (tests compare both; the results are equal)
    wide = env_int("MM_WIDE_INDEX", 0) != 0;
    // MM_XCD_REMAP=0: hardware workgroup order for the tiles (measurements)
    xcd_remap = env_int("MM_XCD_REMAP", 1) != 0 ? 1 : 0;
    // pruned y transforms when the padding is at least half of the line
    {
        // measured on MI355X (256x256x16 cells): slower than the full-length
        // y transforms in FP32 (48 vs 38 us per launch), equal in FP64; opt-in
        const char *e = std::getenv("MM_PRUNE_Y");
        const bool want = e && std::atoi(e) != 0;
        prune_y = want && !g.is2d && g.Py % 2 == 0 && ny <= g.Py / 2 && g.Py / 2 >= 8;
        if (prune_y) {
            pyh.build(g.Py / 2);
            const int lds = mm_max_lds_bytes();
            const size_t cs = sizeof(C2<T>);
            int tk = env_int("MM_TK_YP", (int)(128 / cs));
            while (tk > 1 && ((size_t)tk * pyh.tline > 1024 || lines_lds<MM_SKEW_Y>(tk, pyh.L, cs) > (size_t)lds)) tk /= 2;
            tk_yp = tk;
        }
    }
    MM_REQUIRE((size_t)px.tline <= 1024, "x transform length too large");

    // re-embed the tensor into the radix-friendly padded grid
    const int Px = g.Px, Py = g.Py, Pz = g.Pz;
    auto map_axis = [](int X, int e, int n, int P) -> int {
        if (X < n) return X;
        if (X > e - n) return P - (e - X);
        return -1;
    };
    const size_t plane_full = (size_t)Px * Py * Pz;
    std::vector<double> Nr((size_t)6 * plane_full, 0.0);
    for (int Z = 0; Z < ez; ++Z) {
        const int z2 = map_axis(Z, ez, nz, Pz);
        if (z2 < 0) continue;
        for (int Y = 0; Y < ey; ++Y) {
            const int y2 = map_axis(Y, ey, ny, Py);
            if (y2 < 0) continue;
            for (int X = 0; X < ex; ++X) {
                const int x2 = map_axis(X, ex, nx, Px);
                if (x2 < 0) continue;
                const double *src = N_host + 6 * ((size_t)X + (size_t)ex * ((size_t)Y + (size_t)ey * Z));
                const size_t dst = (size_t)x2 + (size_t)Px * ((size_t)y2 + (size_t)Py * z2);
                for (int e = 0; e < 6; ++e) Nr[e * plane_full + dst] = src[e];
            }
        }
    }

    // Fourier transform of each component in FP64 on the device
    const size_t plane = (size_t)g.Hp * g.Hy2 * g.Hz2;
    kern_elems = 6 * plane;
    MM_HIP(hipMalloc(&kern, sizeof(T) * kern_elems));
    MM_HIP(hipMemset(kern, 0, sizeof(T) * kern_elems));
    double *dN = nullptr;
    double2 *T1 = nullptr, *T2 = nullptr;
    const size_t nspec = (size_t)g.Hp * Py * Pz;
    MM_HIP(hipMalloc(&dN, sizeof(double) * plane_full));
    MM_HIP(hipMalloc(&T1, sizeof(double2) * nspec));
    MM_HIP(hipMalloc(&T2, sizeof(double2) * nspec));
    const int lds = mm_max_lds_bytes();
    const FftPlanDev dx = px.dev<double>(), dy = py.dev<double>(), dz = pz.dev<double>();
    int lpw = std::max(1, 256 / px.tline);
    while (lpw > 1 && lines_lds<MM_SKEW_X>(lpw, px.L, 16) > (size_t)lds) --lpw;
    int tky = 8;
    while (tky > 1 && ((size_t)tky * py.tline > 1024 || lines_lds<MM_SKEW_Y>(tky, py.L, 16) > (size_t)lds)) tky /= 2;
    int tkz = 8;
    while (tkz > 1 && ((size_t)tkz * pz.tline > 1024 || lines_lds<MM_SKEW_Z>(tkz, pz.L, 16) > (size_t)lds)) tkz /= 2;
    const size_t shx = lines_lds<MM_SKEW_X>(lpw, px.L, 16), shy = lines_lds<MM_SKEW_Y>(tky, py.L, 16),
                 shz = lines_lds<MM_SKEW_Z>(tkz, pz.L, 16);
    set_lds(k_xfwd<double, double>, shx);
    set_lds(k_yfwd<double>, shy);
    set_lds(k_zfwd_setup, shz);
    const double scale = 1.0 / ((double)Px * (double)Py * (double)Pz);
    for (int e = 0; e < 6; ++e) {
        MM_HIP(hipMemcpy(dN, Nr.data() + e * plane_full, sizeof(double) * plane_full, hipMemcpyHostToDevice));
        const int nrows = Py * Pz;
        hipLaunchKernelGGL((k_xfwd<double, double>), dim3(((nrows + 1) / 2 + lpw - 1) / lpw), dim3(lpw * px.tline), shx, 0,
                           dN, 0, T1, 0, g.Hx, g.Hp, Py, Pz, nrows, Px, dx, lpw, (const double *)nullptr,
                           (const int *)nullptr);
        MM_HIP(hipGetLastError());
        const unsigned ntyk = (unsigned)((g.Hx + tky - 1) / tky), nbk = ntyk * Pz;
        hipLaunchKernelGGL((k_yfwd<double>), dim3(ntyk, Pz, 1), dim3(tky * py.tline), shy, 0, T1, T2, g.Hx, g.Hp, Py,
                           Pz, dy, tky, 0, mdiv(ntyk, nbk), mdiv((unsigned)Pz, nbk));
        MM_HIP(hipGetLastError());
        if (Pz > 1) {
            hipLaunchKernelGGL(k_zfwd_setup, dim3((g.Hx + tkz - 1) / tkz, Py, 1), dim3(tkz * pz.tline), shz, 0,
                               T2, g.Hx, g.Hp, Py, dz, tkz);
            MM_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL((k_extract_kernel<T>), dim3(grid_1d(plane, 256)), dim3(256), 0, 0, T2, kern + e * plane,
                           g.Hx, g.Hp, Py, g.Hy2, g.Hz2, scale);
        MM_HIP(hipGetLastError());
    }
    MM_HIP(hipDeviceSynchronize());
    (void)hipFree(dN);
    (void)hipFree(T1);
    (void)hipFree(T2);

    MM_HIP(hipMalloc(&S1, sizeof(C2<T>) * 3 * (size_t)nz * ny * g.Hp));
    if (!g.is2d) MM_HIP(hipMalloc(&S2, sizeof(C2<T>) * 3 * (size_t)nz * Py * g.Hp));

    const size_t cs = sizeof(C2<T>);
    set_lds(k_xfwd<T, T>, lines_lds<MM_SKEW_X>(ln.lpw_x, px.L, cs));
    set_lds(k_xinv<T, T>, lines_lds<MM_SKEW_X>(ln.lpw_x, px.L, cs));
    if (g.is2d) {
        set_lds(k_ymi2d<T>, lines_lds<MM_SKEW_Y>(3 * ln.tk_y, py.L, cs));
    } else {
        set_lds(k_yfwd<T>, lines_lds<MM_SKEW_Y>(ln.tk_y, py.L, cs));
        set_lds(k_yinv<T>, lines_lds<MM_SKEW_Y>(ln.tk_y, py.L, cs));
        set_lds(k_zmi<T>, lines_lds<MM_SKEW_Z>(3 * ln.tk_z, pz.L, cs));
    }
}

template <typename T> void ConvImpl<T>::release()
{
    if (kern) (void)hipFree(kern);
    if (S1) (void)hipFree(S1);
    if (S2) (void)hipFree(S2);
    kern = nullptr;
    S1 = S2 = nullptr;
    px.release();
    py.release();
    pz.release();
    pyh.release();
}

// algorithmic HBM bytes per launch (each array read or written once)
template <typename T> static double conv_bytes(const ConvGeom &g, int cls)
{
    const double b = sizeof(T), c = 2.0 * sizeof(T);
    const double N = (double)g.nx * g.ny * g.nz;
    const double s1 = 3.0 * c * g.Hx * g.ny * g.nz;  // S1 planes
    const double s2 = 3.0 * c * g.Hx * g.Py * g.nz;  // S2 planes
    const double K = 6.0 * b * g.Hx * g.Hy2 * g.Hz2;  // real octant kernel
    switch (cls) {
    case KC_XFWD: return 3.0 * b * N + s1;
    case KC_YFWD: return s1 + s2;
    case KC_ZMI: return 2.0 * s2 + K;
    case KC_YINV: return s2 + s1;
    case KC_YMI2D: return 2.0 * s1 + K;
    default: return 0.0;
    }
}

template <typename T> void ConvImpl<T>::forward_x(const T *M, hipStream_t s, const T *M_alt, const int *sel)
{
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const int nrows = 3 * g.ny * g.nz;
    const size_t sh = lines_lds<MM_SKEW_X>(ln.lpw_x, px.L, sizeof(C2<T>));
    mm_with_len(px.L, fft_generic, [&](auto lc) {
        hipLaunchKernelGGL((k_xfwd<T, T, decltype(lc)::value>), dim3(((nrows + 1) / 2 + ln.lpw_x - 1) / ln.lpw_x),
                           dim3(ln.lpw_x * px.tline), sh, s, M, n, S1, (size_t)g.ny * g.nz * g.Hp, g.Hx, g.Hp, g.ny,
                           g.nz, nrows, g.nx, px.dev<T>(), ln.lpw_x, M_alt, sel);
    });
    MM_HIP(hipGetLastError());
    if (klog) klog->mark(s, KC_XFWD, conv_bytes<T>(g, KC_XFWD));
}

template <typename T> void ConvImpl<T>::spectral(hipStream_t s)
{
    const size_t cs = sizeof(C2<T>);
    const size_t plane = (size_t)g.Hp * g.Hy2 * g.Hz2;
    if (g.is2d) {
        const int tk = ln.tk_y;
        const size_t sh = lines_lds<MM_SKEW_Y>(3 * tk, py.L, cs);
        mm_with_len(py.L, fft_generic, [&](auto lc) {
            hipLaunchKernelGGL((k_ymi2d<T, decltype(lc)::value>), dim3((g.Hx + tk - 1) / tk), dim3(3 * tk * py.tline),
                               sh, s, S1, kern, plane, g.Hx, g.Hp, g.Hy2, g.ny, py.dev<T>(), tk);
        });
        MM_HIP(hipGetLastError());
        if (klog) klog->mark(s, KC_YMI2D, conv_bytes<T>(g, KC_YMI2D));
        return;
    }
    const int tk = ln.tk_y;
    const size_t shy = lines_lds<MM_SKEW_Y>(tk, py.L, cs);
    // S2 (3 nz Py rows) is the larger of the two arrays
    const bool y_o32 =
        (size_t)g.Hp * cs < (1u << 24) && (size_t)3 * g.nz * g.Py * g.Hp * cs < (1ull << 32) && !wide;
    const int tkp = tk_yp;
    const size_t shyp = lines_lds<MM_SKEW_Y>(tkp, pyh.L, cs);
    if (prune_y) {
        hipLaunchKernelGGL((k_yfwd_pr<T>), dim3((g.Hx + tkp - 1) / tkp, g.nz, 3), dim3(tkp * pyh.tline), shyp, s, S1, S2,
                           g.Hx, g.Hp, g.ny, g.nz, pyh.dev<T>(), (const C2<T> *)py.plain<T>(), tkp);
    } else {
        mm_with_len(py.L, fft_generic, [&](auto lc) {
            const unsigned ntx = (unsigned)((g.Hx + tk - 1) / tk), nb = ntx * g.nz * 3;
            if (y_o32)
                hipLaunchKernelGGL((k_yfwd<T, decltype(lc)::value, true>), dim3(ntx, g.nz, 3), dim3(tk * py.tline), shy,
                                   s, S1, S2, g.Hx, g.Hp, g.ny, g.nz, py.dev<T>(), tk, xcd_remap, mdiv(ntx, nb, wide),
                                   mdiv((unsigned)g.nz, nb, wide));
            else
                hipLaunchKernelGGL((k_yfwd<T, decltype(lc)::value, false>), dim3(ntx, g.nz, 3), dim3(tk * py.tline),
                                   shy, s, S1, S2, g.Hx, g.Hp, g.ny, g.nz, py.dev<T>(), tk, xcd_remap, mdiv(ntx, nb, wide),
                                   mdiv((unsigned)g.nz, nb, wide));
        });
    }
    MM_HIP(hipGetLastError());
    if (klog) klog->mark(s, KC_YFWD, conv_bytes<T>(g, KC_YFWD));
    const int tz = ln.tk_z;
    const size_t shz = lines_lds<MM_SKEW_Z>(3 * tz, pz.L, cs);
    MM_REQUIRE((size_t)tz * (pz.L / 2 + 1) <= (size_t)ZM_ITEMS * 3 * tz * pz.tline, "z pass items per thread");
    const unsigned ntz = (unsigned)((g.Hx + tz - 1) / tz);
    const unsigned mtz = mdiv(ntz, (size_t)ntz * g.Py, wide);
    const size_t zsb = (size_t)g.Py * g.Hp * cs;
    const bool o32 = zsb < (1u << 24) && 3 * (size_t)g.nz * zsb < (1ull << 32) &&
                     kern_elems * sizeof(T) < (1ull << 32) && !wide;
    mm_with_len(pz.L, fft_generic, [&](auto lc) {
        if (o32)
            hipLaunchKernelGGL((k_zmi<T, decltype(lc)::value, true>), dim3(ntz, g.Py), dim3(3 * tz * pz.tline), shz, s,
                               S2, kern, plane, g.Hx, g.Hp, g.Hy2, g.Py, g.nz, pz.dev<T>(), tz, xcd_remap, mtz,
                               (unsigned)zsb, (unsigned)(plane * sizeof(T)));
        else
            hipLaunchKernelGGL((k_zmi<T, decltype(lc)::value, false>), dim3(ntz, g.Py), dim3(3 * tz * pz.tline), shz,
                               s, S2, kern, plane, g.Hx, g.Hp, g.Hy2, g.Py, g.nz, pz.dev<T>(), tz, xcd_remap, mtz, 0u,
                               0u);
    });
    MM_HIP(hipGetLastError());
    if (klog) klog->mark(s, KC_ZMI, conv_bytes<T>(g, KC_ZMI));
    if (prune_y) {
        hipLaunchKernelGGL((k_yinv_pr<T>), dim3((g.Hx + tkp - 1) / tkp, g.nz, 3), dim3(tkp * pyh.tline), shyp, s, S2, S1,
                           g.Hx, g.Hp, g.ny, g.nz, pyh.dev<T>(), (const C2<T> *)py.plain<T>(), tkp);
    } else {
        mm_with_len(py.L, fft_generic, [&](auto lc) {
            const unsigned ntx = (unsigned)((g.Hx + tk - 1) / tk), nb = ntx * g.nz * 3;
            if (y_o32)
                hipLaunchKernelGGL((k_yinv<T, decltype(lc)::value, true>), dim3(ntx, g.nz, 3), dim3(tk * py.tline), shy,
                                   s, S2, S1, g.Hx, g.Hp, g.ny, g.nz, py.dev<T>(), tk, xcd_remap, mdiv(ntx, nb, wide),
                                   mdiv((unsigned)g.nz, nb, wide));
            else
                hipLaunchKernelGGL((k_yinv<T, decltype(lc)::value, false>), dim3(ntx, g.nz, 3), dim3(tk * py.tline),
                                   shy, s, S2, S1, g.Hx, g.Hp, g.ny, g.nz, py.dev<T>(), tk, xcd_remap, mdiv(ntx, nb, wide),
                                   mdiv((unsigned)g.nz, nb, wide));
        });
    }
    MM_HIP(hipGetLastError());
    if (klog) klog->mark(s, KC_YINV, conv_bytes<T>(g, KC_YINV));
}

template <typename T> void ConvImpl<T>::inverse_x(T *H, hipStream_t s)
{
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const int nrows = 3 * g.ny * g.nz;
    const size_t sh = lines_lds<MM_SKEW_X>(ln.lpw_x, px.L, sizeof(C2<T>));
    mm_with_len(px.L, fft_generic, [&](auto lc) {
        hipLaunchKernelGGL((k_xinv<T, T, decltype(lc)::value>), dim3((nrows + ln.lpw_x - 1) / ln.lpw_x),
                           dim3(ln.lpw_x * px.tline), sh, s, S1, H, n, g.Hx, g.Hp, g.ny, g.nz, nrows, g.nx,
                           px.dev<T>(), ln.lpw_x);
    });
    MM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// batched passes
// ---------------------------------------------------------------------------
template <typename T>
int ConvImpl<T>::batch_x(BxDesc<T> &d, int threads, size_t &lds, const T *y0, const T *y1, const int *sel,
                         const int *stop) const
{
    const size_t cs = sizeof(C2<T>);
    MM_REQUIRE(threads >= px.tline, "batch launch has fewer threads than one x line needs");
    int lpw = std::max(1, threads / px.tline);
    while (lpw > 1 && lines_lds<MM_SKEW_X>(lpw, px.L, cs) > (size_t)mm_max_lds_bytes()) --lpw;
    const int nrows = 3 * g.ny * g.nz;
    d.y0 = y0;
    d.y1 = y1;
    d.sel = sel;
    d.stop = stop;
    d.S1 = S1;
    d.in_cstride = (size_t)g.nx * g.ny * g.nz;
    d.out_cstride = (size_t)g.ny * g.nz * g.Hp;
    d.Hx = g.Hx;
    d.Hp = g.Hp;
    d.ny = g.ny;
    d.nz = g.nz;
    d.nrows = nrows;
    d.nx = g.nx;
    d.lpw = lpw;
    d.tbase = 0;
    d.p = px.template dev<T>();
    lds = lines_lds<MM_SKEW_X>(lpw, px.L, cs);
    return ((nrows + 1) / 2 + lpw - 1) / lpw;
}

template <typename T>
int ConvImpl<T>::batch_y(ByDesc<T> &d, int threads, size_t &lds, bool &o32, const int *stop) const
{
    const size_t cs = sizeof(C2<T>);
    MM_REQUIRE(!g.is2d && threads >= py.tline, "batch y pass: 2-D system or too few threads");
    int tk = std::min(16, floor_pow2(std::max(1, threads / py.tline)));
    while (tk > 1 && lines_lds<MM_SKEW_Y>(tk, py.L, cs) > (size_t)mm_max_lds_bytes()) tk /= 2;
    const unsigned ntx = (unsigned)((g.Hx + tk - 1) / tk), nb = ntx * g.nz * 3;
    d.S1 = S1;
    d.S2 = S2;
    d.stop = stop;
    d.Hx = g.Hx;
    d.Hp = g.Hp;
    d.ny = g.ny;
    d.nz = g.nz;
    d.tk = tk;
    d.ntx = (int)ntx;
    d.tbase = 0;
    d.mtx = mdiv(ntx, nb);
    d.mzl = mdiv((unsigned)g.nz, nb);
    d.p = py.template dev<T>();
    o32 = (size_t)g.Hp * cs < (1u << 24) && (size_t)3 * g.nz * g.Py * g.Hp * cs < (1ull << 32);
    lds = lines_lds<MM_SKEW_Y>(tk, py.L, cs);
    return (int)nb;
}

template <typename T>
int ConvImpl<T>::batch_z(BzDesc<T> &d, int threads, size_t &lds, bool &o32, const int *stop) const
{
    const size_t cs = sizeof(C2<T>);
    MM_REQUIRE(!g.is2d && threads >= 3 * pz.tline, "batch z pass: 2-D system or too few threads");
    // the system's own z tile width (wide for short z lines), within the
    // launch's threads
    int tz = std::min(ln.tk_z, floor_pow2(std::max(1, threads / (3 * pz.tline))));
    while (tz > 1 && lines_lds<MM_SKEW_Z>(3 * tz, pz.L, cs) > (size_t)mm_max_lds_bytes()) tz /= 2;
    MM_REQUIRE((size_t)tz * (pz.L / 2 + 1) <= (size_t)ZM_ITEMS * threads, "z pass items per thread");
    const unsigned ntz = (unsigned)((g.Hx + tz - 1) / tz), nb = ntz * (unsigned)g.Py;
    const size_t plane = (size_t)g.Hp * g.Hy2 * g.Hz2;
    const size_t zsb = (size_t)g.Py * g.Hp * cs;
    d.S2 = S2;
    d.K = kern;
    d.plane = plane;
    d.stop = stop;
    d.Hx = g.Hx;
    d.Hp = g.Hp;
    d.Hy2 = g.Hy2;
    d.Py = g.Py;
    d.nz = g.nz;
    d.tk = tz;
    d.ntx = (int)ntz;
    d.tbase = 0;
    d.mtx = mdiv(ntz, nb);
    d.zsb = (unsigned)zsb;
    d.pb = (unsigned)(plane * sizeof(T));
    d.p = pz.template dev<T>();
    o32 = zsb < (1u << 24) && 3 * (size_t)g.nz * zsb < (1ull << 32) && kern_elems * sizeof(T) < (1ull << 32);
    lds = lines_lds<MM_SKEW_Z>(3 * tz, pz.L, cs);
    return (int)nb;
}

template <typename T> int ConvImpl<T>::batch_2d(B2Desc<T> &d, int threads, size_t &lds, const int *stop) const
{
    const size_t cs = sizeof(C2<T>);
    MM_REQUIRE(g.is2d && threads >= 3 * py.tline, "batch 2-D pass: 3-D system or too few threads");
    int tk = std::min(16, floor_pow2(std::max(1, threads / (3 * py.tline))));
    while (tk > 1 && lines_lds<MM_SKEW_Y>(3 * tk, py.L, cs) > (size_t)mm_max_lds_bytes()) tk /= 2;
    d.S1 = S1;
    d.K = kern;
    d.plane = (size_t)g.Hp * g.Hy2 * g.Hz2;
    d.stop = stop;
    d.Hx = g.Hx;
    d.Hp = g.Hp;
    d.Hy2 = g.Hy2;
    d.ny = g.ny;
    d.tk = tk;
    d.tbase = 0;
    d.p = py.template dev<T>();
    lds = lines_lds<MM_SKEW_Y>(3 * tk, py.L, cs);
    return (g.Hx + tk - 1) / tk;
}

template struct ConvImpl<float>;
template struct ConvImpl<double>;

// The batched kernels: the global workgroup index in XCD order picks the
// system (tsys) and that system's workgroup; descriptor fields are read
// through global-address-space pointer types (gfield) so that the accesses
// through them are global, not flat, operations.
template <typename T> __global__ void kb_xfwd(const BxDesc<T> *__restrict__ D, const int *__restrict__ tsys)
{
    const int gi = block_tile(1);
    const BxDesc<T> &d = D[tsys[gi]];
    if (gfield(d.stop)[0]) return;
    const T *in = gfield(d.sel)[0] ? gfield(d.y1) : gfield(d.y0);
    xfwd_body<T, T, 0>(in, d.in_cstride, gfield(d.S1), d.out_cstride, d.Hx, d.Hp, d.ny, d.nz, d.nrows, d.nx, d.p,
                       d.lpw, gi - d.tbase);
}

template <typename T, bool INV, bool O32>
__global__ void kb_y(const ByDesc<T> *__restrict__ D, const int *__restrict__ tsys)
{
    const int gi = block_tile(1);
    const ByDesc<T> &d = D[tsys[gi]];
    if (gfield(d.stop)[0]) return;
    if (INV)
        yinv_body<T, 0, O32>(gfield(d.S2), gfield(d.S1), d.Hx, d.Hp, d.ny, d.nz, d.p, d.tk, gi - d.tbase, d.ntx,
                             d.mtx, d.mzl);
    else
        yfwd_body<T, 0, O32>(gfield(d.S1), gfield(d.S2), d.Hx, d.Hp, d.ny, d.nz, d.p, d.tk, gi - d.tbase, d.ntx,
                             d.mtx, d.mzl);
}

template <typename T, bool O32> __global__ void kb_zmi(const BzDesc<T> *__restrict__ D, const int *__restrict__ tsys)
{
    const int gi = block_tile(1);
    const BzDesc<T> &d = D[tsys[gi]];
    if (gfield(d.stop)[0]) return;
    zmi_body<T, 0, O32>(gfield(d.S2), gfield(d.K), d.plane, d.Hx, d.Hp, d.Hy2, d.Py, d.nz, d.p, d.tk, 1, gi - d.tbase,
                        d.ntx, d.mtx, d.zsb, d.pb);
}

template <typename T> __global__ void kb_ymi2d(const B2Desc<T> *__restrict__ D, const int *__restrict__ tsys)
{
    const int gi = block_tile(1);
    const B2Desc<T> &d = D[tsys[gi]];
    if (gfield(d.stop)[0]) return;
    ymi2d_body<T, 0>(gfield(d.S1), gfield(d.K), d.plane, d.Hx, d.Hp, d.Hy2, d.ny, d.p, d.tk, gi - d.tbase);
}

template <typename T> void mm_batch_xfwd(const BatchPass &bp, hipStream_t s)
{
    hipLaunchKernelGGL(kb_xfwd<T>, dim3(bp.ntiles), dim3(bp.threads), bp.lds, s, (const BxDesc<T> *)bp.desc,
                       (const int *)bp.tsys);
    MM_HIP(hipGetLastError());
}
template <typename T> void mm_batch_yfwd(const BatchPass &bp, hipStream_t s)
{
    if (bp.o32)
        hipLaunchKernelGGL((kb_y<T, false, true>), dim3(bp.ntiles), dim3(bp.threads), bp.lds, s,
                           (const ByDesc<T> *)bp.desc, (const int *)bp.tsys);
    else
        hipLaunchKernelGGL((kb_y<T, false, false>), dim3(bp.ntiles), dim3(bp.threads), bp.lds, s,
                           (const ByDesc<T> *)bp.desc, (const int *)bp.tsys);
    MM_HIP(hipGetLastError());
}
template <typename T> void mm_batch_yinv(const BatchPass &bp, hipStream_t s)
{
    if (bp.o32)
        hipLaunchKernelGGL((kb_y<T, true, true>), dim3(bp.ntiles), dim3(bp.threads), bp.lds, s,
                           (const ByDesc<T> *)bp.desc, (const int *)bp.tsys);
    else
        hipLaunchKernelGGL((kb_y<T, true, false>), dim3(bp.ntiles), dim3(bp.threads), bp.lds, s,
                           (const ByDesc<T> *)bp.desc, (const int *)bp.tsys);
    MM_HIP(hipGetLastError());
}
template <typename T> void mm_batch_zmi(const BatchPass &bp, hipStream_t s)
{
    if (bp.o32)
        hipLaunchKernelGGL((kb_zmi<T, true>), dim3(bp.ntiles), dim3(bp.threads), bp.lds, s,
                           (const BzDesc<T> *)bp.desc, (const int *)bp.tsys);
    else
        hipLaunchKernelGGL((kb_zmi<T, false>), dim3(bp.ntiles), dim3(bp.threads), bp.lds, s,
                           (const BzDesc<T> *)bp.desc, (const int *)bp.tsys);
    MM_HIP(hipGetLastError());
}
template <typename T> void mm_batch_ymi2d(const BatchPass &bp, hipStream_t s)
{
    hipLaunchKernelGGL(kb_ymi2d<T>, dim3(bp.ntiles), dim3(bp.threads), bp.lds, s, (const B2Desc<T> *)bp.desc,
                       (const int *)bp.tsys);
    MM_HIP(hipGetLastError());
}
template void mm_batch_xfwd<float>(const BatchPass &, hipStream_t);
template void mm_batch_xfwd<double>(const BatchPass &, hipStream_t);
template void mm_batch_yfwd<float>(const BatchPass &, hipStream_t);
template void mm_batch_yfwd<double>(const BatchPass &, hipStream_t);
template void mm_batch_yinv<float>(const BatchPass &, hipStream_t);
template void mm_batch_yinv<double>(const BatchPass &, hipStream_t);
template void mm_batch_zmi<float>(const BatchPass &, hipStream_t);
template void mm_batch_zmi<double>(const BatchPass &, hipStream_t);
template void mm_batch_ymi2d<float>(const BatchPass &, hipStream_t);
template void mm_batch_ymi2d<double>(const BatchPass &, hipStream_t);

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int mm_conv_create(const double *N_host, int nx, int ny, int nz, int ex, int ey, int ez,
                              int precision, mm_conv_t **out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(N_host && out, "null argument");
    MM_REQUIRE(precision == MM_FP64 || precision == MM_FP32, "precision must be MM_FP64 or MM_FP32");
    mm_conv *c = new mm_conv();
    c->precision = precision;
    try {
        if (precision == MM_FP64) {
            c->c64 = new ConvImpl<double>();
            c->c64->init(N_host, nx, ny, nz, ex, ey, ez);
        } else {
            c->c32 = new ConvImpl<float>();
            c->c32->init(N_host, nx, ny, nz, ex, ey, ez);
        }
    } catch (...) {
        if (c->c64) { c->c64->release(); delete c->c64; }
        if (c->c32) { c->c32->release(); delete c->c32; }
        delete c;
        throw;
    }
    *out = c;
    MM_CAPI_END
}

extern "C" int mm_conv_execute(mm_conv_t *conv, const double *M, double *H, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(conv && M && H, "null argument");
    MM_REQUIRE(conv->precision == MM_FP64, "convolution was created for FP32; use mm_conv_execute_f32");
    conv->c64->execute(M, H, (hipStream_t)stream);
    MM_CAPI_END
}

extern "C" int mm_conv_execute_f32(mm_conv_t *conv, const float *M, float *H, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(conv && M && H, "null argument");
    MM_REQUIRE(conv->precision == MM_FP32, "convolution was created for FP64; use mm_conv_execute");
    conv->c32->execute(M, H, (hipStream_t)stream);
    MM_CAPI_END
}

extern "C" int mm_conv_fft_size(const mm_conv_t *conv, int *fft_xyz)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(conv && fft_xyz, "null argument");
    const ConvGeom &g = conv->c64 ? conv->c64->g : conv->c32->g;
    fft_xyz[0] = g.Px;
    fft_xyz[1] = g.Py;
    fft_xyz[2] = g.Pz;
    MM_CAPI_END
}

extern "C" int mm_conv_destroy(mm_conv_t *conv)
{
    MM_CAPI_BEGIN
    if (!conv) return MM_OK;
    if (conv->c64) { conv->c64->release(); delete conv->c64; }
    if (conv->c32) { conv->c32->release(); delete conv->c32; }
    delete conv;
    MM_CAPI_END
}

// ---------------------------------------------------------------------------
// direct-sum convolution
// ---------------------------------------------------------------------------
// H(r) = sum_r' N((r - r') mod E) M(r') over the whole mesh, the tensor in the
// layout of GenerateDemagTensor (6 entries per point, x fastest).  One thread
// per target cell; the workgroup walks the source cells in tiles staged in
// LDS (M of 256 source cells, three components), the source order of every
// thread z, y, x as the reference's loop.
#define MM_SIMPLE_TILE 256
__global__ void __launch_bounds__(MM_SIMPLE_TILE) k_conv_simple(const double *__restrict__ N, int nx, int ny, int nz,
                                                               int ex, int ey, int ez, const double *__restrict__ M,
                                                               double *__restrict__ H)
{
    __shared__ double ms[3][MM_SIMPLE_TILE];
    const int n = nx * ny * nz;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const bool act = i < n;
    const int x = act ? i % nx : 0, y = act ? (i / nx) % ny : 0, z = act ? i / (nx * ny) : 0;
    double hx = 0.0, hy = 0.0, hz = 0.0;
    for (int j0 = 0; j0 < n; j0 += MM_SIMPLE_TILE) {
        const int jl = j0 + (int)threadIdx.x;
        __syncthreads();
        for (int c = 0; c < 3; ++c) ms[c][threadIdx.x] = jl < n ? M[(size_t)c * n + jl] : 0.0;
        __syncthreads();
        if (!act) continue;
        const int cnt = min(MM_SIMPLE_TILE, n - j0);
        int m = j0 % nx, o = j0 / (nx * ny), nn = (j0 / nx) % ny;
        for (int t = 0; t < cnt; ++t) {
            const int X = (x - m + ex) % ex, Y = (y - nn + ey) % ey, Z = (z - o + ez) % ez;
            const double *Np = N + 6 * ((size_t)X + (size_t)ex * ((size_t)Y + (size_t)ey * Z));
            const double Mx = ms[0][t], My = ms[1][t], Mz = ms[2][t];
            hx += Np[0] * Mx + Np[1] * My + Np[2] * Mz;
            hy += Np[1] * Mx + Np[3] * My + Np[4] * Mz;
            hz += Np[2] * Mx + Np[4] * My + Np[5] * Mz;
            if (++m == nx) {
                m = 0;
                if (++nn == ny) {
                    nn = 0;
                    ++o;
                }
            }
        }
    }
    if (act) {
        H[i] = hx;
        H[(size_t)n + i] = hy;
        H[2 * (size_t)n + i] = hz;
    }
}

struct mm_conv_simple {
    double *N = nullptr;
    int nx = 0, ny = 0, nz = 0, ex = 0, ey = 0, ez = 0;
};

extern "C" int mm_conv_simple_create(const double *N_host, int nx, int ny, int nz, int ex, int ey, int ez,
                                     mm_conv_simple_t **out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(N_host && out, "null argument");
    MM_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && ex >= nx && ey >= ny && ez >= nz, "bad sizes");
    MM_REQUIRE((size_t)nx * ny * nz < (1u << 30), "mesh too large for the direct sum");
    mm_conv_simple *c = new mm_conv_simple();
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->ex = ex;
    c->ey = ey;
    c->ez = ez;
    const size_t bytes = sizeof(double) * 6 * (size_t)ex * ey * ez;
    if (hipMalloc(&c->N, bytes) != hipSuccess) {
        delete c;
        MM_REQUIRE(false, "device allocation failed");
    }
    const hipError_t e = hipMemcpy(c->N, N_host, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(c->N);
        delete c;
        MM_HIP(e);
    }
    *out = c;
    MM_CAPI_END
}

extern "C" int mm_conv_simple_execute(mm_conv_simple_t *c, const double *M, double *H, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(c && M && H, "null argument");
    const int n = c->nx * c->ny * c->nz;
    hipLaunchKernelGGL(k_conv_simple, dim3((n + MM_SIMPLE_TILE - 1) / MM_SIMPLE_TILE), dim3(MM_SIMPLE_TILE), 0,
                       (hipStream_t)stream, c->N, c->nx, c->ny, c->nz, c->ex, c->ey, c->ez, M, H);
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_conv_simple_destroy(mm_conv_simple_t *c)
{
    MM_CAPI_BEGIN
    if (!c) return MM_OK;
    if (c->N) (void)hipFree(c->N);
    delete c;
    MM_CAPI_END
}
