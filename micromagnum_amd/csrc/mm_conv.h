// mm_conv.h -- FFT demag convolution objects (host side of mm_conv.hip).
#ifndef MM_CONV_H
#define MM_CONV_H

#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "mm_fft_device.h"
#include "mm_internal.h"

#define MM_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw mm_error(MM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// host plan of one FFT length: radix list, threads per line, twiddle tables
struct FftPlanHost {
    int L = 1;
    int tline = 1;
    std::vector<int> rad;
    void *tw32 = nullptr;  // float2[L], exp(-2 pi i m / L)
    void *tw64 = nullptr;  // double2[L]
    void *twp32 = nullptr; // float2[ntw], per-pass layout (FftPlanDev)
    void *twp64 = nullptr; // double2[ntw]
    int ntw = 0;
    void build(int len);
    void release();
    template <typename T> FftPlanDev dev() const
    {
        FftPlanDev d;
        d.L = L;
        d.tline = tline;
        d.npass = (int)rad.size();
        d.radices = 0u;
        for (int i = 0; i < d.npass; ++i) d.radices |= (unsigned)rad[i] << (4 * i);
        d.tw = sizeof(T) == 4 ? twp32 : twp64;
        d.ntw = ntw;
        d.pow2 = (L & (L - 1)) == 0 ? 1 : 0;
        return d;
    }
    // the plain table exp(-2 pi i m / L), m < L
    template <typename T> const void *plain() const { return sizeof(T) == 4 ? tw32 : tw64; }
};

// Calls f(std::integral_constant<int, LC>()) with LC = L when kernels carry a
// compile-time plan for that length (powers of two 2..2048) and `generic` is
// false, else LC = 0 (runtime plan).
template <typename F> void mm_with_len(int L, bool generic, F &&f)
{
    if (generic) {
        f(std::integral_constant<int, 0>());
        return;
    }
    switch (L) {
    case 2: f(std::integral_constant<int, 2>()); return;
    case 4: f(std::integral_constant<int, 4>()); return;
    case 8: f(std::integral_constant<int, 8>()); return;
    case 16: f(std::integral_constant<int, 16>()); return;
    case 32: f(std::integral_constant<int, 32>()); return;
    case 64: f(std::integral_constant<int, 64>()); return;
    case 128: f(std::integral_constant<int, 128>()); return;
    case 256: f(std::integral_constant<int, 256>()); return;
    case 512: f(std::integral_constant<int, 512>()); return;
    case 1024: f(std::integral_constant<int, 1024>()); return;
    case 2048: f(std::integral_constant<int, 2048>()); return;
    default: f(std::integral_constant<int, 0>()); return;
    }
}

bool mm_is_smooth(int n);                  // n has no prime factor > 5
int mm_fft_size_for(int n, int given);     // radix-friendly padded size for n cells

// Geometry of a convolution: physical (nx,ny,nz), padded FFT sizes (Px,Py,Pz).
struct ConvGeom {
    int nx, ny, nz, Px, Py, Pz, Hx, Hy2, Hz2;
    int Hp;  // row pitch of the spectra and the kernel planes (>= Hx)
    bool is2d;
};

// Tile widths (lines per workgroup) for the strided passes.
struct ConvLaunch {
    int lpw_x;  // x lines per WG (plain x passes)
    int tk_y;   // kx columns per WG, y passes
    int tk_z;   // kx columns per WG, z pass (3 components each)
};

// Optional per-launch event log used by the measurement entry point: after
// every kernel the owner records an event tagged with the kernel class.
enum {
    KC_XFWD = 0, KC_YFWD = 1, KC_ZMI = 2, KC_YINV = 3, KC_YMI2D = 4, KC_STAGE = 5, KC_CTRL = 6, KC_NCLASS = 7
};
struct KernelLog {
    std::vector<hipEvent_t> ev;
    std::vector<int> cls;
    std::vector<double> bytes;
    size_t used = 0;
    void reserve(size_t n)
    {
        while (ev.size() < n) {
            hipEvent_t e;
            MM_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        cls.resize(n);
        bytes.resize(n);
    }
    void mark(hipStream_t s, int c, double b)
    {
        MM_REQUIRE(used < ev.size(), "kernel log full");
        MM_HIP(hipEventRecord(ev[used], s));
        cls[used] = c;
        bytes[used] = b;
        ++used;
    }
    ~KernelLog()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// ---- batched passes: one launch runs one pass for many systems ----------
// Each system of a batch has a descriptor per pass; the launch's global
// workgroup g (in XCD order, xcd_tile) belongs to system tsys[g] and is that
// system's workgroup g - tbase.  A system whose stop flag is set skips the
// pass.  Lengths differ between systems, so the transforms take the runtime
// FFT plans.
template <typename T> struct BxDesc {  // x forward of the current state y[*sel]
    const T *y0, *y1;
    const int *sel, *stop;
    C2<T> *S1;
    size_t in_cstride, out_cstride;
    int Hx, Hp, ny, nz, nrows, nx, lpw, tbase;
    FftPlanDev p;
};
template <typename T> struct ByDesc {  // y forward S1 -> S2 and inverse S2 -> S1
    C2<T> *S1, *S2;
    const int *stop;
    int Hx, Hp, ny, nz, tk, ntx, tbase;
    unsigned mtx, mzl;
    FftPlanDev p;
};
template <typename T> struct BzDesc {  // z forward, multiply, z inverse on S2
    C2<T> *S2;
    const T *K;
    size_t plane;
    const int *stop;
    int Hx, Hp, Hy2, Py, nz, tk, ntx, tbase;
    unsigned mtx, zsb, pb;
    FftPlanDev p;
};
template <typename T> struct B2Desc {  // 2-D: y forward, multiply, y inverse on S1
    C2<T> *S1;
    const T *K;
    size_t plane;
    const int *stop;
    int Hx, Hp, Hy2, ny, tk, tbase;
    FftPlanDev p;
};

// One pass of a batch: descriptors and tile table in device memory, the
// launch's workgroups, threads per workgroup and dynamic LDS.  o32: every
// system's spectra fit the 32-bit offset forms.
struct BatchPass {
    void *desc = nullptr;
    int *tsys = nullptr;
    int ntiles = 0, threads = 0;
    size_t lds = 0;
    bool o32 = true;
};

template <typename T>
struct ConvImpl {
    KernelLog *klog = nullptr;
    ConvGeom g;
    ConvLaunch ln;
    FftPlanHost px, py, pz;
    FftPlanHost pyh;        // length Py/2, for the pruned y transforms
    bool fft_generic = false;  // runtime FFT plans only (no compile-time lengths)
    int xcd_remap = 1;         // XCD-contiguous tile order (xcd_tile)
    bool prune_y = false;
    bool wide = false;         // general index arithmetic only (MM_WIDE_INDEX)
    int tk_yp = 1;          // kx columns per WG of the pruned y transforms
    T *kern = nullptr;      // 6 real planes [e][kz<Hz2][ky<Hy2][kx<Hx], pitch Hp, scaled 1/(PxPyPz)
    C2<T> *S1 = nullptr;    // [c][z<nz][y<ny][kx<Hx], pitch Hp
    C2<T> *S2 = nullptr;    // [c][z<nz][ky<Py][kx<Hx], pitch Hp   (3-D only)
    size_t kern_elems = 0;

    void init(const double *N_host, int nx, int ny, int nz, int ex, int ey, int ez);
    void release();
    // spectral part: pad+x forward .. x inverse up to S1 holding the result
    // spectrum rows; then x inverse into H.
    void forward_x(const T *M, hipStream_t s, const T *M_alt = nullptr, const int *sel = nullptr);
    void spectral(hipStream_t s);  // y fwd, z fwd*mult*inv / 2-D fused, y inv
    void inverse_x(T *H, hipStream_t s);
    void execute(const T *M, T *H, hipStream_t s)
    {
        forward_x(M, s);
        spectral(s);
        inverse_x(H, s);
    }
    // this system's part of a batched pass launched with `threads` threads
    // per workgroup: its descriptor (tbase left to the caller), workgroup
    // count, the dynamic LDS it needs and whether it fits the 32-bit offset
    // forms.  Tile widths follow the batch's thread count.
    int batch_x(BxDesc<T> &d, int threads, size_t &lds, const T *y0, const T *y1, const int *sel,
                const int *stop) const;
    int batch_y(ByDesc<T> &d, int threads, size_t &lds, bool &o32, const int *stop) const;
    int batch_z(BzDesc<T> &d, int threads, size_t &lds, bool &o32, const int *stop) const;
    int batch_2d(B2Desc<T> &d, int threads, size_t &lds, const int *stop) const;
    // threads per workgroup this system's passes need at least
    int batch_threads_x() const { return px.tline; }
    int batch_threads_y() const { return (g.is2d ? 3 : 1) * py.tline; }
    int batch_threads_z() const { return 3 * pz.tline; }
};

// batched launches (all systems of bp in one launch on stream s)
template <typename T> void mm_batch_xfwd(const BatchPass &bp, hipStream_t s);
template <typename T> void mm_batch_yfwd(const BatchPass &bp, hipStream_t s);
template <typename T> void mm_batch_yinv(const BatchPass &bp, hipStream_t s);
template <typename T> void mm_batch_zmi(const BatchPass &bp, hipStream_t s);
template <typename T> void mm_batch_ymi2d(const BatchPass &bp, hipStream_t s);

// device-side FFT kernels reused by the fused stepper
template <typename T> size_t conv_lds_xinv(const ConvImpl<T> &c, int lines) { return (size_t)lines * (c.px.L + 1) * sizeof(C2<T>); }

struct mm_conv {
    int precision;
    ConvImpl<float> *c32 = nullptr;
    ConvImpl<double> *c64 = nullptr;
};

void mm_lds_attr_once();
int mm_max_lds_bytes();

// Launch shapes of a mesh (0: not set, the built-in choice applies): x lines
// per workgroup, kx columns per workgroup of the y and z passes, rows per
// stage workgroup and stage threads.  Tuned shapes live in a process-wide
// table keyed by mesh size, precision and device (the counterpart of the
// reference's FFTW wisdom, Magneto.cpp:151-190).  mm_shape_for gives the
// shape in force for a new convolution or stepper: the tuner's candidate
// while it measures one (this thread), else the table's entry.  The MM_LPW_X,
// MM_TK_Y, MM_TK_Z, MM_RPW and MM_STAGE_THREADS variables override both.
struct MmShape {
    int lpw_x, tk_y, tk_z, rpw, stage_threads;
};
MmShape mm_shape_for(int nx, int ny, int nz, size_t real_bytes);
void mm_shape_set_candidate(const MmShape *s);
void mm_wisdom_put(int nx, int ny, int nz, size_t real_bytes, const MmShape &s);

#endif
