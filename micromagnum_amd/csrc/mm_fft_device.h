// mm_fft_device.h -- LDS-resident Stockham FFT building blocks for gfx950.
//
// One line of length L lives in LDS; `tline` threads of the workgroup work on
// it.  Each pass of radix R (8, 5, 4, 3, 2) reads the R inputs of every
// butterfly into registers, synchronises the workgroup, applies the twiddles
// W_{Ns R}^{r k} from a per-length table and a hand-written size-R DFT, and
// writes the results to their autosorted positions (Stockham), so no
// bit-reversal pass is needed and the line stays in place.  Forward uses
// e^{-i}, inverse e^{+i}; neither is normalised (the 1/(Px Py Pz) factor lives
// in the Fourier kernel, as in the reference's TensorFieldSetup.cpp:58).
#ifndef MM_FFT_DEVICE_H
#define MM_FFT_DEVICE_H

#include <hip/hip_runtime.h>

#define MM_FFT_MAX_PASSES 8

template <typename T> struct cvec;
template <> struct cvec<float> { typedef float2 type; };
template <> struct cvec<double> { typedef double2 type; };

template <typename T> using C2 = typename cvec<T>::type;

template <typename T> __device__ __forceinline__ C2<T> cmk(T r, T i)
{
    C2<T> v;
    v.x = r;
    v.y = i;
    return v;
}
template <typename T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return cmk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return cmk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
    return cmk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ C2<T> cconj(C2<T> a) { return cmk<T>(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> __device__ __forceinline__ C2<T> cmul_mi(C2<T> a)
{
    return INV ? cmk<T>(-a.y, a.x) : cmk<T>(a.y, -a.x);
}

// LDS layout of a line: with skew SK > 0 element j lives at j + (j >> SK)
// (one pad slot per 2^SK elements), a line of length L occupies
// lds_len<SK>(L) slots.  The first radix-8 passes write at strides 8 and 64
// elements, which without the pad puts many lanes of a wave on one LDS bank.
// Measured on MI355X (256x256x16 cells): SK = 3 takes the FP32 z pass from
// 41.7 to 35.0 us and the x transforms from 20.6 to 16.2 us; the y passes,
// whose column tiles already spread over the banks and whose LDS footprint
// sets their occupancy, are faster without it.
#define MM_SKEW_X 3  // x lines: x transforms and the stage kernel
#define MM_SKEW_Y 0  // y column tiles
#define MM_SKEW_Z 3  // z column tiles
template <int SK> __host__ __device__ __forceinline__ int sk(int j) { return SK ? j + (j >> SK) : j; }
template <int SK> __host__ __device__ __forceinline__ int lds_len(int L) { return (SK ? L + (L >> SK) : L) + 1; }

// Workgroups of a launch are dispatched round-robin over gfx950's 8 XCDs,
// each with its own L2: hardware workgroup b runs on XCD b % 8.  xcd_tile(b,
// n) maps the n workgroups of a launch to logical tiles so that every XCD
// takes one contiguous range of tiles, in order; tiles that share cache lines
// (neighbouring rows, the halves of a 128-byte line, mirrored kernel rows)
// then meet in one L2 instead of being fetched once per XCD.
#define MM_NUM_XCD 8
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int x = b % MM_NUM_XCD, i = b / MM_NUM_XCD, q = n / MM_NUM_XCD, r = n % MM_NUM_XCD;
    return x * q + (x < r ? x : r) + i;
}

// linear workgroup index of a 3-D grid (x fastest), remapped by xcd_tile when
// remap is set
__device__ __forceinline__ int block_tile(int remap)
{
    const int n = (int)(gridDim.x * gridDim.y * gridDim.z);
    const int b = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
    return remap ? xcd_tile(b, n) : b;
}

// device-side plan of one transform length (passed by value).  The radix of
// pass p is (radices >> 4p) & 15, so no array is indexed at run time.
// Twiddles are stored per pass: pass p (radix R, Ns = product of the
// earlier radices, Ns > 1) owns (R-1) Ns entries at offset toff_p,
// tw[toff_p + (r-1) Ns + k] = exp(-2 pi i r k / (Ns R)) for r = 1..R-1 and
// k = 0..Ns-1 (the values of the plain table at index r k L/(Ns R)).  The
// lanes of a wave take consecutive k, so each twiddle read of a pass is one
// contiguous run in LDS (the plain table read at stride r hit 2-8 lanes per
// bank).  ntw = sum over passes with Ns > 1 of (R-1) Ns < L entries.
struct FftPlanDev {
    int L;
    int tline;             // threads per line
    int npass;
    unsigned radices;
    const void *tw;        // per-pass twiddles (ntw entries), precision T
    int ntw;
    int pow2;              // L is a power of two (all radices 8, 4, 2)
};

// The kernels copy the plan's twiddle table into LDS next to their lines
// (one read of ntw values per workgroup) before the barrier that precedes
// the first transform, so no pass waits on a global load.
template <typename T, int LC = 0> __device__ __forceinline__ void load_twiddles(C2<T> *twl, const FftPlanDev &p)
{
    const C2<T> *g = (const C2<T> *)p.tw;
    const int n = p.ntw;
    for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = g[i];
}

template <int R> struct Log2R { static const int value = R == 8 ? 3 : (R == 4 ? 2 : 1); };

// ---- size-R DFTs on registers --------------------------------------------
template <typename T, bool INV> __device__ __forceinline__ void dft2(C2<T> *v)
{
    const C2<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
}

template <typename T, bool INV> __device__ __forceinline__ void dft4(C2<T> *v)
{
    const C2<T> t0 = cadd<T>(v[0], v[2]), t1 = csub<T>(v[0], v[2]);
    const C2<T> t2 = cadd<T>(v[1], v[3]), t3 = cmul_mi<T, INV>(csub<T>(v[1], v[3]));
    v[0] = cadd<T>(t0, t2);
    v[2] = csub<T>(t0, t2);
    v[1] = cadd<T>(t1, t3);
    v[3] = csub<T>(t1, t3);
}

template <typename T, bool INV> __device__ __forceinline__ void dft8(C2<T> *v)
{
    C2<T> e[4] = {v[0], v[2], v[4], v[6]};
    C2<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<T, INV>(e);
    dft4<T, INV>(o);
    const T h = (T)0.70710678118654752440084436210484903928483593768847;
    // o[k] *= W8^k
    const C2<T> o1 = INV ? cmk<T>(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                         : cmk<T>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    const C2<T> o2 = cmul_mi<T, INV>(o[2]);
    const C2<T> o3 = INV ? cmk<T>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                         : cmk<T>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd<T>(e[0], o[0]);
    v[4] = csub<T>(e[0], o[0]);
    v[1] = cadd<T>(e[1], o1);
    v[5] = csub<T>(e[1], o1);
    v[2] = cadd<T>(e[2], o2);
    v[6] = csub<T>(e[2], o2);
    v[3] = cadd<T>(e[3], o3);
    v[7] = csub<T>(e[3], o3);
}

template <typename T, bool INV> __device__ __forceinline__ void dft3(C2<T> *v)
{
    const T c = (T)-0.5;
    const T s = INV ? (T)0.86602540378443864676372317075293618347140262690519
                    : (T)-0.86602540378443864676372317075293618347140262690519;
    const C2<T> t = cadd<T>(v[1], v[2]), d = csub<T>(v[1], v[2]);
    const C2<T> m = cmk<T>(v[0].x + c * t.x, v[0].y + c * t.y);
    // i*s*d
    const C2<T> isd = cmk<T>(-s * d.y, s * d.x);
    v[0] = cadd<T>(v[0], t);
    v[1] = cadd<T>(m, isd);
    v[2] = csub<T>(m, isd);
}

template <typename T, bool INV> __device__ __forceinline__ void dft5(C2<T> *v)
{
    const T c1 = (T)0.30901699437494742410229341718281905886436044959787;   // cos(2pi/5)
    const T c2 = (T)-0.80901699437494742410229341718281905886436044959787;  // cos(4pi/5)
    const T s1 = (T)0.95105651629515357211643933337938214340569863412575;   // sin(2pi/5)
    const T s2 = (T)0.58778525229247312916870595463907276859765243764314;   // sin(4pi/5)
    const T sg = INV ? (T)1 : (T)-1;
    const C2<T> b1 = cadd<T>(v[1], v[4]), d1 = csub<T>(v[1], v[4]);
    const C2<T> b2 = cadd<T>(v[2], v[3]), d2 = csub<T>(v[2], v[3]);
    const C2<T> a0 = v[0];
    const C2<T> m1 = cmk<T>(a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y);
    const C2<T> m2 = cmk<T>(a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y);
    const C2<T> n1 = cmk<T>(sg * (s1 * d1.x + s2 * d2.x), sg * (s1 * d1.y + s2 * d2.y)); // times i below
    const C2<T> n2 = cmk<T>(sg * (s2 * d1.x - s1 * d2.x), sg * (s2 * d1.y - s1 * d2.y));
    v[0] = cmk<T>(a0.x + b1.x + b2.x, a0.y + b1.y + b2.y);
    v[1] = cmk<T>(m1.x - n1.y, m1.y + n1.x);
    v[4] = cmk<T>(m1.x + n1.y, m1.y - n1.x);
    v[2] = cmk<T>(m2.x - n2.y, m2.y + n2.x);
    v[3] = cmk<T>(m2.x + n2.y, m2.y - n2.x);
}

template <int R, typename T, bool INV> __device__ __forceinline__ void dftR(C2<T> *v)
{
    if constexpr (R == 2) dft2<T, INV>(v);
    else if constexpr (R == 3) dft3<T, INV>(v);
    else if constexpr (R == 4) dft4<T, INV>(v);
    else if constexpr (R == 5) dft5<T, INV>(v);
    else dft8<T, INV>(v);
}

// Maximum butterflies per thread and pass for radix R when a line of length
// L is spread over tline = ceil(L/EPT) threads: (L/R)/tline <= EPT/R.
template <int R, int EPT> struct MaxB { static const int value = (EPT + R - 1) / R; };

// One Stockham pass over every line of the workgroup.  `active` is false for
// threads with no line; they still take part in both barriers.  Ns is the
// product of the radices of the previous passes; P2: every length involved
// is a power of two and lNs = log2(Ns), so index arithmetic is shifts/masks.
template <int R, int EPT, typename T, bool INV, bool P2, int SK>
__device__ __forceinline__ void fft_pass(C2<T> *line, const C2<T> *__restrict__ tw, int L, int tline, int t,
                                         bool active, int Ns, int lNs, int toff)
{
    constexpr int MB = MaxB<R, EPT>::value;
    const int nb = L / R;
    C2<T> v[MB][R];
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int b = t + q * tline;
        if (active && b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = line[sk<SK>(b + r * nb)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int b = t + q * tline;
        if (active && b < nb) {
            const int k = P2 ? (b & (Ns - 1)) : b % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C2<T> w = tw[toff + (r - 1) * Ns + k];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul<T>(v[q][r], w);
                }
            }
            dftR<R, T, INV>(v[q]);
            const int base = P2 ? (((b >> lNs) << (lNs + Log2R<R>::value)) + k) : (b / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) line[sk<SK>(base + r * Ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// Full transform of one line per thread group.  All threads of the workgroup
// must call this with the same plan (barriers inside).  EPT is the number of
// elements per thread the plan was built for; tw: the plan's twiddle table in
// LDS (load_twiddles).
template <typename T, bool INV, int EPT, bool P2, int SK>
__device__ __forceinline__ void lds_fft_ept(C2<T> *line, const FftPlanDev &p, const C2<T> *tw, int t, bool active)
{
    int Ns = 1, lNs = 0, toff = 0;
    unsigned rs = p.radices;
    for (int ps = 0; ps < p.npass; ++ps, rs >>= 4) {
        const int R = (int)(rs & 15u);
        if (P2) {
            switch (R) {
            case 8: fft_pass<8, EPT, T, INV, true, SK>(line, tw, p.L, p.tline, t, active, Ns, lNs, toff); lNs += 3; break;
            case 4: fft_pass<4, EPT, T, INV, true, SK>(line, tw, p.L, p.tline, t, active, Ns, lNs, toff); lNs += 2; break;
            default: fft_pass<2, EPT, T, INV, true, SK>(line, tw, p.L, p.tline, t, active, Ns, lNs, toff); lNs += 1; break;
            }
        } else {
            switch (R) {
            case 8: fft_pass<8, EPT, T, INV, false, SK>(line, tw, p.L, p.tline, t, active, Ns, 0, toff); break;
            case 5: fft_pass<5, EPT, T, INV, false, SK>(line, tw, p.L, p.tline, t, active, Ns, 0, toff); break;
            case 4: fft_pass<4, EPT, T, INV, false, SK>(line, tw, p.L, p.tline, t, active, Ns, 0, toff); break;
            case 3: fft_pass<3, EPT, T, INV, false, SK>(line, tw, p.L, p.tline, t, active, Ns, 0, toff); break;
            default: fft_pass<2, EPT, T, INV, false, SK>(line, tw, p.L, p.tline, t, active, Ns, 0, toff); break;
            }
        }
        if (Ns > 1) toff += (R - 1) * Ns;
        Ns *= R;
    }
}

#define MM_FFT_EPT 8

template <typename T, bool INV, int SK>
__device__ __forceinline__ void lds_fft(C2<T> *line, const FftPlanDev &p, int t, bool active, const C2<T> *tw)
{
    if (p.pow2) lds_fft_ept<T, INV, MM_FFT_EPT, true, SK>(line, p, tw, t, active);
    else lds_fft_ept<T, INV, MM_FFT_EPT, false, SK>(line, p, tw, t, active);
}

// ---- transforms with the first pass reading and/or the last pass writing
// outside LDS ---------------------------------------------------------------
// A thread takes part in a pass as butterfly owner `t` of a line.  Passes
// that touch global memory use the global mapping (G: line pointer, t,
// active) chosen by the kernel for coalescing; passes inside LDS use the
// LDS mapping (S).  Reading the inputs straight into registers (ld(j) =
// element j) saves the global -> LDS copy and its barrier; writing the last
// pass's outputs straight out (st(j, v)) saves the final LDS round trip.
struct LineMap {
    int t;
    bool active;
};

template <int R, int EPT, typename T, bool INV, bool P2, bool LDIN, bool STOUT, int SK, bool PRE, typename LD,
          typename ST, typename PF>
__device__ __forceinline__ void fft_pass_io(C2<T> *lineG, LineMap g, C2<T> *lineS, LineMap sm_, const C2<T> *tw,
                                            int L, int tline, int Ns, int lNs, int toff, LD &ld, ST &st, PF &pre)
{
    constexpr int MB = MaxB<R, EPT>::value;
    const int nb = L / R;
    C2<T> v[MB][R];
    // a pass that reads through ld or writes through st runs in the global
    // mapping (ld and st address the thread's line of that mapping), others
    // in the LDS mapping
    constexpr bool useG = LDIN || STOUT;
    const LineMap in = useG ? g : sm_;
    C2<T> *lin = useG ? lineG : lineS;
    if constexpr (LDIN) {
        // all items' loads in one block: an item past nb loads element r nb
        // (in range, not used) instead of branching, since a branch per item
        // made the compiler wait for each item's loads before the next
        if (in.active) {
#pragma unroll
            for (int q = 0; q < MB; ++q) {
                const int b = in.t + q * tline;
                const int bb = b < nb ? b : 0;
#pragma unroll
                for (int r = 0; r < R; ++r) v[q][r] = ld(bb + r * nb);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < MB; ++q) {
            const int b = in.t + q * tline;
            if (in.active && b < nb) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[q][r] = lin[sk<SK>(b + r * nb)];
            }
        }
    }
    // first pass: the caller's hook runs while this pass's reads are in
    // flight (kernels store their prefetched twiddles to LDS here)
    if constexpr (PRE) pre();
    if (!LDIN && !STOUT) __syncthreads();  // in place in LDS: every read before any write
    // outputs: a butterfly stays with the thread that read it
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int b = in.t + q * tline;
        if (in.active && b < nb) {
            const int k = P2 ? (b & (Ns - 1)) : b % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C2<T> w = tw[toff + (r - 1) * Ns + k];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul<T>(v[q][r], w);
                }
            }
            dftR<R, T, INV>(v[q]);
            const int base = P2 ? (((b >> lNs) << (lNs + Log2R<R>::value)) + k) : (b / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (STOUT) st(base + r * Ns, v[q][r]);
                else lin[sk<SK>(base + r * Ns)] = v[q][r];
            }
        }
    }
    if (!STOUT) __syncthreads();
}

template <typename T, bool INV, int EPT, bool P2, bool LDIN, bool STOUT, bool FIRST, bool LAST, int SK,
          typename LD, typename ST, typename PF>
__device__ __forceinline__ void fft_pass_sel(int R, C2<T> *lineG, LineMap g, C2<T> *lineS, LineMap s_,
                                             const C2<T> *tw, int L, int tline, int Ns, int lNs, int toff, LD &ld,
                                             ST &st, PF &pre)
{
    constexpr bool I = LDIN && FIRST, O = STOUT && LAST;
    if (P2) {
        switch (R) {
        case 8: fft_pass_io<8, EPT, T, INV, true, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, lNs, toff, ld, st, pre); break;
        case 4: fft_pass_io<4, EPT, T, INV, true, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, lNs, toff, ld, st, pre); break;
        default: fft_pass_io<2, EPT, T, INV, true, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, lNs, toff, ld, st, pre); break;
        }
    } else {
        switch (R) {
        case 8: fft_pass_io<8, EPT, T, INV, false, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, 0, toff, ld, st, pre); break;
        case 5: fft_pass_io<5, EPT, T, INV, false, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, 0, toff, ld, st, pre); break;
        case 4: fft_pass_io<4, EPT, T, INV, false, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, 0, toff, ld, st, pre); break;
        case 3: fft_pass_io<3, EPT, T, INV, false, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, 0, toff, ld, st, pre); break;
        default: fft_pass_io<2, EPT, T, INV, false, I, O, SK, FIRST>(lineG, g, lineS, s_, tw, L, tline, Ns, 0, toff, ld, st, pre); break;
        }
    }
}

// Transform of one line per thread group, all threads of the workgroup
// calling (barriers inside).  LDIN: inputs via ld(j) (else the LDS line holds
// them); STOUT: outputs via st(j, v) (else they end in the LDS line, followed
// by a barrier).  lineG/g: line and mapping of the passes that use ld/st,
// lineS/s: of the LDS passes (the same line in both mappings).
template <typename T, bool INV, bool P2, bool LDIN, bool STOUT, int SK, typename LD, typename ST, typename PF>
__device__ __forceinline__ void fft_line_p2(const FftPlanDev &p, const C2<T> *tw, C2<T> *lineG, LineMap g,
                                            C2<T> *lineS, LineMap s_, LD &ld, ST &st, PF &pre)
{
    constexpr int EPT = MM_FFT_EPT;
    const int np = p.npass;
    if (np == 0) {
        pre();
        if (g.active && g.t == 0) {
            const C2<T> v = LDIN ? ld(0) : lineG[sk<SK>(0)];
            if (STOUT) st(0, v);
            else lineG[sk<SK>(0)] = v;
        }
        if (!STOUT) __syncthreads();
        return;
    }
    int Ns = 1, lNs = 0, toff = 0;
    unsigned rs = p.radices;
    for (int ps = 0; ps < np; ++ps, rs >>= 4) {
        const int R = (int)(rs & 15u);
        // L and tline pass through an empty asm statement each pass, so the
        // compiler cannot hoist the read addresses of every radix and
        // position variant out of the pass loop (keeping them all live costs
        // 150-250 VGPRs and spills; each variant computes its own in place)
        int Lv = p.L, tlv = p.tline;
        asm volatile("" : "+s"(Lv), "+s"(tlv));
        if (ps == 0 && ps == np - 1)
            fft_pass_sel<T, INV, EPT, P2, LDIN, STOUT, true, true, SK>(R, lineG, g, lineS, s_, tw, Lv, tlv, Ns, lNs, toff, ld, st, pre);
        else if (ps == 0)
            fft_pass_sel<T, INV, EPT, P2, LDIN, STOUT, true, false, SK>(R, lineG, g, lineS, s_, tw, Lv, tlv, Ns, lNs, toff, ld, st, pre);
        else if (ps == np - 1)
            fft_pass_sel<T, INV, EPT, P2, LDIN, STOUT, false, true, SK>(R, lineG, g, lineS, s_, tw, Lv, tlv, Ns, lNs, toff, ld, st, pre);
        else
            fft_pass_sel<T, INV, EPT, P2, LDIN, STOUT, false, false, SK>(R, lineG, g, lineS, s_, tw, Lv, tlv, Ns, lNs, toff, ld, st, pre);
        if (Ns > 1) toff += (R - 1) * Ns;
        Ns *= R;
        lNs += R == 8 ? 3 : (R == 4 ? 2 : 1);
    }
}

// ---- compile-time plans for power-of-two lengths ---------------------------
// A kernel instantiated for a fixed power-of-two length LC runs the passes of
// the runtime plan (radix-8 passes, then one radix-4 or radix-2 pass, as
// FftPlanHost::build orders them) with every length, stride and pass count a
// constant: the pass loop unrolls, the radix switch disappears and the index
// arithmetic folds (shorter code, no divisions, fewer live registers).
constexpr int mm_ilog2(int v) { return v <= 1 ? 0 : 1 + mm_ilog2(v >> 1); }

template <int LC> struct FftLen {
    static_assert(LC >= 2 && (LC & (LC - 1)) == 0, "compile-time plans are for powers of two");
    static constexpr int lg = mm_ilog2(LC);
    static constexpr int n8 = lg / 3, rem = lg % 3;
    static constexpr int npass = n8 + (rem ? 1 : 0);
    static constexpr int tline = LC > MM_FFT_EPT ? LC / MM_FFT_EPT : 1;
    static constexpr int radix(int ps) { return ps < n8 ? 8 : (rem == 2 ? 4 : 2); }
    // offset of pass ps in the per-pass twiddle table
    static constexpr int toff(int ps)
    {
        int o = 0, ns = 1;
        for (int q = 0; q < ps; ++q) {
            if (ns > 1) o += (radix(q) - 1) * ns;
            ns *= radix(q);
        }
        return o;
    }
};

// line length and threads per line of a kernel instantiated for LC (0: the
// runtime plan p)
template <int LC> struct PlanC {
    __device__ static constexpr int L(const FftPlanDev &) { return LC; }
    __device__ static constexpr int tline(const FftPlanDev &) { return FftLen<LC>::tline; }
};
template <> struct PlanC<0> {
    __device__ static int L(const FftPlanDev &p) { return p.L; }
    __device__ static int tline(const FftPlanDev &p) { return p.tline; }
};

template <typename T, bool INV, int LC, int PS, int NS, bool LDIN, bool STOUT, int SK, typename LD, typename ST,
          typename PF>
__device__ __forceinline__ void fft_line_c(const C2<T> *tw, C2<T> *lineG, LineMap g, C2<T> *lineS, LineMap s_,
                                           LD &ld, ST &st, PF &pre)
{
    using F = FftLen<LC>;
    if constexpr (PS < F::npass) {
        constexpr int R = F::radix(PS);
        constexpr bool I = LDIN && PS == 0, O = STOUT && PS == F::npass - 1;
        fft_pass_io<R, MM_FFT_EPT, T, INV, true, I, O, SK, PS == 0>(lineG, g, lineS, s_, tw, LC, F::tline, NS,
                                                                    mm_ilog2(NS), F::toff(PS), ld, st, pre);
        fft_line_c<T, INV, LC, PS + 1, NS * R, LDIN, STOUT, SK>(tw, lineG, g, lineS, s_, ld, st, pre);
    }
}

// LC > 0: the compile-time plan of that length (p.L must equal LC); LC == 0:
// the runtime plan p.  SK: the LDS skew of the lines (sk).
struct FftNoPre {
    __device__ void operator()() const {}
};

// pre(): called once in the first pass after its reads are issued and
// before its barrier (FftNoPre: nothing)
template <typename T, bool INV, bool LDIN, bool STOUT, int LC, int SK, typename LD, typename ST,
          typename PF = FftNoPre>
__device__ __forceinline__ void fft_line(const FftPlanDev &p, const C2<T> *tw, C2<T> *lineG, LineMap g, C2<T> *lineS,
                                         LineMap s_, LD ld, ST st, PF pre = PF())
{
    if constexpr (LC > 0) {
        fft_line_c<T, INV, LC, 0, 1, LDIN, STOUT, SK>(tw, lineG, g, lineS, s_, ld, st, pre);
    } else {
        if (p.pow2) fft_line_p2<T, INV, true, LDIN, STOUT, SK>(p, tw, lineG, g, lineS, s_, ld, st, pre);
        else fft_line_p2<T, INV, false, LDIN, STOUT, SK>(p, tw, lineG, g, lineS, s_, ld, st, pre);
    }
}

// Twiddles fetched into registers at kernel start and stored to LDS by the
// first FFT pass's pre() hook, so the first pass's data reads do not wait for
// the twiddle read: TwFetch<T>::fetch(p) / .store(twl).
template <typename T> struct TwFetch {
    static constexpr int N = 2;  // per thread in registers; more go through a loop
    C2<T> v[N];
    const C2<T> *g;
    int n;
    __device__ static TwFetch fetch(const FftPlanDev &p)
    {
        TwFetch f;
        f.g = (const C2<T> *)p.tw;
        f.n = p.ntw;
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const int i = threadIdx.x + q * blockDim.x;
            if (i < f.n) f.v[q] = f.g[i];
        }
        return f;
    }
    __device__ void store(C2<T> *twl) const
    {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const int i = threadIdx.x + q * blockDim.x;
            if (i < n) twl[i] = v[q];
        }
        for (int i = threadIdx.x + N * blockDim.x; i < n; i += blockDim.x) twl[i] = g[i];
    }
};

#endif
