// mm_stepper.hip -- fused Runge-Kutta stepping of the micromagnetic system.
//
// Replaces the Python-driven loop of the reference (RungeKutta.evolve/apply,
// src/magnum/evolver/runge_kutta.py:47-137; LandauLifshitzGilbert
// .calculate_dMdt, landau_lifshitz_gilbert.py:92-125; MicroMagnetics
// StepSizeController, micro_magnetics_stepsize_controller.py:35-73;
// finish_step normalisation, micro_magnetics.py:49-50) by a fixed sequence of
// kernels per attempt:
//
//   for every stage s:   x fwd (stage input) -> y fwd -> z*N*z^-1 -> y inv
//                        -> x inv + EPILOGUE
//   EPILOGUE per cell:   H = H_demag + H_exch + H_aniso + H_ext,
//                        k_s = LLG(M, H) + Zhang-Li(M),
//                        next stage input  ytmp = y + h sum_j b_{s+1,j} k_j,
//                        or on the last stage y_new = normalise(y + h sum c_j k_j),
//                        err = h sum ec_j k_j reduced to max |err|/Ms (FP64)
//   controller (1 thread): accept/reject, next h, swap of the y buffers
//
// y is double-buffered and the buffer index, t and h live in device memory,
// so attempts are enqueued without host round trips and rejected attempts
// simply leave the index unchanged.  Stage inputs alternate between two
// buffers because the epilogue reads neighbours of the current input while
// it writes the next one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <cstddef>
#include <vector>

#include "mm_cell.h"
#include "mm_conv.h"
#include "mm_reduce.h"

// device control block
struct StepCtl {
    double t, h, h_last;
    long long attempts, rejections, accepted, target;
    int cur;                       // index of the current y buffer
    int stop;                      // accepted == target: remaining launches exit
    // relaxation stop rule, checked on the first attempt from every
    // check_every-th accepted state (condition.py:169-200 checks state.step)
    int relax_mode;                // MM_RELAX_NONE / _DMDT / _DEG_PER_NS
    int relax_every;
    int relaxed;                   // the current state satisfied the rule
    int have_last;                 // relax_last holds a previous deg/ns value
    double relax_thr, relax_value, relax_last;
    long long last_check;          // accepted count of the last check
    // end time (RungeKutta.evolve's t_max, runge_kutta.py:74-78): an accepted
    // step past t_max is redone from the same state with h = t_max - t and
    // taken without error check (forced); the controller's h resumes after it
    int tmax_on, forced;
    double t_max, h_saved;
    // stop time without clipping (the solver loop's stop conditions on t,
    // solver.py:110-112 with condition.py:145-157, which name no time of
    // interest): the run stops after the first accepted step that ends at
    // t >= t_stop, and that step keeps its full length
    int tstop_on;
    double t_stop;
};

// a per-cell parameter that is either uniform (p == null) or an array
template <typename T> struct Param {
    const T *p;
    T u;
    __device__ __forceinline__ T at(size_t i) const { return p ? gfield(p)[i] : u; }
};
template <typename T> struct VParam {
    const T *p; // 3n SoA or null
    T u[3];
    __device__ __forceinline__ V3<T> at(size_t n, size_t i) const
    {
        return p ? v3<T>(gfield(p)[i], gfield(p)[n + i], gfield(p)[2 * n + i]) : v3<T>(u[0], u[1], u[2]);
    }
};

enum { FORM_SUM = 0, FORM_SEQ = 1 };

template <typename T> struct SysDev {
    MeshDev m;
    size_t n;
    T dx, dy, dz;
    T wx, wy, wz, coef_ex, mu0, cm2mu0, mu_b, e_charge;
    Param<T> Ms, A, f1, f2, alpha, kuni, kcub, P, xi;
    VParam<T> axis1, axis2, j;
    int has_exch, has_uni, has_cub, has_zee, has_stt, do_precess;
    T hext[3];
    // alternating Zeeman field: hoff + ac_amp sin(t ac_freq + ac_phase)
    int has_ac;
    double hoff[3], ac_amp[3], ac_freq[3], ac_phase[3];
    int Ms_uniform;
    double Ms_u;
    // exchange with uniform, non-zero Ms and uniform A: no masking, no
    // per-neighbour Ms loads; H = exch_f * sum_j w_j (M_j - M_i) with
    // exch_f = 2 A / (mu0 Ms^2), the reference's (2A/(mu0 Ms)) sum (m_j - m_i) w
    int exch_uniform;
    T exch_f;
    // rows of the stage launch in XCD-contiguous ranges (xcd_tile): the
    // exchange stencil's neighbour rows are then read from the same L2
    int xcd_remap;
    // tile_zcol > 0: tiles of rpw rows align with the z planes (tile_zcol =
    // ny / rpw tiles per plane) and consecutive tiles walk a column of
    // planes (same y range, z = 0 .. nz-1) before moving on in y.  The
    // z +- 1 neighbour rows of the exchange stencil are then read by tiles
    // that run next to each other, while still in L2; in plane order they
    // come back only after a whole plane of tiles, which evicts them.
    int tile_zcol;
    // Zeeman field from a host function (has_fn): fn_tab[3 slot + c] is the
    // field of stage record `slot`, evaluated on the host at that
    // evaluation's time before the launch
    int has_fn;
    const double *fn_tab;
    // uniaxial anisotropy with uniform k, Ms (non-zero) and axis:
    // H = uni_c (M.a) a with uni_c = 2 k / (mu0 Ms^2), the reference's
    // 2 k (m.a) / (mu0 Ms) a with m = M / Ms; energy density uni_e |a x M|^2,
    // uni_e = k / Ms^2 (no per-cell divisions)
    int uni_fast;
    T uni_c, uni_e;
    // Zhang-Li torque with uniform P, xi, Ms (non-zero), alpha and j: the
    // factors f1, f2 of fdm_zhangli_cpu.cpp:95-114 and the inverse cell sizes
    // computed once (no per-cell divisions)
    int stt_fast;
    T stt_f1, stt_f2, inv_d[3];
};

// Zhang-Li torque of cell (x,y,z) with the precomputed uniform factors
// (zhangli_cell with its divisions hoisted): gradients central inside,
// one-sided at the faces, zero along axes of one cell
template <typename T>
__device__ __forceinline__ V3<T> zhangli_cell_uniform(const MeshDev &m, int x, int y, int z, size_t i, size_t n,
                                                      const T *inv_d, T f1, T f2, const V3<T> &j,
                                                      const T *__restrict__ M)
{
    const long li = (long)i;
    auto grad = [&](int pos, int dim, long st, T id) -> V3<T> {
        if (dim <= 1) return v3<T>(0, 0, 0);
        long a = li - st, b = li + st;
        T f = id * (T)0.5;
        if (pos == 0) {
            a = li;
            f = id;
        } else if (pos == dim - 1) {
            b = li;
            f = id;
        }
        return v3<T>((M[b] - M[a]) * f, (M[n + b] - M[n + a]) * f, (M[2 * n + b] - M[2 * n + a]) * f);
    };
    const V3<T> gx = grad(x, m.nx, 1, inv_d[0]);
    const V3<T> gy = grad(y, m.ny, m.nx, inv_d[1]);
    const V3<T> gz = grad(z, m.nz, (long)m.nx * m.ny, inv_d[2]);
    const V3<T> jg = v3<T>(j.x * gx.x + j.y * gy.x + j.z * gz.x, j.x * gx.y + j.y * gy.y + j.z * gz.y,
                           j.x * gx.z + j.y * gy.z + j.z * gz.z);
    const V3<T> Mv = v3<T>(M[li], M[n + li], M[2 * n + li]);
    const V3<T> B = cross3(Mv, jg);
    const V3<T> Av = cross3(Mv, B);
    return v3<T>(f1 * Av.x + f2 * B.x, f1 * Av.y + f2 * B.y, f1 * Av.z + f2 * B.z);
}

// exchange field of cell i for uniform Ms and A (see exch_uniform).  Each
// direction's neighbour is the periodic image or, at an open boundary, the
// cell itself (M_i - M_i adds an exact zero), so the six neighbours load
// without branches, in one batch, and sum in the fixed order -x, +x, -y, +y,
// -z, +z.
template <typename T>
__device__ __forceinline__ V3<T> exchange_cell_uniform(const MeshDev &m, int x, int y, int z, size_t i, size_t n,
                                                       T wx, T wy, T wz, T f, const T *__restrict__ M)
{
    const long li = (long)i, sy = m.nx, sz = (long)m.nx * m.ny;
    long j[6];
    j[0] = x > 0 ? li - 1 : (m.px ? li + (m.nx - 1) : li);
    j[1] = x < m.nx - 1 ? li + 1 : (m.px ? li - (m.nx - 1) : li);
    j[2] = y > 0 ? li - sy : (m.py ? li + sy * (m.ny - 1) : li);
    j[3] = y < m.ny - 1 ? li + sy : (m.py ? li - sy * (m.ny - 1) : li);
    j[4] = z > 0 ? li - sz : (m.pz ? li + sz * (m.nz - 1) : li);
    j[5] = z < m.nz - 1 ? li + sz : (m.pz ? li - sz * (m.nz - 1) : li);
    T a[6][3];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        a[q][0] = M[j[q]];
        a[q][1] = M[n + j[q]];
        a[q][2] = M[2 * n + j[q]];
    }
    const V3<T> mi = v3<T>(M[i], M[n + i], M[2 * n + i]);
    const T w[6] = {wx, wx, wy, wy, wz, wz};
    V3<T> s = v3<T>(0, 0, 0);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        s.x += (a[q][0] - mi.x) * w[q];
        s.y += (a[q][1] - mi.y) * w[q];
        s.z += (a[q][2] - mi.z) * w[q];
    }
    return v3<T>(f * s.x, f * s.y, f * s.z);
}

// 32-bit byte-offset access for arrays under 4 GiB: base is a pointer the
// whole workgroup shares (held in SGPRs), off a per-lane byte offset, so the
// access is one global load/store with an SGPR base and a 32-bit VGPR offset
// (no 64-bit address arithmetic per element).
template <typename T> __device__ __forceinline__ T ld32(const T *__restrict__ base, unsigned off)
{
    return *(const T *)((const char *)base + off);
}
template <typename T> __device__ __forceinline__ void st32(T *__restrict__ base, unsigned off, T v)
{
    *(T *)((char *)base + off) = v;
}

// exchange_cell_uniform with byte offsets: ib = i sizeof(T), nb = n sizeof(T)
// (3 nb < 2^32); same neighbours, loads and summation order
template <typename T>
__device__ __forceinline__ V3<T> exchange_cell_uniform32(const MeshDev &m, int x, int y, int z, unsigned ib,
                                                         unsigned nb, T wx, T wy, T wz, T f, const V3<T> &mi,
                                                         const T *__restrict__ M)
{
    constexpr unsigned sx = sizeof(T);
    const unsigned sy = (unsigned)m.nx * sx, sz = (unsigned)m.nx * (unsigned)m.ny * sx;
    unsigned j[6];
    j[0] = x > 0 ? ib - sx : (m.px ? ib + sx * (unsigned)(m.nx - 1) : ib);
    j[1] = x < m.nx - 1 ? ib + sx : (m.px ? ib - sx * (unsigned)(m.nx - 1) : ib);
    j[2] = y > 0 ? ib - sy : (m.py ? ib + sy * (unsigned)(m.ny - 1) : ib);
    j[3] = y < m.ny - 1 ? ib + sy : (m.py ? ib - sy * (unsigned)(m.ny - 1) : ib);
    j[4] = z > 0 ? ib - sz : (m.pz ? ib + sz * (unsigned)(m.nz - 1) : ib);
    j[5] = z < m.nz - 1 ? ib + sz : (m.pz ? ib - sz * (unsigned)(m.nz - 1) : ib);
    T a[6][3];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        a[q][0] = ld32(M, j[q]);
        a[q][1] = ld32(M, j[q] + nb);
        a[q][2] = ld32(M, j[q] + 2 * nb);
    }
    const T w[6] = {wx, wx, wy, wy, wz, wz};
    V3<T> s = v3<T>(0, 0, 0);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        s.x += (a[q][0] - mi.x) * w[q];
        s.y += (a[q][1] - mi.y) * w[q];
        s.z += (a[q][2] - mi.z) * w[q];
    }
    return v3<T>(f * s.x, f * s.y, f * s.z);
}

// zhangli_cell_uniform with byte offsets (as exchange_cell_uniform32)
template <typename T>
__device__ __forceinline__ V3<T> zhangli_cell_uniform32(const MeshDev &m, int x, int y, int z, unsigned ib,
                                                        unsigned nb, const T *inv_d, T f1, T f2, const V3<T> &j,
                                                        const V3<T> &Mv, const T *__restrict__ M)
{
    // one-sided differences at the ends, central inside; a dimension of one
    // cell has a = b (difference 0).  All 18 loads are unconditional and go
    // out together (with a branch per axis the compiler waited for each
    // axis's loads before issuing the next axis's).
    T gv[3][2][3], f[3];
    auto fetch = [&](int d, int pos, int dim, unsigned st, T id) {
        const bool lo = pos == 0, hi = pos == dim - 1;
        const unsigned a = lo ? ib : ib - st, b = hi ? ib : ib + st;
        f[d] = (lo || hi) ? id : id * (T)0.5;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gv[d][0][c] = ld32(M, a + (unsigned)c * nb);
            gv[d][1][c] = ld32(M, b + (unsigned)c * nb);
        }
    };
    constexpr unsigned sx = sizeof(T);
    fetch(0, x, m.nx, sx, inv_d[0]);
    fetch(1, y, m.ny, (unsigned)m.nx * sx, inv_d[1]);
    fetch(2, z, m.nz, (unsigned)m.nx * (unsigned)m.ny * sx, inv_d[2]);
    auto grad = [&](int d) -> V3<T> {
        return v3<T>((gv[d][1][0] - gv[d][0][0]) * f[d], (gv[d][1][1] - gv[d][0][1]) * f[d],
                     (gv[d][1][2] - gv[d][0][2]) * f[d]);
    };
    const V3<T> gx = grad(0), gy = grad(1), gz = grad(2);
    const V3<T> jg = v3<T>(j.x * gx.x + j.y * gy.x + j.z * gz.x, j.x * gx.y + j.y * gy.y + j.z * gz.y,
                           j.x * gx.z + j.y * gy.z + j.z * gz.z);
    const V3<T> B = cross3(Mv, jg);
    const V3<T> Av = cross3(Mv, B);
    return v3<T>(f1 * Av.x + f2 * B.x, f1 * Av.y + f2 * B.y, f1 * Av.z + f2 * B.z);
}

// stored stage vectors one stage combination reads at most (DP54's last
// stage: k_0 .. k_5)
#define MM_STAGE_MAXK 6
// rows of the host-evaluated Zeeman table (stage records: at most 7 stages,
// dM/dt and energy)
#define MM_FN_SLOTS 16

template <typename T> struct StageDev {
    const T *Ys;          // stage input (null: y[cur])
    T *yb0, *yb1;         // the two y buffers
    T *kout;              // k_s (null: not stored)
    T *Hout;              // H_tot (null: not stored)
    // next stage input: out = base + (h*) sum coef_j k_j  (k_s last, from registers)
    int prep;             // 0 none, 1 prepare next stage, 2 final combine
    int form;             // FORM_SUM / FORM_SEQ
    T *prep_out;
    int nk;               // number of stored k_j used (j < s)
    const T *kin[MM_STAGE_MAXK];
    double b[MM_STAGE_MAXK];  // coefficients for kin
    double bself;             // coefficient for k_s
    double ec[MM_STAGE_MAXK], ecself;  // error coefficients (final stage)
    int want_dydt;        // reduce max|k_s| (stage 0)
    int norm_mode;
    // FSAL (dp54, rk23): the last stage takes max|k_0| from kin[0] and then
    // overwrites k_0 with its own k_s for the next attempt (runge_kutta.py:97-101,131-134)
    int dydt_from_k0;
    T *fsal_out;
    // prep == 1 with the demag field: the epilogue also transforms the next
    // stage input along x (rows into S1), so that stage needs no x pass
    int fwd;
    C2<T> *fwd_S1;
    // per-block maxima, rows [0] |err|^2 (final stage), [1] |k0|^2 scaled,
    // [2] |k0|^2, [3] |y|^2, one entry per block of the launch; the
    // controller reduces them (no same-address atomics across the grid)
    double *mpart;
    // node a_s of this stage: the stage is evaluated at t + a_s h (the
    // alternating Zeeman field depends on it)
    double tnode;
    // index of this record (stages 0..ns-1, then dM/dt, then energy): its
    // row of the host-evaluated Zeeman table
    int slot;
};


template <typename T> __device__ __forceinline__ double dotd(const V3<T> &a, const V3<T> &b)
{
    return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z;
}

// per-thread maxima of one stage launch (squared norms, FP64)
struct StageRed {
    double err;   // |err|^2 (/Ms^2 when Ms is not uniform)
    double dydt;  // |k0|^2 (/Ms^2 when Ms is not uniform)
    double draw;  // |k0|^2
    double mraw;  // |y|^2 of the attempt's start state
};

// per-thread sums of M.H per term (energy launches only), compensated
struct EnergyAcc {
    Dd demag, exch, aniso, zee;
};

// Per-cell epilogue: everything after the demag field of one cell.  STT:
// the Zhang-Li term is compiled in; GEN: cubic anisotropy and the exchange
// with per-cell Ms/A are compiled in (without them the epilogue is shorter
// and needs fewer registers).  Without GEN the state arrays are addressed by
// 32-bit byte offsets from their (workgroup-uniform) bases, which the host
// guarantees by taking GEN for state arrays of 4 GiB or more.
template <typename T, bool EN, bool STT, bool GEN, bool P1S = false>
__device__ __forceinline__ void stage_cell(const SysDev<T> &S, const StageDev<T> &G, const StepCtl *ctl,
                                           const T *__restrict__ Y, int x, int y, int z, const V3<T> &Hd,
                                           const V3<T> &hz, StageRed &red, EnergyAcc &ea, T *nxt)
{
    constexpr bool O32 = !GEN;
    const size_t n = S.n;
    const size_t i = (size_t)x + (size_t)S.m.nx * ((size_t)y + (size_t)S.m.ny * z);
    // byte offsets of cell i and of one component plane (O32 only)
    const unsigned ib = O32 ? ((unsigned)x + (unsigned)S.m.nx * ((unsigned)y + (unsigned)S.m.ny * (unsigned)z)) *
                                  (unsigned)sizeof(T)
                            : 0u;
    const unsigned nb = O32 ? (unsigned)n * (unsigned)sizeof(T) : 0u;
    auto ldv = [&](const T *__restrict__ a) -> V3<T> {
        if constexpr (O32) return v3<T>(ld32(a, ib), ld32(a, ib + nb), ld32(a, ib + 2 * nb));
        else return v3<T>(a[i], a[n + i], a[2 * n + i]);
    };
    auto stv = [&](T *__restrict__ a, const V3<T> &v) {
        if constexpr (O32) {
            st32(a, ib, v.x);
            st32(a, ib + nb, v.y);
            st32(a, ib + 2 * nb, v.z);
        } else {
            store3<T>(a, n, i, v);
        }
    };
    // the record's flags and uniform parameters of this cell's path, read in
    // one block (merged scalar loads, one wait) instead of inside each
    // guarded section
    const int Ms_uniform = S.Ms_uniform, has_exch = S.has_exch, has_uni = S.has_uni, uni_fast = S.uni_fast;
    const int has_zee = S.has_zee;
    const T Ms_u = S.Ms.u, wx = S.wx, wy = S.wy, wz = S.wz, exch_f = S.exch_f, uni_c = S.uni_c;
    const T ax0 = S.axis1.u[0], ax1 = S.axis1.u[1], ax2 = S.axis1.u[2];
    const T *const f1p = S.f1.p, *const f2p = S.f2.p;
    const T f1u = S.f1.u, f2u = S.f2.u;
    const T ms = Ms_uniform ? Ms_u : (O32 ? ld32(gfield(S.Ms.p), ib) : S.Ms.at(i));
    const V3<T> Mv = ldv(Y);
    V3<T> H = Hd;
    if (EN) dd_add(ea.demag, dotd<T>(Mv, Hd));
    if (has_exch && ms != (T)0) {
        V3<T> he;
        if constexpr (O32) {
            he = exchange_cell_uniform32<T>(S.m, x, y, z, ib, nb, wx, wy, wz, exch_f, Mv, Y);
        } else {
            he = S.exch_uniform
                     ? exchange_cell_uniform<T>(S.m, x, y, z, i, n, S.wx, S.wy, S.wz, S.exch_f, Y)
                     : exchange_cell<T>(S.m, x, y, z, i, n, S.wx, S.wy, S.wz, S.coef_ex,
                                        S.Ms.p ? gfield(S.Ms.p) : nullptr, S.A.p ? gfield(S.A.p) : nullptr, Y, ms);
        }
        H.x += he.x;
        H.y += he.y;
        H.z += he.z;
        if (EN) dd_add(ea.exch, dotd<T>(Mv, he));
    }
    if (has_uni || (GEN && S.has_cub)) {
        T e0 = (T)0, e1 = (T)0;
        V3<T> ha = v3<T>(0, 0, 0);
        if (GEN && S.has_cub) ha = cubic_cell<T>(Mv, S.axis1.at(n, i), S.axis2.at(n, i), S.kcub.at(i), ms, S.cm2mu0, &e1);
        if (has_uni) {
            V3<T> hu;
            if (uni_fast) {
                const V3<T> ax = v3<T>(ax0, ax1, ax2);
                const T f = uni_c * dot3(Mv, ax);
                hu = v3<T>(f * ax.x, f * ax.y, f * ax.z);
                if (EN) {
                    const V3<T> c = cross3(ax, Mv);
                    e0 = S.uni_e * dot3(c, c);
                }
            } else {
                hu = uniaxial_cell<T>(Mv, S.axis1.at(n, i), S.kuni.at(i), ms, S.mu0, &e0);
            }
            if (EN) dd_add(ea.aniso, (double)e0);
            if (GEN && S.has_cub) {
                ha.x += hu.x;
                ha.y += hu.y;
                ha.z += hu.z;
            } else {
                ha = hu;
            }
        }
        if (EN && GEN && S.has_cub) dd_add(ea.aniso, (double)e1);
        H.x += ha.x;
        H.y += ha.y;
        H.z += ha.z;
    }
    if (has_zee) {
        H.x += hz.x;
        H.y += hz.y;
        H.z += hz.z;
        if (EN) dd_add(ea.zee, (double)Mv.x * hz.x + (double)Mv.y * hz.y + (double)Mv.z * hz.z);
    }
    if (EN) return;
    V3<T> k = llge_cell<T>(Mv, H, f1p ? gfield(f1p)[i] : f1u, f2p ? gfield(f2p)[i] : f2u);
    if (STT && S.has_stt) {
        V3<T> st;
        if (O32 && S.stt_fast)
            st = zhangli_cell_uniform32<T>(S.m, x, y, z, ib, nb, S.inv_d, S.stt_f1, S.stt_f2,
                                           v3<T>(S.j.u[0], S.j.u[1], S.j.u[2]), Mv, Y);
        else if (S.stt_fast)
            st = zhangli_cell_uniform<T>(S.m, x, y, z, i, n, S.inv_d, S.stt_f1, S.stt_f2,
                                         v3<T>(S.j.u[0], S.j.u[1], S.j.u[2]), Y);
        else
            st = zhangli_cell<T>(S.m, x, y, z, i, n, S.dx, S.dy, S.dz, S.do_precess != 0, S.P.at(i), S.xi.at(i), ms,
                                 S.alpha.at(i), S.j.at(n, i), Y, S.mu_b, S.e_charge);
        k.x += st.x;
        k.y += st.y;
        k.z += st.z;
    }
    if (G.kout) stv(gfield(G.kout), k);
    if (G.Hout) stv(gfield(G.Hout), H);
    if (G.want_dydt) {
        const double a = k.x, b = k.y, c = k.z;
        double q = a * a + b * b + c * c;
        red.draw = fmax(red.draw, q);
        red.mraw = fmax(red.mraw, (double)Mv.x * Mv.x + (double)Mv.y * Mv.y + (double)Mv.z * Mv.z);
        if (!S.Ms_uniform) {
            const double s = ms;
            q = s == 0.0 ? 0.0 : q / (s * s);
        }
        red.dydt = fmax(red.dydt, q);
    }
    if (!P1S && !G.prep) return;
    // the combination's inputs (stored stage vectors, the attempt's start
    // state) load in one batch of independent loads before any arithmetic
    // (loading them here rather than at the top keeps them out of the field
    // computation's live registers)
    const T *yc = ctl->cur ? gfield(G.yb1) : gfield(G.yb0);
    T kv[MM_STAGE_MAXK][3], ycv[3];
    {
        const V3<T> v = ldv(yc);
        ycv[0] = v.x;
        ycv[1] = v.y;
        ycv[2] = v.z;
    }
#pragma unroll
    for (int q = 0; q < MM_STAGE_MAXK; ++q)
        if (q < G.nk) {
            const V3<T> v = ldv(gfield(G.kin[q]));
            kv[q][0] = v.x;
            kv[q][1] = v.y;
            kv[q][2] = v.z;
        }
    const double h = ctl->h;
    // stage vectors outermost, components inside: each tableau coefficient is
    // read once per cell (with the components outermost, the guarded loops
    // loaded every coefficient again per component); per component the sums
    // keep the order q = 0, 1, ..., self
    const int nk = G.nk, form = P1S ? FORM_SUM : G.form, prep = P1S ? 1 : G.prep;
    const T kc[3] = {k.x, k.y, k.z};
    T out[3];
    double er[3] = {0.0, 0.0, 0.0};
    if (form == FORM_SUM) {
        T sum[3];
        double e[3];
#pragma unroll
        for (int q = 0; q < MM_STAGE_MAXK; ++q) {
            if (q < nk) {
                const T bq = (T)G.b[q];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T term = bq * kv[q][c];
                    sum[c] = q == 0 ? term : sum[c] + term;
                }
                if (prep == 2) {
                    const double ecq = G.ec[q];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double term = ecq * (double)kv[q][c];
                        e[c] = q == 0 ? term : e[c] + term;
                    }
                }
            }
        }
        const T bself = (T)G.bself;
        const double ecself = G.ecself;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T ts = bself * kc[c];
            const T sc = nk == 0 ? ts : sum[c] + ts;
            out[c] = ycv[c] + (T)h * sc;
            if (prep == 2) {
                const double tse = ecself * (double)kc[c];
                er[c] = h * (nk == 0 ? tse : e[c] + tse);
            }
        }
    } else {
        T v[3] = {ycv[0], ycv[1], ycv[2]};
#pragma unroll
        for (int q = 0; q < MM_STAGE_MAXK; ++q) {
            if (q < nk) {
                const T f = (T)(h * G.b[q]);
                const double fe = prep == 2 ? h * G.ec[q] : 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T kq = kv[q][c];
                    v[c] += kq * f;
                    if (prep == 2) er[c] += (double)kq * fe;
                }
            }
        }
        const T fs = (T)(h * G.bself);
        const double fes = h * G.ecself;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] += kc[c] * fs;
            if (prep == 2) er[c] += (double)kc[c] * fes;
            out[c] = v[c];
        }
    }
    if (prep == 1) {
        stv(gfield(G.prep_out), v3<T>(out[0], out[1], out[2]));
        nxt[0] = out[0];
        nxt[1] = out[1];
        nxt[2] = out[2];
        return;
    }
    // final stage: error norm and normalised new state into the other buffer
    if (G.dydt_from_k0) {
        const V3<T> k0 = ldv(gfield(G.kin[0]));
        const double a = k0.x, b = k0.y, c = k0.z;
        double q = a * a + b * b + c * c;
        red.draw = fmax(red.draw, q);
        const double y0 = ycv[0], y1 = ycv[1], y2 = ycv[2];
        red.mraw = fmax(red.mraw, y0 * y0 + y1 * y1 + y2 * y2);
        if (!S.Ms_uniform) {
            const double s = ms;
            q = s == 0.0 ? 0.0 : q / (s * s);
        }
        red.dydt = fmax(red.dydt, q);
    }
    if (G.fsal_out) stv(gfield(G.fsal_out), k);
    {
        double q = er[0] * er[0] + er[1] * er[1] + er[2] * er[2];
        if (!S.Ms_uniform) {
            const double s = ms;
            q = s == 0.0 ? 0.0 : q / (s * s);
        }
        red.err = fmax(red.err, q);
    }
    T *yn = ctl->cur ? gfield(G.yb0) : gfield(G.yb1);
    double norm;
    if (G.norm_mode == 0) {
        norm = float_cast_norm((double)out[0], (double)out[1], (double)out[2]);
    } else {
        norm = fp64_norm((double)out[0], (double)out[1], (double)out[2]);
    }
    if (norm != 0.0) {
        const double sc = (double)ms / norm;
        out[0] = (T)(out[0] * sc);
        out[1] = (T)(out[1] * sc);
        out[2] = (T)(out[2] * sc);
    }
    stv(yn, v3<T>(out[0], out[1], out[2]));
}

// x inverse of the demag spectrum rows + epilogue.  rpw rows per block, all
// three components of a row in the same block.  has_demag: 0 no stray field,
// 1 inverse x transform of S1 rows in LDS, 2 H_demag read from Hdem (lines
// too long for three in one workgroup's LDS).  EN: instead of the stage
// epilogue, per-block compensated sums of M.H per term: the pair of term q
// at epart[2q nblk + blk] (s) and epart[(2q+1) nblk + blk] (e).  blk: this
// workgroup's index among the nblk of the system's launch; remap: order the
// tiles by xcd_tile(blk, nblk) (the single-system launch; a batched launch
// orders its global workgroups before it picks the system).
template <typename T, bool EN, int LC, bool STT, bool GEN, bool P1S = false>
__device__ __forceinline__ void stage_body(const SysDev<T> *__restrict__ Sp, const StageDev<T> *__restrict__ Gp,
                                           StepCtl *ctl, const C2<T> *__restrict__ S1, const FftPlanDev &p, int Hx,
                                           int Hp, int rpw, int has_demag, double *__restrict__ epart,
                                           const T *__restrict__ Hdem, int blk, int nblk, bool remap)
{
    // the system and stage records live in device memory (read where used)
    // rather than in the kernel arguments: by value they need ~160 SGPRs,
    // and the compiler spilled most of them to VGPR lanes
    const SysDev<T> &S = *Sp;
    const StageDev<T> &G = *Gp;
    constexpr int SK = MM_SKEW_X;
    // stages of an attempt (prep != 0) do nothing once the run has stopped;
    // evaluations outside attempts (dM/dt, energies) always run
    if (!EN && (P1S || G.prep) && ctl->stop) return;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const int nx = S.m.nx, ny = S.m.ny, nz = S.m.nz;
    const int nrows = ny * nz;
    int tile = (remap && S.xcd_remap) ? xcd_tile(blk, nblk) : blk;
    if (S.tile_zcol > 0) {
        const int ty = tile / nz;
        tile = (tile - ty * nz) * S.tile_zcol + ty;
    }
    const int row0 = tile * rpw;
    const T *Y = G.Ys ? gfield(G.Ys) : (ctl->cur ? gfield(G.yb1) : gfield(G.yb0));
    const int L = PlanC<LC>::L(p), LP = lds_len<SK>(L), TL = PlanC<LC>::tline(p);
    // Rows are transformed in pairs, two real rows per complex line: line
    // l = (component c, pair q) holds rows gr0 = row0 + 2q (real part) and
    // gr0 + 1 (imaginary part).  With has_demag == 1, rpw is even.
    const int np = rpw >> 1;
    C2<T> *twl = sm + (size_t)3 * np * LP;
    if (has_demag == 1) {
        const TwFetch<T> twf = TwFetch<T>::fetch(p);
        // inverse: Z = X0 + i X1 with both half spectra extended Hermitian
        // (X[L-k] = conj X[k]; X[0], X[L/2] taken real), so that Re and Im of
        // the inverse transform are the two rows
        const int l = threadIdx.x / TL, t = threadIdx.x % TL;
        const int c = np == 1 ? l : l / (np > 0 ? np : 1), q = np == 1 ? 0 : l % (np > 0 ? np : 1);
        const int gr0 = row0 + 2 * q;
        const bool active = (l < 3 * np) && gr0 < nrows;
        const bool has1 = gr0 + 1 < nrows;
        C2<T> *line = sm + (size_t)l * LP;
        const C2<T> *src0 = S1 + ((size_t)c * nrows + (active ? gr0 : 0)) * Hp;
        const C2<T> *src1 = src0 + (has1 ? Hp : 0);
        const LineMap m = {t, active};
        fft_line<T, true, true, false, LC, SK>(
            p, twl, line, m, line, m,
            [&](int k) {
                const bool lo = k < Hx;
                const int kk = lo ? k : L - k;
                // both loads unconditional (src1 is src0 for an absent second
                // row) and the cases folded into a factor f (0 at k = 0 and
                // L/2, +1 below Hx, -1 above; f*v is exact): with branches
                // around the loads, the compiler waited for every element's
                // loads before issuing the next element's
                const C2<T> a = src0[kk];
                const C2<T> b = src1[kk];
                const T bx = has1 ? b.x : (T)0, by = has1 ? b.y : (T)0;
                const T f = ((k == 0) | (2 * k == L)) ? (T)0 : (lo ? (T)1 : (T)-1);
                return cmk<T>(a.x - f * by, bx + f * a.y);
            },
            [](int, C2<T>) {}, [&] { twf.store(twl); });
    }
    StageRed red = {0.0, 0.0, 0.0, 0.0};
    EnergyAcc ea = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    const bool fused = !EN && has_demag == 1 && G.fwd;
    // Zeeman field of this evaluation (alternating_field.py:57-84), at the
    // stage time t + a_s h
    V3<T> hz = v3<T>(S.hext[0], S.hext[1], S.hext[2]);
    if (S.has_zee && S.has_ac) {
        const double ts = ctl->t + G.tnode * ctl->h;
        hz = v3<T>((T)(S.hoff[0] + S.ac_amp[0] * sin(ts * S.ac_freq[0] + S.ac_phase[0])),
                   (T)(S.hoff[1] + S.ac_amp[1] * sin(ts * S.ac_freq[1] + S.ac_phase[1])),
                   (T)(S.hoff[2] + S.ac_amp[2] * sin(ts * S.ac_freq[2] + S.ac_phase[2])));
    } else if (S.has_zee && S.has_fn) {
        const double *f = gfield(S.fn_tab) + 3 * G.slot;
        hz = v3<T>((T)f[0], (T)f[1], (T)f[2]);
    }
    // rows of the workgroup in (y, z): from row0 once, then stepped (no
    // per-cell integer division)
    int yr = row0 % ny, zr = row0 / ny;
    for (int r = 0; r < rpw; ++r, ++yr) {
        if (yr == ny) {
            yr = 0;
            ++zr;
        }
        const int gr = row0 + r;
        const int q = r >> 1, hi = r & 1;
        for (int x = threadIdx.x; x < nx; x += blockDim.x) {
            if (gr >= nrows) {
                // the absent second row of a pair transforms as zeros
                if (fused)
                    for (int c = 0; c < 3; ++c) reinterpret_cast<T *>(&sm[(c * np + q) * LP + sk<SK>(x)])[hi] = (T)0;
                continue;
            }
            V3<T> Hd = v3<T>(0, 0, 0);
            if (has_demag == 1) {
                T h3[3];
                for (int c = 0; c < 3; ++c) h3[c] = reinterpret_cast<const T *>(&sm[(c * np + q) * LP + sk<SK>(x)])[hi];
                Hd = v3<T>(h3[0], h3[1], h3[2]);
            } else if (has_demag == 2) {
                const size_t i = (size_t)gr * nx + x;
                Hd = v3<T>(Hdem[i], Hdem[S.n + i], Hdem[2 * S.n + i]);
            }
            T nxt[3];
            stage_cell<T, EN, STT, GEN, P1S>(S, G, ctl, Y, x, yr, zr, Hd, hz, red, ea, nxt);
            // the half slot (row, x) of the three lines is this thread's
            // alone: the next stage input replaces H_demag in place
            if (fused)
                for (int c = 0; c < 3; ++c) reinterpret_cast<T *>(&sm[(c * np + q) * LP + sk<SK>(x)])[hi] = nxt[c];
        }
    }
    if (fused) {
        for (int l = 0; l < 3 * np; ++l)
            for (int x = nx + threadIdx.x; x < L; x += blockDim.x) sm[(size_t)l * LP + sk<SK>(x)] = cmk<T>((T)0, (T)0);
        __syncthreads();
        {
            const int l = threadIdx.x / TL, t = threadIdx.x % TL;
            const bool active = (l < 3 * np) && row0 + 2 * (np == 1 ? 0 : l % np) < nrows;
            C2<T> *line = sm + (size_t)l * LP;
            const LineMap m = {t, active};
            fft_line<T, false, false, false, LC, SK>(
                p, twl, line, m, line, m, [](int) { return cmk<T>((T)0, (T)0); }, [](int, C2<T>) {});
        }
        // forward: Z = FFT(x0 + i x1); X0[k] = (Z[k] + conj Z[L-k]) / 2,
        // X1[k] = (Z[k] - conj Z[L-k]) / 2i for k < Hx
        for (int l = 0; l < 3 * np; ++l) {
            const int c = l / np, q = l % np;
            const int gr0 = row0 + 2 * q;
            if (gr0 >= nrows) continue;
            const bool has1 = gr0 + 1 < nrows;
            const C2<T> *ln = sm + (size_t)l * LP;
            C2<T> *d0 = gfield(G.fwd_S1) + ((size_t)c * nrows + gr0) * Hp;
            for (int k = threadIdx.x; k < Hx; k += blockDim.x) {
                const C2<T> zk = ln[sk<SK>(k)], zm = ln[sk<SK>(k == 0 ? 0 : L - k)];
                d0[k] = cmk<T>((zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5);
                if (has1) d0[(size_t)Hp + k] = cmk<T>((zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5);
            }
        }
    }
    __shared__ double wred[4][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)((blockDim.x + 63) >> 6);
    if (EN) {
        __shared__ double wsum[8][16];
        const Dd a = wave_sum_dd(ea.demag), b = wave_sum_dd(ea.exch), c = wave_sum_dd(ea.aniso),
                 d = wave_sum_dd(ea.zee);
        if (lane == 0) {
            wsum[0][w] = a.s;
            wsum[1][w] = a.e;
            wsum[2][w] = b.s;
            wsum[3][w] = b.e;
            wsum[4][w] = c.s;
            wsum[5][w] = c.e;
            wsum[6][w] = d.s;
            wsum[7][w] = d.e;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            const int q = threadIdx.x;
            Dd acc = {wsum[2 * q][0], wsum[2 * q + 1][0]};
            for (int v = 1; v < nw; ++v) {
                const Dd o = {wsum[2 * q][v], wsum[2 * q + 1][v]};
                acc = dd_merge(acc, o);
            }
            epart[(size_t)(2 * q) * nblk + blk] = acc.s;
            epart[(size_t)(2 * q + 1) * nblk + blk] = acc.e;
        }
        return;
    }
    if ((!P1S && (G.prep == 2 || G.dydt_from_k0)) || G.want_dydt) {
        const double a = wave_max(red.err), b = wave_max(red.dydt), c = wave_max(red.draw), d = wave_max(red.mraw);
        if (lane == 0) {
            wred[0][w] = a;
            wred[1][w] = b;
            wred[2][w] = c;
            wred[3][w] = d;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double ea0 = wred[0][0], eb = wred[1][0], ec = wred[2][0], ed = wred[3][0];
            for (int q = 1; q < nw; ++q) {
                ea0 = fmax(ea0, wred[0][q]);
                eb = fmax(eb, wred[1][q]);
                ec = fmax(ec, wred[2][q]);
                ed = fmax(ed, wred[3][q]);
            }
            double *mp = gfield(G.mpart);
            if (!P1S && G.prep == 2) mp[blk] = ea0;
            if (G.want_dydt || (!P1S && G.dydt_from_k0)) {
                mp[(size_t)nblk + blk] = eb;
                mp[2 * (size_t)nblk + blk] = ec;
                mp[3 * (size_t)nblk + blk] = ed;
            }
        }
    }
}

// P1S: the record is an inner stage with a summed combination (prep 1,
// FORM_SUM, no FSAL reads), so the final-stage code and its record reads
// are compiled out (RKF45 / CC45: five of the six stage launches).  In FP32
// that kernel needs 55-59 VGPRs and is held to 7 waves per SIMD by its 106
// SGPRs; asking for 8 (SGPRs 78, the rest spilled to VGPR lanes, no
// scratch) lets a CU hold 8 of its 256-thread workgroups, one round for
// 256x256x16 cells: stage 32.4 -> 30.7 us (MI355X, same-run A/B).
template <typename T, bool EN, int LC, bool STT, bool GEN, bool P1S = false>
__global__ void __attribute__((amdgpu_waves_per_eu((P1S && sizeof(T) == 4) ? 8 : 1))) k_stage(const SysDev<T> *__restrict__ Sp, const StageDev<T> *__restrict__ Gp, StepCtl *ctl,
                        const C2<T> *__restrict__ S1, FftPlanDev p, int Hx, int Hp, int rpw, int has_demag,
                        double *__restrict__ epart, const T *__restrict__ Hdem)
{
    stage_body<T, EN, LC, STT, GEN, P1S>(Sp, Gp, ctl, S1, p, Hx, Hp, rpw, has_demag, epart, Hdem, (int)blockIdx.x,
                                    (int)gridDim.x, true);
}

// Relaxation rule on the state the attempt started from; true: stop there
// (the attempt is discarded).  MM_RELAX_DMDT compares scaled_abs_max(dM/dt,
// Ms) with the threshold; MM_RELAX_DEG_PER_NS is the reference's Relaxed
// condition (condition.py:169-200) on deg_per_ns
// (landau_lifshitz_gilbert.py:127-131) with state.h = the last accepted h.
__device__ bool relax_check(StepCtl *ctl, double dmax_scaled, double draw2, double mraw2)
{
    if (!ctl->relax_mode || ctl->last_check == ctl->accepted) return false;
    if (ctl->accepted % ctl->relax_every != 0) return false;
    ctl->last_check = ctl->accepted;
    bool relaxed;
    double v;
    if (ctl->relax_mode == 1) {
        v = dmax_scaled;
        relaxed = v < ctl->relax_thr;
    } else {
        const double dm = sqrt(draw2);
        const double mm = sqrt(mraw2);
        const double h = ctl->h_last;
        const double deg_per_step = (180.0 / 3.14159265358979323846) * atan2(dm * h, mm);
        v = 1e-9 * deg_per_step / h;
        relaxed = ctl->have_last && v < ctl->relax_last && v <= ctl->relax_thr;
        ctl->relax_last = v;
        ctl->have_last = 1;
    }
    ctl->relax_value = v;
    if (relaxed) {
        ctl->relaxed = 1;
        ctl->stop = 1;
    }
    return relaxed;
}

__device__ __forceinline__ void accept_step(StepCtl *ctl, double h)
{
    ctl->t += h;
    ctl->h_last = h;
    ctl->cur = 1 - ctl->cur;
    ctl->accepted += 1;
    if (ctl->accepted >= ctl->target || (ctl->tmax_on && ctl->t >= ctl->t_max) ||
        (ctl->tstop_on && ctl->t >= ctl->t_stop))
        ctl->stop = 1;
}

// MicroMagneticsStepSizeController.adjust_stepsize
// (micro_magnetics_stepsize_controller.py:35-73) + RungeKutta.evolve retry
// and overshoot handling (runge_kutta.py:56-86).  Non-adaptive (rk4): every
// attempt is accepted with the fixed h.
// mode: 1 adaptive; 0 fixed step (rk4), last step clipped to t_max; 2 fixed
// step without clipping (euler.py:27-36 ignores t_max)
__device__ __forceinline__ void controller_update(StepCtl &c, const double m[4], int order, double eps_abs,
                                                  double eps_rel, int Ms_uniform, double Ms_u, int mode)
{
    const double h = c.h;
    double dmax = sqrt(m[1]);
    if (Ms_uniform) dmax /= Ms_u;
    const bool relaxed = relax_check(&c, dmax, m[2], m[3]);
    if (relaxed) return;
    c.attempts += 1;
    if (c.forced) {
        c.forced = 0;
        accept_step(&c, h);
        c.h = c.h_saved;
        return;
    }
    int accept = 1;
    double hn = h;
    if (mode == 1) {
        double e_abs = sqrt(m[0]);
        if (Ms_uniform) e_abs /= Ms_u;
        const double HEADROOM = 0.85, MINS = 0.2, MAXS = 5.0, MINH = 1e-15, MAXH = 1e-10;
        const double den = h * dmax;
        if (den == 0.0) {
            hn = 5e-14;
        } else {
            const double e_rel = e_abs / den;
            accept = (e_abs <= eps_abs) && (e_rel <= eps_rel);
            const double sa = HEADROOM * pow(eps_abs / e_abs, 1.0 / (order + 1.0));
            const double sr = HEADROOM * pow(eps_rel / e_rel, 1.0 / order);
            double sc = sa < sr ? sa : sr;
            if (sc < MINS) sc = MINS;
            if (sc > MAXS) sc = MAXS;
            hn = h * sc;
            if (hn < MINH) hn = MINH;
            if (hn > MAXH) hn = MAXH;
            if (hn == MINH) accept = 1;
        }
    }
    if (accept) {
        if (mode != 2 && c.tmax_on && c.t + h > c.t_max) {
            c.forced = 1;
            c.h_saved = hn;
            c.h = c.t_max - c.t;
            return;
        }
        accept_step(&c, h);
    } else {
        c.rejections += 1;
    }
    c.h = hn;
}

// Step-size control after an attempt (one workgroup): the maxima of the
// per-block partials of the attempt's stage launches, then controller_update
// on the control block.
__device__ __forceinline__ void controller_body(StepCtl *ctl, const double *__restrict__ mpart, int nb, int order,
                                                double eps_abs, double eps_rel, int Ms_uniform, double Ms_u, int mode)
{
    // The control block (thread 0) and the partials load together; the stop
    // flag is looked at only when the maxima are there (after a stopped run
    // the reduction's result is simply not used), so the kernel waits for
    // one memory round trip instead of two.
    StepCtl c;
    if (threadIdx.x == 0) c = *ctl;
    // maxima of the per-block partials of this attempt's stage launches
    double m[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int q = 0; q < 4; ++q) m[q] = fmax(m[q], mpart[(size_t)q * nb + b]);
    __shared__ double red[4][16];
    for (int q = 0; q < 4; ++q) {
        const double w = wave_max(m[q]);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x != 0 || c.stop) return;
    for (int q = 0; q < 4; ++q) {
        double v = red[q][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, red[q][w]);
        m[q] = v;
    }
    // the control block is updated in registers and written back once
    controller_update(c, m, order, eps_abs, eps_rel, Ms_uniform, Ms_u, mode);
    *ctl = c;
}

__global__ void __launch_bounds__(1024) k_controller(StepCtl *ctl, const double *__restrict__ mpart, int nb,
                                                    int order, double eps_abs, double eps_rel, int Ms_uniform,
                                                    double Ms_u, int mode)
{
    controller_body(ctl, mpart, nb, order, eps_abs, eps_rel, Ms_uniform, Ms_u, mode);
}

// ytmp_1 = y; ytmp_1 += k_0 * (h b10) for attempts that start from a FSAL k_0
template <typename T>
__global__ void k_prep_first(size_t m3, const T *__restrict__ y0, const T *__restrict__ y1, const T *__restrict__ k0,
                             T *__restrict__ out, const StepCtl *ctl, double b10)
{
    if (ctl->stop) return;
    const T *y = ctl->cur ? y1 : y0;
    const T f = (T)(ctl->h * b10);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m3; i += (size_t)gridDim.x * blockDim.x)
        out[i] = y[i] + k0[i] * f;
}

// k_prep_first fused with the x forward transform of its result: rows
// (c, z, y) of ytmp_1 = y + k_0 (h b10), written to `out` and, transformed
// along x (zero-extended to L), to S1 rows (pitch Hp)
template <typename T, int LC = 0>
__global__ void k_xfwd_prep(const T *__restrict__ y0, const T *__restrict__ y1, const T *__restrict__ k0,
                            T *__restrict__ out, C2<T> *__restrict__ S1, const StepCtl *ctl, double b10, int nx,
                            int ny, int nz, int Hx, int Hp, FftPlanDev p, int lpw)
{
    constexpr int SK = MM_SKEW_X;
    if (ctl->stop) return;
    extern __shared__ char smem_raw[];
    C2<T> *sm = reinterpret_cast<C2<T> *>(smem_raw);
    const T *y = ctl->cur ? y1 : y0;
    const T f = (T)(ctl->h * b10);
    const int L = PlanC<LC>::L(p), LP = lds_len<SK>(L), TL = PlanC<LC>::tline(p);
    const int nrows = 3 * ny * nz;
    // two real rows per complex line (rows 2 g and 2 g + 1, real and
    // imaginary part); S1 rows and `out` rows follow the row index
    const int l = threadIdx.x / TL, t = threadIdx.x % TL;
    const int g0 = 2 * (blockIdx.x * lpw + l);
    const bool active = l < lpw && g0 < nrows;
    const bool has1 = g0 + 1 < nrows;
    C2<T> *line = sm + (size_t)l * LP;
    C2<T> *twl = sm + (size_t)lpw * LP;
    const TwFetch<T> twf = TwFetch<T>::fetch(p);
    const size_t base0 = (size_t)(active ? g0 : 0) * nx;
    const size_t base1 = base0 + nx;
    const LineMap m = {t, active};
    fft_line<T, false, true, false, LC, SK>(
        p, twl, line, m, line, m,
        [&](int x) {
            T v0 = (T)0, v1 = (T)0;
            if (x < nx) {
                v0 = y[base0 + x] + k0[base0 + x] * f;
                out[base0 + x] = v0;
                if (has1) {
                    v1 = y[base1 + x] + k0[base1 + x] * f;
                    out[base1 + x] = v1;
                }
            }
            return cmk<T>(v0, v1);
        },
        [](int, C2<T>) {}, [&] { twf.store(twl); });
    // X0[k] = (Z[k] + conj Z[L-k]) / 2, X1[k] = (Z[k] - conj Z[L-k]) / 2i
    for (int l2 = 0; l2 < lpw; ++l2) {
        const int h0 = 2 * (blockIdx.x * lpw + l2);
        if (h0 >= nrows) break;
        const C2<T> *ln = sm + (size_t)l2 * LP;
        C2<T> *d0 = S1 + (size_t)h0 * Hp;
        const bool has1 = h0 + 1 < nrows;
        for (int k = threadIdx.x; k < Hx; k += blockDim.x) {
            const C2<T> zk = ln[sk<SK>(k)], zm = ln[sk<SK>(k == 0 ? 0 : L - k)];
            d0[k] = cmk<T>((zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5);
            if (has1) d0[(size_t)Hp + k] = cmk<T>((zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5);
        }
    }
}

template <typename T> __global__ void k_cast_in(size_t n, const double *__restrict__ a, T *__restrict__ b)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (T)a[i];
}
template <typename T> __global__ void k_cast_out(size_t n, const T *__restrict__ a, double *__restrict__ b)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}
// per-block compensated sums of the three components of y[cur]: component
// c's pair at part[2c MM_RED_BLOCKS + b] and part[(2c+1) MM_RED_BLOCKS + b]
template <typename T> __global__ void k_sum3_sel(size_t n, const T *y0, const T *y1, const StepCtl *ctl, double *part)
{
    const T *A = ctl->cur ? y1 : y0;
    Dd s0 = {0.0, 0.0}, s1 = {0.0, 0.0}, s2 = {0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        dd_add(s0, (double)A[i]);
        dd_add(s1, (double)A[n + i]);
        dd_add(s2, (double)A[2 * n + i]);
    }
    block_sum_dd_store(s0, part, MM_RED_BLOCKS);
    block_sum_dd_store(s1, part + 2 * MM_RED_BLOCKS, MM_RED_BLOCKS);
    block_sum_dd_store(s2, part + 4 * MM_RED_BLOCKS, MM_RED_BLOCKS);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct StepperBase {
    virtual ~StepperBase() {}
    virtual void set_m(const double *M_host) = 0;
    virtual void get_m(double *M_host) = 0;
    virtual void dmdt(double *dM_host, double *H_host) = 0;
    virtual void run(long long steps) = 0;
    virtual void sync() = 0;
    virtual void stats(double *s5) = 0;
    virtual void set_time(double t, double h_next) = 0;
    virtual long long step_count() = 0;
    virtual void average(double *a3) = 0;
    virtual hipStream_t stream() = 0;
    virtual void profile(int attempts, double *out) = 0;
    virtual void energy(double *E5) = 0;
    virtual void relax(int criterion, double thr, int every, long long max_steps, double *info3) = 0;
    virtual void run_until(double t_end, long long max_steps, double t_stop) = 0;
    virtual void set_field_fn(mm_field_fn_t fn, void *user) = 0;
    virtual MmShape launch_shape() const = 0;
    // ensemble driving: arm(steps) sets the target, then rounds of
    // enqueue_some() on the stepper's stream and refresh() after the stream
    // is synchronised, while pending()
    virtual void arm(long long steps) = 0;
    virtual bool pending() const = 0;
    virtual void enqueue_some(long long batch) = 0;
    virtual void refresh() = 0;
};

static bool all_equal(const double *a, size_t n)
{
    for (size_t i = 1; i < n; ++i)
        if (a[i] != a[0]) return false;
    return true;
}

template <typename T> struct Stepper : public StepperBase {
    mm_system_desc_t d;
    size_t n = 0;
    hipStream_t s = nullptr;
    ConvImpl<T> *conv = nullptr;
    SysDev<T> sys;
    mm_tableau_t tab;
    std::vector<void *> allocs;
    T *yb[2] = {nullptr, nullptr};
    T *k[8] = {nullptr};
    T *tmp[2] = {nullptr, nullptr};
    T *Hbuf = nullptr;
    double *dbuf = nullptr;
    double *part = nullptr;
    double *epart = nullptr;
    double *mpart = nullptr;
    StepCtl *ctl = nullptr;
    int rpw = 1;
    size_t stage_lds = 0;
    bool split_x = false;   // inverse x transform as a separate pass
    T *Hdem = nullptr;      // H_demag of the split mode
    int grid_rows = 1;
    int threads = 256;

    T *dalloc(size_t count)
    {
        void *p = nullptr;
        MM_HIP(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
        allocs.push_back(p);
        return (T *)p;
    }

    Param<T> param(const double *h, double fallback)
    {
        Param<T> p;
        p.p = nullptr;
        p.u = (T)fallback;
        if (!h) return p;
        if (all_equal(h, n)) {
            p.u = (T)h[0];
            return p;
        }
        T *dp = dalloc(n);
        upload(dp, h, n);
        p.p = dp;
        return p;
    }

    VParam<T> vparam(const double *h)
    {
        VParam<T> p;
        p.p = nullptr;
        p.u[0] = p.u[1] = p.u[2] = (T)0;
        if (!h) return p;
        if (all_equal(h, n) && all_equal(h + n, n) && all_equal(h + 2 * n, n)) {
            for (int c = 0; c < 3; ++c) p.u[c] = (T)h[c * n];
            return p;
        }
        T *dp = dalloc(3 * n);
        upload(dp, h, 3 * n);
        p.p = dp;
        return p;
    }

    void upload(T *dst, const double *src, size_t cnt)
    {
        MM_HIP(hipMemcpyAsync(dbuf, src, sizeof(double) * cnt, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_cast_in<T>, dim3(1024), dim3(256), 0, s, cnt, dbuf, dst);
        MM_HIP(hipGetLastError());
        MM_HIP(hipStreamSynchronize(s));
    }

    void download(double *dst, const T *src, size_t cnt)
    {
        hipLaunchKernelGGL(k_cast_out<T>, dim3(1024), dim3(256), 0, s, cnt, src, dbuf);
        MM_HIP(hipGetLastError());
        MM_HIP(hipMemcpyAsync(dst, dbuf, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
        MM_HIP(hipStreamSynchronize(s));
    }

    bool own_stream = true;

    // ext: a stream shared with other steppers (ensembles), else an own one
    void init(const mm_system_desc_t &desc, hipStream_t ext = nullptr)
    {
        d = desc;
        MM_REQUIRE(d.nx > 0 && d.ny > 0 && d.nz > 0, "bad mesh size");
        MM_REQUIRE(d.Ms_host && d.alpha_host, "Ms and alpha are required");
        n = (size_t)d.nx * d.ny * d.nz;
        if (ext) {
            s = ext;
            own_stream = false;
        } else {
            MM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        }
        MM_HIP(hipMalloc(&dbuf, sizeof(double) * 3 * n));
        MM_HIP(hipMalloc(&part, sizeof(double) * 6 * MM_RED_BLOCKS));
        MM_HIP(hipMalloc(&ctl, sizeof(StepCtl)));
        int rc = mm_tableau(d.method, &tab);
        if (rc != MM_OK) throw mm_error(rc, mm_last_error());

        // physical parameters; LLG factors as landau_lifshitz_gilbert.py:133-158
        const double mu0 = MM_MU0_VALUE;
        std::vector<double> f1(n), f2(n);
        for (size_t i = 0; i < n; ++i) {
            const double a = d.alpha_host[i], ms = d.Ms_host[i];
            if (ms != 0.0) {
                const double gp = MM_GAMMA / (1.0 + a * a);
                f1[i] = d.do_precess ? -gp : 0.0;
                f2[i] = -a * gp / ms;
            } else {
                f1[i] = 0.0;
                f2[i] = 0.0;
            }
        }
        std::memset(&sys, 0, sizeof(sys));
        {
            const char *e = std::getenv("MM_XCD_REMAP");
            sys.xcd_remap = e ? (std::atoi(e) != 0) : 1;
        }
        sys.m.nx = d.nx;
        sys.m.ny = d.ny;
        sys.m.nz = d.nz;
        sys.m.px = d.px != 0;
        sys.m.py = d.py != 0;
        sys.m.pz = d.pz != 0;
        sys.n = n;
        sys.dx = (T)d.dx;
        sys.dy = (T)d.dy;
        sys.dz = (T)d.dz;
        sys.wx = (T)(1.0 / (d.dx * d.dx));
        sys.wy = (T)(1.0 / (d.dy * d.dy));
        sys.wz = (T)(1.0 / (d.dz * d.dz));
        sys.coef_ex = (T)(2.0 / mu0);
        sys.mu0 = (T)mu0;
        sys.cm2mu0 = (T)(-2.0 / mu0);
        sys.mu_b = (T)MM_MU_BOHR;
        sys.e_charge = (T)MM_E_CHARGE;
        sys.Ms_uniform = all_equal(d.Ms_host, n);
        sys.Ms_u = d.Ms_host[0];
        // exchange and Zhang-Li index Ms/A through arrays in exchange_cell:
        // keep real arrays for them (uniform values still use scalars elsewhere)
        {
            T *dp = dalloc(n);
            upload(dp, d.Ms_host, n);
            sys.Ms.p = dp;
            sys.Ms.u = (T)d.Ms_host[0];
        }
        sys.alpha = param(d.alpha_host, 0.0);
        sys.f1 = param(f1.data(), 0.0);
        sys.f2 = param(f2.data(), 0.0);
        sys.has_exch = d.A_host != nullptr;
        if (sys.has_exch) {
            T *dp = dalloc(n);
            upload(dp, d.A_host, n);
            sys.A.p = dp;
            sys.exch_uniform = sys.Ms_uniform && sys.Ms_u != 0.0 && all_equal(d.A_host, n);
            if (sys.exch_uniform) sys.exch_f = (T)(2.0 * d.A_host[0] / (mu0 * sys.Ms_u * sys.Ms_u));
        }
        sys.has_uni = d.k_uni_host != nullptr;
        sys.has_cub = d.k_cub_host != nullptr;
        if (sys.has_uni) sys.kuni = param(d.k_uni_host, 0.0);
        if (sys.has_cub) sys.kcub = param(d.k_cub_host, 0.0);
        if (sys.has_uni || sys.has_cub) {
            MM_REQUIRE(d.axis1_host, "anisotropy needs axis1");
            sys.axis1 = vparam(d.axis1_host);
        }
        if (sys.has_cub) {
            MM_REQUIRE(d.axis2_host, "cubic anisotropy needs axis2");
            sys.axis2 = vparam(d.axis2_host);
        }
        sys.uni_fast = sys.has_uni && !sys.kuni.p && !sys.axis1.p && sys.Ms_uniform && sys.Ms_u != 0.0;
        if (sys.uni_fast) {
            const double k = d.k_uni_host[0], ms = sys.Ms_u;
            sys.uni_c = (T)(2.0 * k / (mu0 * ms * ms));
            sys.uni_e = (T)(k / (ms * ms));
        }
        sys.has_zee = d.use_zeeman;
        sys.has_ac = d.use_zeeman && d.use_ac;
        for (int c = 0; c < 3; ++c) {
            sys.hoff[c] = d.h_ext[c];
            sys.ac_amp[c] = d.h_amp[c];
            sys.ac_freq[c] = d.h_freq[c];
            sys.ac_phase[c] = d.h_phase[c];
        }
        for (int c = 0; c < 3; ++c) sys.hext[c] = (T)d.h_ext[c];
        sys.has_stt = d.P_host != nullptr;
        sys.do_precess = d.do_precess;
        if (sys.has_stt) {
            MM_REQUIRE(d.xi_host && d.j_host, "spin torque needs xi and j");
            sys.P = param(d.P_host, 0.0);
            sys.xi = param(d.xi_host, 0.0);
            sys.j = vparam(d.j_host);
            sys.stt_fast = !sys.P.p && !sys.xi.p && !sys.j.p && !sys.alpha.p && sys.Ms_uniform && sys.Ms_u != 0.0;
            if (sys.stt_fast) {
                // fdm_zhangli_cpu.cpp:95-114 in double
                const double P = d.P_host[0], xi = d.xi_host[0], al = d.alpha_host[0], ms = sys.Ms_u;
                const double bj = P * MM_MU_BOHR / (MM_E_CHARGE * ms * (1.0 + xi * xi));
                const double bjp = bj / (1.0 + al * al);
                sys.stt_f1 = (T)(d.do_precess ? -bjp / (ms * ms) * (1.0 + al * xi) : -bjp / (ms * ms) * (al * xi));
                sys.stt_f2 = (T)(d.do_precess ? -bjp / ms * (xi - al) : -bjp / ms * (-al));
                sys.inv_d[0] = (T)(1.0 / d.dx);
                sys.inv_d[1] = (T)(1.0 / d.dy);
                sys.inv_d[2] = (T)(1.0 / d.dz);
            }
        }

        // launch shape in force for this mesh (tuner candidate or table entry)
        const MmShape wshape = mm_shape_for(d.nx, d.ny, d.nz, sizeof(T));
        if (d.N_host) {
            conv = new ConvImpl<T>();
            conv->init(d.N_host, d.nx, d.ny, d.nz, d.ex, d.ey, d.ez);
            const int tl = conv->px.tline;
            const size_t cs = sizeof(C2<T>), L = conv->px.L;
            split_x = 3 * tl > 1024 || (3 * lds_len<MM_SKEW_X>((int)L) + L) * cs > (size_t)mm_max_lds_bytes();
            if (const char *e = std::getenv("MM_SPLIT_X")) split_x = split_x || std::atoi(e) != 0;
            if (split_x) {
                // the inverse x transform runs as its own pass into Hdem
                rpw = std::max(1, 256 / d.nx);
                threads = 256;
                stage_lds = 0;
                Hdem = dalloc(3 * n);
            } else {
                // rows go through the x transforms in pairs (two real rows per
                // complex line): np pairs, 3 np lines, rpw = 2 np rows per block
                int np = std::max(1, 192 / (3 * tl));
                if (wshape.rpw > 0) np = std::max(1, wshape.rpw / 2);
                if (const char *e = std::getenv("MM_RPW")) np = std::max(1, std::atoi(e) / 2);
                // a tuned, imported or environment row count is cut to what one
                // workgroup holds: LDS and 1024 threads (3 np lines of tl)
                while (np > 1 && (((size_t)3 * np * lds_len<MM_SKEW_X>((int)L) + L) * cs > (size_t)mm_max_lds_bytes() ||
                                  3 * np * tl > 1024))
                    --np;
                rpw = 2 * np;
                threads = 3 * np * tl;
                stage_lds = ((size_t)3 * np * lds_len<MM_SKEW_X>((int)L) + L) * cs;
            }
        } else {
            rpw = std::max(1, 256 / d.nx);
            threads = 256;
            stage_lds = 0;
        }
        threads = std::max(64, ((threads + 63) / 64) * 64);
        if (conv && !split_x) {
            // two epilogue cells per thread, the two rows of a pair at one x
            // (threads beyond the 3 np x transform lines idle in the FFT
            // passes).  Measured on MI355X, 256x256x16 cells: 256 threads per
            // 2-row block against 512 (one cell each), stage 44.8 vs 49.4 us
            // FP32 and 74.9 vs 83.8 us FP64.  MM_STAGE_THREADS overrides.
            const int fft_threads = threads;
            const int epi = std::max(fft_threads, std::min(1024, (rpw * d.nx / 2 + 63) / 64 * 64));
            // between the x transforms' and the epilogue's thread counts, the
            // one that keeps the most threads resident per CU (ties: more
            // threads).  Measured on MI355X: 1024x1024x2 FP32, 768 threads
            // (two workgroups per CU) against 1024 (one), stage 108 -> 85 us,
            // +10 % per step; the other workloads keep the epilogue's count.
            threads = epi;
            for (int t = fft_threads, best = -1; t <= epi; t += 64) {
                const int res = stage_occupancy(t) * t;
                if (res >= best) {
                    best = res;
                    threads = t;
                }
            }
            if (wshape.stage_threads > 0)
                threads = std::min(1024, std::max(fft_threads, (wshape.stage_threads + 63) / 64 * 64));
            if (const char *e = std::getenv("MM_STAGE_THREADS"))
                threads = std::min(1024, std::max(fft_threads, (std::atoi(e) + 63) / 64 * 64));
        }
        grid_rows = (d.ny * d.nz + rpw - 1) / rpw;
        {
            const char *e = std::getenv("MM_TILE_ZCOL");
            const bool zcol = e ? std::atoi(e) != 0 : true;
            sys.tile_zcol = (zcol && d.nz > 1 && d.ny % rpw == 0) ? d.ny / rpw : 0;
        }
        MM_HIP(hipMalloc(&epart, sizeof(double) * 8 * grid_rows));
        MM_HIP(hipMalloc(&mpart, sizeof(double) * 4 * grid_rows));
        MM_HIP(hipMemset(mpart, 0, sizeof(double) * 4 * grid_rows));
        {
            void *ft = nullptr;
            MM_HIP(hipMalloc(&ft, sizeof(double) * 3 * MM_FN_SLOTS));
            allocs.push_back(ft);
            fn_tab = (double *)ft;
            MM_HIP(hipMemset(fn_tab, 0, sizeof(double) * 3 * MM_FN_SLOTS));
            MM_HIP(hipHostMalloc((void **)&fn_host, sizeof(double) * 3 * MM_FN_SLOTS, hipHostMallocDefault));
        }

        for (int b = 0; b < 2; ++b) yb[b] = dalloc(3 * n);
        for (int q = 0; q < tab.num_steps; ++q) k[q] = dalloc(3 * n);
        tmp[0] = dalloc(3 * n);
        tmp[1] = dalloc(3 * n);
        Hbuf = dalloc(3 * n);
        StepCtl c0;
        std::memset(&c0, 0, sizeof(c0));
        c0.h = d.h0;
        c0.h_last = d.h0;
        MM_HIP(hipMemcpy(ctl, &c0, sizeof(c0), hipMemcpyHostToDevice));
        MM_HIP(hipMemset(yb[0], 0, sizeof(T) * 3 * n));
        build_tables();
    }

    void release()
    {
        for (int g = 0; g < 2; ++g)
            for (int w = 0; w < 2; ++w)
                if (gexec[g][w]) {
                    (void)hipGraphExecDestroy(gexec[g][w]);
                    gexec[g][w] = nullptr;
                }
        if (conv) {
            conv->release();
            delete conv;
            conv = nullptr;
        }
        for (void *p : allocs) (void)hipFree(p);
        allocs.clear();
        if (dbuf) (void)hipFree(dbuf);
        if (part) (void)hipFree(part);
        if (epart) (void)hipFree(epart);
        if (mpart) (void)hipFree(mpart);
        if (gtab) (void)hipFree(gtab);
        if (sysd) (void)hipFree(sysd);
        gtab = nullptr;
        sysd = nullptr;
        if (ctl) (void)hipFree(ctl);
        if (fn_host) (void)hipHostFree(fn_host);
        fn_host = nullptr;
        if (s && own_stream) (void)hipStreamDestroy(s);
        dbuf = part = epart = mpart = nullptr;
        ctl = nullptr;
        s = nullptr;
    }

    ~Stepper() override { release(); }

    hipStream_t stream() override { return s; }

    KernelLog *klog = nullptr;

    // algorithmic bytes of one fused stage launch: demag rows (S1) plus every
    // state array the epilogue reads or writes once (parameters excluded)
    double stage_bytes(const StageDev<T> &G) const
    {
        const double V = 3.0 * sizeof(T) * (double)n;
        double b = conv ? 3.0 * 2.0 * sizeof(T) * conv->g.Hx * d.ny * d.nz : 0.0;
        if (split_x) b += 2.0 * V;                 // Hdem written by the x pass, read here
        b += V;                                    // stage input
        if (G.prep && G.Ys) b += V;                // y (same array as the input on stage 0)
        b += V * G.nk;                             // stored stage vectors
        if (G.kout) b += V;
        if (G.Hout) b += V;
        if (G.prep) b += V;                        // next stage input or new state
        if (G.fwd) b += 3.0 * 2.0 * sizeof(T) * conv->g.Hx * d.ny * d.nz;  // its x transform
        return b;
    }

    // one stage: dM/dt at stage input Ys (null = y[cur])
    int demag_mode() const { return conv ? (split_x ? 2 : 1) : 0; }
    // x transform length inside the stage kernel (0: it runs none)
    int stage_len() const { return demag_mode() == 1 ? conv->px.L : 0; }

    // the stage kernel instantiated for the x transform length and the
    // physics present: STT (Zhang-Li) and GEN (cubic anisotropy or exchange
    // with per-cell Ms/A); the energy launch (EN) has no spin torque term
    template <bool EN, int LC, bool STT, bool GEN, bool P1S = false>
    void stage_launch(const StageDev<T> *Gd, double *ep)
    {
        hipLaunchKernelGGL((k_stage<T, EN, LC, STT, GEN, P1S>), dim3(grid_rows), dim3(threads), stage_lds, s, sysd, Gd,
                           ctl, conv ? conv->S1 : nullptr, conv ? conv->px.template dev<T>() : FftPlanDev(),
                           conv ? conv->g.Hx : 0, conv ? conv->g.Hp : 0, rpw, demag_mode(), ep, Hdem);
    }
    // Gd: the stage record in device memory (gtab)
    // the stage kernel's physics class: STT (Zhang-Li) and GEN (cubic
    // anisotropy, exchange with per-cell Ms/A, or state arrays of 4 GiB or
    // more, which the other instantiation addresses by 32-bit byte offsets)
    void physics_class(bool &stt, bool &gen) const
    {
        const bool big = 3.0 * (double)n * sizeof(T) >= 4294967296.0;
        stt = sys.has_stt != 0;
        gen = sys.has_cub || (sys.has_exch && !sys.exch_uniform) || big;
    }

    // workgroups of the stage kernel (this system's instantiation) that one
    // CU holds at thr threads, with the stage's dynamic LDS
    int stage_occupancy(int thr)
    {
        bool stt, gen;
        physics_class(stt, gen);
        int nb = 0;
        mm_with_len(stage_len(), conv && conv->fft_generic, [&](auto lc) {
            constexpr int LC = decltype(lc)::value;
            const void *f = stt ? (gen ? reinterpret_cast<const void *>(k_stage<T, false, LC, true, true>)
                                       : reinterpret_cast<const void *>(k_stage<T, false, LC, true, false>))
                                : (gen ? reinterpret_cast<const void *>(k_stage<T, false, LC, false, true>)
                                       : reinterpret_cast<const void *>(k_stage<T, false, LC, false, false>));
            MM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, thr, stage_lds));
        });
        return nb;
    }

    // p1s: the record is an inner stage with a summed combination (the
    // classes without spin torque have a kernel compiled for it)
    template <bool EN> void stage_kernel(const StageDev<T> *Gd, double *ep, bool p1s = false)
    {
        bool stt, gen;
        physics_class(stt, gen);
        stt = stt && !EN;
        mm_with_len(stage_len(), conv && conv->fft_generic, [&](auto lc) {
            constexpr int LC = decltype(lc)::value;
            if (stt) {
                if (gen) stage_launch<EN, LC, !EN, true>(Gd, ep);
                else stage_launch<EN, LC, !EN, false>(Gd, ep);
            } else {
                if (gen) {
                    if constexpr (!EN) {
                        if (p1s) {
                            stage_launch<EN, LC, false, true, true>(Gd, ep);
                            return;
                        }
                    }
                    stage_launch<EN, LC, false, true>(Gd, ep);
                } else {
                    if constexpr (!EN) {
                        if (p1s) {
                            stage_launch<EN, LC, false, false, true>(Gd, ep);
                            return;
                        }
                    }
                    stage_launch<EN, LC, false, false>(Gd, ep);
                }
            }
        });
    }

    // need_fwd: S1 does not yet hold the x transform of the stage input
    void launch_stage(const T *Ys, const StageDev<T> &G, const StageDev<T> *Gd, bool need_fwd = true)
    {
        if (conv) {
            conv->klog = klog;
            if (need_fwd) {
                if (Ys) conv->forward_x(Ys, s);
                else conv->forward_x(yb[0], s, yb[1], &ctl->cur);
            }
            conv->spectral(s);
            if (split_x) conv->inverse_x(Hdem, s);
        }
        stage_kernel<false>(Gd, nullptr, G.prep == 1 && G.form == FORM_SUM && !G.dydt_from_k0 && !G.fsal_out);
        MM_HIP(hipGetLastError());
        if (klog) klog->mark(s, KC_STAGE, stage_bytes(G));
    }

    StageDev<T> blank_stage(const T *Ys) const
    {
        StageDev<T> G;
        std::memset(&G, 0, sizeof(G));
        G.Ys = Ys;
        G.yb0 = yb[0];
        G.yb1 = yb[1];
        G.norm_mode = d.normalize_mode;
        G.mpart = mpart;
        return G;
    }

    // the record of stage st of an attempt (the same for attempts that start
    // from a carried-over k_0: those skip stage 0)
    StageDev<T> stage_record(int st) const
    {
        const int ns = tab.num_steps;
        const bool six = ns == 6;
        const bool fsal = tab.fsal != 0;
        const T *Ys = st == 0 ? nullptr : tmp[(st - 1) & 1];
        StageDev<T> G = blank_stage(Ys);
        G.tnode = tab.a[st];
        G.kout = (st < ns - 1) ? k[st] : nullptr;
        G.want_dydt = (st == 0) && !fsal;
        if (conv && !split_x && st < ns - 1) {
            G.fwd = 1;
            G.fwd_S1 = conv->S1;
        }
        G.form = six ? FORM_SUM : FORM_SEQ;
        if (st < ns - 1) {
            G.prep = 1;
            G.prep_out = tmp[st & 1];
            G.nk = 0;
            for (int q = 0; q < st; ++q) {
                G.kin[G.nk] = k[q];
                G.b[G.nk++] = tab.b[st + 1][q];
            }
            G.bself = tab.b[st + 1][st];
            if (!six) {
                // sequential form of runge_kutta.py:107-111 / runge_kutta_4.py keeps
                // zero coefficients out (RK4 only adds k_s with h*b)
                int m2 = 0;
                for (int q = 0; q < G.nk; ++q)
                    if (G.b[q] != 0.0) {
                        G.kin[m2] = G.kin[q];
                        G.b[m2++] = G.b[q];
                    }
                G.nk = m2;
            }
        } else {
            G.prep = 2;
            G.nk = 0;
            const bool skip1 = six && tab.c[1] == 0.0 && tab.ec[1] == 0.0;
            for (int q = 0; q < st; ++q) {
                if (q == 1 && skip1) continue;
                G.kin[G.nk] = k[q];
                G.b[G.nk] = tab.c[q];
                G.ec[G.nk++] = tab.ec[q];
            }
            G.bself = tab.c[st];
            G.ecself = tab.ec[st];
            if (fsal) {
                G.dydt_from_k0 = 1;
                G.fsal_out = k[0];
            }
        }
        return G;
    }

    // device copies of the stage records: stages 0..ns-1, then dM/dt, then
    // the energy launch; and of the system record.  Uploaded once at set-up,
    // so captured graphs and eager launches read the same fixed addresses.
    std::vector<StageDev<T>> htab;
    StageDev<T> *gtab = nullptr;
    SysDev<T> *sysd = nullptr;
    void build_tables()
    {
        const int ns = tab.num_steps;
        htab.clear();
        for (int st = 0; st < ns; ++st) htab.push_back(stage_record(st));
        StageDev<T> G = blank_stage(nullptr);
        G.kout = tmp[0];
        G.Hout = Hbuf;
        htab.push_back(G);                      // dM/dt
        htab.push_back(blank_stage(nullptr));   // energy
        for (size_t q = 0; q < htab.size(); ++q) htab[q].slot = (int)q;
        MM_REQUIRE((int)htab.size() <= MM_FN_SLOTS, "too many stage records");
        for (const StageDev<T> &r : htab)
            MM_REQUIRE(r.nk >= 0 && r.nk <= MM_STAGE_MAXK, "stage combination reads too many stage vectors");
        MM_HIP(hipMalloc(&gtab, sizeof(StageDev<T>) * htab.size()));
        MM_HIP(hipMemcpy(gtab, htab.data(), sizeof(StageDev<T>) * htab.size(), hipMemcpyHostToDevice));
        MM_HIP(hipMalloc(&sysd, sizeof(SysDev<T>)));
        MM_HIP(hipMemcpy(sysd, &sys, sizeof(SysDev<T>), hipMemcpyHostToDevice));
    }

    bool have_fsal = false;

    // one attempt of the tableau from y[cur] with h from the control block
    void enqueue_attempt()
    {
        const int ns = tab.num_steps;
        const bool fsal = tab.fsal != 0;
        int st0 = 0;
        bool s1_ready = false;  // S1 holds the x transform of the next stage input
        if (fsal && have_fsal) {
            // k_0 is the last stage of the previous attempt: only ytmp_1 is needed
            if (conv) {
                const FftPlanHost &px = conv->px;
                const int lpw = conv->ln.lpw_x;
                const int nrows = 3 * d.ny * d.nz;
                const size_t sh = ((size_t)lpw * lds_len<MM_SKEW_X>(px.L) + px.L) * sizeof(C2<T>);
                mm_with_len(px.L, conv->fft_generic, [&](auto lc) {
                    hipLaunchKernelGGL((k_xfwd_prep<T, decltype(lc)::value>), dim3(((nrows + 1) / 2 + lpw - 1) / lpw),
                                       dim3(lpw * px.tline), sh, s, yb[0], yb[1], k[0], tmp[0], conv->S1, ctl,
                                       tab.b[1][0], d.nx, d.ny, d.nz, conv->g.Hx, conv->g.Hp, px.template dev<T>(),
                                       lpw);
                });
                MM_HIP(hipGetLastError());
                if (klog)
                    klog->mark(s, KC_XFWD, 3.0 * 3.0 * sizeof(T) * n + 3.0 * 2.0 * sizeof(T) * conv->g.Hx * d.ny * d.nz);
                s1_ready = true;
            } else {
                const size_t m3 = 3 * n;
                hipLaunchKernelGGL(k_prep_first<T>, dim3((int)std::min<size_t>((m3 + 255) / 256, 16384)), dim3(256), 0,
                                   s, m3, yb[0], yb[1], k[0], tmp[0], ctl, tab.b[1][0]);
                MM_HIP(hipGetLastError());
                if (klog) klog->mark(s, KC_CTRL, 3.0 * 3.0 * sizeof(T) * n);
            }
            st0 = 1;
        }
        for (int st = st0; st < ns; ++st) {
            const StageDev<T> &G = htab[st];
            launch_stage(G.Ys, G, gtab + st, !s1_ready);
            s1_ready = G.fwd != 0;
        }
        if (fsal) have_fsal = true;
        hipLaunchKernelGGL(k_controller, dim3(1), dim3(1024), 0, s, ctl, mpart, grid_rows, tab.order, d.eps_abs,
                           d.eps_rel, sys.Ms_uniform, sys.Ms_u, ctl_mode());
        MM_HIP(hipGetLastError());
        if (klog) klog->mark(s, KC_CTRL, 0.0);
    }

    // controller mode: 1 adaptive, 0 fixed step clipped to t_max (rk4), 2
    // fixed step without clipping (euler)
    int ctl_mode() const { return d.method == MM_RK4 ? 0 : (d.method == MM_EULER ? 2 : 1); }

    // Enqueue `attempts` attempts with an event after every kernel; per
    // kernel class: [avg ms per launch, launches, avg algorithmic bytes].
    void profile(int attempts, double *out) override
    {
        KernelLog log;
        const int per = tab.num_steps * 5 + 2;
        log.reserve((size_t)attempts * per + 1);
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        c.target = c.accepted + (1ll << 40);
        c.stop = 0;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
        klog = &log;
        log.mark(s, -1, 0.0);
        try {
            for (int q = 0; q < attempts; ++q) enqueue_attempt();
        } catch (...) {
            klog = nullptr;
            if (conv) conv->klog = nullptr;
            throw;
        }
        klog = nullptr;
        if (conv) conv->klog = nullptr;
        MM_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < 3 * KC_NCLASS; ++k) out[k] = 0.0;
        for (size_t i = 1; i < log.used; ++i) {
            float ms = 0.f;
            MM_HIP(hipEventElapsedTime(&ms, log.ev[i - 1], log.ev[i]));
            const int k = log.cls[i];
            out[3 * k + 0] += ms;
            out[3 * k + 1] += 1.0;
            out[3 * k + 2] += log.bytes[i];
        }
        for (int k = 0; k < KC_NCLASS; ++k)
            if (out[3 * k + 1] > 0) {
                out[3 * k + 0] /= out[3 * k + 1];
                out[3 * k + 2] /= out[3 * k + 1];
            }
        // leave the control block as run() expects it
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        c.target = c.accepted;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
    }

    void set_m(const double *M_host) override
    {
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        upload(yb[c.cur], M_host, 3 * n);
        have_fsal = false;
    }

    void get_m(double *M_host) override
    {
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        download(M_host, yb[c.cur], 3 * n);
    }

    void dmdt(double *dM_host, double *H_host) override
    {
        const int slot = tab.num_steps;
        if (ffn) eval_field_at_t(slot);
        launch_stage(nullptr, htab[slot], gtab + slot);
        if (dM_host) download(dM_host, tmp[0], 3 * n);
        if (H_host) download(H_host, Hbuf, 3 * n);
    }

    // E_stray, E_exch, E_aniso, E_ext, E_tot at the current state
    void energy(double *E5) override
    {
        if (ffn) eval_field_at_t(tab.num_steps + 1);
        if (conv) {
            conv->klog = nullptr;
            conv->forward_x(yb[0], s, yb[1], &ctl->cur);
            conv->spectral(s);
            if (split_x) conv->inverse_x(Hdem, s);
        }
        stage_kernel<true>(gtab + tab.num_steps + 1, epart);
        MM_HIP(hipGetLastError());
        std::vector<double> h((size_t)8 * grid_rows);
        MM_HIP(hipMemcpyAsync(h.data(), epart, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
        MM_HIP(hipStreamSynchronize(s));
        double sum[4];
        for (int q = 0; q < 4; ++q) {
            Dd acc = {0.0, 0.0};
            for (int b = 0; b < grid_rows; ++b) {
                const Dd o = {h[(size_t)(2 * q) * grid_rows + b], h[(size_t)(2 * q + 1) * grid_rows + b]};
                acc = dd_merge(acc, o);
            }
            sum[q] = dd_value(acc);
        }
        // the reference's front-end evaluates these with MU0 = 4 pi 1e-7 in
        // Python doubles (constants.py:21)
        const double mu0 = 4.0 * 3.14159265358979323846 * 1e-7;
        const double V = d.dx * d.dy * d.dz;
        E5[0] = conv ? -mu0 / 2.0 * V * sum[0] : 0.0;
        E5[1] = sys.has_exch ? -mu0 / 2.0 * V * sum[1] : 0.0;
        E5[2] = (sys.has_uni || sys.has_cub) ? sum[2] * V : 0.0;
        E5[3] = sys.has_zee ? -mu0 * V * sum[3] : 0.0;
        E5[4] = E5[0] + E5[1] + E5[2] + E5[3];
    }

    // step until the relaxation rule holds on an accepted state or max_steps
    // accepted steps were made; attempts are enqueued in batches between
    // reads of the control block
    void relax(int criterion, double thr, int every, long long max_steps, double *info3) override
    {
        MM_REQUIRE(criterion == MM_RELAX_DMDT || criterion == MM_RELAX_DEG_PER_NS, "unknown relaxation criterion");
        MM_REQUIRE(every >= 1 && max_steps >= 0, "bad relaxation arguments");
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        const long long a0 = c.accepted;
        c.target = c.accepted + max_steps;
        c.stop = max_steps == 0 ? 1 : 0;
        c.tmax_on = 0;
        c.tstop_on = 0;
        c.relax_mode = criterion;
        c.relax_every = every;
        c.relax_thr = thr;
        c.relaxed = 0;
        c.have_last = 0;
        c.relax_value = 0.0;
        c.last_check = -1;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
        drive(c, 32);
        info3[0] = (double)(c.accepted - a0);
        info3[1] = c.relaxed ? 1.0 : 0.0;
        info3[2] = c.relax_value;
        // a discarded attempt may have overwritten the FSAL vector
        if (c.relaxed) have_fsal = false;
        c.relax_mode = 0;
        c.relaxed = 0;
        c.stop = 0;
        c.target = c.accepted;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
    }

    // enqueue attempts in batches until the device sets stop; c holds the
    // control block on entry (already on the device) and on return
    void drive(StepCtl &c, long long batch)
    {
        while (ffn && !c.stop) {
            // one attempt at a time: its stage times t + a_s h follow the
            // step size the device chose
            eval_stage_fields(c);
            launch_attempt();
            MM_HIP(hipStreamSynchronize(s));
            MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        }
        while (!c.stop) {
            const long long q = std::max(1ll, std::min(batch, c.target - c.accepted));
            launch_attempts(q);
            MM_HIP(hipStreamSynchronize(s));
            MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        }
    }

    // Attempts as replays of captured HIP graphs (graphs for attempts that
    // compute k_0, graphs for FSAL attempts that reuse it): every kernel
    // argument is fixed at set-up and all step state lives in device memory,
    // so the recorded launches are valid for every attempt.  A graph holds
    // one attempt or graph_attempts of them in a row (fewer graph launches,
    // so fewer gaps between them; attempts after the run has stopped do
    // nothing).  MM_STEPPER_GRAPHS=0 launches the kernels one by one instead;
    // MM_GRAPH_ATTEMPTS sets the attempts per long graph.
    hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int use_graphs = -1;
    int graph_attempts = 4;

    void launch_attempt() { launch_attempts(1); }

    void launch_attempts(long long q)
    {
        if (use_graphs < 0) {
            const char *e = std::getenv("MM_STEPPER_GRAPHS");
            use_graphs = (e && std::atoi(e) == 0) ? 0 : 1;
            const char *ga = std::getenv("MM_GRAPH_ATTEMPTS");
            if (ga) graph_attempts = std::max(1, std::min(64, std::atoi(ga)));
        }
        if (!use_graphs || klog) {
            for (long long r = 0; r < q; ++r) enqueue_attempt();
            return;
        }
        while (q > 0) {
            // whole long graphs, then single attempts for the rest
            const int w = (graph_attempts > 1 && q >= graph_attempts) ? 1 : 0;
            const int na = w ? graph_attempts : 1;
            const int g = (tab.fsal && have_fsal) ? 1 : 0;
            if (!gexec[g][w]) {
                hipGraph_t graph = nullptr;
                MM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    for (int r = 0; r < na; ++r) enqueue_attempt();
                } catch (...) {
                    (void)hipStreamEndCapture(s, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                MM_HIP(hipStreamEndCapture(s, &graph));
                const hipError_t e = hipGraphInstantiate(&gexec[g][w], graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                MM_HIP(e);
            } else if (tab.fsal) {
                have_fsal = true;  // what enqueue_attempt records for the next attempt
            }
            MM_HIP(hipGraphLaunch(gexec[g][w], s));
            q -= na;
        }
    }

    // Zeeman field from a host function (AlternatingField with H_ext_fn,
    // alternating_field.py:60-75): evaluated at every evaluation time and
    // uploaded to fn_tab before the launch that uses it
    mm_field_fn_t ffn = nullptr;
    void *fuser = nullptr;
    double *fn_tab = nullptr;
    double *fn_host = nullptr;  // pinned staging rows, rewritten only with the stream idle

    MmShape launch_shape() const override
    {
        MmShape r = {0, 0, 0, 0, 0};
        if (conv) {
            r.lpw_x = conv->ln.lpw_x;
            r.tk_y = conv->ln.tk_y;
            r.tk_z = conv->ln.tk_z;
        }
        if (conv && !split_x) {
            r.rpw = rpw;
            r.stage_threads = threads;
        }
        return r;
    }

    void set_field_fn(mm_field_fn_t fn, void *user) override
    {
        // alternating_field.py:68-69: offs, amp, freq, phase must be zero
        // when a function is given
        MM_REQUIRE(!fn || !d.use_zeeman, "a Zeeman field function excludes h_ext and the alternating field parameters");
        MM_HIP(hipStreamSynchronize(s));
        ffn = fn;
        fuser = user;
        sys.has_fn = fn ? 1 : 0;
        sys.fn_tab = fn_tab;
        sys.has_zee = fn ? 1 : d.use_zeeman;
        MM_HIP(hipMemcpy(sysd, &sys, sizeof(SysDev<T>), hipMemcpyHostToDevice));
    }

    // rows slot0 .. slot0+nslots-1 of fn_host to the device table, ordered
    // on the stream before the launches that read them (the caller has
    // synchronised the stream before filling fn_host)
    void upload_fields(int slot0, int nslots)
    {
        MM_HIP(hipMemcpyAsync(fn_tab + 3 * slot0, fn_host + 3 * slot0, sizeof(double) * 3 * nslots,
                              hipMemcpyHostToDevice, s));
    }

    // the stage times of the next attempt, as the stage kernel forms them
    // (t + a_s h, runge_kutta.py:115); the stream is idle
    void eval_stage_fields(const StepCtl &c)
    {
        const int ns = tab.num_steps;
        for (int st = 0; st < ns; ++st) ffn(c.t + htab[st].tnode * c.h, fn_host + 3 * st, fuser);
        upload_fields(0, ns);
    }

    // dM/dt and energies: the current time
    void eval_field_at_t(int slot)
    {
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        ffn(c.t, fn_host + 3 * slot, fuser);
        upload_fields(slot, 1);
    }

    StepCtl cached;

    void arm(long long steps) override
    {
        MM_REQUIRE(!ffn, "a Zeeman field function is not supported in ensembles");
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&cached, ctl, sizeof(cached), hipMemcpyDeviceToHost));
        cached.target = cached.accepted + std::max(0ll, steps);
        cached.stop = steps > 0 ? 0 : 1;
        cached.tmax_on = 0;
        cached.tstop_on = 0;
        MM_HIP(hipMemcpy(ctl, &cached, sizeof(cached), hipMemcpyHostToDevice));
    }

    bool pending() const override { return !cached.stop; }

    void enqueue_some(long long batch) override
    {
        const long long q = std::max(1ll, std::min(batch, cached.target - cached.accepted));
        launch_attempts(q);
    }

    void refresh() override { MM_HIP(hipMemcpy(&cached, ctl, sizeof(cached), hipMemcpyDeviceToHost)); }

    void run(long long steps) override
    {
        if (steps <= 0) return;
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        c.target = c.accepted + steps;
        c.stop = 0;
        c.tmax_on = 0;
        c.tstop_on = 0;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
        drive(c, 1ll << 40);
    }

    // accepted steps until t reaches t_end (last step clipped), a step ends
    // at t >= t_stop (not clipped; t_stop = inf: no such stop) or max_steps
    void run_until(double t_end, long long max_steps, double t_stop) override
    {
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        const bool ts = std::isfinite(t_stop);
        c.target = c.accepted + max_steps;
        c.stop = (max_steps <= 0 || c.t >= t_end || (ts && c.t >= t_stop)) ? 1 : 0;
        c.tmax_on = 1;
        c.t_max = t_end;
        c.forced = 0;
        c.tstop_on = ts ? 1 : 0;
        c.t_stop = ts ? t_stop : 0.0;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
        drive(c, 64);
        c.tmax_on = 0;
        c.tstop_on = 0;
        c.forced = 0;
        c.stop = 0;
        c.target = c.accepted;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
    }

    void sync() override { MM_HIP(hipStreamSynchronize(s)); }

    void stats(double *s5) override
    {
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        s5[0] = c.t;
        s5[1] = c.h_last;
        s5[2] = c.h;
        s5[3] = (double)c.attempts;
        s5[4] = (double)c.rejections;
    }

    long long step_count() override
    {
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        return c.accepted;
    }

    // state.t and the evolver's next trial step (runge_kutta.py:55-58,81-82)
    void set_time(double t, double h_next) override
    {
        StepCtl c;
        MM_HIP(hipStreamSynchronize(s));
        MM_HIP(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
        c.t = t;
        if (h_next > 0.0) c.h = h_next;
        MM_HIP(hipMemcpy(ctl, &c, sizeof(c), hipMemcpyHostToDevice));
    }

    void average(double *a3) override
    {
        const int grid = (int)std::max<size_t>(1, std::min<size_t>(MM_RED_BLOCKS, (n + 255) / 256));
        hipLaunchKernelGGL(k_sum3_sel<T>, dim3(grid), dim3(MM_RED_THREADS), 0, s, n, yb[0], yb[1], ctl, part);
        MM_HIP(hipGetLastError());
        std::vector<double> h((size_t)6 * MM_RED_BLOCKS);
        MM_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 6 * MM_RED_BLOCKS, hipMemcpyDeviceToHost, s));
        MM_HIP(hipStreamSynchronize(s));
        for (int c = 0; c < 3; ++c) {
            Dd acc = {0.0, 0.0};
            for (int b = 0; b < grid; ++b) {
                const Dd o = {h[(size_t)(2 * c) * MM_RED_BLOCKS + b], h[(size_t)(2 * c + 1) * MM_RED_BLOCKS + b]};
                acc = dd_merge(acc, o);
            }
            a3[c] = dd_value(acc) / (double)n;
        }
    }
};

struct mm_stepper {
    StepperBase *impl;
};

static StepperBase *make_stepper(const mm_system_desc_t &d)
{
    MM_REQUIRE(d.precision == MM_FP64 || d.precision == MM_FP32, "bad precision");
    if (d.precision == MM_FP64) {
        auto *p = new Stepper<double>();
        try {
            p->init(d);
        } catch (...) {
            delete p;
            throw;
        }
        return p;
    }
    auto *p = new Stepper<float>();
    try {
        p->init(d);
    } catch (...) {
        delete p;
        throw;
    }
    return p;
}

extern "C" int mm_stepper_create(const mm_system_desc_t *desc, mm_stepper_t **out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(desc && out, "null argument");
    mm_stepper *st = new mm_stepper();
    try {
        st->impl = make_stepper(*desc);
    } catch (...) {
        delete st;
        throw;
    }
    *out = st;
    MM_CAPI_END
}

extern "C" int mm_stepper_launch_shape(mm_stepper_t *st, mm_launch_shape_t *out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && out, "null argument");
    const MmShape r = st->impl->launch_shape();
    out->lpw_x = r.lpw_x;
    out->tk_y = r.tk_y;
    out->tk_z = r.tk_z;
    out->rpw = r.rpw;
    out->stage_threads = r.stage_threads;
    MM_CAPI_END
}

// Milliseconds per accepted step of a stepper for d built with launch shape
// cand (its resolved shape into *got), from a fixed pseudo-random state.
// Launch shapes change which thread computes what, not the arithmetic, so
// every candidate takes the same steps.
static double time_shape(const mm_system_desc_t &d, const MmShape &cand, int steps, MmShape *got)
{
    mm_shape_set_candidate(&cand);
    std::unique_ptr<StepperBase> st;
    try {
        st.reset(make_stepper(d));
    } catch (...) {
        mm_shape_set_candidate(nullptr);
        throw;
    }
    mm_shape_set_candidate(nullptr);
    *got = st->launch_shape();
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    std::vector<double> M(3 * n);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto uni = [&x]() {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    for (size_t i = 0; i < n; ++i) {
        double v[3], q = 0.0;
        do {
            for (int c = 0; c < 3; ++c) v[c] = uni();
            q = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        } while (q < 1e-4 || q > 1.0);
        const double ms = d.Ms_host ? d.Ms_host[i] : 0.0;
        for (int c = 0; c < 3; ++c) M[(size_t)c * n + i] = ms * v[c] / std::sqrt(q);
    }
    st->set_m(M.data());
    st->run(2);
    st->sync();
    const auto t0 = std::chrono::steady_clock::now();
    st->run(steps);
    st->sync();
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
}

extern "C" int mm_autotune(const mm_system_desc_t *desc, int steps, mm_launch_shape_t *best, double *ms_per_step)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(desc && best, "null argument");
    MM_REQUIRE(steps >= 1, "steps must be positive");
    MM_REQUIRE(desc->Ms_host && desc->alpha_host, "Ms and alpha are required");
    // the built-in choice first, then one knob at a time around the best so far
    MmShape cur;
    double cur_ms = time_shape(*desc, MmShape{0, 0, 0, 0, 0}, steps, &cur);
    std::vector<MmShape> seen = {cur};
    auto same = [](const MmShape &a, const MmShape &b) {
        return a.lpw_x == b.lpw_x && a.tk_y == b.tk_y && a.tk_z == b.tk_z && a.rpw == b.rpw &&
               a.stage_threads == b.stage_threads;
    };
    auto consider = [&](MmShape cand) {
        for (const MmShape &q : seen)
            if (same(q, cand)) return;
        MmShape got;
        double ms;
        try {
            ms = time_shape(*desc, cand, steps, &got);
        } catch (const mm_error &) {
            return;  // a shape this system cannot launch (the built-in one could)
        }
        for (const MmShape &q : seen)
            if (same(q, got)) return;  // clamped to a shape already measured
        seen.push_back(got);
        if (ms < 0.99 * cur_ms) {
            cur = got;
            cur_ms = ms;
        }
    };
    if (cur.lpw_x > 0) {
        const int base = cur.lpw_x;
        for (int v : {std::max(1, base / 2), base * 2}) {
            MmShape c = cur;
            c.lpw_x = v;
            consider(c);
        }
        for (int v : {4, 8, 16}) {
            MmShape c = cur;
            c.tk_y = v;
            consider(c);
        }
        for (int v : {8, 16, 32}) {
            MmShape c = cur;
            c.tk_z = v;
            consider(c);
        }
    }
    if (cur.rpw > 0) {
        for (int v : {2, 4}) {
            MmShape c = cur;
            c.rpw = v;
            c.stage_threads = 0;
            consider(c);
        }
        const int base = cur.stage_threads;
        for (int v : {base / 2, base * 2}) {
            MmShape c = cur;
            c.stage_threads = std::min(1024, std::max(64, v));
            consider(c);
        }
    }
    mm_wisdom_put(desc->nx, desc->ny, desc->nz, desc->precision == MM_FP32 ? 4 : 8, cur);
    best->lpw_x = cur.lpw_x;
    best->tk_y = cur.tk_y;
    best->tk_z = cur.tk_z;
    best->rpw = cur.rpw;
    best->stage_threads = cur.stage_threads;
    if (ms_per_step) *ms_per_step = cur_ms;
    MM_CAPI_END
}

extern "C" int mm_stepper_set_m(mm_stepper_t *st, const double *M_host)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && M_host, "null argument");
    st->impl->set_m(M_host);
    MM_CAPI_END
}

extern "C" int mm_stepper_get_m(mm_stepper_t *st, double *M_host)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && M_host, "null argument");
    st->impl->get_m(M_host);
    MM_CAPI_END
}

extern "C" int mm_stepper_dmdt(mm_stepper_t *st, double *dMdt_host, double *Htot_host)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st, "null argument");
    st->impl->dmdt(dMdt_host, Htot_host);
    MM_CAPI_END
}

extern "C" int mm_stepper_run(mm_stepper_t *st, int steps)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && steps >= 0, "bad argument");
    st->impl->run(steps);
    MM_CAPI_END
}

extern "C" int mm_stepper_run_until(mm_stepper_t *st, double t_end, long long max_steps)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st, "null argument");
    st->impl->run_until(t_end, max_steps, INFINITY);
    MM_CAPI_END
}

extern "C" int mm_stepper_run_until_stop(mm_stepper_t *st, double t_end, double t_stop, long long max_steps)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st, "null argument");
    MM_REQUIRE(!std::isnan(t_stop) && !std::isnan(t_end), "NaN time");
    st->impl->run_until(t_end, max_steps, t_stop);
    MM_CAPI_END
}

extern "C" int mm_stepper_sync(mm_stepper_t *st)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st, "null argument");
    st->impl->sync();
    MM_CAPI_END
}

extern "C" int mm_stepper_stats(mm_stepper_t *st, double *stats5)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && stats5, "null argument");
    st->impl->stats(stats5);
    MM_CAPI_END
}

extern "C" int mm_stepper_step_count(mm_stepper_t *st, long long *steps)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && steps, "null argument");
    *steps = st->impl->step_count();
    MM_CAPI_END
}

extern "C" int mm_stepper_set_time(mm_stepper_t *st, double t, double h_next)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && std::isfinite(t) && std::isfinite(h_next) && h_next >= 0.0, "bad argument");
    st->impl->set_time(t, h_next);
    MM_CAPI_END
}

extern "C" int mm_stepper_average_m(mm_stepper_t *st, double *avg3)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && avg3, "null argument");
    st->impl->average(avg3);
    MM_CAPI_END
}

extern "C" void *mm_stepper_stream(mm_stepper_t *st) { return st ? (void *)st->impl->stream() : nullptr; }

extern "C" int mm_stepper_kernel_profile(mm_stepper_t *st, int attempts, double *out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && out && attempts > 0, "bad argument");
    st->impl->profile(attempts, out);
    MM_CAPI_END
}

extern "C" int mm_stepper_set_field_fn(mm_stepper_t *st, mm_field_fn_t fn, void *user)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st, "null argument");
    st->impl->set_field_fn(fn, user);
    MM_CAPI_END
}

extern "C" int mm_stepper_energy(mm_stepper_t *st, double *E5)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && E5, "null argument");
    st->impl->energy(E5);
    MM_CAPI_END
}

extern "C" int mm_stepper_relax(mm_stepper_t *st, int criterion, double threshold, int check_every,
                                long long max_steps, double *info3)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(st && info3, "null argument");
    st->impl->relax(criterion, threshold, check_every, max_steps, info3);
    MM_CAPI_END
}

// ---------------------------------------------------------------------------
// ensembles: independent systems stepped together, spread over a few shared
// streams so that the small launches of different systems overlap; every
// system keeps its own control block (own h, accept/reject, t)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// batched ensembles: every pass of an attempt is one launch for all systems
// ---------------------------------------------------------------------------
// Per-system records of the batched stage and controller launches (the conv
// passes' records are in mm_conv.h).  The global workgroup index in XCD order
// picks the system through the pass's tile table, as in the conv passes.
template <typename T> struct BStageDesc {
    const SysDev<T> *S;
    const StageDev<T> *G0;  // the system's stage records; the launch picks stage st
    StepCtl *ctl;
    const C2<T> *S1;
    int Hx, Hp, rpw, nb, tbase;
    FftPlanDev p;
};

struct BCtlDesc {
    StepCtl *ctl;
    const double *mpart;
    int nb, order, Ms_uniform, mode;
    double eps_abs, eps_rel, Ms_u;
};

template <typename T, bool STT, bool GEN, bool P1S = false>
__global__ void kb_stage(const BStageDesc<T> *__restrict__ D, const int *__restrict__ tsys, int st)
{
    const int gi = block_tile(1);
    const BStageDesc<T> &d = D[tsys[gi]];
    stage_body<T, false, 0, STT, GEN, P1S>(gfield(d.S), gfield(d.G0) + st, gfield(d.ctl), gfield(d.S1), d.p, d.Hx, d.Hp,
                                      d.rpw, 1, nullptr, nullptr, gi - d.tbase, d.nb, false);
}

// one workgroup per system
__global__ void __launch_bounds__(256) kb_controller(const BCtlDesc *__restrict__ D)
{
    const BCtlDesc &d = D[blockIdx.x];
    controller_body(gfield(d.ctl), gfield(d.mpart), d.nb, d.order, d.eps_abs, d.eps_rel, d.Ms_uniform, d.Ms_u,
                    d.mode);
}

// out[0] = number of systems whose run has not stopped
__global__ void kb_running(StepCtl *const *__restrict__ ctls, int n, int *__restrict__ out)
{
    int c = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) c += gfield(ctls[i])->stop ? 0 : 1;
    __shared__ int sred[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sred[w];
        out[0] = t;
    }
}

struct BatchRunnerBase {
    virtual ~BatchRunnerBase() {}
    virtual void run(long long steps) = 0;
};

// Systems of one precision and one method without FSAL, all with the demag
// field in the fused stage (no split x pass).  An attempt of every system is
// one x forward launch, per stage the y, z (or 2-D) and stage launches (one
// per physics class present), and one controller launch: 2 + 4 s launches
// for the whole ensemble (s stages), recorded once as a HIP graph.  Systems
// that reached their step target skip the passes.
template <typename T> struct BatchRunner : BatchRunnerBase {
    std::vector<Stepper<T> *> sys;  // not owned
    hipStream_t s = nullptr;
    int ns = 0;
    bool any3d = false, any2d = false;
    BatchPass bx, by, bz, b2;
    struct StageGroup {
        BatchPass bp;
        bool stt, gen;
    };
    std::vector<StageGroup> stg;
    BCtlDesc *cdesc = nullptr;
    StepCtl **ctls = nullptr;
    int *running = nullptr;
    int *running_host = nullptr;
    std::vector<void *> allocs;
    hipGraphExec_t gexec = nullptr;

    static int round64(int v) { return std::max(64, (v + 63) / 64 * 64); }

    template <typename D> void upload_pass(BatchPass &bp, const std::vector<D> &desc, const std::vector<int> &tsys)
    {
        void *dd = nullptr, *dt = nullptr;
        MM_HIP(hipMalloc(&dd, sizeof(D) * std::max<size_t>(1, desc.size())));
        allocs.push_back(dd);
        MM_HIP(hipMalloc(&dt, sizeof(int) * std::max<size_t>(1, tsys.size())));
        allocs.push_back(dt);
        MM_HIP(hipMemcpy(dd, desc.data(), sizeof(D) * desc.size(), hipMemcpyHostToDevice));
        MM_HIP(hipMemcpy(dt, tsys.data(), sizeof(int) * tsys.size(), hipMemcpyHostToDevice));
        bp.desc = dd;
        bp.tsys = (int *)dt;
        bp.ntiles = (int)tsys.size();
        MM_REQUIRE(bp.lds <= (size_t)mm_max_lds_bytes(), "batched pass needs more LDS than the device has");
        MM_REQUIRE(bp.threads <= 1024, "batched pass needs more than 1024 threads per workgroup");
    }

    BatchRunner(const std::vector<Stepper<T> *> &systems)
    {
        sys = systems;
        MM_REQUIRE(!sys.empty(), "empty batch");
        MM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        ns = sys[0]->tab.num_steps;
        const int n = (int)sys.size();
        // threads per workgroup of each pass: the largest a system's own
        // launch takes
        int tx = 64, ty = 64, tz = 64, t2 = 64, tsg = 64;
        for (Stepper<T> *q : sys) {
            const ConvImpl<T> &c = *q->conv;
            MM_REQUIRE(q->tab.num_steps == ns && !q->tab.fsal && q->d.method == sys[0]->d.method,
                       "batched systems share a method without FSAL");
            MM_REQUIRE(!q->split_x, "batched systems run the fused x transform");
            tx = std::max(tx, c.ln.lpw_x * c.px.tline);
            if (c.g.is2d) {
                any2d = true;
                t2 = std::max(t2, 3 * c.ln.tk_y * c.py.tline);
            } else {
                any3d = true;
                ty = std::max(ty, c.ln.tk_y * c.py.tline);
                tz = std::max(tz, 3 * c.ln.tk_z * c.pz.tline);
            }
            tsg = std::max(tsg, q->threads);
        }
        bx.threads = round64(tx);
        by.threads = round64(ty);
        bz.threads = round64(tz);
        b2.threads = round64(t2);
        // per-system records and tile tables
        std::vector<BxDesc<T>> dx;
        std::vector<ByDesc<T>> dy;
        std::vector<BzDesc<T>> dz;
        std::vector<B2Desc<T>> d2;
        std::vector<int> tx_s, ty_s, tz_s, t2_s;
        auto add_tiles = [](std::vector<int> &t, int count, int sysi) {
            const int base = (int)t.size();
            t.insert(t.end(), (size_t)count, sysi);
            return base;
        };
        for (int i = 0; i < n; ++i) {
            Stepper<T> *q = sys[i];
            const ConvImpl<T> &c = *q->conv;
            const int *sel = (const int *)((char *)q->ctl + offsetof(StepCtl, cur));
            const int *stop = (const int *)((char *)q->ctl + offsetof(StepCtl, stop));
            size_t lds = 0;
            bool o32 = true;
            BxDesc<T> a;
            int t = c.batch_x(a, bx.threads, lds, q->yb[0], q->yb[1], sel, stop);
            a.tbase = add_tiles(tx_s, t, i);
            bx.lds = std::max(bx.lds, lds);
            dx.push_back(a);
            if (c.g.is2d) {
                B2Desc<T> e;
                t = c.batch_2d(e, b2.threads, lds, stop);
                e.tbase = add_tiles(t2_s, t, (int)d2.size());
                b2.lds = std::max(b2.lds, lds);
                d2.push_back(e);
            } else {
                ByDesc<T> e;
                t = c.batch_y(e, by.threads, lds, o32, stop);
                e.tbase = add_tiles(ty_s, t, (int)dy.size());
                by.lds = std::max(by.lds, lds);
                by.o32 = by.o32 && o32;
                dy.push_back(e);
                BzDesc<T> f;
                t = c.batch_z(f, bz.threads, lds, o32, stop);
                f.tbase = add_tiles(tz_s, t, (int)dz.size());
                bz.lds = std::max(bz.lds, lds);
                bz.o32 = bz.o32 && o32;
                dz.push_back(f);
            }
        }
        upload_pass(bx, dx, tx_s);
        if (any3d) {
            upload_pass(by, dy, ty_s);
            upload_pass(bz, dz, tz_s);
        }
        if (any2d) upload_pass(b2, d2, t2_s);
        // stage launches, one per physics class present
        for (int cls = 0; cls < 4; ++cls) {
            const bool stt = cls & 1, gen = (cls >> 1) & 1;
            std::vector<BStageDesc<T>> ds;
            std::vector<int> ts;
            StageGroup gr;
            gr.stt = stt;
            gr.gen = gen;
            gr.bp.threads = round64(tsg);
            for (int i = 0; i < n; ++i) {
                Stepper<T> *q = sys[i];
                bool a, b;
                q->physics_class(a, b);
                if (a != stt || b != gen) continue;
                BStageDesc<T> e;
                e.S = q->sysd;
                e.G0 = q->gtab;
                e.ctl = q->ctl;
                e.S1 = q->conv->S1;
                e.Hx = q->conv->g.Hx;
                e.Hp = q->conv->g.Hp;
                e.rpw = q->rpw;
                e.nb = q->grid_rows;
                e.p = q->conv->px.template dev<T>();
                e.tbase = add_tiles(ts, q->grid_rows, (int)ds.size());
                gr.bp.lds = std::max(gr.bp.lds, q->stage_lds);
                ds.push_back(e);
            }
            if (ds.empty()) continue;
            upload_pass(gr.bp, ds, ts);
            stg.push_back(gr);
        }
        // controllers and the running count
        std::vector<BCtlDesc> dc;
        std::vector<StepCtl *> cp;
        for (Stepper<T> *q : sys) {
            BCtlDesc e;
            e.ctl = q->ctl;
            e.mpart = q->mpart;
            e.nb = q->grid_rows;
            e.order = q->tab.order;
            e.Ms_uniform = q->sys.Ms_uniform;
            e.mode = q->d.method == MM_RK4 ? 0 : (q->d.method == MM_EULER ? 2 : 1);
            e.eps_abs = q->d.eps_abs;
            e.eps_rel = q->d.eps_rel;
            e.Ms_u = q->sys.Ms_u;
            dc.push_back(e);
            cp.push_back(q->ctl);
        }
        MM_HIP(hipMalloc(&cdesc, sizeof(BCtlDesc) * dc.size()));
        allocs.push_back(cdesc);
        MM_HIP(hipMemcpy(cdesc, dc.data(), sizeof(BCtlDesc) * dc.size(), hipMemcpyHostToDevice));
        MM_HIP(hipMalloc(&ctls, sizeof(StepCtl *) * cp.size()));
        allocs.push_back(ctls);
        MM_HIP(hipMemcpy(ctls, cp.data(), sizeof(StepCtl *) * cp.size(), hipMemcpyHostToDevice));
        MM_HIP(hipMalloc(&running, sizeof(int)));
        allocs.push_back(running);
        MM_HIP(hipHostMalloc((void **)&running_host, sizeof(int), hipHostMallocDefault));
    }

    ~BatchRunner() override
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (void *p : allocs) (void)hipFree(p);
        if (running_host) (void)hipHostFree(running_host);
        if (s) (void)hipStreamDestroy(s);
    }

    void stage_launch(const StageGroup &gr, int st)
    {
        const BStageDesc<T> *D = (const BStageDesc<T> *)gr.bp.desc;
        const dim3 grid(gr.bp.ntiles), block(gr.bp.threads);
        if (gr.stt) {
            if (gr.gen) hipLaunchKernelGGL((kb_stage<T, true, true>), grid, block, gr.bp.lds, s, D, gr.bp.tsys, st);
            else hipLaunchKernelGGL((kb_stage<T, true, false>), grid, block, gr.bp.lds, s, D, gr.bp.tsys, st);
        } else {
            // inner stages of a six-stage tableau (summed combination, no
            // FSAL in a batch): the kernel compiled for them
            const bool p1s = ns == 6 && st < ns - 1;
            if (gr.gen)
                hipLaunchKernelGGL((kb_stage<T, false, true>), grid, block, gr.bp.lds, s, D, gr.bp.tsys, st);
            else if (p1s)
                hipLaunchKernelGGL((kb_stage<T, false, false, true>), grid, block, gr.bp.lds, s, D, gr.bp.tsys, st);
            else
                hipLaunchKernelGGL((kb_stage<T, false, false>), grid, block, gr.bp.lds, s, D, gr.bp.tsys, st);
        }
        MM_HIP(hipGetLastError());
    }

    // one attempt of every system (the records of stage st < ns - 1 also
    // transform the next stage input along x, so only stage 0 has an x pass)
    void enqueue_attempt()
    {
        mm_batch_xfwd<T>(bx, s);
        for (int st = 0; st < ns; ++st) {
            if (any3d) {
                mm_batch_yfwd<T>(by, s);
                mm_batch_zmi<T>(bz, s);
                mm_batch_yinv<T>(by, s);
            }
            if (any2d) mm_batch_ymi2d<T>(b2, s);
            for (const StageGroup &gr : stg) stage_launch(gr, st);
        }
        hipLaunchKernelGGL(kb_controller, dim3((unsigned)sys.size()), dim3(256), 0, s, (const BCtlDesc *)cdesc);
        MM_HIP(hipGetLastError());
    }

    void launch_attempt()
    {
        if (!gexec) {
            hipGraph_t graph = nullptr;
            MM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_attempt();
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            MM_HIP(hipStreamEndCapture(s, &graph));
            const hipError_t e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            MM_HIP(e);
        }
        MM_HIP(hipGraphLaunch(gexec, s));
    }

    int count_running()
    {
        hipLaunchKernelGGL(kb_running, dim3(1), dim3(256), 0, s, (StepCtl *const *)ctls, (int)sys.size(), running);
        MM_HIP(hipGetLastError());
        MM_HIP(hipMemcpyAsync(running_host, running, sizeof(int), hipMemcpyDeviceToHost, s));
        MM_HIP(hipStreamSynchronize(s));
        return *running_host;
    }

    // `steps` accepted steps of every system (the systems were armed)
    void run(long long steps) override
    {
        if (steps <= 0) return;
        for (Stepper<T> *q : sys) MM_HIP(hipStreamSynchronize(q->s));
        // attempts between two looks at the running count: every system needs
        // at least `steps`, then the rejected ones
        long long batch = steps;
        while (count_running() > 0) {
            for (long long r = 0; r < batch; ++r) launch_attempt();
            batch = std::max(1ll, std::min(8ll, steps / 8));
        }
    }
};

struct mm_batch {
    std::vector<StepperBase *> sims;
    std::vector<hipStream_t> streams;
    std::vector<BatchRunnerBase *> runners;
    std::vector<char> in_runner;  // system i is stepped by a runner
};

static void batch_free(mm_batch *b)
{
    for (BatchRunnerBase *r : b->runners) delete r;
    for (StepperBase *p : b->sims) delete p;
    for (hipStream_t st : b->streams) (void)hipStreamDestroy(st);
    delete b;
}

// batched runners for the systems that can take them: per precision and
// method, systems with the demag field in the fused stage and no FSAL
template <typename T> static void batch_group(mm_batch *b, int precision)
{
    std::map<int, std::vector<int>> by_method;
    for (int i = 0; i < (int)b->sims.size(); ++i) {
        auto *q = dynamic_cast<Stepper<T> *>(b->sims[i]);
        if (!q || q->d.precision != precision) continue;
        if (!q->conv || q->split_x || q->tab.fsal) continue;
        by_method[q->d.method].push_back(i);
    }
    for (auto &kv : by_method) {
        std::vector<Stepper<T> *> v;
        for (int i : kv.second) v.push_back(dynamic_cast<Stepper<T> *>(b->sims[i]));
        b->runners.push_back(new BatchRunner<T>(v));
        for (int i : kv.second) b->in_runner[i] = 1;
    }
}

extern "C" int mm_batch_create(const mm_system_desc_t *descs, int count, int nstreams, mm_batch_t **out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(descs && out && count > 0, "bad argument");
    mm_batch *b = new mm_batch();
    try {
        const int ns = std::max(1, std::min(nstreams > 0 ? nstreams : 4, count));
        for (int q = 0; q < ns; ++q) {
            hipStream_t st = nullptr;
            MM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            b->streams.push_back(st);
        }
        for (int i = 0; i < count; ++i) {
            const mm_system_desc_t &d = descs[i];
            MM_REQUIRE(d.precision == MM_FP64 || d.precision == MM_FP32, "bad precision");
            hipStream_t st = b->streams[i % ns];
            if (d.precision == MM_FP64) {
                auto *p = new Stepper<double>();
                b->sims.push_back(p);
                p->init(d, st);
            } else {
                auto *p = new Stepper<float>();
                b->sims.push_back(p);
                p->init(d, st);
            }
        }
        b->in_runner.assign(b->sims.size(), 0);
        // MM_BATCH_LAUNCH=0: every system on its own graphs over the streams
        const char *e = std::getenv("MM_BATCH_LAUNCH");
        if (!(e && std::atoi(e) == 0)) {
            batch_group<double>(b, MM_FP64);
            batch_group<float>(b, MM_FP32);
        }
    } catch (...) {
        batch_free(b);
        throw;
    }
    *out = b;
    MM_CAPI_END
}

static StepperBase *batch_at(mm_batch *b, int index)
{
    MM_REQUIRE(b && index >= 0 && index < (int)b->sims.size(), "bad ensemble index");
    return b->sims[index];
}

extern "C" int mm_batch_set_m(mm_batch_t *b, int index, const double *M_host)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(M_host, "null argument");
    batch_at(b, index)->set_m(M_host);
    MM_CAPI_END
}

extern "C" int mm_batch_get_m(mm_batch_t *b, int index, double *M_host)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(M_host, "null argument");
    batch_at(b, index)->get_m(M_host);
    MM_CAPI_END
}

extern "C" int mm_batch_stats(mm_batch_t *b, int index, double *stats5)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(stats5, "null argument");
    batch_at(b, index)->stats(stats5);
    MM_CAPI_END
}

extern "C" int mm_batch_average_m(mm_batch_t *b, int index, double *avg3)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(avg3, "null argument");
    batch_at(b, index)->average(avg3);
    MM_CAPI_END
}

extern "C" int mm_batch_run(mm_batch_t *b, int steps)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(b && steps >= 0, "bad argument");
    for (StepperBase *p : b->sims) p->arm(steps);
    for (BatchRunnerBase *r : b->runners) r->run(steps);
    for (;;) {
        bool any = false;
        for (size_t i = 0; i < b->sims.size(); ++i)
            if (!b->in_runner[i] && b->sims[i]->pending()) {
                b->sims[i]->enqueue_some(1ll << 40);
                any = true;
            }
        if (!any) break;
        for (hipStream_t st : b->streams) MM_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < b->sims.size(); ++i)
            if (!b->in_runner[i]) b->sims[i]->refresh();
    }
    for (StepperBase *p : b->sims) p->refresh();
    MM_CAPI_END
}

extern "C" int mm_batch_destroy(mm_batch_t *b)
{
    MM_CAPI_BEGIN
    if (b) batch_free(b);
    MM_CAPI_END
}

extern "C" int mm_stepper_destroy(mm_stepper_t *st)
{
    MM_CAPI_BEGIN
    if (!st) return MM_OK;
    delete st->impl;
    delete st;
    MM_CAPI_END
}
