/*
 * mm_oracle.c -- TEST INFRASTRUCTURE ONLY (see mm_oracle.h).
 *
 * FP64 CPU restatement of the reference CPU path of MicroMagnum's hot path.
 * Each function names the reference file:line whose behaviour it restates.
 * Compiled with -O2 -ffp-contract=off (no FMA contraction, IEEE order of
 * operations as written).  Optional OpenMP threading over independent lines
 * and cells; every reduction stays sequential so results do not depend on
 * the thread count.
 *
 * What pins each part (the tests under tests/ hold the checks):
 *   FFT        : textbook DFT by brute force, numpy.fft (pocketfft).
 *   conv       : direct sum orc_conv_simple on small grids, H1_stray.omf.
 *   exchange   : H1_exch.omf, uniform-M => 0 for every PBC combination.
 *   anisotropy : H1_unianiso.omf, closed-form energies.
 *   llge       : exact known-answer vector of llge_test.py.
 *   zhangli    : <dM> regression value of spintorque_test.py, zero gradient.
 *   RK         : tableau order conditions, convergence order on scalar ODEs.
 *   controller : fixed-point error ratios 0.85^(p+1) / 0.85^p under power-law
 *                errors, scale clamps, step bounds, forced accept, zero
 *                denominator; closed-loop step size against the acceptance
 *                boundary on uniform precession (every method).
 *   FSAL reuse : the whole retry sequence of dp54 on one spin, emulated
 *                with the published coefficients and the stale k_0.
 *   normalize  : mode 0 against numpy float32 arithmetic, bit for bit.
 *   relax      : stop step against the closed-form damped single spin.
 */
#include "mm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* constants: src/magneto/mmm/constants.h:26-40                              */
/* ------------------------------------------------------------------------ */
static const long double ORC_PI_L = 3.141592653589793238462643383279502884197169399L;
#define ORC_MU0 ((double)(4.0L * ORC_PI_L * (long double)1e-7))
static const double ORC_MU_BOHR = 9.2740154e-24;
static const double ORC_E_CHARGE = 1.602176487e-19;
static const double ORC_GAMMA = 2.211e5; /* src/magnum/micromagnetics/constants.py:33 */

double orc_mu0(void) { return ORC_MU0; }

/* number of OpenMP threads used by the parallel loops (results do not depend on it) */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* 1-D complex DFT: mixed-radix Cooley-Tukey (decimation in time), radices  */
/* 4, 2, 3, 5 and a direct O(p^2) DFT for any other prime factor.  Stands   */
/* in for FFTW's unnormalised transforms (Transformer_CPU.cpp:29-175):      */
/* forward e^{-i}, inverse e^{+i}.                                           */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

typedef struct {
    int n;
    cplx *w; /* w[j] = exp(-2 pi i j / n), rounded from long double */
} fft_plan;

static fft_plan *plan_create(int n)
{
    fft_plan *p = (fft_plan *)malloc(sizeof(fft_plan));
    p->n = n;
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) {
        long double ang = 2.0L * ORC_PI_L * (long double)j / (long double)n;
        p->w[j].re = (double)cosl(ang);
        p->w[j].im = (double)(-sinl(ang));
    }
    return p;
}

static void plan_destroy(fft_plan *p)
{
    if (!p) return;
    free(p->w);
    free(p);
}

static int pick_factor(int n)
{
    if (n % 4 == 0 && n != 8) return 4; /* 8 = 2*4 keeps the radix-4 inner */
    if (n % 2 == 0) return 2;
    if (n % 3 == 0) return 3;
    if (n % 5 == 0) return 5;
    for (int f = 7; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

/* out[0..n) = DFT_n of in[0], in[istride], ... ; tw(j) = w[(j*(N/n)) mod N] */
static void fft_rec(const cplx *in, int istride, cplx *out, int n, const fft_plan *pl,
                    int sign, cplx *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    const int p = pick_factor(n);
    const int m = n / p;
    const int N = pl->n;
    const int wstep = N / n;
    for (int q = 0; q < p; ++q)
        fft_rec(in + (size_t)q * istride, istride * p, out + (size_t)q * m, m, pl, sign, tmp);
    for (int k = 0; k < m; ++k) {
        /* twiddled inputs t_q = Y_q[k] * W_n^{q k} */
        for (int q = 0; q < p; ++q) {
            const cplx y = out[(size_t)q * m + k];
            const long idx = ((long)q * k) % n * wstep;
            const double wr = pl->w[idx].re;
            const double wi = sign < 0 ? pl->w[idx].im : -pl->w[idx].im;
            tmp[q].re = y.re * wr - y.im * wi;
            tmp[q].im = y.re * wi + y.im * wr;
        }
        for (int s = 0; s < p; ++s) {
            double ar = 0.0, ai = 0.0;
            for (int q = 0; q < p; ++q) {
                const long idx = ((long)q * s % p) * m * wstep;
                const double wr = pl->w[idx].re;
                const double wi = sign < 0 ? pl->w[idx].im : -pl->w[idx].im;
                ar += tmp[q].re * wr - tmp[q].im * wi;
                ai += tmp[q].re * wi + tmp[q].im * wr;
            }
            tmp[p + s].re = ar;
            tmp[p + s].im = ai;
        }
        for (int s = 0; s < p; ++s) out[(size_t)s * m + k] = tmp[p + s];
    }
}

/* in-place transform of buf[0..n) using scratch of >= n + 2*maxfactor entries */
static void fft_exec(const fft_plan *pl, cplx *buf, cplx *scratch, int sign)
{
    const int n = pl->n;
    if (n <= 1) return;
    cplx *out = scratch;
    cplx *tmp = scratch + n;
    fft_rec(buf, 1, out, n, pl, sign, tmp);
    memcpy(buf, out, sizeof(cplx) * (size_t)n);
}

static size_t fft_scratch_len(int n) { return (size_t)3 * (size_t)(n > 0 ? n : 1) + 16; }

void orc_fft(double *re_im, int n, int sign)
{
    fft_plan *pl = plan_create(n);
    cplx *scratch = (cplx *)malloc(sizeof(cplx) * fft_scratch_len(n));
    fft_exec(pl, (cplx *)re_im, scratch, sign);
    free(scratch);
    plan_destroy(pl);
}

/* ------------------------------------------------------------------------ */
/* FFT convolution                                                           */
/*   ctor / kernel: SymmetricMatrixVectorConvolution_FFT.cpp:31-65,          */
/*                  TensorFieldSetup.cpp:35-114 (full complex fftn, keep     */
/*                  x < Px/2+1, scale 1/(Px Py Pz))                          */
/*   execute      : MatrixVectorConvolution_FFT.cpp:171-240 pass order       */
/*                  pad -> x R2C (Ny*Nz rows) -> y (Nz*Hx lines, zero pad)   */
/*                  -> z (Hx*Py lines, zero pad) -> 3x3 multiply             */
/*                  (cpu_multiplication.cpp:56-91) -> z inverse -> keep      */
/*                  z<Nz -> y inverse -> keep y<Ny -> x C2R -> keep x<Nx.    */
/*   The transposes of the reference only move data; this restatement       */
/*   gathers each line directly, which gives the same arithmetic.           */
/* ------------------------------------------------------------------------ */
struct orc_conv {
    int nx, ny, nz, Px, Py, Pz, Hx;
    double *Nre, *Nim;  /* 6*Hx*Py*Pz, e + 6*(kx + Hx*(ky + Py*kz)) */
    fft_plan *plx, *ply, *plz;
    cplx *S[3];         /* Hx*Py*Pz, kx + Hx*(ky + Py*kz) */
};

#define SIDX(c, kx, ky, kz) ((size_t)(kx) + (size_t)(c)->Hx * ((size_t)(ky) + (size_t)(c)->Py * (size_t)(kz)))

orc_conv *orc_conv_create(const double *N, int nx, int ny, int nz, int Px, int Py, int Pz)
{
    orc_conv *c = (orc_conv *)calloc(1, sizeof(orc_conv));
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->Px = Px; c->Py = Py; c->Pz = Pz;
    c->Hx = Px / 2 + 1;
    c->plx = plan_create(Px);
    c->ply = plan_create(Py);
    c->plz = plan_create(Pz);
    const size_t nsp = (size_t)c->Hx * Py * Pz;
    for (int e = 0; e < 3; ++e) c->S[e] = (cplx *)malloc(sizeof(cplx) * nsp);
    c->Nre = (double *)malloc(sizeof(double) * 6 * nsp);
    c->Nim = (double *)malloc(sizeof(double) * 6 * nsp);

    /* full complex 3-D transform of each tensor component */
    const size_t ntot = (size_t)Px * Py * Pz;
    cplx *T = (cplx *)malloc(sizeof(cplx) * ntot);
    int maxP = Px > Py ? Px : Py; if (Pz > maxP) maxP = Pz;
    cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)maxP);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * fft_scratch_len(maxP));
    const double scale = 1.0 / ((double)Px * (double)Py * (double)Pz);
    for (int e = 0; e < 6; ++e) {
        for (size_t i = 0; i < ntot; ++i) { T[i].re = N[e + 6 * i]; T[i].im = 0.0; }
        for (int z = 0; z < Pz; ++z)
            for (int y = 0; y < Py; ++y)
                fft_exec(c->plx, T + (size_t)Px * (y + (size_t)Py * z), scr, -1);
        for (int z = 0; z < Pz; ++z)
            for (int x = 0; x < Px; ++x) {
                for (int y = 0; y < Py; ++y) line[y] = T[x + (size_t)Px * (y + (size_t)Py * z)];
                fft_exec(c->ply, line, scr, -1);
                for (int y = 0; y < Py; ++y) T[x + (size_t)Px * (y + (size_t)Py * z)] = line[y];
            }
        for (int y = 0; y < Py; ++y)
            for (int x = 0; x < Px; ++x) {
                for (int z = 0; z < Pz; ++z) line[z] = T[x + (size_t)Px * (y + (size_t)Py * z)];
                fft_exec(c->plz, line, scr, -1);
                for (int z = 0; z < Pz; ++z) T[x + (size_t)Px * (y + (size_t)Py * z)] = line[z];
            }
        for (int z = 0; z < Pz; ++z)
            for (int y = 0; y < Py; ++y)
                for (int x = 0; x < c->Hx; ++x) {
                    const cplx v = T[x + (size_t)Px * (y + (size_t)Py * z)];
                    const size_t o = e + 6 * SIDX(c, x, y, z);
                    c->Nre[o] = scale * v.re;
                    c->Nim[o] = scale * v.im;
                }
    }
    free(T); free(line); free(scr);
    return c;
}

void orc_conv_get_kernel(orc_conv *c, double *out_re, double *out_im)
{
    const size_t n = (size_t)6 * c->Hx * c->Py * c->Pz;
    memcpy(out_re, c->Nre, sizeof(double) * n);
    memcpy(out_im, c->Nim, sizeof(double) * n);
}

void orc_conv_destroy(orc_conv *c)
{
    if (!c) return;
    for (int e = 0; e < 3; ++e) free(c->S[e]);
    free(c->Nre); free(c->Nim);
    plan_destroy(c->plx); plan_destroy(c->ply); plan_destroy(c->plz);
    free(c);
}

/* res = a*b + c*d + e*f over complex numbers, cpu_multiplication.cpp:23-39 */
static inline void mul3(double *rr, double *ri,
                        double ar, double ai, double br, double bi,
                        double cr, double ci, double dr, double di,
                        double er, double ei, double fr, double fi)
{
    double r = ar * br - ai * bi, i = ai * br + ar * bi;
    r += cr * dr - ci * di; i += ci * dr + cr * di;
    r += er * fr - ei * fi; i += ei * fr + er * fi;
    *rr = r; *ri = i;
}

void orc_conv_execute(orc_conv *c, const double *M, double *H)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, Px = c->Px, Py = c->Py, Pz = c->Pz, Hx = c->Hx;
    const size_t n = (size_t)nx * ny * nz;
    int maxP = Px > Py ? Px : Py; if (Pz > maxP) maxP = Pz;

    for (int comp = 0; comp < 3; ++comp) {
        cplx *S = c->S[comp];
        const double *Mc = M + comp * n;
        /* pad + x R2C over Ny*Nz rows */
#pragma omp parallel
        {
            cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)maxP);
            cplx *scr = (cplx *)malloc(sizeof(cplx) * fft_scratch_len(maxP));
#pragma omp for collapse(2) schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y) {
                    for (int x = 0; x < Px; ++x) {
                        line[x].re = x < nx ? Mc[x + (size_t)nx * (y + (size_t)ny * z)] : 0.0;
                        line[x].im = 0.0;
                    }
                    fft_exec(c->plx, line, scr, -1);
                    for (int kx = 0; kx < Hx; ++kx) S[SIDX(c, kx, y, z)] = line[kx];
                }
            /* y forward over Nz*Hx lines, zero-padded y >= Ny */
#pragma omp for collapse(2) schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int kx = 0; kx < Hx; ++kx) {
                    for (int y = 0; y < Py; ++y) {
                        if (y < ny) line[y] = S[SIDX(c, kx, y, z)];
                        else { line[y].re = 0.0; line[y].im = 0.0; }
                    }
                    fft_exec(c->ply, line, scr, -1);
                    for (int y = 0; y < Py; ++y) S[SIDX(c, kx, y, z)] = line[y];
                }
            /* z forward over Hx*Py lines, zero-padded z >= Nz */
            if (Pz > 1) {
#pragma omp for collapse(2) schedule(static)
                for (int ky = 0; ky < Py; ++ky)
                    for (int kx = 0; kx < Hx; ++kx) {
                        for (int z = 0; z < Pz; ++z) {
                            if (z < nz) line[z] = S[SIDX(c, kx, ky, z)];
                            else { line[z].re = 0.0; line[z].im = 0.0; }
                        }
                        fft_exec(c->plz, line, scr, -1);
                        for (int z = 0; z < Pz; ++z) S[SIDX(c, kx, ky, z)] = line[z];
                    }
            }
            free(line); free(scr);
        }
    }

    /* 3x3 symmetric complex multiply at every Fourier point */
    {
        const size_t nsp = (size_t)Hx * Py * Pz;
        cplx *Sx = c->S[0], *Sy = c->S[1], *Sz = c->S[2];
#pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)nsp; ++q) {
            const double *Nr = c->Nre + 6 * (size_t)q, *Ni = c->Nim + 6 * (size_t)q;
            const double xr = Sx[q].re, xi = Sx[q].im;
            const double yr = Sy[q].re, yi = Sy[q].im;
            const double zr = Sz[q].re, zi = Sz[q].im;
            mul3(&Sx[q].re, &Sx[q].im, xr, xi, Nr[0], Ni[0], yr, yi, Nr[1], Ni[1], zr, zi, Nr[2], Ni[2]);
            mul3(&Sy[q].re, &Sy[q].im, xr, xi, Nr[1], Ni[1], yr, yi, Nr[3], Ni[3], zr, zi, Nr[4], Ni[4]);
            mul3(&Sz[q].re, &Sz[q].im, xr, xi, Nr[2], Ni[2], yr, yi, Nr[4], Ni[4], zr, zi, Nr[5], Ni[5]);
        }
    }

    for (int comp = 0; comp < 3; ++comp) {
        cplx *S = c->S[comp];
        double *Hc = H + comp * n;
#pragma omp parallel
        {
            cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)maxP);
            cplx *scr = (cplx *)malloc(sizeof(cplx) * fft_scratch_len(maxP));
            if (Pz > 1) {
#pragma omp for collapse(2) schedule(static)
                for (int ky = 0; ky < Py; ++ky)
                    for (int kx = 0; kx < Hx; ++kx) {
                        for (int z = 0; z < Pz; ++z) line[z] = S[SIDX(c, kx, ky, z)];
                        fft_exec(c->plz, line, scr, +1);
                        for (int z = 0; z < nz; ++z) S[SIDX(c, kx, ky, z)] = line[z];
                    }
            }
#pragma omp for collapse(2) schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int kx = 0; kx < Hx; ++kx) {
                    for (int y = 0; y < Py; ++y) line[y] = S[SIDX(c, kx, y, z)];
                    fft_exec(c->ply, line, scr, +1);
                    for (int y = 0; y < ny; ++y) S[SIDX(c, kx, y, z)] = line[y];
                }
            /* x C2R: Hermitian extension, complex inverse, real part */
#pragma omp for collapse(2) schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y) {
                    for (int kx = 0; kx < Hx; ++kx) line[kx] = S[SIDX(c, kx, y, z)];
                    for (int kx = 1; kx < Hx; ++kx) {
                        const int m = Px - kx;
                        if (m >= Hx && m < Px) { line[m].re = line[kx].re; line[m].im = -line[kx].im; }
                    }
                    fft_exec(c->plx, line, scr, +1);
                    for (int x = 0; x < nx; ++x) Hc[x + (size_t)nx * (y + (size_t)ny * z)] = line[x].re;
                }
            free(line); free(scr);
        }
    }
}

/* SymmetricMatrixVectorConvolution_Simple.cpp:36-80 */
void orc_conv_simple(const double *N, int nx, int ny, int nz, int Px, int Py, int Pz,
                     const double *M, double *H)
{
    const size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for collapse(3) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                double hx = 0.0, hy = 0.0, hz = 0.0;
                for (int o = 0; o < nz; ++o)
                    for (int q = 0; q < ny; ++q)
                        for (int m = 0; m < nx; ++m) {
                            const int X = (x - m + Px) % Px;
                            const int Y = (y - q + Py) % Py;
                            const int Z = (z - o + Pz) % Pz;
                            const double *Ne = N + 6 * ((size_t)X + (size_t)Px * ((size_t)Y + (size_t)Py * Z));
                            const size_t j = m + (size_t)nx * (q + (size_t)ny * o);
                            const double mx = M[j], my = M[n + j], mz = M[2 * n + j];
                            hx += Ne[0] * mx + Ne[1] * my + Ne[2] * mz;
                            hy += Ne[1] * mx + Ne[3] * my + Ne[4] * mz;
                            hz += Ne[2] * mx + Ne[4] * my + Ne[5] * mz;
                        }
                const size_t i = x + (size_t)nx * (y + (size_t)ny * z);
                H[i] = hx; H[n + i] = hy; H[2 * n + i] = hz;
            }
}

/* ------------------------------------------------------------------------ */
/* exchange: exchange_cpu.cpp:60-139 (open) and :141-235 (periodic)          */
/* ------------------------------------------------------------------------ */
double orc_exchange(int nx, int ny, int nz, double dx, double dy, double dz,
                    int px, int py, int pz, const double *Ms, const double *A,
                    const double *M, double *H)
{
    const size_t n = (size_t)nx * ny * nz;
    const long nxy = (long)nx * ny;
    const double w[3] = {1.0 / (dx * dx), 1.0 / (dy * dy), 1.0 / (dz * dz)};
    const int per[3] = {px, py, pz};
    const int dim[3] = {nx, ny, nz};
    const long stride[3] = {1, nx, nxy};
    const double MU0 = ORC_MU0;
    double *esum = (double *)malloc(sizeof(double) * n);

#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const long i = x + (long)nx * (y + (long)ny * z);
                const double ms = Ms[i];
                if (ms == 0.0) {
                    H[i] = H[n + i] = H[2 * n + i] = 0.0;
                    esum[i] = 0.0;
                    continue;
                }
                const int pos[3] = {x, y, z};
                const double mi[3] = {M[i] / ms, M[n + i] / ms, M[2 * n + i] / ms};
                double s[3] = {0.0, 0.0, 0.0};
                for (int ax = 0; ax < 3; ++ax) {
                    for (int side = -1; side <= 1; side += 2) {
                        const int p = pos[ax] + side;
                        long j;
                        if (p < 0 || p >= dim[ax]) {
                            if (!per[ax]) continue;
                            j = i + side * stride[ax] - side * stride[ax] * dim[ax];
                        } else {
                            j = i + side * stride[ax];
                        }
                        const double msj = Ms[j];
                        if (msj == 0.0) continue;
                        for (int c = 0; c < 3; ++c)
                            s[c] += (M[c * n + j] / msj - mi[c]) * w[ax];
                    }
                }
                const double f = (2.0 / MU0) * A[i];
                double hv[3];
                for (int c = 0; c < 3; ++c) hv[c] = f * s[c] / ms;
                H[i] = hv[0]; H[n + i] = hv[1]; H[2 * n + i] = hv[2];
                esum[i] = mi[0] * hv[0] + mi[1] * hv[1] + mi[2] * hv[2];
            }
    double energy = 0.0;
    for (size_t i = 0; i < n; ++i) energy += esum[i];
    free(esum);
    return energy * (-MU0 / 2.0 * dx * dy * dz);
}

/* ------------------------------------------------------------------------ */
/* anisotropy: anisotropy_cpu.cpp:29-73 (uniaxial), :75-119 (cubic)         */
/* ------------------------------------------------------------------------ */
double orc_uniaxial_anisotropy(int n, const double *axis, const double *k,
                               const double *Ms, const double *M, double *H)
{
    const double MU0 = ORC_MU0;
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        const double ms = Ms[i], ki = k[i];
        if (ms == 0.0 || ki == 0.0) { H[i] = H[n + i] = H[2 * n + i] = 0.0; continue; }
        const double ax = axis[i], ay = axis[n + i], az = axis[2 * n + i];
        const double sx = M[i] / ms, sy = M[n + i] / ms, sz = M[2 * n + i] / ms;
        const double d = sx * ax + sy * ay + sz * az;
        const double f = 2.0 * ki * d / ms / MU0;
        H[i] = f * ax; H[n + i] = f * ay; H[2 * n + i] = f * az;
        const double cx = ay * sz - az * sy, cy = az * sx - ax * sz, cz = ax * sy - ay * sx;
        e += ki * (cx * cx + cy * cy + cz * cz);
    }
    return e;
}

double orc_cubic_anisotropy(int n, const double *axis1, const double *axis2,
                            const double *k, const double *Ms, const double *M, double *H)
{
    const double MU0 = ORC_MU0;
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        const double ms = Ms[i];
        if (ms == 0.0) { H[i] = H[n + i] = H[2 * n + i] = 0.0; continue; }
        const double ki = k[i];
        /* Vector3d normalize(M, 1.0): M * (1/|M|) */
        const double mx0 = M[i], my0 = M[n + i], mz0 = M[2 * n + i];
        const double len = sqrt(mx0 * mx0 + my0 * my0 + mz0 * mz0);
        const double sc = len == 0.0 ? 1.0 : 1.0 / len; /* Vector3d::normalize keeps a zero vector */
        const double mx = mx0 * sc, my = my0 * sc, mz = mz0 * sc;
        const double u1x = axis1[i], u1y = axis1[n + i], u1z = axis1[2 * n + i];
        const double u2x = axis2[i], u2y = axis2[n + i], u2z = axis2[2 * n + i];
        const double u3x = u1y * u2z - u1z * u2y, u3y = u1z * u2x - u1x * u2z, u3z = u1x * u2y - u1y * u2x;
        const double a1 = u1x * mx + u1y * my + u1z * mz, a1s = a1 * a1;
        const double a2 = u2x * mx + u2y * my + u2z * mz, a2s = a2 * a2;
        const double a3 = u3x * mx + u3y * my + u3z * mz, a3s = a3 * a3;
        const double f = (-2.0 / MU0) * ki / ms;
        const double c1 = (a2s + a3s) * a1, c2 = (a1s + a3s) * a2, c3 = (a1s + a2s) * a3;
        H[i] = f * (c1 * u1x + c2 * u2x + c3 * u3x);
        H[n + i] = f * (c1 * u1y + c2 * u2y + c3 * u3y);
        H[2 * n + i] = f * (c1 * u1z + c2 * u2z + c3 * u3z);
        e += ki * (a1s * a2s + a1s * a3s + a2s * a3s);
    }
    return e;
}

/* ------------------------------------------------------------------------ */
/* LLG right-hand side: llge_cpu.cpp:25-45                                   */
/* ------------------------------------------------------------------------ */
void orc_llge(int n, const double *f1, const double *f2, const double *M,
              const double *H, double *dM)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const double mx = M[i], my = M[n + i], mz = M[2 * n + i];
        const double hx = H[i], hy = H[n + i], hz = H[2 * n + i];
        const double ax = my * hz - mz * hy, ay = mz * hx - mx * hz, az = mx * hy - my * hx;
        const double bx = my * az - mz * ay, by = mz * ax - mx * az, bz = mx * ay - my * ax;
        dM[i] = f1[i] * ax + f2[i] * bx;
        dM[n + i] = f1[i] * ay + f2[i] * by;
        dM[2 * n + i] = f1[i] * az + f2[i] * bz;
    }
}

/* ------------------------------------------------------------------------ */
/* Zhang-Li spin torque: fdm_zhangli_cpu.cpp:57-200                          */
/* ------------------------------------------------------------------------ */
void orc_zhangli(int nx, int ny, int nz, double dx, double dy, double dz, int do_precess,
                 const double *P, const double *xi, const double *Ms, const double *alpha,
                 const double *j, const double *M, double *dM)
{
    const long n = (long)nx * ny * nz;
    const long nxy = (long)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const long k = x + (long)nx * (y + (long)ny * z);
                const double ms = Ms[k];
                if (ms == 0.0) { dM[k] = dM[n + k] = dM[2 * n + k] = 0.0; continue; }
                const double al = alpha[k], p = P[k], x_i = xi[k];
                const double bj = p * ORC_MU_BOHR / (ORC_E_CHARGE * ms * (1.0 + x_i * x_i));
                const double bjp = bj / (1.0 + al * al);
                double f1, f2;
                if (do_precess) {
                    f1 = -bjp / (ms * ms) * (1.0 + al * x_i);
                    f2 = -bjp / ms * (x_i - al);
                } else {
                    f1 = -bjp / (ms * ms) * (al * x_i);
                    f2 = -bjp / ms * (-al);
                }
                double g[3][3]; /* g[axis][comp] = dM/d(axis) */
                const int pos[3] = {x, y, z};
                const int dim[3] = {nx, ny, nz};
                const long st[3] = {1, nx, nxy};
                const double del[3] = {dx, dy, dz};
                for (int ax = 0; ax < 3; ++ax) {
                    for (int c = 0; c < 3; ++c) {
                        if (dim[ax] <= 1) { g[ax][c] = 0.0; continue; }
                        const double m0 = M[c * n + k];
                        if (pos[ax] == 0) g[ax][c] = (M[c * n + k + st[ax]] - m0) / del[ax];
                        else if (pos[ax] == dim[ax] - 1) g[ax][c] = (m0 - M[c * n + k - st[ax]]) / del[ax];
                        else g[ax][c] = (M[c * n + k + st[ax]] - M[c * n + k - st[ax]]) / (2 * del[ax]);
                    }
                }
                const double jx = j[k], jy = j[n + k], jz = j[2 * n + k];
                double jg[3];
                for (int c = 0; c < 3; ++c) jg[c] = jx * g[0][c] + jy * g[1][c] + jz * g[2][c];
                const double mx = M[k], my = M[n + k], mz = M[2 * n + k];
                const double Bx = my * jg[2] - mz * jg[1], By = mz * jg[0] - mx * jg[2], Bz = mx * jg[1] - my * jg[0];
                const double Ax = my * Bz - mz * By, Ay = mz * Bx - mx * Bz, Az = mx * By - my * Bx;
                dM[k] = f1 * Ax + f2 * Bx;
                dM[n + k] = f1 * Ay + f2 * By;
                dM[2 * n + k] = f1 * Az + f2 * Bz;
            }
}

/* ------------------------------------------------------------------------ */
/* reductions: ScaledAbsMax.cpp:33-68, CPUDevice.cpp:192-313                 */
/* ------------------------------------------------------------------------ */
double orc_abs_max(int n, const double *M)
{
    double mx = -DBL_MAX;
    for (int i = 0; i < n; ++i) {
        const double a = M[i], b = M[n + i], c = M[2 * n + i];
        const double s = a * a + b * b + c * c;
        if (s > mx) mx = s;
    }
    return sqrt(mx);
}

double orc_scaled_abs_max(int n, const double *M, const double *s, int s_uniform)
{
    if (s_uniform) return orc_abs_max(n, M) / s[0];
    double mx = 0.0;
    for (int i = 0; i < n; ++i) {
        const double si = s[i];
        if (si == 0.0) continue;
        const double a = M[i], b = M[n + i], c = M[2 * n + i];
        const double q = (a * a + b * b + c * c) / (si * si);
        if (q > mx) mx = q;
    }
    return sqrt(mx);
}

double orc_dot_sum(int n, const double *A, const double *B)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += A[i] * B[i] + A[n + i] * B[n + i] + A[2 * n + i] * B[2 * n + i];
    return s;
}

void orc_average(int n, const double *M, double out[3])
{
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += M[c * n + i];
        out[c] = s / n;
    }
}

/* CPUDevice.cpp:244-271 (mode 0).  The reference's `std::sqrt(e0*e0+e1*e1+
 * e2*e2)` has a float argument (e0..e2 are float, CPUDevice.cpp:255-258), so
 * it is the float overload: the norm is the correctly rounded float square
 * root of the float sum, then widened to double for `len / norm`. */
void orc_normalize(int n, double *M, const double *len, int mode)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double norm;
        if (mode == 0) {
            const float e0 = (float)M[i], e1 = (float)M[n + i], e2 = (float)M[2 * n + i];
            const float s = e0 * e0 + e1 * e1 + e2 * e2;
            norm = (double)sqrtf(s);
        } else {
            const double a = M[i], b = M[n + i], c = M[2 * n + i];
            norm = sqrt(a * a + b * b + c * c);
        }
        if (norm == 0.0) continue;
        const double sc = len[i] / norm;
        M[i] *= sc; M[n + i] *= sc; M[2 * n + i] *= sc;
    }
}

/* ------------------------------------------------------------------------ */
/* Runge-Kutta: tableaus.py:54-139, runge_kutta_cpu.cpp:28-194,              */
/* runge_kutta.py:89-137                                                     */
/* ------------------------------------------------------------------------ */
static void tab_clear(orc_tableau *t) { memset(t, 0, sizeof(*t)); }

void orc_tableau_get(int method, orc_tableau *t)
{
    tab_clear(t);
    if (method == ORC_RKF45) {
        t->num_steps = 6; t->order = 4; t->fsal = 0;
        const double a[6] = {0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
        const double b[6][5] = {{0}, {1.0 / 4.0}, {3.0 / 32.0, 9.0 / 32.0},
                                {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0},
                                {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0},
                                {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}};
        const double c4[6] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0};
        const double c5[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
        for (int i = 0; i < 6; ++i) {
            t->a[i] = a[i]; t->c[i] = c4[i]; t->ec[i] = c4[i] - c5[i];
            for (int j = 0; j < 5; ++j) t->b[i][j] = b[i][j];
        }
    } else if (method == ORC_CC45) {
        t->num_steps = 6; t->order = 4; t->fsal = 0;
        const double a[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
        const double b[6][5] = {{0}, {1.0 / 5.0}, {3.0 / 40.0, 9.0 / 40.0},
                                {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0},
                                {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0},
                                {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
        const double c4[6] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0};
        const double c5[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
        for (int i = 0; i < 6; ++i) {
            t->a[i] = a[i]; t->c[i] = c4[i]; t->ec[i] = c4[i] - c5[i];
            for (int j = 0; j < 5; ++j) t->b[i][j] = b[i][j];
        }
    } else if (method == ORC_DP54) {
        t->num_steps = 7; t->order = 5; t->fsal = 1;
        const double a[7] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
        const double b[7][6] = {{0}, {1.0 / 5.0}, {3.0 / 40.0, 9.0 / 40.0},
                                {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
                                {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
                                {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
                                {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
        const double c4[7] = {5179.0 / 57600.0, 0.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
        const double c5[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
        for (int i = 0; i < 7; ++i) {
            t->a[i] = a[i]; t->c[i] = c5[i]; t->ec[i] = c4[i] - c5[i];
            for (int j = 0; j < 6; ++j) t->b[i][j] = b[i][j];
        }
    } else if (method == ORC_RK23) {
        t->num_steps = 4; t->order = 3; t->fsal = 1;
        const double a[4] = {0.0, 0.5, 3.0 / 4.0, 1.0};
        const double b[4][3] = {{0}, {1.0 / 2.0}, {0.0, 3.0 / 4.0}, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}};
        const double c2[4] = {7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0};
        const double c3[4] = {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0, 0.0};
        for (int i = 0; i < 4; ++i) {
            t->a[i] = a[i]; t->c[i] = c3[i]; t->ec[i] = c2[i] - c3[i];
            for (int j = 0; j < 3; ++j) t->b[i][j] = b[i][j];
        }
    } else if (method == ORC_EULER) { /* euler.py:27-36: y += h dy/dt */
        t->num_steps = 1; t->order = 1; t->fsal = 0;
        t->c[0] = 1.0;
    } else { /* classical RK4, fixed step (runge_kutta_4.py:28-59, stage 4 at h) */
        t->num_steps = 4; t->order = 4; t->fsal = 0;
        t->a[1] = 0.5; t->a[2] = 0.5; t->a[3] = 1.0;
        t->b[1][0] = 0.5; t->b[2][1] = 0.5; t->b[3][2] = 1.0;
        t->c[0] = 1.0 / 6.0; t->c[1] = 2.0 / 6.0; t->c[2] = 2.0 / 6.0; t->c[3] = 1.0 / 6.0;
    }
}

/* ytmp for stage `step` (1 <= step < num_steps).
 * 6-stage tableaus: runge_kutta_cpu.cpp:28-107, ytmp = y + h*(sum_j b_sj k_j).
 * other tableaus  : runge_kutta.py:107-111, ytmp = y; ytmp += k_j*(h*b_sj). */
void orc_rk_prepare_step(int step, double h, const orc_tableau *t, double *const *k,
                         const double *y, double *ytmp, int n)
{
    const long m = 3L * n;
    if (t->num_steps == 6) {
        for (long i = 0; i < m; ++i) {
            double s = t->b[step][0] * k[0][i];
            for (int j = 1; j < step; ++j) s += t->b[step][j] * k[j][i];
            ytmp[i] = y[i] + h * s;
        }
    } else {
        for (long i = 0; i < m; ++i) ytmp[i] = y[i];
        for (int j = 0; j < step; ++j) {
            const double f = h * t->b[step][j];
            for (long i = 0; i < m; ++i) ytmp[i] += k[j][i] * f;
        }
    }
}

/* runge_kutta_cpu.cpp:139-194 (6 stages, c1==ec1==0 shortcut keeps the
 * same arithmetic minus the k1 term); runge_kutta.py:125-129 otherwise. */
void orc_rk_combine_result(double h, const orc_tableau *t, double *const *k,
                           double *y, double *y_err, int n)
{
    const long m = 3L * n;
    if (t->num_steps == 6) {
        const int skip1 = (t->c[1] == 0.0 && t->ec[1] == 0.0);
        for (long i = 0; i < m; ++i) {
            double s = t->c[0] * k[0][i], e = t->ec[0] * k[0][i];
            for (int j = 1; j < 6; ++j) {
                if (j == 1 && skip1) continue;
                s += t->c[j] * k[j][i];
                e += t->ec[j] * k[j][i];
            }
            y[i] = y[i] + h * s;
            y_err[i] = h * e;
        }
    } else {
        for (long i = 0; i < m; ++i) y_err[i] = 0.0;
        for (int j = 0; j < t->num_steps; ++j) {
            const double fc = h * t->c[j], fe = h * t->ec[j];
            for (long i = 0; i < m; ++i) {
                y[i] += k[j][i] * fc;
                y_err[i] += k[j][i] * fe;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* micromagnetic system                                                      */
/* ------------------------------------------------------------------------ */

/* landau_lifshitz_gilbert.py:133-158 */
void orc_llge_factors(const orc_system *s, double *f1, double *f2)
{
    const int n = s->nx * s->ny * s->nz;
    for (int i = 0; i < n; ++i) {
        const double a = s->alpha[i], ms = s->Ms[i];
        if (ms != 0.0) {
            const double gp = ORC_GAMMA / (1.0 + a * a);
            f1[i] = -gp;
            f2[i] = -a * gp / ms;
        } else {
            f1[i] = 0.0; f2[i] = 0.0;
        }
    }
}

/* H_tot = 0 + H_stray + H_exch + H_aniso + H_ext (landau_lifshitz_gilbert.py:92-101,
 * module order of the list given to create_solver; anisotropy_field.py:41-66). */
/* Zeeman field component c at time t: ExternalField = AlternatingField
 * (external_field.py:21-37, alternating_field.py:57-84) */
static double zeeman_at(const orc_system *s, int c, double t)
{
    if (s->h_fn) {
        double h[3];
        s->h_fn(t, h, s->h_user);
        return h[c];
    }
    if (!s->use_ac) return s->h_ext[c];
    return s->h_ext[c] + s->h_amp[c] * sin(t * s->h_freq[c] + s->h_phase[c]);
}

void orc_h_tot(const orc_system *s, const double *M, double *H_tot) { orc_h_tot_at(s, M, 0.0, H_tot); }

void orc_h_tot_at(const orc_system *s, const double *M, double t, double *H_tot)
{
    const int n = s->nx * s->ny * s->nz;
    const long m = 3L * n;
    double *tmp = (double *)malloc(sizeof(double) * m);
    double *tmp2 = (double *)malloc(sizeof(double) * m);
    for (long i = 0; i < m; ++i) H_tot[i] = 0.0;
    if (s->conv) {
        orc_conv_execute(s->conv, M, tmp);
        for (long i = 0; i < m; ++i) H_tot[i] += tmp[i];
    }
    if (s->use_exchange) {
        orc_exchange(s->nx, s->ny, s->nz, s->dx, s->dy, s->dz, s->px, s->py, s->pz, s->Ms, s->A, M, tmp);
        for (long i = 0; i < m; ++i) H_tot[i] += tmp[i];
    }
    if (s->k_uni || s->k_cub) {
        if (s->k_uni && s->k_cub) {
            orc_uniaxial_anisotropy(n, s->axis1, s->k_uni, s->Ms, M, tmp2);
            orc_cubic_anisotropy(n, s->axis1, s->axis2, s->k_cub, s->Ms, M, tmp);
            for (long i = 0; i < m; ++i) tmp[i] += tmp2[i];
        } else if (s->k_uni) {
            orc_uniaxial_anisotropy(n, s->axis1, s->k_uni, s->Ms, M, tmp);
        } else {
            orc_cubic_anisotropy(n, s->axis1, s->axis2, s->k_cub, s->Ms, M, tmp);
        }
        for (long i = 0; i < m; ++i) H_tot[i] += tmp[i];
    }
    if (s->use_zeeman) {
        for (int c = 0; c < 3; ++c) {
            const double hz = zeeman_at(s, c, t);
            for (int i = 0; i < n; ++i) H_tot[c * (long)n + i] += hz;
        }
    }
    free(tmp); free(tmp2);
}

/* landau_lifshitz_gilbert.py:108-125 */
void orc_dmdt(const orc_system *s, const double *f1, const double *f2,
              const double *M, double *dMdt, double *H_tot_out)
{
    orc_dmdt_at(s, f1, f2, M, 0.0, dMdt, H_tot_out);
}

void orc_dmdt_at(const orc_system *s, const double *f1, const double *f2, const double *M, double t,
                 double *dMdt, double *H_tot_out)
{
    const int n = s->nx * s->ny * s->nz;
    const long m = 3L * n;
    double *H = H_tot_out ? H_tot_out : (double *)malloc(sizeof(double) * m);
    orc_h_tot_at(s, M, t, H);
    orc_llge(n, f1, f2, M, H, dMdt);
    if (s->P) {
        double *st = (double *)malloc(sizeof(double) * m);
        orc_zhangli(s->nx, s->ny, s->nz, s->dx, s->dy, s->dz, s->do_precess,
                    s->P, s->xi, s->Ms, s->alpha, s->j, M, st);
        for (long i = 0; i < m; ++i) dMdt[i] += st[i];
        free(st);
    }
    if (!H_tot_out) free(H);
}

/* micro_magnetics_stepsize_controller.py:35-73 */
static int controller(double eps_abs, double eps_rel, double h, int order,
                      double max_y_err, double max_dydt, double *h_next)
{
    const double HEADROOM = 0.85, MINS = 0.2, MAXS = 5.0, MINH = 1e-15, MAXH = 1e-10;
    const double den = h * max_dydt;
    if (den == 0.0) { *h_next = 5e-14; return 1; }
    const double e_abs = max_y_err;
    const double e_rel = e_abs / den;
    int accept = (e_abs <= eps_abs) && (e_rel <= eps_rel);
    const double sa = HEADROOM * pow(eps_abs / e_abs, 1.0 / (order + 1.0));
    const double sr = HEADROOM * pow(eps_rel / e_rel, 1.0 / order);
    double sc = sa < sr ? sa : sr;
    if (sc < MINS) sc = MINS;
    if (sc > MAXS) sc = MAXS;
    double hn = h * sc;
    if (hn < MINH) hn = MINH;
    if (hn > MAXH) hn = MAXH;
    if (hn == MINH) accept = 1;
    *h_next = hn;
    return accept;
}

int orc_controller(double eps_abs, double eps_rel, double h, int order, double max_y_err,
                   double max_dydt, double *h_next)
{
    return controller(eps_abs, eps_rel, h, order, max_y_err, max_dydt, h_next);
}

/* Relaxation rule (mm_stepper_relax semantics): checked before a step whose
 * accepted-step count is a multiple of `every`, from the k_0 the step's first
 * attempt uses (dM/dt at M, or the carried-over FSAL vector).
 * criterion 1: scaled_abs_max(k_0, Ms) < thr; criterion 2: Relaxed(thr)
 * (condition.py:169-200) on deg_per_ns (landau_lifshitz_gilbert.py:127-131). */
typedef struct {
    int criterion, every;
    double thr;
    int have_last, relaxed;
    double last, value;
} relax_state;

static int relax_test(relax_state *r, const orc_system *s, int n, const double *k0,
                      const double *M, double h_last)
{
    double v;
    int relaxed;
    if (r->criterion == 1) {
        v = orc_scaled_abs_max(n, k0, s->Ms, s->Ms_uniform);
        relaxed = v < r->thr;
    } else {
        const double dm = orc_abs_max(n, k0), mm = orc_abs_max(n, M);
        const double deg_per_step = (180.0 / 3.14159265358979323846) * atan2(dm * h_last, mm);
        v = 1e-9 * deg_per_step / h_last;
        relaxed = r->have_last && v < r->last && v <= r->thr;
        r->last = v;
        r->have_last = 1;
    }
    r->value = v;
    if (relaxed) r->relaxed = 1;
    return relaxed;
}

/* RungeKutta.apply (runge_kutta.py:89-137), or one classical RK4 step
 * (runge_kutta_4.py:28-59 with stage 4 at y + h k3): one attempt from M with
 * step h.  k0_given: k[0] already holds this attempt's k_0.  FSAL methods
 * (dp54, rk23) take k_0 from the last stage of the most recent attempt,
 * accepted or not (runge_kutta.py:97-101,131-134). */
static void apply_attempt(const orc_system *s, int method, const orc_tableau *t, const double *f1,
                          const double *f2, double **k, double *M, double *ytmp, double *yerr,
                          double *fsal, int *have_fsal, double t0, double h, int k0_given, int n)
{
    const long m = 3L * n;
    if (!k0_given) {
        if (method != ORC_RK4 && t->fsal && *have_fsal) memcpy(k[0], fsal, sizeof(double) * m);
        else orc_dmdt_at(s, f1, f2, M, t0, k[0], NULL);
    }
    if (method == ORC_RK4) {
        for (int sg = 1; sg < 4; ++sg) {
            orc_rk_prepare_step(sg, h, t, k, M, ytmp, n);
            orc_dmdt_at(s, f1, f2, ytmp, t0 + h * t->a[sg], k[sg], NULL);
        }
        for (int sg = 0; sg < 4; ++sg) {
            const double f = h * (sg == 0 || sg == 3 ? 1.0 : 2.0) / 6.0;
            for (long i = 0; i < m; ++i) M[i] += k[sg][i] * f;
        }
        return;
    }
    for (int sg = 1; sg < t->num_steps; ++sg) {
        orc_rk_prepare_step(sg, h, t, k, M, ytmp, n);
        orc_dmdt_at(s, f1, f2, ytmp, t0 + h * t->a[sg], k[sg], NULL);
    }
    orc_rk_combine_result(h, t, k, M, yerr, n);
    if (t->fsal) {
        memcpy(fsal, k[t->num_steps - 1], sizeof(double) * m);
        *have_fsal = 1;
    }
}

/* RungeKutta.evolve (runge_kutta.py:47-87) per accepted step, then
 * finish_step = normalize(Ms) (micro_magnetics.py:49-50); classical RK4 with
 * the fixed h0 for ORC_RK4.  t_max (tmax_on): an accepted step past it is
 * redone from the same state with h = t_max - t (runge_kutta.py:74-78) and
 * stepping ends once t >= t_max.  tmax_on == 2: no clipping, stepping ends
 * after the first step that ends at t >= t_max (the solver loop of
 * solver.py:110-112 with TimeGreaterEq, condition.py:145-157, which names no
 * time of interest, so evolve runs with t_max = 1e100). */
static long run_impl(const orc_system *s, int method, double eps_abs, double eps_rel,
                     double h0, long steps, double *M, double *stats, double *avg_out,
                     relax_state *rx, int tmax_on, double t_max)
{
    const int n = s->nx * s->ny * s->nz;
    const long m = 3L * n;
    orc_tableau t;
    orc_tableau_get(method, &t);
    double *f1 = (double *)malloc(sizeof(double) * n);
    double *f2 = (double *)malloc(sizeof(double) * n);
    orc_llge_factors(s, f1, f2);
    double *k[7];
    for (int i = 0; i < 7; ++i) k[i] = (double *)malloc(sizeof(double) * m);
    double *y0 = (double *)malloc(sizeof(double) * m);
    double *ytmp = (double *)malloc(sizeof(double) * m);
    double *yerr = (double *)malloc(sizeof(double) * m);
    double *fsal = (double *)malloc(sizeof(double) * m);
    int have_fsal = 0;
    double tnow = 0.0, h_try = h0, h_last = h0, h_next = h0;
    long attempts = 0, rejections = 0;

    long st;
    for (st = 0; st < steps; ++st) {
        if (tmax_on && tnow >= t_max) break;
        int k0_ready = 0;
        if (rx && st % rx->every == 0) {
            if (method != ORC_RK4 && t.fsal && have_fsal) memcpy(k[0], fsal, sizeof(double) * m);
            else orc_dmdt_at(s, f1, f2, M, tnow, k[0], NULL);
            k0_ready = 1;
            if (relax_test(rx, s, n, k[0], M, h_last)) break;
        }
        memcpy(y0, M, sizeof(double) * m);
        for (;;) {
            const int fixed = method == ORC_RK4 || method == ORC_EULER;
            double h = fixed ? h0 : h_try;
            attempts++;
            apply_attempt(s, method, &t, f1, f2, k, M, ytmp, yerr, fsal, &have_fsal, tnow, h, k0_ready, n);
            k0_ready = 0;
            double hn = h;
            int acc = 1;
            if (!fixed) {
                const double max_err = orc_scaled_abs_max(n, yerr, s->Ms, s->Ms_uniform);
                const double max_dydt = orc_scaled_abs_max(n, k[0], s->Ms, s->Ms_uniform);
                acc = controller(eps_abs, eps_rel, h, t.order, max_err, max_dydt, &hn);
            }
            if (!acc) {
                rejections++;
                memcpy(M, y0, sizeof(double) * m);
                h_try = hn;
                continue;
            }
            if (tmax_on == 1 && method != ORC_EULER && tnow + h > t_max) {
                h = t_max - tnow;
                memcpy(M, y0, sizeof(double) * m);
                attempts++;
                apply_attempt(s, method, &t, f1, f2, k, M, ytmp, yerr, fsal, &have_fsal, tnow, h, 0, n);
            }
            tnow += h; h_last = h; h_next = hn;
            h_try = hn;
            break;
        }
        orc_normalize(n, M, s->Ms, s->normalize_mode);
        if (avg_out) orc_average(n, M, avg_out + 3 * st);
    }
    if (stats) {
        stats[0] = tnow; stats[1] = h_last; stats[2] = h_next;
        stats[3] = (double)attempts; stats[4] = (double)rejections;
    }
    for (int i = 0; i < 7; ++i) free(k[i]);
    free(f1); free(f2); free(y0); free(ytmp); free(yerr); free(fsal);
    return st;
}

void orc_run(const orc_system *s, int method, double eps_abs, double eps_rel,
             double h0, int steps, double *M, double *stats, double *avg_out)
{
    run_impl(s, method, eps_abs, eps_rel, h0, steps, M, stats, avg_out, NULL, 0, 0.0);
}

long orc_run_until(const orc_system *s, int method, double eps_abs, double eps_rel, double h0,
                   double t_end, long max_steps, double *M, double *stats)
{
    return run_impl(s, method, eps_abs, eps_rel, h0, max_steps, M, stats, NULL, NULL, 1, t_end);
}

long orc_run_to_time(const orc_system *s, int method, double eps_abs, double eps_rel, double h0,
                     double t_stop, long max_steps, double *M, double *stats)
{
    return run_impl(s, method, eps_abs, eps_rel, h0, max_steps, M, stats, NULL, NULL, 2, t_stop);
}

long orc_relax(const orc_system *s, int method, double eps_abs, double eps_rel, double h0,
               int criterion, double threshold, int check_every, long max_steps,
               double *M, double *stats, double *info3)
{
    relax_state r;
    memset(&r, 0, sizeof(r));
    r.criterion = criterion;
    r.every = check_every < 1 ? 1 : check_every;
    r.thr = threshold;
    const long done = run_impl(s, method, eps_abs, eps_rel, h0, max_steps, M, stats, NULL, &r, 0, 0.0);
    if (info3) {
        info3[0] = (double)done;
        info3[1] = (double)r.relaxed;
        info3[2] = r.value;
    }
    return done;
}

/* Term energies at M: stray_field.py:51, exchange_field.py:58,
 * anisotropy_field.py:84-87 (uniaxial + cubic summed), external_field.py:33-35,
 * E_tot landau_lifshitz_gilbert.py:103-106; MU0 as the Python front-end's
 * 4.0*math.pi*1e-7 (constants.py:21). */
void orc_energy(const orc_system *s, const double *M, double *E5) { orc_energy_at(s, M, 0.0, E5); }

void orc_energy_at(const orc_system *s, const double *M, double t, double *E5)
{
    const int n = s->nx * s->ny * s->nz;
    const long m = 3L * n;
    const double mu0 = 4.0 * 3.14159265358979323846 * 1e-7;
    const double V = s->dx * s->dy * s->dz;
    double *tmp = (double *)malloc(sizeof(double) * m);
    for (int q = 0; q < 5; ++q) E5[q] = 0.0;
    if (s->conv) {
        orc_conv_execute(s->conv, M, tmp);
        E5[0] = -mu0 / 2.0 * V * orc_dot_sum(n, M, tmp);
    }
    if (s->use_exchange) {
        orc_exchange(s->nx, s->ny, s->nz, s->dx, s->dy, s->dz, s->px, s->py, s->pz, s->Ms, s->A, M, tmp);
        E5[1] = -mu0 / 2.0 * V * orc_dot_sum(n, M, tmp);
    }
    if (s->k_uni || s->k_cub) {
        double e = 0.0;
        if (s->k_uni) e += orc_uniaxial_anisotropy(n, s->axis1, s->k_uni, s->Ms, M, tmp);
        if (s->k_cub) e += orc_cubic_anisotropy(n, s->axis1, s->axis2, s->k_cub, s->Ms, M, tmp);
        E5[2] = e * V;
    }
    if (s->use_zeeman) {
        for (int c = 0; c < 3; ++c) {
            const double hz = zeeman_at(s, c, t);
            for (int i = 0; i < n; ++i) tmp[c * (long)n + i] = hz;
        }
        E5[3] = -mu0 * V * orc_dot_sum(n, M, tmp);
    }
    E5[4] = E5[0] + E5[1] + E5[2] + E5[3];
    free(tmp);
}
