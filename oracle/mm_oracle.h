/*
 * mm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of MicroMagnum's effective-field + Runge-Kutta hot
 * path.  It is the checker for the HIP library: only tests/, the smoke check
 * and bench.py's CPU-baseline leg load it.  Nothing in the product links it.
 *
 * Layout conventions (same as the reference's matty storage,
 * src/magneto/matrix/Shape.cpp:64-75, VectorMatrix_accessor.cpp:28-41):
 *   scalar field   : double[n],      n = nx*ny*nz, index x + nx*(y + ny*z)
 *   vector field   : double[3*n],    SoA: [x-comp n | y-comp n | z-comp n]
 *   demag tensor   : double[6*Px*Py*Pz], component fastest
 *                    (xx,xy,xz,yy,yz,zz), as GenerateDemagTensor returns it
 *                    (src/magneto/mmm/demag/tensor.cpp:40).
 */
#ifndef MM_ORACLE_H
#define MM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (src/magneto/mmm/constants.h:26-40) ---- */
double orc_mu0(void);
void orc_set_threads(int n);

/* ---- 1-D complex DFT, unnormalised; sign -1 forward, +1 inverse ---- */
void orc_fft(double *re_im /* interleaved, 2*n */, int n, int sign);

/* ---- FFT convolution (src/magneto/math/conv/MatrixVectorConvolution_FFT.cpp:110-240) ---- */
typedef struct orc_conv orc_conv;
orc_conv *orc_conv_create(const double *N, int nx, int ny, int nz, int Px, int Py, int Pz);
void orc_conv_execute(orc_conv *c, const double *M, double *H);
void orc_conv_destroy(orc_conv *c);
/* Fourier kernel as stored by the reference (re/im planes, x < Px/2+1), natural
 * xyz order: out_re/out_im each 6*Hx*Py*Pz, index e + 6*(kx + Hx*(ky + Py*kz)). */
void orc_conv_get_kernel(orc_conv *c, double *out_re, double *out_im);

/* ---- direct-sum convolution (SymmetricMatrixVectorConvolution_Simple.cpp:36-80) ---- */
void orc_conv_simple(const double *N, int nx, int ny, int nz, int Px, int Py, int Pz,
                     const double *M, double *H);

/* ---- local field terms ---- */
double orc_exchange(int nx, int ny, int nz, double dx, double dy, double dz,
                    int px, int py, int pz, const double *Ms, const double *A,
                    const double *M, double *H);
double orc_uniaxial_anisotropy(int n, const double *axis, const double *k,
                               const double *Ms, const double *M, double *H);
double orc_cubic_anisotropy(int n, const double *axis1, const double *axis2,
                            const double *k, const double *Ms, const double *M, double *H);
void orc_llge(int n, const double *f1, const double *f2, const double *M,
              const double *H, double *dM);
void orc_zhangli(int nx, int ny, int nz, double dx, double dy, double dz, int do_precess,
                 const double *P, const double *xi, const double *Ms, const double *alpha,
                 const double *j, const double *M, double *dM);

/* ---- reductions / normalisation ---- */
double orc_scaled_abs_max(int n, const double *M, const double *s, int s_uniform);
double orc_abs_max(int n, const double *M);
double orc_dot_sum(int n, const double *A, const double *B);
void orc_average(int n, const double *M, double out[3]);
/* mode 0: reference CPU semantics (norm from float-cast components, zero
 * vector kept); mode 1: plain FP64 norm. */
void orc_normalize(int n, double *M, const double *len, int mode);

/* ---- Runge-Kutta tableaus and stage algebra ---- */
/* ORC_EULER: the reference's Euler evolver (euler.py:27-36): y += h dy/dt with
 * the fixed step h0, then finish_step; it does not clip its last step to
 * t_max, so a run to t_end ends at the first multiple of h0 at or past it */
enum { ORC_RKF45 = 0, ORC_CC45 = 1, ORC_DP54 = 2, ORC_RK23 = 3, ORC_RK4 = 4, ORC_EULER = 5 };
typedef struct {
    int num_steps, order, fsal;
    double a[7], b[7][7], c[7], ec[7];
} orc_tableau;
void orc_tableau_get(int method, orc_tableau *t);
void orc_rk_prepare_step(int step, double h, const orc_tableau *t, double *const *k,
                         const double *y, double *ytmp, int n);
void orc_rk_combine_result(double h, const orc_tableau *t, double *const *k,
                           double *y, double *y_err, int n);

/* ---- micromagnetic system: H_tot and dM/dt in module order ---- */
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
    int px, py, pz;               /* periodic flags (exchange only)                 */
    const double *Ms, *A, *alpha; /* per-cell                                      */
    int Ms_uniform;               /* reference Field.isUniform() of Ms              */
    orc_conv *conv;               /* NULL: no stray field                           */
    int use_exchange;
    const double *k_uni, *k_cub, *axis1, *axis2; /* NULL k: term absent             */
    int use_zeeman;
    double h_ext[3];
    const double *P, *xi, *j;     /* NULL P: no spin torque                         */
    int do_precess;
    int normalize_mode;           /* 0 reference float-norm, 1 fp64                 */
    /* AlternatingField (alternating_field.py:57-84): with use_ac the Zeeman
     * field is h_ext + h_amp * sin(t * h_freq + h_phase) per component, at
     * the time t of the evaluation (stage s of an RK attempt from t0 with
     * step h: t0 + a_s h, runge_kutta.py:115) */
    int use_ac;
    double h_amp[3], h_freq[3], h_phase[3];
    /* AlternatingField with a user function (alternating_field.py:60-75):
     * h_fn(t, h3, h_user) gives the uniform field at t; it replaces h_ext
     * and the alternating parameters (use_zeeman must be set) */
    void (*h_fn)(double t, double *h3, void *user);
    void *h_user;
} orc_system;

void orc_llge_factors(const orc_system *s, double *f1, double *f2);
void orc_h_tot(const orc_system *s, const double *M, double *H_tot);
void orc_dmdt(const orc_system *s, const double *f1, const double *f2,
              const double *M, double *dMdt, double *H_tot_out /* may be NULL */);
/* the same at time t (orc_h_tot / orc_dmdt evaluate at t = 0) */
void orc_h_tot_at(const orc_system *s, const double *M, double t, double *H_tot);
void orc_dmdt_at(const orc_system *s, const double *f1, const double *f2, const double *M, double t,
                 double *dMdt, double *H_tot_out);

/* Adaptive RK (rkf45/cc45/dp54/rk23) driven by MicroMagneticsStepSizeController,
 * or fixed-step classical RK4.  Performs `steps` accepted steps in place on M.
 * stats: [0]=t, [1]=h (last accepted), [2]=h_next, [3]=attempts, [4]=rejections. */
void orc_run(const orc_system *s, int method, double eps_abs, double eps_rel,
             double h0, int steps, double *M, double *stats,
             double *avg_out /* 3*steps or NULL: <M> after each accepted step */);

/* Run until the relaxation rule holds or max_steps accepted steps (the rule
 * of mm_stepper_relax: criterion 1 scaled_abs_max(dM/dt, Ms) < threshold,
 * 2 the reference's Relaxed condition on deg_per_ns).  info3: [0] steps
 * made, [1] relaxed, [2] last checked value.  Returns the steps made. */
long orc_relax(const orc_system *s, int method, double eps_abs, double eps_rel, double h0,
               int criterion, double threshold, int check_every, long max_steps,
               double *M, double *stats, double *info3);

/* Run until t_end (the last accepted step redone with h = t_end - t as in
 * runge_kutta.py:74-78) or max_steps accepted steps.  Returns the steps made. */
long orc_run_until(const orc_system *s, int method, double eps_abs, double eps_rel, double h0,
                   double t_end, long max_steps, double *M, double *stats);
/* steps until the first accepted step that ends at t >= t_stop, none clipped
 * (solve(Time(t_stop)), solver.py:110-112); returns the accepted steps */
long orc_run_to_time(const orc_system *s, int method, double eps_abs, double eps_rel, double h0,
                     double t_stop, long max_steps, double *M, double *stats);

/* MicroMagneticsStepSizeController.adjust_stepsize
 * (micro_magnetics_stepsize_controller.py:35-73) on the two maxima it takes
 * from scaled_abs_max: returns accept, writes the next step size. */
int orc_controller(double eps_abs, double eps_rel, double h, int order, double max_y_err,
                   double max_dydt, double *h_next);

/* E_stray, E_exch, E_aniso, E_ext, E_tot at M [J]. */
void orc_energy(const orc_system *s, const double *M, double *E5);
void orc_energy_at(const orc_system *s, const double *M, double t, double *E5);

#ifdef __cplusplus
}
#endif
#endif
