/*
 * mm_hotpath.h -- C ABI of libmmhip.so, the MI355X (gfx950) implementation of
 * MicroMagnum's effective-field and Runge-Kutta hot path.
 *
 * Conventions
 *   - Every function returns MM_OK (0) or an error code; mm_last_error()
 *     returns the message of the last failure on the calling thread.  The
 *     reference raises RuntimeError from C++ exceptions
 *     (src/magneto/bindings/magneto.i:31-42); a binding maps a non-zero code
 *     to the same exception.
 *   - Array arguments are DEVICE pointers unless the name ends in _host.
 *     Layout is the reference's (src/magneto/matrix/Shape.cpp:64-75):
 *     x fastest, index x + nx*(y + ny*z); vector fields are SoA, three
 *     consecutive planes of n = nx*ny*nz values (x, y, z components).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Functions that return a scalar (energies, norms) synchronise the
 *     stream; the others only enqueue work.
 *   - FP64 entry points compute in double.  *_f32 entry points take float
 *     arrays and compute in float (the mixed-precision bulk path).
 */
#ifndef MM_HOTPATH_H
#define MM_HOTPATH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MM_OK = 0,
    MM_ERR_INVALID = 1,     /* bad argument (shape, pointer, option)       */
    MM_ERR_RUNTIME = 2,     /* internal failure                            */
    MM_ERR_HIP = 3,         /* HIP runtime error                           */
    MM_ERR_UNSUPPORTED = 4  /* option the library does not implement        */
};

/* padding policies: src/magneto/bindings/mmm/demag.i:27-34 */
enum {
    MM_PADDING_DISABLE = 0,
    MM_PADDING_ROUND_2 = 1,
    MM_PADDING_ROUND_4 = 2,
    MM_PADDING_ROUND_8 = 3,
    MM_PADDING_ROUND_POT = 4,
    MM_PADDING_SMALL_PRIME_FACTORS = 5
};

/* precision of a convolution object / stepper */
enum { MM_FP64 = 0, MM_FP32 = 1 };

/* Runge-Kutta methods: src/magnum/evolver/tableaus.py:54-139 (+ classical RK4,
 * src/magnum/evolver/runge_kutta_4.py:28-59) */
/* MM_EULER: the reference's Euler evolver (src/magnum/evolver/euler.py:27-36),
 * y += h dy/dt with the fixed step h0; as there, its last step is not clipped
 * to the end time of mm_stepper_run_until. */
enum { MM_RKF45 = 0, MM_CC45 = 1, MM_DP54 = 2, MM_RK23 = 3, MM_RK4 = 4, MM_EULER = 5 };

/* ---------------------------------------------------------------- library */
/* Last error message of this thread (never NULL). */
const char *mm_last_error(void);
/* Select the HIP device and initialise the library on it.  Replaces
 * magneto.initialize + enableCuda (src/magneto/Magneto.cpp:56-73,153-176). */
int mm_init(int device);
int mm_device_alloc(size_t bytes, void **out);
int mm_device_free(void *ptr);
int mm_memcpy_htod(void *dst, const void *src_host, size_t bytes, void *stream);
int mm_memcpy_dtoh(void *dst_host, const void *src, size_t bytes, void *stream);
int mm_synchronize(void *stream);
/* Device-to-device streaming copy (bytes a multiple of 16). */
int mm_device_copy(void *dst, const void *src, size_t bytes, void *stream);

/* ------------------------------------------------------- demag tensor setup */
/* round_tensor_dimension (src/magneto/mmm/demag/tensor_round.cpp:90-106) */
int mm_round_tensor_dimension(int dim, int periodic, int padding, int *out);
/* GenerateDemagTensor (src/magneto/mmm/demag/demag_tensor.cpp:116-186).
 * exp_out receives (ex, ey, ez): padded per `padding` along open axes, the
 * mesh size along periodic ones.  N_host (may be NULL to query the size)
 * receives 6*ex*ey*ez doubles, component fastest (xx,xy,xz,yy,yz,zz),
 * already sign-flipped so that H = N (*) M.  Periodic axes sum `repeat`
 * images per axis cell by cell plus the far-image (infinity) correction of
 * tensor.cpp:66-160, which the environment variable
 * MAGNUM_DEMAG_NO_INFINITY_CORRECTION switches off as in the reference.
 * cache_dir (may be NULL): a field cached there under the reference's file
 * name (demag_tensor.cpp:52-112, raw doubles) is loaded instead of computed;
 * a computation longer than 30 s (or any, with MM_DEMAG_CACHE_ALWAYS set) is
 * saved there.  MAGNUM_DEMAG_GARBAGE skips the computation (zeros).
 * nthreads <= 0: one per host core. */
int mm_generate_demag_tensor(int nx, int ny, int nz, double dx, double dy, double dz,
                             int px, int py, int pz, int repeat, int padding,
                             const char *cache_dir, int nthreads, double *N_host, int *exp_out);

/* GenerateDemagTensor for open boundaries with the GPU doing the work
 * (demag_tensor.cpp:116-186; tensor.cpp:35-191).  One GPU thread per entry of
 * the positive octant evaluates the six components in FP64 and writes their
 * mirror images; entries whose branch is a finite difference (the near field
 * and the partly expanded shells, where FP64 loses ~6 log10(r) digits to
 * cancellation at distance r cells) or whose branch choice is within 1e-12 of
 * a threshold are evaluated on the host in long double, bit-identical to
 * mm_generate_demag_tensor -- unless near_on_gpu, which keeps them in FP64 on
 * the GPU.  Output as mm_generate_demag_tensor (padding as there, periodic
 * axes not taken).  info4 (may be NULL): [0] entries done on the GPU, [1] on
 * the host, [2] GPU seconds (incl. the copy), [3] host seconds.  Uses the
 * device selected by mm_init. */
int mm_generate_demag_tensor_gpu(int nx, int ny, int nz, double dx, double dy, double dz, int padding,
                                 int near_on_gpu, int nthreads, double *N_host, int *exp_out,
                                 double *info4);
/* The open-boundary tensor's value at (i, j, k), 0 <= i < nx etc., of the
 * padded field mm_generate_demag_tensor returns (six components, long
 * double arithmetic as there), for checking sampled entries of large meshes. */
int mm_demag_tensor_entry(int i, int j, int k, double dx, double dy, double dz, double *v6);

/* ------------------------------------------------------ FFT convolution */
/* SymmetricMatrixVectorConvolution_FFT(N, nx, ny, nz)
 * (src/magneto/math/conv/SymmetricMatrixVectorConvolution_FFT.cpp:31-65).
 * N_host: tensor as returned by mm_generate_demag_tensor for the padded size
 * (ex, ey, ez).  precision MM_FP64 or MM_FP32 selects the arithmetic of the
 * transforms; the Fourier kernel is computed in FP64 either way. */
typedef struct mm_conv mm_conv_t;
int mm_conv_create(const double *N_host, int nx, int ny, int nz, int ex, int ey, int ez,
                   int precision, mm_conv_t **out);
/* execute(M, H) (MatrixVectorConvolution_FFT.cpp:110-375): H = N (*) M. */
int mm_conv_execute(mm_conv_t *conv, const double *M, double *H, void *stream);
int mm_conv_execute_f32(mm_conv_t *conv, const float *M, float *H, void *stream);
/* padded FFT sizes actually used (radix-friendly; result-neutral for open
 * boundaries) */
int mm_conv_fft_size(const mm_conv_t *conv, int *fft_xyz);
int mm_conv_destroy(mm_conv_t *conv);

/* SymmetricMatrixVectorConvolution_Simple
 * (src/magneto/math/conv/SymmetricMatrixVectorConvolution_Simple.cpp:23-80,
 * bindings/math.i:35-41): the direct sum H(r) = sum_r' N((r - r') mod E) M(r')
 * over the mesh, FP64, O(n^2) -- what the reference's StrayFieldCalculator
 * uses below 32 cells on the CPU (stray_field_calculator.py:109-121).  N_host:
 * the tensor as for mm_conv_create (6 entries per point of the (ex, ey, ez)
 * grid). */
typedef struct mm_conv_simple mm_conv_simple_t;
int mm_conv_simple_create(const double *N_host, int nx, int ny, int nz, int ex, int ey, int ez,
                          mm_conv_simple_t **out);
int mm_conv_simple_execute(mm_conv_simple_t *conv, const double *M, double *H, void *stream);
int mm_conv_simple_destroy(mm_conv_simple_t *conv);

/* ------------------------------------------------------ local field terms */
/* exchange (src/magneto/mmm/exchange/exchange.cpp:31-77; bindings/mmm/exchange.i:24-39).
 * energy (may be NULL) receives -mu0/2 * dV * sum_i m_i.H_i. */
int mm_exchange(int nx, int ny, int nz, double dx, double dy, double dz,
                int px, int py, int pz, const double *Ms, const double *A,
                const double *M, double *H, double *energy, void *stream);
int mm_exchange_f32(int nx, int ny, int nz, double dx, double dy, double dz,
                    int px, int py, int pz, const float *Ms, const float *A,
                    const float *M, float *H, void *stream);
/* uniaxial_anisotropy(axis, k, Ms, M, H) (anisotropy.cpp:33-58;
 * bindings/mmm/anisotropy.i:24-39); energy = sum_i k |a x m|^2 */
int mm_uniaxial_anisotropy(int n, const double *axis, const double *k, const double *Ms,
                           const double *M, double *H, double *energy, void *stream);
/* cubic_anisotropy(axis1, axis2, k, Ms, M, H) (anisotropy.cpp:60-88) */
int mm_cubic_anisotropy(int n, const double *axis1, const double *axis2, const double *k,
                        const double *Ms, const double *M, double *H, double *energy,
                        void *stream);
/* llge(f1, f2, M, H, dM) (src/magneto/mmm/llge/llge.cpp:33-59) */
int mm_llge(int n, const double *f1, const double *f2, const double *M, const double *H,
            double *dM, void *stream);
/* fdm_zhangli (src/magneto/mmm/spintorque/fdm_zhangli.cpp:33-53) */
int mm_fdm_zhangli(int nx, int ny, int nz, double dx, double dy, double dz, int do_precess,
                   const double *P, const double *xi, const double *Ms, const double *alpha,
                   const double *j, const double *M, double *dM, void *stream);

/* ------------------------------------------------------ reductions */
/* scaled_abs_max(M, scale) (src/magneto/math/ScaledAbsMax.cpp:52-68):
 * uniform scale -> absMax(M)/scale[0]; else max over scale!=0 of |M|/scale. */
int mm_scaled_abs_max(int n, const double *M, const double *scale, int scale_uniform,
                      double *out, void *stream);
/* No counterpart in the reference.  mm_scaled_abs_max without allocating:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * result goes to the device pointer out_dev (one double) on `stream`; the
 * call neither allocates nor synchronises.  The per-block partial maxima live
 * in a file-scope device array of 1024 doubles (8 KiB) that each GPU holds
 * until the process ends; any n up to INT_MAX fits it (at most 1024 blocks
 * stride over the cells).  Calls on one GPU share that array: calls that
 * overlap on one GPU race. */
int mm_scaled_abs_max_noalloc(int n, const double *M, const double *scale, int scale_uniform,
                              double *out_dev, void *stream);
/* VectorMatrix::absMax / dotSum / average (VectorMatrix.cpp:187-274) */
int mm_abs_max(int n, const double *M, double *out, void *stream);
int mm_dot_sum(int n, const double *A, const double *B, double *out, void *stream);
int mm_average(int n, const double *M, double *out3, void *stream);
/* VectorMatrix::normalize(Matrix len) (CPUDevice.cpp:244-271).
 * mode 0: reference CPU semantics (|M| from float-cast components, zero
 * vectors kept); mode 1: FP64 norm. */
int mm_normalize(int n, double *M, const double *len, int mode, void *stream);

/* ------------------------------------------------------ Runge-Kutta algebra */
typedef struct {
    int num_steps, order, fsal;
    double a[8], b[8][8], c[8], ec[8];
} mm_tableau_t;
/* ButcherTableau for a method (src/magneto/evolver/runge_kutta.h:31-72). */
int mm_tableau(int method, mm_tableau_t *out);
/* rk_prepare_step (runge_kutta.cpp:32-48): ytmp = y + h*sum_{j<step} b[step][j] k_j.
 * k: host array of num_steps device pointers (entries >= step may be NULL). */
int mm_rk_prepare_step(int step, double h, const mm_tableau_t *tab, int n,
                       const double *const *k, const double *y, double *ytmp, void *stream);
/* rk_combine_result (runge_kutta.cpp:50-96): y += h*sum c_j k_j; y_err = h*sum ec_j k_j */
int mm_rk_combine_result(double h, const mm_tableau_t *tab, int n, const double *const *k,
                         double *y, double *y_err, void *stream);

/* ------------------------------------------------------ fused stepper */
/* Whole LLG system on the device: H_tot = H_stray + H_exch + H_aniso + H_ext
 * (module order of the reference's LandauLifshitzGilbert.calculate_H_tot,
 * landau_lifshitz_gilbert.py:92-101), dM/dt = LLG + Zhang-Li, stepped by an
 * RK method with MicroMagneticsStepSizeController
 * (micro_magnetics_stepsize_controller.py:35-73) and M renormalised to Ms
 * after every accepted step (micro_magnetics.py:49-50).  Accept/reject and
 * the step size stay on the device; one call runs many steps. */
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
    int precision;          /* MM_FP64 or MM_FP32 (FP32 bulk, FP64 step control) */
    int method;             /* MM_RKF45 ... MM_RK4                                */
    double eps_abs, eps_rel;
    double h0;              /* initial / fixed step size [s]                      */
    int normalize_mode;     /* 0 reference float-norm, 1 FP64                     */
    int do_precess;
    /* host arrays of n (scalars) or 3n (vectors, SoA); NULL = term absent
     * (Ms and alpha are required; A NULL = no exchange). */
    const double *Ms_host, *alpha_host, *A_host;
    const double *k_uni_host, *k_cub_host, *axis1_host, *axis2_host;
    const double *P_host, *xi_host, *j_host;
    int use_zeeman;
    double h_ext[3];        /* uniform external field [A/m]                       */
    const double *N_host;   /* demag tensor (6*ex*ey*ez) or NULL = no stray field */
    int ex, ey, ez;
    int px, py, pz;         /* periodic axes (exchange neighbours wrap; the tensor
                               must be generated with the same flags)           */
    /* AlternatingField (alternating_field.py:57-84): with use_ac the Zeeman
     * field is h_ext + h_amp * sin(t * h_freq + h_phase) per component, at the
     * time of each evaluation (RK stage s of an attempt from t with step h:
     * t + a_s h, runge_kutta.py:115).  use_zeeman must be set too. */
    int use_ac;
    double h_amp[3], h_freq[3], h_phase[3];
} mm_system_desc_t;

typedef struct mm_stepper mm_stepper_t;
int mm_stepper_create(const mm_system_desc_t *desc, mm_stepper_t **out);
int mm_stepper_set_m(mm_stepper_t *st, const double *M_host);
int mm_stepper_get_m(mm_stepper_t *st, double *M_host);
/* One evaluation of dM/dt and H_tot at the current M (host outputs, may be NULL). */
int mm_stepper_dmdt(mm_stepper_t *st, double *dMdt_host, double *Htot_host);
/* Run `steps` accepted steps (adaptive methods retry rejected attempts on
 * the device).  Returns when they are done. */
int mm_stepper_run(mm_stepper_t *st, int steps);
/* Run until t reaches t_end or max_steps accepted steps were made, with the
 * reference's overshoot rule (runge_kutta.py:74-78): an accepted step that
 * would pass t_end is redone from the same state with h = t_end - t and
 * taken without error check; the controller's step size resumes after it.
 * This is what solve(Time(t_end)) does step by step. */
int mm_stepper_run_until(mm_stepper_t *st, double t_end, long long max_steps);
/* As mm_stepper_run_until, and also stop after the first accepted step that
 * ends at t >= t_stop, without shortening that step (t_stop = +inf: no such
 * stop).  This is the solver loop of src/magnum/solver/solver.py:110-112 for a
 * stop condition on t that names no time of interest, such as
 * TimeGreaterEq (src/magnum/solver/condition.py:145-157): solve(Time(t))
 * steps with step_with_t_max(1e100) until state.t >= t, so its last step
 * overshoots t.  t_end remains the time of interest that steps are clipped
 * to (step handlers' get_time_of_interest, solver.py:76-82). */
int mm_stepper_run_until_stop(mm_stepper_t *st, double t_end, double t_stop, long long max_steps);
int mm_stepper_sync(mm_stepper_t *st);
/* stats: [0]=t [1]=h of last accepted step [2]=next h [3]=attempts [4]=rejections */
int mm_stepper_stats(mm_stepper_t *st, double *stats5);
/* Accepted steps since creation (the reference's state.step when the
 * stepper started from step 0). */
int mm_stepper_step_count(mm_stepper_t *st, long long *steps);
/* Set the simulation time t and, if h_next > 0, the step size of the next
 * attempt: the state.t and the next trial step that RungeKutta.evolve keeps
 * on the state (src/magnum/evolver/runge_kutta.py:55-58,81-82), for a caller
 * whose state was changed outside the stepper. */
int mm_stepper_set_time(mm_stepper_t *st, double t, double h_next);
/* <M> per component (FP64 sums). */
int mm_stepper_average_m(mm_stepper_t *st, double *avg3);
void *mm_stepper_stream(mm_stepper_t *st);
/* Measurement: enqueue `attempts` attempts with a HIP event after every
 * kernel on the stepper's stream.  out[3*k+0..2] for kernel class k =
 * 0 x forward, 1 y forward, 2 z (forward, multiply, inverse), 3 y inverse,
 * 4 2-D y pass, 5 fused x inverse + stage epilogue, 6 step controller:
 * average ms per launch, launches, average algorithmic HBM bytes per launch. */
enum { MM_KCLASS_COUNT = 7 };
int mm_stepper_kernel_profile(mm_stepper_t *st, int attempts, double *out);
/* Energies at the current M [J] (FP64 sums of FP64 or FP32 products):
 * E5[0] E_stray = -mu0/2 V sum M.H_stray   (stray_field.py:51)
 * E5[1] E_exch  = -mu0/2 V sum M.H_exch    (exchange_field.py:58)
 * E5[2] E_aniso = V sum of the uniaxial (k |a x m|^2) and cubic energy
 *                 densities                (anisotropy_field.py:84-87,
 *                 anisotropy_cpu.cpp:64-65,114; with both terms present the
 *                 reference's E_aniso is None, here it is their sum)
 * E5[3] E_ext   = -mu0 V sum M.H_ext       (external_field.py:33-35)
 * E5[4] E_tot   = sum of the present terms (landau_lifshitz_gilbert.py:103-106)
 * with mu0 = 4 pi 1e-7 as the Python front-end uses it (constants.py:21). */
int mm_stepper_energy(mm_stepper_t *st, double *E5);
/* Relaxation: run until the rule holds on an accepted state, or until
 * max_steps more accepted steps.  The rule is checked once per accepted
 * state whose step count (accepted steps since creation, as state.step) is
 * a multiple of check_every, from the k_0 = dM/dt of the attempt that
 * starts there (for FSAL methods the carried-over k_0); a state that
 * satisfies it is kept and that attempt is discarded.
 *   MM_RELAX_DMDT:       scaled_abs_max(dM/dt, Ms) < threshold [1/s]
 *                        (1e7 /s = 0.01 rad/ns);
 *   MM_RELAX_DEG_PER_NS: the reference's Relaxed(threshold, check_every)
 *                        (condition.py:169-200): deg_per_ns =
 *                        1e-9 (180/pi) atan2(|dM/dt|max h, |M|max) / h
 *                        (landau_lifshitz_gilbert.py:127-131, h = last
 *                        accepted step) <= threshold and below the previous
 *                        check's value.
 * info3: [0] accepted steps made, [1] 1 if relaxed else 0, [2] value of the
 * last check. */
enum { MM_RELAX_DMDT = 1, MM_RELAX_DEG_PER_NS = 2 };
int mm_stepper_relax(mm_stepper_t *st, int criterion, double threshold, int check_every,
                     long long max_steps, double *info3);
/* ExternalField / AlternatingField with a user function `H_ext_fn`
 * (src/magnum/micromagnetics/alternating_field.py:60-75, A = fn(t) returned
 * as a uniform 3-vector): fn is called on the host with every evaluation
 * time -- the stage times t + a_s h of each attempt, and t for
 * mm_stepper_dmdt and mm_stepper_energy -- and writes the field [A/m] into
 * h3; it takes the place of h_ext.  The description must then have no other
 * Zeeman parameters (use_zeeman = 0): the reference raises ValueError when
 * offs/amp/freq/phase are set beside fn (alternating_field.py:68-69).  While
 * a function is set each attempt waits for the previous one, since its stage
 * times depend on the step size the device chose.  fn = NULL removes it.
 * Ensembles (mm_batch_*) do not take a function. */
typedef void (*mm_field_fn_t)(double t, double *h3, void *user);
int mm_stepper_set_field_fn(mm_stepper_t *st, mm_field_fn_t fn, void *user);
int mm_stepper_destroy(mm_stepper_t *st);

/* ---------------------------------------------- launch shapes ("wisdom") */
/* The reference keeps FFTW's tuned plans ("wisdom") in a file that
 * magneto.initialize/deinitialize import and export
 * (src/magneto/Magneto.cpp:151-190).  Here the tuned choices are launch
 * shapes: x transform lines per workgroup, kx columns per workgroup of the y
 * and z passes, rows per stage workgroup and threads of the stage launch
 * (0 = the built-in choice).  A process-wide table keyed by mesh size,
 * precision and device holds them; convolutions and steppers created later
 * for that key use the entry (the MM_LPW_X, MM_TK_Y, MM_TK_Z, MM_RPW and
 * MM_STAGE_THREADS environment variables take precedence).  Launch shapes do
 * not change results: every shape computes the same values. */
typedef struct {
    int lpw_x, tk_y, tk_z, rpw, stage_threads;
} mm_launch_shape_t;
/* the shape a stepper runs with */
int mm_stepper_launch_shape(mm_stepper_t *st, mm_launch_shape_t *out);
/* Measure `steps` accepted steps of steppers for desc with the built-in shape
 * and with variations of one knob at a time (coordinate descent), record the
 * fastest in the table and return it with its time per step [ms].  Takes
 * seconds: one stepper set-up per candidate. */
int mm_autotune(const mm_system_desc_t *desc, int steps, mm_launch_shape_t *best, double *ms_per_step);
/* Text file: '#' comments, then one entry per line,
 * "<nx>x<ny>x<nz> <fp32|fp64> <arch>/<CUs> lpw_x tk_y tk_z rpw stage_threads".
 * Import adds (replacing equal keys) all entries of a valid file and nothing
 * from an invalid one (the reference logs an invalid FFTW wisdom file and
 * goes on without it). */
int mm_wisdom_import(const char *path, int *n_entries);
int mm_wisdom_export(const char *path);
int mm_wisdom_forget(void);

/* ------------------------------------------------------ ensembles */
/* Many independent systems (each a full mm_system_desc_t: own mesh, material,
 * method, precision) stepped together: the batched variant of the fused
 * stepper.  The reference steps one system per process; an ensemble of its
 * scripts maps to one mm_batch.  Systems are spread over `nstreams` shared
 * streams (<= 0: 4), each keeps its own step control, and mm_batch_run makes
 * `steps` accepted steps in every system.  Results per system equal those of
 * mm_stepper_* on the same description. */
typedef struct mm_batch mm_batch_t;
int mm_batch_create(const mm_system_desc_t *descs, int count, int nstreams, mm_batch_t **out);
int mm_batch_set_m(mm_batch_t *b, int index, const double *M_host);
int mm_batch_get_m(mm_batch_t *b, int index, double *M_host);
int mm_batch_run(mm_batch_t *b, int steps);
int mm_batch_stats(mm_batch_t *b, int index, double *stats5);
int mm_batch_average_m(mm_batch_t *b, int index, double *avg3);
int mm_batch_destroy(mm_batch_t *b);

#ifdef __cplusplus
}
#endif
#endif
