#!/usr/bin/env python3
"""LLG Runge-Kutta throughput on one MI355X.

Headline workload (the reference's FP64-vs-mixed block case): 256x256x16
cells of 2x2x2 nm, open boundaries, Ms = 1.1e6 A/m, A = 1.5e-11 J/m,
uniaxial K = 5e5 J/m^3 along z, uniform Zeeman field of 20 mT along x, alpha
= 0.02, seeded random unit magnetisation (PCG64 seed 0x5EED), RKF45 with
eps_abs = eps_rel = 1e-4.  Terms: demag (FFT convolution) + exchange +
uniaxial + Zeeman.

One step = one accepted RKF45 step of the whole grid (rejected attempts are
counted and reported as attempted steps).  `value` is the mixed-precision
path (FP32 bulk, FP64 step control); the plain FP64 path is run on the same
workload and reported next to it.  Inputs are resident in HBM before timing.

roofline: the dominant kernel class (largest share of GPU time), algorithmic
bytes per launch / average launch time from HIP events recorded on the
stepper's stream around every kernel.  The step-level figure of the canonical
byte model (6 B_conv + 36 V per RKF45 attempt) is reported beside it.
cpu_baseline: the CPU oracle (the FP64 restatement of the reference CPU path)
on a bounded sample of the same workload, on this host's cores.

errors: field, energy and <m>/M after a step sample, of both GPU paths
against the FP64 oracle on the same inputs.

workloads: the other four workloads of the benchmark set, each timed on a
bounded sample (thin film RK4, spin-torque film DP54 with its vortex-core
trajectory, the 512-system ensemble, the 512x512x64 grid), with their own
errors against the oracle or the FP64 path.  --no-extra skips them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LLG RK steps/s & Mcell-updates/s, FP64 vs mixed-FP32, % HBM roofline"
HBM_PEAK = 8.0e12
MU0 = 4.0e-7 * np.pi  # src/magnum/micromagnetics/constants.py:21

SHAPE = (256, 256, 16)
DELTA = (2e-9, 2e-9, 2e-9)
PARAMS = dict(Ms=1.1e6, alpha=0.02, A=1.5e-11, k_uni=5e5, axis1=(0.0, 0.0, 1.0),
              h_ext=(0.02 / MU0, 0.0, 0.0))
TOL = 1e-4

# thin film: demag + exchange, RK4 h = 1e-13 s, 1000 steps, <m> every 10
FILM_SHAPE, FILM_DELTA = (128, 32, 1), (4e-9, 4e-9, 3e-9)
FILM_PARAMS = dict(Ms=8e5, alpha=0.5, A=1.3e-11)
# spin-torque film: Zhang-Li current along x, vortex with a Gaussian core
# (width 10 nm) plus 1 % noise, DP54 tol 1e-5
STT_SHAPE, STT_DELTA = (1024, 1024, 2), (3e-9, 3e-9, 2e-9)
STT_PARAMS = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, P=0.7, xi=0.05, j=(1e12, 0.0, 0.0))
STT_TOL = 1e-5
# grid: permalloy with cubic anisotropy, alpha 1, relaxation from random m
GRID_SHAPE, GRID_DELTA = (512, 512, 64), (2e-9, 2e-9, 2e-9)
GRID_PARAMS = dict(Ms=8e5, alpha=1.0, A=1.3e-11, k_cub=-5e3, axis1=(1.0, 0.0, 0.0), axis2=(0.0, 1.0, 0.0))
# relaxation stop of the grid workload: max |dm/dt| < 1e-2 rad/ns
GRID_RELAX = 1e7


def round4(n):
    """the reference's ROUND_4 padding of an open axis (tensor_round.cpp:90-106)"""
    p = 2 * n - 1
    return p if n == 1 else (p + 3) // 4 * 4


def model_bytes(shape, b, method):
    """Canonical algorithmic bytes of one field evaluation (B_conv) and of one
    attempt of `method` (ROUND_4 padding whatever the build pads to)."""
    nx, ny, nz = shape
    Px, Py, Pz = round4(nx), round4(ny), round4(nz)
    c = 2 * b
    N = nx * ny * nz
    Hx = Px // 2 + 1
    B_conv = 3 * (b * N + 4 * c * Hx * ny * nz + 4 * c * Hx * Py * nz * (1 if Pz > 1 else 0)) \
        + 6 * b * Hx * (Py // 2 + 1) * (Pz // 2 + 1)
    V = 3 * b * N
    step = {"rkf45": 6 * B_conv + 36 * V, "dp54": 6 * B_conv + 45 * V, "rk4": 4 * B_conv + 17 * V}[method]
    return B_conv, step


def canonical_bytes(shape, P, b):
    """B_conv and the per-attempt RKF45 bytes of the canonical model (P: the
    FFT size, only its x extent enters through Hx)."""
    nx, ny, nz = shape
    Px, Py, Pz = P
    c = 2 * b
    N = nx * ny * nz
    Hx = Px // 2 + 1
    B_conv = 3 * (b * N + 4 * c * Hx * ny * nz + 4 * c * Hx * Py * nz * (1 if Pz > 1 else 0)) \
        + 6 * b * Hx * (Py // 2 + 1) * (Pz // 2 + 1)
    V = 3 * b * N
    return B_conv, 6 * B_conv + 36 * V


def initial_m(shape, Ms, seed=0x5EED):
    nx, ny, nz = shape
    rng = np.random.Generator(np.random.PCG64(seed))
    m = rng.standard_normal((3, nz, ny, nx))
    m /= np.sqrt((m ** 2).sum(axis=0))
    return m * Ms


def stt_initial_m():
    """vortex (in-plane curl, core +z, Gaussian width 10 nm) at the film's
    centre plus 1 % seeded noise, renormalised"""
    from micromagnum_amd.vortex import gaussian_core_vortex
    Ms = STT_PARAMS["Ms"]
    centre = (STT_SHAPE[0] * STT_DELTA[0] / 2, STT_SHAPE[1] * STT_DELTA[1] / 2)
    M = gaussian_core_vortex(STT_SHAPE, STT_DELTA, centre, 10e-9, Ms)
    rng = np.random.Generator(np.random.PCG64(0x5EED))
    M = M + 0.01 * Ms * rng.standard_normal(M.shape)
    return M * (Ms / np.sqrt((M ** 2).sum(axis=0)))


def ensemble_specs(count=512, precision="fp32"):
    """The ensemble workload's seeded draw (PCG64 0x5EED): nx, ny ~ U{32..256},
    nz ~ U{1..4}, Ms ~ U[4e5, 1.4e6], A ~ U[5e-12, 2e-11], K ~ U[0, 1e6]
    along a random axis, cells U[2, 5] nm, alpha 0.5, RKF45; and the seeded
    random initial states."""
    rng = np.random.Generator(np.random.PCG64(0x5EED))
    nx = rng.integers(32, 257, size=count)
    ny = rng.integers(32, 257, size=count)
    nz = rng.integers(1, 5, size=count)
    Ms = rng.uniform(4e5, 1.4e6, count)
    A = rng.uniform(5e-12, 2e-11, count)
    K = rng.uniform(0.0, 1e6, count)
    ax = rng.standard_normal((count, 3))
    ax /= np.linalg.norm(ax, axis=1)[:, None]
    d = rng.uniform(2e-9, 5e-9, (count, 3))
    specs, M0 = [], []
    for i in range(count):
        shape = (int(nx[i]), int(ny[i]), int(nz[i]))
        specs.append(dict(shape=shape, delta=tuple(float(v) for v in d[i]), Ms=float(Ms[i]), alpha=0.5,
                          A=float(A[i]), k_uni=float(K[i]), axis1=tuple(float(v) for v in ax[i]),
                          method="rkf45", eps_abs=TOL, eps_rel=TOL, precision=precision, tensor="gpu"))
        m = rng.standard_normal((3, shape[2], shape[1], shape[0]))
        m /= np.sqrt((m ** 2).sum(axis=0))
        M0.append(m * Ms[i])
    return specs, M0


def host_threads():
    """threads for the CPU oracle: the host's CPU share (OMP_NUM_THREADS where
    the environment sets it, else every core this process may run on)"""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(aff, int(env))) if env and env.isdigit() else aff, aff


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def oracle_mod():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def run_gpu(precision, N, M0, steps, warmup):
    from micromagnum_amd.solver import Stepper
    st = Stepper(SHAPE, DELTA, N=N, precision=precision, method="rkf45", eps_abs=TOL, eps_rel=TOL, **PARAMS)
    st.set_m(M0)
    st.run(warmup)
    st.sync()
    s0 = st.stats()
    t0 = time.perf_counter()
    st.run(steps)
    st.sync()
    t1 = time.perf_counter()
    s1 = st.stats()
    prof = st.kernel_profile(attempts=6)
    st.set_m(M0)
    dM, H = st.dmdt()
    E = st.energy()
    res = dict(seconds=t1 - t0, attempts=s1["attempts"] - s0["attempts"],
               rejections=s1["rejections"] - s0["rejections"], sim_time=s1["t"] - s0["t"],
               h=s1["h"], profile=prof, H=H, E=E)
    st.close()
    return res


def gpu_sample(precision, N, M0, steps):
    """M and stats after `steps` RKF45 steps from M0 (the CPU sample's run)"""
    from micromagnum_amd.solver import Stepper
    st = Stepper(SHAPE, DELTA, N=N, precision=precision, method="rkf45", eps_abs=TOL, eps_rel=TOL, h0=1e-13,
                 **PARAMS)
    st.set_m(M0)
    st.run(steps)
    M, s = st.get_m(), st.stats()
    st.close()
    return M, s


def cpu_baseline(N, M0, steps, threads):
    O = oracle_mod()
    O.set_threads(threads)
    sysO = O.System(SHAPE, DELTA, N=N, **PARAMS)
    t0 = time.perf_counter()
    M, st = sysO.run(M0, method="rkf45", steps=steps, eps_abs=TOL, eps_rel=TOL, h0=1e-13)
    dt = time.perf_counter() - t0
    return dict(seconds=dt, steps=steps, attempts=st["attempts"], M=M, stats=st)


def calibration_copies(nbytes=256 << 20, reps=5):
    """k_copy_calib over nbytes, reps times: reads and writes exactly nbytes each,
    the known transfer against which FETCH_SIZE / WRITE_SIZE are scaled."""
    import ctypes
    from micromagnum_amd import _lib
    from micromagnum_amd.device import empty
    a = empty((nbytes // 8,))
    b = empty((nbytes // 8,))
    L = _lib.device_lib()
    for _ in range(reps):
        _lib.check(L.mm_device_copy(b.ptr, a.ptr, ctypes.c_size_t(nbytes), None))
    _lib.check(L.mm_synchronize(None))


def load_traffic(kernel_class, precision):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_block256x256x16_%s.json" % (kernel_class, precision))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


# ------------------------------------------------------------------ workloads
def bench_thin_film(threads):
    """RK4 h = 1e-13 s, 1000 steps, <m> read every 10 steps (the reads are
    part of the timed work), FP64 and FP32; <m> trajectory against the FP64
    oracle's, which is also the CPU sample (whole run, all host threads)."""
    from micromagnum_amd import magneto as mg
    from micromagnum_amd.solver import Stepper
    N = mg.GenerateDemagTensor(*FILM_SHAPE, *FILM_DELTA)
    M0 = initial_m(FILM_SHAPE, FILM_PARAMS["Ms"])
    Ms = FILM_PARAMS["Ms"]
    n = int(np.prod(FILM_SHAPE))
    out, avg = {}, {}
    for prec in ("fp64", "fp32"):
        st = Stepper(FILM_SHAPE, FILM_DELTA, N=N, precision=prec, method="rk4", h0=1e-13, **FILM_PARAMS)
        st.set_m(M0)
        st.run(10)  # warm-up: graph capture
        st.set_m(M0)
        st.set_time(0.0)
        st.sync()
        a = []
        t0 = time.perf_counter()
        for _ in range(100):
            st.run(10)
            a.append(st.average_m())
        dt = time.perf_counter() - t0
        st.close()
        avg[prec] = np.array(a)
        _, b = model_bytes(FILM_SHAPE, 8 if prec == "fp64" else 4, "rk4")
        out[prec] = {"steps_per_s": round(1000 / dt, 1), "mcell_updates_per_s": round(1000 * n / dt / 1e6, 2),
                     "step_model_roofline_frac": round(b * 1000 / dt / HBM_PEAK, 5)}
    O = oracle_mod()
    O.set_threads(threads)
    t0 = time.perf_counter()
    _, so = O.System(FILM_SHAPE, FILM_DELTA, N=N, **FILM_PARAMS).run(M0, method="rk4", steps=1000, h0=1e-13,
                                                                       record_avg=True)
    dt = time.perf_counter() - t0
    ref = so["avg"][9::10]
    out["m_traj_err_fp64_vs_oracle"] = float(np.abs(avg["fp64"] - ref).max() / Ms)
    out["m_traj_err_fp32_vs_oracle"] = float(np.abs(avg["fp32"] - ref).max() / Ms)
    out["cpu_oracle"] = {"steps_per_s": round(1000 / dt, 2), "cores": threads, "sample": "the whole 1000-step run"}
    return out


def bench_stt_film(threads, steps):
    """DP54 tol 1e-5 from the vortex + noise state: `steps` timed steps after
    60 warm-up steps (the noise is damped first), FP32 and FP64; the core
    trajectory at common times (run_until) in both precisions; the field of
    the initial state against the oracle."""
    from micromagnum_amd.solver import Stepper
    from micromagnum_amd.vortex import find_core
    from micromagnum_amd import magneto as mg
    N = mg.GenerateDemagTensor(*STT_SHAPE, *STT_DELTA, device="gpu")
    M0 = stt_initial_m()
    n = int(np.prod(STT_SHAPE))
    out, cores, H = {}, {}, {}
    times = [k * 2e-12 for k in range(1, 6)]
    for prec in ("fp32", "fp64"):
        st = Stepper(STT_SHAPE, STT_DELTA, N=N, precision=prec, method="dp54", eps_abs=STT_TOL, eps_rel=STT_TOL,
                     **STT_PARAMS)
        st.set_m(M0)
        H[prec] = st.dmdt()[1]
        st.run(60)
        st.sync()
        s0 = st.stats()
        t0 = time.perf_counter()
        st.run(steps)
        st.sync()
        dt = time.perf_counter() - t0
        s1 = st.stats()
        att = s1["attempts"] - s0["attempts"]
        _, b = model_bytes(STT_SHAPE, 8 if prec == "fp64" else 4, "dp54")
        out[prec] = {"steps_per_s": round(steps / dt, 1), "attempted_steps_per_s": round(att / dt, 1),
                     "mcell_updates_per_s": round(steps * n / dt / 1e6, 1),
                     "step_model_roofline_frac": round(b * att / dt / HBM_PEAK, 4)}
        # core trajectory: from M0 to t = 2, 4, ..., 10 ps
        st.set_m(M0)
        st.set_time(0.0, 1e-13)
        c = []
        for t in times:
            st.run_until(t, max_steps=100000)
            c.append(find_core(st.get_m(), STT_DELTA))
        cores[prec] = np.array(c)
        st.close()
    O = oracle_mod()
    O.set_threads(threads)
    Ho = O.System(STT_SHAPE, STT_DELTA, N=N, **STT_PARAMS).h_tot(M0)
    out["field_err_fp64_vs_oracle"] = rel(H["fp64"], Ho)
    out["field_err_fp32_vs_oracle"] = rel(H["fp32"], Ho)
    out["core_trajectory_nm"] = {"t_ps": [round(t * 1e12, 3) for t in times],
                                 "fp64": (cores["fp64"] * 1e9).round(4).tolist(),
                                 "fp32": (cores["fp32"] * 1e9).round(4).tolist()}
    out["core_err_fp32_vs_fp64_nm"] = float(np.abs(cores["fp32"] - cores["fp64"]).max() * 1e9)
    return out


def bench_grid(steps):
    """RKF45 from seeded random m, `steps` timed steps after 5 warm-up steps
    in each precision; the relaxation rule (max |dm/dt| < 1e-2 rad/ns) is
    evaluated on the device at the end of the sample, and the energies of the
    two precisions are compared at the same step."""
    from micromagnum_amd.solver import Stepper
    from micromagnum_amd import magneto as mg
    N = mg.GenerateDemagTensor(*GRID_SHAPE, *GRID_DELTA, device="gpu")
    M0 = initial_m(GRID_SHAPE, GRID_PARAMS["Ms"])
    n = int(np.prod(GRID_SHAPE))
    out, E = {}, {}
    for prec in ("fp32", "fp64"):
        st = Stepper(GRID_SHAPE, GRID_DELTA, N=N, precision=prec, method="rkf45", eps_abs=TOL, eps_rel=TOL,
                     **GRID_PARAMS)
        st.set_m(M0)
        st.run(5)
        st.sync()
        s0 = st.stats()
        t0 = time.perf_counter()
        st.run(steps)
        st.sync()
        dt = time.perf_counter() - t0
        s1 = st.stats()
        att = s1["attempts"] - s0["attempts"]
        dM, _ = st.dmdt()
        E[prec] = st.energy()["E_tot"]
        st.close()
        _, b = model_bytes(GRID_SHAPE, 8 if prec == "fp64" else 4, "rkf45")
        dmax = float(np.sqrt((dM ** 2).sum(axis=0)).max() / GRID_PARAMS["Ms"])
        out[prec] = {"steps_per_s": round(steps / dt, 2), "attempted_steps_per_s": round(att / dt, 2),
                     "mcell_updates_per_s": round(steps * n / dt / 1e6, 1),
                     "step_model_roofline_frac": round(b * att / dt / HBM_PEAK, 4),
                     "max_dmdt_over_Ms_per_s": dmax, "E_tot": E[prec], "relaxed": dmax < GRID_RELAX}
    out["energy_err_fp32_vs_fp64"] = abs(E["fp32"] - E["fp64"]) / abs(E["fp64"])
    out["note"] = ("relaxation to 1e-2 rad/ns takes far more steps than a bench sample; the sample's "
                   "max |dm/dt| / Ms is reported")
    return out


def bench_ensemble(steps):
    """The 512-system draw, FP32 with FP64 step control, `steps` timed RKF45
    steps of every system after 5 warm-up steps."""
    from micromagnum_amd.solver import Ensemble
    specs, M0 = ensemble_specs()
    cells = sum(int(np.prod(s["shape"])) for s in specs)
    t0 = time.perf_counter()
    ens = Ensemble(specs)
    setup = time.perf_counter() - t0
    for i, M in enumerate(M0):
        ens.set_m(i, M)
    ens.run(5)
    a0 = sum(ens.stats(i)["attempts"] for i in range(len(specs)))
    t0 = time.perf_counter()
    ens.run(steps)
    dt = time.perf_counter() - t0
    att = [ens.stats(i)["attempts"] for i in range(len(specs))]
    ens.close()
    per_attempt = [model_bytes(s["shape"], 4, "rkf45")[1] for s in specs]
    # model bytes of the attempts made in the timed region
    a1 = sum(att)
    bytes_timed = sum(per_attempt) * (a1 - a0) / len(specs)
    return {"systems": len(specs), "cells": cells, "ensemble_steps_per_s": round(steps / dt, 2),
            "mcell_updates_per_s": round(steps * cells / dt / 1e6, 1),
            "step_model_bytes": int(sum(per_attempt)),
            "step_model_roofline_frac": round(bytes_timed / dt / HBM_PEAK, 4), "setup_s": round(setup, 1)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=16, help="RKF45 steps in the CPU-baseline sample")
    ap.add_argument("--cpu-steps-1t", type=int, default=2, help="RKF45 steps in the single-thread CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the host's CPU share (see host_threads)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other workloads")
    ap.add_argument("--calib", action="store_true",
                    help="also run 5 copies of 256 MiB (k_copy_calib) for hardware-counter calibration")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the per-GPU kernels would run unchanged on each GPU's shard); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from micromagnum_amd.magneto import GenerateDemagTensor
    from micromagnum_amd.device import gpu_available
    if not gpu_available():
        print("bench.py needs a HIP GPU and the built libmmhip.so", file=sys.stderr)
        sys.exit(1)

    if args.calib:
        calibration_copies()

    N = GenerateDemagTensor(*SHAPE, *DELTA)
    P = (N.shape[2], N.shape[1], N.shape[0])
    M0 = initial_m(SHAPE, PARAMS["Ms"])
    ncell = SHAPE[0] * SHAPE[1] * SHAPE[2]
    threads, aff = host_threads()
    if args.cpu_threads:
        threads = args.cpu_threads

    r32 = run_gpu("fp32", N, M0, args.steps, args.warmup)
    r64 = None if args.no_fp64 else run_gpu("fp64", N, M0, args.steps, args.warmup)

    steps_s = args.steps / r32["seconds"]
    attempts_s = r32["attempts"] / r32["seconds"]
    Bc32, Bstep32 = canonical_bytes(SHAPE, P, 4)
    Bc64, Bstep64 = canonical_bytes(SHAPE, P, 8)

    prof = r32["profile"]
    dom = max(prof.items(), key=lambda kv: kv[1]["ms"] * kv[1]["launches"])
    dom_name, dom_p = dom
    achieved = dom_p["bytes"] / (dom_p["ms"] * 1e-3)
    gpu_time_attempt = sum(v["ms"] * v["launches"] for v in prof.values()) / 6.0

    out = {
        "metric": METRIC,
        "value": round(steps_s, 3),
        "unit": "RKF45 steps/s (accepted, mixed FP32/FP64, 256x256x16 cells)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r32["seconds"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk / f64 step control",
        "data": "synthetic (seeded random unit magnetisation, PCG64 0x5EED)",
        "config": {"workload": "block 256x256x16 cells of 2 nm, demag+exchange+uniaxial+Zeeman, "
                               "RKF45 eps 1e-4", "cells": ncell, "fft_size": list(P),
                   "parallelism": "single GPU"},
        "attempted_steps_per_s": round(attempts_s, 3),
        "rejections": r32["rejections"],
        "mcell_updates_per_s": round(steps_s * ncell / 1e6, 1),
        "step_model_bytes": Bstep32,
        "step_model_roofline_frac": round(Bstep32 * attempts_s / HBM_PEAK, 4),
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": load_traffic(dom_name, "fp32"),
                     "bytes_per_launch": dom_p["bytes"], "avg_launch_ms": round(dom_p["ms"], 5)},
        "kernels_fp32": {k: {"avg_ms": round(v["ms"], 5), "launches_per_attempt": v["launches"] / 6.0,
                             "GB_s": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["bytes"] else None}
                         for k, v in prof.items()},
        "gpu_busy_ms_per_attempt": round(gpu_time_attempt, 4),
    }
    if r64 is not None:
        s64 = args.steps / r64["seconds"]
        a64 = r64["attempts"] / r64["seconds"]
        out["fp64"] = {"steps_per_s": round(s64, 3), "attempted_steps_per_s": round(a64, 3),
                       "mcell_updates_per_s": round(s64 * ncell / 1e6, 1),
                       "step_model_roofline_frac": round(Bstep64 * a64 / HBM_PEAK, 4),
                       "kernels": {k: {"avg_ms": round(v["ms"], 5), "launches_per_attempt": v["launches"] / 6.0}
                                   for k, v in r64["profile"].items()}}
        out["fp32_over_fp64"] = round(steps_s / s64, 3)

    if not args.no_cpu_baseline:
        O = oracle_mod()
        O.set_threads(threads)
        sysO = O.System(SHAPE, DELTA, N=N, **PARAMS)
        dMo, Ho = sysO.dmdt(M0)
        Eo = sysO.energy(M0)
        cb = cpu_baseline(N, M0, args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": round(cb["steps"] / cb["seconds"], 5),
                               "unit": "RKF45 steps/s (accepted)", "cores": threads, "kind": "port",
                               "nproc": os.cpu_count(), "cpu_share": aff,
                               "sample": "%d RKF45 steps (%d attempts) of the same 256x256x16 workload, "
                                         "FP64 oracle, %.1f s" % (cb["steps"], cb["attempts"], cb["seconds"])}
        if args.cpu_steps_1t > 0:
            # the reference's default single-thread configuration
            c1 = cpu_baseline(N, M0, args.cpu_steps_1t, 1)
            out["cpu_baseline"]["single_thread"] = {
                "value": round(c1["steps"] / c1["seconds"], 5), "cores": 1,
                "sample": "%d RKF45 steps (%d attempts), FP64 oracle, %.1f s" % (c1["steps"], c1["attempts"],
                                                                                c1["seconds"])}
        out["gpu_over_cpu"] = round(steps_s / out["cpu_baseline"]["value"], 1)
        # errors against the FP64 oracle: H_tot and energies at the initial
        # state; M and <m> after the CPU sample's steps
        err = {}
        Ms = PARAMS["Ms"]
        for prec, r in (("fp32", r32), ("fp64", r64)):
            if r is None:
                continue
            Mg, sg = gpu_sample(prec, N, M0, args.cpu_steps)
            e_terms = {k: abs(r["E"][k] - Eo[k]) / abs(Eo[k]) for k in Eo if Eo[k] != 0.0}
            err[prec] = {"field": rel(r["H"], Ho), "energy_max_term": max(e_terms.values()),
                         "energy_tot": abs(r["E"]["E_tot"] - Eo["E_tot"]) / abs(Eo["E_tot"]),
                         "m_after_%d_steps_max_cell" % args.cpu_steps: float(np.abs(Mg - cb["M"]).max() / Ms),
                         "avg_m_after_%d_steps" % args.cpu_steps:
                             float(np.abs(Mg.reshape(3, -1).mean(axis=1) - cb["M"].reshape(3, -1).mean(axis=1)).max()
                                   / Ms),
                         "same_attempts": sg["attempts"] == cb["stats"]["attempts"]}
        out["errors_vs_fp64_oracle"] = err
        out["field_err_fp32_vs_fp64"] = rel(r32["H"], r64["H"]) if r64 is not None else None

    if not args.no_extra:
        wl = {}
        for name, fn in (("thin_film_128x32x1_rk4", lambda: bench_thin_film(threads)),
                         ("stt_film_1024x1024x2_dp54", lambda: bench_stt_film(threads, 200)),
                         ("grid_512x512x64_rkf45", lambda: bench_grid(20)),
                         ("ensemble_512_rkf45", lambda: bench_ensemble(20))):
            t0 = time.perf_counter()
            try:
                wl[name] = fn()
            except Exception as e:  # one workload's failure does not hide the others
                wl[name] = {"error": "%s: %s" % (type(e).__name__, e)}
            wl[name]["wall_s"] = round(time.perf_counter() - t0, 1)
            print("bench: %s done in %.1f s" % (name, wl[name]["wall_s"]), file=sys.stderr, flush=True)
        out["workloads"] = wl
    print(json.dumps(out))


if __name__ == "__main__":
    main()
