#!/usr/bin/env python3
"""LLG Runge-Kutta throughput on one MI355X.

Workload (the reference's FP64-vs-mixed block case): 256x256x16 cells of
2x2x2 nm, open boundaries, Ms = 1.1e6 A/m, A = 1.5e-11 J/m, uniaxial K = 5e5
J/m^3 along z, uniform Zeeman field of 20 mT along x, alpha = 0.02, seeded
random unit magnetisation (PCG64 seed 0x5EED), RKF45 with eps_abs = eps_rel =
1e-4.  Terms: demag (FFT convolution) + exchange + uniaxial + Zeeman.

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


def canonical_bytes(shape, P, b):
    """B_conv and the per-attempt RKF45 bytes of the canonical model."""
    nx, ny, nz = shape
    Px, Py, Pz = P
    c = 2 * b
    N = nx * ny * nz
    Hx = Px // 2 + 1
    B_conv = 3 * (b * N + 4 * c * Hx * ny * nz + 4 * c * Hx * Py * nz * (1 if Pz > 1 else 0)) \
        + 6 * b * Hx * (Py // 2 + 1) * (Pz // 2 + 1)
    V = 3 * b * N
    return B_conv, 6 * B_conv + 36 * V


def initial_m(shape, Ms):
    nx, ny, nz = shape
    rng = np.random.Generator(np.random.PCG64(0x5EED))
    m = rng.standard_normal((3, nz, ny, nx))
    m /= np.sqrt((m ** 2).sum(axis=0))
    return m * Ms


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
    res = dict(seconds=t1 - t0, attempts=s1["attempts"] - s0["attempts"],
               rejections=s1["rejections"] - s0["rejections"], sim_time=s1["t"] - s0["t"],
               h=s1["h"], profile=prof, H=H)
    st.close()
    return res


def cpu_baseline(N, M0, steps, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_threads(threads)
    sysO = O.System(SHAPE, DELTA, N=N, **PARAMS)
    M = M0
    t0 = time.perf_counter()
    M, st = sysO.run(M, method="rkf45", steps=steps, eps_abs=TOL, eps_rel=TOL, h0=1e-13)
    dt = time.perf_counter() - t0
    return dict(seconds=dt, steps=steps, attempts=st["attempts"])


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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=16, help="RKF45 steps in the CPU-baseline sample")
    ap.add_argument("--cpu-steps-1t", type=int, default=2, help="RKF45 steps in the single-thread CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host cores (capped at 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
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
                       "step_model_roofline_frac": round(Bstep64 * a64 / HBM_PEAK, 4)}
        out["fp32_over_fp64"] = round(steps_s / s64, 3)
        H32, H64 = r32["H"], r64["H"]
        # H_tot of the initial state, FP32 path against the FP64 path
        out["field_err_fp32_vs_fp64"] = float(np.linalg.norm(H32 - H64) / np.linalg.norm(H64))

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cb = cpu_baseline(N, M0, args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": round(cb["steps"] / cb["seconds"], 5),
                               "unit": "RKF45 steps/s (accepted)", "cores": threads, "kind": "port",
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
    print(json.dumps(out))


if __name__ == "__main__":
    main()
