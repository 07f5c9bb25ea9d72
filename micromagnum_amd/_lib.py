"""ctypes binding of libmmhip.so (the C ABI declared in include/mm_hotpath.h).

The shared library is built in-tree by ``micromagnum_amd.build.build()``.
Nothing here falls back to another implementation: a missing library or a
missing GPU raises.  Device memory comes from the library itself
(micromagnum_amd.device), so a process holds a single HIP runtime.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmmhip.so")

MM_OK = 0
MM_FP64, MM_FP32 = 0, 1
MM_RKF45, MM_CC45, MM_DP54, MM_RK23, MM_RK4, MM_EULER = 0, 1, 2, 3, 4, 5
METHODS = {"rkf45": MM_RKF45, "cc45": MM_CC45, "dp54": MM_DP54, "rk23": MM_RK23, "rk4": MM_RK4,
           "euler": MM_EULER}
MM_RELAX_DMDT, MM_RELAX_DEG_PER_NS = 1, 2

PADDING_DISABLE = 0
PADDING_ROUND_2 = 1
PADDING_ROUND_4 = 2
PADDING_ROUND_8 = 3
PADDING_ROUND_POT = 4
PADDING_SMALL_PRIME_FACTORS = 5

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int
_d = ctypes.c_double


class Tableau(ctypes.Structure):
    _fields_ = [("num_steps", _i), ("order", _i), ("fsal", _i),
                ("a", _d * 8), ("b", (_d * 8) * 8), ("c", _d * 8), ("ec", _d * 8)]


class SystemDesc(ctypes.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i),
                ("dx", _d), ("dy", _d), ("dz", _d),
                ("precision", _i), ("method", _i),
                ("eps_abs", _d), ("eps_rel", _d), ("h0", _d),
                ("normalize_mode", _i), ("do_precess", _i),
                ("Ms_host", _dp), ("alpha_host", _dp), ("A_host", _dp),
                ("k_uni_host", _dp), ("k_cub_host", _dp), ("axis1_host", _dp), ("axis2_host", _dp),
                ("P_host", _dp), ("xi_host", _dp), ("j_host", _dp),
                ("use_zeeman", _i), ("h_ext", _d * 3),
                ("N_host", _dp), ("ex", _i), ("ey", _i), ("ez", _i),
                ("px", _i), ("py", _i), ("pz", _i),
                ("use_ac", _i), ("h_amp", _d * 3), ("h_freq", _d * 3), ("h_phase", _d * 3)]


class LaunchShape(ctypes.Structure):
    _fields_ = [("lpw_x", _i), ("tk_y", _i), ("tk_z", _i), ("rpw", _i), ("stage_threads", _i)]


# mm_field_fn_t: void (*)(double t, double *h3, void *user)
FIELD_FN = ctypes.CFUNCTYPE(None, _d, _dp, _vp)

# name -> (restype, argtypes); every symbol of include/mm_hotpath.h
SIGNATURES = {
    "mm_last_error": (ctypes.c_char_p, []),
    "mm_init": (_i, [_i]),
    "mm_device_alloc": (_i, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "mm_device_free": (_i, [_vp]),
    "mm_memcpy_htod": (_i, [_vp, _vp, ctypes.c_size_t, _vp]),
    "mm_memcpy_dtoh": (_i, [_vp, _vp, ctypes.c_size_t, _vp]),
    "mm_synchronize": (_i, [_vp]),
    "mm_device_copy": (_i, [_vp, _vp, ctypes.c_size_t, _vp]),
    "mm_round_tensor_dimension": (_i, [_i, _i, _i, _ip]),
    "mm_generate_demag_tensor": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _i, _i, _i, ctypes.c_char_p, _i, _dp, _ip]),
    "mm_generate_demag_tensor_gpu": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _i, _dp, _ip, _dp]),
    "mm_demag_tensor_entry": (_i, [_i, _i, _i, _d, _d, _d, _dp]),
    "mm_conv_create": (_i, [_dp, _i, _i, _i, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "mm_conv_execute": (_i, [_vp, _vp, _vp, _vp]),
    "mm_conv_execute_f32": (_i, [_vp, _vp, _vp, _vp]),
    "mm_conv_fft_size": (_i, [_vp, _ip]),
    "mm_conv_destroy": (_i, [_vp]),
    "mm_conv_simple_create": (_i, [_dp, _i, _i, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "mm_conv_simple_execute": (_i, [_vp, _vp, _vp, _vp]),
    "mm_conv_simple_destroy": (_i, [_vp]),
    "mm_exchange": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _i, _vp, _vp, _vp, _vp, _dp, _vp]),
    "mm_exchange_f32": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "mm_uniaxial_anisotropy": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _dp, _vp]),
    "mm_cubic_anisotropy": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _dp, _vp]),
    "mm_llge": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mm_fdm_zhangli": (_i, [_i, _i, _i, _d, _d, _d, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mm_scaled_abs_max": (_i, [_i, _vp, _vp, _i, _dp, _vp]),
    "mm_scaled_abs_max_noalloc": (_i, [_i, _vp, _vp, _i, _vp, _vp]),
    "mm_abs_max": (_i, [_i, _vp, _dp, _vp]),
    "mm_dot_sum": (_i, [_i, _vp, _vp, _dp, _vp]),
    "mm_average": (_i, [_i, _vp, _dp, _vp]),
    "mm_normalize": (_i, [_i, _vp, _vp, _i, _vp]),
    "mm_tableau": (_i, [_i, ctypes.POINTER(Tableau)]),
    "mm_rk_prepare_step": (_i, [_i, _d, ctypes.POINTER(Tableau), _i, ctypes.POINTER(_vp), _vp, _vp, _vp]),
    "mm_rk_combine_result": (_i, [_d, ctypes.POINTER(Tableau), _i, ctypes.POINTER(_vp), _vp, _vp, _vp]),
    "mm_stepper_create": (_i, [ctypes.POINTER(SystemDesc), ctypes.POINTER(_vp)]),
    "mm_stepper_set_m": (_i, [_vp, _dp]),
    "mm_stepper_get_m": (_i, [_vp, _dp]),
    "mm_stepper_dmdt": (_i, [_vp, _dp, _dp]),
    "mm_stepper_run": (_i, [_vp, _i]),
    "mm_stepper_run_until": (_i, [_vp, _d, ctypes.c_longlong]),
    "mm_stepper_run_until_stop": (_i, [_vp, _d, _d, ctypes.c_longlong]),
    "mm_stepper_sync": (_i, [_vp]),
    "mm_stepper_stats": (_i, [_vp, _dp]),
    "mm_stepper_set_time": (_i, [_vp, _d, _d]),
    "mm_stepper_step_count": (_i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "mm_stepper_average_m": (_i, [_vp, _dp]),
    "mm_stepper_stream": (_vp, [_vp]),
    "mm_stepper_kernel_profile": (_i, [_vp, _i, _dp]),
    "mm_stepper_energy": (_i, [_vp, _dp]),
    "mm_stepper_relax": (_i, [_vp, _i, _d, _i, ctypes.c_longlong, _dp]),
    "mm_stepper_set_field_fn": (_i, [_vp, FIELD_FN, _vp]),
    "mm_stepper_destroy": (_i, [_vp]),
    "mm_stepper_launch_shape": (_i, [_vp, ctypes.POINTER(LaunchShape)]),
    "mm_autotune": (_i, [ctypes.POINTER(SystemDesc), _i, ctypes.POINTER(LaunchShape), _dp]),
    "mm_wisdom_import": (_i, [ctypes.c_char_p, _ip]),
    "mm_wisdom_export": (_i, [ctypes.c_char_p]),
    "mm_wisdom_forget": (_i, []),
    "mm_batch_create": (_i, [ctypes.POINTER(SystemDesc), _i, _i, ctypes.POINTER(_vp)]),
    "mm_batch_set_m": (_i, [_vp, _i, _dp]),
    "mm_batch_get_m": (_i, [_vp, _i, _dp]),
    "mm_batch_run": (_i, [_vp, _i]),
    "mm_batch_stats": (_i, [_vp, _i, _dp]),
    "mm_batch_average_m": (_i, [_vp, _i, _dp]),
    "mm_batch_destroy": (_i, [_vp]),
}

_LIB = None
_INIT = False


def load():
    """Load libmmhip.so and declare every entry point (no GPU needed)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmmhip.so is not built (%s); run micromagnum_amd.build.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != MM_OK:
        msg = load().mm_last_error().decode("utf-8", "replace")
        raise RuntimeError("libmmhip error %d: %s" % (rc, msg))


def device_lib():
    """The library with its HIP device initialised (MM_DEVICE, default 0);
    raises when no device is visible.  torch is never touched here: the
    library brings its own HIP runtime."""
    global _INIT
    L = load()
    if not _INIT:
        dev = int(os.environ.get("MM_DEVICE", "0"))
        check(L.mm_init(dev))
        _INIT = True
    return L
