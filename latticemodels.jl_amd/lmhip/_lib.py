"""ctypes binding of liblmhip.so (declarations in include/lmhip.h).

The library is required: importing the package without a built library or on
a machine without a HIP device raises instead of falling back to any CPU path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liblmhip.so")

LMH_OK, LMH_EINVAL, LMH_ENOMEM, LMH_EDEVICE, LMH_ESTATE, LMH_ERANGE = range(6)
(KERNEL_HORNER_LEFT, KERNEL_HORNER_RIGHT, KERNEL_BUILD_H, KERNEL_TAYLOR_ORDER, KERNEL_OBSERVABLES,
 KERNEL_PROJECTOR, KERNEL_FUSED, KERNEL_TRANSPOSE, KERNEL_ADJOINT, KERNEL_ADJOINT_REDUCE, KERNEL_KRYLOV) = range(11)
STEP_AUTO, STEP_PER_PASS, STEP_FUSED_T, STEP_FUSED_N = range(4)
SERIES_DENSITY, SERIES_CURRENTS, SERIES_REGION = 1, 2, 3
FIELD_LANDAU, FIELD_SYMMETRIC, FIELD_POINT_AXIAL, FIELD_POINT_SINGULAR = 1, 2, 3, 4

# lmh_progress_fn: (user, done, total)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64)

_lib = None


class LMHError(RuntimeError):
    """A device, allocation or ordering failure reported by liblmhip."""

    def __init__(self, code, msg):
        super().__init__("lmhip error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc == LMH_OK:
        return
    msg = lib().lmh_last_error().decode(errors="replace")
    if rc == LMH_EINVAL:
        # the reference raises ArgumentError for invalid input
        raise ValueError(msg)
    raise LMHError(rc, msg)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError("liblmhip.so is not built; run `python latticemodels.jl_amd/build.py`")
    L = ctypes.CDLL(LIBPATH)
    vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    pvp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "lmh_version": ([], ctypes.c_int),
        "lmh_last_error": ([], ctypes.c_char_p),
        "lmh_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "lmh_create": ([pvp, ctypes.c_int], ctypes.c_int),
        "lmh_destroy": ([vp], ctypes.c_int),
        "lmh_set_geometry": ([vp, i64, i64, vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lmh_set_onsite": ([vp, vp], ctypes.c_int),
        "lmh_set_site_coords": ([vp, vp], ctypes.c_int),
        "lmh_set_option": ([vp, ctypes.c_int, i64], ctypes.c_int),
        "lmh_get_option": ([vp, ctypes.c_int, ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_autotune": ([vp, vp, ctypes.c_int32, ctypes.POINTER(dbl)], ctypes.c_int),
        "lmh_set_field": ([vp, i64, vp, vp], ctypes.c_int),
        "lmh_update_solver": ([vp, dbl, ctypes.c_int, dbl, dbl, ctypes.POINTER(i32), ctypes.POINTER(dbl)], ctypes.c_int),
        "lmh_solver_order": ([vp, ctypes.POINTER(i32), ctypes.POINTER(dbl)], ctypes.c_int),
        "lmh_state_create": ([vp, pvp], ctypes.c_int),
        "lmh_state_destroy": ([vp], ctypes.c_int),
        "lmh_state_upload": ([vp, vp, i64], ctypes.c_int),
        "lmh_state_download": ([vp, vp, i64], ctypes.c_int),
        "lmh_state_copy": ([vp, vp], ctypes.c_int),
        "lmh_state_device_ptr": ([vp, pvp, ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_state_set_entries": ([vp, i64, vp, vp, vp], ctypes.c_int),
        "lmh_state_get_entries": ([vp, i64, vp, vp, vp], ctypes.c_int),
        "lmh_state_idempotency": ([vp, ctypes.POINTER(dbl)], ctypes.c_int),
        "lmh_step": ([vp, vp], ctypes.c_int),
        "lmh_set_csc_pattern": ([vp, i64, vp, vp, ctypes.c_int32], ctypes.c_int),
        "lmh_batch_member_kernel": ([vp, i64, ctypes.POINTER(i32)], ctypes.c_int),
        "lmh_set_csc_values": ([vp, i64, vp], ctypes.c_int),
        "lmh_currents_region_noalloc": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lmh_observables": ([vp, vp, vp, vp], ctypes.c_int),
        "lmh_get_hamiltonian": ([vp, vp, vp], ctypes.c_int),
        "lmh_sync": ([vp], ctypes.c_int),
        "lmh_densitymatrix": ([vp, vp, dbl, dbl, ctypes.POINTER(i64), vp], ctypes.c_int),
        "lmh_diagonalize": ([vp, vp], ctypes.c_int),
        "lmh_projector": ([vp, vp, vp, ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_eigensystem_free": ([vp], ctypes.c_int),
        "lmh_currents_region": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lmh_series_create": ([vp, ctypes.c_int, i64, vp, vp, pvp], ctypes.c_int),
        "lmh_series_destroy": ([vp], ctypes.c_int),
        "lmh_series_record": ([vp, vp, dbl], ctypes.c_int),
        "lmh_series_length": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_series_differentiate": ([vp], ctypes.c_int),
        "lmh_series_integrate": ([vp], ctypes.c_int),
        "lmh_series_download": ([vp, vp, vp], ctypes.c_int),
        "lmh_set_step_mode": ([vp, ctypes.c_int], ctypes.c_int),
        "lmh_set_profiling": ([vp, ctypes.c_int], ctypes.c_int),
        "lmh_get_profile": ([vp, ctypes.c_int, ctypes.POINTER(i64), ctypes.POINTER(dbl)], ctypes.c_int),
        "lmh_get_stream": ([vp, pvp], ctypes.c_int),
        "lmh_kets_create": ([vp, i64, pvp], ctypes.c_int),
        "lmh_kets_destroy": ([vp], ctypes.c_int),
        "lmh_kets_upload": ([vp, vp, i64], ctypes.c_int),
        "lmh_kets_download": ([vp, vp, i64], ctypes.c_int),
        "lmh_kets_step": ([vp, vp], ctypes.c_int),
        "lmh_krylov_step": ([vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                             ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_kets_observables": ([vp, vp, i64, vp, vp], ctypes.c_int),
        "lmh_tape_begin": ([vp, vp, i64], ctypes.c_int),
        "lmh_tape_end": ([vp], ctypes.c_int),
        "lmh_tape_length": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_tape_backward": ([vp, vp, vp, vp], ctypes.c_int),
        "lmh_tape_checkpoints": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_value_classes": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64)], ctypes.c_int),
        "lmh_set_progress": ([vp, PROGRESS_FN, vp], ctypes.c_int),
        "lmh_hamiltonian_derivative": ([vp, i64, vp], ctypes.c_int),
        "lmh_batch_create": ([pvp, i64, vp, vp], ctypes.c_int),
        "lmh_batch_destroy": ([vp], ctypes.c_int),
        "lmh_batch_step": ([vp, i64, ctypes.c_int], ctypes.c_int),
        "lmh_batch_mean_density": ([vp, vp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names of the C-ABI entry points bound here."""
    lib()
    return [n for n in dir(_lib) if n.startswith("lmh_")]


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)

(OPT_STEP_MODE, OPT_RIGHT_PASS, OPT_LEFT_COLS, OPT_RIGHT_COLS, OPT_FUSED_MAX_N, OPT_EIGENSOLVER, OPT_FUSED_VARIANT,
 OPT_ADJOINT_MODE, OPT_LEFT_PASS, OPT_CHECKPOINTS, OPT_KRYLOV_GRAPHS, OPT_VALUE_CLASSES,
 OPT_FUSED_LAUNCHED, OPT_EIGEN_USED, OPT_BATCH_STRIP, OPT_DEFER, OPT_ASYNC_ORDER, OPT_STEP_COUNT,
 OPT_STEP_TERMS, OPT_ASYNC_STEPS, OPT_ASYNC_REDONE) = range(1, 22)
FUSED_STRIP = 32      # fused variant of the strip kernel (grid-stencil H)
