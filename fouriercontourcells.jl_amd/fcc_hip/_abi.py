"""ctypes mirror of include/fcc.h and the loader of libfcc_hip.so.

The structures match the C declarations field for field; both the device
library (libfcc_hip.so) and any other implementation of the same C ABI are
driven through them.
"""
import ctypes
import os

FCC_MAX_TYPES = 4
FCC_MAX_TERMS = 8
FCC_MAX_MODES = 31
FCC_MAX_THETA = 8000

FCC_OK = 0
FCC_ERR_ARGUMENT = -1
FCC_ERR_DIMENSION = -2
FCC_ERR_UNSUPPORTED = -3
FCC_ERR_DEVICE = -4
FCC_ERR_STATE = -5

FCC_FP64 = 0
FCC_FP32 = 1
FCC_MIXED_FP16 = 2
FCC_MIXED_BF16 = 3

FCC_TERM_TENSION = 1
FCC_TERM_PERIMETER_ELASTICITY = 2
FCC_TERM_BENDING = 3
FCC_TERM_OVERLAP = 16
FCC_TERM_LINEAR_REPULSION = 17
FCC_TERM_HARMONIC_BOND = 32

FCC_FLAG_DEBUG_BUFFERS = 1
FCC_FLAG_TWO_LANES = 2
FCC_FLAG_TWO_KERNELS = 4

FCC_NOALLOC_MAX_NOISE = 262144

PRECISIONS = {
    "fp64": FCC_FP64,
    "fp32": FCC_FP32,
    "mixed": FCC_MIXED_FP16,
    "mixed_fp16": FCC_MIXED_FP16,
    "mixed_bf16": FCC_MIXED_BF16,
}


class FccBodyType(ctypes.Structure):
    _fields_ = [
        ("v0", ctypes.c_double),
        ("Dr", ctypes.c_double),
        ("Dt", ctypes.c_double),
        ("mur", ctypes.c_double),
        ("mutheta", ctypes.c_double),
        ("muab", ctypes.c_double),
        ("parent_drag", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FccTerm(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n_on", ctypes.c_int32),
        ("on_types", ctypes.c_int32 * FCC_MAX_TYPES),
        ("p", ctypes.c_double * 2),
        ("J", ctypes.c_double * (FCC_MAX_TYPES * FCC_MAX_TYPES)),
        ("Jrel", ctypes.c_double * (FCC_MAX_TYPES * FCC_MAX_TYPES)),
    ]


class FccConfig(ctypes.Structure):
    _fields_ = [
        ("n_bodies", ctypes.c_int64),
        ("n_modes", ctypes.c_int32),
        ("n_theta", ctypes.c_int32),
        ("n_types", ctypes.c_int32),
        ("n_terms", ctypes.c_int32),
        ("types", FccBodyType * FCC_MAX_TYPES),
        ("terms", FccTerm * FCC_MAX_TERMS),
        ("dt", ctypes.c_double),
        ("Lx", ctypes.c_double),
        ("Ly", ctypes.c_double),
        ("skin", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("precision", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


FCC_OBS_MAX_BINS = 2046


class FccObserverConfig(ctypes.Structure):
    _fields_ = [
        ("nbins", ctypes.c_int32),
        ("max_samples", ctypes.c_int32),
        ("pmin", ctypes.c_double),
        ("pmax", ctypes.c_double),
    ]


class FccLaunch(ctypes.Structure):
    _fields_ = [
        ("update_block", ctypes.c_int32),
        ("xcd_order", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("lanes", ctypes.c_int32),
        ("kernels", ctypes.c_int32),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class FccBodies(ctypes.Structure):
    _fields_ = [
        ("x", _dp),
        ("y", _dp),
        ("theta", _dp),
        ("imgx", _ip),
        ("imgy", _ip),
        ("coef", _dp),
        ("R0", _dp),
        ("type", _ip),
        ("parent", _ip),
    ]


# Entry points of the C ABI with their argument types; `prefix` selects the
# implementation ("fcc_" for libfcc_hip.so).
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_SIGNATURES = {
    "create": [ctypes.POINTER(FccConfig), ctypes.POINTER(FccBodies), ctypes.POINTER(_vp)],
    "destroy": [_vp],
    "set_bodies": [_vp, ctypes.POINTER(FccBodies)],
    "get_bodies": [_vp, ctypes.POINTER(FccBodies)],
    "set_dt": [_vp, ctypes.c_double],
    "set_bonds": [_vp, _i64, _ip],
    "step": [_vp, _i64],
    "update_neighbors": [_vp, _i32],
    "compute_gradients": [_vp, _i32],
    "get_gradients": [_vp, _dp, _dp, _dp, _dp, _dp, _dp],
    "get_contour_gradients": [_vp, _dp, _dp, _dp, _dp],
    "get_geometry": [_vp, _dp, _dp, _dp, _dp, _dp],
    "shape_index": [_vp, _dp, _dp, _dp],
    "contours": [_vp, _dp, _dp],
    "get_celllist": [_vp, _ip, _ip, _ip, _ip, _ip, _ip],
    "get_neighbors": [_vp, _ip, _ip, _ip],
    "set_neighbor_capacity": [_vp, _i32],
    "get_counters": [_vp, ctypes.POINTER(_i64), _dp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "philox_noise": [ctypes.c_uint64, ctypes.c_uint64, _i64, _i32, _vp],
}

# Entry points only the device library has.
_DEVICE_ONLY = {
    "run": [_vp, _i64],
    "philox4x32": [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), _i64,
                   ctypes.POINTER(ctypes.c_uint32)],
    "philox_noise_noalloc": [ctypes.c_uint64, ctypes.c_uint64, _i64, _i32, _vp],
    "abi_version": [],
    "time_step_kernels": [_vp, _i64, _dp, _dp],
    "run_timed": [_vp, _i64, _dp, _dp],
    "time_kernels": [_vp, _i64, _dp, _dp],
    "time_rebuild": [_vp, _i64, _dp],
    "ensemble_create": [_i32, ctypes.POINTER(FccConfig), ctypes.POINTER(FccBodies), ctypes.POINTER(_vp)],
    "ensemble_destroy": [_vp],
    "ensemble_member": [_vp, _i32, ctypes.POINTER(_vp)],
    "ensemble_run": [_vp, _i64],
    "ensemble_run_timed": [_vp, _i64, _dp, _dp],
    "ensemble_get_launch": [_vp, ctypes.POINTER(FccLaunch)],
    "ensemble_set_launch": [_vp, ctypes.POINTER(FccLaunch)],
    "ensemble_autotune": [_vp, _i64, ctypes.POINTER(FccLaunch), _dp],
    "frame_request": [_vp],
    "frame_read": [_vp, _dp, _dp, ctypes.POINTER(_i64)],
    "checkpoint_save": [_vp, _vp, _i64, ctypes.POINTER(_i64)],
    "checkpoint_load": [_vp, _vp, _i64],
    "observer_start": [_vp, ctypes.POINTER(FccObserverConfig)],
    "observe": [_vp],
    "run_observed": [_vp, _i64, _i64],
    "observer_read": [_vp, _ip, ctypes.POINTER(_i64), _dp, _dp, _dp, ctypes.POINTER(ctypes.c_uint64)],
    "ensemble_observer_start": [_vp, ctypes.POINTER(FccObserverConfig)],
    "ensemble_observe": [_vp],
    "ensemble_run_observed": [_vp, _i64, _i64],
}

# Every symbol include/fcc.h declares.
HEADER_SYMBOLS = sorted(
    ["fcc_" + n for n in _SIGNATURES] + ["fcc_" + n for n in _DEVICE_ONLY] + ["fcc_last_error"]
)


class Engine:
    """A loaded implementation of the C ABI (function table + error text)."""

    def __init__(self, lib, prefix, name):
        self.lib = lib
        self.prefix = prefix
        self.name = name
        for fname, args in list(_SIGNATURES.items()) + list(_DEVICE_ONLY.items()):
            sym = prefix + fname
            if not hasattr(lib, sym):
                continue
            f = getattr(lib, sym)
            f.argtypes = args
            f.restype = ctypes.c_int
            setattr(self, fname, f)
        err = getattr(lib, prefix + "last_error")
        err.argtypes = []
        err.restype = ctypes.c_char_p
        self._last_error = err

    def last_error(self):
        return self._last_error().decode()


def library_path():
    """The in-tree build (csrc/libfcc_hip.so), or the file named by
    FCC_HIP_LIBRARY (another build of the same C ABI)."""
    other = os.environ.get("FCC_HIP_LIBRARY")
    if other:
        return other
    here = os.path.dirname(os.path.abspath(__file__))
    return os.path.join(os.path.dirname(here), "csrc", "libfcc_hip.so")


_ENGINE = None


def hip_engine():
    """The device library. Raises if it has not been built: there is no
    host fallback for the time step."""
    global _ENGINE
    if _ENGINE is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"libfcc_hip.so not found at {path}: build it first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7
        # (same soname as /opt/rocm's), and its stack only initialises on that
        # copy. Loading torch first makes libfcc_hip.so bind to it as well, so
        # torch and this library can share the process in any import order.
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        _ENGINE = Engine(ctypes.CDLL(path), "fcc_", "hip")
    return _ENGINE
