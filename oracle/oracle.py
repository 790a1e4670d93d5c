"""ctypes loader of the FP64 CPU oracle (liboracle_fcc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check of
__graft_entry__.py and by the cpu_baseline leg of bench.py, as the checker.
The oracle implements the same C ABI shape as libfcc_hip.so (prefix fcco_),
so the host mirror's Simulation can drive either.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_fcc.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "fouriercontourcells.jl_amd"))

from fcc_hip import _abi  # noqa: E402

_ENGINE = None


def build(force=False):
    """Compile liboracle_fcc.so with the Makefile next to this file."""
    args = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def engine():
    global _ENGINE
    if _ENGINE is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        e = _abi.Engine(lib, "fcco_", "oracle")
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        lib.fcco_set_center_gradients.argtypes = [ctypes.c_void_p, dp, dp]
        lib.fcco_integrate_noise.argtypes = [ctypes.c_void_p, dp]
        lib.fcco_build_lists.argtypes = [ctypes.c_int64, dp, dp, dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int32, ip, ip, ip, ip, ip, ip, ip, ip, ip]
        lib.fcco_philox4x32.argtypes = [u32p, u32p, u32p]
        lib.fcco_philox4x32.restype = None
        lib.fcco_uniform_open64.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        lib.fcco_uniform_open64.restype = ctypes.c_double
        lib.fcco_uniform_open32.argtypes = [ctypes.c_uint32]
        lib.fcco_uniform_open32.restype = ctypes.c_float
        lib.fcco_minimum_image.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.fcco_minimum_image.restype = ctypes.c_double
        lib.fcco_wrap.argtypes = [ctypes.c_double, ctypes.c_double, ip]
        lib.fcco_wrap.restype = ctypes.c_double
        lib.fcco_set_threads.argtypes = [ctypes.c_int]
        lib.fcco_set_threads.restype = None
        _ENGINE = e
    return _ENGINE


def set_threads(n):
    engine().lib.fcco_set_threads(int(n))


def philox4x32(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    engine().lib.fcco_philox4x32(c, k, o)
    return tuple(o)


def uniform_open64(a, b):
    return engine().lib.fcco_uniform_open64(a, b)


def uniform_open32(a):
    return engine().lib.fcco_uniform_open32(a)


def minimum_image(d, L):
    return engine().lib.fcco_minimum_image(d, L)


def wrap(x, L):
    s = ctypes.c_int32()
    v = engine().lib.fcco_wrap(x, L, ctypes.byref(s))
    return v, s.value


def set_center_gradients(sim, gx, gy):
    gx = np.ascontiguousarray(gx, dtype=np.float64)
    gy = np.ascontiguousarray(gy, dtype=np.float64)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    engine().lib.fcco_set_center_gradients(sim._h, dp(gx), dp(gy))


def integrate_noise(sim, xi):
    xi = np.asfortranarray(xi, dtype=np.float64)
    engine().lib.fcco_integrate_noise(sim._h, xi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def build_lists(x, y, rmax, Lx, Ly, cutoff, rng, capacity=32):
    """Cell list + neighbour list from given positions (1-based outputs)."""
    n = len(x)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    rmax = np.ascontiguousarray(rmax, dtype=np.float64)
    ncx_f = int(np.floor(Lx / cutoff))
    ncy_f = int(np.floor(Ly / cutoff))
    nce = (ncx_f if ncx_f >= 3 else 1) * (ncy_f if ncy_f >= 3 else 1)
    cellid = np.zeros(n, np.int32)
    perm = np.zeros(n, np.int32)
    start = np.zeros(nce, np.int32)
    stop = np.zeros(nce, np.int32)
    count = np.zeros(n, np.int32)
    idx = np.zeros((capacity, n), np.int32, order="F")
    ncx, ncy, mx = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rc = engine().lib.fcco_build_lists(n, dp(x), dp(y), dp(rmax), Lx, Ly, cutoff, rng, capacity,
                                       ctypes.byref(ncx), ctypes.byref(ncy), ip(cellid), ip(perm),
                                       ip(start), ip(stop), ip(count), ip(idx), ctypes.byref(mx))
    assert rc == 0
    return dict(ncx=ncx.value, ncy=ncy.value, cellid=cellid, perm=perm, start=start, stop=stop,
                count=count, idx=idx, max_count=mx.value)
