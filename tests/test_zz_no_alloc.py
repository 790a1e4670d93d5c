"""fcc_philox_noise_noalloc (the entry point that allocates nothing) gives the
words of fcc_philox_noise bit for bit, on two calls in a row (the second with
a smaller input, so stale words of the first would show), in FP64 and FP32,
and rejects sizes beyond its fixed buffer."""
import ctypes
import sys

import numpy as np
import pytest

from fcc_hip import _abi

pytestmark = pytest.mark.gpu


def noise(fn, seed, draw, n, dtype, prec):
    out = np.full(3 * n, np.nan, dtype=dtype)
    rc = fn(seed, draw, n, prec, out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_noise_equals_allocating_entry_point(hip, prec):
    print("noalloc noise", prec, flush=True)
    sys.stdout.flush()
    code = _abi.PRECISIONS[prec]
    dtype = np.float64 if prec == "fp64" else np.float32
    for seed, draw, n in ((1234, 7, _abi.FCC_NOALLOC_MAX_NOISE), (99, (1 << 32) + 5, 1000)):
        rc_a, a = noise(hip.philox_noise, seed, draw, n, dtype, code)
        rc_b, b = noise(hip.philox_noise_noalloc, seed, draw, n, dtype, code)
        assert rc_a == 0 and rc_b == 0, hip.last_error()
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert np.all(np.isfinite(b))


def test_noalloc_noise_rejects_sizes_beyond_its_buffer(hip):
    out = np.zeros(3, dtype=np.float64)
    rc = hip.philox_noise_noalloc(1, 1, _abi.FCC_NOALLOC_MAX_NOISE + 1, _abi.FCC_FP64,
                                  out.ctypes.data_as(ctypes.c_void_p))
    assert rc == _abi.FCC_ERR_ARGUMENT
    assert "exceeds" in hip.last_error()
    assert hip.philox_noise_noalloc(1, 1, 0, _abi.FCC_FP64, out.ctypes.data_as(ctypes.c_void_p)) == 0
