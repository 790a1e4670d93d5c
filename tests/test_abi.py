"""The C-ABI library loads and exports every symbol include/fcc.h declares
(no device calls), and the host mirror validates like the reference."""
import ctypes
import os
import re

import numpy as np
import pytest

import fcc_hip as F
from fcc_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = ""
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text += open(os.path.join(ROOT, "include", h)).read()
    return sorted(set(re.findall(r"\b(fcc_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_match_bindings():
    assert header_symbols() == _abi.HEADER_SYMBOLS


def test_header_constants_match_bindings():
    text = "".join(open(os.path.join(ROOT, "include", h)).read()
                   for h in sorted(os.listdir(os.path.join(ROOT, "include"))) if h.endswith(".h"))
    defs = dict(re.findall(r"#define (FCC_[A-Z0-9_]+) (\d+)\b", text))
    shared = [k for k in defs if hasattr(_abi, k)]
    assert "FCC_MAX_THETA" in shared and "FCC_MAX_MODES" in shared
    for k in shared:
        assert getattr(_abi, k) == int(defs[k]), k


def test_library_exports_header_symbols():
    path = _abi.library_path()
    assert os.path.exists(path), "libfcc_hip.so not built"
    lib = ctypes.CDLL(path)
    for sym in header_symbols():
        assert hasattr(lib, sym), sym
    lib.fcc_abi_version.restype = ctypes.c_int
    version = int(re.search(r"#define FCC_ABI_VERSION (\d+)",
                            open(os.path.join(ROOT, "include", "fcc.h")).read()).group(1))
    assert lib.fcc_abi_version() == version


def test_struct_layout():
    # sizes of the C structs as compiled by the C compiler (oracle build)
    assert ctypes.sizeof(_abi.FccBodyType) == 56
    assert ctypes.sizeof(_abi.FccTerm) == 8 + 16 + 16 + 2 * 8 * 16
    assert ctypes.sizeof(_abi.FccBodies) == 9 * 8
    assert ctypes.sizeof(_abi.FccLaunch) == 20


def test_host_validation_errors():
    with pytest.raises(F.ArgumentError):
        F.BodyType(drag="cytoplasm")
    with pytest.raises(F.ArgumentError):
        F.hexagonal_monolayer(4, 3)
    with pytest.raises(F.DimensionMismatch):
        F.Bodies([0.0, 1.0], [0.0], [0.0, 0.0], np.zeros((11, 2)), [1.0, 1.0])
    with pytest.raises(F.DimensionMismatch):
        F.Overlap([[1.0, 2.0]]).tables()


def test_library_validates_before_touching_the_device():
    # argument checks of fcc_create / fcc_ensemble_create return before any
    # HIP call (FCC_ERR_ARGUMENT / FCC_ERR_UNSUPPORTED with a message)
    path = _abi.library_path()
    assert os.path.exists(path), "libfcc_hip.so not built"
    eng = _abi.Engine(ctypes.CDLL(path), "fcc_", "hip")
    h = ctypes.c_void_p()
    assert eng.create(None, None, ctypes.byref(h)) == _abi.FCC_ERR_ARGUMENT
    assert eng.ensemble_create(0, None, None, ctypes.byref(h)) == _abi.FCC_ERR_ARGUMENT
    assert "replica" in eng.last_error()
    b, box = F.hexagonal_monolayer(4, 4)
    bad = F.api._config(b, box, (F.BodyType(),), (F.Tension(0.01),), 0.005, _abi.FCC_MAX_THETA + 1, "fp32", 1,
                        0.3, 0)
    st = b.struct()
    assert eng.create(ctypes.byref(bad), ctypes.byref(st), ctypes.byref(h)) == _abi.FCC_ERR_UNSUPPORTED
    assert str(_abi.FCC_MAX_THETA) in eng.last_error()
    bad.n_theta = 10  # must exceed 2·maxmode (src/basis.jl:38)
    assert eng.create(ctypes.byref(bad), ctypes.byref(st), ctypes.byref(h)) == _abi.FCC_ERR_ARGUMENT
    cfgs = (_abi.FccConfig * 2)()
    structs = (_abi.FccBodies * 2)()
    b8, box8 = F.hexagonal_monolayer(4, 4, maxmode=8)
    cfgs[0] = F.api._config(b, box, (F.BodyType(),), (F.Tension(0.01),), 0.005, 40, "fp32", 1, 0.3, 0)
    cfgs[1] = F.api._config(b8, box8, (F.BodyType(),), (F.Tension(0.01),), 0.005, 40, "fp32", 2, 0.3, 0)
    structs[0], structs[1] = b.struct(), b8.struct()
    assert eng.ensemble_create(2, cfgs, structs, ctypes.byref(h)) == _abi.FCC_ERR_UNSUPPORTED
    assert "replica 1" in eng.last_error()
    assert eng.ensemble_destroy(None) == _abi.FCC_OK
    assert eng.destroy(None) == _abi.FCC_OK
    # every entry point taking a handle rejects a null one before any HIP call
    obs = _abi.FccObserverConfig(16, 4, 3.5, 4.0)
    n = ctypes.c_int64()
    for rc in (eng.set_bonds(None, 0, None), eng.observer_start(None, ctypes.byref(obs)), eng.observe(None),
               eng.run_observed(None, 10, 5), eng.observer_read(None, None, None, None, None, None, None),
               eng.contours(None, None, None), eng.frame_request(None), eng.frame_read(None, None, None, None),
               eng.checkpoint_save(None, None, 0, ctypes.byref(n)), eng.checkpoint_load(None, None, 0),
               eng.time_rebuild(None, 1, None), eng.ensemble_observer_start(None, ctypes.byref(obs)),
               eng.ensemble_observe(None), eng.ensemble_run_observed(None, 10, 5)):
        assert rc == _abi.FCC_ERR_ARGUMENT


def test_hexagonal_monolayer_geometry():
    # src/initial.jl:16-24: box 2r nx × √3 r ny with N π R0² = packing Lx Ly
    b, box = F.hexagonal_monolayer(32, 32, packing=0.95)
    assert len(b) == 1024
    assert abs(1024 * np.pi / (box.Lx * box.Ly) - 0.95) < 1e-12
    assert abs(box.Lx / box.Ly - 2 / np.sqrt(3)) < 1e-12
    assert np.all((b.x >= 0) & (b.x < box.Lx) & (b.y >= 0) & (b.y < box.Ly))
    assert np.all((b.theta > 0) & (b.theta < 2 * np.pi))


def test_with_nuclei_and_vcat():
    # test/test_parents.jl:160-178
    coef = np.zeros((11, 2))
    coef[0] = 1
    a = F.Bodies([1.0, 2.0], [1.0, 2.0], np.zeros(2), coef, np.ones(2), parent=[1, 1])
    b = F.Bodies([3.0, 4.0], [3.0, 4.0], np.zeros(2), coef, np.ones(2), type=2, parent=[2, 2])
    c = F.vcat(a, b)
    assert list(c.parent) == [1, 1, 4, 4] and list(c.type) == [1, 1, 2, 2]
    cells, box = F.hexagonal_monolayer(4, 4)
    nb = F.with_nuclei(cells, ratio=0.4)
    assert len(nb) == 32
    assert list(nb.parent) == list(range(1, 17)) * 2
    np.testing.assert_allclose(nb.R0[16:], 0.4 * cells.R0)
