"""The Julia side of the boundary, checked without Julia (not installed
here): every `ccall` of the package extension and of FCCHIP names a function
include/*.h declares and libfcc_hip.so exports, with the declared number of
arguments; the Julia mirrors of the C structs have the fields of the ctypes
mirror (whose layout test_abi.py pins); FCCHIP's package files and the
registration lines for the reference's Project.toml name the same package."""
import ctypes
import os
import re

import pytest

from fcc_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "fouriercontourcells.jl_amd", "julia")
EXT = os.path.join(JL, "ext", "FourierContourCellsHIPExt.jl")
PKG = os.path.join(JL, "FCCHIP")
# FourierContourCells.jl's package uuid (its Project.toml)
FCC_UUID = "199c25b0-0c5e-4658-8424-08d56a2832c7"


def split_top(s):
    """Split at commas outside braces and brackets."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "{[(":
            depth += 1
        elif ch in "}])":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def header_arities():
    text = ""
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text += open(os.path.join(ROOT, "include", h)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(fcc_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else len(split_top(args))
    return out


def julia_ccalls(path):
    text = open(path).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(fcc_[a-z0-9_]+),\s*libfcc_hip\),\s*([A-Za-z{}]+),\s*\(([^()]*)\)", text):
        calls.append((m.group(1), len(split_top(m.group(3)))))
    return calls


def test_every_ccall_is_a_declared_export_with_its_arity():
    decl = header_arities()
    lib = ctypes.CDLL(_abi.library_path())
    calls = julia_ccalls(EXT) + julia_ccalls(os.path.join(PKG, "src", "FCCHIP.jl"))
    assert len(calls) > 30
    for name, n in calls:
        assert name in decl, name
        assert decl[name] == n, f"{name}: {n} arguments in Julia, {decl[name]} in the header"
        assert hasattr(lib, name), name
    # the time-step entry points are all bound
    names = {c[0] for c in calls}
    for must in ("fcc_create", "fcc_destroy", "fcc_step", "fcc_run", "fcc_compute_gradients", "fcc_get_gradients",
                 "fcc_update_neighbors", "fcc_shape_index", "fcc_get_bodies", "fcc_get_counters", "fcc_set_dt"):
        assert must in names, must


@pytest.mark.parametrize("jl,cstruct", [("FccBodyType", _abi.FccBodyType), ("FccTerm", _abi.FccTerm),
                                        ("FccConfig", _abi.FccConfig), ("FccBodies", _abi.FccBodies),
                                        ("FccObserverConfig", _abi.FccObserverConfig),
                                        ("FccLaunch", _abi.FccLaunch)])
def test_julia_structs_mirror_the_c_structs(jl, cstruct):
    text = open(EXT).read()
    m = re.search(r"struct " + jl + r"\n(.*?)\nend", text, re.S)
    assert m, jl
    fields = [f.split("::")[0].strip() for line in m.group(1).split("\n") for f in line.split(";") if "::" in f]
    assert fields == [f[0] for f in cstruct._fields_]


def test_constants_match_the_header():
    h = open(os.path.join(ROOT, "include", "fcc.h")).read()
    ext = open(EXT).read()
    pkg = open(os.path.join(PKG, "src", "FCCHIP.jl")).read()
    abi = int(re.search(r"#define FCC_ABI_VERSION (\d+)", h).group(1))
    assert int(re.search(r"const ABI_VERSION = (\d+)", pkg).group(1)) == abi
    for name in ("FCC_ERR_ARGUMENT", "FCC_ERR_DIMENSION", "FCC_ERR_UNSUPPORTED"):
        assert int(re.search(name + r" = Cint\((-?\d+)\)", ext).group(1)) == getattr(_abi, name)
    assert int(re.search(r"FCC_FLAG_DEBUG_BUFFERS = Int32\((\d+)\)", ext).group(1)) == _abi.FCC_FLAG_DEBUG_BUFFERS
    prec = dict(re.findall(r"(Float64|Float32|:mixed_bf16|:mixed_fp16|:mixed) => (\d)", ext))
    assert prec == {"Float64": "0", "Float32": "1", ":mixed": "2", ":mixed_fp16": "2", ":mixed_bf16": "3"}
    kinds = [int(k) for k in re.findall(r"return FccTerm\((\d+), length\(on\)", ext)]
    assert sorted(kinds) == sorted([_abi.FCC_TERM_TENSION, _abi.FCC_TERM_PERIMETER_ELASTICITY,
                                                     _abi.FCC_TERM_BENDING, _abi.FCC_TERM_OVERLAP,
                                                     _abi.FCC_TERM_LINEAR_REPULSION, _abi.FCC_TERM_HARMONIC_BOND])


def test_package_and_registration_files():
    import tomli
    proj = tomli.load(open(os.path.join(PKG, "Project.toml"), "rb"))
    assert proj["name"] == "FCCHIP"
    assert proj["deps"]["FourierContourCells"] == FCC_UUID
    reg = tomli.load(open(os.path.join(JL, "Project-additions.toml"), "rb"))
    assert reg["weakdeps"]["FCCHIP"] == proj["uuid"]
    assert reg["extensions"]["FourierContourCellsHIPExt"] == "FCCHIP"
    assert os.path.basename(EXT) == "FourierContourCellsHIPExt.jl"
    ext = open(EXT).read()
    assert re.search(r"^module FourierContourCellsHIPExt$", ext, re.M)
    pkg = open(os.path.join(PKG, "src", "FCCHIP.jl")).read()
    # what the extension takes from FCCHIP is defined there
    used = re.findall(r"(?:using|import) FCCHIP: ([^\n]*(?:\n    [^\n]*)*)", ext)
    names = [n.strip() for u in used for n in u.replace("\n", " ").split(",") if n.strip()]
    assert len(names) >= 10
    for n in names:
        assert re.search(r"(struct|function|const) " + re.escape(n) + r"(?=[\s({])", pkg), n


def test_fallback_rules_are_in_the_constructor():
    # hook!, an explicit layout, non-native terms, unknown keywords and
    # FCC_ERR_UNSUPPORTED go to the reference path with every argument
    ext = open(EXT).read()
    ctor = ext[ext.index("function FourierContourCells.Simulation"):ext.index("handle(sim::HIPSimulation) =")]
    for rule in ("hook! === nothing ||", "layout === nothing ||", "isempty(kwargs) ||",
                 "all(!isnothing, terms) ||", "rc == FCC_ERR_UNSUPPORTED && return ref("):
        assert rule in ctor, rule
    ref = ext[ext.index("    ref(why) = reference_simulation("):]
    ref = ref[:ref.index("\n    isempty(kwargs)")]
    for kw in ("T", "types", "energies", "bonds", "dt", "nθ", "seed", "skin", "hook!", "layout", "kwargs..."):
        assert re.search(r"[ (;,]" + re.escape(kw) + r"[,)\s]", ref), kw
