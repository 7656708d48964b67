"""The kernels' deterministic powers, and the C ABI surface (CPU only)."""
import ctypes
import os
import re
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle.pyoracle import lib as olib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _exact_pow(x, e):
    getcontext().prec = 60
    return (Decimal(e) * Decimal(x).ln()).exp()


def test_pow_routines_are_correctly_rounded_on_random_inputs():
    # hf_pow.h claims ~correct rounding for pow(x, 4.0/3.0) and pow(x, 4.0/3.0 - 1.0);
    # check against 60-digit decimal arithmetic on log-uniform inputs.
    rng = np.random.default_rng(11)
    xs = 10.0 ** rng.uniform(-12, 6, 400)
    L = olib()
    for e, fn in ((4.0 / 3.0, L.hfo_pow_4_3), (4.0 / 3.0 - 1.0, L.hfo_pow_1_3m)):
        for x in xs:
            got = fn(float(x))
            ex = _exact_pow(float(x), e)
            ulp = np.spacing(got)
            assert abs(Decimal(got) - ex) <= Decimal(ulp) * Decimal("0.5000001"), (x, e)


def test_pow_routines_agree_with_glibc_within_one_ulp():
    rng = np.random.default_rng(5)
    xs = 10.0 ** rng.uniform(-9, 4, 20000)
    L = olib()
    for e, fn in ((4.0 / 3.0, L.hfo_pow_4_3), (4.0 / 3.0 - 1.0, L.hfo_pow_1_3m)):
        a = np.array([fn(float(x)) for x in xs])
        b = np.array([L.hfo_libm_pow(float(x), e) for x in xs])
        d = np.abs(a - b) / np.spacing(b)
        assert d.max() <= 1.0
        assert np.mean(a != b) < 5e-3  # glibc 2.35 pow misses correct rounding rarely


def test_pow_special_values():
    L = olib()
    assert L.hfo_pow_4_3(0.0) == 0.0 and L.hfo_pow_1_3m(0.0) == 0.0
    assert L.hfo_pow_4_3(8.0) == pow(8.0, 4.0 / 3.0)
    assert L.hfo_pow_4_3(float("inf")) == float("inf")
    assert np.isnan(L.hfo_pow_4_3(-1.0))
    assert L.hfo_pow_1_3m(5e-324) == pow(5e-324, 4.0 / 3.0 - 1.0)


def _pow35(L):
    L.hfo_pow_3_5.restype = ctypes.c_double
    L.hfo_pow_3_5.argtypes = [ctypes.c_double]
    L.hfo_pow_m2_5.restype = ctypes.c_double
    L.hfo_pow_m2_5.argtypes = [ctypes.c_double]
    return ((3.0 / 5.0, L.hfo_pow_3_5), (3.0 / 5.0 - 1.0, L.hfo_pow_m2_5))


def test_pow_3_5_routines_are_correctly_rounded():
    # the Discharge-onto-dry ghost depth x^(3/5) and its Dual tangent power,
    # against 60-digit decimal arithmetic, over the whole double range
    rng = np.random.default_rng(13)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 300), 10.0 ** rng.uniform(-3, 3, 300),
                         [5e-324, 2.2e-308, 1e-310, 1.7e308, 1.0, 2.0, 31.999999999999996, 1024.0]])
    L = olib()
    for e, fn in _pow35(L):
        for x in xs:
            got = fn(float(x))
            ex = _exact_pow(float(x), e)
            ulp = np.spacing(got)
            assert abs(Decimal(got) - ex) <= Decimal(ulp) * Decimal("0.5000001"), (x, e)


def test_pow_3_5_agrees_with_glibc_within_one_ulp():
    rng = np.random.default_rng(6)
    xs = 10.0 ** rng.uniform(-9, 4, 20000)
    L = olib()
    for e, fn in _pow35(L):
        a = np.array([fn(float(x)) for x in xs])
        b = np.array([L.hfo_libm_pow(float(x), e) for x in xs])
        assert (np.abs(a - b) / np.spacing(b)).max() <= 1.0
        assert np.mean(a != b) < 5e-3


def test_pow_3_5_special_values_and_exponent_gap():
    L = olib()
    (_, p35), (_, pm25) = _pow35(L)
    assert p35(0.0) == 0.0 and pm25(0.0) == float("inf") and pm25(-0.0) == float("inf")
    assert p35(float("inf")) == float("inf") and pm25(float("inf")) == 0.0
    assert np.isnan(p35(-1.0)) and np.isnan(pm25(-2.0)) and np.isnan(p35(float("nan")))
    # 32^(3/5) = 8 and 32^(-2/5) = 1/4 exactly, but the double exponent is
    # 3/5 - 2^-53/5, so the powers fall just below: one ulp under 8 and 1/4.
    assert p35(32.0) == float.fromhex("0x1.fffffffffffffp+2")
    assert pm25(32.0) == float.fromhex("0x1.fffffffffffffp-3")
    assert p35(1.0) == 1.0 and pm25(1.0) == 1.0


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "hydroflux_amd.h")).read()
    return sorted(set(re.findall(r"\b(hf2d_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from hydroflux_amd import _lib

    names = _declared_symbols()
    assert len(names) >= 20
    L = ctypes.CDLL(_lib.LIB_PATH)  # loading needs no GPU
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_library_is_gfx950_code_object():
    from hydroflux_amd import _lib

    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_size_limits_rejected_before_any_device_call():
    """Grid sides of 2^30 cells or more and batches of more than 16384 slots
    are refused by argument checks that run before the library touches the
    device (so this runs without a GPU); the handle stays null."""
    from hydroflux_amd import _lib

    L = _lib.lib()
    h = ctypes.c_void_p()
    side = 1 << 30
    assert L.hf2d_create(side, 8, 1.0, 1.0, 0, ctypes.byref(h)) == 2  # HF2D_INVALID_CONFIG
    assert L.hf2d_create(8, side, 1.0, 1.0, 0, ctypes.byref(h)) == 2
    assert L.hf2d_create(8, 8, 0.0, 1.0, 0, ctypes.byref(h)) == 2
    assert L.hf2d_create(8, 8, 1.0, 1.0, 99, ctypes.byref(h)) == 6  # unknown kind: HF2D_INVALID_ARGUMENT
    assert h.value is None
    for slots in (0, 16385):
        assert L.hf2d_batch_create(2, 1 << 20, slots, ctypes.byref(h)) == 6
    assert L.hf2d_batch_create(7, 1 << 20, 4, ctypes.byref(h)) == 6  # the fast kind is single-grid only
    assert h.value is None


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    from hydroflux_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(ImportError):
        _lib.lib()


def test_mesh_validation_mirrors_reference_panics():
    from hydroflux_amd import InvalidConfig, Mesh2D, ShapeMismatch

    with pytest.raises(InvalidConfig):
        Mesh2D(np.zeros((2, 2)), 0.0, 1.0, 0.03)
    with pytest.raises(InvalidConfig):
        Mesh2D(np.zeros((2, 2)), 1.0, -1.0, 0.03)
    with pytest.raises(InvalidConfig):
        Mesh2D(np.zeros((2, 2)), 1.0, 1.0, -0.01)
    with pytest.raises(ShapeMismatch):
        Mesh2D(np.zeros((2, 2)), 1.0, 1.0, np.zeros((3, 2)))
    with pytest.raises(InvalidConfig):
        Mesh2D(np.zeros((0, 2)), 1.0, 1.0, 0.0)


def _header_functions():
    """{name: [(param name, C type)], return type} of every entry point the
    header declares."""
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "hydroflux_amd.h")).read(), flags=re.S)
    out = {}
    for ret, name, params in re.findall(r"^\s*((?:const\s+)?\w+\s*\*?)\s*(hf2d_\w+)\s*\(([^;]*?)\)\s*;", txt,
                                        flags=re.M | re.S):
        ps = []
        for p in (q for q in params.split(",") if q.strip()):
            p = " ".join(p.split())
            arr = re.search(r"\[\d+\]$", p) is not None
            p = re.sub(r"\[\d+\]$", "", p)
            const = p.startswith("const ")
            base, pname = p[6:].replace("*", " ").split() if const else p.replace("*", " ").split()
            ps.append((pname, (const, base, p.count("*") + (1 if arr else 0))))
        out[name] = (ps, " ".join(ret.split()))
    return out


def _rust_type(const, base, stars):
    t = {"int32_t": "i32", "int64_t": "i64", "double": "f64", "int": "c_int"}.get(base, base)
    for k in range(stars):
        t = ("*const " if const and k == 0 else "*mut ") + t
    return t


def test_integration_rust_block_declares_every_entry_point():
    """INTEGRATION.md's hydroflux-hip-sys block declares every function of
    include/hydroflux_amd.h, with the header's parameter names, the Rust
    types of the C types, and the config struct's fields in order."""
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = md[md.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    rust = {}
    for name, params, ret in re.findall(r"pub fn (hf2d_\w+)\(([^)]*)\)\s*->\s*([^;]+);", block, flags=re.S):
        ps = [tuple(x.strip() for x in p.split(":", 1)) for p in params.split(",") if p.strip()]
        rust[name] = (ps, " ".join(ret.split()))
    header = _header_functions()
    assert set(rust) == set(header), set(rust) ^ set(header)
    ret_map = {"int": "c_int", "const char*": "*const c_char"}
    for name, (ps, ret) in header.items():
        rps, rret = rust[name]
        assert rret == ret_map[ret], name
        assert [p[0] for p in rps] == [p[0] for p in ps], name
        assert [p[1] for p in rps] == [_rust_type(*t) for _, t in ps], name
    # struct hf2d_config field order and names
    htxt = open(os.path.join(ROOT, "include", "hydroflux_amd.h")).read()
    hcfg = re.search(r"typedef struct hf2d_config \{(.*?)\} hf2d_config;", htxt, flags=re.S).group(1)
    hfields = re.findall(r"^\s*\w+\s+(\w+);", hcfg, flags=re.M)
    rcfg = re.search(r"pub struct hf2d_config \{(.*?)\}", md, flags=re.S).group(1)
    assert re.findall(r"pub (\w+):", rcfg) == hfields
