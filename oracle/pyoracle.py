"""ctypes view of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the
product package. The method names mirror hydroflux_amd.solver2d.Simulation so
one test body can drive either backend.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.hfo_create.restype = C.c_void_p
        L.hfo_create.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int]
        L.hfo_destroy.argtypes = [C.c_void_p]
        L.hfo_set_pow_mode.argtypes = [C.c_int]
        L.hfo_set_threads.argtypes = [C.c_int]
        L.hfo_set_mesh.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.hfo_set_state.argtypes = [C.c_void_p, _D, _D]
        L.hfo_get_state.argtypes = [C.c_void_p, _D, _D]
        L.hfo_set_boundaries.argtypes = [C.c_void_p, C.POINTER(C.c_int), _D]
        L.hfo_max_wave_speeds.argtypes = [C.c_void_p, _D, _D]
        L.hfo_cfl_time_step.restype = C.c_double
        L.hfo_cfl_time_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.hfo_forward_euler_step.argtypes = [C.c_void_p, C.c_double]
        L.hfo_ssprk2_step.argtypes = [C.c_void_p, C.c_double]
        L.hfo_manning_friction_step.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.hfo_apply_rain.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.hfo_apply_point_sources.argtypes = [C.c_void_p, _I64, _I64, _D, C.c_int64, C.c_double]
        L.hfo_configure.restype = C.c_int
        L.hfo_configure.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double,
                                    C.c_int, C.c_double, C.c_int64, C.c_double,
                                    _I64, _I64, _D, C.c_int64]
        L.hfo_step.restype = C.c_int
        L.hfo_step.argtypes = [C.c_void_p, C.c_double, _D]
        L.hfo_run_until.restype = C.c_int
        L.hfo_run_until.argtypes = [C.c_void_p, C.c_double]
        L.hfo_run_steps.restype = C.c_int
        L.hfo_run_steps.argtypes = [C.c_void_p, C.c_int64]
        L.hfo_time.argtypes = [C.c_void_p, _D, _I64]
        L.hfo_set_series.argtypes = [C.c_void_p, C.c_int, C.c_int, _D, _D, C.c_double, C.c_double, C.c_double]
        L.hfo_set_gauges.argtypes = [C.c_void_p, C.c_int, _I64, _I64]
        L.hfo_gauge_values.restype = C.c_int64
        L.hfo_gauge_values.argtypes = [C.c_void_p, _D, C.c_int64]
        L.hfo_set_time.argtypes = [C.c_void_p, C.c_double, C.c_int64]
        for name in ("hfo_pow_4_3", "hfo_pow_1_3m"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        L.hfo_libm_pow.restype = C.c_double
        L.hfo_libm_pow.argtypes = [C.c_double, C.c_double]
        L.hfo_dualn_probe.argtypes = [C.c_double, C.c_double, C.c_double, _D]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def set_pow_mode(mode: int) -> None:
    """0: glibc pow (as the Rust reference); 1: the kernels' correctly rounded pow."""
    lib().hfo_set_pow_mode(int(mode))


def set_threads(n: int) -> None:
    lib().hfo_set_threads(int(n))


def set_dualn_div_mode(mode: int) -> None:
    """DualN quotient value: 0 the reference's a * (1 / b); 1 a / b (a
    test switch that makes DualN follow the f64 value path)."""
    lib().hfo_set_dualn_div_mode(int(mode))


def dualn_probe(a, b, c):
    """The reference's DualN<3> unit-test function and its three Dual passes:
    (vec value, vec tangents[3], [(dual value, dual tangent)] * 3)."""
    out = np.zeros(10)
    lib().hfo_dualn_probe(float(a), float(b), float(c), _p(out))
    return out[0], out[1:4], [(out[4 + 2 * k], out[5 + 2 * k]) for k in range(3)]


class OracleSim:
    """One grid on the CPU restatement. States are (n_rows, n_cols, 3) float64;
    Dual tangents the same shape; DualN<N> (tangents=N in 2, 4, 8, 16) tangents
    (n_rows, n_cols, 3, N) for the state and (n_rows, n_cols, N) for bed and
    Manning."""

    def __init__(self, n_rows, n_cols, dx, dy, dual=False, tangents=None):
        self.n_rows, self.n_cols, self.dx, self.dy = int(n_rows), int(n_cols), float(dx), float(dy)
        if tangents is not None and tangents not in (2, 4, 8, 16):
            raise ValueError("tangents must be 2, 4, 8 or 16")
        self.n_tan = int(tangents) if tangents else (1 if dual else 0)
        self.dual = self.n_tan > 0
        self._h = lib().hfo_create(self.n_rows, self.n_cols, self.dx, self.dy, self.n_tan)

    def _tshape(self, fields=None):
        base = (self.n_rows, self.n_cols) if fields is None else (self.n_rows, self.n_cols, fields)
        return base if self.n_tan == 1 else base + (self.n_tan,)

    def close(self):
        if self._h:
            lib().hfo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, bed, manning, bed_tangent=None, manning_tangent=None):
        n = self.n_rows * self.n_cols
        bed = np.ascontiguousarray(np.broadcast_to(np.asarray(bed, np.float64), (self.n_rows, self.n_cols)))
        man = np.ascontiguousarray(np.broadcast_to(np.asarray(manning, np.float64), (self.n_rows, self.n_cols)))
        bt = None if bed_tangent is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(bed_tangent, np.float64), self._tshape()))
        mt = None if manning_tangent is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(manning_tangent, np.float64), self._tshape()))
        assert bed.size == n
        lib().hfo_set_mesh(self._h, _p(bed), _p(bt), _p(man), _p(mt))

    def set_state(self, state, tangent=None):
        s = np.ascontiguousarray(state, np.float64).reshape(self.n_rows, self.n_cols, 3)
        t = None if tangent is None else np.ascontiguousarray(tangent, np.float64).reshape(self._tshape(3))
        lib().hfo_set_state(self._h, _p(s), _p(t))

    def get_state(self):
        s = np.empty((self.n_rows, self.n_cols, 3), np.float64)
        t = np.empty(self._tshape(3)) if self.dual else None
        lib().hfo_get_state(self._h, _p(s), _p(t))
        return (s, t) if self.dual else s

    def set_boundaries(self, kinds, params=(0.0, 0.0, 0.0, 0.0)):
        k = (C.c_int * 4)(*[int(x) for x in kinds])
        p = np.asarray(params, np.float64)
        lib().hfo_set_boundaries(self._h, k, _p(p))

    def max_wave_speeds(self):
        sx, sy = C.c_double(), C.c_double()
        lib().hfo_max_wave_speeds(self._h, C.byref(sx), C.byref(sy))
        return sx.value, sy.value

    def cfl_time_step(self, cfl, with_bcs=False):
        return lib().hfo_cfl_time_step(self._h, float(cfl), int(with_bcs))

    def forward_euler_step(self, dt):
        lib().hfo_forward_euler_step(self._h, float(dt))

    def ssprk2_step(self, dt):
        lib().hfo_ssprk2_step(self._h, float(dt))

    def manning_friction_step(self, dt, dry_tol):
        lib().hfo_manning_friction_step(self._h, float(dt), float(dry_tol))

    def apply_rain(self, rate, dt):
        lib().hfo_apply_rain(self._h, float(rate), float(dt))

    def apply_point_sources(self, sources, dt):
        rows = np.asarray([s[0] for s in sources], np.int64)
        cols = np.asarray([s[1] for s in sources], np.int64)
        q = np.asarray([s[2] for s in sources], np.float64)
        lib().hfo_apply_point_sources(self._h, rows.ctypes.data_as(_I64), cols.ctypes.data_as(_I64),
                                      _p(q), len(q), float(dt))

    def configure(self, cfl=0.4, integrator=1, cfl_with_bcs=True, max_dt=60.0, friction=True,
                  friction_dry_tol=1e-6, max_steps=10_000_000, rain_rate=0.0, sources=()):
        rows = np.asarray([s[0] for s in sources], np.int64)
        cols = np.asarray([s[1] for s in sources], np.int64)
        q = np.asarray([s[2] for s in sources], np.float64)
        rc = lib().hfo_configure(self._h, float(cfl), int(integrator), int(cfl_with_bcs), float(max_dt),
                                 int(friction), float(friction_dry_tol), int(max_steps), float(rain_rate),
                                 rows.ctypes.data_as(_I64), cols.ctypes.data_as(_I64), _p(q), len(q))
        return rc

    def step(self, dt_cap=float("inf")):
        dt = C.c_double()
        rc = lib().hfo_step(self._h, float(dt_cap), C.byref(dt))
        return rc, dt.value

    def run_until(self, t_target):
        return lib().hfo_run_until(self._h, float(t_target))

    def set_forcing_series(self, which, times, values, time_div=1.0, scale_mul=1.0, scale_div=1.0):
        t = np.ascontiguousarray(times, np.float64)
        v = np.ascontiguousarray(values, np.float64)
        lib().hfo_set_series(self._h, {"rain": 0, "sources": 1}[which], int(t.size), _p(t), _p(v),
                             float(time_div), float(scale_mul), float(scale_div))

    def set_gauges(self, cells, capacity=None):
        rows = np.ascontiguousarray([c[0] for c in cells], np.int64)
        cols = np.ascontiguousarray([c[1] for c in cells], np.int64)
        self._n_gauges = len(rows)
        lib().hfo_set_gauges(self._h, len(rows), rows.ctypes.data_as(_I64), cols.ctypes.data_as(_I64))

    def read_gauges(self):
        n = lib().hfo_gauge_values(self._h, None, 0)
        buf = np.zeros(n)
        lib().hfo_gauge_values(self._h, _p(buf), n)
        buf = buf.reshape(-1, self._n_gauges + 1)
        return buf[:, 0].copy(), buf[:, 1:].copy()

    def run_steps(self, n):
        return lib().hfo_run_steps(self._h, int(n))

    def time(self):
        t, s = C.c_double(), C.c_int64()
        lib().hfo_time(self._h, C.byref(t), C.byref(s))
        return t.value, s.value
