"""ctypes view of the CPU build of the stage-kernel tile code (libemu.so).

TEST INFRASTRUCTURE ONLY. Same method names as oracle.pyoracle.OracleSim and
hydroflux_amd.Solver2D so the parity tests can drive all three.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_LIBS = {}


def lib(asan: bool = False):
    name = "libemu_asan.so" if asan else "libemu.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"emulator library missing: {path}")
        L = C.CDLL(path)
        L.emu_create.restype = C.c_void_p
        L.emu_create.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        L.emu_destroy.argtypes = [C.c_void_p]
        L.emu_set_mesh.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.emu_set_state.argtypes = [C.c_void_p, _D, _D]
        L.emu_get_state.argtypes = [C.c_void_p, _D, _D]
        L.emu_set_boundaries.argtypes = [C.c_void_p, C.POINTER(C.c_int), _D]
        L.emu_configure.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                    C.c_int64, C.c_double, _I64, _I64, _D, C.c_int64]
        L.emu_ssprk2_step.argtypes = [C.c_void_p, C.c_double]
        L.emu_forward_euler_step.argtypes = [C.c_void_p, C.c_double]
        L.emu_run_until.restype = C.c_int
        L.emu_run_until.argtypes = [C.c_void_p, C.c_double]
        L.emu_run_steps.restype = C.c_int
        L.emu_run_steps.argtypes = [C.c_void_p, C.c_int64]
        L.emu_cfl_time_step.restype = C.c_double
        L.emu_cfl_time_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.emu_time.argtypes = [C.c_void_p, _D, _I64]
        L.emu_manning_friction_step.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.emu_apply_rain.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.emu_apply_point_sources.argtypes = [C.c_void_p, _I64, _I64, _D, C.c_int64, C.c_double]
        L.emu_set_active_tiles.argtypes = [C.c_void_p, C.c_int]
        L.emu_tiles_skipped.restype = C.c_int64
        L.emu_tiles_skipped.argtypes = []
        L.emu_set_chunked.argtypes = [C.c_void_p, C.c_int]
        L.emu_tile_runs.restype = C.c_int64
        L.emu_tile_runs.argtypes = [C.c_void_p, C.POINTER(C.c_ubyte), C.c_int64]
        _LIBS[name] = L
    return _LIBS[name]


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _grid(x, nr, nc):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), (nr, nc)))


class EmuSim:
    def __init__(self, n_rows, n_cols, dx, dy, dual=False, tile=None, asan=False):
        self.n_rows, self.n_cols = int(n_rows), int(n_cols)
        self.dual = bool(dual)
        tw, th = tile if tile else ((16, 16) if dual else (32, 16))
        self._L = lib(asan)
        self._h = self._L.emu_create(self.n_rows, self.n_cols, float(dx), float(dy), int(dual), tw, th)

    def set_active_tiles(self, on: bool):
        self._L.emu_set_active_tiles(self._h, 1 if on else 0)

    def tiles_skipped(self) -> int:
        return int(self._L.emu_tiles_skipped())

    def set_chunked(self, on: bool):
        """run_until in the device driver's chunks of steps (trailing no-op
        steps after the target) or stopping at the target."""
        self._L.emu_set_chunked(self._h, 1 if on else 0)

    def tile_runs(self) -> np.ndarray:
        """Per-tile counts of consecutive quiet steps of the active-tile list."""
        buf = np.zeros(1 << 16, np.uint8)
        n = int(self._L.emu_tile_runs(self._h, buf.ctypes.data_as(C.POINTER(C.c_ubyte)), buf.size))
        return buf[:n].copy()

    def close(self):
        if self._h:
            self._L.emu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, bed, manning, bed_tangent=None, manning_tangent=None):
        nr, nc = self.n_rows, self.n_cols
        b, m = _grid(bed, nr, nc), _grid(manning, nr, nc)
        bt = None if bed_tangent is None else _grid(bed_tangent, nr, nc)
        mt = None if manning_tangent is None else _grid(manning_tangent, nr, nc)
        self._L.emu_set_mesh(self._h, _p(b), _p(bt), _p(m), _p(mt))

    def set_state(self, state, tangent=None):
        s = np.ascontiguousarray(state, np.float64).reshape(self.n_rows, self.n_cols, 3)
        t = None if tangent is None else np.ascontiguousarray(tangent, np.float64).reshape(s.shape)
        self._L.emu_set_state(self._h, _p(s), _p(t))

    def get_state(self):
        s = np.empty((self.n_rows, self.n_cols, 3), np.float64)
        t = np.empty_like(s) if self.dual else None
        self._L.emu_get_state(self._h, _p(s), _p(t))
        return (s, t) if self.dual else s

    def set_boundaries(self, kinds, params=(0.0, 0.0, 0.0, 0.0)):
        k = (C.c_int * 4)(*[int(x) for x in kinds])
        p = np.asarray(params, np.float64)
        self._L.emu_set_boundaries(self._h, k, _p(p))

    def cfl_time_step(self, cfl, with_bcs=False):
        return self._L.emu_cfl_time_step(self._h, float(cfl), int(with_bcs))

    def ssprk2_step(self, dt):
        self._L.emu_ssprk2_step(self._h, float(dt))

    def forward_euler_step(self, dt):
        self._L.emu_forward_euler_step(self._h, float(dt))

    def manning_friction_step(self, dt, dry_tol):
        self._L.emu_manning_friction_step(self._h, float(dt), float(dry_tol))

    def apply_rain(self, rate, dt):
        self._L.emu_apply_rain(self._h, float(rate), float(dt))

    def apply_point_sources(self, sources, dt):
        rows = np.asarray([s[0] for s in sources], np.int64)
        cols = np.asarray([s[1] for s in sources], np.int64)
        q = np.asarray([s[2] for s in sources], np.float64)
        self._L.emu_apply_point_sources(self._h, rows.ctypes.data_as(_I64), cols.ctypes.data_as(_I64), _p(q),
                                        len(q), float(dt))

    def configure(self, cfl=0.4, integrator=1, cfl_with_bcs=True, max_dt=60.0, friction=True,
                  friction_dry_tol=1e-6, max_steps=10_000_000, rain_rate=0.0, sources=()):
        rows = np.asarray([s[0] for s in sources], np.int64)
        cols = np.asarray([s[1] for s in sources], np.int64)
        q = np.asarray([s[2] for s in sources], np.float64)
        self._L.emu_configure(self._h, float(cfl), int(integrator), int(cfl_with_bcs), float(max_dt), int(friction),
                              float(friction_dry_tol), int(max_steps), float(rain_rate),
                              rows.ctypes.data_as(_I64), cols.ctypes.data_as(_I64), _p(q), len(q))
        return 0

    def run_until(self, t):
        return self._L.emu_run_until(self._h, float(t))

    def run_steps(self, n):
        return self._L.emu_run_steps(self._h, int(n))

    def time(self):
        t, s = C.c_double(), C.c_int64()
        self._L.emu_time(self._h, C.byref(t), C.byref(s))
        return t.value, s.value


def add_rounded_stochastically(h, dh, keys):
    """The FP32 form's stochastically rounded h + dh (host build of the
    kernels' routine), elementwise over float32 arrays and uint32 keys."""
    L = lib()
    if not hasattr(L, "_sr_ready"):
        L.emu_add_rounded_stochastically.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L._sr_ready = True
    h = np.ascontiguousarray(h, np.float32)
    dh = np.ascontiguousarray(dh, np.float32)
    keys = np.ascontiguousarray(keys, np.uint32)
    out = np.empty_like(h)
    L.emu_add_rounded_stochastically(h.ctypes.data, dh.ctypes.data, keys.ctypes.data, h.size, out.ctypes.data)
    return out
