"""Batched scenarios on one GPU (hf2d_batch_* in include/hydroflux_amd.h).

A Batch holds many independent flood scenarios of different grid sizes in
one HBM arena. Each scenario runs to its own end time with its own dt
(Simulation::run_until semantics, sim.rs:231-246, plus the configured rain
and point sources); all of them advance together, one launch per RK stage
over all their tiles, the step replayed from a hipGraph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._lib import HfConfig, HfPointSource, lib
from .solver2d import H_DRY, _check, _sources

_D = C.POINTER(C.c_double)
KINDS = {"f64": 0, "dual": 1, "f32": 2}


class BatchFull(RuntimeError):
    """No free slot or not enough arena memory for the scenario (HF2D_CAPACITY)."""


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def footprint(kind: str, n_rows: int, n_cols: int, n_sources: int = 1) -> int:
    out = C.c_int64()
    _check(lib().hf2d_batch_footprint(KINDS[kind], int(n_rows), int(n_cols), int(n_sources), C.byref(out)))
    return out.value


@dataclass
class ScenarioSpec:
    """Inputs of one scenario (f64 host arrays)."""

    bed: np.ndarray
    manning: np.ndarray
    dx: float
    state: Optional[np.ndarray] = None          # (n_rows, n_cols, 3); default dry
    manning_tangent: Optional[np.ndarray] = None
    state_tangent: Optional[np.ndarray] = None
    bc_kinds: Sequence[int] = (1, 1, 1, 0)
    bc_params: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    cfl: float = 0.4
    max_dt: float = 1.0
    rain_rate: float = 0.0
    sources: Sequence = ()
    t_end: float = 600.0
    friction_dry_tol: float = H_DRY
    max_steps: int = 10_000_000
    dy: Optional[float] = None

    @property
    def shape(self):
        return self.bed.shape

    @property
    def n_cells(self):
        return int(self.bed.size)


class Batch:
    def __init__(self, kind: str = "f32", capacity_bytes: int = 8 << 30, max_slots: int = 1024):
        self.kind = kind
        self.max_slots = int(max_slots)
        h = C.c_void_p()
        _check(lib().hf2d_batch_create(KINDS[kind], int(capacity_bytes), self.max_slots, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().hf2d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc):
        if rc == 7:
            raise BatchFull(lib().hf2d_batch_last_error(self._h).decode())
        if rc:
            from .solver2d import _ERRORS, SimError

            raise _ERRORS.get(rc, SimError)(lib().hf2d_batch_last_error(self._h).decode())

    def add(self, spec: ScenarioSpec) -> int:
        nr, nc = spec.bed.shape
        bed = np.ascontiguousarray(spec.bed, np.float64)
        man = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.manning, np.float64), (nr, nc)))
        mt = None if spec.manning_tangent is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(spec.manning_tangent, np.float64), (nr, nc)))
        st = np.zeros((nr, nc, 3)) if spec.state is None else np.ascontiguousarray(spec.state, np.float64)
        tg = None if spec.state_tangent is None else np.ascontiguousarray(spec.state_tangent, np.float64)
        kinds = (C.c_int32 * 4)(*[int(x) for x in spec.bc_kinds])
        params = np.asarray(spec.bc_params, np.float64)
        cfg = HfConfig(float(spec.cfl), 1, 1, float(spec.max_dt), 1, 0, float(spec.friction_dry_tol),
                       int(spec.max_steps), float(spec.rain_rate))
        arr, k = _sources(spec.sources)
        slot = C.c_int32()
        dy = spec.dx if spec.dy is None else spec.dy
        rc = lib().hf2d_batch_add(self._h, nr, nc, float(spec.dx), float(dy), _p(bed), _p(man), _p(mt), _p(st),
                                  _p(tg), kinds, _p(params), C.byref(cfg), arr, k, float(spec.t_end),
                                  C.byref(slot))
        self._err(rc)
        return slot.value

    def remove(self, slot: int):
        self._err(lib().hf2d_batch_remove(self._h, int(slot)))

    def step(self, n: int, graph: bool = True):
        self._err(lib().hf2d_batch_step(self._h, int(n), 1 if graph else 0))

    def synchronize(self):
        self._err(lib().hf2d_batch_synchronize(self._h))

    def idle(self) -> bool:
        """True when every enqueued step has finished (does not wait)."""
        v = C.c_int32()
        self._err(lib().hf2d_batch_query(self._h, C.byref(v)))
        return v.value != 0

    def poll(self):
        """Per slot: (t, steps, status, done, used) arrays."""
        n = self.max_slots
        t = np.zeros(n)
        steps = np.zeros(n, np.int64)
        status = np.zeros(n, np.int32)
        done = np.zeros(n, np.int32)
        used = np.zeros(n, np.int32)
        i32 = C.POINTER(C.c_int32)
        self._err(lib().hf2d_batch_poll(self._h, n, _p(t), steps.ctypes.data_as(C.POINTER(C.c_int64)),
                                        status.ctypes.data_as(i32), done.ctypes.data_as(i32),
                                        used.ctypes.data_as(i32)))
        return t, steps, status, done, used

    def download(self, slot: int, shape, finished: bool = False):
        """State (and Dual tangents) of a slot. finished=True: the slot has
        reached its end time, and its state is copied without waiting for
        the steps still enqueued (hf2d_batch_download_done)."""
        nr, nc = shape
        s = np.empty((nr, nc, 3))
        t = np.empty_like(s) if self.kind == "dual" else None
        fn = lib().hf2d_batch_download_done if finished else lib().hf2d_batch_download
        self._err(fn(self._h, int(slot), _p(s), _p(t)))
        return (s, t) if self.kind == "dual" else s

    def free_bytes(self):
        lg, tot = C.c_int64(), C.c_int64()
        self._err(lib().hf2d_batch_free_bytes(self._h, C.byref(lg), C.byref(tot)))
        return lg.value, tot.value

    def configs(self):
        n = C.c_int32()
        self._err(lib().hf2d_batch_config_count(self._h, C.byref(n)))
        out = []
        for i in range(n.value):
            w, h, t = C.c_int32(), C.c_int32(), C.c_int32()
            self._err(lib().hf2d_batch_config_info(self._h, i, C.byref(w), C.byref(h), C.byref(t)))
            out.append((w.value, h.value, t.value))
        return out

    def set_config(self, index: int):
        self._err(lib().hf2d_batch_set_config(self._h, int(index)))

    def autotune(self, steps: int = 8):
        """Time every tile configuration on the resident mix; keep the fastest."""
        n = len(self.configs())
        ms = np.zeros(n)
        best = C.c_int32()
        self._err(lib().hf2d_batch_autotune(self._h, int(steps), C.byref(best), _p(ms)))
        return best.value, ms
