"""Host-side mirror of hydroflux solver-2d's step API over the gfx950 library.

Names, argument meaning and error behaviour follow the reference's public
API (solver-2d/src/lib.rs:74-91):

- ``Boundary`` / ``Boundaries2D``      boundary.rs:80-160
- ``Mesh2D``                           geometry.rs:31-113 (f64 or Dual fields)
- ``PointSource``                      source.rs:36-44
- ``Grid2D``                           a device-resident state + workspace; its
  methods are the free functions ``max_wave_speeds``, ``cfl_time_step(_with_bcs)``,
  ``forward_euler_step``, ``ssprk2_step`` (update.rs:113-1287),
  ``manning_friction_step``, ``apply_rain``, ``apply_point_sources``
  (source.rs:52-165) applied to that state
- ``SimulationConfig`` / ``Simulation`` / ``SimError``   sim.rs:36-246

States cross the boundary as float64 arrays of shape (n_rows, n_cols, 3)
holding (h, hu, hv) per cell (the reference's Array2<Conserved2D>); a Dual
grid carries a second array of the same shape with the tangents.
Everything computes on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from ._lib import HfConfig, HfPointSource, lib

GRAVITY = 9.81   # lib.rs:43
H_DRY = 1.0e-6   # lib.rs:63
H_VEL = 1.0e-5   # lib.rs:72

_D = C.POINTER(C.c_double)


# ------------------------------------------------------------------ errors
class SimError(RuntimeError):
    """sim.rs:81-115. ``code`` is the C ABI status."""

    code = 0


class ShapeMismatch(SimError, ValueError):
    code = 1


class InvalidConfig(SimError, ValueError):
    code = 2


class DegenerateDt(SimError):
    code = 3


class StepBudgetExhausted(SimError):
    code = 4


class DeviceError(SimError):
    code = 5


_ERRORS = {1: ShapeMismatch, 2: InvalidConfig, 3: DegenerateDt, 4: StepBudgetExhausted, 5: DeviceError,
           6: ValueError}


def _check(rc: int, handle=None) -> None:
    if rc == 0:
        return
    msg = ""
    if handle:
        raw = lib().hf2d_last_error(handle)
        msg = raw.decode() if raw else ""
    raise _ERRORS.get(rc, SimError)(msg or f"hydroflux_amd status {rc}")


# ------------------------------------------------------------------ boundaries
@dataclass(frozen=True)
class Boundary:
    """boundary.rs:80-106. Use the constructors below."""

    kind: int
    param: float = 0.0

    @staticmethod
    def transmissive() -> "Boundary":
        return Boundary(0)

    @staticmethod
    def wall() -> "Boundary":
        return Boundary(1)

    @staticmethod
    def discharge(q: float) -> "Boundary":
        return Boundary(2, float(q))

    @staticmethod
    def depth(h: float) -> "Boundary":
        return Boundary(3, float(h))


Boundary.TRANSMISSIVE = Boundary(0)
Boundary.WALL = Boundary(1)


@dataclass(frozen=True)
class Boundaries2D:
    """boundary.rs:131-160 (sides N, S, W, E)."""

    north: Boundary = Boundary.WALL
    south: Boundary = Boundary.WALL
    west: Boundary = Boundary.WALL
    east: Boundary = Boundary.WALL

    def kinds(self):
        return [self.north.kind, self.south.kind, self.west.kind, self.east.kind]

    def params(self):
        return [self.north.param, self.south.param, self.west.param, self.east.param]


Boundaries2D.WALLS = Boundaries2D()
Boundaries2D.TRANSMISSIVE = Boundaries2D(Boundary.TRANSMISSIVE, Boundary.TRANSMISSIVE,
                                         Boundary.TRANSMISSIVE, Boundary.TRANSMISSIVE)


# ------------------------------------------------------------------ mesh
@dataclass
class Mesh2D:
    """geometry.rs:31-113. ``bed_tangent``/``manning_tangent`` make it a
    Mesh2DG<Dual> (None means tangent 0)."""

    bed: np.ndarray
    dx: float
    dy: float
    manning: np.ndarray
    bed_tangent: Optional[np.ndarray] = None
    manning_tangent: Optional[np.ndarray] = None

    def __post_init__(self):
        # geometry.rs:58-76 / 90-113 panics -> InvalidConfig / ShapeMismatch
        self.bed = np.ascontiguousarray(self.bed, np.float64)
        if not (self.dx > 0.0):
            raise InvalidConfig(f"dx must be strictly positive (got {self.dx})")
        if not (self.dy > 0.0):
            raise InvalidConfig(f"dy must be strictly positive (got {self.dy})")
        if self.bed.ndim != 2 or self.bed.size == 0:
            raise InvalidConfig("mesh must have at least one cell")
        man = np.asarray(self.manning, np.float64)
        if man.ndim == 0:
            man = np.full(self.bed.shape, float(man))
        if man.shape != self.bed.shape:
            raise ShapeMismatch(f"manning field shape {man.shape} must match bed shape {self.bed.shape}")
        if not np.all(man >= 0.0):
            raise InvalidConfig("Manning n must be non-negative at every cell")
        self.manning = np.ascontiguousarray(man)
        for name in ("bed_tangent", "manning_tangent"):
            v = getattr(self, name)
            if v is not None:
                v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), self.bed.shape))
                setattr(self, name, v)

    @classmethod
    def new(cls, bed, dx, dy, manning: float):
        return cls(bed, dx, dy, np.full(np.shape(bed), float(manning)))

    @classmethod
    def with_manning_field(cls, bed, dx, dy, manning, **tangents):
        return cls(bed, dx, dy, manning, **tangents)

    @property
    def n_rows(self) -> int:
        return self.bed.shape[0]

    @property
    def n_cols(self) -> int:
        return self.bed.shape[1]


@dataclass(frozen=True)
class PointSource:
    """source.rs:36-44."""

    row: int
    col: int
    q_mass: float


def _sources(srcs):
    srcs = list(srcs or [])
    arr = (HfPointSource * max(len(srcs), 1))()
    for k, s in enumerate(srcs):
        if isinstance(s, PointSource):
            arr[k] = HfPointSource(int(s.row), int(s.col), float(s.q_mass))
        else:
            arr[k] = HfPointSource(int(s[0]), int(s[1]), float(s[2]))
    return arr, len(srcs)


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


# ------------------------------------------------------------------ device grid
class Grid2D:
    """One device-resident grid: state U, mesh, boundaries and workspace.

    The methods apply the reference's free functions to the state held on the
    device; ``set_state``/``get_state`` are the only host<->device copies.
    """

    KINDS = {"f64": 0, "dual": 1, "f32": 2, "dual2": 3, "dual4": 4, "dual8": 5, "dual16": 6, "f64fast": 7}
    TANGENTS = {"f64": 0, "dual": 1, "f32": 0, "dual2": 2, "dual4": 4, "dual8": 8, "dual16": 16, "f64fast": 0}

    def __init__(self, n_rows: int, n_cols: int, dx: float, dy: float, dual: bool = False, kind: str = None):
        kind = kind or ("dual" if dual else "f64")
        if kind not in self.KINDS:
            raise ValueError(f"kind must be one of {sorted(self.KINDS)}")
        h = C.c_void_p()
        rc = lib().hf2d_create(int(n_rows), int(n_cols), float(dx), float(dy), self.KINDS[kind], C.byref(h))
        _check(rc)
        self._h = h
        self.n_rows, self.n_cols, self.dx, self.dy = int(n_rows), int(n_cols), float(dx), float(dy)
        self.kind = kind
        self.n_tan = self.TANGENTS[kind]
        self.dual = self.n_tan > 0
        self._n_gauges = 0

    def _tshape(self, fields=None):
        """Tangent array shape: Dual like the values; DualN<N> with a last axis N."""
        base = (self.n_rows, self.n_cols) if fields is None else (self.n_rows, self.n_cols, fields)
        return base if self.n_tan == 1 else base + (self.n_tan,)

    @classmethod
    def for_mesh(cls, mesh: Mesh2D, dual: Optional[bool] = None) -> "Grid2D":
        if dual is None:
            dual = mesh.bed_tangent is not None or mesh.manning_tangent is not None
        g = cls(mesh.n_rows, mesh.n_cols, mesh.dx, mesh.dy, dual)
        g.set_mesh(mesh.bed, mesh.manning, mesh.bed_tangent, mesh.manning_tangent)
        return g

    def close(self):
        if getattr(self, "_h", None):
            lib().hf2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _grid(self, x):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), (self.n_rows, self.n_cols)))

    # -- data movement
    def set_mesh(self, bed, manning, bed_tangent=None, manning_tangent=None):
        b, m = self._grid(bed), self._grid(manning)
        tan = (lambda x: np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), self._tshape()))) \
            if self.n_tan > 1 else self._grid
        bt = None if bed_tangent is None else tan(bed_tangent)
        mt = None if manning_tangent is None else tan(manning_tangent)
        _check(lib().hf2d_upload_mesh(self._h, _p(b), _p(bt), _p(m), _p(mt)), self._h)

    def set_state(self, state, tangent=None):
        s = np.ascontiguousarray(state, np.float64)
        if s.shape != (self.n_rows, self.n_cols, 3):
            raise ShapeMismatch(f"states shape {s.shape[:2]} must match mesh ({self.n_rows}, {self.n_cols})")
        t = None
        if tangent is not None:
            t = np.ascontiguousarray(tangent, np.float64)
            if t.shape != self._tshape(3):
                raise ShapeMismatch(f"tangent shape must be {self._tshape(3)}")
        _check(lib().hf2d_upload_state(self._h, _p(s), _p(t)), self._h)

    def get_state(self):
        s = np.empty((self.n_rows, self.n_cols, 3), np.float64)
        t = np.empty(self._tshape(3)) if self.dual else None
        _check(lib().hf2d_download_state(self._h, _p(s), _p(t)), self._h)
        return (s, t) if self.dual else s

    def set_boundaries(self, kinds_or_bcs, params=(0.0, 0.0, 0.0, 0.0)):
        if isinstance(kinds_or_bcs, Boundaries2D):
            kinds, params = kinds_or_bcs.kinds(), kinds_or_bcs.params()
        else:
            kinds = list(kinds_or_bcs)
        k = (C.c_int32 * 4)(*[int(x) for x in kinds])
        p = np.asarray(params, np.float64)
        _check(lib().hf2d_set_boundaries(self._h, k, _p(p)), self._h)

    # -- reference free functions on the device state
    def max_wave_speeds(self):
        sx, sy = C.c_double(), C.c_double()
        _check(lib().hf2d_max_wave_speeds(self._h, C.byref(sx), C.byref(sy)), self._h)
        return sx.value, sy.value

    def cfl_time_step(self, cfl: float, with_bcs: bool = False) -> float:
        dt = C.c_double()
        _check(lib().hf2d_cfl_time_step(self._h, float(cfl), 1 if with_bcs else 0, C.byref(dt)), self._h)
        return dt.value

    def cfl_time_step_with_bcs(self, cfl: float) -> float:
        return self.cfl_time_step(cfl, True)

    def forward_euler_step(self, dt: float):
        _check(lib().hf2d_forward_euler_step(self._h, float(dt)), self._h)

    def ssprk2_step(self, dt: float):
        _check(lib().hf2d_ssprk2_step(self._h, float(dt)), self._h)

    def manning_friction_step(self, dt: float, dry_tol: float):
        _check(lib().hf2d_manning_friction_step(self._h, float(dt), float(dry_tol)), self._h)

    def apply_rain(self, rate: float, dt: float):
        _check(lib().hf2d_apply_rain(self._h, float(rate), float(dt)), self._h)

    def apply_point_sources(self, sources: Sequence, dt: float):
        arr, k = _sources(sources)
        _check(lib().hf2d_apply_point_sources(self._h, arr, k, float(dt)), self._h)

    # -- device-clocked driver
    def configure(self, cfl=0.4, integrator=1, cfl_with_bcs=True, max_dt=60.0, friction=True,
                  friction_dry_tol=H_DRY, max_steps=10_000_000, rain_rate=0.0, sources=(), active_tiles=True):
        """SimulationConfig + forcing. active_tiles=False turns off the
        compacted active-tile list (same results; for timing comparisons);
        "always" uses it on grids of any size (by default only above 512
        tiles)."""
        if active_tiles not in (True, False, "always"):
            raise ValueError("active_tiles must be True, False or 'always'")
        opts = 2 if active_tiles == "always" else 0 if active_tiles else 1
        cfg = HfConfig(float(cfl), int(integrator), 1 if cfl_with_bcs else 0, float(max_dt),
                       1 if friction else 0, opts, float(friction_dry_tol), int(max_steps),
                       float(rain_rate))
        arr, k = _sources(sources)
        rc = lib().hf2d_configure(self._h, C.byref(cfg), arr, k)
        _check(rc, self._h)
        return rc

    def step(self, dt_cap: float = math.inf) -> float:
        dt = C.c_double()
        _check(lib().hf2d_step(self._h, float(dt_cap), C.byref(dt)), self._h)
        return dt.value

    def run_until(self, t_target: float):
        _check(lib().hf2d_run_until(self._h, float(t_target)), self._h)

    def run_steps(self, n: int):
        _check(lib().hf2d_run_steps(self._h, int(n)), self._h)

    def enqueue_steps(self, n: int):
        _check(lib().hf2d_enqueue_steps(self._h, int(n)), self._h)

    def synchronize(self):
        _check(lib().hf2d_synchronize(self._h), self._h)

    def time(self):
        t, s = C.c_double(), C.c_int64()
        _check(lib().hf2d_get_time(self._h, C.byref(t), C.byref(s)), self._h)
        return t.value, s.value

    def set_time(self, t: float, steps: int):
        _check(lib().hf2d_set_time(self._h, float(t), int(steps)), self._h)

    def set_forcing_series(self, which: str, times, values, time_div=1.0, scale_mul=1.0, scale_div=1.0):
        """Piecewise-linear forcing evaluated on the device at t + dt of every
        driver step: value = (interp(t / time_div) * scale_mul) / scale_div.
        which = "rain" (rate, m/s) or "sources" (q_mass of every configured
        point source, m^3/s); times=() removes it."""
        k = {"rain": 0, "sources": 1}[which]
        t = np.ascontiguousarray(times, np.float64)
        v = np.ascontiguousarray(values, np.float64)
        if t.shape != v.shape or t.ndim != 1:
            raise ValueError("times and values must be 1-D arrays of one length")
        _check(lib().hf2d_set_forcing_series(self._h, k, int(t.size), _p(t) if t.size else None,
                                             _p(v) if v.size else None, float(time_div), float(scale_mul),
                                             float(scale_div)), self._h)

    def set_gauges(self, cells, capacity: int):
        """Record (t, h at each (row, col) of `cells`) after every driver
        step, into a device buffer of `capacity` rows; the current state is
        the first row."""
        rows = np.ascontiguousarray([c[0] for c in cells], np.int64)
        cols = np.ascontiguousarray([c[1] for c in cells], np.int64)
        self._n_gauges = len(rows)
        i64 = C.POINTER(C.c_int64)
        _check(lib().hf2d_set_gauges(self._h, len(rows), rows.ctypes.data_as(i64), cols.ctypes.data_as(i64),
                                     int(capacity)), self._h)

    def read_gauges(self):
        """(times (n_rows,), depths (n_rows, n_gauges)) recorded so far."""
        n = C.c_int64()
        _check(lib().hf2d_read_gauges(self._h, 0, None, C.byref(n)), self._h)
        buf = np.zeros((n.value, self._n_gauges + 1))
        if n.value:
            _check(lib().hf2d_read_gauges(self._h, n.value, _p(buf), C.byref(n)), self._h)
        return buf[:, 0].copy(), buf[:, 1:].copy()

    def total_volume(self, noalloc: bool = False) -> float:
        """sum(h) dx dy of the current state (FP64, compensated). noalloc:
        through hf2d_total_volume_noalloc (library-owned scratch, no device
        allocation; same bits)."""
        v = C.c_double()
        fn = lib().hf2d_total_volume_noalloc if noalloc else lib().hf2d_total_volume
        _check(fn(self._h, C.byref(v)), self._h)
        return v.value

    def set_kernel_timing(self, on: bool):
        _check(lib().hf2d_set_kernel_timing(self._h, 1 if on else 0), self._h)

    def kernel_timing(self):
        ms = (C.c_double * 2)()
        n = C.c_int64()
        _check(lib().hf2d_kernel_timing(self._h, ms, C.byref(n)), self._h)
        return ms[0], ms[1], n.value


# ------------------------------------------------------------------ driver
class Integrator:
    """sim.rs:37-46."""

    FORWARD_EULER = 0
    SSP_RK2 = 1


@dataclass
class SimulationConfig:
    """sim.rs:49-77."""

    cfl: float = 0.4
    integrator: int = Integrator.SSP_RK2
    boundaries: Boundaries2D = field(default_factory=lambda: Boundaries2D.WALLS)
    max_steps: int = 10_000_000
    max_dt: float = 60.0


class Simulation:
    """sim.rs:119-246: owns mesh, state, workspace and clock (all on the GPU).

    ``rain_rate`` and ``sources`` are the forcing the reference's operational
    callers apply after every step (uk_ea_test8a_official.rs:206-217); here
    they run inside the same device step.
    """

    def __init__(self, mesh: Mesh2D, states, config: SimulationConfig = None, state_tangent=None,
                 rain_rate: float = 0.0, sources: Sequence = ()):
        config = config or SimulationConfig()
        st = np.asarray(states, np.float64)
        if st.shape[:2] != (mesh.n_rows, mesh.n_cols):
            raise ShapeMismatch(f"states shape {st.shape[:2]} must match mesh ({mesh.n_rows}, {mesh.n_cols})")
        if not (config.cfl > 0.0 and config.cfl <= 1.0):
            raise InvalidConfig(f"cfl must be in (0, 1], got {config.cfl}")
        if not (config.max_dt > 0.0):
            raise InvalidConfig(f"max_dt must be positive, got {config.max_dt}")
        self.mesh = mesh
        self.config = config
        self.grid = Grid2D.for_mesh(mesh)
        self.grid.set_state(st, state_tangent)
        self.grid.set_boundaries(config.boundaries)
        self._rain, self._sources = float(rain_rate), list(sources)
        self._configure()

    def _configure(self):
        c = self.config
        self.grid.configure(cfl=c.cfl, integrator=c.integrator, cfl_with_bcs=True, max_dt=c.max_dt,
                            friction=True, friction_dry_tol=H_DRY, max_steps=c.max_steps,
                            rain_rate=self._rain, sources=self._sources)

    def time(self) -> float:
        return self.grid.time()[0]

    def steps(self) -> int:
        return self.grid.time()[1]

    def states(self):
        return self.grid.get_state()

    def set_states(self, states, tangent=None):
        """Write access (states_mut, sim.rs:178-181)."""
        self.grid.set_state(states, tangent)

    def set_boundaries(self, bcs: Boundaries2D):
        self.config.boundaries = bcs
        self.grid.set_boundaries(bcs)

    def step(self, dt_cap: float = math.inf) -> float:
        return self.grid.step(dt_cap)

    def run_until(self, t_target: float) -> None:
        self.grid.run_until(t_target)


def max_depth_exceeds_relief(states, mesh: Mesh2D, factor: float) -> bool:
    """sim.rs:265-271."""
    h = np.asarray(states)[..., 0]
    max_depth = max(0.0, float(np.max(h))) if h.size else 0.0
    relief = max(float(np.max(mesh.bed)) - float(np.min(mesh.bed)), 1.0)
    return max_depth > factor * relief
