"""Reference test/fixture scenarios, rebuilt from the reference's own setups.

Each builder returns a Case with the mesh, the initial state and the
boundaries exactly as the reference example or test constructs them (file:line
cited), and `hand_loop` replays the reference's caller loop on any backend
exposing the Solver2D-style methods (the oracle, the CPU emulator of the
kernels, or the GPU library).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

G = 9.81
WALL, TRANS, DISCHARGE, DEPTH = 1, 0, 2, 3
DATA = "papers/01_review/figures/data"


@dataclass
class Case:
    name: str
    nr: int
    nc: int
    dx: float
    dy: float
    bed: np.ndarray
    manning: np.ndarray
    state: np.ndarray
    bc_kinds: tuple
    bc_params: tuple = (0.0, 0.0, 0.0, 0.0)
    extra: dict = field(default_factory=dict)

    def load(self, sim, manning_tangent=None, bed_tangent=None, state_tangent=None):
        sim.set_mesh(self.bed, self.manning, bed_tangent=bed_tangent, manning_tangent=manning_tangent)
        if state_tangent is not None:
            sim.set_state(self.state, state_tangent)
        else:
            sim.set_state(self.state)
        sim.set_boundaries(self.bc_kinds, self.bc_params)
        return sim


def hand_loop(sim, t_end, cfl=0.4, friction_tol=None, max_steps=None, with_bcs=False):
    """`while t < t_end { dt = cfl_time_step(..).min(t_end - t); ssprk2_step(dt);
    [manning_friction_step(dt, tol)]; t += dt }` (gen_verification_data.rs:67-72)."""
    t = 0.0
    k = 0
    while t < t_end:
        if max_steps is not None and k >= max_steps:
            break
        dt = min(sim.cfl_time_step(cfl, with_bcs), t_end - t)
        if not (math.isfinite(dt) and dt > 0):
            break
        sim.ssprk2_step(dt)
        if friction_tol is not None:
            sim.manning_friction_step(dt, friction_tol)
        t += dt
        k += 1
    return t, k


def stoker(n_cols=400, n_rows=3):
    """Stoker dam-break on a dry bed (gen_verification_data.rs:46-82)."""
    h_l, x_dam, length = 1.0, 50.0, 100.0
    dx = length / n_cols
    st = np.zeros((n_rows, n_cols, 3))
    x = (np.arange(n_cols) + 0.5) * dx
    st[:, x < x_dam, 0] = h_l
    return Case(f"stoker_{n_rows}x{n_cols}", n_rows, n_cols, dx, dx, np.zeros((n_rows, n_cols)),
                np.zeros((n_rows, n_cols)), st, (WALL, WALL, TRANS, TRANS), extra={"t_end": 4.0})


def stoker_depth(x, t, h_l=1.0, x_dam=50.0):
    c_l = math.sqrt(G * h_l)
    xi = (x - x_dam) / t
    if xi <= -c_l:
        return h_l
    if xi >= 2.0 * c_l:
        return 0.0
    return max((2.0 * c_l - xi) ** 2 / (9.0 * G), 0.0)


def thacker(n=80):
    """Thacker planar oscillation in a paraboloid (gen_verification_data.rs:130-170)."""
    h0, a, b = 0.1, 1.0, 0.1
    half = 1.25
    dx = 2.0 * half / n
    omega = math.sqrt(2.0 * G * h0) / a
    period = 2.0 * math.pi / omega
    x = -half + (np.arange(n) + 0.5) * dx
    bed = np.empty((n, n))
    st = np.zeros((n, n, 3))
    u0 = -b * omega * math.sin(0.0)
    v0 = b * omega * math.cos(0.0)
    for i in range(n):
        for j in range(n):
            bed[i, j] = h0 * ((x[j] * x[j] + x[i] * x[i]) / (a * a) - 1.0)
            xc, yc = b * math.cos(0.0), b * math.sin(0.0)
            r2 = (x[j] - xc) ** 2 + (x[i] - yc) ** 2
            h = max(h0 / (a * a) * (a * a - r2), 0.0)
            st[i, j] = (h, h * u0, h * v0)
    return Case(f"thacker_{n}", n, n, dx, dx, bed, np.zeros((n, n)), st, (WALL, WALL, WALL, WALL),
                extra={"t_end": period / 2.0})


def macdonald():
    """MacDonald uniform flow: Discharge in, Depth out, friction (gen_verification_data.rs:86-118)."""
    q, slope, manning = 1.0, 0.01, 0.03
    h_n = (manning * q / math.sqrt(slope)) ** (3.0 / 5.0)
    nr, nc, dx = 5, 50, 1.0
    bed = np.array([[-(j * dx * slope) for j in range(nc)] for _ in range(nr)])
    st = np.zeros((nr, nc, 3))
    st[..., 0] = h_n
    st[..., 1] = q
    u = q / h_n
    c = math.sqrt(G * h_n)
    t_end = 2.0 * (nc * dx) / (u + c)
    return Case("macdonald", nr, nc, dx, dx, bed, np.full((nr, nc), manning), st,
                (WALL, WALL, DISCHARGE, DEPTH), (0.0, 0.0, q, h_n), extra={"t_end": t_end, "h_n": h_n})


_BUILDINGS = [(150.0, 170.0, 20.0, 40.0), (150.0, 170.0, 60.0, 80.0), (200.0, 220.0, 40.0, 60.0),
              (250.0, 270.0, 20.0, 40.0), (250.0, 270.0, 60.0, 80.0), (320.0, 360.0, 35.0, 65.0)]


def uk_ea_t6():
    """UK EA test 6 stand-in: urban dam-break through buildings (gen_verification_data.rs:196-253)."""
    nx, ny = 250, 50
    dx, dy = 500.0 / nx, 100.0 / ny
    bed = np.zeros((ny, nx))
    st = np.zeros((ny, nx, 3))
    for i in range(ny):
        for j in range(nx):
            x, y = (j + 0.5) * dx, (i + 0.5) * dy
            inside = any(xa <= x <= xb and ya <= y <= yb for xa, xb, ya, yb in _BUILDINGS)
            if inside:
                bed[i, j] = 5.0
            elif x < 100.0:
                st[i, j, 0] = 2.0
    return Case("uk_ea_t6", ny, nx, dx, dy, bed, np.full((ny, nx), 0.03), st, (WALL, WALL, WALL, TRANS),
                extra={"t_end": 30.0})


def inverse_manning_channel():
    """3x60 dam-break channel of the inverse-Manning demo (m1_inverse_manning_demo.rs:42-108)."""
    nr, nc = 3, 60
    st = np.zeros((nr, nc, 3))
    st[:, : nc // 2, 0] = 1.0
    st[:, nc // 2:, 0] = 0.1
    return Case("inverse_manning", nr, nc, 1.0, 1.0, np.zeros((nr, nc)), np.zeros((nr, nc)), st,
                (WALL, WALL, TRANS, TRANS), extra={"t_end": 20.0})


def dam_break_256(n=256):
    """256x256 flat frictionless dam-break: dx 0.1, h 1.0 for x < 12.8 m, 0.1 elsewhere, walls."""
    dx = 0.1
    st = np.zeros((n, n, 3))
    x = (np.arange(n) + 0.5) * dx
    st[:, x < 12.8, 0] = 1.0
    st[:, x >= 12.8, 0] = 0.1
    return Case(f"dam_break_{n}", n, n, dx, dx, np.zeros((n, n)), np.zeros((n, n)), st, (WALL, WALL, WALL, WALL))


def bumpy_wet_dry(nr=37, nc=53, seed=7, dry_frac=0.35):
    """Adversarial mix for kernel parity: rough bed, islands, dry patches, momentum, varied Manning."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:nr, 0:nc]
    bed = 0.6 * np.sin(x / 5.0) * np.cos(y / 7.0) + 0.3 * rng.standard_normal((nr, nc))
    eta = 0.5 + 0.2 * np.exp(-((x - nc / 3) ** 2 + (y - nr / 2) ** 2) / 40.0)
    h = np.maximum(eta - bed, 0.0)
    h[rng.random((nr, nc)) < dry_frac * 0.3] = 0.0
    h[:, : nc // 5] = np.where(rng.random((nr, nc // 5)) < 0.5, 0.0, h[:, : nc // 5])
    st = np.zeros((nr, nc, 3))
    st[..., 0] = h
    st[..., 1] = h * 0.3 * rng.standard_normal((nr, nc))
    st[..., 2] = h * 0.3 * rng.standard_normal((nr, nc))
    # thin films with stale momentum (caller-supplied states, update.rs:1196-1205)
    film = rng.random((nr, nc)) < 0.03
    st[film, 0] = 5e-6
    st[film, 1] = 1e-6
    man = rng.choice([0.0, 0.03, 0.05, 0.08], size=(nr, nc))
    return Case(f"bumpy_{nr}x{nc}", nr, nc, 0.7, 1.3, bed, man, st, (WALL, TRANS, WALL, TRANS))


def mostly_dry(nr=44, nc=60, seed=5, west=DISCHARGE):
    """Dry-domain mix for the dry-tile skip: a dry rough bed with one pond,
    a Discharge (or Depth) inflow on the west edge, dry cells carrying stale momentum,
    depths exactly at and just above H_DRY, negative zeros, and films that
    a source or the pond reaches only after some steps."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:nr, 0:nc]
    bed = 0.02 * x + 0.3 * np.sin(y / 4.0) + 0.05 * rng.standard_normal((nr, nc))
    st = np.zeros((nr, nc, 3))
    pond = (x - 0.7 * nc) ** 2 + (y - 0.3 * nr) ** 2 < 20
    st[pond, 0] = 0.4
    st[pond, 1] = 0.05
    dry = ~pond
    stale = dry & (rng.random((nr, nc)) < 0.2)
    st[stale, 0] = 3e-7
    st[stale, 1] = 2e-7 * rng.standard_normal(np.count_nonzero(stale))
    st[stale, 2] = -1e-7
    st[nr // 2, nc // 2, 0] = 1e-6                                 # exactly H_DRY: dry
    st[(nr // 2 + 7) % nr, (nc // 2 - 9) % nc, 0] = 1.000001e-6   # just wet: keeps its tiles active
    st[3 % nr, 40 % nc, :] = -0.0
    man = rng.choice([0.0, 0.03, 0.05], size=(nr, nc))
    return Case(f"mostly_dry_{nr}x{nc}", nr, nc, 1.0, 1.0, bed, man, st, (WALL, WALL, west, TRANS),
                (0.0, 0.0, 0.05, 0.0))
