"""Closed forms and invariants from the reference's unit tests, on the oracle.

These pin the parts of the restatement the golden CSVs leave open (single
faces, single cells, exact conservation), each chosen so a plausible slip
(dropped term, wrong sign or index, swapped operand) fails it.
"""
import math

import numpy as np
import pytest

import cases
from oracle.pyoracle import OracleSim, set_pow_mode, set_threads

G = 9.81


@pytest.fixture(autouse=True)
def _mode():
    set_threads(4)
    set_pow_mode(0)


def _sim(nr, nc, bed=0.0, man=0.0, dx=1.0, dy=1.0, dual=False, **tangents):
    s = OracleSim(nr, nc, dx, dy, dual=dual)
    s.set_mesh(np.broadcast_to(np.asarray(bed, float), (nr, nc)), man, **tangents)
    return s


def test_max_wave_speeds_textbook():
    # update.rs:1373-1381
    s = _sim(2, 2)
    st = np.zeros((2, 2, 3))
    st[0, 0] = (1.0, 2.0, 3.0)
    s.set_state(st)
    sx, sy = s.max_wave_speeds()
    c = math.sqrt(G)
    assert sx == pytest.approx(2.0 + c, abs=1e-12) and sy == pytest.approx(3.0 + c, abs=1e-12)


def test_cfl_dry_domain_is_infinite():
    s = _sim(4, 5)
    s.set_state(np.zeros((4, 5, 3)))
    assert s.cfl_time_step(0.5) == math.inf


def test_lake_at_rest_flat_exact():
    # update.rs:1393-1408
    s = _sim(10, 12)
    st = np.zeros((10, 12, 3))
    st[..., 0] = 2.0
    s.set_state(st)
    s.set_boundaries([0, 0, 0, 0])
    for _ in range(100):
        s.forward_euler_step(0.01)
    out = s.get_state()
    assert np.all(out[..., 0] == 2.0) and np.all(out[..., 1:] == 0.0)


@pytest.mark.parametrize("bed_kind", ["x", "y", "diag", "gauss"])
def test_lake_at_rest_sloped_beds(bed_kind):
    # update.rs:1410-1500 (1e-10)
    nr, nc = 12, 20
    y, x = np.mgrid[0:nr, 0:nc].astype(float)
    bed = {"x": -0.05 * x, "y": -0.05 * y, "diag": -0.03 * x - 0.04 * y,
           "gauss": 1.5 * np.exp(-((x - 9.5) ** 2 + (y - 5.5) ** 2) / 5.76)}[bed_kind]
    eta = 5.0
    s = _sim(nr, nc, bed)
    st = np.zeros((nr, nc, 3))
    st[..., 0] = eta - bed
    s.set_state(st)
    s.set_boundaries([1, 1, 1, 1])
    for _ in range(50):
        s.ssprk2_step(s.cfl_time_step(0.4))
    out = s.get_state()
    assert np.max(np.abs(out[..., 0] + bed - eta)) < 1e-10
    assert np.max(np.abs(out[..., 1:])) < 1e-10


def test_emerged_island_lake_at_rest():
    # update.rs:1522-1605: emerged bump, eta below its crest, no spurious momentum
    n = 21
    y, x = np.mgrid[0:n, 0:n].astype(float)
    bed = 2.0 * np.exp(-((x - 10) ** 2 + (y - 10) ** 2) / 10.0)
    eta = 1.0
    st = np.zeros((n, n, 3))
    st[..., 0] = np.maximum(eta - bed, 0.0)
    s = _sim(n, n, bed)
    s.set_state(st)
    s.set_boundaries([1, 1, 1, 1])
    for _ in range(100):
        s.ssprk2_step(s.cfl_time_step(0.4))
    out = s.get_state()
    wet = st[..., 0] > 0
    assert np.max(np.abs(out[wet, 0] + bed[wet] - eta)) < 1e-10
    assert np.max(np.abs(out[..., 1:])) < 1e-10


def test_draining_cell_regression_and_mass():
    # update.rs:1607-1644
    s = _sim(3, 3)
    st = np.zeros((3, 3, 3))
    h0 = 0.05
    st[0, 1] = (h0, 0.0, h0 * 5.0)
    s.set_state(st)
    m0 = st[..., 0].sum()
    s.forward_euler_step(0.25)
    out = s.get_state()
    assert out[0, 1, 0] <= 1e-6
    assert out[1, 1, 0] > 0.5 * h0
    assert abs(out[..., 0].sum() - m0) < 1e-12


H_DRY = 1.0e-6  # update.rs: the depth below which a cell is dry


def _ulps(x):
    return float(np.spacing(np.float64(x)))


@pytest.mark.parametrize("dt", [0.2, 0.5, 1.0])
def test_alpha_rescaling_closed_form_on_a_draining_chain(dt):
    """Liang-Marche outflow limiter (update.rs:953-1053) where it binds.

    One row between walls: A (h 0.05, u 5) upstream of B (h 0.02, u 5), then
    dry cells, with a dt at which both would empty. alpha = (h - H_DRY) / out
    counts outflows only and scales each face by its upstream cell's alpha,
    so exactly the available depth of each wet cell moves one cell
    downstream: A keeps H_DRY, B receives all of A's (ending at h_A), the
    first dry cell receives all of B's. Rounding aside, nothing else is
    free: using the receiving cell's alpha, counting inflow in `out`, or
    dropping H_DRY from the available depth each moves these values by far
    more than the few ulps allowed."""
    s = _sim(1, 5)
    st = np.zeros((1, 5, 3))
    st[0, 0] = (0.05, 0.25, 0.0)
    st[0, 1] = (0.02, 0.10, 0.0)
    s.set_state(st)
    s.set_boundaries([1, 1, 1, 1])
    s.forward_euler_step(dt)
    h = s.get_state()[0, :, 0]
    tol = 4 * _ulps(0.05)
    assert abs(h[0] - H_DRY) <= tol, h
    assert abs(h[1] - 0.05) <= tol, h
    assert abs(h[2] - (0.02 - H_DRY)) <= tol, h
    assert h[3] == 0.0 and h[4] == 0.0
    assert abs(h.sum() - 0.07) <= tol


@pytest.mark.parametrize("angle", [0.0, 0.7, 1.9, 3.3, 5.1])
def test_alpha_rescaling_isolated_cell_empties_to_h_dry(angle):
    """An isolated wet cell among dry ones has only outflows: where the
    limiter binds (update.rs:953-982) its new depth is h - alpha*out =
    H_DRY to rounding, and the water it loses is what its neighbours gain."""
    h0, speed = 0.05, 6.0
    s = _sim(3, 3)
    st = np.zeros((3, 3, 3))
    st[1, 1] = (h0, h0 * speed * math.cos(angle), h0 * speed * math.sin(angle))
    s.set_state(st)
    s.set_boundaries([1, 1, 1, 1])
    s.forward_euler_step(0.5)
    out = s.get_state()[..., 0]
    assert abs(out[1, 1] - H_DRY) <= 4 * _ulps(h0), out[1, 1]
    assert abs(out.sum() - h0) <= 4 * _ulps(h0)
    assert np.all(out >= 0.0)


def test_wetting_front_conserves_mass_closed_box():
    # update.rs:1689-1717
    n = 30
    st = np.zeros((n, n, 3))
    st[:, : n // 2, 0] = 1.0
    s = _sim(n, n)
    s.set_state(st)
    m0 = st[..., 0].sum()
    for _ in range(300):
        s.forward_euler_step(s.cfl_time_step(0.4))
    out = s.get_state()
    assert np.all(np.isfinite(out[..., 0])) and np.all(out[..., 0] >= 0)
    assert abs(out[..., 0].sum() - m0) <= 1e-12 * m0 * 10


def test_radial_symmetry_x_vs_y():
    # radial_dam_break.rs:224-231: transposing the problem transposes the answer
    n = 25
    y, x = np.mgrid[0:n, 0:n].astype(float)
    h = 1.0 + 1.0 * (((x - 12) ** 2 + (y - 12) ** 2) < 16)
    st = np.stack([h, np.zeros_like(h), np.zeros_like(h)], -1)
    a = _sim(n, n)
    a.set_state(st)
    for _ in range(20):
        a.ssprk2_step(0.02)
    out = a.get_state()
    assert np.max(np.abs(out[..., 0] - out[..., 0].T)) < 1e-10
    assert np.max(np.abs(out[..., 1] - out[..., 2].T)) < 1e-10


def test_friction_one_step_analytic():
    # source.rs:260-279: hu /= 1 + dt g n^2 |U| / h^(4/3)
    s = _sim(1, 1, man=0.03)
    h, hu, hv, dt = 0.5, 0.4, -0.3, 0.7
    s.set_state(np.array([[[h, hu, hv]]]))
    s.manning_friction_step(dt, 1e-6)
    speed = math.hypot(hu, hv) / h
    f = 1.0 + dt * G * 0.03 ** 2 * speed / h ** (4.0 / 3.0)
    out = s.get_state()[0, 0]
    assert out[1] == pytest.approx(hu / f, rel=1e-14) and out[2] == pytest.approx(hv / f, rel=1e-14)
    # direction preserved exactly up to rounding
    assert out[2] / out[1] == pytest.approx(hv / hu, rel=1e-15)


def test_friction_skips_dry_and_zero_n():
    s = _sim(1, 2, man=np.array([[0.0, 0.05]]))
    s.set_state(np.array([[[1.0, 1.0, 1.0], [1e-7, 1e-8, 0.0]]]))
    s.manning_friction_step(1.0, 1e-6)
    out = s.get_state()
    assert np.array_equal(out[0, 0], [1.0, 1.0, 1.0]) and np.array_equal(out[0, 1], [1e-7, 1e-8, 0.0])


def test_rain_and_point_source_increments():
    # source.rs:479-631
    s = _sim(2, 3, dx=2.0, dy=0.5)
    st = np.zeros((2, 3, 3))
    st[0, 0] = (0.2, 0.1, 0.0)
    st[1, 2] = (1e-3, 0.0, 0.0)
    s.set_state(st)
    s.apply_rain(1e-3, 2.0)
    out = s.get_state()
    assert np.allclose(out[..., 0], st[..., 0] + 2e-3, rtol=0, atol=0)
    s.apply_rain(-5e-3, 0.5)  # evaporation dries cells that would go negative
    out2 = s.get_state()
    assert out2[0, 1, 0] == 0.0 and np.all(out2[0, 1, 1:] == 0.0)  # 2e-3 - 2.5e-3 < 0 -> DRY
    assert out2[1, 2, 0] == pytest.approx(3e-3 - 2.5e-3) and out2[0, 0, 0] == pytest.approx(0.2 + 2e-3 - 2.5e-3)
    s.apply_point_sources([(0, 0, 0.3), (0, 0, -10.0), (1, 1, 0.5)], 0.5)
    out3 = s.get_state()
    assert out3[0, 0, 0] == 0.0 and np.all(out3[0, 0, 1:] == 0.0)  # sink clamps and zeroes momentum
    assert out3[1, 1, 0] == pytest.approx(out2[1, 1, 0] + 0.5 * 0.5 / 1.0)


def test_forward_euler_dual_mass_tangent():
    # update.rs:1953-1992: d(mass)/d(h_init) = n_cells exactly
    s = _sim(4, 5, dual=True)
    st = np.zeros((4, 5, 3))
    st[..., 0] = 2.0
    tan = np.zeros_like(st)
    tan[..., 0] = 1.0
    s.set_state(st, tan)
    s.forward_euler_step(0.01)
    v, t = s.get_state()
    assert v[..., 0].sum() == pytest.approx(40.0, abs=1e-10)
    assert t[..., 0].sum() == pytest.approx(20.0, abs=1e-10)


def test_ssprk2_dual_bed_seed_analytic():
    # update.rs:1994-2038: d(mass)/d(z0) = -36
    z = 0.5
    s = _sim(6, 6, bed=np.full((6, 6), z), dual=True, bed_tangent=np.ones((6, 6)))
    st = np.zeros((6, 6, 3))
    st[..., 0] = 3.0 - z
    tan = np.zeros_like(st)
    tan[..., 0] = -1.0
    s.set_state(st, tan)
    s.ssprk2_step(0.005)
    _, t = s.get_state()
    assert t[..., 0].sum() == pytest.approx(-36.0, abs=1e-8)


def test_dual_friction_tangent_matches_central_difference():
    # source.rs:420-473: seed on n
    def run(n, dual):
        s = _sim(1, 1, man=n, dual=dual, manning_tangent=(1.0 if dual else None))
        s.set_state(np.array([[[0.4, 0.3, 0.1]]]))
        s.manning_friction_step(0.5, 1e-6)
        return s.get_state()
    eps = 1e-6
    fd = (run(0.05 + eps, False)[0, 0, 1] - run(0.05 - eps, False)[0, 0, 1]) / (2 * eps)
    v, t = run(0.05, True)
    assert t[0, 0, 1] == pytest.approx(fd, rel=1e-7)


def test_ghost_discharge_on_dry_manning_normal_depth():
    # boundary.rs:193-219: a dry inner cell on a sloped bed gets h_n = (n|q|/sqrt(S))^(3/5)
    nr, nc, slope, n, q = 3, 6, 0.01, 0.03, 0.5
    bed = np.tile(-slope * np.arange(nc, dtype=float), (nr, 1))
    s = _sim(nr, nc, bed, man=n)
    s.set_state(np.zeros((nr, nc, 3)))
    s.set_boundaries([1, 1, 2, 0], [0, 0, q, 0])
    dt = s.cfl_time_step(0.4, with_bcs=True)
    h_n = (n * (q / math.sqrt(slope))) ** 0.6
    c = math.sqrt(G * h_n)
    assert math.isfinite(dt)
    assert dt == pytest.approx(0.4 / ((q / h_n + c) / 1.0 + c / 1.0), rel=1e-12)
    s.ssprk2_step(dt)
    assert s.get_state()[:, 0, 0].min() > 0.0  # the inflow entered
