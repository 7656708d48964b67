"""GPU: device-side forcing series and gauges, and UK EA Test 4 on its
official geometry.

The series and gauges are checked bit for bit against the oracle's driver
(whose series equal the reference's loops written out on the host, see
test_ascii_grid_and_series.py). The full 18000 s Test 4 run on the GPU must
then reproduce the reference's own published comparison against
LISFLOOD-FP (benchmarks/data/uk_ea/test4/results_hydroflux.md): 53,610
steps, and RMSE, peak bias and arrival time at the six control points to
the printed digits, since the step is the reference's arithmetic.
"""
import os
import sys

import numpy as np
import pytest

import cases
from hydroflux_amd.stage import compare_point, parse_reference_stage
from oracle.pyoracle import OracleSim, set_pow_mode, set_threads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "uk_ea", "test4")
DEM = os.path.join(GOLDEN, "ea4-5m.dem.gz")

# results_hydroflux.md: (reference peak, RMSE, peak bias, arrival dt) per control point
PUBLISHED = {
    "ea4-5m-dg2.stage": [(0.3251, 0.0040, -0.0046, 0.0), (0.2745, 0.0009, -0.0007, 0.0),
                         (0.2275, 0.0007, 0.0001, 0.0), (0.2002, 0.0006, 0.0001, 60.2),
                         (0.1800, 0.0006, 0.0001, 59.8), (0.1755, 0.0006, 0.0002, 0.0)],
    "ea4-1m-acc.stage": [(0.3293, 0.0078, -0.0088, 0.0), (0.2771, 0.0031, -0.0033, 99.7),
                         (0.2288, 0.0015, -0.0012, 0.0), (0.2010, 0.0011, -0.0007, 0.0),
                         (0.1805, 0.0010, -0.0004, 0.0), (0.1761, 0.0011, -0.0004, 0.0)],
}
PUBLISHED_STEPS = 53_610


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _mode():
    set_threads(16)
    set_pow_mode(1)
    yield
    set_pow_mode(0)


@pytest.mark.parametrize("kind", ["f64", "dual"])
def test_series_and_gauges_bitwise_with_oracle(kind):
    _say(f"series + gauges {kind}")
    from hydroflux_amd import Grid2D

    c = cases.mostly_dry(70, 90, 4, west=cases.WALL)
    c.bc_kinds = (cases.TRANS, cases.TRANS, cases.WALL, cases.TRANS)
    kw = dict(manning_tangent=c.manning) if kind == "dual" else {}
    inlet = [(34, 0), (35, 0), (36, 0)]
    sims = [c.load(Grid2D(c.nr, c.nc, c.dx, c.dy, kind=kind), **kw),
            c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=kind == "dual"), **kw)]
    for s in sims:
        s.configure(cfl=0.4, max_dt=0.5, sources=[(r, cc, 0.0) for r, cc in inlet] + [(60, 70, 0.02)])
        s.set_forcing_series("sources", (0.0, 0.5, 2.0, 4.0, 6.0), (0.0, 0.0, 1.0, 1.0, 0.0), 1.0, 0.3, 3.0)
        s.set_forcing_series("rain", (0.0, 0.009, 0.01, 0.05, 0.0501, 1.0), (0.0, 0.0, 400.0, 400.0, 0.0, 0.0),
                             60.0, 1e-3, 3600.0)
        s.set_gauges([(35, 1), (35, 10), (60, 70)], 10_000)
    sims[0].run_until(7.0)
    assert sims[1].run_until(7.0) == 0
    assert sims[0].time() == sims[1].time()
    gt, gh = sims[0].read_gauges()
    ot, oh = sims[1].read_gauges()
    assert np.array_equal(gt, ot) and np.array_equal(gh, oh)
    a, r = sims[0].get_state(), sims[1].get_state()
    if kind == "dual":
        assert np.array_equal(a[0], r[0]) and np.array_equal(a[1], r[1], equal_nan=True)
    else:
        assert np.array_equal(a, r)


@pytest.mark.parametrize("kind", ["f64", "f32", "dual", "dual4"])
def test_gauge_rows_equal_the_state_after_each_step(kind):
    """Every scalar kind: each recorded row is (clock, depth value at the
    gauge cells) of the state after that step, one row per step, stepped one
    at a time and through the captured multi-step graphs; rows past the
    capacity are dropped."""
    _say(f"gauge rows vs state {kind}")
    from hydroflux_amd import Grid2D

    c = cases.mostly_dry(40, 56, 7, west=cases.WALL)
    cells = [(20, 1), (20, 9), (33, 40), (0, 0)]
    g = c.load(Grid2D(c.nr, c.nc, c.dx, c.dy, kind=kind))
    g.configure(cfl=0.4, max_dt=0.5, sources=[(20, 0, 0.0), (33, 40, 0.01)])
    g.set_forcing_series("sources", (0.0, 0.5, 2.0, 40.0), (0.0, 0.0, 1.0, 1.0), 1.0, 0.3, 1.0)
    g.set_forcing_series("rain", (0.0, 1.0, 2.0), (0.0, 200.0, 200.0), 1.0, 1e-3, 3600.0)
    g.set_gauges(cells, 64)

    def depth_now():
        s = g.get_state()
        h = (s[0] if isinstance(s, tuple) else s)[..., 0]
        return np.array([h[r, cc] for r, cc in cells], np.float64)

    expect_t, expect_h = [g.time()[0]], [depth_now()]
    for _ in range(5):
        g.step()
        expect_t.append(g.time()[0])
        expect_h.append(depth_now())
    g.run_steps(40)  # graph replays
    t, steps = g.time()
    assert steps == 45
    gt, gh = g.read_gauges()
    assert gt.shape == (46,) and gh.shape == (46, len(cells))
    assert np.array_equal(gt[:6], expect_t) and np.array_equal(gh[:6], np.array(expect_h))
    assert gt[-1] == t and np.array_equal(gh[-1], depth_now())
    assert np.all(np.diff(gt) > 0) and np.any(gh[-1] > 0.0)
    g.run_steps(30)
    gt, gh = g.read_gauges()
    assert gt.shape == (64,) and g.time()[1] == 75
    g.close()


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_gauge_rows_with_more_gauges_than_threads(kind):
    """700 gauges: the row is written by more columns than the recording
    workgroup has threads (512), so some threads write several."""
    _say(f"700 gauges {kind}")
    from hydroflux_amd import Grid2D

    c = cases.mostly_dry(40, 56, 9, west=cases.WALL)
    rng = np.random.default_rng(3)
    cells = [(int(r), int(cc)) for r, cc in zip(rng.integers(0, c.nr, 700), rng.integers(0, c.nc, 700))]
    g = c.load(Grid2D(c.nr, c.nc, c.dx, c.dy, kind=kind))
    g.configure(cfl=0.4, max_dt=0.5, rain_rate=2e-5, sources=[(20, 0, 0.2)])
    g.set_gauges(cells, 100)
    g.run_steps(3)   # eager
    h3 = g.get_state()[..., 0]
    g.run_steps(32)  # graph replays
    h = g.get_state()[..., 0]
    gt, gh = g.read_gauges()
    assert gh.shape == (36, 700) and gt[-1] == g.time()[0]
    assert np.array_equal(gh[3], [h3[r, cc] for r, cc in cells])
    assert np.array_equal(gh[-1], [h[r, cc] for r, cc in cells])
    g.close()


def _oracle_test4(t_end):
    from hydroflux_amd import uk_ea
    from hydroflux_amd.ascii_grid import read_ascii_grid

    bed, hdr = read_ascii_grid(DEM)
    o = OracleSim(hdr.nrows, hdr.ncols, hdr.cellsize, hdr.cellsize)
    o.set_mesh(bed, np.full(bed.shape, uk_ea.MANNING_N))
    o.set_state(np.zeros((hdr.nrows, hdr.ncols, 3)))
    o.set_boundaries([uk_ea.TRANSMISSIVE, uk_ea.TRANSMISSIVE, uk_ea.WALL, uk_ea.TRANSMISSIVE])
    rows = hdr.rows_overlapping_y_range(*uk_ea.INLET_Y)
    o.configure(cfl=uk_ea.CFL, max_dt=uk_ea.MAX_DT, sources=[(r, 0, 0.0) for r in rows])
    t, f = zip(*uk_ea.HYDROGRAPH)
    o.set_forcing_series("sources", t, f, 1.0, uk_ea.PEAK_Q_M3S, float(len(rows)))
    o.set_gauges([hdr.cell_at(x, y) for x, y in uk_ea.CONTROL_POINTS])
    assert o.run_until(t_end) == 0
    return o


def test_uk_ea_test4_first_1200s_bitwise_with_oracle():
    _say("UK EA test 4, 0-1200 s, gpu vs oracle")
    from hydroflux_amd import uk_ea

    g, _ = uk_ea.ea_test4_grid(DEM)
    g.run_until(1200.0)
    o = _oracle_test4(1200.0)
    assert g.time() == o.time()
    gt, gh = g.read_gauges()
    ot, oh = o.read_gauges()
    assert np.array_equal(gt, ot) and np.array_equal(gh, oh)
    assert np.array_equal(g.get_state(), o.get_state())
    assert gh[-1, 0] > 0.05  # the inflow has reached control point 1


def test_uk_ea_test4_full_run_reproduces_published_comparison():
    _say("UK EA test 4, full 18000 s")
    from hydroflux_amd import uk_ea

    g, _ = uk_ea.ea_test4_grid(DEM)
    g.run_until(uk_ea.SIM_TIME)
    t, steps = g.time()
    assert t == uk_ea.SIM_TIME and steps == PUBLISHED_STEPS
    times, depths = g.read_gauges()
    assert len(times) == steps + 1
    for name, rows in PUBLISHED.items():
        ref = parse_reference_stage(os.path.join(GOLDEN, name))
        for p, (peak, rmse, bias, arrival) in enumerate(rows):
            cmp = compare_point(ref, p, list(times), list(depths[:, p]), 0.02)
            assert abs(cmp.reference_peak - peak) <= 0.5e-4 + 1e-12, (name, p)
            assert abs(cmp.rmse - rmse) <= 0.5e-4 + 1e-12, (name, p, cmp.rmse)
            assert abs(cmp.peak_bias - bias) <= 0.5e-4 + 1e-12, (name, p, cmp.peak_bias)
            assert cmp.sim_arrival is not None and cmp.reference_arrival is not None
            assert abs((cmp.sim_arrival - cmp.reference_arrival) - arrival) <= 0.05 + 1e-9, (name, p)


# The reference's run of Test 8A with its current step (after the
# source-term change of docs/bug-report-2026-07-boundary-slope-instability.md
# section 10), papers/01_review/ROADMAP_REVISION_EMS.md:222-229: 141,573
# steps, depths at control points 1, 2, 4, 7 (peak) and 3 (peak / final) to
# 4 decimals. (The earlier table in test8a_glasgow/results_hydroflux.md,
# 142,074 steps, predates that change.)
T8A = os.path.join(os.path.dirname(__file__), "golden", "uk_ea", "test8a")
POSTFIX_8A_STEPS = 141_573
POSTFIX_8A_PEAKS = {0: 0.5534, 1: 0.2477, 2: 0.7340, 3: 0.1629, 6: 0.2427}
POSTFIX_8A_FINAL_PT3 = 0.7339


def _oracle_test8a(t_end, pow_mode):
    from hydroflux_amd import uk_ea
    from hydroflux_amd.ascii_grid import read_ascii_grid

    bed, hdr = read_ascii_grid(os.path.join(T8A, "ea8-2m.dem.gz"))
    man, _ = read_ascii_grid(os.path.join(T8A, "ea8-2m.n.gz"))
    set_pow_mode(pow_mode)
    o = OracleSim(hdr.nrows, hdr.ncols, hdr.cellsize, hdr.cellsize)
    o.set_mesh(bed, man)
    o.set_state(np.zeros((hdr.nrows, hdr.ncols, 3)))
    o.set_boundaries([uk_ea.WALL] * 4)
    cell = hdr.cell_at(*uk_ea.T8_POINT_XY)
    o.configure(cfl=uk_ea.CFL, max_dt=uk_ea.T8_MAX_DT, sources=[(cell[0], cell[1], 0.0)])
    t, v = zip(*uk_ea.T8_RAIN_MIN)
    o.set_forcing_series("rain", t, v, 60.0, 1e-3, 3600.0)
    t, v = zip(*uk_ea.T8_POINT_MIN)
    o.set_forcing_series("sources", t, v, 60.0, hdr.cellsize, 1.0)
    o.set_gauges([hdr.cell_at(x, y) for x, y in uk_ea.T8_CONTROL_POINTS])
    assert o.run_until(t_end) == 0
    return o


def test_uk_ea_test8a_first_300s_bitwise_with_oracle():
    """Urban topography, spatially varying Manning, the rain pulse and the
    start of the inflow, bit for bit. (Against glibc's pow instead of the
    kernels' correctly rounded one, single cells already differ by
    centimetres after 300 s here: one-ulp friction differences flip cells
    across the H_VEL / H_DRY switches in the thin roof and street films, so
    only identical arithmetic gives a meaningful comparison.)"""
    _say("UK EA test 8A, 0-300 s, gpu vs oracle")
    from hydroflux_amd import uk_ea

    g, _ = uk_ea.ea_test8a_grid(os.path.join(T8A, "ea8-2m.dem.gz"), os.path.join(T8A, "ea8-2m.n.gz"))
    g.run_until(300.0)
    o = _oracle_test8a(300.0, 1)
    assert g.time() == o.time()
    assert np.array_equal(g.get_state(), o.get_state())
    assert np.array_equal(g.read_gauges()[1], o.read_gauges()[1])


def test_uk_ea_test8a_full_run_reproduces_reference_run():
    _say("UK EA test 8A, full 18000 s")
    from hydroflux_amd import uk_ea

    g, _ = uk_ea.ea_test8a_grid(os.path.join(T8A, "ea8-2m.dem.gz"), os.path.join(T8A, "ea8-2m.n.gz"))
    g.run_until(uk_ea.SIM_TIME)
    t, steps = g.time()
    times, depths = g.read_gauges()
    peaks = np.maximum(depths.max(axis=0), 0.0)
    print(f"[test8a] steps {steps}, peaks {peaks.tolist()}, final pt3 {depths[-1, 2]!r}", flush=True)
    assert t == uk_ea.SIM_TIME and steps == POSTFIX_8A_STEPS
    assert len(times) == steps + 1
    # one unit of the printed 4th decimal
    for p, ref in POSTFIX_8A_PEAKS.items():
        assert abs(peaks[p] - ref) <= 1e-4, (p, peaks[p], ref)
    assert abs(depths[-1, 2] - POSTFIX_8A_FINAL_PT3) <= 1e-4
