"""The stage-kernel tile code, run on the CPU, against the oracle (bit for bit).

tests/emu builds hydroflux_amd/csrc/hf2d_tile.h with g++ and runs every
barrier-separated phase thread by thread with LDS bounds checks. This checks
the kernel logic (halo sizes, tile-edge and domain-edge handling, the device
driver's clock and CFL slots) without a GPU; the GPU tests then check the
same code on gfx950.
"""
import numpy as np
import pytest

import cases
from emu.pyemu import EmuSim
from oracle.pyoracle import OracleSim, set_pow_mode, set_threads


@pytest.fixture(autouse=True)
def _mode():
    set_threads(8)
    set_pow_mode(1)  # the kernels' pow
    yield
    set_pow_mode(0)


TILES = [(64, 24), (32, 24), (32, 16), (24, 16), (24, 8), (8, 4), (4, 8)]
CASES = [
    (cases.stoker(100), None),
    (cases.thacker(40), None),
    (cases.macdonald(), 1e-9),
    (cases.bumpy_wet_dry(), 1e-6),
    (cases.bumpy_wet_dry(5, 3, 1), 1e-6),
    (cases.bumpy_wet_dry(9, 1, 2), 1e-6),
]


@pytest.mark.parametrize("tile", TILES, ids=lambda t: f"{t[0]}x{t[1]}")
@pytest.mark.parametrize("case,fr", CASES, ids=lambda c: getattr(c, "name", "f"))
def test_hand_loop_bitwise(tile, case, fr):
    o = case.load(OracleSim(case.nr, case.nc, case.dx, case.dy))
    e = case.load(EmuSim(case.nr, case.nc, case.dx, case.dy, tile=tile))
    for _ in range(15):
        dto, dte = o.cfl_time_step(0.4), e.cfl_time_step(0.4)
        assert dto == dte
        o.ssprk2_step(dto)
        e.ssprk2_step(dte)
        if fr is not None:
            o.manning_friction_step(dto, fr)
            e.manning_friction_step(dte, fr)
    assert np.array_equal(o.get_state(), e.get_state())


@pytest.mark.parametrize("tile", TILES, ids=lambda t: f"{t[0]}x{t[1]}")
@pytest.mark.parametrize("integrator", [1, 0], ids=["ssprk2", "euler"])
def test_fused_driver_bitwise(tile, integrator):
    c = cases.bumpy_wet_dry(23, 41, 4)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy))
    e = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, tile=tile))
    src = [(c.nr // 2, c.nc // 3, 0.7), (1, 1, -0.2), (c.nr // 2, c.nc // 3, 0.1)]
    for s in (o, e):
        s.configure(cfl=0.45, integrator=integrator, max_dt=0.05, friction_dry_tol=1e-6, rain_rate=1e-3,
                    sources=src)
    assert o.run_until(0.6) == e.run_until(0.6) == 0
    assert o.time() == e.time()
    assert np.array_equal(o.get_state(), e.get_state())


@pytest.mark.parametrize("tile", [(16, 16), (8, 4), (24, 16), (16, 8)], ids=lambda t: f"{t[0]}x{t[1]}")
def test_dual_driver_bitwise(tile):
    c = cases.bumpy_wet_dry(19, 27, 9)
    c.manning = np.maximum(c.manning, 0.03)
    rng = np.random.default_rng(3)
    mt, bt = c.manning.copy(), 0.01 * rng.standard_normal((c.nr, c.nc))
    stt = 0.1 * rng.standard_normal((c.nr, c.nc, 3))
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=True), manning_tangent=mt, bed_tangent=bt,
               state_tangent=stt)
    e = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, dual=True, tile=tile), manning_tangent=mt, bed_tangent=bt,
               state_tangent=stt)
    for s in (o, e):
        s.configure(cfl=0.4, max_dt=0.5, rain_rate=2e-4, sources=[(1, 2, 0.3)])
        s.run_steps(20)
    (vo, to), (ve, te) = o.get_state(), e.get_state()
    assert np.array_equal(vo, ve) and np.array_equal(to, te)


def test_driver_errors_match_oracle():
    # dry domain with an unbounded max_dt: DegenerateDt (sim.rs:209-215)
    c = cases.stoker(20)
    for cls in (OracleSim, EmuSim):
        s = cls(c.nr, c.nc, c.dx, c.dy)
        s.set_mesh(c.bed, 0.0)
        s.set_state(np.zeros_like(c.state))
        s.configure(max_dt=float("inf"))
        assert s.run_steps(1) == 3
    # step budget (sim.rs:234-240)
    for cls in (OracleSim, EmuSim):
        s = c.load(cls(c.nr, c.nc, c.dx, c.dy))
        s.configure(max_steps=3)
        assert s.run_until(100.0) == 4
        assert s.time()[1] == 3


@pytest.mark.parametrize("tile", [(8, 4), (4, 8), (16, 8), (32, 16), (32, 24), (24, 8)], ids=lambda t: f"{t[0]}x{t[1]}")
@pytest.mark.parametrize("west", [cases.DISCHARGE, cases.DEPTH], ids=["discharge", "depth"])
def test_active_tile_list_bitwise(tile, west):
    """The compacted active-tile list leaves settled dry tiles out; results
    stay the oracle's bits and equal a run over every tile."""
    c = cases.mostly_dry(60, 84, 5, west=west)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy))
    on = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, tile=tile))
    off = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, tile=tile))
    off.set_active_tiles(False)
    src = [(c.nr - 6, 8, 0.02), (30, 50, 0.01)]
    for s in (o, on, off):
        s.configure(cfl=0.4, max_dt=0.2, sources=src)
    before = on.tiles_skipped()
    assert on.run_steps(30) == 0
    skipped = on.tiles_skipped() - before
    assert off.run_steps(30) == 0
    assert on.tiles_skipped() - before == skipped  # the run over every tile skipped nothing
    # the state set from the host mid-run restarts the list from its wet flags
    st = on.get_state()
    st[40, 20] = (0.3, 0.0, 0.0)
    on.set_state(st)
    off.set_state(st)
    o.run_steps(30)
    o.set_state(st)
    for s in (o, on, off):
        s.run_steps(25)
    if tile[0] * tile[1] < 512:  # on this 60 x 84 grid, larger tiles all touch the edge or the water
        assert skipped > 0
    assert o.time() == on.time() == off.time()
    assert np.array_equal(on.get_state(), off.get_state())
    assert np.array_equal(o.get_state(), on.get_state())


def test_active_tile_list_not_used_with_rain():
    c = cases.mostly_dry(40, 48, 2)
    e = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, tile=(8, 4)))
    e.configure(cfl=0.4, max_dt=0.2, rain_rate=1e-7)
    before = e.tiles_skipped()
    e.run_steps(10)
    assert e.tiles_skipped() == before


def test_active_tile_list_dual_runs_every_tile():
    """Dual: a tangent at h == 0 halves at each max(h, 0) tie, so no tile is
    a fixed point and the list is not used; the results are the oracle's."""
    c = cases.mostly_dry(36, 40, 9)
    rng = np.random.default_rng(4)
    stt = 0.1 * rng.standard_normal((c.nr, c.nc, 3))
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=True), manning_tangent=c.manning, state_tangent=stt)
    e = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, dual=True, tile=(8, 4)), manning_tangent=c.manning,
               state_tangent=stt)
    for s in (o, e):
        s.configure(cfl=0.4, max_dt=0.2, sources=[(30, 5, 0.01)])
    before = e.tiles_skipped()
    o.run_steps(25)
    e.run_steps(25)
    assert e.tiles_skipped() == before
    ov, ot = o.get_state()
    ev, et = e.get_state()
    # the Discharge-on-dry ghost gives NaN tangents along the inflow edge in both
    assert np.array_equal(ov, ev) and np.array_equal(ot, et, equal_nan=True)


@pytest.mark.parametrize("tile", [(8, 4), (16, 8)], ids=lambda t: f"{t[0]}x{t[1]}")
def test_active_tile_list_run_until_split_across_calls(tile):
    """run_until split over several calls: the no-op steps that end each
    call's last chunk of device steps (the clock has reached its target) do
    not count as quiet steps of the active-tile list, so its counts equal
    those of a driver that starts no step that will not run. The first call
    ends after one step, right after the counts were reset. Bounds-checked
    tile code; results are the oracle's bits."""
    c = cases.mostly_dry(60, 84, 5, west=cases.WALL)
    src = [(c.nr - 6, 8, 0.02)]
    chunked = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, tile=tile))
    exact = c.load(EmuSim(c.nr, c.nc, c.dx, c.dy, tile=tile))
    exact.set_chunked(False)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy))
    for s in (chunked, exact, o):
        s.configure(cfl=0.4, max_dt=0.2, sources=src)
    for t in (0.01, 0.9, 1.7, 2.2, 3.0, 3.05, 4.0):
        for s in (chunked, exact, o):
            assert s.run_until(t) == 0
        assert np.array_equal(chunked.tile_runs(), exact.tile_runs())
    assert chunked.tile_runs().max() == 2  # the list was in use, with settled tiles
    assert chunked.time() == exact.time() == o.time()
    assert np.array_equal(chunked.get_state(), o.get_state())
    assert np.array_equal(exact.get_state(), o.get_state())


def test_fp32_stochastic_rounding_exact_in_expectation():
    """The FP32 form's rounding of its depth additions (hf2d_tile.h,
    add_rounded_stochastically), host build: every result is one of the two
    FP32 neighbours of the exact sum, exact sums stay exact, and over many
    keys the mean error vanishes where round-to-nearest has a fixed bias
    (rain of 9.5e-7 m on cells 2-4 m deep: 3.99 ulps, rounded up to 4 on
    every cell)."""
    from emu.pyemu import add_rounded_stochastically as add_sr

    rng = np.random.default_rng(11)
    n = 200_000
    h = rng.uniform(2.0, 4.0, n).astype(np.float32)
    dh = np.full(n, 9.5e-7, np.float32)
    keys = np.arange(n, dtype=np.uint32) * np.uint32(2654435761)
    out = add_sr(h, dh, keys)
    exact = h.astype(np.float64) + dh.astype(np.float64)
    rn = h + dh  # float32 round to nearest
    other = np.nextafter(rn, np.where(exact > rn, np.float32(np.inf), np.float32(-np.inf)))
    assert np.all((out == rn) | ((out == other) & (exact != rn.astype(np.float64))))
    err_rn = float(np.mean(rn.astype(np.float64) - exact))
    err_sr = float(np.mean(out.astype(np.float64) - exact))
    assert abs(err_rn) > 1e-9  # the bias being removed: about 3.7e-9 m per addition
    assert abs(err_sr) < 3e-10, (err_sr, err_rn)
    # sums that are exact in FP32 stay exact; subtractions (depth updates) too
    ex_h = rng.uniform(0.25, 0.5, 1000).astype(np.float32)
    ex_dh = np.float32(2.0 ** -20) * rng.integers(-64, 64, 1000).astype(np.float32)
    assert np.array_equal(add_sr(ex_h, ex_dh, np.arange(1000, dtype=np.uint32)), ex_h + ex_dh)
    neg = rng.uniform(0.5, 1.0, n).astype(np.float32)
    dneg = -rng.uniform(1e-8, 1e-6, n).astype(np.float32)
    out = add_sr(neg, dneg, keys)
    exact = neg.astype(np.float64) + dneg.astype(np.float64)
    assert abs(float(np.mean(out.astype(np.float64) - exact))) < 3e-10
