"""GPU parity: the gfx950 library through its C ABI against the oracle.

The f64 and Dual steps are compared bit for bit with the oracle running the
same arithmetic (pow_mode 1 = the kernels' correctly rounded pow; every other
operation is IEEE-exact on both sides and -ffp-contract=off on both). The
reference-faithful oracle (glibc pow) is then compared at the tolerances the
task states (1e-10 m / 1e-10 m^2/s), and the reference's published vectors
are checked directly on the GPU output.
"""
import math
import os
import sys

import numpy as np
import pytest

import cases
from oracle.pyoracle import OracleSim, set_pow_mode, set_threads

pytestmark = pytest.mark.gpu


def _hf():
    import hydroflux_amd

    return hydroflux_amd


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _mode():
    set_threads(16)
    set_pow_mode(1)
    yield
    set_pow_mode(0)


def _pair(case, dual=False, **tangents):
    g = case.load(_hf().Grid2D(case.nr, case.nc, case.dx, case.dy, dual=dual), **tangents)
    o = case.load(OracleSim(case.nr, case.nc, case.dx, case.dy, dual=dual), **tangents)
    return g, o


def _csv(golden_dir, name):
    return np.loadtxt(os.path.join(golden_dir, name), delimiter=",", skiprows=1)


def test_device_arithmetic_contract_is_bit_exact():
    # sqrt and / must be correctly rounded on the device, and the kernels' pow
    # must equal the host build of the same routine, for the step to be bitwise.
    _say("math probe")
    import ctypes as C

    from hydroflux_amd._lib import lib
    from oracle.pyoracle import lib as olib

    rng = np.random.default_rng(1)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 200000), rng.uniform(0, 20, 100000),
                        np.array([0.0, 1.0, 2.0, 4.0, 8.0, 1e-12, 5e-324, np.inf])])
    y = 10.0 ** rng.uniform(-50, 50, x.size)
    out = np.empty(4 * x.size)
    D = C.POINTER(C.c_double)
    assert lib().hf2d_math_probe(x.ctypes.data_as(D), y.ctypes.data_as(D), x.size, out.ctypes.data_as(D)) == 0
    out = out.reshape(-1, 4)
    assert np.array_equal(out[:, 0], np.sqrt(x))
    assert np.array_equal(out[:, 1], x / y)
    sub = slice(0, None, 37)
    L = olib()
    assert np.array_equal(out[sub, 2], np.array([L.hfo_pow_4_3(float(v)) for v in x[sub]]))
    assert np.array_equal(out[sub, 3], np.array([L.hfo_pow_1_3m(float(v)) for v in x[sub]]))


@pytest.mark.parametrize("name", ["stoker400", "stoker100", "thacker80", "macdonald", "uk_ea_t6"])
def test_reference_loops_bitwise_and_golden(name, golden_dir):
    _say(f"reference loop {name}")
    case, fr, csv, row, dec = {
        "stoker400": (cases.stoker(400), None, "verif_stoker.csv", 1, 6),
        "stoker100": (cases.stoker(100), None, "verif_stoker_coarse.csv", 1, 6),
        "thacker80": (cases.thacker(80), None, "verif_thacker.csv", 40, 6),
        "macdonald": (cases.macdonald(), 1e-9, "verif_macdonald.csv", 2, 6),
        "uk_ea_t6": (cases.uk_ea_t6(), 1e-9, "verif_uk_ea_t6.csv", None, 5),
    }[name]
    g, o = _pair(case)
    tg, kg = cases.hand_loop(g, case.extra["t_end"], friction_tol=fr)
    to, ko = cases.hand_loop(o, case.extra["t_end"], friction_tol=fr)
    assert (tg, kg) == (to, ko)
    sg = g.get_state()
    assert np.array_equal(sg, o.get_state())
    h = sg[..., 0].reshape(-1) if row is None else sg[row, :, 0]
    assert np.array_equal(np.round(h, dec), _csv(golden_dir, csv)[:, 2])


def test_dam_break_256_500_steps_fused_driver():
    # 256x256 flat frictionless dam-break, dx 0.1, walls, CFL 0.5, 500 steps,
    # device-clocked driver (no host dt round trip) vs the oracle's loop.
    _say("dam-break 256^2 x 500 steps")
    case = cases.dam_break_256()
    g, o = _pair(case)
    for s in (g, o):
        s.configure(cfl=0.5, cfl_with_bcs=False, max_dt=math.inf, friction=False)
    g.run_steps(500)
    assert o.run_steps(500) == 0
    assert g.time() == o.time()
    sg, so = g.get_state(), o.get_state()
    assert np.array_equal(sg, so)
    m0 = case.state[..., 0].sum()
    assert abs(sg[..., 0].sum() - m0) / m0 < 1e-12


def test_reference_pow_oracle_within_tolerance():
    # Against the reference-faithful oracle (glibc pow): 1e-10 on h, hu, hv.
    _say("bumpy driver vs glibc-pow oracle")
    c = cases.bumpy_wet_dry(61, 83, 21)
    g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy))
    set_pow_mode(0)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy))
    for s in (g, o):
        s.configure(cfl=0.4, max_dt=0.1, rain_rate=2e-4, sources=[(30, 10, 0.4)])
    g.run_steps(60)
    o.run_steps(60)
    d = np.abs(g.get_state() - o.get_state())
    assert d.max() <= 1e-10, d.max()


@pytest.mark.parametrize("seed", [21, 22])
def test_driver_with_forcing_bitwise(seed):
    _say(f"bumpy driver rain+sources seed {seed}")
    c = cases.bumpy_wet_dry(77, 131, seed)
    g, o = _pair(c)
    src = [(c.nr // 2, c.nc // 3, 0.7), (1, 1, -0.2), (c.nr // 2, c.nc // 3, 0.1), (c.nr - 1, c.nc - 1, 2.0)]
    for s in (g, o):
        s.configure(cfl=0.45, max_dt=0.05, friction_dry_tol=1e-6, rain_rate=1e-3, sources=src)
    g.run_until(1.5)
    assert o.run_until(1.5) == 0
    assert g.time() == o.time()
    assert np.array_equal(g.get_state(), o.get_state())


@pytest.mark.parametrize("west", [cases.DISCHARGE, cases.DEPTH], ids=["discharge", "depth"])
@pytest.mark.parametrize("kind", ["f64", "dual", "f32"])
def test_active_tile_list_bitwise(kind, west):
    """Mostly dry domain (Discharge or Depth inflow onto dry cells, pond,
    stale films, point sources): the compacted active-tile list gives the
    same bits as a run over every tile, and the oracle's (f64 and Dual;
    Dual runs every tile). The Discharge-onto-dry ghost depth
    (n |q| / sqrt(S))^(3/5) uses the deterministic power on both sides."""
    _say(f"active-tile list {kind} west={west}")
    c = cases.mostly_dry(150, 230, 3, west=west)
    hf = _hf()
    kw = {}
    if kind == "dual":
        rng = np.random.default_rng(8)
        kw = dict(manning_tangent=c.manning, state_tangent=0.1 * rng.standard_normal((c.nr, c.nc, 3)))
    sims = [c.load(hf.Grid2D(c.nr, c.nc, c.dx, c.dy, kind=kind), **kw) for _ in range(2)]
    src = [(c.nr - 9, 11, 0.05), (20, c.nc - 30, 0.02)]
    for s, on in zip(sims, ("always", False)):
        s.configure(cfl=0.4, max_dt=0.2, rain_rate=0.0, sources=src, active_tiles=on)
        s.run_steps(60)
    assert sims[0].time() == sims[1].time()
    a, b = sims[0].get_state(), sims[1].get_state()
    if kind == "dual":
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1], equal_nan=True)
    else:
        assert np.array_equal(a, b), "list on != list off"
    if kind != "f32":
        o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=kind == "dual"), **kw)
        o.configure(cfl=0.4, max_dt=0.2, rain_rate=0.0, sources=src)
        o.run_steps(60)
        r = o.get_state()
        assert o.time() == sims[0].time()
        if kind == "dual":
            assert np.array_equal(a[0], r[0]) and np.array_equal(a[1], r[1], equal_nan=True)
        else:
            assert np.array_equal(a, r), "gpu != oracle"


@pytest.mark.parametrize("rain", [0.0, 2e-4], ids=["list", "rain"])
def test_graph_replays_interleaved_with_host_calls_bitwise(rain):
    """The driver replays captured 16-step graphs; eager steps, host-side
    state changes and reconfiguration between replays must not leave a
    stale graph or a stale active-tile list."""
    _say(f"graph replays rain={rain}")
    c = cases.mostly_dry(90, 130, 6, west=cases.DEPTH)
    g, o = _pair(c)
    src = [(c.nr - 9, 11, 0.05)]
    g.configure(cfl=0.4, max_dt=0.2, rain_rate=rain, sources=src, active_tiles="always")
    o.configure(cfl=0.4, max_dt=0.2, rain_rate=rain, sources=src)
    for s in (g, o):
        s.run_steps(40)           # 2 replays + 8 eager steps
    assert np.array_equal(g.get_state(), o.get_state())
    for s in (g, o):
        s.apply_rain(1e-3, 0.5)   # host-side state change between replays
        s.run_steps(37)
    assert g.time() == o.time() and np.array_equal(g.get_state(), o.get_state())
    st = g.get_state()
    st[45, 60] = (0.4, 0.1, 0.0)
    g.set_state(st)
    g.configure(cfl=0.35, max_dt=0.1, rain_rate=rain, sources=src + [(50, 70, 0.01)], active_tiles="always")
    o.set_state(st)
    o.configure(cfl=0.35, max_dt=0.1, rain_rate=rain, sources=src + [(50, 70, 0.01)])
    for s in (g, o):
        s.run_until(s.time()[0] + 3.0)
    assert g.time() == o.time() and np.array_equal(g.get_state(), o.get_state())
    g.set_kernel_timing(True)     # eager path with events, then graphs again
    o.run_steps(20)
    g.run_steps(20)
    g.set_kernel_timing(False)
    for s in (g, o):
        s.run_steps(48)
    assert g.time() == o.time() and np.array_equal(g.get_state(), o.get_state())


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_dualn_bitwise(n):
    """Vector-mode DualN<N> (autograd/src/dual_n.rs) on the GPU: the fused
    driver with rain and point sources, then a host-side loop of
    cfl_time_step / ssprk2_step / friction; values and all N tangents (bed,
    Manning and state seeds) bit for bit with the oracle."""
    _say(f"dualn<{n}>")
    c = cases.bumpy_wet_dry(41, 57, 12)
    c.manning = np.maximum(c.manning, 0.03)
    rng = np.random.default_rng(n)
    kw = dict(manning_tangent=0.01 * rng.standard_normal((c.nr, c.nc, n)),
              bed_tangent=0.01 * rng.standard_normal((c.nr, c.nc, n)),
              state_tangent=0.1 * rng.standard_normal((c.nr, c.nc, 3, n)))
    g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy, kind=f"dual{n}"), **kw)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, tangents=n), **kw)
    for s in (g, o):
        s.configure(cfl=0.4, max_dt=0.05, rain_rate=2e-4, sources=[(5, 7, 0.2)])
        s.run_steps(20)
    assert g.time() == o.time()
    (gv, gt), (ov, ot) = g.get_state(), o.get_state()
    assert np.array_equal(gv, ov) and np.array_equal(gt, ot, equal_nan=True)
    for _ in range(10):
        dg, do = g.cfl_time_step(0.4), o.cfl_time_step(0.4)
        assert dg == do
        g.ssprk2_step(dg)
        o.ssprk2_step(do)
        g.manning_friction_step(dg, 1e-6)
        o.manning_friction_step(do, 1e-6)
    (gv, gt), (ov, ot) = g.get_state(), o.get_state()
    assert gt.shape == (c.nr, c.nc, 3, n)
    assert np.array_equal(gv, ov) and np.array_equal(gt, ot, equal_nan=True)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_dualn_large_tiles_and_groups_bitwise(n):
    """DualN<2> on its 24 x 8 tiles, DualN<4> on its 16 x 8 tiles, DualN<8>
    and DualN<16> as 2 and 4 launches of the DualN<4> kernel per stage, one
    group of 4 tangent components each. 520 x 600 (1,625 tiles of 24 x 8,
    2,475 of 16 x 8): the fused driver with rain and a point source, then
    the host loop, values and all N tangents bit for bit with the oracle."""
    _say(f"dualn<{n}> default tiles 520x600")
    c = cases.bumpy_wet_dry(520, 600, 3)
    c.manning = np.maximum(c.manning, 0.03)
    rng = np.random.default_rng(100 + n)
    kw = dict(manning_tangent=0.01 * rng.standard_normal((c.nr, c.nc, n)),
              state_tangent=0.1 * rng.standard_normal((c.nr, c.nc, 3, n)))
    g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy, kind=f"dual{n}"), **kw)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, tangents=n), **kw)
    for s in (g, o):
        s.configure(cfl=0.4, max_dt=0.05, rain_rate=2e-4, sources=[(260, 300, 0.5)])
        s.run_steps(6)
    assert g.time() == o.time()
    (gv, gt), (ov, ot) = g.get_state(), o.get_state()
    assert np.array_equal(gv, ov) and np.array_equal(gt, ot, equal_nan=True)
    for _ in range(3):
        dg, do = g.cfl_time_step(0.4), o.cfl_time_step(0.4)
        assert dg == do
        g.ssprk2_step(dg)
        o.ssprk2_step(do)
        g.manning_friction_step(dg, 1e-6)
        o.manning_friction_step(do, 1e-6)
    (gv, gt), (ov, ot) = g.get_state(), o.get_state()
    assert np.array_equal(gv, ov) and np.array_equal(gt, ot, equal_nan=True)


@pytest.mark.parametrize("kind,nr,nc", [("f64", 520, 600), ("f64", 260, 300), ("f64", 150, 100), ("dual", 520, 600),
                                         ("dual", 150, 100)])
def test_default_mid_and_small_grid_tiles_bitwise(kind, nr, nc):
    """A grid takes the default f64 stage tiles (32 x 24) when they number
    at least 160, else 32 x 16 tiles when those number at least 128, else
    16 x 8: 520 x 600 (418 default tiles), 260 x 300 (110 default, 170 of
    32 x 16) and 150 x 100 (40 of 32 x 16) run the three kernels, each bit
    for bit with the oracle through the fused driver with friction, rain
    and point sources and the host loop. Dual takes 24 x 16 tiles from 160
    of them (520 x 600: 726) and 16 x 8 below (150 x 100: 70 of 24 x 16)."""
    _say(f"large-tile kernels {kind} {nr}x{nc}")
    c = cases.bumpy_wet_dry(nr, nc, 3)
    kw = {"manning_tangent": np.where(c.manning > 0.04, 1.0, 0.0)} if kind == "dual" else {}
    g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy, kind=kind), **kw)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=kind == "dual"), **kw)
    for s in (g, o):
        s.configure(cfl=0.4, max_dt=0.05, rain_rate=3e-4, sources=[(nr // 2, 7, 0.3), (10, nc - 9, -0.01)])
        s.run_steps(6)
        dt = s.cfl_time_step(0.4)
        s.ssprk2_step(dt)
        s.manning_friction_step(dt, 1e-6)
    assert g.time() == o.time()
    assert np.array_equal(g.get_state(), o.get_state())


RANDOM_SHAPES = [(23, 17), (150, 100), (260, 300), (520, 600), (61, 389), (300, 41), (97, 211), (400, 200)]


@pytest.mark.parametrize("seed", range(len(RANDOM_SHAPES)))
def test_randomized_configurations_bitwise(seed):
    """Seeded random configurations across the three f64 tile shapes and
    the Dual kind: boundary kinds on each side (Discharge q, Depth h),
    rain, point sources and sinks, friction dry tolerance. The fused driver,
    then the host loop (boundary-aware CFL, SSP-RK2, friction, rain, point
    sources), bit for bit with the oracle."""
    rng = np.random.default_rng(4000 + seed)
    nr, nc = RANDOM_SHAPES[seed]
    kind = "dual" if seed % 3 == 2 else "f64"
    c = cases.bumpy_wet_dry(nr, nc, seed)
    kinds = tuple(int(k) for k in rng.choice([cases.WALL, cases.TRANS, cases.DISCHARGE, cases.DEPTH], size=4))
    c.bc_kinds = kinds
    c.bc_params = tuple(0.4 if k == cases.DISCHARGE else (0.7 if k == cases.DEPTH else 0.0) for k in kinds)
    rain = float(rng.choice([0.0, 2e-4]))
    srcs = [(int(rng.integers(nr)), int(rng.integers(nc)), float(rng.uniform(-0.05, 0.3)))
            for _ in range(int(rng.integers(0, 3)))]
    tol = float(rng.choice([1e-6, 1e-9]))
    _say(f"random config {seed}: {kind} {nr}x{nc} bcs {kinds} rain {rain} sources {len(srcs)} tol {tol}")
    kw = dict(manning_tangent=c.manning) if kind == "dual" else {}
    g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy, kind=kind), **kw)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=kind == "dual"), **kw)
    for s in (g, o):
        s.configure(cfl=0.4, max_dt=0.05, rain_rate=rain, sources=srcs, friction_dry_tol=tol)
        s.run_steps(12)
        for _ in range(3):
            dt = s.cfl_time_step(0.4, True)
            s.ssprk2_step(dt)
            s.manning_friction_step(dt, tol)
            s.apply_rain(rain, dt)
            s.apply_point_sources(srcs, dt)
    assert g.time() == o.time()
    if kind == "dual":
        (gv, gt), (ov, ot) = g.get_state(), o.get_state()
        assert np.array_equal(gv, ov) and np.array_equal(gt, ot, equal_nan=True)
    else:
        assert np.array_equal(g.get_state(), o.get_state())


@pytest.mark.parametrize("between", ["synchronize", "enqueue_sync_set_state"])
def test_synchronize_does_not_validate_stale_cfl_slots(between):
    """hf2d_synchronize records the driver's CFL maxima only after enqueued
    driver steps: after set_state, a synchronize must leave the next driver
    call to recompute them (a stale, zeroed slot would give dt = the whole
    remaining time)."""
    _say(f"synchronize after state change: {between}")
    c = cases.bumpy_wet_dry(40, 52, 5)
    runs = []
    for sync in (False, True):
        g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy))
        g.configure(cfl=0.4, max_dt=1e9)
        if sync and between == "enqueue_sync_set_state":
            g.enqueue_steps(3)
            g.set_state(c.state)
            g.set_time(0.0, 0)
        if sync:
            g.synchronize()
        g.run_until(2.0)
        runs.append((g.time(), g.get_state()))
        g.close()
    assert runs[0][0] == runs[1][0] and runs[0][0][1] > 5
    assert np.array_equal(runs[0][1], runs[1][1])


def test_forward_euler_driver_bitwise():
    _say("forward euler driver")
    c = cases.bumpy_wet_dry(40, 70, 5)
    g, o = _pair(c)
    for s in (g, o):
        s.configure(cfl=0.4, integrator=0, max_dt=0.05, rain_rate=5e-4)
    g.run_steps(40)
    o.run_steps(40)
    assert np.array_equal(g.get_state(), o.get_state())


def test_dual_driver_bitwise_with_all_tangents():
    _say("dual driver")
    c = cases.bumpy_wet_dry(45, 67, 9)
    c.manning = np.maximum(c.manning, 0.03)
    rng = np.random.default_rng(3)
    tang = dict(manning_tangent=c.manning.copy(), bed_tangent=0.01 * rng.standard_normal((c.nr, c.nc)),
                state_tangent=0.1 * rng.standard_normal((c.nr, c.nc, 3)))
    g, o = _pair(c, dual=True, **tang)
    for s in (g, o):
        s.configure(cfl=0.4, max_dt=0.5, rain_rate=2e-4, sources=[(1, 2, 0.3)])
        s.run_steps(40)
    (vg, tg), (vo, to) = g.get_state(), o.get_state()
    assert np.array_equal(vg, vo) and np.array_equal(tg, to)


def test_dual_inverse_manning_trace_on_gpu(golden_dir):
    # m1_inverse_manning_demo.rs:42-108 run entirely on the GPU (Dual kind)
    _say("dual inverse-manning newton trace")
    ref = _csv(golden_dir, "m1_inverse_manning.csv")
    c = cases.inverse_manning_channel()

    def centroid(n_val, dual):
        g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy, dual=dual), manning_tangent=(1.0 if dual else None))
        g.set_mesh(c.bed, n_val, manning_tangent=(1.0 if dual else None))
        cases.hand_loop(g, 20.0, friction_tol=1e-6, max_steps=5000)
        out = g.get_state()
        h = (out[0] if dual else out)[1, :, 0]
        ht = out[1][1, :, 0] if dual else np.zeros_like(h)
        num = den = dnum = dden = 0.0
        for j in range(c.nc):
            x = j + 0.5
            num, dnum = num + h[j] * x, dnum + ht[j] * x
            den, dden = den + h[j], dden + ht[j]
        return num / den, (dnum * den - num * dden) / (den * den)

    target, _ = centroid(0.04, False)
    n = 0.08
    for it in range(8):
        cv, dc = centroid(n, True)
        err = cv - target
        assert f"{cv:.10f}" == f"{ref[it, 2]:.10f}"
        assert float(f"{dc:.10e}") == ref[it, 4]
        n = min(max(n - err / dc if abs(dc) > 1e-12 else n, 1e-4), 1.0)


@pytest.mark.parametrize("kind", ["f64", "f64fast"])
def test_outflow_limiter_closed_form_on_gpu(kind):
    """The Liang-Marche limiter where it binds (update.rs:953-1053) on the
    device: on a one-row chain of two draining cells between walls each
    cell's available depth (h - 1e-6) moves exactly one cell downstream, and
    an isolated draining cell ends at 1e-6 (to 4 ulp of the depths; the
    fast kind's reciprocal divisions to 1e-15 m); the bitwise kind also
    equals the oracle bit for bit."""
    _say(f"outflow limiter closed form {kind}")
    h_dry = 1.0e-6
    tol = 4 * float(np.spacing(0.05)) if kind == "f64" else 1e-15
    st = np.zeros((1, 5, 3))
    st[0, 0] = (0.05, 0.25, 0.0)
    st[0, 1] = (0.02, 0.10, 0.0)
    iso = np.zeros((3, 3, 3))
    iso[1, 1] = (0.05, 0.05 * 6.0 * math.cos(1.9), 0.05 * 6.0 * math.sin(1.9))
    for state, dt in ((st, 0.5), (iso, 0.5)):
        nr, nc = state.shape[:2]
        outs = []
        sims = [_hf().Grid2D(nr, nc, 1.0, 1.0, kind=kind)]
        if kind == "f64":
            set_pow_mode(1)
            sims.append(OracleSim(nr, nc, 1.0, 1.0))
        for s in sims:
            s.set_mesh(np.zeros((nr, nc)), 0.0)
            s.set_state(state)
            s.set_boundaries([1, 1, 1, 1])
            s.forward_euler_step(dt)
            outs.append(s.get_state())
        h = outs[0][..., 0]
        if nr == 1:
            assert abs(h[0, 0] - h_dry) <= tol and abs(h[0, 1] - 0.05) <= tol, h
            assert abs(h[0, 2] - (0.02 - h_dry)) <= tol and h[0, 3] == 0.0 and h[0, 4] == 0.0, h
        else:
            assert abs(h[1, 1] - h_dry) <= tol and np.all(h >= 0.0), h
        assert abs(h.sum() - state[..., 0].sum()) <= tol
        if len(outs) == 2:
            assert np.array_equal(outs[0], outs[1])


def test_lake_at_rest_and_mass_on_gpu():
    _say("lake at rest + mass")
    n = 64
    y, x = np.mgrid[0:n, 0:n].astype(float)
    bed = 1.2 * np.exp(-((x - 30) ** 2 + (y - 33) ** 2) / 60.0) + 0.01 * x
    eta = 1.0
    st = np.zeros((n, n, 3))
    st[..., 0] = np.maximum(eta - bed, 0.0)
    g = _hf().Grid2D(n, n, 1.0, 1.0)
    g.set_mesh(bed, 0.03)
    g.set_state(st)
    g.set_boundaries([1, 1, 1, 1])
    g.configure(cfl=0.4, max_dt=10.0)
    g.run_steps(200)
    out = g.get_state()
    wet = st[..., 0] > 0
    assert np.max(np.abs(out[wet, 0] + bed[wet] - eta)) <= 1e-13
    assert np.max(np.abs(out[..., 1:])) <= 1e-12
    assert abs(out[..., 0].sum() - st[..., 0].sum()) / st[..., 0].sum() <= 1e-12


def test_errors_through_the_abi():
    _say("error paths")
    hf = _hf()
    c = cases.stoker(20)
    g = hf.Grid2D(c.nr, c.nc, c.dx, c.dy)
    g.set_mesh(c.bed, 0.0)
    g.set_state(np.zeros_like(c.state))
    g.configure(max_dt=math.inf)
    with pytest.raises(hf.DegenerateDt):
        g.run_steps(1)
    g.set_state(c.state)
    g.configure(max_steps=3)
    with pytest.raises(hf.StepBudgetExhausted):
        g.run_until(100.0)
    assert g.time()[1] == 3
    with pytest.raises(hf.InvalidConfig):
        g.configure(cfl=1.5)
    with pytest.raises(hf.ShapeMismatch):
        g.set_state(np.zeros((c.nr + 1, c.nc, 3)))
    with pytest.raises(hf.ShapeMismatch):
        g.apply_point_sources([(c.nr, 0, 1.0)], 0.1)
    # a NaN injected into one cell surfaces as DegenerateDt? No: f64::max ignores
    # NaN (update.rs:132), so the reference keeps stepping; so does the device.
    st = c.state.copy()
    st[1, 3, 1] = np.nan
    g.set_state(st)
    g.set_time(0.0, 0)
    g.configure(max_dt=0.01)
    g.run_steps(2)
    assert g.time()[1] == 2


@pytest.mark.parametrize("change", ["set_boundaries", "configure_dual", "kernel_timing", "gauges"])
def test_graph_dropped_while_enqueued_steps_run(change):
    """hf2d_enqueue_steps returns with graph replays in flight; a call that
    drops the captured graph right after it (new boundaries, a Dual
    reconfiguration without sources, kernel timing, gauges) must not
    destroy it under the running replays. Results stay the oracle's bits."""
    _say(f"graph dropped with replays in flight: {change}")
    dual = change == "configure_dual"
    c = cases.bumpy_wet_dry(384, 448, 17)
    kw = dict(manning_tangent=c.manning) if dual else {}
    g = c.load(_hf().Grid2D(c.nr, c.nc, c.dx, c.dy, dual=dual), **kw)
    o = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=dual), **kw)
    for s in (g, o):
        s.configure(cfl=0.4, max_dt=0.05, rain_rate=2e-4)
    g.enqueue_steps(64)
    o.run_steps(64)
    if change == "set_boundaries":
        for s in (g, o):
            s.set_boundaries([cases.WALL, cases.TRANS, cases.WALL, cases.TRANS])
    elif change == "configure_dual":
        for s in (g, o):
            s.configure(cfl=0.35, max_dt=0.05, rain_rate=2e-4)
    elif change == "kernel_timing":
        g.set_kernel_timing(True)
    else:
        g.set_gauges([(10, 10), (200, 300)], capacity=64)
    g.synchronize()
    for s in (g, o):
        s.run_steps(40)
    assert g.time() == o.time()
    if dual:
        (gv, gt), (ov, ot) = g.get_state(), o.get_state()
        assert np.array_equal(gv, ov) and np.array_equal(gt, ot, equal_nan=True)
    else:
        assert np.array_equal(g.get_state(), o.get_state())
    g.close()
