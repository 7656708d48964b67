"""Pins the oracle's Dual / DualN<4> rain and point sources (CPU only).

The reference's apply_rain and apply_point_sources take f64 states only
(solver-2d/src/source.rs:13-17, 52-101); the oracle runs them on Dual and
DualN states as well, so that the driver can carry tangents through the
forcing the reference's callers apply after every step. Nothing in the
reference holds a vector for that extension, so these tests pin it by what
the mathematics fixes:

- the forcing is a constant in the differentiated parameter, so on a cell
  that stays wet the value equals the f64 routine's and the tangents pass
  through bit for bit;
- a cell clamped dry (evaporation with h + dh <= 0, source.rs:95-96; a sink
  deeper than the cell, source.rs:63-67) is locally constant, so its value
  and every tangent are exactly zero;
- over a whole run with friction, rain, an inflow and a sink the tangents
  equal central finite differences of f64 runs at the perturbed parameter
  (the reference's own check pattern for the step, update.rs:1953-2038, and
  for friction, source.rs:421-473), for a Manning-multiplier seed and an
  initial-depth seed.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleSim, set_pow_mode


@pytest.fixture(autouse=True)
def _glibc_pow():
    set_pow_mode(0)
    yield
    set_pow_mode(0)


def _states(rng, nr, nc, n_tan):
    st = np.zeros((nr, nc, 3))
    st[..., 0] = rng.uniform(0.0, 0.02, (nr, nc))
    st[..., 1:] = 0.01 * rng.standard_normal((nr, nc, 2))
    st[0, 0, 0] = 0.0  # a dry cell
    shape = (nr, nc, 3) if n_tan == 1 else (nr, nc, 3, n_tan)
    tg = rng.standard_normal(shape)
    return st, tg


def _sim(nr, nc, n_tan):
    return OracleSim(nr, nc, 1.0, 1.0, dual=(n_tan == 1), tangents=(None if n_tan == 1 else n_tan))


@pytest.mark.parametrize("n_tan", [1, 4], ids=["dual", "dualn4"])
@pytest.mark.parametrize("rate", [1.4e-5, -2.0e-4], ids=["rain", "evaporation"])
def test_rain_passes_tangents_through_and_zeroes_the_dry_clamp(n_tan, rate):
    rng = np.random.default_rng(21)
    nr, nc, dt = 12, 17, 40.0
    st, tg = _states(rng, nr, nc, n_tan)
    dh = rate * dt
    if rate < 0:  # one cell lands exactly on h + dh == 0 (clamped: new_h <= 0)
        st[3, 4, 0] = -dh
    f = OracleSim(nr, nc, 1.0, 1.0)
    f.set_state(st)
    f.apply_rain(rate, dt)
    ref = f.get_state()
    d = _sim(nr, nc, n_tan)
    d.set_state(st, tg)
    d.apply_rain(rate, dt)
    val, tan = d.get_state()
    assert np.array_equal(val, ref)  # the value path is the f64 routine's
    clamped = st[..., 0] + dh <= 0.0
    assert (rate > 0) == (not clamped.any())
    assert np.all(val[clamped] == 0.0)
    assert np.all(tan[clamped] == 0.0)
    assert np.array_equal(tan[~clamped], tg[~clamped])


@pytest.mark.parametrize("n_tan", [1, 4], ids=["dual", "dualn4"])
def test_point_sources_pass_tangents_through_and_zero_the_sink_clamp(n_tan):
    rng = np.random.default_rng(22)
    nr, nc, dt = 9, 11, 2.0
    st, tg = _states(rng, nr, nc, n_tan)
    st[2, 3, 0] = 0.3
    st[5, 5, 0] = 0.004
    st[7, 1, 0] = 0.5
    # inflow on a wet cell; a sink deeper than its cell; a sink that leaves
    # water; a sink and an inflow on one cell, applied in list order
    srcs = [(2, 3, 0.1), (5, 5, -0.01), (4, 6, -0.0001), (7, 1, -0.15), (7, 1, 0.025)]
    f = OracleSim(nr, nc, 1.0, 1.0)
    f.set_state(st)
    f.apply_point_sources(srcs, dt)
    ref = f.get_state()
    d = _sim(nr, nc, n_tan)
    d.set_state(st, tg)
    d.apply_point_sources(srcs, dt)
    val, tan = d.get_state()
    assert np.array_equal(val, ref)
    assert val[5, 5, 0] == 0.0 and np.all(val[5, 5] == 0.0)  # the sink clamp
    assert np.all(tan[5, 5] == 0.0)
    touched = np.zeros((nr, nc), bool)
    for r, c, _ in srcs:
        touched[r, c] = True
    keep = touched.copy()
    keep[5, 5] = False
    assert np.array_equal(tan[keep], tg[keep])  # wet after every source: tangents unchanged
    assert np.array_equal(tan[~touched], tg[~touched])
    assert np.array_equal(val[~touched], st[~touched])


def test_point_source_exact_tie_follows_the_dual_max_rule():
    """h + dh == 0 exactly: max(x, 0) at a tie averages the tangents
    (autograd/src/dual.rs:130-155), so h keeps half its tangent at value 0
    and the momentum is zeroed (value and tangent)."""
    st = np.zeros((1, 2, 3))
    st[0, 0] = (0.25, 0.1, -0.2)
    tg = np.zeros((1, 2, 3))
    tg[0, 0] = (3.0, 5.0, 7.0)
    d = OracleSim(1, 2, 2.0, 2.0, dual=True)
    d.set_state(st, tg)
    d.apply_point_sources([(0, 0, -0.5)], 2.0)  # dh = -0.5 * 2 / 4 = -0.25
    val, tan = d.get_state()
    assert np.all(val[0, 0] == 0.0)
    assert tan[0, 0, 0] == 1.5 and tan[0, 0, 1] == 0.0 and tan[0, 0, 2] == 0.0


def _wet_case(nr=24, nc=30):
    # a wet domain whose surface is nowhere flat: the minmod limiter is only
    # piecewise smooth, and on a lake at rest its face differences are zero
    # to rounding, where a perturbation flips its choice
    y, x = np.mgrid[0:nr, 0:nc]
    bed = 0.004 * x + 0.05 * np.sin(x / 3.0) * np.cos(y / 4.0)
    st = np.zeros((nr, nc, 3))
    eta = 0.6 + 0.03 * np.sin(x / 5.0 + 0.3) * np.cos(y / 6.0 + 0.2) + 0.1 * (1 - np.tanh((x - nc / 3) / 3.0)) / 2
    st[..., 0] = np.maximum(eta - bed, 0.0)
    land = ((x // 8) + (y // 7)) % 3
    man = np.choose(land, [0.03, 0.05, 0.08])
    return bed, st, man, land


def _hand_loop(sim, dts, rate, srcs):
    # the reference callers' order: step, friction, rain, point sources
    for dt in dts:
        sim.ssprk2_step(dt)
        sim.manning_friction_step(dt, 1e-6)
        sim.apply_rain(rate, dt)
        sim.apply_point_sources(srcs, dt)


@pytest.mark.parametrize("n_tan,seed", [(1, 0), (1, 1), (4, None)], ids=["dual-manning", "dual-depth", "dualn4"])
def test_forced_run_tangents_equal_central_differences(n_tan, seed):
    """30 SSP-RK2 steps with friction, 50 mm/h rain, a 0.8 m3/s inflow and a
    0.05 m3/s sink on a wet 24 x 30 grid. Seeds: (0) a global Manning
    multiplier m (n = m n0); (1) the amplitude of a Gaussian bump of initial
    depth; DualN<4> also (2) the Manning n of one land class and (3) a bump
    of initial depth at the inflow cell. Tangents are compared with
    (f(p + e) - f(p - e)) / 2e of f64 runs over the same dt sequence."""
    bed, st0, man, land = _wet_case()
    nr, nc = bed.shape
    rate, srcs = 50e-3 / 3600.0, [(12, 2, 0.8), (6, 25, -0.05)]
    base = OracleSim(nr, nc, 1.0, 1.0)
    base.set_mesh(bed, man)
    base.set_state(st0)
    base.set_boundaries([1, 1, 1, 0])
    dts = [0.4 * base.cfl_time_step(0.4, True)] * 30
    y, x = np.mgrid[0:nr, 0:nc]
    bump = np.exp(-((x - 15.0) ** 2 + (y - 11.0) ** 2) / 18.0)
    seeds = [("manning", man), ("depth", bump)]
    if n_tan == 4:
        seeds += [("manning", np.where(land == 2, man, 0.0)), ("depth", np.exp(-((x - 2.0) ** 2 + (y - 12.0) ** 2) / 8.0))]
    else:
        seeds = [seeds[seed]]

    # automatic derivatives
    d = OracleSim(nr, nc, 1.0, 1.0, dual=(n_tan == 1), tangents=(None if n_tan == 1 else n_tan))
    mt = np.zeros((nr, nc) if n_tan == 1 else (nr, nc, n_tan))
    stt = np.zeros((nr, nc, 3) if n_tan == 1 else (nr, nc, 3, n_tan))
    for k, (kind, field) in enumerate(seeds):
        if kind == "manning":
            if n_tan == 1:
                mt[...] = field
            else:
                mt[..., k] = field
        else:
            if n_tan == 1:
                stt[..., 0] = field
            else:
                stt[..., 0, k] = field
    d.set_mesh(bed, man, manning_tangent=mt)
    d.set_state(st0, stt)
    d.set_boundaries([1, 1, 1, 0])
    _hand_loop(d, dts, rate, srcs)
    val, tan = d.get_state()

    def run(man_p, st_p):
        s = OracleSim(nr, nc, 1.0, 1.0)
        s.set_mesh(bed, man_p)
        s.set_state(st_p)
        s.set_boundaries([1, 1, 1, 0])
        _hand_loop(s, dts, rate, srcs)
        return s.get_state()

    ref = run(man, st0)
    assert np.max(np.abs(val - ref)) <= 1e-12  # values: Dual bit-identical, DualN's a*(1/b) within rounding
    # Manning enters smoothly: the central difference is good to ~1e-7 at
    # eps = 1e-5. A depth perturbation moves face differences across the
    # minmod limiter's switch points (|a| = |b|), where the run is only
    # piecewise smooth: that error grows with eps (4e-6 at 1e-7, 4e-4 at
    # 1e-6), so depth seeds take a small eps and a looser bound.
    for k, (kind, field) in enumerate(seeds):
        eps, tol = (1e-5, 1e-6) if kind == "manning" else (2e-8, 2e-5)
        if kind == "manning":
            plus, minus = run(man + eps * field, st0), run(man - eps * field, st0)
        else:
            sp, sm = st0.copy(), st0.copy()
            sp[..., 0] += eps * field
            sm[..., 0] -= eps * field
            plus, minus = run(man, sp), run(man, sm)
        fd = (plus - minus) / (2 * eps)
        ad = tan if n_tan == 1 else tan[..., k]
        scale = np.max(np.abs(ad), axis=(0, 1))
        assert np.all(scale > 0)
        err = np.max(np.abs(ad - fd), axis=(0, 1)) / scale
        assert np.all(err < tol), (kind, k, err)
