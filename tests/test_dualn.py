"""Vector-mode DualN<N> (autograd/src/dual_n.rs) in the oracle.

DualN follows its own arithmetic (combine / chain, division through the
reciprocal, 0.5/s sqrt factor, tie averaging as 0.5 a + 0.5 b), so a DualN
run's values are not bit-equal to the f64 run's. What pins it:

- the reference's own unit test of DualN (components match independent
  seeded Dual passes, at its tolerances 1e-14 / 1e-12) and its special
  cases (sqrt at 0, ties, the quotient rule per component);
- in the solver (m1_forward_scaling.rs's zonal Manning setup: dam break,
  transmissive, fixed dt, friction), with the oracle's quotient switched to
  f64's value a / b, DualN values are the f64 run's bit for bit and each
  tangent equals the Dual run seeded on the same parameter to 1e-13;
- with the reference's quotient a * (1 / b), the values stay within 1e-12
  of f64, but the tangents can differ from the seeded Dual runs cell by
  cell: where the value paths differ in the last bit, a max/min near-tie or
  a minmod |a| < |b| comparison resolves differently, the tangent of a
  selection is the selected side's (or the tie average), and the
  difference is carried along by the flow (up to 0.4 % of the tangent
  scale here, at 30-50 % of cells for 4 and 8 zones);
- bit for bit, the GPU kernels against this oracle (tests/test_gpu_parity.py).
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleSim, dualn_probe, set_dualn_div_mode, set_pow_mode, set_threads


@pytest.fixture(autouse=True)
def _mode():
    set_threads(8)
    set_pow_mode(0)
    yield


def test_components_match_independent_scalar_passes():
    val, tan, scalar = dualn_probe(1.7, 0.035, -0.42)
    for k, (sv, sd) in enumerate(scalar):
        assert abs(val - sv) < 1e-14
        assert abs(tan[k] - sd) < 1e-12, (k, tan[k], sd)
    assert np.all(np.isfinite(tan)) and np.any(tan != 0.0)


def _zonal_case(n_params, n=32, steps=60, dt=0.01, dual=None, seed_zone=None):
    """m1_forward_scaling.rs: flat bed, h = 1 left half / 0.1 right half,
    transmissive edges, Manning 0.03 in n_params column zones, SSP-RK2 +
    friction(1e-6) with a fixed dt; objective = total volume."""
    h = np.where(np.arange(n) < n // 2, 1.0, 0.1)[None, :].repeat(n, 0)
    st = np.zeros((n, n, 3))
    st[..., 0] = h
    st[..., 1] = 0.05 * np.sin(np.arange(n))[:, None]
    zone = (np.arange(n) * n_params) // n
    kw = {}
    if dual == "n":
        sim = OracleSim(n, n, 1.0, 1.0, tangents=n_params)
        mt = np.zeros((n, n, n_params))
        for k in range(n_params):
            mt[:, zone == k, k] = 1.0
        kw = dict(manning_tangent=mt)
    elif dual == "1":
        sim = OracleSim(n, n, 1.0, 1.0, dual=True)
        kw = dict(manning_tangent=np.where(zone == seed_zone, 1.0, 0.0)[None, :].repeat(n, 0))
    else:
        sim = OracleSim(n, n, 1.0, 1.0)
    sim.set_mesh(np.zeros((n, n)), np.full((n, n), 0.03), **kw)
    sim.set_state(st)
    sim.set_boundaries([0, 0, 0, 0])
    for _ in range(steps):
        sim.ssprk2_step(dt)
        sim.manning_friction_step(dt, 1e-6)
    return sim.get_state()


@pytest.mark.parametrize("n_params", [2, 4, 8, 16])
def test_dualn_tangents_equal_seeded_dual_runs_on_the_f64_value_path(n_params):
    """With the quotient value a / b (the f64 path), DualN values are the f64
    run's bit for bit and each tangent is the seeded Dual run's to rounding:
    this pins every tangent rule of DualN against Dual's."""
    set_dualn_div_mode(1)
    try:
        v64 = _zonal_case(n_params)
        vn, tn = _zonal_case(n_params, dual="n")
        assert np.array_equal(vn, v64)
        for k in range(n_params):
            vd, td = _zonal_case(n_params, dual="1", seed_zone=k)
            assert np.max(np.abs(tn[..., k] - td)) <= 1e-13 * np.max(np.abs(td)), k
    finally:
        set_dualn_div_mode(0)


@pytest.mark.parametrize("n_params", [2, 4, 8, 16])
def test_dualn_reference_value_path(n_params):
    """The reference's DualN quotient a * (1 / b) moves values by the last
    bit only; the tangents then follow the branches that value path takes
    (see the module docstring), so they agree with the seeded Dual runs in
    aggregate, not cell by cell."""
    v64 = _zonal_case(n_params)
    vn, tn = _zonal_case(n_params, dual="n")
    assert np.max(np.abs(vn - v64)) <= 1e-12 * max(1.0, np.max(np.abs(v64)))
    vd, td_all = _zonal_case(1, dual="1", seed_zone=0)
    rel = np.linalg.norm(tn.sum(axis=-1) - td_all) / np.linalg.norm(td_all)
    assert rel <= 1e-2, rel
