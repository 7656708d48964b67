"""The benchmark workloads at their stated sizes, on the GPU.

- 1024 x 1024 flood DEM (fBm, 50 m relief, slope, carved valley; 4 Manning
  classes; rain 50 mm/h; 100 m^3/s inflow; walls N/S/W, open E) in f64 and
  in Dual seeded on a global Manning multiplier: the fused driver for 200
  steps, bit for bit with the oracle (its OpenMP rows are the serial
  restatement's bits), clock and step count included.
"""
import sys

import numpy as np
import pytest

from oracle.pyoracle import OracleSim, set_pow_mode, set_threads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _mode():
    set_threads(16)
    set_pow_mode(1)
    yield
    set_pow_mode(0)


def _load(sim, sc, dual):
    sim.set_mesh(sc.bed, sc.manning, manning_tangent=(sc.manning if dual else None))
    if dual:
        sim.set_state(sc.dry_state(), np.zeros_like(sc.dry_state()))
    else:
        sim.set_state(sc.dry_state())
    sim.set_boundaries(list(sc.bc_kinds))
    sim.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6, rain_rate=sc.rain_rate,
                  sources=sc.sources)
    return sim


@pytest.mark.parametrize("kind", ["f64", "dual"])
def test_flood_1024_bitwise(kind):
    from hydroflux_amd import Grid2D
    from hydroflux_amd.scenarios import flood_scenario

    _say(f"flood 1024x1024 {kind}, 200 driver steps")
    sc = flood_scenario(n=1024, dx=5.0)
    dual = kind == "dual"
    g = _load(Grid2D(sc.n_rows, sc.n_cols, sc.dx, sc.dx, kind=kind), sc, dual)
    o = _load(OracleSim(sc.n_rows, sc.n_cols, sc.dx, sc.dx, dual=dual), sc, dual)
    g.run_steps(200)
    o.run_steps(200)
    assert g.time() == o.time()
    assert g.time()[1] == 200
    if dual:
        (gv, gt), (ov, ot) = g.get_state(), o.get_state()
        assert np.array_equal(gv, ov)
        assert np.array_equal(gt, ot, equal_nan=True)
        assert np.count_nonzero(gt[..., 0]) > 0.1 * gt[..., 0].size  # the tangent is carried through the run
    else:
        gv = g.get_state()
        assert np.array_equal(gv, o.get_state())
    assert np.count_nonzero(gv[..., 0] > 1e-6) > 0.9 * gv[..., 0].size  # rain wets the domain
    g.close()
