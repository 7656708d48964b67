"""Dual tangents of the GPU step against finite differences of GPU f64 runs.

The flood construction at 256 x 256 (seeded fBm DEM with slope and carved
valley, Manning n from 4 land-cover classes in Voronoi patches, 50 mm/h rain,
100 m^3/s inflow, walls N/S/W, open E) with friction, warm-started at 0.5 m
of water. A hand loop over the base run's dt sequence, as the reference's
own derivative checks are written (update.rs:1953-2038, source.rs:421-473):
SSP-RK2, friction, rain, point sources per step. Dual seeds: a global
Manning multiplier (n = m n0) and the Manning n of one land-cover class. The
tangents after 100 steps equal (f(n0 + e s) - f(n0 - e s)) / 2e of two f64
GPU runs.

The DEM here has 2 m of relief. At the flood workload's 50 m the run is not
differentiable at the base point: the two perturbed runs differ by ~4 cm in
some cells whatever e is (1e-5 or 1e-7), so no finite difference converges
there; at 2 m the difference shrinks in proportion to e and the central
difference agrees with the tangent to ~1e-7.
"""
import sys

import numpy as np
import pytest

from oracle.pyoracle import set_pow_mode

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _mode():
    set_pow_mode(1)
    yield
    set_pow_mode(0)


@pytest.mark.parametrize("seed_kind", ["global_multiplier", "land_class_2"])
def test_dual_tangents_equal_central_differences_on_gpu(seed_kind):
    from hydroflux_amd import Grid2D
    from hydroflux_amd.scenarios import flood_scenario

    _say(f"Dual vs central differences, flood 256x256 relief 2 m, seed {seed_kind}")
    n, steps, eps = 256, 100, 1e-7
    sc = flood_scenario(n=n, dx=5.0, relief=2.0)
    st0 = sc.dry_state()
    st0[..., 0] = 0.5
    seed = sc.manning if seed_kind == "global_multiplier" else np.where(sc.land_class == 2, sc.manning, 0.0)

    def grid(kind, man):
        g = Grid2D(n, n, sc.dx, sc.dx, kind=kind)
        g.set_mesh(sc.bed, man, manning_tangent=(seed if kind == "dual" else None))
        if kind == "dual":
            g.set_state(st0, np.zeros_like(st0))
        else:
            g.set_state(st0)
        g.set_boundaries(list(sc.bc_kinds))
        return g

    base = grid("f64", sc.manning)
    base.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6, rain_rate=sc.rain_rate,
                   sources=sc.sources)
    dts = [base.step() for _ in range(steps)]
    driver_state = base.get_state()
    base.close()

    def loop(g):
        for dt in dts:
            g.ssprk2_step(dt)
            g.manning_friction_step(dt, 1e-6)
            g.apply_rain(sc.rain_rate, dt)
            g.apply_point_sources(sc.sources, dt)
        out = g.get_state()
        g.close()
        return out

    val, tan = loop(grid("dual", sc.manning))
    ref = loop(grid("f64", sc.manning))
    assert np.array_equal(ref, driver_state)  # the hand loop is the driver's step
    assert np.array_equal(val, ref)  # Dual values are the f64 run's bits
    plus = loop(grid("f64", sc.manning + eps * seed))
    minus = loop(grid("f64", sc.manning - eps * seed))
    fd = (plus - minus) / (2 * eps)
    for f in range(3):
        a, g = tan[..., f], fd[..., f]
        assert np.all(np.isfinite(a))
        l2 = np.linalg.norm(a - g) / np.linalg.norm(a)
        mx = np.max(np.abs(a - g)) / np.max(np.abs(a))
        _say(f"  field {f}: L2 rel {l2:.2e}, max rel {mx:.2e}")
        assert l2 < 1e-5 and mx < 1e-4, (f, l2, mx)
