"""GPU: the FP32 precision form against the FP64 step.

The FP32 form (FP32 state and fluxes, an FP64 bed datum per 64 x 64 block,
FP64 clock) has no bitwise counterpart. It is held to the accuracy it has on
the flood workload (fBm terrain 2000 m above datum, 50 mm/h rain, 100 m3/s
inflow, Manning friction): after the FP64 run's 600 steps, the FP32 run
advanced to the same simulated time must match its depths, wet extent and
water volume within fixed bounds. A wrong face, slope or forcing term moves
these by orders of magnitude more than FP32 rounding does.
"""
import sys

import numpy as np
import pytest

from hydroflux_amd.metrics import depth_metrics
from hydroflux_amd.scenarios import flood_scenario

pytestmark = pytest.mark.gpu


def _grid(sc, kind):
    from hydroflux_amd import Grid2D

    g = Grid2D(sc.n_rows, sc.n_cols, sc.dx, sc.dx, kind=kind)
    g.set_mesh(sc.bed, sc.manning)
    g.set_state(sc.dry_state())
    g.set_boundaries(list(sc.bc_kinds))
    g.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6, rain_rate=sc.rain_rate,
                sources=sc.sources)
    return g


@pytest.mark.parametrize("n", [384, 200])
def test_f32_form_tracks_f64_on_the_flood_workload(n):
    print(f"[gpu-case] FP32 vs FP64 flood {n}", flush=True)
    sc = flood_scenario(n=n, dx=5.0, offset=2000.0)
    g64 = _grid(sc, "f64")
    g64.run_steps(600)
    t, _ = g64.time()
    h64, v64 = g64.get_state()[..., 0], g64.total_volume()
    g64.close()
    g32 = _grid(sc, "f32")
    g32.run_until(t)
    t32, _ = g32.time()
    m = depth_metrics(g32.get_state()[..., 0], h64, sc.dx, sc.dx, vol_test=g32.total_volume(), vol_ref=v64)
    g32.close()
    print(f"[gpu-case] t {t} / {t32}: {m}", flush=True)
    sys.stdout.flush()
    assert abs(t32 - t) <= 1e-9 * t
    assert m["wet_cells_ref"] > 1000
    # measured on MI355X: L1 6.3e-5 / 1.2e-4 m, peak 0.005 / 0.015 m,
    # IoU 0.925 / 0.952, volume 1.0e-5 / 2.3e-5 (n = 384 / 200)
    assert m["depth_L1_m"] <= 5e-4
    assert m["peak_depth_err_m"] <= 5e-2
    assert m["extent_iou_1cm"] >= 0.85
    assert m["mass_balance_rel_err"] <= 1e-4
