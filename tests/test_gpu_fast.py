"""The fast FP64 kind (HF2D_F64_FAST) against the bitwise FP64 kind on the GPU.

The fast kind computes the same step with FMA contraction, reciprocal-and-
Newton division and Newton square root (about one ulp per operation), so it
is held to tolerances, not bits:

- the 256 x 256 dam break (flat frictionless bed, walls, CFL 0.5): depth
  and discharge within 1e-10 of the bitwise run after 10, 100 and 500 steps,
  mass conserved to 1e-12;
- lake at rest over a bumpy bed with an emerged island: the surface and the
  discharge stay at rest to 1e-13 over 200 steps;
- a smooth wet run with friction, rain, an inflow and a sink through the
  fused driver: within 1e-10 m and 1e-10 m^2/s of the bitwise run after 50
  steps, with the same clock to 1e-12 s.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _dam(kind):
    from hydroflux_amd import Grid2D

    n = 256
    g = Grid2D(n, n, 0.1, 0.1, kind=kind)
    st = np.zeros((n, n, 3))
    st[..., 0] = np.where((np.arange(n) + 0.5) * 0.1 < 12.8, 1.0, 0.1)[None, :]
    g.set_mesh(np.zeros((n, n)), np.zeros((n, n)))
    g.set_state(st)
    g.set_boundaries([1, 1, 1, 1])
    g.configure(cfl=0.5, max_dt=1e9, friction=False)
    return g


def test_fast_dam_break_within_1e10_of_bitwise():
    _say("fast f64: dam break 256^2, 500 steps")
    a, b = _dam("f64"), _dam("f64fast")
    v0 = b.total_volume()
    done = 0
    for n in (10, 100, 500):
        a.run_steps(n - done)
        b.run_steps(n - done)
        done = n
        sa, sb = a.get_state(), b.get_state()
        dh = float(np.max(np.abs(sa - sb)[..., 0]))
        dq = float(np.max(np.abs(sa - sb)[..., 1:]))
        _say(f"  {n} steps: max |dh| {dh:.2e}, max |dq| {dq:.2e}")
        assert dh <= 1e-10 and dq <= 1e-10
        assert abs(a.time()[0] - b.time()[0]) <= 1e-12
    assert abs(b.total_volume() - v0) <= 1e-12 * v0
    a.close()
    b.close()


def test_fast_lake_at_rest():
    from hydroflux_amd import Grid2D

    _say("fast f64: lake at rest, 200 steps")
    n = 96
    y, x = np.mgrid[0:n, 0:n]
    bed = 0.4 * np.exp(-((x - 40.0) ** 2 + (y - 50.0) ** 2) / 60.0) + 0.1 * np.sin(x / 7.0) * np.cos(y / 5.0)
    bed[20:26, 60:66] = 1.5  # an island above the surface
    st = np.zeros((n, n, 3))
    st[..., 0] = np.maximum(1.0 - bed, 0.0)
    g = Grid2D(n, n, 1.0, 1.0, kind="f64fast")
    g.set_mesh(bed, np.full((n, n), 0.03))
    g.set_state(st)
    g.set_boundaries([1, 1, 1, 1])
    g.configure(cfl=0.4, max_dt=1e9, friction=True)
    g.run_steps(200)
    out = g.get_state()
    wet = st[..., 0] > 0
    eta = out[..., 0] + bed
    assert np.max(np.abs(eta[wet] - 1.0)) <= 1e-13
    assert np.max(np.abs(out[..., 1:])) <= 1e-13
    assert np.all(out[~wet, 0] == 0.0)
    g.close()


def test_fast_forced_run_within_1e10_of_bitwise():
    from hydroflux_amd import Grid2D

    _say("fast f64: smooth forced run, 50 driver steps")
    n = 160
    y, x = np.mgrid[0:n, 0:n]
    bed = 0.002 * x + 0.05 * np.sin(x / 9.0) * np.cos(y / 11.0)
    st = np.zeros((n, n, 3))
    st[..., 0] = 0.6 + 0.05 * np.sin(x / 13.0 + 0.4) - bed
    man = np.where((x // 40 + y // 40) % 2 == 0, 0.03, 0.06)
    runs = {}
    for kind in ("f64", "f64fast"):
        g = Grid2D(n, n, 2.0, 2.0, kind=kind)
        g.set_mesh(bed, man)
        g.set_state(st)
        g.set_boundaries([1, 1, 1, 0])
        g.configure(cfl=0.4, max_dt=0.5, friction=True, rain_rate=50e-3 / 3600.0,
                    sources=[(80, 2, 1.5), (40, 120, -0.2)])
        g.run_steps(50)
        runs[kind] = (g.time(), g.get_state())
        g.close()
    (ta, sa), (tb, sb) = runs["f64"], runs["f64fast"]
    assert ta[1] == tb[1] == 50 and abs(ta[0] - tb[0]) <= 1e-12
    dh, dq = float(np.max(np.abs(sa - sb)[..., 0])), float(np.max(np.abs(sa - sb)[..., 1:]))
    _say(f"  max |dh| {dh:.2e}, max |dq| {dq:.2e}")
    assert dh <= 1e-10 and dq <= 1e-10
