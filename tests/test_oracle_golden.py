"""Pins the CPU restatement (oracle) to the reference's own published vectors.

Each test rebuilds the reference example that wrote a CSV under
papers/01_review/figures/data (tests/golden) and requires the oracle to
reproduce every printed digit. These are the first acceptance gate: the GPU
path is checked against the oracle, so the oracle is checked against the
reference's data.
"""
import math
import os

import numpy as np
import pytest

import cases
from oracle.pyoracle import OracleSim, set_pow_mode, set_threads


@pytest.fixture(autouse=True)
def _oracle_mode():
    set_threads(8)
    set_pow_mode(0)  # glibc pow, as the Rust reference
    yield
    set_pow_mode(0)


def _csv(golden_dir, name):
    return np.loadtxt(os.path.join(golden_dir, name), delimiter=",", skiprows=1)


def _run(case, **kw):
    sim = case.load(OracleSim(case.nr, case.nc, case.dx, case.dy))
    t, k = cases.hand_loop(sim, case.extra["t_end"], **kw)
    return sim, t, k


@pytest.mark.parametrize("n_cols,fname", [(400, "verif_stoker.csv"), (100, "verif_stoker_coarse.csv")])
def test_stoker_profile_matches_reference_csv(golden_dir, n_cols, fname):
    # gen_verification_data.rs:46-82 / gen_stoker_coarse.rs:36-75
    sim, t, _ = _run(cases.stoker(n_cols))
    h = sim.get_state()[1, :, 0]
    ref = _csv(golden_dir, fname)
    assert np.array_equal(np.round(h, 6), ref[:, 2])
    # the analytical column is the Stoker solution (dam_break_on_dry.rs:57-82)
    an = [cases.stoker_depth(x, 4.0) for x in ref[:, 0]]
    assert np.allclose(an, ref[:, 1], atol=5e-7)


def test_thacker_profile_and_step_count(golden_dir):
    # gen_verification_data.rs:130-170; 388 steps (manuscript.md:504)
    sim, t, k = _run(cases.thacker(80))
    assert k == 388
    h = sim.get_state()[40, :, 0]
    assert np.array_equal(np.round(h, 6), _csv(golden_dir, "verif_thacker.csv")[:, 2])


def test_macdonald_discharge_depth_friction(golden_dir):
    # gen_verification_data.rs:86-118: Discharge in, Depth out, Manning friction
    c = cases.macdonald()
    sim, _, _ = _run(c, friction_tol=1e-9)
    ref = _csv(golden_dir, "verif_macdonald.csv")
    assert np.array_equal(np.round(sim.get_state()[2, :, 0], 6), ref[:, 2])
    assert round(c.extra["h_n"], 6) == ref[0, 1]


def test_uk_ea_t6_full_field(golden_dir):
    # gen_verification_data.rs:196-253: full 250x50 depth field, 5 decimals
    sim, _, k = _run(cases.uk_ea_t6(), friction_tol=1e-9)
    ref = _csv(golden_dir, "verif_uk_ea_t6.csv")
    h = sim.get_state()[..., 0].reshape(-1)
    assert np.array_equal(np.round(h, 5), ref[:, 2])


def test_stoker_400x10_step_count(golden_dir):
    # m2_hydroflux_wallclock.rs: 400x10, dx 0.5, t_end 8 s -> 343 steps
    nr, nc, dx = 10, 400, 0.5
    st = np.zeros((nr, nc, 3))
    st[:, (np.arange(nc) + 0.5) * dx < 100.0, 0] = 1.0
    sim = OracleSim(nr, nc, dx, dx)
    sim.set_mesh(np.zeros((nr, nc)), 0.0)
    sim.set_state(st)
    sim.set_boundaries([1, 1, 0, 0])
    _, k = cases.hand_loop(sim, 8.0)
    ref = np.genfromtxt(os.path.join(golden_dir, "m2_anuga_wallclock.csv"), delimiter=",", names=True,
                        dtype=None, encoding="utf-8")
    assert k == int(ref["n_steps"][0]) == 343


@pytest.mark.parametrize("n", [32, 64])
def test_thacker_convergence_steps_and_errors(golden_dir, n):
    # gen_convergence.rs:42-90: exact step counts and relative L1/L2 (4 digits)
    ref = _csv(golden_dir, "convergence_thacker.csv")
    row = ref[ref[:, 0] == n][0]
    h0, a, b = 0.1, 1.0, 0.1
    half = 1.25
    dx = 2.0 * half / n
    omega = math.sqrt(2.0 * cases.G * h0) / a
    t_end = math.pi / omega
    x = -half + (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    bed = h0 * ((X * X + Y * Y) / (a * a) - 1.0)

    def th(t):
        xc, yc = b * math.cos(omega * t), b * math.sin(omega * t)
        return np.maximum(h0 / (a * a) * (a * a - ((X - xc) ** 2 + (Y - yc) ** 2)), 0.0)

    h = th(0.0)
    st = np.stack([h, np.zeros_like(h), h * (b * omega)], axis=-1)
    sim = OracleSim(n, n, dx, dx)
    sim.set_mesh(bed, 0.0)
    sim.set_state(st)
    sim.set_boundaries([1, 1, 1, 1])
    _, k = cases.hand_loop(sim, t_end)
    assert k == int(row[4])
    h_an = th(t_end)
    e = np.abs(sim.get_state()[..., 0] - h_an)
    l1 = np.sum(e) / np.sum(h_an)
    l2 = math.sqrt(np.sum(e * e) / np.sum(h_an * h_an))
    assert abs(l1 - row[2]) / row[2] < 5e-6
    assert abs(l2 - row[3]) / row[3] < 5e-6


def _centroid_dual(n_val, dual=True):
    c = cases.inverse_manning_channel()
    sim = OracleSim(c.nr, c.nc, c.dx, c.dy, dual=dual)
    sim.set_mesh(c.bed, n_val, manning_tangent=(1.0 if dual else None))
    sim.set_state(c.state)
    sim.set_boundaries(c.bc_kinds)
    cases.hand_loop(sim, 20.0, friction_tol=1e-6, max_steps=5000)
    out = sim.get_state()
    h = out[0][1, :, 0] if dual else out[1, :, 0]
    x = np.arange(c.nc) + 0.5
    num = den = 0.0
    dnum = dden = 0.0
    ht = out[1][1, :, 0] if dual else np.zeros_like(h)
    for j in range(c.nc):  # m1_inverse_manning_demo.rs:88-100, Dual arithmetic in order
        num, dnum = num + h[j] * x[j], dnum + ht[j] * x[j]
        den, dden = den + h[j], dden + ht[j]
    return num / den, (dnum * den - num * dden) / (den * den)


def test_dual_inverse_manning_newton_trace(golden_dir):
    # m1_inverse_manning_demo.rs:42-108: the only Dual vector the reference holds.
    ref = _csv(golden_dir, "m1_inverse_manning.csv")
    c_target, _ = _centroid_dual(0.04, dual=False)
    n = 0.08
    for it in range(ref.shape[0]):
        cv, dc = _centroid_dual(n)
        err = cv - c_target
        step = -err / dc if abs(dc) > 1e-12 else 0.0
        assert f"{n:.10f}" == f"{ref[it, 1]:.10f}"
        assert f"{cv:.10f}" == f"{ref[it, 2]:.10f}"
        assert float(f"{err * err:.10e}") == ref[it, 3]
        assert float(f"{dc:.10e}") == ref[it, 4]
        n = min(max(n + step, 1e-4), 1.0)
