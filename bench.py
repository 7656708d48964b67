"""Benchmark: SSP-RK2 flood step on one MI355X — cell-updates/s and HBM
roofline for FP64 and the FP32 precision form, the 256-scenario batch, and
the continuous-batching serving loop's p50/p99 latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 4096] [--no-cpu]
                    [--quick] [--sections ...] [--acc-steps 20000] [--serve-requests 64]

Headline (`value`): the rain-on-grid flood construction — seeded fBm DEM
(H 0.8, 50 m relief) + 0.01 slope + 8 m carved valley, +2000 m bed offset,
Manning n from 4 land-cover classes in Voronoi patches, initially dry,
50 mm/h rain, 100 m^3/s point inflow, walls N/S/W and open E — at
4096 x 4096 cells, dx 5 m, in FP64: the largest single-GPU plain-FP64
configuration, with a footprint (1.1 GB) well beyond the 256 MB Infinity
Cache, so HBM bandwidth is what the roofline fraction is taken against.
Each step is one SSP-RK2 step with Manning friction, rain and the point
source, dt from the device-side CFL reduction (no host round trip), after W
warm-up steps that wet the domain. Inputs are resident in HBM before the
timed region.

`roofline` is the final stage kernel's: algorithmic HBM bytes over its
HIP-event launch time against 8 TB/s, and its VALU issue (lane-instructions
per launch from the committed SQ counters over the same launch time) against
the vector issue peak; `bound` names the larger fraction. The streaming copy
probe's rate is the attainable HBM ceiling beside the 8 TB/s peak.

Beside it in the same JSON line (unless --quick):
  fast_f64        the fast FP64 kind (contraction, reciprocal/Newton division
                  and square root) on the same workload, and its drift from
                  the bitwise kind on the 256^2 dam break;
  precision_form  the same 4096^2 workload in the FP32 precision form
                  (throughput, roofline), and its depth accuracy against the
                  FP64 run at the same simulated times, at 100, 1000, 5000 and
                  --acc-steps FP64 steps, next to the FP64 run's own
                  sensitivity to a 1e-9 m bed perturbation (the conditioning
                  floor of the workload);
  c1_1024         the 1024^2 workload (no offset) in FP64 and Dual;
  batch           256 scenarios of mixed size (sides log-uniform in
                  [128, 2048], aspect U(1, 4)), each with its own DEM, land
                  cover, rain and dt: FP32 form and FP64, autotuned tile
                  shape, against serial per-scenario FP64 launches;
  serving         continuous batching of requests drawn from that mix (FP32
                  arena holding about a third of it) at a sweep of Poisson
                  arrival rates up to saturation: p50/p99 latency-to-completion
                  and aggregate cell-updates/s; and the mix with 25 % Dual
                  requests (tangents on one land-cover class's Manning n);
  dry_domain      the compacted active-tile list on the 4096^2 workload
                  without rain (inflow only), on and off;
  uk_ea           UK EA Tests 4 and 8A on their official geometries, full
                  18000 s runs, wall time;
  cpu_baseline    the C++ restatement of the reference step, 1 thread pinned
                  to one core: the 256^2 dam break in full, the 1024^2 flood
                  over 3 + 50 steps, and its Dual cost ratio.

One JSON line on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s 2D SSP-RK2 (FP64 vs FP32) + %HBM peak; batched p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell for one launch (each field read or written once):
#   FP64 first stage : read h, hu, hv, z; write U(1)                 = 7 x 8  = 56
#   FP64 final stage : read U(1), z, U^n, n; write U^{n+1}          = 11 x 8 = 88
#   FP32 first stage : read h, hu, hv, z deviation; write U(1)       = 7 x 4  = 28
#   FP32 final stage : read U(1), z dev., U^n, n; write U^{n+1}     = 11 x 4 = 44
ALG_BYTES = {
    "f64": {"stage1": 56.0, "final": 88.0, "step": 144.0},
    "f32": {"stage1": 28.0, "final": 44.0, "step": 72.0},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_grid(sc, kind="f64", bed=None):
    from hydroflux_amd import Grid2D

    g = Grid2D(sc.n_rows, sc.n_cols, sc.dx, sc.dx, kind=kind)
    g.set_mesh(sc.bed if bed is None else bed, sc.manning,
               manning_tangent=(sc.manning if kind == "dual" else None))
    g.set_state(sc.dry_state())
    g.set_boundaries(list(sc.bc_kinds))
    g.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6, rain_rate=sc.rain_rate,
                sources=sc.sources)
    return g


def time_steps(g, warmup, steps):
    """Wall time of `steps` enqueued steps after `warmup` steps, and the
    per-launch averages of the two stage kernels from HIP events recorded
    on the grid's own stream."""
    g.run_steps(warmup)
    g.synchronize()
    g.set_kernel_timing(True)
    t0 = time.perf_counter()
    g.enqueue_steps(steps)
    g.synchronize()
    t1 = time.perf_counter()
    s1, s2, cnt = g.kernel_timing()
    g.set_kernel_timing(False)
    return (t1 - t0), s1 / max(cnt, 1), s2 / max(cnt, 1)


# Vector issue peaks of MI355X from its published vector rates, one FMA
# (2 flops) per lane-instruction: FP64 78.6 TFLOP/s -> 39.3 T lane-ops/s
# (256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz); FP32 157.3 TFLOP/s -> 78.6 T.
# The FP32 form's kernels issue above 39.3 T lane-instructions/s (no packed
# FP32 instructions are emitted), so the FP64 figure does not bound them.
VALU_PEAK_LANE_OPS = {"f64": 78.6e12 / 2, "f32": 157.3e12 / 2}


def _counters(key):
    """VALU wave-instructions per launch of a stage kernel from the committed
    rocprofv3 SQ counter summary (profiles/sq_counters.json)."""
    try:
        with open(os.path.join(ROOT, "profiles", "sq_counters.json")) as f:
            return json.load(f)["kernels"].get(key)
    except (OSError, ValueError, KeyError):
        return None


def roofline(kind, cells, t_first_ms, t_final_ms, traffic, kernel, counter_key=None):
    """HBM roofline of the final stage (algorithmic bytes / launch time vs
    8 TB/s) and its VALU issue (counted lane-instructions / launch time vs
    the vector issue peak); `bound` names the larger of the two fractions."""
    ab = ALG_BYTES["f32" if kind == "f32" else "f64"]
    achieved = ab["final"] * cells / (t_final_ms / 1e3) / 1e9
    out = {
        "bound": "hbm",
        "kernel": kernel,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "alg_bytes_per_cell": ab["final"],
        "cells_per_launch": cells,
        "avg_launch_ms": t_final_ms,
        "first_stage": {"alg_bytes_per_cell": ab["stage1"], "avg_launch_ms": t_first_ms,
                        "achieved": ab["stage1"] * cells / (t_first_ms / 1e3) / 1e9},
    }
    c = _counters(counter_key) if counter_key else None
    if c:
        lane = c["valu_wave_instr_per_launch"] * 64.0
        issue = lane / (t_final_ms / 1e3)
        peak = VALU_PEAK_LANE_OPS["f32" if kind == "f32" else "f64"]
        out["valu"] = {"lane_instr_per_cell": lane / cells, "achieved_lane_ops_per_s": issue,
                       "peak_lane_ops_per_s": peak, "issue_frac": issue / peak,
                       "source": "profiles/sq_counters.json (SQ_INSTS_VALU per launch) / the live launch time"}
        if out["valu"]["issue_frac"] > out["frac"]:
            out["bound"] = "valu"
    return out


def cpu_baseline():
    """The CPU restatement of the reference step (the oracle), single-threaded
    and pinned to one host core, as the reference runs by default (its rayon
    feature is off): the 256 x 256 dam break in full (500 steps, CFL 0.5,
    walls, cfl_time_step each step), the 1024 x 1024 flood workload over 3
    warm-up + 50 timed steps (solver-2d/examples/m2_perf_large_grid.rs:52-73)
    and the same in Dual over 3 + 10 steps for the f64/Dual cost ratio
    (m1_timing_f64_vs_dual.rs). The reference's published single-thread
    figures (282-294 ns per cell-step, i5-13420H, papers/01_review/
    manuscript.md:684-685; Dual/f64 2.01x, manuscript.md:699) are given beside
    it; they come from other hardware."""
    from hydroflux_amd.scenarios import flood_scenario
    from oracle.pyoracle import OracleSim, set_pow_mode, set_threads

    allowed = sorted(os.sched_getaffinity(0))
    core = allowed[0]
    os.sched_setaffinity(0, {core})
    try:
        set_threads(1)
        set_pow_mode(0)
        # 256^2 dam break, 500 steps
        n = 256
        o = OracleSim(n, n, 0.1, 0.1)
        st = np.zeros((n, n, 3))
        st[..., 0] = np.where((np.arange(n) + 0.5) * 0.1 < 12.8, 1.0, 0.1)[None, :]
        o.set_mesh(np.zeros((n, n)), np.zeros((n, n)))
        o.set_state(st)
        o.set_boundaries([1, 1, 1, 1])
        o.configure(cfl=0.5, cfl_with_bcs=False, max_dt=1e30, friction=False)
        t0 = time.perf_counter()
        o.run_steps(500)
        w0 = time.perf_counter() - t0
        o.close()
        # 1024^2 flood, f64 and Dual
        sc = flood_scenario(n=1024, dx=5.0)
        rates = {}
        for kind, steps in (("f64", 50), ("dual", 10)):
            o = OracleSim(sc.n_rows, sc.n_cols, sc.dx, sc.dx, dual=(kind == "dual"))
            o.set_mesh(sc.bed, sc.manning, manning_tangent=(sc.manning if kind == "dual" else None))
            if kind == "dual":
                o.set_state(sc.dry_state(), np.zeros_like(sc.dry_state()))
            else:
                o.set_state(sc.dry_state())
            o.set_boundaries(list(sc.bc_kinds))
            o.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6,
                        rain_rate=sc.rain_rate, sources=sc.sources)
            o.run_steps(3)
            t0 = time.perf_counter()
            o.run_steps(steps)
            rates[kind] = (time.perf_counter() - t0) / (sc.n_cells * steps)
            o.close()
    finally:
        os.sched_setaffinity(0, set(allowed))
    return {
        "value": 1.0 / rates["f64"],
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": "C++ restatement of the reference step (the oracle), 1 thread pinned to one core: flood DEM "
                  "1024x1024 (rain, 100 m3/s inflow, friction), 3 warm-up + 50 timed SSP-RK2 steps",
        "ns_per_cell_step": rates["f64"] * 1e9,
        "dam_break_256_500_steps": {"wall_s": w0, "ns_per_cell_step": w0 / (n * n * 500) * 1e9},
        "dual_ns_per_cell_step": rates["dual"] * 1e9,
        "dual_over_f64": rates["dual"] / rates["f64"],
        "published_reference": {"ns_per_cell_step": [282, 294], "dual_over_f64": 2.01,
                                "hardware": "Intel i5-13420H, 1 thread",
                                "source": "papers/01_review/manuscript.md:684-685, 699"},
        "pinned_core": core,
        "host_cpu": _cpu_model(),
        "host_cores": len(allowed),
    }


def _cpu_model():
    """The host CPU model as lscpu names it, with its family / model /
    stepping numbers and clock where lscpu gives them (a virtualised host
    may name only the product line); else /proc/cpuinfo's model name."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                f[k.strip()] = v.strip()
        if "Model name" in f:
            extra = [f"{k.lower()} {f[k]}" for k in ("CPU family", "Model", "Stepping", "CPU max MHz", "BogoMIPS")
                     if f.get(k)]
            return f["Model name"] + (f" ({', '.join(extra)})" if extra else "")
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _traffic(key):
    """HBM bytes per launch of a stage kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated)."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(key)
    except (OSError, ValueError):
        return None


def precision_form(sc, g64, args):
    """FP32 form on the headline workload: throughput, roofline, and depth
    accuracy against the FP64 grid `g64` (already advanced warmup + steps)
    at FP64 step counts 100, 1000, 5000 and --acc-steps: the FP32 form, the
    fast FP64 kind and an FP64 run on the bed perturbed by 1e-9 m
    (N(0, 1) * 1e-9, the workload's own conditioning) are each run to the
    FP64 run's simulated time at every checkpoint and compared with it
    there."""
    from hydroflux_amd.metrics import depth_metrics

    cells = sc.n_cells
    log("[bench] FP32 precision form: timing")
    g32 = build_grid(sc, "f32")
    wall, t1, t2 = time_steps(g32, args.warmup, args.steps)
    out = {
        "value": cells * args.steps / wall,
        "ms_per_step": wall / args.steps * 1e3,
        "dtype": "f32 state/flux, f64 bed datum per 64x64 block, f64 dt/time/mass",
        "roofline": roofline("f32", cells, t1, t2, _traffic("f32_final_stage_bytes_per_launch"),
                             "stage_kernel<float, final>", "f32_final"),
    }
    g32.close()
    if args.acc_steps > 0:
        done = g64.time()[1]
        marks = sorted(set(m for m in (100, 1000, 5000, args.acc_steps) if done <= m <= args.acc_steps))
        log(f"[bench] accuracy: FP64 to {marks} steps, FP32 and perturbed FP64 to the same times")
        g32 = build_grid(sc, "f32")
        gf = build_grid(sc, "f64fast")
        rng = np.random.default_rng(12345)
        gp = build_grid(sc, "f64", bed=sc.bed + 1e-9 * rng.standard_normal(sc.bed.shape))
        curve = []
        for m in marks:
            g64.run_steps(m - g64.time()[1])
            t_ref, n_ref = g64.time()
            s64 = g64.get_state()[..., 0]
            v64 = g64.total_volume()
            log(f"[bench]   FP64 at {n_ref} steps")
            g32.run_until(t_ref)
            log(f"[bench]   FP32 at t = {t_ref:.0f} s")
            gp.run_until(t_ref)
            gf.run_until(t_ref)
            row = {"fp64_steps": n_ref, "t_simulated_s": t_ref, "fp32_steps": g32.time()[1],
                   "fp32_vs_fp64": depth_metrics(g32.get_state()[..., 0], s64, sc.dx, sc.dx,
                                                 vol_test=g32.total_volume(), vol_ref=v64),
                   "fast_fp64_steps": gf.time()[1],
                   "fast_fp64_vs_fp64": depth_metrics(gf.get_state()[..., 0], s64, sc.dx, sc.dx,
                                                      vol_test=gf.total_volume(), vol_ref=v64),
                   "fp64_bed_perturbed_1e-9m_vs_fp64": depth_metrics(gp.get_state()[..., 0], s64, sc.dx, sc.dx,
                                                                     vol_test=gp.total_volume(), vol_ref=v64)}
            curve.append(row)
            log(f"[bench]   {n_ref} steps, t = {t_ref:.0f} s: FP32 Linf {row['fp32_vs_fp64']['depth_Linf_m']:.3g} m, "
                f"fast FP64 Linf {row['fast_fp64_vs_fp64']['depth_Linf_m']:.3g} m, "
                f"perturbed FP64 Linf {row['fp64_bed_perturbed_1e-9m_vs_fp64']['depth_Linf_m']:.3g} m")
        g32.close()
        gf.close()
        gp.close()
        out["accuracy"] = curve
    return out


def fast_f64(sc, args, bitwise_value):
    """The fast FP64 kind (HF2D_F64_FAST: contraction, reciprocal/Newton
    division and square root, primitive-variable faces) on the headline
    workload, and its drift from
    the bitwise kind on the 256 x 256 dam break (flat frictionless bed,
    walls, CFL 0.5) after 10, 100, 1,000 and 20,000 steps, against the
    1e-10 budget."""
    from hydroflux_amd import Grid2D

    log("[bench] fast FP64 kind: timing")
    g = build_grid(sc, "f64fast")
    wall, t1, t2 = time_steps(g, args.warmup, args.steps)
    g.close()
    out = {"value": sc.n_cells * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
           "over_bitwise_fp64": sc.n_cells * args.steps / wall / bitwise_value,
           "roofline": roofline("f64", sc.n_cells, t1, t2, _traffic("f64fast_final_stage_bytes_per_launch"),
                                "stage_kernel<double, 32, 12, 512, final> (fast build, one workgroup per tile)",
                                "f64fast_final")}
    n = 256

    def dam(kind):
        d = Grid2D(n, n, 0.1, 0.1, kind=kind)
        st = np.zeros((n, n, 3))
        st[..., 0] = np.where((np.arange(n) + 0.5) * 0.1 < 12.8, 1.0, 0.1)[None, :]
        d.set_mesh(np.zeros((n, n)), np.zeros((n, n)))
        d.set_state(st)
        d.set_boundaries([1, 1, 1, 1])
        d.configure(cfl=0.5, max_dt=1e9, friction=False)
        return d

    a, b = dam("f64"), dam("f64fast")
    v0 = b.total_volume()
    drift, done = [], 0
    for m in (10, 100, 1000, 20000):
        a.run_steps(m - done)
        b.run_steps(m - done)
        done = m
        sa, sb = a.get_state(), b.get_state()
        dh = float(np.max(np.abs(sa - sb)[..., 0]))
        drift.append({"steps": m, "max_abs_h_m": dh, "max_abs_q_m2s": float(np.max(np.abs(sa - sb)[..., 1:])),
                      "time_diff_s": abs(a.time()[0] - b.time()[0]), "within_1e-10": dh <= 1e-10,
                      "mass_rel_err": abs(b.total_volume() - v0) / v0})
    a.close()
    b.close()
    out["drift_vs_bitwise_dam_break_256"] = drift
    log(f"[bench] fast FP64: {out['value'] / 1e9:.2f} G cell-updates/s ({out['over_bitwise_fp64']:.2f}x), "
        f"dam-break drift after 1000 / 20000 steps {drift[-2]['max_abs_h_m']:.2g} / {drift[-1]['max_abs_h_m']:.2g} m")
    return out


def c1_1024(args):
    """The 1024^2 workload (no bed offset) in FP64 and Dual (tangent seeded
    on a global Manning multiplier: n-tangent = n)."""
    from hydroflux_amd.scenarios import flood_scenario

    sc = flood_scenario(n=1024, dx=5.0)
    out = {}
    for kind in ("f64", "dual"):
        g = build_grid(sc, kind)
        w, _, _ = time_steps(g, 10, 50)
        out[kind] = sc.n_cells * 50 / w
        g.close()
    out["dual_over_f64_time"] = out["f64"] / out["dual"]
    return out


def dry_domain(args):
    """The compacted active-tile list (the reference's cell-mask early skip
    at tile granularity): the 4096^2 flood construction without rain — only
    the 100 m^3/s inflow wets the valley — timed with the list on and off
    after a warm-up that spreads the inflow. Same results either way;
    cell-updates count every cell of the grid."""
    from hydroflux_amd.scenarios import flood_scenario

    sc = flood_scenario(n=args.size, dx=5.0, offset=2000.0, rain_mm_h=0.0)
    out = {"workload": f"flood DEM {args.size}x{args.size}, +2000 m, no rain, 100 m3/s inflow, "
                       f"after {args.dry_warmup} steps"}
    for skip in (True, False):
        g = build_grid(sc)
        g.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6, rain_rate=0.0,
                    sources=sc.sources, active_tiles=skip)
        wall, t1, t2 = time_steps(g, args.dry_warmup, args.steps)
        h = g.get_state()[..., 0]
        key = "skip" if skip else "no_skip"
        out[key] = {"value": sc.n_cells * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
                    "first_stage_ms": t1, "final_stage_ms": t2}
        out["wet_cells"] = int(np.count_nonzero(h > 1e-6))
        out["wet_fraction"] = out["wet_cells"] / h.size
        g.close()
    out["speedup"] = out["skip"]["value"] / out["no_skip"]["value"]
    log(f"[bench] dry domain: list {out['skip']['value'] / 1e9:.2f} vs every tile "
        f"{out['no_skip']['value'] / 1e9:.2f} G cell-updates/s ({out['wet_cells']} wet cells)")
    return out


def uk_ea(args):
    """The UK EA official benchmarks end to end on the GPU (device-side
    hydrograph / hyetograph series and gauges; 18000 simulated s each):
    Test 4 (400 x 200, 5 m) and Test 8A (Glasgow, 199 x 481, 2 m). The
    reference's recorded CPU wall times of the same runs, on its authors'
    machine (papers/01_review/ROADMAP_REVISION_EMS.md:211-229), are given for
    context only. Each test also runs in the FP32 precision form, and in FP64
    with the bed perturbed by up to 1e-9 m (the run's own conditioning): both
    are compared with the FP64 run (final depth field and the control-point
    series)."""
    from hydroflux_amd import uk_ea as U
    from hydroflux_amd.ascii_grid import read_ascii_grid
    from hydroflux_amd.metrics import depth_metrics

    g4 = os.path.join(ROOT, "tests", "golden", "uk_ea", "test4")
    g8 = os.path.join(ROOT, "tests", "golden", "uk_ea", "test8a")
    dem4 = os.path.join(g4, "ea4-5m.dem.gz")
    dem8, man8 = os.path.join(g8, "ea8-2m.dem.gz"), os.path.join(g8, "ea8-2m.n.gz")

    def make4(kind, pert):
        g, _ = U.ea_test4_grid(dem4, kind=kind)
        if pert:
            bed, _ = read_ascii_grid(dem4)
            g.set_mesh(bed + pert * np.random.default_rng(5).uniform(-1, 1, bed.shape), U.MANNING_N)
        return g

    def make8(kind, pert):
        g, _ = U.ea_test8a_grid(dem8, man8, kind=kind)
        if pert:
            bed, _ = read_ascii_grid(dem8)
            man, _ = read_ascii_grid(man8)
            g.set_mesh(bed + pert * np.random.default_rng(5).uniform(-1, 1, bed.shape), man)
        return g

    out = {}
    for name, make, ref_s, dx in (("test4", make4, 1136.0, 5.0), ("test8a", make8, 5289.0, 2.0)):
        runs = {}
        for kind, pert in (("f64", 0.0), ("f32", 0.0), ("f64", 1e-9)):
            g = make(kind, pert)
            cells = g.n_rows * g.n_cols
            g.synchronize()
            t0 = time.perf_counter()
            t_end = args.ukea_time if args.ukea_time > 0 else U.SIM_TIME
            g.run_until(t_end)
            g.synchronize()
            wall = time.perf_counter() - t0
            steps = g.time()[1]
            times, depths = g.read_gauges()
            runs[(kind, pert)] = (g.get_state()[..., 0].copy(), g.total_volume(), np.asarray(times),
                                  np.asarray(depths), steps, wall)
            g.close()
            log(f"[bench] UK EA {name} {kind}{' bed+1e-9' if pert else ''}: {steps} steps in {wall:.2f} s "
                f"({wall / steps * 1e6:.1f} us/step)")
        h64, v64, t64, d64, steps, wall = runs[("f64", 0.0)]
        r = {"grid": list(h64.shape), "simulated_s": t_end, "steps": steps, "wall_s": wall,
             "cell_updates_per_s": h64.size * steps / wall, "us_per_step": wall / steps * 1e6,
             "reference_cpu_wall_s": ref_s}
        for key, label in ((("f32", 0.0), "fp32_form_vs_fp64"), (("f64", 1e-9), "fp64_bed_perturbed_1e-9m_vs_fp64")):
            h, v, tt, dd, st, wl = runs[key]
            m = depth_metrics(h, h64, dx, dx, vol_test=v, vol_ref=v64)
            m["steps"] = st
            m["control_point_max_abs_diff_m"] = max(
                float(np.max(np.abs(np.interp(t64, tt, dd[:, p]) - d64[:, p]))) for p in range(d64.shape[1]))
            m["control_point_peak_max_abs_diff_m"] = max(
                abs(float(dd[:, p].max() - d64[:, p].max())) for p in range(d64.shape[1]))
            if key[0] == "f32":
                m["us_per_step"] = wl / st * 1e6
            r[label] = m
        out[name] = r
    return out


def huasco_day(args):
    """The Huasco reach from its GeoTIFF inputs (200 x 67 cells at 30 m),
    one simulated day on the GPU: the sealed, source-free f64 run of
    huasco_closed_domain.rs --no-inflow (the reference's cross-validation
    against SynxFlow) and the Dual run of huasco_calibrate_tree.rs (value
    and d h / d n_tree in one pass). SynxFlow's time for its run of the same
    day, on an RTX 3050 (docs/xval-synxflow-huasco-results.md:126), is given
    for context."""
    from hydroflux_amd import huasco as H
    from hydroflux_amd.ascii_grid import read_ascii_grid

    d = os.path.join(ROOT, "tests", "golden", "huasco")
    inp = H.load_inputs(os.path.join(d, "huasco_subset_dem.tif"), os.path.join(d, "huasco_subset_landcover.tif"),
                        os.path.join(d, "huasco_subset_acc.tif"))
    syn, _ = read_ascii_grid(os.path.join(d, "synxflow_noflow_h_86400.asc.gz"))
    out = {"grid": list(inp.mesh.bed.shape), "synxflow_gpu_s_context": 6.5}
    for name, make in (("sealed_f64", lambda: H.closed_domain_grid(inp, no_inflow=True)[0]),
                       ("calibration_dual", lambda: H.calibration_grid(inp, H.N_TREE_INIT))):
        g = make()
        g.synchronize()
        t0 = time.perf_counter()
        g.run_until(H.SECONDS_PER_DAY)
        g.synchronize()
        wall = time.perf_counter() - t0
        steps = g.time()[1]
        st = g.get_state()
        h = (st[0] if isinstance(st, tuple) else st)[..., 0]
        out[name] = {"steps": steps, "wall_s": wall, "us_per_step": wall / steps * 1e6,
                     "cell_updates_per_s": h.size * steps / wall}
        if name == "sealed_f64":
            out[name]["vs_synxflow"] = H.xval_metrics(h, syn)
        g.close()
        log(f"[bench] Huasco {name}: {steps} steps in {wall:.2f} s ({wall / steps * 1e6:.1f} us/step)")
    return out


def _zonal_pass(n, steps, kind, n_params, seed_zone=None):
    """solver-2d/examples/m1_forward_scaling.rs on the GPU: dam break (h = 1
    left half, 0.1 right half), transmissive edges, Manning 0.03 in
    n_params vertical stripes, `steps` SSP-RK2 + friction(1e-6) steps at a
    fixed dt of 0.01 s, host loop as in the example. Returns the wall time
    and the depth tangent field (Dual: n x n; DualN<P>: n x n x P)."""
    from hydroflux_amd import Grid2D

    st = np.zeros((n, n, 3))
    st[..., 0] = np.where(np.arange(n) < n // 2, 1.0, 0.1)[None, :]
    zone = (np.arange(n) * n_params) // n
    g = Grid2D(n, n, 1.0, 1.0, kind=kind)
    mt = None
    if kind == "dual":
        mt = np.where(zone == seed_zone, 1.0, 0.0)[None, :].repeat(n, 0)
    elif kind != "f64":
        mt = np.zeros((n, n, n_params))
        for k in range(n_params):
            mt[:, zone == k, k] = 1.0
    g.set_mesh(np.zeros((n, n)), np.full((n, n), 0.03), manning_tangent=mt)
    g.set_state(st)
    g.set_boundaries([0, 0, 0, 0])
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.ssprk2_step(0.01)
        g.manning_friction_step(0.01, 1e-6)
    g.synchronize()
    wall = time.perf_counter() - t0
    out = g.get_state()
    g.close()
    if kind == "f64":
        return wall, None
    return wall, out[1][..., 0] if kind == "dual" else out[1][..., 0, :]


def forward_scaling(args):
    """Cost of the full gradient over P zonal Manning parameters relative to
    the f64 primal, as solver-2d/examples/m1_forward_scaling.rs defines it:
    P seeded Dual passes (per-parameter cost = cost / primal / P) and one
    DualN<P> pass (per-parameter cost = (cost / primal - 1) / P), on the
    example's 64^2 x 200-step workload and on the same construction at
    2048^2 x 20 steps. The reference's figures on its CPU: 2.05-2.18x per
    parameter with Dual (papers/01_review/figures/data/m1_forward_scaling.csv:2-6),
    0.88x with DualN (papers/01_review/manuscript.md:406-412).
    tangent_rel_diff: max over the P components of the DualN depth tangent
    field against the seeded Dual run's, relative to the latter's largest
    entry (the two value paths differ in the last bit, so cells where a
    near-tie resolves differently differ)."""
    out = {}
    for n, steps in ((64, 200), (2048, 20)):
        _zonal_pass(n, steps, "f64", 1)  # warm
        prim = min(_zonal_pass(n, steps, "f64", 1)[0] for _ in range(3))
        rows = []
        for p in (1, 2, 4, 8, 16):
            fields, wall_d = [], 0.0
            for k in range(p):
                w, td = _zonal_pass(n, steps, "dual", p, k)
                wall_d += w
                fields.append(td)
            row = {"P": p, "dual_passes_s": wall_d, "dual_cost_vs_primal": wall_d / prim,
                   "dual_per_param": wall_d / prim / p}
            if p > 1:
                _zonal_pass(n, steps, f"dual{p}", p)  # warm
                w, tn = _zonal_pass(n, steps, f"dual{p}", p)
                diff = max(float(np.max(np.abs(tn[..., k] - fields[k]))) / max(float(np.max(np.abs(fields[k]))),
                                                                                1e-300) for k in range(p))
                row.update({"dualn_s": w, "dualn_cost_vs_primal": w / prim, "dualn_per_param": (w / prim - 1.0) / p,
                            "tangent_rel_diff": diff})
            rows.append(row)
            log(f"[bench] forward scaling {n}^2 P={p}: Dual {row['dual_per_param']:.2f}x / param"
                + (f", DualN {row['dualn_per_param']:.2f}x / param" if p > 1 else ""))
        out[f"{n}x{n}x{steps}"] = {
            "primal_s": prim, "rows": rows,
            "dual_per_param_mean": float(np.mean([r["dual_per_param"] for r in rows])),
            "dualn_per_param_mean": float(np.mean([r["dualn_per_param"] for r in rows if "dualn_per_param" in r]))}
    out["reference_cpu"] = {"grid": "64x64x200", "dual_per_param": [2.0481, 2.1392, 2.1777, 2.1176, 2.1796],
                            "dualn_per_param_mean": 0.88}
    return out


def batch_bench(args):
    from hydroflux_amd import Grid2D
    from hydroflux_amd.batch import Batch, ScenarioSpec
    from hydroflux_amd.scenarios import scenario_mix

    t0 = time.time()
    mix = scenario_mix(256)
    cells = sum(s.n_cells for s in mix)
    log(f"[bench] batch: 256 scenarios, {cells / 1e6:.1f} M cells (generated in {time.time() - t0:.1f} s)")

    def specs():
        return [ScenarioSpec(bed=s.bed, manning=s.manning, dx=s.dx, bc_kinds=s.bc_kinds, rain_rate=s.rain_rate,
                             sources=s.sources, t_end=1e9, max_dt=s.max_dt, cfl=s.cfl) for s in mix]

    out = {"scenarios": len(mix), "cells": cells}
    for kind in ("f32", "f64"):
        b = Batch(kind, capacity_bytes=24 << 30, max_slots=512)
        for sp in specs():
            b.add(sp)
        b.step(8)
        b.synchronize()
        best, ms = b.autotune(6)
        b.step(16)
        b.synchronize()
        n = 64
        t0 = time.perf_counter()
        b.step(n)
        b.synchronize()
        w = time.perf_counter() - t0
        cfgs = b.configs()
        out[kind] = {"value": cells * n / w, "ms_per_step": w / n * 1e3, "tile": cfgs[best], "tile_index": int(best),
                     "autotune_ms_per_step": {f"{c[0]}x{c[1]}x{c[2]}": float(m) for c, m in zip(cfgs, ms)}}
        b.close()
        log(f"[bench] batch {kind}: {cells * n / w / 1e9:.2f} G cell-updates/s, tile {cfgs[best]}")
    # serial per-scenario plain FP64 launches (one grid and driver call per scenario)
    tot_c, tot_w = 0, 0.0
    for s in mix:
        g = Grid2D(s.n_rows, s.n_cols, s.dx, s.dx)
        g.set_mesh(s.bed, s.manning)
        g.set_state(s.dry_state())
        g.set_boundaries(list(s.bc_kinds))
        g.configure(cfl=s.cfl, max_dt=s.max_dt, rain_rate=s.rain_rate, sources=s.sources)
        g.run_steps(8)
        g.synchronize()
        t0 = time.perf_counter()
        g.run_steps(16)
        tot_w += time.perf_counter() - t0
        tot_c += s.n_cells * 16
        g.close()
    out["serial_f64"] = tot_c / tot_w
    out["f32_over_serial_f64"] = out["f32"]["value"] / out["serial_f64"]
    out["f64_over_serial_f64"] = out["f64"]["value"] / out["serial_f64"]
    log(f"[bench] serial f64: {out['serial_f64'] / 1e9:.2f} G cell-updates/s")
    out["f32_accuracy"] = batch_accuracy(mix)
    return out, mix


def batch_accuracy(mix, n_pick=16, t_end=120.0):
    """Depth accuracy of the FP32 batch against the FP64 batch on the mix:
    n_pick scenarios spread over its sizes (sorted by cell count), each run
    to t_end simulated seconds in an FP64 batch, an FP32 batch and an FP64
    batch on beds perturbed by up to 1e-9 m (each scenario's own
    conditioning); per-scenario depth metrics against the FP64 run, reported
    as the worst and the median over the picked scenarios."""
    from hydroflux_amd.batch import Batch, ScenarioSpec
    from hydroflux_amd.metrics import depth_metrics

    order = np.argsort([s.n_cells for s in mix])
    pick = [mix[int(order[int(round(k * (len(mix) - 1) / (n_pick - 1)))])] for k in range(n_pick)]
    rng = np.random.default_rng(0xACC)
    perts = [1e-9 * rng.uniform(-1.0, 1.0, s.bed.shape) for s in pick]

    def run(kind, perturbed):
        b = Batch(kind, capacity_bytes=16 << 30, max_slots=64)
        slots = [b.add(ScenarioSpec(bed=s.bed + (p if perturbed else 0.0), manning=s.manning, dx=s.dx,
                                    bc_kinds=s.bc_kinds, rain_rate=s.rain_rate, sources=s.sources, t_end=t_end,
                                    max_dt=s.max_dt, cfl=s.cfl)) for s, p in zip(pick, perts)]
        while True:
            b.step(64)
            b.synchronize()
            t, steps, status, done, used = b.poll()
            if all(done[sl] or status[sl] for sl in slots):
                break
        res = []
        for s, sl in zip(pick, slots):
            assert status[sl] == 0 and t[sl] == t_end, (kind, s.n_rows, s.n_cols, status[sl], t[sl])
            res.append(b.download(sl, (s.n_rows, s.n_cols))[..., 0].copy())
        b.close()
        return res

    t0 = time.perf_counter()
    h64, h32, hp = run("f64", False), run("f32", False), run("f64", True)
    rows = []
    for s, a, b32, bp in zip(pick, h64, h32, hp):
        m32 = depth_metrics(b32, a, s.dx, s.dx)
        mp = depth_metrics(bp, a, s.dx, s.dx)
        rows.append({"grid": [s.n_rows, s.n_cols], "fp32_Linf_m": m32["depth_Linf_m"], "fp32_L1_m": m32["depth_L1_m"],
                     "fp32_iou": m32["extent_iou_1cm"], "fp32_mass": m32["mass_balance_rel_err"],
                     "fp64_perturbed_Linf_m": mp["depth_Linf_m"], "fp64_perturbed_L1_m": mp["depth_L1_m"],
                     "fp64_perturbed_mass": mp["mass_balance_rel_err"]})
    li = np.array([r["fp32_Linf_m"] for r in rows])
    lp = np.array([r["fp64_perturbed_Linf_m"] for r in rows])
    out = {"scenarios": n_pick, "t_end_s": t_end,
           "fp32_Linf_m": {"max": float(li.max()), "median": float(np.median(li))},
           "fp64_bed_perturbed_1e-9m_Linf_m": {"max": float(lp.max()), "median": float(np.median(lp))},
           "fp32_within_1mm": int(np.sum(li <= 1e-3)),
           "fp32_L1_m_max": float(max(r["fp32_L1_m"] for r in rows)),
           "fp32_iou_min": float(min(r["fp32_iou"] for r in rows)),
           "fp32_mass_rel_err_max": float(max(abs(r["fp32_mass"]) for r in rows)),
           "fp64_perturbed_L1_m_max": float(max(r["fp64_perturbed_L1_m"] for r in rows)),
           "fp64_perturbed_mass_rel_err_max": float(max(abs(r["fp64_perturbed_mass"]) for r in rows)),
           "per_scenario": rows, "wall_s": time.perf_counter() - t0}
    log(f"[bench] batch FP32 vs FP64 on {n_pick} mix scenarios to {t_end:.0f} s: Linf max {li.max():.3g} m "
        f"(median {np.median(li):.3g}); perturbed FP64 Linf max {lp.max():.3g} m (median {np.median(lp):.3g}); "
        f"{out['fp32_within_1mm']} of {n_pick} within 1 mm")
    return out


def serving_bench(args, mix, f32_tile, batch_f32_value):
    """Continuous batching over requests drawn from the 256-scenario mix,
    each running for its own simulated duration (--serve-duration, scaled
    down from the 600-7200 s of the serving configuration so that a sweep
    fits the bench), admitted by HBM footprint into an FP32 arena that holds
    about a third of the mix at once:

    - FP32 requests at a sweep of Poisson arrival rates up to saturation
      (--serve-requests each; at rate inf all 256): p50/p99 latency to
      completion and aggregate cell-updates/s, the last against the batch
      engine running the whole mix resident;
    - the serving configuration's mix, 25 % of the requests with Dual
      tangents on one land-cover class's Manning n, at rate inf."""
    from hydroflux_amd.batch import ScenarioSpec
    from hydroflux_amd.serving import Request, Server, poisson_arrivals

    rng = np.random.default_rng(0x5E4F)
    lo, hi = args.serve_duration
    durations = rng.uniform(lo, hi, len(mix))

    def spec(i, dual):
        s = mix[i]
        return ScenarioSpec(bed=s.bed, manning=s.manning, dx=s.dx, bc_kinds=s.bc_kinds, rain_rate=s.rain_rate,
                            sources=s.sources, t_end=float(durations[i]), max_dt=s.max_dt, cfl=s.cfl,
                            manning_tangent=((s.land_class == 2).astype(np.float64) if dual else None))

    cap = int(args.serve_capacity_gb * (1 << 30))
    out = {"duration_s": [lo, hi], "f32_arena_bytes": cap, "f32_rates": []}

    def serve(reqs, label):
        srv = Server({"f32": cap, "dual": 6 << 30}, max_slots=512, chunk_steps=args.serve_chunk, keep_results=False,
                     tile_config={"f32": f32_tile})
        seen = {"n": 0, "t": time.perf_counter()}

        def progress(r):  # a line at most every 20 s: long runs at full durations stay visible
            seen["n"] += 1
            if time.perf_counter() - seen["t"] >= 20.0:
                seen["t"] = time.perf_counter()
                log(f"[bench]   serving {label}: {seen['n']} of {len(reqs)} requests done")

        rep = srv.run(reqs, on_done=progress)
        srv.close()
        s = rep.summary()
        log(f"[bench] serving {label}: {len(reqs)} requests, p50 {s['latency_p50_s']:.2f} s, "
            f"p99 {s['latency_p99_s']:.2f} s, {s['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s")
        return s

    subset = rng.choice(len(mix), size=min(args.serve_requests, len(mix)), replace=False)
    for rate in args.serve_rates:
        idx = range(len(mix)) if not np.isfinite(rate) else subset
        arr = poisson_arrivals(len(idx), rate=rate, seed=7)
        reqs = [Request(k, float(a), spec(i, False), "f32") for k, (i, a) in enumerate(zip(idx, arr))]
        s = serve(reqs, f"FP32 rate {rate}")
        s["rate_req_per_s"] = None if not np.isfinite(rate) else rate
        out["f32_rates"].append(s)
    sat = out["f32_rates"][-1]
    out["f32_saturated_over_batch_engine"] = sat["cell_updates_per_s"] / batch_f32_value
    reqs = [Request(k, 0.0, spec(int(i), k % 4 == 3), "dual" if k % 4 == 3 else "f32")
            for k, i in enumerate(subset)]
    out["mixed_25pct_dual_rate_inf"] = serve(reqs, "25 % Dual, rate inf")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline FP64 line only (and cpu_baseline)")
    ap.add_argument("--sections", default="fast,precision,c1,dry,ukea,huasco,forward,batch,serving",
                    help="comma list of extra sections: fast, precision, c1, dry, ukea, huasco, forward, batch, "
                         "serving")
    ap.add_argument("--dry-warmup", type=int, default=500)
    ap.add_argument("--ukea-time", type=float, default=0.0,
                    help="simulated seconds of the UK EA runs (0: the full 18000 s)")
    ap.add_argument("--acc-steps", type=int, default=20000,
                    help="FP64 steps of the FP32 accuracy curve (checkpoints 100, 1000, 5000 and this; 0: skip)")
    ap.add_argument("--serve-requests", type=int, default=64, help="requests per finite arrival rate")
    ap.add_argument("--serve-duration", type=float, nargs=2, default=(30.0, 120.0),
                    help="simulated seconds per request, uniform in [lo, hi]")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=(8.0, 16.0, 32.0, 64.0, 128.0, float("inf")))
    ap.add_argument("--serve-capacity-gb", type=float, default=1.25, help="FP32 serving arena (GiB)")
    ap.add_argument("--serve-chunk", type=int, default=32, help="steps per serving chunk")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the step runs on one GPU; --gpus {args.gpus} is not supported "
              "(a single grid would shard by row strips with halo exchange, scenarios by footprint)",
              file=sys.stderr)
        sys.exit(2)

    from hydroflux_amd._lib import lib
    from hydroflux_amd.scenarios import flood_scenario

    n = args.size
    log(f"[bench] building {n}x{n} flood scenario")
    sc = flood_scenario(n=n, dx=5.0, offset=2000.0)
    g = build_grid(sc)
    log(f"[bench] FP64: warmup {args.warmup} + timed {args.steps} steps")
    wall, t_first, t_final = time_steps(g, args.warmup, args.steps)
    cells = sc.n_cells
    value = cells * args.steps / wall
    log(f"[bench] {value / 1e9:.3f} G cell-updates/s, {wall / args.steps * 1e3:.3f} ms/step, "
        f"stage1 {t_first:.3f} ms, final {t_final:.3f} ms")

    ms = C.c_double()
    copy_bytes = 1 << 30
    lib().hf2d_copy_probe(copy_bytes, 20, C.byref(ms))
    copy_gbs = 2 * copy_bytes / (ms.value / 1e3) / 1e9
    log(f"[bench] copy probe {copy_gbs:.0f} GB/s")

    rf = roofline("f64", cells, t_first, t_final, _traffic("final_stage_bytes_per_launch"),
                  "stage_persistent_kernel<double, final> (final SSP-RK2 stage + combine + forcing + CFL reduction)",
                  "f64_final")
    rf["step_hbm_frac"] = value * ALG_BYTES["f64"]["step"] / 1e9 / HBM_PEAK_GBS
    rf["copy_probe_GBs"] = copy_gbs
    rf["copy_probe_frac_of_peak"] = copy_gbs / HBM_PEAK_GBS
    rf["step_frac_of_copy_probe"] = value * ALG_BYTES["f64"]["step"] / 1e9 / copy_gbs
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"flood DEM {n}x{n} dx 5 m, +2000 m bed offset, FP64 SSP-RK2 + Manning friction + "
                        f"50 mm/h rain + 100 m3/s inflow, device-side CFL dt",
            "grid": [sc.n_rows, sc.n_cols],
            "cells": cells,
            "precision": "fp64 (bit-identical to the CPU restatement)",
            "parallelism": "single GPU",
        },
        "roofline": rf,
    }
    sections = set() if args.quick else set(x for x in args.sections.split(",") if x)
    if "fast" in sections:
        out["fast_f64"] = fast_f64(sc, args, value)
    if "precision" in sections:
        out["precision_form"] = precision_form(sc, g, args)
        out["precision_form"]["over_fp64"] = out["precision_form"]["value"] / value
    g.close()
    if "c1" in sections:
        log("[bench] 1024^2 FP64 and Dual")
        out["c1_1024"] = c1_1024(args)
    if "dry" in sections:
        out["dry_domain"] = dry_domain(args)
    if "ukea" in sections:
        out["uk_ea"] = uk_ea(args)
    if "huasco" in sections:
        out["huasco"] = huasco_day(args)
    if "forward" in sections:
        out["forward_scaling"] = forward_scaling(args)
    if "batch" in sections or "serving" in sections:
        out["batch"], mix = batch_bench(args)
        if "serving" in sections:
            sv = serving_bench(args, mix, out["batch"]["f32"]["tile_index"], out["batch"]["f32"]["value"])
            sv["serial_f64"] = out["batch"]["serial_f64"]
            sv["f32_saturated_over_serial_f64"] = sv["f32_rates"][-1]["cell_updates_per_s"] / sv["serial_f64"]
            out["serving"] = sv
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, 1 thread pinned: dam break 256^2 x 500, flood 1024^2 f64 and Dual)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
