"""Benchmark: SSP-RK2 flood step on one MI355X, cell-updates/s and HBM roofline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 4096] [--no-cpu]

Workload (config.workload): the rain-on-grid flood construction — seeded fBm
DEM (H 0.8, 50 m relief) + 0.01 slope + 8 m carved valley, +2000 m bed
offset, Manning n from 4 land-cover classes in Voronoi patches, initially
dry, 50 mm/h rain, 100 m^3/s point inflow, walls N/S/W and open E — at
4096 x 4096 cells, dx 5 m, in FP64: the largest single-GPU plain-FP64
configuration, and a footprint (1.1 GB) well beyond the 256 MB Infinity
Cache, so HBM bandwidth is what is measured. Each step is one SSP-RK2 step
with Manning friction, rain and the point source, dt from the device-side
CFL reduction (no host round trip), after W warm-up steps that wet the
domain. Inputs are resident in HBM before the timed region.

One JSON line on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s 2D SSP-RK2 (FP64 vs FP32) + %HBM peak; batched p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell for one launch, FP64 (8 B per value):
#   first stage : read h, hu, hv, z; write U(1)                    = 7 x 8 = 56
#   final stage : read U(1), z, U^n, n; write U^{n+1}             = 11 x 8 = 88
ALG_BYTES = {"stage1": 56.0, "final": 88.0, "step": 144.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_grid(sc):
    from hydroflux_amd import Grid2D

    g = Grid2D(sc.n_rows, sc.n_cols, sc.dx, sc.dx)
    g.set_mesh(sc.bed, sc.manning)
    g.set_state(sc.dry_state())
    g.set_boundaries(list(sc.bc_kinds))
    g.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6, rain_rate=sc.rain_rate,
                sources=sc.sources)
    return g


def time_steps(g, warmup, steps):
    g.run_steps(warmup)
    g.synchronize()
    g.set_kernel_timing(True)
    t0 = time.perf_counter()
    g.enqueue_steps(steps)
    g.synchronize()
    t1 = time.perf_counter()
    s1, s2, cnt = g.kernel_timing()
    g.set_kernel_timing(False)
    return (t1 - t0), s1, s2, cnt


def cpu_baseline(n=1024, warmup=3, steps=20):
    """The CPU restatement of the reference step, single-threaded as the
    reference runs by default, on a bounded sample of the same workload
    construction at 1024 x 1024."""
    from hydroflux_amd.scenarios import flood_scenario
    from oracle.pyoracle import OracleSim, set_pow_mode, set_threads

    sc = flood_scenario(n=n, dx=5.0)
    set_threads(1)
    set_pow_mode(0)
    o = OracleSim(sc.n_rows, sc.n_cols, sc.dx, sc.dx)
    o.set_mesh(sc.bed, sc.manning)
    o.set_state(sc.dry_state())
    o.set_boundaries(list(sc.bc_kinds))
    o.configure(cfl=sc.cfl, max_dt=sc.max_dt, friction=True, friction_dry_tol=1e-6, rain_rate=sc.rain_rate,
                sources=sc.sources)
    o.run_steps(warmup)
    t0 = time.perf_counter()
    o.run_steps(steps)
    dt = time.perf_counter() - t0
    return {
        "value": sc.n_cells * steps / dt,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"C++ restatement of the reference step (oracle), 1 thread, flood DEM {n}x{n}, "
                  f"{warmup} warm-up + {steps} timed SSP-RK2+friction+rain+inflow steps ({dt:.1f} s)",
        "ns_per_cell_step": dt / (sc.n_cells * steps) * 1e9,
        "host_cpu": _cpu_model(),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _traffic():
    """HBM bytes per launch of the final stage kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated)."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get("final_stage_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--secondary", action="store_true", help="also time 1024^2 f64 and Dual")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a single grid would shard by row strips with halo exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from hydroflux_amd._lib import lib
    from hydroflux_amd.scenarios import flood_scenario

    n = args.size
    log(f"[bench] building {n}x{n} flood scenario")
    sc = flood_scenario(n=n, dx=5.0, offset=2000.0)
    g = build_grid(sc)
    log(f"[bench] warmup {args.warmup} + timed {args.steps} steps")
    wall, s1_ms, s2_ms, cnt = time_steps(g, args.warmup, args.steps)
    cells = sc.n_cells
    value = cells * args.steps / wall
    t_final = s2_ms / max(cnt, 1) / 1e3
    t_first = s1_ms / max(cnt, 1) / 1e3
    achieved = ALG_BYTES["final"] * cells / t_final / 1e9
    log(f"[bench] {value / 1e9:.3f} G cell-updates/s, {wall / args.steps * 1e3:.3f} ms/step, "
        f"stage1 {t_first * 1e3:.3f} ms, final {t_final * 1e3:.3f} ms")

    import ctypes as C

    ms = C.c_double()
    copy_bytes = 1 << 30
    lib().hf2d_copy_probe(copy_bytes, 20, C.byref(ms))
    copy_gbs = 2 * copy_bytes / (ms.value / 1e3) / 1e9
    log(f"[bench] copy probe {copy_gbs:.0f} GB/s")

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
        "roofline": {
            "bound": "hbm",
            "kernel": "stage_kernel<double, final> (final SSP-RK2 stage + combine + forcing + CFL reduction)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": _traffic(),
            "alg_bytes_per_cell": ALG_BYTES["final"],
            "avg_launch_ms": t_final * 1e3,
            "first_stage": {"alg_bytes_per_cell": ALG_BYTES["stage1"], "avg_launch_ms": t_first * 1e3,
                            "achieved": ALG_BYTES["stage1"] * cells / t_first / 1e9},
            "step_hbm_frac": value * ALG_BYTES["step"] / 1e9 / HBM_PEAK_GBS,
            "copy_probe_GBs": copy_gbs,
        },
    }
    if args.secondary:
        for dual in (False, True):
            sc1 = flood_scenario(n=1024, dx=5.0)
            from hydroflux_amd import Grid2D

            g1 = Grid2D(1024, 1024, 5.0, 5.0, dual=dual)
            g1.set_mesh(sc1.bed, sc1.manning, manning_tangent=(sc1.manning if dual else None))
            g1.set_state(sc1.dry_state())
            g1.set_boundaries(list(sc1.bc_kinds))
            g1.configure(cfl=0.4, max_dt=1.0, rain_rate=sc1.rain_rate, sources=sc1.sources)
            w1, _, _, _ = time_steps(g1, 10, 50)
            out.setdefault("secondary", {})["dual_1024" if dual else "f64_1024"] = 1024 * 1024 * 50 / w1
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, 1 thread, 1024^2 sample)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
