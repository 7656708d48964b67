"""Control-point (stage) series: LISFLOOD-FP .stage reference files and the
comparison the reference's UK EA loops report (RMSE, peak bias, arrival).

Host-side analysis of gauge records (Grid2D.set_gauges / read_gauges),
restating solver-2d/examples/uk_ea_common/mod.rs: parse_reference_stage,
interp (piecewise linear, clamped at both ends), compare_point (the
simulated series interpolated onto the reference's times; RMSE over the
reference rows; peak bias = sim peak - reference peak, with peaks folded
from 0.0; arrival = first reference time whose depth exceeds a threshold).
"""
from __future__ import annotations

import bisect
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple


@dataclass
class ReferenceStage:
    points: List[Tuple[float, float]]
    times: List[float]
    depths: List[List[float]]


def parse_reference_stage(path: str) -> ReferenceStage:
    with open(path) as f:
        lines = iter(f.read().splitlines())
    for line in lines:
        if "stage,x,y,elev" in line:
            break
    points = []
    for line in lines:
        toks = line.split()
        if len(toks) < 3:
            break
        try:
            points.append((float(toks[1]), float(toks[2])))
        except ValueError:
            break
    for line in lines:
        if "Time;" in line:
            break
    times, depths = [], []
    for line in lines:
        try:
            vals = [float(t) for t in line.split()]
        except ValueError:
            continue
        if len(vals) != len(points) + 1:
            continue
        times.append(vals[0])
        depths.append(vals[1:])
    if not points or not times:
        raise ValueError(f"no control points or data rows parsed from {path}")
    return ReferenceStage(points, times, depths)


def interp(times: Sequence[float], values: Sequence[float], t: float) -> float:
    if t <= times[0]:
        return values[0]
    if t >= times[-1]:
        return values[-1]
    idx = bisect.bisect_right(times, t)  # partition_point(|ti| ti <= t)
    t0, t1 = times[idx - 1], times[idx]
    v0, v1 = values[idx - 1], values[idx]
    frac = (t - t0) / (t1 - t0)
    return v0 + frac * (v1 - v0)


@dataclass
class PointComparison:
    rmse: float
    peak_bias: float
    reference_peak: float
    reference_arrival: Optional[float]
    sim_arrival: Optional[float]


def _arrival(times, values, threshold):
    for t, v in zip(times, values):
        if v > threshold:
            return t
    return None


def compare_point(reference: ReferenceStage, point_idx: int, sim_times: Sequence[float],
                  sim_values: Sequence[float], arrival_threshold: float) -> PointComparison:
    ref = [row[point_idx] for row in reference.depths]
    sim = [interp(sim_times, sim_values, t) for t in reference.times]
    mse = 0.0
    for r, s in zip(ref, sim):
        mse += (r - s) ** 2
    mse /= len(ref)
    ref_peak = 0.0
    for r in ref:
        ref_peak = max(ref_peak, r)
    sim_peak = 0.0
    for s in sim:
        sim_peak = max(sim_peak, s)
    return PointComparison(math.sqrt(mse), sim_peak - ref_peak, ref_peak,
                           _arrival(reference.times, ref, arrival_threshold),
                           _arrival(reference.times, sim, arrival_threshold))
