"""The Huasco reach (200 x 67 cells at 30 m) from GeoTIFF inputs.

Builders for the reference's sealed-domain runs on the real terrain of its
application chapter:

- huasco_closed_domain.rs: DEM + ESA WorldCover land cover
  (mesh_from_geotiff_with_landcover), warm start h_n = Manning normal depth
  of Q_DAY1 on cells whose flow accumulation exceeds ACC_THRESHOLD, every
  edge a wall, CFL 0.4 with the boundary-aware dt capped by the time left,
  SSP-RK2, friction with dry tolerance 1e-9, and the inflow point source
  (q = 0 with --no-inflow) for one day;
- huasco_calibrate_tree.rs: the same sealed, source-free day with a fixed
  warm depth H_WARM, the tree class's Manning n as the Dual variable and no
  point source; its objective is the squared depth misfit against the
  SynxFlow field over cells wet in either field.

xval_metrics() computes the cross-validation measures of the reference's
comparison with SynxFlow (papers/01_review/manuscript.md:1009-1031) on the
wet mask h > H_THR in either field.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Tuple

import numpy as np

from .geotiff import GeoTransform, esa_worldcover_to_manning, mesh_from_geotiff_with_landcover, read_geotiff
from .solver2d import Grid2D, Mesh2D

# huasco_closed_domain.rs:39-49
N_REF = 0.04
ACC_THRESHOLD = 1_000_000.0
SLOPE_MEAN = 0.0074
CFL = 0.4
SECONDS_PER_DAY = 86_400.0
INFLOW_ROW, INFLOW_COL = 135, 66
Q_DAY1 = 17.5
FRICTION_DRY_TOL = 1.0e-9
WALL = 1
# huasco_calibrate_tree.rs:64-70
H_WARM = 0.457
H_THR = 0.01
N_TREE_INIT = 0.060
TREE_CLASS = 10


def manning_normal_depth(q_m3s: float, n: float, slope: float, cell_width_m: float) -> float:
    """huasco_closed_domain.rs:51-54: (n (q / w) / sqrt(S))^(3/5), with the
    platform pow as the reference's powf."""
    q_per_w = q_m3s / cell_width_m
    return (n * q_per_w / math.sqrt(slope)) ** (3.0 / 5.0)


@dataclass
class HuascoInputs:
    mesh: Mesh2D
    transform: GeoTransform
    codes: np.ndarray  # land-cover classes (uint8)
    acc: np.ndarray  # flow accumulation (float64)


def load_inputs(dem_path, landcover_path, acc_path) -> HuascoInputs:
    mesh, transform = mesh_from_geotiff_with_landcover(dem_path, landcover_path, esa_worldcover_to_manning)
    codes = read_geotiff(landcover_path).data.astype(np.uint8)
    acc = read_geotiff(acc_path).data.astype(np.float64)
    return HuascoInputs(mesh, transform, codes, acc)


def warm_state(inp: HuascoInputs, h_warm: float) -> np.ndarray:
    """h_warm on the channel cells (flow accumulation > ACC_THRESHOLD), dry elsewhere."""
    st = np.zeros(inp.mesh.bed.shape + (3,))
    st[..., 0] = np.where(inp.acc > ACC_THRESHOLD, h_warm, 0.0)
    return st


def closed_domain_grid(inp: HuascoInputs, no_inflow: bool = True, kind: str = "f64") -> Tuple[Grid2D, float]:
    """huasco_closed_domain.rs on the GPU, configured for run_until(SECONDS_PER_DAY).
    Returns the grid and the warm-start depth."""
    m = inp.mesh
    h_warm = manning_normal_depth(Q_DAY1, N_REF, SLOPE_MEAN, m.dx)
    g = Grid2D(m.n_rows, m.n_cols, m.dx, m.dy, kind=kind)
    g.set_mesh(m.bed, m.manning)
    g.set_state(warm_state(inp, h_warm))
    g.set_boundaries([WALL] * 4)
    g.configure(cfl=CFL, max_dt=math.inf, friction=True, friction_dry_tol=FRICTION_DRY_TOL,
                sources=[(INFLOW_ROW, INFLOW_COL, 0.0 if no_inflow else Q_DAY1)])
    return g, h_warm


def tree_manning(codes: np.ndarray, n_tree: float) -> Tuple[np.ndarray, np.ndarray]:
    """huasco_calibrate_tree.rs:75-90: the land-cover table with the tree
    class's n free; returns (n field, its tangent d n / d n_tree)."""
    lut = np.array([esa_worldcover_to_manning(c) for c in range(256)])
    lut[TREE_CLASS] = n_tree
    return lut[codes], (codes == TREE_CLASS).astype(np.float64)


def calibration_grid(inp: HuascoInputs, n_tree: float, kind: str = "dual") -> Grid2D:
    """huasco_calibrate_tree.rs:122-145 (forward::<Dual>) on the GPU,
    configured for run_until(SECONDS_PER_DAY)."""
    m = inp.mesh
    man, man_t = tree_manning(inp.codes, n_tree)
    g = Grid2D(m.n_rows, m.n_cols, m.dx, m.dy, kind=kind)
    g.set_mesh(m.bed, man, manning_tangent=man_t if kind == "dual" else None)
    g.set_state(warm_state(inp, H_WARM))
    g.set_boundaries([WALL] * 4)
    g.configure(cfl=CFL, max_dt=math.inf, friction=True, friction_dry_tol=FRICTION_DRY_TOL)
    return g


def calibration_terms(h: np.ndarray, dh: np.ndarray, target: np.ndarray) -> Tuple[float, float, float]:
    """huasco_calibrate_tree.rs:166-178: loss, dJ/dn and the Gauss-Newton
    curvature, accumulated over the cells in row-major order, skipping cells
    dry (<= H_THR) in both fields."""
    loss = grad = curv = 0.0
    for hv, dv, tv in zip(h.ravel().tolist(), dh.ravel().tolist(), target.ravel().tolist()):
        if hv <= H_THR and tv <= H_THR:
            continue
        r = hv - tv
        loss += r * r
        grad += 2.0 * r * dv
        curv += 2.0 * dv * dv
    return loss, grad, curv


def xval_metrics(h: np.ndarray, other: np.ndarray) -> dict:
    """Depth agreement on the wet mask (h > H_THR in either field):
    RMSE, MAE, bias (other - h), peaks, the Critical Success Index of the two
    wet masks, and the total-volume difference of `other` relative to h."""
    a, b = h > H_THR, other > H_THR
    mask = a | b
    r = other[mask] - h[mask]
    return {
        "cells": int(mask.sum()),
        "rmse": float(np.sqrt(np.mean(r * r))),
        "mae": float(np.mean(np.abs(r))),
        "bias": float(np.mean(r)),
        "peak": float(h.max()),
        "peak_other": float(other.max()),
        "csi": float((a & b).sum() / mask.sum()),
        "volume_rel": float((other.sum() - h.sum()) / h.sum()),
    }
