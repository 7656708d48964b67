"""Accuracy of a reduced-precision run against an FP64 run of the same scenario.

The FP32 precision form reports, against the FP64 result at the same
simulated time: L1 and L-infinity depth error, peak-depth error, flood-extent
intersection-over-union at a depth threshold (1 cm), and the relative
mass-balance error (total water volume).
"""
from __future__ import annotations

import numpy as np


def depth_metrics(h_test, h_ref, dx, dy, wet_threshold=0.01, vol_test=None, vol_ref=None):
    h_test = np.asarray(h_test, np.float64)
    h_ref = np.asarray(h_ref, np.float64)
    d = np.abs(h_test - h_ref)
    wet_t = h_test > wet_threshold
    wet_r = h_ref > wet_threshold
    union = np.count_nonzero(wet_t | wet_r)
    inter = np.count_nonzero(wet_t & wet_r)
    if vol_test is None:
        vol_test = float(np.sum(h_test)) * dx * dy
    if vol_ref is None:
        vol_ref = float(np.sum(h_ref)) * dx * dy
    return {
        "depth_L1_m": float(np.mean(d)),
        "depth_Linf_m": float(np.max(d)),
        "peak_depth_err_m": float(abs(np.max(h_test) - np.max(h_ref))),
        "extent_iou_1cm": (inter / union) if union else 1.0,
        "mass_balance_rel_err": abs(vol_test - vol_ref) / vol_ref if vol_ref else 0.0,
        "peak_depth_m": float(np.max(h_ref)),
        "wet_cells_ref": int(np.count_nonzero(wet_r)),
    }
