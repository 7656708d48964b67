"""hf2d_total_volume_noalloc against hf2d_total_volume (GPU, through the C ABI).

The allocation-free variant keeps its partial sums in a library-owned
device array shared by every call on the GPU; two calls in a row, the second
on a smaller grid, must each give the allocating entry point's bits (the
second call must not see the first call's partials), for the f64, FP32 and
Dual kinds.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _grid(kind, nr, nc, seed):
    from hydroflux_amd import Grid2D

    rng = np.random.default_rng(seed)
    g = Grid2D(nr, nc, 2.0, 3.0, kind=kind)
    st = np.zeros((nr, nc, 3))
    st[..., 0] = rng.uniform(0.0, 2.0, (nr, nc)) * (rng.random((nr, nc)) < 0.8)
    g.set_mesh(np.zeros((nr, nc)), np.full((nr, nc), 0.03))
    if kind == "dual":
        g.set_state(st, rng.standard_normal((nr, nc, 3)))
    else:
        g.set_state(st)
    return g, st


def _volumes(g):
    from hydroflux_amd._lib import lib

    a, b = C.c_double(), C.c_double()
    assert lib().hf2d_total_volume(g._h, C.byref(a)) == 0
    assert lib().hf2d_total_volume_noalloc(g._h, C.byref(b)) == 0
    return a.value, b.value


@pytest.mark.parametrize("kind", ["f64", "f32", "dual"])
def test_total_volume_noalloc_equals_allocating_entry_point(kind):
    _say(f"total_volume_noalloc {kind}: 700 x 900, then 37 x 11")
    big, st_big = _grid(kind, 700, 900, 1)
    small, st_small = _grid(kind, 37, 11, 2)
    a1, b1 = _volumes(big)
    a2, b2 = _volumes(small)
    assert a1 == b1 and a2 == b2
    exact = [float(np.sum(s[..., 0].astype(np.float32 if kind == "f32" else np.float64), dtype=np.float64)) * 6.0
             for s in (st_big, st_small)]
    assert abs(b1 - exact[0]) <= 1e-12 * exact[0] and abs(b2 - exact[1]) <= 1e-12 * exact[1]
    assert b2 != b1
    big.close()
    small.close()
