"""UK Environment Agency 2D benchmark Test 4 (Neelz & Pender 2013, report
SC120002 section 4.5) on its official 5 m geometry, set up as the
reference's operational loop runs it (solver-2d/examples/
uk_ea_test4_official.rs):

- flat 400 x 200 DEM (ESRI ASCII `ea4-5m.dem.gz`), Manning n = 0.05;
- west edge a wall with a 20 m inlet (y in [990, 1010]) fed by one point
  source per overlapping edge row; the trapezoidal hydrograph (0 until
  300 s, ramp to 20 m^3/s at 3600 s, hold to 14400 s, ramp to 0 at 18000 s)
  is split evenly over the inlet rows, q = (frac(t_new) * 20) / n_rows;
- north, south and east edges transmissive (FREE);
- Simulation driver: SSP-RK2, CFL 0.4 with boundary ghosts, friction with
  H_DRY, max_dt 15 s, run until 18000 s; the sources use Q(t_new) after
  each step;
- depths recorded after every step at the six official control points.

Here the hydrograph is a device-side forcing series and the control-point
depths are device-side gauges, so the whole run stays on the GPU.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .ascii_grid import read_ascii_grid
from .solver2d import Grid2D

MANNING_N = 0.05
CFL = 0.4
MAX_DT = 15.0
SIM_TIME = 18_000.0
INLET_Y = (990.0, 1010.0)
PEAK_Q_M3S = 20.0
HYDROGRAPH = ((0.0, 0.0), (300.0, 0.0), (3600.0, 1.0), (14400.0, 1.0), (18000.0, 0.0))
CONTROL_POINTS = ((50.0, 1000.0), (100.0, 1000.0), (200.0, 1000.0), (300.0, 1000.0), (400.0, 1000.0),
                  (300.0, 1300.0))
TRANSMISSIVE, WALL = 0, 1


def ea_test4_grid(dem_path: str, kind: str = "f64", gauge_capacity: int = 60_000) -> Tuple[Grid2D, List[Tuple[int, int]]]:
    """A Grid2D set up for Test 4 at t = 0 (gauges armed), and the control
    cells (row, col)."""
    bed, hdr = read_ascii_grid(dem_path)
    if not np.all(bed == 0.0):
        raise ValueError("Test 4 bed must be flat (report section 4.5.1)")
    g = Grid2D(hdr.nrows, hdr.ncols, hdr.cellsize, hdr.cellsize, kind=kind)
    g.set_mesh(bed, MANNING_N)
    g.set_state(np.zeros((hdr.nrows, hdr.ncols, 3)))
    g.set_boundaries([TRANSMISSIVE, TRANSMISSIVE, WALL, TRANSMISSIVE])  # N, S, W, E
    rows = hdr.rows_overlapping_y_range(*INLET_Y)
    if not rows:
        raise ValueError("inlet y-range fell outside the grid")
    g.configure(cfl=CFL, max_dt=MAX_DT, sources=[(r, 0, 0.0) for r in rows])
    t, f = zip(*HYDROGRAPH)
    g.set_forcing_series("sources", t, f, time_div=1.0, scale_mul=PEAK_Q_M3S, scale_div=float(len(rows)))
    cells = []
    for x, y in CONTROL_POINTS:
        c = hdr.cell_at(x, y)
        if c is None:
            raise ValueError(f"control point ({x}, {y}) falls outside the grid")
        cells.append(c)
    g.set_gauges(cells, gauge_capacity)
    return g, cells


# --------------------------------------------------------------- Test 8A
# UK EA Test 8A (report SC120002 section 4.9; Glasgow urban, official 2 m
# geometry) as solver-2d/examples/uk_ea_test8a_official.rs runs it: 481 x 199
# DEM with buildings as raised bed, Manning from the official raster (0.02
# roads, 0.05 elsewhere), all edges walls, uniform rain 400 mm/h over
# t in [1, 4] min, a point inflow QVAR(t) * cellsize at (264894, 664750),
# CFL 0.4, max_dt 10 s, 18000 s; rain(t_new) then the inflow after each step.
T8_MAX_DT = 10.0
T8_POINT_XY = (264_894.0, 664_750.0)
T8_RAIN_MIN = ((0.0, 0.0), (0.9999, 0.0), (1.0, 400.0), (4.0, 400.0), (4.0001, 0.0), (300.0, 0.0))
T8_POINT_MIN = ((0.0, 0.0), (20.0, 0.0), (23.0, 0.03125), (25.0, 0.125), (27.0, 0.3125), (29.0, 0.65625),
                (31.0, 1.21875), (33.0, 1.78125), (35.0, 2.3125), (37.0, 2.5), (39.0, 2.5), (41.0, 2.3125),
                (43.0, 1.78125), (45.0, 1.21875), (47.0, 0.65625), (49.0, 0.3125), (51.0, 0.125),
                (53.0, 0.03125), (55.0, 0.0), (57.0, 0.0), (75.0, 0.0), (120.0, 0.0), (300.0, 0.0))
T8_CONTROL_POINTS = ((264_680.0, 664_582.0), (264_536.0, 664_668.0), (264_354.0, 664_490.0),
                     (264_200.0, 664_556.0), (264_332.0, 664_564.0), (264_572.0, 664_556.0),
                     (264_708.0, 664_702.0), (264_306.0, 664_650.0), (264_220.0, 664_614.0))


def ea_test8a_grid(dem_path: str, manning_path: str, kind: str = "f64",
                   gauge_capacity: int = 150_000) -> Tuple[Grid2D, List[Tuple[int, int]]]:
    """A Grid2D set up for Test 8A at t = 0 (gauges armed), and the control
    cells (row, col)."""
    bed, hdr = read_ascii_grid(dem_path)
    man, mhdr = read_ascii_grid(manning_path)
    if (mhdr.ncols, mhdr.nrows) != (hdr.ncols, hdr.nrows):
        raise ValueError("DEM and Manning grids must share geometry")
    if not np.all(np.isfinite(bed)):
        raise ValueError("Test 8A DEM has no NODATA cells")
    g = Grid2D(hdr.nrows, hdr.ncols, hdr.cellsize, hdr.cellsize, kind=kind)
    g.set_mesh(bed, man)
    g.set_state(np.zeros((hdr.nrows, hdr.ncols, 3)))
    g.set_boundaries([WALL, WALL, WALL, WALL])
    cell = hdr.cell_at(*T8_POINT_XY)
    if cell is None:
        raise ValueError("point inflow falls outside the grid")
    g.configure(cfl=CFL, max_dt=T8_MAX_DT, sources=[(cell[0], cell[1], 0.0)])
    t, v = zip(*T8_RAIN_MIN)
    g.set_forcing_series("rain", t, v, time_div=60.0, scale_mul=1e-3, scale_div=3600.0)
    t, v = zip(*T8_POINT_MIN)
    g.set_forcing_series("sources", t, v, time_div=60.0, scale_mul=hdr.cellsize, scale_div=1.0)
    cells = [hdr.cell_at(x, y) for x, y in T8_CONTROL_POINTS]
    if any(c is None for c in cells):
        raise ValueError("a control point falls outside the grid")
    g.set_gauges(cells, gauge_capacity)
    return g, cells
