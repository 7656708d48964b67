"""ESRI ASCII grid input, LISFLOOD-FP stage files, and the driver's forcing
series and gauges, on the CPU.

The reader follows solver-2d/src/ascii_grid.rs; its unit tests' sample and
expectations are restated here. The forcing series of the oracle's driver
(evaluated at t_new, scaled as the reference's UK EA loops scale it) is
checked bit for bit against a literal transcription of those loops: step,
then apply_rain(rate(t_new)) and apply_point_sources(Q(t_new)) from the
host with the loop's own arithmetic.
"""
import gzip
import os

import numpy as np
import pytest

import cases
from hydroflux_amd.ascii_grid import (AsciiGridHeaderError, AsciiGridParseError, AsciiGridValueCountError,
                                      parse_ascii_grid, read_ascii_grid)
from hydroflux_amd.stage import interp, parse_reference_stage
from oracle.pyoracle import OracleSim, set_pow_mode, set_threads

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "uk_ea", "test4")

# ascii_grid.rs tests: SAMPLE
SAMPLE = """ncols         4
nrows         3
xllcorner     0.0
yllcorner     0.0
cellsize      10
NODATA_value  -9999
1 2 3 4
5 6 7 8
9 10 11 12
"""


def test_parses_header_fields():
    _, h = parse_ascii_grid(SAMPLE)
    assert (h.ncols, h.nrows, h.cellsize, h.nodata_value) == (4, 3, 10.0, -9999.0)


def test_row_zero_is_the_first_data_row_after_header():
    data, _ = parse_ascii_grid(SAMPLE)
    assert list(data[0]) == [1, 2, 3, 4] and list(data[2]) == [9, 10, 11, 12]


def test_rejects_wrong_value_count_and_out_of_order_header_and_bad_token():
    with pytest.raises(AsciiGridValueCountError):
        parse_ascii_grid(SAMPLE.replace("9 10 11 12\n", ""))
    with pytest.raises(AsciiGridHeaderError):
        parse_ascii_grid(SAMPLE.replace("ncols         4\n", "", 1))
    with pytest.raises(AsciiGridParseError):
        parse_ascii_grid(SAMPLE.replace("7", "seven"))


def test_wrapped_rows_and_case_insensitive_keys():
    text = SAMPLE.replace("NODATA_value", "nodata_VALUE").replace("5 6 7 8\n", "5 6\n7 8\n")
    data, h = parse_ascii_grid(text)
    assert h.nodata_value == -9999.0 and list(data[1]) == [5, 6, 7, 8]


def test_cell_at_and_rows_overlapping_y_range():
    _, h = parse_ascii_grid(SAMPLE)
    assert h.cell_at(5.0, 5.0) == (2, 0)       # south-west cell -> last array row
    assert h.cell_at(35.0, 25.0) == (0, 3)
    assert h.cell_at(-1.0, 5.0) is None and h.cell_at(5.0, 1000.0) is None
    assert sorted(h.rows_overlapping_y_range(12.0, 22.0)) == [0, 1]


def test_reads_the_official_test4_dem(tmp_path):
    bed, h = read_ascii_grid(os.path.join(GOLDEN, "ea4-5m.dem.gz"))
    assert bed.shape == (400, 200) and h.cellsize == 5.0 and np.all(bed == 0.0)
    assert h.rows_overlapping_y_range(990.0, 1010.0) == [201, 200, 199, 198]
    assert h.cell_at(50.0, 1000.0) == (199, 10) and h.cell_at(300.0, 1300.0) == (139, 60)
    # plain (not gzipped) file gives the same grid
    p = tmp_path / "ea4.asc"
    p.write_text(gzip.open(os.path.join(GOLDEN, "ea4-5m.dem.gz"), "rt").read())
    bed2, h2 = read_ascii_grid(str(p))
    assert h2 == h and np.array_equal(bed2, bed)


def test_reference_stage_files_parse():
    for name, n in (("ea4-5m-dg2.stage", 6), ("ea4-1m-acc.stage", 6)):
        r = parse_reference_stage(os.path.join(GOLDEN, name))
        assert len(r.points) == n and r.points[0] == (50.0, 1000.0) and r.points[5] == (300.0, 1300.0)
        assert r.times[0] == 5.0 and r.times[-1] > 17_900 and all(len(row) == n for row in r.depths)
        assert np.all(np.diff(r.times) > 0)


def test_interp_clamps_and_is_linear():
    t, v = [0.0, 300.0, 3600.0, 14400.0, 18000.0], [0.0, 0.0, 1.0, 1.0, 0.0]
    assert interp(t, v, -5.0) == 0.0 and interp(t, v, 20000.0) == 0.0
    assert interp(t, v, 300.0) == 0.0 and interp(t, v, 3600.0) == 1.0
    assert interp(t, v, 1950.0) == (1950.0 - 300.0) / (3600.0 - 300.0)
    assert interp(t, v, 16200.0) == 0.5


def _loop_case():
    c = cases.mostly_dry(30, 44, 4, west=cases.WALL)
    c.bc_kinds = (cases.TRANS, cases.TRANS, cases.WALL, cases.TRANS)
    return c


@pytest.mark.parametrize("dual", [False, True], ids=["f64", "dual"])
def test_oracle_series_equal_the_reference_loops(dual):
    """Device-style series in the oracle's driver == the UK EA loops written
    out on the host (uk_ea_test4_official.rs, uk_ea_test8a_official.rs)."""
    set_threads(4)
    set_pow_mode(1)
    try:
        c = _loop_case()
        kw = {}
        if dual:
            kw = dict(manning_tangent=c.manning)
        hyd_t, hyd_f = (0.0, 0.5, 2.0, 4.0, 6.0), (0.0, 0.0, 1.0, 1.0, 0.0)      # hydrograph [s]
        rain_t, rain_v = (0.0, 0.009, 0.01, 0.05, 0.0501, 1.0), (0.0, 0.0, 400.0, 400.0, 0.0, 0.0)  # [min, mm/h]
        inlet = [(14, 0), (15, 0), (16, 0)]
        peak = 0.3
        a = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=dual), **kw)
        a.configure(cfl=0.4, max_dt=0.5, sources=[(r, cc, 0.0) for r, cc in inlet])
        a.set_forcing_series("sources", hyd_t, hyd_f, 1.0, peak, float(len(inlet)))
        a.set_forcing_series("rain", rain_t, rain_v, 60.0, 1e-3, 3600.0)
        a.set_gauges([(15, 1), (20, 30)])
        assert a.run_until(6.5) == 0

        b = c.load(OracleSim(c.nr, c.nc, c.dx, c.dy, dual=dual), **kw)
        b.configure(cfl=0.4, max_dt=0.5)
        t = 0.0
        hs = []
        while t < 6.5:
            rc, dt = b.step(6.5 - t)
            assert rc == 0
            t = b.time()[0]
            rain = interp(rain_t, rain_v, t / 60.0) * 1e-3 / 3600.0
            b.apply_rain(rain, dt)
            q = interp(hyd_t, hyd_f, t) * peak / len(inlet)
            b.apply_point_sources([(r, cc, q) for r, cc in inlet], dt)
            st = b.get_state()
            hs.append((t, (st[0] if dual else st)[15, 1, 0]))
        assert a.time() == b.time()
        sa, sb = a.get_state(), b.get_state()
        if dual:
            assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1], equal_nan=True)
        else:
            assert np.array_equal(sa, sb)
        gt, gh = a.read_gauges()
        assert gt[0] == 0.0 and len(gt) == len(hs) + 1
        assert np.array_equal(gt[1:], [x[0] for x in hs]) and np.array_equal(gh[1:, 0], [x[1] for x in hs])
    finally:
        set_pow_mode(0)
