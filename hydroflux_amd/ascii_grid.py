"""ESRI ASCII grid input (.asc / .dem, optionally gzip-compressed).

Host-side reader of the rasters the reference's operational loops feed to
the step (solver-2d/src/ascii_grid.rs:101-160, read_ascii_grid): a
six-line header `ncols, nrows, xllcorner, yllcorner, cellsize,
NODATA_value` in that order (keys compared case-insensitively), then
`nrows * ncols` whitespace-separated values in row-major order, the first
data row being the north edge (array row 0). The header helpers map
world coordinates to (row, col) cells as the reference does
(ascii_grid.rs:48-74). Errors follow the reference's AsciiGridError kinds.
"""
from __future__ import annotations

import gzip
import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np


class AsciiGridError(ValueError):
    """Malformed ESRI ASCII grid (ascii_grid.rs AsciiGridError)."""


class AsciiGridHeaderError(AsciiGridError):
    """Missing, misordered or unparsable header line (AsciiGridError::Header)."""


class AsciiGridValueCountError(AsciiGridError):
    """Number of data values differs from nrows * ncols (AsciiGridError::ValueCount)."""


class AsciiGridParseError(AsciiGridError):
    """A token that is not a number (AsciiGridError::Parse)."""


def _as_usize(x: float) -> int:
    # Rust `f64 as usize`: truncation toward zero, saturating at 0, NaN -> 0
    if not math.isfinite(x):
        return 0 if (math.isnan(x) or x < 0) else 2**64 - 1
    return max(0, int(x))


@dataclass(frozen=True)
class AsciiGridHeader:
    ncols: int
    nrows: int
    xllcorner: float
    yllcorner: float
    cellsize: float
    nodata_value: float

    def cell_at(self, x: float, y: float) -> Optional[Tuple[int, int]]:
        """(row, col) of the cell containing world point (x, y), row 0 at the
        north edge; None outside the grid (ascii_grid.rs:48-60)."""
        col = (x - self.xllcorner) / self.cellsize
        row_from_south = (y - self.yllcorner) / self.cellsize
        if col < 0.0 or row_from_south < 0.0:
            return None
        col = _as_usize(math.floor(col))
        row_from_south = _as_usize(math.floor(row_from_south))
        if col >= self.ncols or row_from_south >= self.nrows:
            return None
        return (self.nrows - 1 - row_from_south, col)

    def rows_overlapping_y_range(self, y_min: float, y_max: float) -> List[int]:
        """Array rows whose y-extent overlaps [y_min, y_max), north-first
        numbering, in south-to-north order (ascii_grid.rs:66-74)."""
        lo = _as_usize(max(math.floor((y_min - self.yllcorner) / self.cellsize), 0.0))
        hi = min(_as_usize(math.ceil((y_max - self.yllcorner) / self.cellsize)), self.nrows)
        return [self.nrows - 1 - r for r in range(lo, hi)]


_KEYS = ("ncols", "nrows", "xllcorner", "yllcorner", "cellsize", "NODATA_value")


def _header_value(line: Optional[str], key: str) -> float:
    if line is None:
        raise AsciiGridHeaderError(f"missing {key}")
    parts = line.split()
    if not parts:
        raise AsciiGridHeaderError(f"empty line, expected {key}")
    if parts[0].lower() != key.lower():
        raise AsciiGridHeaderError(f"expected key {key!r}, got {parts[0]!r} (line: {line!r})")
    if len(parts) < 2:
        raise AsciiGridHeaderError(f"missing value for {key}")
    try:
        return float(parts[1])
    except ValueError:
        raise AsciiGridParseError(parts[1]) from None


def parse_ascii_grid(text: str) -> Tuple[np.ndarray, AsciiGridHeader]:
    lines = text.splitlines()
    vals = []
    for k, key in enumerate(_KEYS):
        vals.append(_header_value(lines[k] if k < len(lines) else None, key))
    ncols, nrows = _as_usize(vals[0]), _as_usize(vals[1])
    tokens = " ".join(lines[len(_KEYS):]).split()
    try:
        data = np.array([float(t) for t in tokens], dtype=np.float64)
    except ValueError:
        bad = next(t for t in tokens if not _is_number(t))
        raise AsciiGridParseError(bad) from None
    if data.size != ncols * nrows:
        raise AsciiGridValueCountError(f"expected {ncols * nrows} data values (ncols * nrows), found {data.size}")
    header = AsciiGridHeader(ncols, nrows, vals[2], vals[3], vals[4], vals[5])
    return data.reshape(nrows, ncols), header


def _is_number(t: str) -> bool:
    try:
        float(t)
        return True
    except ValueError:
        return False


def read_ascii_grid(path: str) -> Tuple[np.ndarray, AsciiGridHeader]:
    """Read an ESRI ASCII grid into a row-major (nrows, ncols) float64 array
    plus its header; gunzips when the path ends in .gz (ascii_grid.rs:101)."""
    if str(path).endswith(".gz"):
        with gzip.open(path, "rt") as f:
            text = f.read()
    else:
        with open(path) as f:
            text = f.read()
    return parse_ascii_grid(text)
