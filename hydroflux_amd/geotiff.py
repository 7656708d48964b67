"""GeoTIFF input and output for the step's meshes and depth fields.

Host-side counterpart of the reference's GeoTIFF bridge
(solver-2d/src/io.rs:84-240): a DEM raster becomes a Mesh2D with a uniform
Manning n (mesh_from_geotiff, io.rs:84-105) or with a Manning field mapped
from a land-cover raster (mesh_from_geotiff_with_landcover, io.rs:133-164;
ESA WorldCover table esa_worldcover_to_manning, io.rs:170-208), and a depth
field goes back out as a single-band float32 GeoTIFF with dry cells as
no-data (depth_raster_from_states / write_depth_geotiff, io.rs:210-240).

The reference reads and writes rasters through the surtgis crate, which
wraps libtiff and is not part of this repository; the TIFF codec here is
written from the TIFF 6.0 specification (baseline structure, LZW section
13, PackBits section 9, predictor 2 section 14), Adobe's Deflate and
floating-point predictor (predictor 3) technical notes, BigTIFF, and the
GeoTIFF 1.0 tags (ModelPixelScale, ModelTiepoint, ModelTransformation,
GTRasterTypeGeoKey) plus GDAL's GDAL_NODATA ASCII tag.

Conventions (as io.rs documents them): the affine transform maps
(col, row) to (x, y) with a negative pixel_height for north-up images;
array row 0 is the first raster row (the north edge); dx = |pixel_width|,
dy = |pixel_height|.
"""
from __future__ import annotations

import math
import struct
import zlib
from dataclasses import dataclass
from typing import Callable, Optional, Tuple

import numpy as np

from .solver2d import H_DRY, Mesh2D


# --------------------------------------------------------------- errors
class IoError(ValueError):
    """io.rs IoError: conditions caused by the user's files."""


class TiffFormatError(IoError):
    """Malformed or unsupported TIFF (surtgis' file-format errors, IoError::Surtgis)."""


class ShapeMismatch(IoError):
    """DEM and land-cover rasters differ in shape (IoError::ShapeMismatch)."""

    def __init__(self, dem, landcover):
        super().__init__(f"DEM shape {tuple(dem)} and landcover shape {tuple(landcover)} must match")
        self.dem = tuple(dem)
        self.landcover = tuple(landcover)


class InvalidPixelSize(IoError):
    """Pixel size from the geotransform not strictly positive (IoError::InvalidPixelSize)."""

    def __init__(self, dx, dy):
        super().__init__(f"pixel size must be strictly positive, got dx = {dx}, dy = {dy}")
        self.dx = dx
        self.dy = dy


class EmptyRaster(IoError):
    """The raster has no cells (IoError::EmptyRaster)."""


# --------------------------------------------------------------- raster
@dataclass(frozen=True)
class GeoTransform:
    """GDAL-style affine transform: x = origin_x + col * pixel_width + row * row_rotation,
    y = origin_y + col * col_rotation + row * pixel_height."""

    origin_x: float
    origin_y: float
    pixel_width: float
    pixel_height: float
    row_rotation: float = 0.0
    col_rotation: float = 0.0

    def cell_center(self, row: int, col: int) -> Tuple[float, float]:
        """Map coordinates of the centre of cell (row, col)
        (export_huasco_inputs.rs's inflow-point arithmetic)."""
        x = self.origin_x + (col + 0.5) * self.pixel_width + (row + 0.5) * self.row_rotation
        y = self.origin_y + (col + 0.5) * self.col_rotation + (row + 0.5) * self.pixel_height
        return x, y


@dataclass
class Raster:
    """A single-band raster: data[row, col], its transform and no-data value."""

    data: np.ndarray
    transform: GeoTransform
    nodata: Optional[float] = None

    @property
    def shape(self) -> Tuple[int, int]:
        return tuple(self.data.shape)


# --------------------------------------------------------------- TIFF tags
T_WIDTH, T_LENGTH, T_BITS, T_COMPRESSION, T_PHOTOMETRIC = 256, 257, 258, 259, 262
T_STRIP_OFFSETS, T_SPP, T_ROWS_PER_STRIP, T_STRIP_BYTES = 273, 277, 278, 279
T_PLANAR, T_PREDICTOR, T_TILE_W, T_TILE_L, T_TILE_OFFSETS, T_TILE_BYTES = 284, 317, 322, 323, 324, 325
T_SAMPLE_FORMAT = 339
T_PIXEL_SCALE, T_TIEPOINT, T_TRANSFORMATION, T_GEOKEYS = 33550, 33922, 34264, 34735
T_GDAL_NODATA = 42113
GT_RASTER_TYPE_KEY, PIXEL_IS_AREA, PIXEL_IS_POINT = 1025, 1, 2

# type code -> (struct format, size)
_TYPES = {1: ("B", 1), 2: ("c", 1), 3: ("H", 2), 4: ("I", 4), 5: ("I", 8), 6: ("b", 1), 7: ("B", 1),
          8: ("h", 2), 9: ("i", 4), 10: ("i", 8), 11: ("f", 4), 12: ("d", 8), 16: ("Q", 8), 17: ("q", 8),
          18: ("Q", 8)}


def _read_ifd(buf: bytes, bo: str, big: bool):
    """The first image file directory as {tag: tuple | str}."""
    if big:
        if len(buf) < 16 or struct.unpack(bo + "HH", buf[4:8]) != (8, 0):
            raise TiffFormatError("bad BigTIFF header")
        off = struct.unpack(bo + "Q", buf[8:16])[0]
        cfmt, csz, esz, vsz = "Q", 8, 20, 8
    else:
        off = struct.unpack(bo + "I", buf[4:8])[0]
        cfmt, csz, esz, vsz = "H", 2, 12, 4
    if off + csz > len(buf):
        raise TiffFormatError("IFD offset beyond end of file")
    n = struct.unpack(bo + cfmt, buf[off:off + csz])[0]
    p = off + csz
    if p + n * esz > len(buf):
        raise TiffFormatError("IFD entries beyond end of file")
    tags = {}
    for k in range(n):
        e = buf[p + k * esz:p + (k + 1) * esz]
        tag, typ = struct.unpack(bo + "HH", e[:4])
        cnt = struct.unpack(bo + ("Q" if big else "I"), e[4:4 + (8 if big else 4)])[0]
        if typ not in _TYPES:
            continue  # unknown types are skipped, as readers must
        fmt, sz = _TYPES[typ]
        total = sz * cnt
        vb = e[esz - vsz:]
        if total <= vsz:
            raw = vb[:total]
        else:
            o = struct.unpack(bo + ("Q" if big else "I"), vb)[0]
            if o + total > len(buf):
                raise TiffFormatError(f"tag {tag} data beyond end of file")
            raw = buf[o:o + total]
        if typ == 2:
            tags[tag] = raw.split(b"\0", 1)[0].decode("latin-1")
        elif typ in (5, 10):
            v = struct.unpack(bo + fmt * (2 * cnt), raw)
            tags[tag] = tuple(v[2 * i] / v[2 * i + 1] if v[2 * i + 1] else math.nan for i in range(cnt))
        else:
            tags[tag] = struct.unpack(bo + fmt * cnt, raw)
    return tags


def _lzw_decode(data: bytes, expected: int) -> bytes:
    """TIFF LZW (TIFF 6.0 section 13): codes packed most significant bit
    first, 9 to 12 bits wide; 256 clears the table, 257 ends the strip; the
    width grows one code early (when the next free code reaches 2^w - 1)."""
    if len(data) >= 2 and data[0] == 0 and (data[1] & 1):
        raise TiffFormatError("pre-6.0 (LSB-first) LZW is not supported")
    out = bytearray()
    table = [bytes((i,)) for i in range(256)] + [b"", b""]
    width, mask = 9, 511
    acc = nacc = 0
    prev = None
    pos, n = 0, len(data)
    while True:
        while nacc < width:
            if pos >= n:
                return bytes(out)  # no EOI: end of data
            acc = ((acc << 8) | data[pos]) & 0xFFFFFFFF
            pos += 1
            nacc += 8
        code = (acc >> (nacc - width)) & mask
        nacc -= width
        if code == 256:
            del table[258:]
            width, mask = 9, 511
            prev = None
            continue
        if code == 257:
            break
        if prev is None:
            if code >= 256:
                raise TiffFormatError("LZW: first code after a clear is not a literal")
            entry = table[code]
        elif code < len(table):
            entry = table[code]
            table.append(prev + entry[:1])
        elif code == len(table):
            entry = prev + prev[:1]
            table.append(entry)
        else:
            raise TiffFormatError("LZW: code beyond the table")
        out += entry
        if len(out) >= expected:
            break
        prev = entry
        if len(table) >= mask and width < 12:
            width += 1
            mask = (1 << width) - 1
    return bytes(out)


def _packbits_decode(data: bytes) -> bytes:
    """PackBits (TIFF 6.0 section 9)."""
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        c = data[i]
        i += 1
        if c < 128:
            out += data[i:i + c + 1]
            i += c + 1
        elif c > 128:
            if i >= n:
                break
            out += bytes((data[i],)) * (257 - c)
            i += 1
    return bytes(out)


def _decompress(raw: bytes, compression: int, expected: int) -> bytes:
    if compression == 1:
        return raw
    if compression == 5:
        return _lzw_decode(raw, expected)
    if compression in (8, 32946):
        try:
            return zlib.decompress(raw)
        except zlib.error as e:
            raise TiffFormatError(f"Deflate: {e}") from None
    if compression == 32773:
        return _packbits_decode(raw)
    raise TiffFormatError(f"unsupported TIFF compression {compression}")


def _dtype(bits: int, fmt: int, bo: str) -> np.dtype:
    kind = {1: "u", 2: "i", 3: "f"}.get(fmt)
    if kind is None or bits not in (8, 16, 32, 64) or (kind == "f" and bits < 32):
        raise TiffFormatError(f"unsupported sample type: {bits} bits, SampleFormat {fmt}")
    return np.dtype(f"{bo}{kind}{bits // 8}")


def _block_array(raw: bytes, rows: int, cols: int, spp: int, dt: np.dtype, predictor: int) -> np.ndarray:
    """One decoded strip or tile as a (rows, cols, spp) array."""
    nbytes = rows * cols * spp * dt.itemsize
    if len(raw) < nbytes:
        raw = raw + bytes(nbytes - len(raw))  # short final strips are zero-filled
    raw = raw[:nbytes]
    if predictor == 3:  # floating-point predictor: byte deltas, then per-byte planes MSB first
        b = np.frombuffer(raw, np.uint8).reshape(rows, cols * spp * dt.itemsize)
        b = np.cumsum(b, axis=1, dtype=np.uint8)
        planes = b.reshape(rows, dt.itemsize, cols * spp)
        be = np.ascontiguousarray(planes.transpose(0, 2, 1)).reshape(-1)
        return np.frombuffer(be.tobytes(), dt.newbyteorder(">")).reshape(rows, cols, spp)
    a = np.frombuffer(raw, dt).reshape(rows, cols, spp)
    if predictor == 2:  # horizontal differencing, modulo 2^bits (float samples: of their bit patterns, as libtiff)
        ut = np.dtype(f"{dt.byteorder}u{dt.itemsize}")
        a = np.cumsum(a.view(ut), axis=1, dtype=ut).view(dt)
    elif predictor != 1:
        raise TiffFormatError(f"unsupported TIFF predictor {predictor}")
    return a


def _geotransform(tags) -> GeoTransform:
    raster_type = PIXEL_IS_AREA
    keys = tags.get(T_GEOKEYS)
    if keys and len(keys) >= 4:
        for k in range(keys[3]):
            e = keys[4 + 4 * k:8 + 4 * k]
            if len(e) == 4 and e[0] == GT_RASTER_TYPE_KEY and e[1] == 0:
                raster_type = e[3]
    if T_TRANSFORMATION in tags:
        m = tags[T_TRANSFORMATION]
        if len(m) < 16:
            raise TiffFormatError("ModelTransformation needs 16 values")
        gt = GeoTransform(m[3], m[7], m[0], m[5], m[1], m[4])
    elif T_PIXEL_SCALE in tags and T_TIEPOINT in tags:
        sx, sy = tags[T_PIXEL_SCALE][:2]
        i, j, _, x, y, _ = tags[T_TIEPOINT][:6]
        gt = GeoTransform(x - i * sx, y + j * sy, sx, -sy)
    else:
        gt = GeoTransform(0.0, 0.0, 1.0, 1.0)  # no georeferencing: unit pixels
    if raster_type == PIXEL_IS_POINT:  # tie points name pixel centres: shift to the corner
        gt = GeoTransform(gt.origin_x - 0.5 * gt.pixel_width - 0.5 * gt.row_rotation,
                          gt.origin_y - 0.5 * gt.col_rotation - 0.5 * gt.pixel_height,
                          gt.pixel_width, gt.pixel_height, gt.row_rotation, gt.col_rotation)
    return gt


def parse_geotiff(buf: bytes, band: int = 0) -> Raster:
    """Decode one band of the first image of a (Geo)TIFF held in memory."""
    if len(buf) < 8 or buf[:2] not in (b"II", b"MM"):
        raise TiffFormatError("not a TIFF file")
    bo = "<" if buf[:2] == b"II" else ">"
    magic = struct.unpack(bo + "H", buf[2:4])[0]
    if magic not in (42, 43):
        raise TiffFormatError(f"bad TIFF magic {magic}")
    tags = _read_ifd(buf, bo, magic == 43)
    try:
        width, length = tags[T_WIDTH][0], tags[T_LENGTH][0]
    except KeyError:
        raise TiffFormatError("missing ImageWidth / ImageLength") from None
    spp = tags.get(T_SPP, (1,))[0]
    bits = tags.get(T_BITS, (1,))
    if len(set(bits)) != 1:
        raise TiffFormatError("samples of different sizes are not supported")
    fmt = tags.get(T_SAMPLE_FORMAT, (1,))
    dt = _dtype(bits[0], fmt[0], bo)
    compression = tags.get(T_COMPRESSION, (1,))[0]
    predictor = tags.get(T_PREDICTOR, (1,))[0]
    planar = tags.get(T_PLANAR, (1,))[0]
    if not 0 <= band < spp:
        raise TiffFormatError(f"band {band} out of range (the file has {spp})")
    bspp = 1 if planar == 2 else spp  # samples per block pixel
    out = np.zeros((length, width), dt.newbyteorder("="))
    if T_TILE_OFFSETS in tags:
        tw, tl = tags[T_TILE_W][0], tags[T_TILE_L][0]
        offs, cnts = tags[T_TILE_OFFSETS], tags[T_TILE_BYTES]
        across, down = -(-width // tw), -(-length // tl)
        per_band = across * down
        need = per_band * (spp if planar == 2 else 1)
        if len(offs) < need or len(cnts) < need:
            raise TiffFormatError("too few tiles")
        for ty in range(down):
            for tx in range(across):
                k = ty * across + tx + (band * per_band if planar == 2 else 0)
                raw = _decompress(buf[offs[k]:offs[k] + cnts[k]], compression, tl * tw * bspp * dt.itemsize)
                a = _block_array(raw, tl, tw, bspp, dt, predictor)
                r0, c0 = ty * tl, tx * tw
                r1, c1 = min(r0 + tl, length), min(c0 + tw, width)
                out[r0:r1, c0:c1] = a[:r1 - r0, :c1 - c0, 0 if planar == 2 else band]
    else:
        if T_STRIP_OFFSETS not in tags:
            raise TiffFormatError("neither strips nor tiles")
        rps = min(tags.get(T_ROWS_PER_STRIP, (length,))[0], length)
        offs, cnts = tags[T_STRIP_OFFSETS], tags.get(T_STRIP_BYTES)
        per_band = -(-length // rps)
        need = per_band * (spp if planar == 2 else 1)
        if cnts is None or len(offs) < need or len(cnts) < need:
            raise TiffFormatError("too few strips")
        for s in range(per_band):
            k = s + (band * per_band if planar == 2 else 0)
            r0 = s * rps
            rows = min(rps, length - r0)
            raw = _decompress(buf[offs[k]:offs[k] + cnts[k]], compression, rows * width * bspp * dt.itemsize)
            a = _block_array(raw, rows, width, bspp, dt, predictor)
            out[r0:r0 + rows] = a[:, :, 0 if planar == 2 else band]
    nodata = None
    if T_GDAL_NODATA in tags:
        try:
            nodata = float(tags[T_GDAL_NODATA].strip())
        except ValueError:
            raise TiffFormatError(f"GDAL_NODATA is not a number: {tags[T_GDAL_NODATA]!r}") from None
    return Raster(out, _geotransform(tags), nodata)


def read_geotiff(path, band: int = 0) -> Raster:
    """Read one band of a GeoTIFF (surtgis_core::io::read_geotiff). The data
    keep the file's sample type; callers convert (a float32 DEM widens to
    f64 exactly)."""
    with open(path, "rb") as f:
        return parse_geotiff(f.read(), band)


# --------------------------------------------------------------- writer
def write_geotiff(raster: Raster, path, dtype=None) -> None:
    """Write a single-band, uncompressed GeoTIFF (surtgis_core::io::write_geotiff
    with its default options): floating-point data as float32, integer data
    in their own type; georeferencing as ModelPixelScale + ModelTiepoint
    (ModelTransformation when the transform is rotated), PixelIsArea; the
    no-data value, if any, as GDAL_NODATA."""
    data = np.asarray(raster.data)
    if data.ndim != 2:
        raise IoError("a raster is two-dimensional")
    if dtype is None:
        dtype = np.float32 if data.dtype.kind == "f" else data.dtype
    dt = np.dtype(dtype).newbyteorder("<")
    if dt.kind not in "uif":
        raise IoError(f"cannot write samples of type {dt}")
    arr = np.ascontiguousarray(data.astype(dt))
    length, width = arr.shape
    row_bytes = width * dt.itemsize
    rps = max(1, min(length, (1 << 16) // max(row_bytes, 1)))
    strips = [arr[r:r + rps].tobytes() for r in range(0, length, rps)]
    gt = raster.transform
    entries = [
        (T_WIDTH, 4, [width]), (T_LENGTH, 4, [length]), (T_BITS, 3, [dt.itemsize * 8]),
        (T_COMPRESSION, 3, [1]), (T_PHOTOMETRIC, 3, [1]), (T_STRIP_OFFSETS, 4, [0] * len(strips)),
        (T_SPP, 3, [1]), (T_ROWS_PER_STRIP, 4, [rps]), (T_STRIP_BYTES, 4, [len(s) for s in strips]),
        (T_PLANAR, 3, [1]), (T_SAMPLE_FORMAT, 3, [{"u": 1, "i": 2, "f": 3}[dt.kind]]),
    ]
    if gt.row_rotation == 0.0 and gt.col_rotation == 0.0:
        entries += [(T_PIXEL_SCALE, 12, [gt.pixel_width, -gt.pixel_height, 0.0]),
                    (T_TIEPOINT, 12, [0.0, 0.0, 0.0, gt.origin_x, gt.origin_y, 0.0])]
    else:
        entries += [(T_TRANSFORMATION, 12, [gt.pixel_width, gt.row_rotation, 0.0, gt.origin_x,
                                            gt.col_rotation, gt.pixel_height, 0.0, gt.origin_y,
                                            0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0])]
    entries.append((T_GEOKEYS, 3, [1, 1, 0, 1, GT_RASTER_TYPE_KEY, 0, 1, PIXEL_IS_AREA]))
    if raster.nodata is not None:
        entries.append((T_GDAL_NODATA, 2, repr(float(raster.nodata)).encode() + b"\0"))
    entries.sort(key=lambda e: e[0])
    fmt = {2: "c", 3: "H", 4: "I", 12: "d"}
    size = {2: 1, 3: 2, 4: 4, 12: 8}

    def payload(typ, vals):
        return vals if typ == 2 else struct.pack("<" + fmt[typ] * len(vals), *vals)

    ifd_off = 8
    ifd_len = 2 + 12 * len(entries) + 4
    extra_off = ifd_off + ifd_len
    blobs = []
    for tag, typ, vals in entries:
        n = len(vals)
        if size[typ] * n > 4:
            blobs.append((tag, extra_off))
            extra_off += size[typ] * n + (size[typ] * n) % 2
    data_off = extra_off
    offsets = []
    for s in strips:
        offsets.append(data_off)
        data_off += len(s)
    if data_off >= 1 << 32:
        raise IoError("raster too large for a classic TIFF")
    entries = [(t, ty, offsets if t == T_STRIP_OFFSETS else v) for t, ty, v in entries]
    out = bytearray(b"II" + struct.pack("<HI", 42, ifd_off))
    out += struct.pack("<H", len(entries))
    extra = bytearray()
    for tag, typ, vals in entries:
        n = len(vals)
        p = payload(typ, vals)
        if len(p) <= 4:
            out += struct.pack("<HHI", tag, typ, n) + p + bytes(4 - len(p))
        else:
            out += struct.pack("<HHI", tag, typ, n) + struct.pack("<I", ifd_off + ifd_len + len(extra))
            extra += p + bytes(len(p) % 2)
    out += struct.pack("<I", 0)
    out += extra
    for s in strips:
        out += s
    with open(path, "wb") as f:
        f.write(bytes(out))


# --------------------------------------------------------------- mesh bridge
def _pixel_size(t: GeoTransform) -> Tuple[float, float]:
    dx, dy = abs(t.pixel_width), abs(t.pixel_height)
    if not (dx > 0.0 and dy > 0.0):
        raise InvalidPixelSize(dx, dy)
    return dx, dy


def mesh_from_geotiff(path, manning: float) -> Tuple[Mesh2D, GeoTransform]:
    """io.rs:84-105: a DEM GeoTIFF as a Mesh2D with uniform Manning n, and
    the DEM's transform."""
    dem = read_geotiff(path)
    dx, dy = _pixel_size(dem.transform)
    bed = dem.data.astype(np.float64)
    if bed.size == 0:
        raise EmptyRaster("raster is empty")
    return Mesh2D.new(bed, dx, dy, manning), dem.transform


def _as_u8(a: np.ndarray) -> np.ndarray:
    """Rust `as u8` on the land-cover samples: truncation toward zero,
    saturating at 0 and 255, NaN -> 0."""
    if a.dtype == np.uint8:
        return a
    f = np.nan_to_num(np.asarray(a, np.float64), nan=0.0, posinf=255.0, neginf=0.0)
    return np.clip(np.trunc(f), 0, 255).astype(np.uint8)


def mesh_from_geotiff_with_landcover(dem_path, landcover_path,
                                     landcover_to_manning: Callable[[int], float]) -> Tuple[Mesh2D, GeoTransform]:
    """io.rs:133-164: DEM + land-cover GeoTIFFs as a Mesh2D whose Manning n
    at each cell is landcover_to_manning(code); the DEM's transform is the
    grid's."""
    dem = read_geotiff(dem_path)
    lc = read_geotiff(landcover_path)
    if dem.shape != lc.shape:
        raise ShapeMismatch(dem.shape, lc.shape)
    dx, dy = _pixel_size(dem.transform)
    bed = dem.data.astype(np.float64)
    if bed.size == 0:
        raise EmptyRaster("raster is empty")
    codes = _as_u8(lc.data)
    lut = np.array([float(landcover_to_manning(c)) for c in range(256)])
    return Mesh2D.with_manning_field(bed, dx, dy, lut[codes]), dem.transform


ESA_WORLDCOVER_MANNING = {10: 0.100, 20: 0.060, 30: 0.040, 40: 0.035, 50: 0.015, 60: 0.025, 70: 0.030,
                          80: 0.030, 90: 0.050, 95: 0.100, 100: 0.045}
ESA_WORLDCOVER_FALLBACK = 0.040


def esa_worldcover_to_manning(code: int) -> float:
    """io.rs:170-208: ESA WorldCover class -> Manning n (fallback 0.040)."""
    return ESA_WORLDCOVER_MANNING.get(int(code), ESA_WORLDCOVER_FALLBACK)


def depth_raster_from_states(states, transform: GeoTransform, nodata: Optional[float] = None) -> Raster:
    """io.rs:210-233: the depth band of a state array (rows, cols, 3); with a
    no-data value, cells with h <= H_DRY are written as it."""
    h = np.asarray(states, np.float64)[..., 0]
    if nodata is not None:
        h = np.where(h <= H_DRY, float(nodata), h)
    return Raster(np.array(h), transform, None if nodata is None else float(nodata))


def write_depth_geotiff(path, states, transform: GeoTransform, nodata: Optional[float] = None) -> None:
    """io.rs:235-240: depth_raster_from_states written as float32."""
    write_geotiff(depth_raster_from_states(states, transform, nodata), path)
