"""Tilt-series file formats the reference's caller reads and writes.

* MRC (mode 2, float32; 1024-byte header + ``nsymbt`` bytes of extended
  header), read as ``data[section][row][column]`` = (tilts, slices, bins),
  as ``mrcfile`` presents it to h5utilities.readh5 (h5utilities.py:369-373).
* ``.rawtlt`` (whitespace-separated degrees) or ``.csv`` (first column,
  degrees) angle files next to the MRC file, converted to radians
  (h5utilities.py:378-399).
* ``write_mrc`` writes a mode-2 volume the way the caller stores each
  reconstructed channel (Reconstruction_coupled.py:487-507).
"""
from __future__ import annotations

import csv
import os
import struct

import numpy as np

_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}


def read_mrc(path: str) -> np.ndarray:
    """Return the MRC volume as an array shaped (nz, ny, nx)."""
    with open(path, "rb") as fh:
        head = fh.read(1024)
        if len(head) < 1024:
            raise IOError(f"{path}: truncated MRC header")
        nx, ny, nz, mode = struct.unpack("<4i", head[:16])
        byteorder = "<"
        if not (0 < nx < 1 << 20 and 0 < ny < 1 << 20 and 0 < nz < 1 << 20):
            nx, ny, nz, mode = struct.unpack(">4i", head[:16])
            byteorder = ">"
        if mode not in _MODES:
            raise IOError(f"{path}: unsupported MRC mode {mode}")
        nsymbt = struct.unpack(byteorder + "i", head[92:96])[0]
        fh.seek(1024 + max(nsymbt, 0))
        dt = np.dtype(_MODES[mode]).newbyteorder(byteorder)
        data = np.fromfile(fh, dtype=dt, count=nx * ny * nz)
    if data.size != nx * ny * nz:
        raise IOError(f"{path}: expected {nx * ny * nz} values, found {data.size}")
    return data.reshape(nz, ny, nx).astype(dt.newbyteorder("="), copy=False)


def write_mrc(path: str, data: np.ndarray) -> None:
    """Write a 3D float32 volume (nz, ny, nx) as a mode-2 MRC file."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    if data.ndim == 2:
        data = data[None]
    nz, ny, nx = data.shape
    head = bytearray(1024)
    struct.pack_into("<4i", head, 0, nx, ny, nz, 2)
    struct.pack_into("<3i", head, 28, nx, ny, nz)          # mx, my, mz
    struct.pack_into("<3f", head, 40, float(nx), float(ny), float(nz))  # cell
    struct.pack_into("<3f", head, 52, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", head, 64, 1, 2, 3)             # axis order
    struct.pack_into("<3f", head, 76, float(data.min()), float(data.max()),
                     float(data.mean()))
    head[208:212] = b"MAP "
    head[212:216] = bytes([0x44, 0x44, 0x00, 0x00])        # little-endian stamp
    struct.pack_into("<f", head, 216, float(data.std()))
    with open(path, "wb") as fh:
        fh.write(head)
        fh.write(data.tobytes())


def read_angles(mrc_path: str) -> np.ndarray:
    """Angles in radians from the .rawtlt (preferred) or .csv file next to
    ``mrc_path``."""
    base = mrc_path[:-4] if mrc_path.endswith(".mrc") else mrc_path
    raw, csvp = base + ".rawtlt", base + ".csv"
    if os.path.isfile(raw):
        with open(raw) as fh:
            vals = [float(t) for t in fh.read().split()]
    elif os.path.isfile(csvp):
        with open(csvp) as fh:
            vals = [float(row[0]) for row in csv.reader(fh, delimiter=",") if row]
    else:
        raise IOError(f"no angle file {raw} or {csvp}")
    return np.asarray(vals, dtype=np.float64) / 180.0 * np.pi


def read_tilt_series(mrc_path: str):
    """(data[tilt, slice, bin] float64, angles in radians)."""
    return read_mrc(mrc_path).astype(np.float64), read_angles(mrc_path)
