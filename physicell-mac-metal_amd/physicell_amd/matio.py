"""MATLAB Level-4 (.mat) matrices: the reference's snapshot format.

    name, M = read_mat(path)        # M: numpy array (rows, cols)
    write_mat(path, name, M)

Layout (src/matlab_writer.cpp:5-49): five little-endian uint32 (type, rows,
cols, imaginary flag, name length incl. NUL), the name, then the values
column-major.  Type 0 is little-endian full double; 10 is float (read only).
Snapshots (src/output_writer.cpp:295-453): cells.mat has one column per cell
(row 0 id, rows 1-3 position, row 4 total volume, ...); microenvironment0.mat
has one column per voxel (rows 0-2 centre, row 3 volume, rows 4.. densities).
"""
import struct

import numpy as np

CELL_ROW_ID, CELL_ROW_X, CELL_ROW_Y, CELL_ROW_Z, CELL_ROW_TOTAL_VOLUME = 0, 1, 2, 3, 4
CELL_ROW_TYPE, CELL_ROW_CYCLE_MODEL, CELL_ROW_PHASE = 5, 6, 7
CELL_ROW_VX, CELL_ROW_VY, CELL_ROW_VZ, CELL_ROW_PRESSURE = 17, 18, 19, 20
CELL_ROW_DEAD, CELL_ROW_RADIUS = 26, 37
ME_ROW_X, ME_ROW_Y, ME_ROW_Z, ME_ROW_VOLUME, ME_FIRST_SUBSTRATE = 0, 1, 2, 3, 4


def read_mat(path):
    with open(path, "rb") as f:
        head = f.read(20)
        if len(head) != 20:
            raise ValueError(f"{path}: header truncated")
        typ, rows, cols, imag, nlen = struct.unpack("<5I", head)
        if (typ // 1000) % 10 != 0 or typ % 10 != 0 or imag != 0:
            raise ValueError(f"{path}: only little-endian real full matrices are supported (type {typ})")
        prec = (typ // 10) % 10
        name = f.read(nlen).rstrip(b"\0").decode("latin1")
        dt = {0: "<f8", 1: "<f4"}.get(prec)
        if dt is None:
            raise ValueError(f"{path}: unsupported precision {prec}")
        data = np.frombuffer(f.read(rows * cols * np.dtype(dt).itemsize), dtype=dt)
        if data.size != rows * cols:
            raise ValueError(f"{path}: data truncated")
    return name, data.astype(np.float64).reshape(cols, rows).T


def write_mat(path, name, m):
    m = np.asarray(m, np.float64)
    rows, cols = m.shape
    nb = name.encode("latin1") + b"\0"
    with open(path, "wb") as f:
        f.write(struct.pack("<5I", 0, rows, cols, 0, len(nb)))
        f.write(nb)
        f.write(np.ascontiguousarray(m.T).astype("<f8").tobytes())
