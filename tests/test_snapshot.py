""".mat snapshots: the library's Level-4 writer and snapshot writer against
the reference's own writers (src/matlab_writer.cpp, src/output_writer.cpp,
compiled from the reference sources into oracle/_ref), byte for byte, and
the numpy reader/writer.  Host only."""
import ctypes
import os

import numpy as np
import pytest

from physicell_amd import hip as H
from physicell_amd import matio

REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_path.so")


def ref_lib():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    lib = ctypes.CDLL(REF)
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    lib.ref_write_matlab4.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u32, u32, vp]
    lib.ref_write_snapshot.argtypes = ([ctypes.c_char_p, ctypes.c_int, u32] + [vp] * 14 + [vp] * 3 + [u32] +
                                       [vp] * 4 + [u32] * 3 + [ctypes.c_float] * 6 + [vp])
    return lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_matlab4_bytes_match_reference(tmp_path):
    lib = H.load()
    ref = ref_lib()
    rng = np.random.default_rng(7)
    for rows, cols, name in [(7, 5, "cells"), (1, 1, "x"), (4, 300, "multiscale_microenvironment")]:
        m = rng.standard_normal((rows, cols))
        colmajor = np.ascontiguousarray(m.T).reshape(-1)
        a, b, c = tmp_path / "a.mat", tmp_path / "b.mat", tmp_path / "c.mat"
        assert lib.pc_write_matlab4(str(a).encode(), name.encode(), rows, cols, p(colmajor)) == 0
        assert ref.ref_write_matlab4(str(b).encode(), name.encode(), rows, cols, p(colmajor)) == 0
        matio.write_mat(str(c), name, m)
        assert a.read_bytes() == b.read_bytes() == c.read_bytes()
        nm, back = matio.read_mat(str(a))
        assert nm == name and np.array_equal(back, m)
    assert lib.pc_write_matlab4(str(tmp_path / "e.mat").encode(), b"e", 0, 3, p(np.zeros(3))) != 0


def f32(a):
    """values exactly representable in FP32 (the reference's cell store)"""
    return np.asarray(a, np.float32).astype(np.float64)


def test_snapshot_bytes_match_reference(tmp_path):
    lib = H.load()
    ref = ref_lib()
    rng = np.random.default_rng(11)
    n, ns, nx, ny, nz = 9, 2, 5, 4, 3
    V = nx * ny * nz
    cols = {k: f32(rng.uniform(lo, hi, n)) for k, lo, hi in [
        ("x", -90, 90), ("y", -70, 70), ("z", -40, 40), ("total_volume", 2000, 3000), ("nuclear_volume", 400, 600),
        ("radius", 7, 9), ("nuclear_radius", 4, 6), ("vx", -1, 1), ("vy", -1, 1), ("vz", -1, 1),
        ("pressure", 0, 2), ("adhesion", 0.2, 0.6), ("repulsion", 5, 15), ("rel_max_adhesion", 1.1, 1.4)]}
    ints = {"cell_type": np.zeros(n, np.uint32), "current_phase": rng.integers(0, 3, n).astype(np.uint32),
            "is_alive": np.array([1, 1, 0, 1, 1, 1, 0, 1, 1], np.uint32)}
    rates = {k: rng.uniform(0, 10, n * ns) for k in ("secretion", "uptake", "saturation", "net_export")}
    density = f32(rng.uniform(0, 40, ns * V))
    grid = dict(dx=20.0, dy=20.0, dz=20.0, x_min=-100.0, y_min=-80.0, z_min=-30.0)
    snap = H.Snapshot()
    snap.n_cells, snap.n_substrates = n, ns
    keep = []
    for k, a in list(cols.items()) + list(ints.items()) + list(rates.items()):
        a = np.ascontiguousarray(a)
        keep.append(a)
        setattr(snap, k, a.ctypes.data)
    snap.nx, snap.ny, snap.nz = nx, ny, nz
    for k, v in grid.items():
        setattr(snap, k, v)
    snap.density = density.ctypes.data
    ours, theirs = tmp_path / "ours", tmp_path / "ref"
    ours.mkdir(); theirs.mkdir()
    assert lib.pc_write_snapshot(str(ours).encode(), 3, ctypes.byref(snap)) == 0
    d32 = density.astype(np.float32)
    ref.ref_write_snapshot(str(theirs).encode(), 3, n,
                           *[p(cols[k]) for k in ("x", "y", "z", "total_volume", "nuclear_volume", "radius",
                                                  "nuclear_radius", "vx", "vy", "vz", "pressure", "adhesion",
                                                  "repulsion", "rel_max_adhesion")],
                           *[p(ints[k]) for k in ("cell_type", "current_phase", "is_alive")], ns,
                           *[p(rates[k]) for k in ("secretion", "uptake", "saturation", "net_export")],
                           nx, ny, nz, *[grid[k] for k in ("dx", "dy", "dz", "x_min", "y_min", "z_min")], p(d32))
    # microenvironment: identical bytes
    fname = "output00000003_microenvironment0.mat"
    assert (ours / fname).read_bytes() == (theirs / fname).read_bytes()
    # cells: identical header and rows, except row 10 (cytoplasmic volume =
    # total - nuclear), which the reference subtracts in FP32 and the library
    # in FP64 (exact for these inputs); equal to FP32 rounding
    fname = "output00000003_cells.mat"
    a, b = (ours / fname).read_bytes(), (theirs / fname).read_bytes()
    assert len(a) == len(b) and a[:20 + 6] == b[:20 + 6]
    na, ma = matio.read_mat(str(ours / fname))
    nb, mb = matio.read_mat(str(theirs / fname))
    assert na == nb and ma.shape == mb.shape
    rows = np.arange(ma.shape[0]) != 10
    assert np.array_equal(ma[rows], mb[rows]), sorted(set(np.argwhere(ma != mb)[:, 0].tolist()))
    cyto = cols["total_volume"] - cols["nuclear_volume"]
    assert np.array_equal(ma[10], cyto)
    assert np.array_equal(mb[10], cyto.astype(np.float32).astype(np.float64))
    name, cells = matio.read_mat(str(ours / "output00000003_cells.mat"))
    assert name == "cells" and cells.shape[1] == n
    assert np.array_equal(cells[matio.CELL_ROW_X], cols["x"])
    assert np.array_equal(cells[matio.CELL_ROW_DEAD], 1.0 - ints["is_alive"])
    name, me = matio.read_mat(str(ours / "output00000003_microenvironment0.mat"))
    assert me.shape == (4 + ns, V)
    assert np.array_equal(me[matio.ME_FIRST_SUBSTRATE:].reshape(-1), density)
