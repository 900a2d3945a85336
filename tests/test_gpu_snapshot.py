"""pc_sim_write_snapshot: the .mat files hold the device state (positions,
velocities, pressure, densities) exactly, in the reference's row schema."""
import numpy as np
import pytest

from physicell_amd import hip as H
from physicell_amd import matio
from physicell_amd import scenarios as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("builder", ["config0", "small3d"])
def test_sim_snapshot_holds_state(tmp_path, builder):
    spec, cells = S.BUILDERS[builder]()
    sim = H.Simulation(spec, cells)
    sim.run_periods(3)
    sim.wait()
    sim.write_snapshot(tmp_path, 12)
    name, m = matio.read_mat(str(tmp_path / "output00000012_cells.mat"))
    assert name == "cells" and m.shape[1] == cells.n
    for row, f in [(matio.CELL_ROW_X, H.F_POS_X), (matio.CELL_ROW_Y, H.F_POS_Y), (matio.CELL_ROW_Z, H.F_POS_Z),
                   (matio.CELL_ROW_VX, H.F_VEL_X), (matio.CELL_ROW_PRESSURE, H.F_PRESSURE),
                   (matio.CELL_ROW_RADIUS, H.F_RADIUS)]:
        assert np.array_equal(m[row], sim.field(f)), row
    assert np.array_equal(m[matio.CELL_ROW_ID], np.arange(cells.n))
    name, me = matio.read_mat(str(tmp_path / "output00000012_microenvironment0.mat"))
    assert name == "multiscale_microenvironment"
    assert me.shape == (4 + len(spec.substrates), spec.n_voxels)
    assert np.array_equal(me[matio.ME_FIRST_SUBSTRATE:].reshape(-1), sim.density())
    # voxel centres (FP32 as the reference computes them)
    x0 = np.float32(spec.x_min) + (np.arange(spec.nx, dtype=np.float32) + np.float32(0.5)) * np.float32(spec.dx)
    assert np.array_equal(me[matio.ME_ROW_X][:spec.nx], x0.astype(np.float64))
    sim.close()


def oracle_snapshot(folder, frame, spec, o):
    """The oracle's state through the library's host writer (pc_write_snapshot)."""
    import ctypes
    a = o.arr
    n, ns = len(a["x"]), len(spec.substrates)
    cols = {"x": a["x"], "y": a["y"], "z": a["z"], "total_volume": a["volume"], "radius": a["radius"],
            "vx": a["vx"], "vy": a["vy"], "vz": a["vz"], "pressure": a["pressure"], "adhesion": a["adh"],
            "repulsion": a["rep"], "rel_max_adhesion": a["rel_adh"], "nuclear_volume": np.full(n, 540.0),
            "nuclear_radius": np.full(n, np.cbrt(3 * 540.0 / (4 * np.pi))), "cell_type": np.zeros(n, np.uint32),
            "current_phase": np.zeros(n, np.uint32), "is_alive": a["alive"], "secretion": a["sec"],
            "uptake": a["upt"], "saturation": a["sat"], "net_export": a["net"]}
    snap = H.Snapshot()
    snap.n_cells, snap.n_substrates = n, ns
    keep = []
    for k, v in cols.items():
        v = np.ascontiguousarray(v)
        keep.append(v)
        setattr(snap, k, v.ctypes.data)
    snap.nx, snap.ny, snap.nz = spec.nx, spec.ny, spec.nz
    for k in ("dx", "dy", "dz", "x_min", "y_min", "z_min"):
        setattr(snap, k, float(getattr(spec, k)))
    rho = np.ascontiguousarray(a["rho"])
    snap.density = rho.ctypes.data
    assert H.load().pc_write_snapshot(str(folder).encode(), frame, ctypes.byref(snap)) == 0


def test_config0_snapshots_pass_the_parity_diff(tmp_path):
    """Config 0 over its full 60 min (6001 substeps, 600 mechanics steps) on
    the device and in the oracle, both written as snapshots and compared with
    the parity harness's semantics (physicell_amd.snapdiff) at the FP64 bar:
    positions to 1e-10 um relative to the domain, volume exact, densities to
    1e-10 relative."""
    import oracle as O
    from physicell_amd import snapdiff as D
    spec, cells = S.config0()
    sim = H.Simulation(spec, cells)
    counts = sim.run(60.0)
    sim.wait()
    o = O.OracleSim(spec, cells)
    assert o.run(60.0) == tuple(counts)
    ours, ref = tmp_path / "ours", tmp_path / "ref"
    ours.mkdir(); ref.mkdir()
    sim.write_snapshot(ours, 5)
    oracle_snapshot(ref, 5, spec, o)
    sim.close()
    scale = float(np.abs(o.arr["rho"]).max())
    tol = {"position": 1e-10 * 500, "volume": 0.0, "phase": 0, "cell_count_abs": 0,
           "substrate_linf": 1e-10 * scale, "substrate_rmse": 1e-10 * scale}
    cells_d, micro_d = D.diff_frame(str(ref), str(ours), 5, tol)
    ok, text = D.report(5, cells_d, micro_d)
    print(text, flush=True)
    assert ok and cells_d.matched == cells.n
    assert max(f.rel for f in micro_d.fields) <= 1e-10


def csr(counts, idx_padded):
    counts = np.asarray(counts, np.int64)
    off = np.concatenate([[0], np.cumsum(counts)])
    used = np.arange(idx_padded.shape[1])[None, :] < counts[:, None]
    return off, idx_padded[used]


def device_neighbors(spec, cells, x, y, z, cap=128):
    """Force-predicate neighbour lists (visiting order) of positions x, y, z on
    the device hash (hc_build_hash + hc_neighbor_lists)."""
    import dataclasses
    lib = H.load()
    ctx = H.Context(0)
    c = dataclasses.replace(cells, x=np.asarray(x), y=np.asarray(y), z=np.asarray(z))
    cb = H.CellBuffer(ctx, c, len(spec.substrates))
    m, mg, nb, n = H.mech_of(spec), H.mech_grid_of(spec), spec.n_mech_voxels, c.n
    hb = ctx.alloc(lib.hc_hash_bytes(nb, cb.max_cells))
    ctx._check(lib.hc_build_hash(ctx.h, cb.soa.h, hb.h, m, mg, n, cb.max_cells, nb), "hash")
    cnt_b, idx_b = ctx.alloc(4 * n), ctx.alloc(4 * n * cap)
    ctx._check(lib.hc_neighbor_lists(ctx.h, cb.soa.h, hb.h, m, n, cb.max_cells, nb, cap, cnt_b.h, idx_b.h), "nbr")
    cnt = cnt_b.download(np.uint32, n)
    idx = idx_b.download(np.uint32, n * cap).reshape(n, cap)
    for b in (cnt_b, idx_b, hb, cb.soa, cb.rates):
        b.free()
    ctx.close()
    assert cnt.max() <= cap
    return csr(cnt, idx)


def oracle_neighbors(spec, x, y, z, cap=128):
    import oracle as O
    om = O.mech(spec)
    keys = O.mech_keys(x, y, z, om)
    off, srt = O.hash_sort(keys, spec.n_mech_voxels)
    cnt, idx = O.neighbors(x, y, z, keys, off, srt, om, cap)
    assert cnt.max() <= cap
    return csr(cnt, idx)


@pytest.mark.parametrize("builder", ["config0", "config1"])
def test_runs_pass_the_parity_harness_exactly(tmp_path, builder):
    """Config 0 and config 1 over their full 60 min (6001 substeps, 600
    mechanics steps) on the device and in the oracle, each written as a
    snapshot plus the exact-mode files (voxel and mechanics keys, bins, the
    hash-sorted list, the neighbour lists of the final positions), judged by
    the parity harness (parity/diff.py --exact): integer outputs equal,
    positions and densities at the FP64 bar (1e-10 relative)."""
    import os
    import oracle as O
    from physicell_amd import snapdiff as D
    spec, cells = S.BUILDERS[builder]()
    print(f"parity harness: {builder}, 60 min", flush=True)
    sim = H.Simulation(spec, cells)
    counts = sim.run(60.0)
    sim.wait()
    o = O.OracleSim(spec, cells)
    assert o.run(60.0, nthreads=max(1, min(16, os.cpu_count() or 1))) == tuple(counts)
    ours, ref = tmp_path / "ours", tmp_path / "ref"
    ours.mkdir(); ref.mkdir()
    sim.write_snapshot(ours, 5)
    oracle_snapshot(ref, 5, spec, o)
    pos = [sim.field(f) for f in (H.F_POS_X, H.F_POS_Y, H.F_POS_Z)]
    off, srt = sim.hash()
    noff, nidx = device_neighbors(spec, cells, *pos)
    D.write_exact(ours, 5, np.arange(cells.n), sim.field(H.F_VOXEL_IDX), sim.field(H.F_MECH_VOXEL), off, srt,
                  noff, nidx)
    a = o.arr
    roff, ridx = oracle_neighbors(spec, a["x"], a["y"], a["z"])
    D.write_exact(ref, 5, np.arange(cells.n), a["voxel"], a["mech_key"], o.off, o.srt[:cells.n], roff, ridx)
    sim.close()
    scale = float(np.abs(a["rho"]).max())
    ext = max(abs(spec.x_min), abs(spec.x_max), abs(spec.y_min), abs(spec.y_max), abs(spec.z_min), abs(spec.z_max))
    tol = {"position": 1e-10 * ext, "volume": 0.0, "phase": 0, "cell_count_abs": 0,
           "substrate_linf": 1e-10 * scale, "substrate_rmse": 1e-10 * scale}
    ok, text = D.compare(str(ref), str(ours), [5], tol, exact=True)
    print(text, flush=True)
    assert ok
    assert len(nidx) > cells.n                      # the lists are not empty
