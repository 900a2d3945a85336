"""Parity on the benched workload: 200^3 voxels x 4 substrates, 1,000,000
agents in a jittered close-packed sphere (the configuration bench.py times).

* FP64 against the FP64 oracle over 2 mechanics periods (20 diffusion
  substeps with secretion, 2 mechanics steps): voxel keys, mechanics keys, bin
  offsets and the sorted list bit-exact; density, positions, velocities and
  pressure within 1e-10 relative.  This runs, at full size, the 16-segment x
  staging and the 8-segment y/z windows of 200-node lines, the FP32 mark pass
  of the FP64 force kernel on these coordinates, and the secretion plan over
  1M agents.
* Neighbour sets of the 1M agents (force predicates, visit order) bit-exact
  against the oracle.
* FP32-mixed against FP64 on the device over the workload's full 200
  mechanics steps (2000 substeps): per substrate max-abs and relative error
  (max |mixed - fp64| / max |fp64|, and pointwise over values above 1e-3 of
  the field's maximum), positions max-abs in um.  Bar: 1e-4 relative, 1e-3 um.
"""
import os

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H
from physicell_amd import scenarios as S

pytestmark = pytest.mark.gpu
REL = 1e-10


@pytest.fixture(scope="module")
def config2():
    return S.config2()


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    floor = max(float(np.max(np.abs(b))), 1e-300) * 1e-14
    return float(np.max(np.where(d == 0, 0.0, d / np.maximum(np.abs(b), floor)))) if a.size else 0.0


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def test_config2_fp64_two_periods_vs_oracle(config2):
    spec, cells = config2
    print(f"config2 FP64 vs oracle: 2 periods, {cells.n} agents, oracle on {threads()} threads", flush=True)
    sim = H.Simulation(spec, cells)
    max_time = 19 * spec.dt_diff           # 20 substeps (src/main.mm:275 loop condition)
    counts = sim.run(max_time)
    ref = O.OracleSim(spec, cells)
    counts_r = ref.run(max_time, nthreads=threads())
    assert counts == counts_r == (20, 2, 0), (counts, counts_r)
    assert np.array_equal(sim.field(H.F_VOXEL_IDX), ref.arr["voxel"])
    assert np.array_equal(sim.field(H.F_MECH_VOXEL), ref.arr["mech_key"])
    off, srt = sim.hash()
    assert np.array_equal(off, ref.off), "bin offsets"
    assert np.array_equal(srt, ref.srt[:cells.n]), "sorted list"
    V = spec.n_voxels
    rho, rho_r = sim.density(), ref.arr["rho"]
    for s in range(len(spec.substrates)):
        e = rel_err(rho[s * V:(s + 1) * V], rho_r[s * V:(s + 1) * V])
        print(f"  substrate {s}: max rel err {e:.3e}", flush=True)
        assert e <= REL, (s, e)
    for f, k in ((H.F_POS_X, "x"), (H.F_POS_Y, "y"), (H.F_POS_Z, "z"), (H.F_VEL_X, "vx"), (H.F_VEL_Y, "vy"),
                 (H.F_VEL_Z, "vz"), (H.F_PRESSURE, "pressure")):
        got = sim.field(f)
        e = rel_err(got, ref.arr[k])
        print(f"  {k}: max rel err {e:.3e}, bit-identical={np.array_equal(got, ref.arr[k])}", flush=True)
        assert e <= REL, (k, e)
    sim.close()


def test_config2_neighbor_sets_vs_oracle(config2):
    spec, cells = config2
    lib = H.load()
    ctx = H.Context(0)
    n = cells.n
    cb = H.CellBuffer(ctx, cells, len(spec.substrates))
    m, mg = H.mech_of(spec), H.mech_grid_of(spec)
    nb = spec.n_mech_voxels
    hb = ctx.alloc(lib.hc_hash_bytes(nb, cb.max_cells))
    ctx._check(lib.hc_build_hash(ctx.h, cb.soa.h, hb.h, m, mg, n, cb.max_cells, nb), "hash")
    mx = H.c_u32()
    ctx._check(lib.hc_hash_max_bin(ctx.h, hb.h, nb, H.ctypes.byref(mx)), "max bin")
    om = O.mech(spec)
    keys = O.mech_keys(cells.x, cells.y, cells.z, om)
    off_r, srt_r = O.hash_sort(keys, nb)
    assert mx.value == int(np.diff(off_r.astype(np.int64)).max())
    cap = 64
    cnt_b = ctx.alloc(4 * n)
    idx_b = ctx.alloc(4 * n * cap)
    ctx._check(lib.hc_neighbor_lists(ctx.h, cb.soa.h, hb.h, m, n, cb.max_cells, nb, cap, cnt_b.h, idx_b.h), "nbr")
    cnt = cnt_b.download(np.uint32, n)
    idx = idx_b.download(np.uint32, n * cap).reshape(n, cap)
    cnt_r, idx_r = O.neighbors(cells.x, cells.y, cells.z, keys, off_r, srt_r, om, cap)
    print(f"config2 neighbour sets: {int(cnt.sum())} pairs, max {int(cnt.max())} per agent, max bin {mx.value}",
          flush=True)
    assert cnt.max() <= cap
    assert np.array_equal(cnt, cnt_r)
    used = np.arange(cap)[None, :] < cnt[:, None]
    assert np.array_equal(idx[used], idx_r[used])
    for b in (cnt_b, idx_b, hb, cb.soa, cb.rates):
        b.free()
    ctx.close()


def test_config2_mixed_vs_fp64_full_run(config2):
    spec, cells = config2
    periods = 200
    print(f"config2 mixed vs FP64 on the device: {periods} mechanics steps", flush=True)
    a = H.Simulation(spec, cells)
    a.run_periods(periods)
    a.wait()
    rho_a = a.density()
    pos_a = [a.field(f) for f in (H.F_POS_X, H.F_POS_Y, H.F_POS_Z)]
    a.close()
    b = H.Simulation(spec, cells, mixed=True)
    b.run_periods(periods)
    b.wait()
    rho_b = b.density()
    pos_b = [b.field(f) for f in (H.F_POS_X, H.F_POS_Y, H.F_POS_Z)]
    b.close()
    V = spec.n_voxels
    for s in range(len(spec.substrates)):
        fa, fb = rho_a[s * V:(s + 1) * V], rho_b[s * V:(s + 1) * V]
        d = np.abs(fb - fa)
        scale = float(np.max(np.abs(fa)))
        big = np.abs(fa) > 1e-3 * scale
        pointwise = float(np.max(d[big] / np.abs(fa[big])))
        print(f"  substrate {s}: max |fp64| {scale:.4e}, max abs {d.max():.3e}, rel {d.max() / scale:.3e}, "
              f"pointwise rel (|v| > 1e-3 max) {pointwise:.3e}", flush=True)
        assert d.max() / scale <= 1e-4
    pe = [float(np.max(np.abs(pb - pa))) for pa, pb in zip(pos_a, pos_b)]
    print(f"  positions: max abs x/y/z {pe[0]:.3e} / {pe[1]:.3e} / {pe[2]:.3e} um", flush=True)
    assert max(pe) <= 1e-3
