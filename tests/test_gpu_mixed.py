"""FP32-mixed form against the FP64 oracle (GPU, through the C-ABI).

Bar (north_star): substrate fields within 1e-4 relative and positions within
1e-3 um of FP64 over whole runs; voxel keys and bins are computed from FP64
positions in both forms.  Each test prints max-abs and max-relative error.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H
from physicell_amd import scenarios as S

pytestmark = pytest.mark.gpu


def errs(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(d.max()), float(d.max() / scale)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("nx,ny,nz", [(200, 30, 20), (50, 50, 1), (34, 66, 67), (530, 3, 12), (7, 9, 5)])
def test_fp32_substep_vs_fp64(ctx, nx, ny, nz):
    spec, _ = S.small3d(nx, ny, nz, n_cells=1, dirichlet=(True, False))
    if nz == 1:
        spec.use_2d = True
        spec.z_min, spec.z_max = -10, 10
    lib = H.load()
    g = H.grid_of(spec, fp32=True)
    subs = H.substrates_of(spec)
    tb = ctx.alloc(lib.hc_thomas_table_bytes(g))
    ctx._check(lib.hc_build_thomas_tables(ctx.h, g, subs, tb.h), "tables")
    rng = np.random.default_rng(2)
    rho0 = rng.uniform(0.0, 50.0, len(spec.substrates) * spec.n_voxels)
    d = ctx.alloc(rho0.size * 4)
    d.upload(rho0.astype(np.float32))
    ref = rho0.copy()
    print(f"fp32 substeps {nx}x{ny}x{nz}", flush=True)
    for _ in range(20):
        ctx._check(lib.hc_diffusion_substep(ctx.h, d.h, g, subs, tb.h), "substep")
        O.diffusion_substep(ref, O.grid(spec), O.subs(spec))
    got = d.download(np.float32, rho0.size)
    ea, er = errs(got, ref)
    print(f"  max abs {ea:.3e}  max rel {er:.3e}")
    assert er < 1e-5
    # compensated totals agree with an exactly rounded sum of the same values
    tot = np.zeros(len(spec.substrates))
    ctx._check(lib.hc_substrate_totals(ctx.h, d.h, g, H._ptr(tot)), "totals")
    for s in range(len(spec.substrates)):
        exact = math.fsum(got[s * spec.n_voxels:(s + 1) * spec.n_voxels].astype(np.float64))
        assert abs(tot[s] - exact) <= 1e-15 * max(1.0, abs(exact)), (tot[s], exact)
    d.free(); tb.free()


@pytest.mark.parametrize("builder,max_time", [("config0", 60.0), ("config1", 60.0)])
def test_mixed_whole_run_vs_fp64(builder, max_time):
    spec, cells = S.BUILDERS[builder]()
    print(f"mixed run {builder} {max_time} min", flush=True)
    sim = H.Simulation(spec, cells, mixed=True)
    counts = sim.run(max_time)
    ref = O.OracleSim(spec, cells)
    assert counts == ref.run(max_time, nthreads=max(1, min(16, os.cpu_count() or 1)))
    ea, er = errs(sim.density(), ref.arr["rho"])
    print(f"  density: max abs {ea:.3e} max rel {er:.3e}")
    assert er <= 1e-4
    pe = max(errs(sim.field(f), ref.arr[k])[0] for f, k in ((H.F_POS_X, "x"), (H.F_POS_Y, "y"), (H.F_POS_Z, "z")))
    print(f"  positions: max abs {pe:.3e} um")
    assert pe <= 1e-3
    # totals: compensated device sum equals an exactly rounded host sum of the field
    tot = sim.totals()
    rho = sim.density()
    V = spec.n_voxels
    for s in range(len(spec.substrates)):
        exact = math.fsum(rho[s * V:(s + 1) * V])
        assert abs(tot[s] - exact) <= 1e-15 * max(1.0, abs(exact))
    sim.close()


def test_mixed_first_hash_matches_fp64(ctx):
    """Keys and bins come from FP64 positions in both forms: identical before
    the positions diverge."""
    spec, cells = S.config1()
    lib = H.load()
    cb = H.CellBuffer(ctx, cells, len(spec.substrates))
    nb = spec.n_mech_voxels
    hb = ctx.alloc(lib.hc_hash_bytes(nb, cb.max_cells))
    ctx._check(lib.hc_build_hash(ctx.h, cb.soa.h, hb.h, H.mech_of(spec, fp32=True), H.mech_grid_of(spec), cells.n,
                                 cb.max_cells, nb), "hash")
    off = np.empty(nb + 1, np.uint32)
    srt = np.empty(cells.n, np.uint32)
    ctx._check(lib.hc_hash_download(ctx.h, hb.h, nb, cb.max_cells, cells.n, H._ptr(off), H._ptr(srt)), "dl")
    keys = O.mech_keys(cells.x, cells.y, cells.z, O.mech(spec))
    o2, s2 = O.hash_sort(keys, nb)
    assert np.array_equal(off, o2) and np.array_equal(srt, s2)
    ctx._check(lib.hc_forces(ctx.h, cb.soa.h, hb.h, H.mech_of(spec, fp32=True), cells.n, cb.max_cells, nb), "f32")
    vx = cb.field(H.F_VEL_X)
    rx, _, _, _ = O.forces(cells.x, cells.y, cells.z, cells.radius, cells.rep, cells.adh, cells.rel_adh, keys, o2, s2,
                           O.mech(spec))
    ea, er = errs(vx, rx)
    print(f"fp32 forces: max abs {ea:.3e} max rel {er:.3e}")
    assert er < 1e-4
    for b in (hb, cb.soa, cb.rates):
        b.free()


def _fp32_forces(ctx, spec, cells, mode):
    """FP32 forces of one hash build: mode 1 = x-trimmed mark pass with a
    list, mode 2 = the chunked pass over whole rows."""
    lib = H.load()
    cb = H.CellBuffer(ctx, cells, len(spec.substrates))
    nb = spec.n_mech_voxels
    hb = ctx.alloc(lib.hc_hash_bytes(nb, cb.max_cells))
    mech = H.mech_of(spec, fp32=True)
    mech.forces_fp32 = mode
    ctx._check(lib.hc_build_hash(ctx.h, cb.soa.h, hb.h, mech, H.mech_grid_of(spec), cells.n, cb.max_cells, nb),
               "hash")
    ctx._check(lib.hc_forces(ctx.h, cb.soa.h, hb.h, mech, cells.n, cb.max_cells, nb), "forces")
    out = [cb.field(f) for f in (H.F_VEL_X, H.F_VEL_Y, H.F_VEL_Z, H.F_PRESSURE)]
    for b in (hb, cb.soa, cb.rates):
        b.free()
    return out


def _crowded(seed=7):
    """small3d with a knot of 300 agents in a few 30 um bins (bins above the
    16-member sort cap, slabs holding many members) beside sparse agents."""
    spec, cells = S.small3d(nx=16, ny=12, nz=9, n_cells=600, seed=seed)
    rng = np.random.default_rng(seed)
    k = 300
    cells.x[:k] = 95.0 + rng.uniform(-20.0, 20.0, k)
    cells.y[:k] = 118.0 + rng.uniform(-12.0, 12.0, k)
    cells.z[:k] = 61.0 + rng.uniform(-6.0, 6.0, k)
    return spec, cells


@pytest.mark.parametrize("builder", ["config1", "small3d", "crowded", "config0"])
def test_fp32_trimmed_list_equals_chunked_pass(ctx, builder):
    """The mark pass walks only the x interval of each row that can interact
    (records ordered by x slab inside their bins); the chunked pass walks
    every member of the 9 rows.  Pairs outside the interval add +0, so the
    two must agree bit for bit."""
    spec, cells = {"config1": S.config1, "small3d": S.small3d, "crowded": _crowded, "config0": S.config0}[builder]()
    a = _fp32_forces(ctx, spec, cells, 1)
    b = _fp32_forces(ctx, spec, cells, 2)
    nz = int(np.count_nonzero(a[3]))
    print(f"{builder}: {cells.n} agents, {nz} under pressure", flush=True)
    assert nz > cells.n // 10
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
