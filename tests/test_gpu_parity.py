"""GPU parity: the HIP path through the C-ABI against the FP64 oracle.

Bar: integer/index outputs (voxel keys, bins, neighbour sets) bit-exact;
FP64 fields, positions, velocities within 1e-10 relative (the kernels and the
oracle evaluate the same operation sequences with -ffp-contract=off, so the
results are expected to be bit-identical and the check also reports that).
"""
import os

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H
from physicell_amd import scenarios as S

pytestmark = pytest.mark.gpu
REL = 1e-10


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(b), 1e-300)
    d = np.abs(a - b)
    return float(np.max(np.where(d == 0, 0.0, d / np.maximum(scale, np.max(np.abs(b)) * 1e-14)))) if a.size else 0.0


def assert_close(a, b, what, rel=REL):
    e = rel_err(a, b)
    print(f"{what}: max rel err {e:.3e}, bit-identical={np.array_equal(a, b)}", flush=True)
    assert e <= rel, f"{what}: rel err {e} > {rel}"


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def make_density(spec, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 50.0, len(spec.substrates) * spec.n_voxels)


def tables_for(ctx, spec):
    lib = H.load()
    g = H.grid_of(spec)
    subs = H.substrates_of(spec)
    tb = ctx.alloc(max(8, lib.hc_thomas_table_bytes(g)))
    ctx._check(lib.hc_build_thomas_tables(ctx.h, g, subs, tb.h), "tables")
    return g, subs, tb


GRIDS = [
    # (nx, ny, nz, dirichlet flags)  ragged and special-case line lengths
    (12, 10, 8, (True, False)),
    (50, 50, 1, (True, False)),
    (3, 5, 4, (True, True)),       # n == 3 along x
    (2, 7, 3, (False, True)),      # n == 2 along x, n == 3 along z
    (67, 33, 5, (False, False)),
    (17, 4, 19, (True, False)),
    (220, 16, 10, (True, False)),   # x: 4 segments of 55 (64-wide register window)
    (34, 66, 67, (False, True)),    # segment lengths 8, 16 and 17 (window 16 / 32)
    (259, 12, 9, (True, True)),     # x: 16 segments of 17 (window 32), y/z short
    (530, 3, 12, (True, False)),    # x too long for the partitioned window -> serial
    (6, 150, 140, (True, False)),   # FP64 y/z longer than 130: 8 segments (window 32)
    (5, 300, 6, (False, True)),     # FP64 y: 8 segments of 37 (window 64)
]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("nx,ny,nz,dirs", GRIDS)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_single_sweep(ctx, nx, ny, nz, dirs, axis, mode):
    if nz == 1 and axis == 2:
        pytest.skip("no z sweep in 2-D")
    spec, _ = S.small3d(nx, ny, nz, n_cells=1, dirichlet=dirs)
    if nz == 1:
        spec.use_2d = True
        spec.z_min, spec.z_max = -10, 10
    lib = H.load()
    g, subs, tb = tables_for(ctx, spec)
    rho0 = make_density(spec, 11 + axis)
    d = ctx.alloc(rho0.nbytes)
    d.upload(rho0)
    print(f"sweep axis={axis} grid={nx}x{ny}x{nz} mode={mode}", flush=True)
    ctx._check(lib.hc_set_sweep_mode(ctx.h, mode), "mode")
    ctx._check(lib.hc_diffusion_sweep(ctx.h, axis, d.h, g, subs, tb.h, 0), "sweep")
    got = d.download(np.float64, rho0.size)
    ref = rho0.copy()
    O.sweep(axis, ref, O.grid(spec), O.subs(spec))
    assert_close(got, ref, f"sweep{axis} {nx}x{ny}x{nz} mode {mode}")
    if mode == 1:
        assert np.array_equal(got, ref), "serial mode is bit-identical to the oracle"
    ctx._check(lib.hc_set_sweep_mode(ctx.h, 0), "mode")
    d.free(); tb.free()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("nx,ny,nz,dirs", GRIDS)
def test_substep_with_dirichlet(ctx, nx, ny, nz, dirs, mode):
    spec, _ = S.small3d(nx, ny, nz, n_cells=1, dirichlet=dirs)
    if nz == 1:
        spec.use_2d = True
        spec.z_min, spec.z_max = -10, 10
    lib = H.load()
    g, subs, tb = tables_for(ctx, spec)
    rho0 = make_density(spec, 3)
    d = ctx.alloc(rho0.nbytes)
    d.upload(rho0)
    print(f"substep grid={nx}x{ny}x{nz} mode={mode}", flush=True)
    ctx._check(lib.hc_set_sweep_mode(ctx.h, mode), "mode")
    ref = rho0.copy()
    for _ in range(5):
        ctx._check(lib.hc_diffusion_substep(ctx.h, d.h, g, subs, tb.h), "substep")
        O.diffusion_substep(ref, O.grid(spec), O.subs(spec))
    got = d.download(np.float64, rho0.size)
    assert_close(got, ref, f"5 substeps {nx}x{ny}x{nz} mode {mode}")
    ctx._check(lib.hc_set_sweep_mode(ctx.h, 0), "mode")
    # standalone Dirichlet is idempotent after the fused substep
    ctx._check(lib.hc_dirichlet(ctx.h, d.h, g, subs), "dirichlet")
    got2 = d.download(np.float64, rho0.size)
    assert np.array_equal(got, got2)
    d.free(); tb.free()


@pytest.mark.parametrize("fp32", [False, True])
def test_substep_orders_agree(ctx, fp32):
    """Axis-major and substrate-major sweep orders give the same field, bit
    for bit (the substrates' sweeps are independent)."""
    spec, _ = S.small3d(44, 30, 26, n_cells=1, ns=2, dirichlet=(True, False))
    lib = H.load()
    g = H.grid_of(spec, fp32)
    subs = H.substrates_of(spec)
    tb = ctx.alloc(max(8, lib.hc_thomas_table_bytes(g)))
    ctx._check(lib.hc_build_thomas_tables(ctx.h, g, subs, tb.h), "tables")
    rho0 = make_density(spec, 5).astype(np.float32 if fp32 else np.float64)
    out = {}
    for order in (1, 2):
        d = ctx.alloc(rho0.nbytes)
        d.upload(rho0)
        ctx._check(lib.hc_set_substep_order(ctx.h, order), "order")
        for _ in range(4):
            ctx._check(lib.hc_diffusion_substep(ctx.h, d.h, g, subs, tb.h), "substep")
        out[order] = d.download(rho0.dtype, rho0.size)
        d.free()
    ctx._check(lib.hc_set_substep_order(ctx.h, 0), "order")
    print(f"substep orders fp32={fp32}: bit-identical={np.array_equal(out[1], out[2])}", flush=True)
    assert np.array_equal(out[1], out[2])
    tb.free()


def cells_buffer(ctx, spec, cells):
    return H.CellBuffer(ctx, cells, len(spec.substrates))


def crowded():
    """About 12 agents per 30 um bin at random positions: bins of up to 16
    members are sorted by one thread each, larger ones by the per-agent rank;
    this population has both."""
    return S.small3d(12, 10, 8, n_cells=4000, seed=11)


MECH_BUILDERS = dict(S.BUILDERS, crowded=crowded)


@pytest.mark.parametrize("builder", ["config0", "small3d", "config1", "crowded"])
def test_voxel_keys_hash_neighbors_forces_ab2(ctx, builder):
    spec, cells = MECH_BUILDERS[builder]()
    lib = H.load()
    n = cells.n
    cb = cells_buffer(ctx, spec, cells)
    print(f"mech parity {builder} n={n}", flush=True)
    # voxel keys (diffusion grid)
    g = H.grid_of(spec)
    ctx._check(lib.hc_voxel_keys(ctx.h, cb.soa.h, g, n, cb.max_cells), "voxel keys")
    vk = cb.field(H.F_VOXEL_IDX)
    assert np.array_equal(vk, O.voxel_keys(cells.x, cells.y, cells.z, O.grid(spec)))
    # hash
    m, mg = H.mech_of(spec), H.mech_grid_of(spec)
    nb = spec.n_mech_voxels
    hb = ctx.alloc(lib.hc_hash_bytes(nb, cb.max_cells))
    ctx._check(lib.hc_build_hash(ctx.h, cb.soa.h, hb.h, m, mg, n, cb.max_cells, nb), "hash")
    off = np.empty(nb + 1, np.uint32)
    srt = np.empty(n, np.uint32)
    ctx._check(lib.hc_hash_download(ctx.h, hb.h, nb, cb.max_cells, n, H._ptr(off), H._ptr(srt)), "hash dl")
    om = O.mech(spec)
    keys_ref = O.mech_keys(cells.x, cells.y, cells.z, om)
    off_ref, srt_ref = O.hash_sort(keys_ref, nb)
    assert np.array_equal(cb.field(H.F_MECH_VOXEL), keys_ref)
    assert np.array_equal(off, off_ref)
    assert np.array_equal(srt, srt_ref)
    assert np.array_equal(np.diff(off_ref.astype(np.int64)), np.bincount(keys_ref, minlength=nb))
    if builder == "crowded":
        sizes = np.diff(off_ref.astype(np.int64))
        assert sizes.max() > 16 and ((sizes > 0) & (sizes <= 16)).any(), "both bin paths"
    # neighbour lists (visit order) vs oracle
    cap = 128
    cnt_b = ctx.alloc(4 * n)
    idx_b = ctx.alloc(4 * n * cap)
    ctx._check(lib.hc_neighbor_lists(ctx.h, cb.soa.h, hb.h, m, n, cb.max_cells, nb, cap, cnt_b.h, idx_b.h), "nbr")
    cnt = cnt_b.download(np.uint32, n)
    idx = idx_b.download(np.uint32, n * cap).reshape(n, cap)
    cnt_r, idx_r = O.neighbors(cells.x, cells.y, cells.z, keys_ref, off_ref, srt_ref, om, cap)
    assert np.array_equal(cnt, cnt_r)
    assert cnt.max() <= cap
    for i in range(n):
        assert np.array_equal(idx[i, :cnt[i]], idx_r[i, :cnt[i]])
    # forces
    ctx._check(lib.hc_forces(ctx.h, cb.soa.h, hb.h, m, n, cb.max_cells, nb), "forces")
    vx_r, vy_r, vz_r, p_r = O.forces(cells.x, cells.y, cells.z, cells.radius, cells.rep, cells.adh,
                                     cells.rel_adh, keys_ref, off_ref, srt_ref, om)
    for f, r, nm in ((H.F_VEL_X, vx_r, "vx"), (H.F_VEL_Y, vy_r, "vy"), (H.F_VEL_Z, vz_r, "vz"),
                     (H.F_PRESSURE, p_r, "pressure")):
        assert_close(cb.field(f), r, f"{builder} {nm}")
    # AB2 twice (second step uses v_prev)
    x, y, z = (np.array(a, np.float64) for a in (cells.x, cells.y, cells.z))
    pvx, pvy, pvz = np.zeros(n), np.zeros(n), np.zeros(n)
    for step in range(2):
        ctx._check(lib.hc_integrate(ctx.h, cb.soa.h, mg, spec.dt_mech, n, cb.max_cells), "ab2")
        O.integrate(x, y, z, vx_r, vy_r, vz_r, pvx, pvy, pvz, om, spec.dt_mech)
    assert_close(cb.field(H.F_POS_X), x, f"{builder} x")
    assert_close(cb.field(H.F_POS_Y), y, f"{builder} y")
    assert_close(cb.field(H.F_POS_Z), z, f"{builder} z")
    assert_close(cb.field(H.F_PREV_VX), vx_r, f"{builder} prev vx")
    for b in (cnt_b, idx_b, hb, cb.soa, cb.rates):
        b.free()


def test_hash_overflow_report_keeps_every_agent(ctx):
    """The reference's hash keeps at most 64 agents per bucket and drops the
    rest (shaders/mechanics.metal:79-84).  This build drops none and reports
    the largest bin: one 30 um bin holding 150 agents plus a sparse background;
    hc_hash_max_bin returns the true count, every agent appears exactly once
    in the sorted list, the bins equal the oracle's, and the forces on the
    crowded bin's members equal the oracle's."""
    spec, cells = S.small3d(12, 10, 8, n_cells=400, seed=21)
    rng = np.random.default_rng(22)
    crowd = 150
    # bin (3, 2, 1) of the 30 um grid spans [90, 120) x [60, 90) x [30, 60)
    lo = np.array([90.0, 60.0, 30.0])
    p = lo + 0.5 + rng.random((crowd, 3)) * 29.0
    cells.x[:crowd], cells.y[:crowd], cells.z[:crowd] = p[:, 0], p[:, 1], p[:, 2]
    # keep the background out of that bin
    bg = slice(crowd, None)
    inside = ((cells.x[bg] >= 90) & (cells.x[bg] < 120) & (cells.y[bg] >= 60) & (cells.y[bg] < 90)
              & (cells.z[bg] >= 30) & (cells.z[bg] < 60))
    cells.x[bg][inside] += 60.0
    lib = H.load()
    n = cells.n
    cb = cells_buffer(ctx, spec, cells)
    m, mg = H.mech_of(spec), H.mech_grid_of(spec)
    nb = spec.n_mech_voxels
    hb = ctx.alloc(lib.hc_hash_bytes(nb, cb.max_cells))
    ctx._check(lib.hc_build_hash(ctx.h, cb.soa.h, hb.h, m, mg, n, cb.max_cells, nb), "hash")
    mx = H.c_u32()
    ctx._check(lib.hc_hash_max_bin(ctx.h, hb.h, nb, H.ctypes.byref(mx)), "max bin")
    om = O.mech(spec)
    keys_ref = O.mech_keys(cells.x, cells.y, cells.z, om)
    sizes = np.bincount(keys_ref, minlength=nb)
    print(f"overflow: max bin {mx.value}, oracle {sizes.max()}", flush=True)
    assert sizes.max() == crowd and mx.value == crowd
    off = np.empty(nb + 1, np.uint32)
    srt = np.empty(n, np.uint32)
    ctx._check(lib.hc_hash_download(ctx.h, hb.h, nb, cb.max_cells, n, H._ptr(off), H._ptr(srt)), "hash dl")
    off_ref, srt_ref = O.hash_sort(keys_ref, nb)
    assert np.array_equal(off, off_ref) and np.array_equal(srt, srt_ref)
    assert np.array_equal(np.sort(srt), np.arange(n, dtype=np.uint32)), "every agent exactly once"
    big = int(np.argmax(sizes))
    assert np.array_equal(srt[off[big]:off[big + 1]], np.arange(crowd, dtype=np.uint32))
    ctx._check(lib.hc_forces(ctx.h, cb.soa.h, hb.h, m, n, cb.max_cells, nb), "forces")
    vx_r, vy_r, vz_r, p_r = O.forces(cells.x, cells.y, cells.z, cells.radius, cells.rep, cells.adh,
                                     cells.rel_adh, keys_ref, off_ref, srt_ref, om)
    for f, r, nm in ((H.F_VEL_X, vx_r, "vx"), (H.F_VEL_Y, vy_r, "vy"), (H.F_VEL_Z, vz_r, "vz"),
                     (H.F_PRESSURE, p_r, "pressure")):
        assert_close(cb.field(f), r, f"overflow {nm}")
    for b in (hb, cb.soa, cb.rates):
        b.free()


def test_secretion_crowded_voxels(ctx):
    """Many cells per voxel, dead cells, net export, out-of-range voxels."""
    spec, cells = S.small3d(6, 5, 4, n_cells=2000, seed=9)
    rng = np.random.default_rng(4)
    n, ns = cells.n, len(spec.substrates)
    cells.sec[:] = rng.uniform(0, 10, (n, ns))
    cells.upt[:] = rng.uniform(0, 10, (n, ns))
    cells.sat[:] = rng.uniform(0, 40, (n, ns))
    cells.net[:] = rng.uniform(-2000, 1000, (n, ns))
    cells.volume[:] = rng.uniform(1000, 4000, n)
    cells.alive[rng.random(n) < 0.1] = 0
    lib = H.load()
    cb = cells_buffer(ctx, spec, cells)
    g = H.grid_of(spec)
    ctx._check(lib.hc_voxel_keys(ctx.h, cb.soa.h, g, n, cb.max_cells), "keys")
    vox = cb.field(H.F_VOXEL_IDX)
    rho0 = make_density(spec, 8)
    d = ctx.alloc(rho0.nbytes)
    d.upload(rho0)
    plan = ctx.alloc(lib.hc_secretion_plan_bytes(g, cb.max_cells))
    ctx._check(lib.hc_secretion(ctx.h, d.h, cb.soa.h, cb.rates.h, g, spec.dt_diff, n, cb.max_cells, plan.h), "sec")
    for _ in range(3):
        ctx._check(lib.hc_secretion_apply(ctx.h, d.h, plan.h, g), "apply")
    got = d.download(np.float64, rho0.size)
    ref = rho0.copy()
    for _ in range(4):
        O.secretion(ref, O.grid(spec), spec.dt_diff, vox, cells.volume, cells.sec.ravel(), cells.upt.ravel(),
                    cells.sat.ravel(), cells.net.ravel(), cells.alive)
    assert_close(got, ref, "secretion x4 crowded")
    assert (got >= 0).all()
    for b in (d, plan, cb.soa, cb.rates):
        b.free()


@pytest.mark.parametrize("builder,max_time", [("config0", 60.0), ("small3d", 2.0), ("config1", 60.0)])
def test_whole_run_vs_oracle(builder, max_time):
    spec, cells = S.BUILDERS[builder]()
    print(f"whole run {builder} {max_time} min", flush=True)
    sim = H.Simulation(spec, cells)
    counts = sim.run(max_time)
    ref = O.OracleSim(spec, cells)
    counts_r = ref.run(max_time, nthreads=max(1, min(16, os.cpu_count() or 1)))
    assert counts == counts_r
    assert_close(sim.density(), ref.arr["rho"], f"{builder} density after {counts[0]} substeps")
    for f, k in ((H.F_POS_X, "x"), (H.F_POS_Y, "y"), (H.F_POS_Z, "z"), (H.F_VEL_X, "vx"), (H.F_PRESSURE, "pressure")):
        assert_close(sim.field(f), ref.arr[k], f"{builder} {k}")
    assert np.array_equal(sim.field(H.F_VOXEL_IDX), ref.arr["voxel"])
    assert np.array_equal(sim.field(H.F_MECH_VOXEL), ref.arr["mech_key"])
    off, srt = sim.hash()
    assert np.array_equal(off, ref.off)
    assert np.array_equal(srt, ref.srt[:cells.n])
    sim.close()


def test_graph_replay_matches_eager():
    spec, cells = S.small3d(20, 18, 16, n_cells=1500, seed=3)
    a = H.Simulation(spec, cells)
    b = H.Simulation(spec, cells)
    for rep in range(3):
        a.run_periods(2, use_graph=False)
        b.run_periods(2, use_graph=True)
        b.wait()
        da, db = a.density(), b.density()
        assert np.array_equal(da, db), f"replay {rep}"
        assert np.array_equal(a.field(H.F_POS_X), b.field(H.F_POS_X))
    a.close(); b.close()


@pytest.mark.parametrize("dt_pheno,periods", [(6.0, 70), (0.25, 12)])
def test_run_periods_phenotype_clock_matches_run(dt_pheno, periods):
    """pc_sim_run_periods fires the phenotype step (voxel refresh + secretion
    re-plan) where the step loop of src/main.mm:394-412 fires it: P periods
    equal pc_sim_run over the same P*k substeps bit for bit, eager and graph,
    serial and overlapped, with the periods split over two calls.  dt_pheno
    6.0 fires after the mechanics step of period 60; 0.25 fires inside
    periods (substeps 25, 50, 75, ...)."""
    import dataclasses
    spec, cells = S.small3d(16, 14, 12, n_cells=3000, seed=4)
    spec = dataclasses.replace(spec, dt_pheno=dt_pheno)
    k = 10
    print(f"run_periods phenotype clock dt_pheno={dt_pheno} periods={periods}", flush=True)
    ref = H.Simulation(spec, cells)
    keys0 = ref.field(H.F_VOXEL_IDX)
    counts = ref.run((periods * k - 1) * spec.dt_diff)
    assert counts[0] == periods * k and counts[1] == periods
    assert counts[2] == int(periods * k * spec.dt_diff / dt_pheno + 1e-9)
    keys = ref.field(H.F_VOXEL_IDX)
    assert not np.array_equal(keys, keys0), "the voxel refresh must change keys for this case to test anything"
    rho, pos = ref.density(), [ref.field(f) for f in (H.F_POS_X, H.F_POS_Y, H.F_POS_Z)]
    ref.close()
    for use_graph in (False, True):
        for overlap in (False, True):
            sim = H.Simulation(spec, cells)
            sim.set_overlap(overlap)
            first = periods // 2 + 1
            sim.run_periods(first, use_graph=use_graph)
            sim.run_periods(periods - first, use_graph=use_graph)
            sim.wait()
            what = f"graph={use_graph} overlap={overlap}"
            assert np.array_equal(sim.field(H.F_VOXEL_IDX), keys), what
            assert np.array_equal(sim.density(), rho), what
            for f, p in zip((H.F_POS_X, H.F_POS_Y, H.F_POS_Z), pos):
                assert np.array_equal(sim.field(f), p), what
            sim.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_overlapped_mechanics_matches_serial(mixed):
    """The mechanics step on a second stream, concurrent with the substeps,
    gives the serial order's results bit for bit (graph and eager)."""
    spec, cells = S.small3d(30, 24, 20, n_cells=2500, seed=9)
    a = H.Simulation(spec, cells, mixed=mixed)
    b = H.Simulation(spec, cells, mixed=mixed)
    a.set_overlap(False)
    b.set_overlap(True)
    for use_graph in (False, True, True):
        a.run_periods(3, use_graph=use_graph)
        b.run_periods(3, use_graph=use_graph)
        b.wait()
        assert np.array_equal(a.density(), b.density())
        for f in (H.F_POS_X, H.F_POS_Y, H.F_POS_Z, H.F_VEL_X, H.F_PRESSURE):
            assert np.array_equal(a.field(f), b.field(f))
    a.close(); b.close()


def test_errors_are_reported(ctx):
    lib = H.load()
    spec, cells = S.small3d(8, 8, 8, n_cells=10)
    g = H.grid_of(spec)
    subs = H.substrates_of(spec)
    small = ctx.alloc(16)
    tb = ctx.alloc(8)
    assert lib.hc_diffusion_sweep(ctx.h, 0, small.h, g, subs, tb.h, 0) == -1
    assert b"too small" in lib.hc_last_error(ctx.h)
    assert lib.hc_diffusion_sweep(ctx.h, 5, small.h, g, subs, tb.h, 0) == -1
    cb = cells_buffer(ctx, spec, cells)
    assert lib.hc_voxel_keys(ctx.h, cb.soa.h, g, cb.max_cells + 1, cb.max_cells) == -1
    plan = ctx.alloc(64)
    assert lib.hc_secretion_apply(ctx.h, small.h, plan.h, g) != 0
    for b in (small, tb, plan, cb.soa, cb.rates):
        b.free()
