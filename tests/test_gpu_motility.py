"""Motility on the device (hc_motility) against the oracle's or_motility on
the same inputs, over consecutive calls (the step counter in device memory
advances by one per call, so every call draws fresh numbers and the oracle,
given the same step, draws the same ones): resample decisions and elapsed
times bit for bit, vectors and velocities within 1e-12 relative (the device's
sin/cos/sqrt against the host's); FP64 and FP32 fields.  Then the
Simulation with motility on: no motile cell (or zero speed) changes nothing,
runs are reproducible, graph replay equals eager, and the seed matters."""
import ctypes

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H, scenarios as S

pytestmark = pytest.mark.gpu

POS = (H.F_POS_X, H.F_POS_Y, H.F_POS_Z)
VEL = (H.F_VEL_X, H.F_VEL_Y, H.F_VEL_Z)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context()
    yield c
    c.close()


def motility_state(n, ns, seed, dt):
    r = np.random.default_rng(seed)
    st = {
        "persistence": np.where(r.uniform(size=n) < 0.3, 0.5 * dt, r.uniform(2 * dt, 20 * dt, n)),
        "elapsed": np.zeros(n),
        "bias": np.where(r.uniform(size=n) < 0.2, 0.0, r.uniform(0, 1, n)),
        "vec": [np.zeros(n) for _ in range(3)],
        "motile": (r.uniform(size=n) < 0.85).astype(np.uint32),
        "r2d": (r.uniform(size=n) < 0.4).astype(np.uint32),
        "chem_index": r.integers(0, ns, n).astype(np.int32),
        "chem_dir": np.where(r.uniform(size=n) < 0.5, 1, -1).astype(np.int32),
        "sens": r.uniform(-2, 2, n * ns),
    }
    return st


def pack(lib, st, n, M, ns, seed):
    buf = np.zeros(lib.hc_motility_bytes(M, ns), np.uint8)
    buf[:16] = np.array([0, seed], np.uint64).view(np.uint8)

    def put(field, a):
        o = lib.hc_motility_field_offset(field, M, ns)
        b = np.ascontiguousarray(a).view(np.uint8)
        buf[o:o + b.size] = b

    put(H.M_PERSISTENCE, st["persistence"]); put(H.M_ELAPSED, st["elapsed"]); put(H.M_BIAS, st["bias"])
    for k in range(3):
        put(H.M_VEC_X + k, st["vec"][k])
    put(H.M_MOTILE, st["motile"]); put(H.M_RESTRICT_2D, st["r2d"])
    put(H.M_CHEM_INDEX, st["chem_index"]); put(H.M_CHEM_DIR, st["chem_dir"]); put(H.M_SENSITIVITY, st["sens"])
    return buf


def read(lib, mot, field, n, M, ns, dtype=np.float64):
    return mot.download(dtype, n, lib.hc_motility_field_offset(field, M, ns))


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("use_sens", [0, 1])
def test_motility_matches_oracle(ctx, fp32, use_sens):
    lib = H.load()
    spec, cells = S.small3d(14, 11, 9, n_cells=2000, seed=21 + use_sens)
    ns, n = len(spec.substrates), cells.n
    M = n + 37
    print(f"motility fp32={fp32} sensitivity={use_sens} n={n} ns={ns}", flush=True)
    cb = H.CellBuffer(ctx, cells, ns, max_cells=M)
    r = np.random.default_rng(5)
    vel = [r.uniform(-1, 1, n) for _ in range(3)]
    speed = r.uniform(0.2, 2.0, n)
    for f, v in zip(VEL, vel):
        cb.set_field(f, v)
    cb.set_field(H.F_MOT_SPEED, speed)
    V = spec.nx * spec.ny * spec.nz
    rho = r.uniform(0, 40, ns * V)
    if fp32:
        rho = rho.astype(np.float32).astype(np.float64)
    dens = ctx.alloc(rho.size * (4 if fp32 else 8))
    dens.upload(rho.astype(np.float32) if fp32 else rho)
    dt = 0.1
    seed = 1234
    st = motility_state(n, ns, 8, dt)
    mot = ctx.alloc(lib.hc_motility_bytes(M, ns))
    mot.upload(pack(lib, st, n, M, ns, seed))
    g = H.grid_of(spec, fp32)
    og = O.grid(spec)
    alive = np.asarray(cells.alive, np.uint32)
    x, y, z = (np.ascontiguousarray(np.asarray(a, np.float64)) for a in (cells.x, cells.y, cells.z))
    ov = [v.copy() for v in vel]
    pers, el, mv = st["persistence"].copy(), st["elapsed"].copy(), [v.copy() for v in st["vec"]]
    for step in range(4):
        ctx._check(lib.hc_motility(ctx.h, cb.soa.h, mot.h, dens.h, ctypes.byref(g), dt, n, M, use_sens),
                   "hc_motility")
        ctx.wait()
        O.motility(x, y, z, *ov, alive, speed, og, rho, dt, seed, step, pers, el, st["bias"], *mv,
                   st["motile"], st["r2d"], st["chem_index"], st["chem_dir"], st["sens"] if use_sens else None)
        d_el = read(lib, mot, H.M_ELAPSED, n, M, ns)
        assert np.array_equal(d_el, el), f"step {step}: resample decisions differ"
        for k in range(3):
            d_mv = read(lib, mot, H.M_VEC_X + k, n, M, ns)
            np.testing.assert_allclose(d_mv, mv[k], rtol=1e-12, atol=1e-13, err_msg=f"step {step} vec {k}")
            np.testing.assert_allclose(cb.field(VEL[k]), ov[k], rtol=1e-12, atol=1e-12,
                                       err_msg=f"step {step} velocity {k}")
        resampled = (alive == 1) & (st["motile"] == 1) & (el == 0.0)
        print(f"  step {step}: {resampled.sum()} of {n} resampled", flush=True)
    hdr = mot.download(np.uint64, 2)
    assert hdr[0] == 4 and hdr[1] == seed
    for b in (cb.soa, cb.rates, dens, mot):
        b.free()


def run_sim(spec, cells, periods, use_graph=False, motility=None, mixed=False):
    sim = H.Simulation(spec, cells, mixed=mixed)
    if motility is not None:
        sim.set_motility(**motility)
    sim.run_periods(periods, use_graph=use_graph)
    sim.wait()
    out = {f: sim.field(f) for f in POS + VEL}
    out["rho"] = sim.density()
    sim.close()
    return out


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


@pytest.mark.parametrize("mixed", [False, True])
def test_simulation_with_motility(mixed):
    spec, cells = S.small3d(24, 20, 16, n_cells=1500, seed=13)
    n = cells.n
    r = np.random.default_rng(2)
    base = run_sim(spec, cells, 3, mixed=mixed)
    # nothing motile, or motile at zero speed: the same run as without motility
    assert same(run_sim(spec, cells, 3, mixed=mixed, motility=dict(seed=1)), base)
    zero = dict(seed=1, is_motile=np.ones(n, np.uint32), speed=np.zeros(n))
    assert same(run_sim(spec, cells, 3, mixed=mixed, motility=zero), base)
    mot = dict(seed=77, is_motile=np.ones(n, np.uint32), persistence=r.uniform(0.05, 1.0, n),
               bias=r.uniform(0, 1, n), speed=np.full(n, 0.5), restrict_2d=np.zeros(n, np.uint32),
               chem_index=np.zeros(n, np.int32), chem_dir=np.ones(n, np.int32))
    a = run_sim(spec, cells, 4, mixed=mixed, motility=mot)
    b = run_sim(spec, cells, 4, mixed=mixed, motility=mot)
    c = run_sim(spec, cells, 4, use_graph=True, motility=mot, mixed=mixed)
    assert same(a, b), "not reproducible"
    assert same(a, c), "graph replay differs from eager"
    d = run_sim(spec, cells, 4, mixed=mixed, motility=dict(mot, seed=78))
    assert not np.array_equal(a[H.F_POS_X], d[H.F_POS_X])
    moved = np.abs(a[H.F_POS_X] - run_sim(spec, cells, 4, mixed=mixed)[H.F_POS_X])
    assert moved.max() > 1e-3
