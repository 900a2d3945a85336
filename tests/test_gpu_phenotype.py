"""Phenotype-side SoA mutation on the device (hc_volume_update, hc_divide,
hc_remove) against the oracle (or_volume_update, or_divide, or_remove) on
the same inputs: every copied, halved or inherited value bit for bit, the
volume compartments bit for bit, radii within 1e-15 relative (the device's
cbrt against the host's), drawn positions and oncoprotein levels within
1e-12; removal moves every buffer (cell SoA, rates, volume and motility
state) exactly as the sequential loop does."""
import ctypes

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H

from test_phenotype import make_cells

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context()
    yield c
    c.close()


def p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Dev:
    """The four device buffers of a cell population, filled from host state."""

    def __init__(self, ctx, state, M, ns, mot=None):
        lib = H.load()
        self.ctx, self.M, self.ns = ctx, M, ns
        f64, u32, vol, rates = state
        self.cells = ctx.alloc(lib.hc_cells_bytes(M))
        self.cells.upload(np.ascontiguousarray(f64).reshape(-1))
        self.cells.upload(np.ascontiguousarray(u32).reshape(-1), offset=lib.hc_cells_field_offset(23, M))
        self.rates = ctx.alloc(lib.hc_rates_bytes(M, ns))
        self.rates.upload(np.ascontiguousarray(rates).reshape(-1))
        self.vol = ctx.alloc(lib.hc_volume_bytes(M))
        self.vol.upload(np.array([0, 99], np.uint64))
        self.vol.upload(np.ascontiguousarray(vol).reshape(-1), offset=lib.hc_volume_field_offset(0, M))
        self.mot = None
        if mot is not None:
            from test_gpu_motility import pack
            self.mot = ctx.alloc(lib.hc_motility_bytes(M, ns))
            self.mot.upload(pack(lib, mot, M, M, ns, 1))

    def state(self):
        lib = H.load()
        M, ns = self.M, self.ns
        f64 = self.cells.download(np.float64, 23 * M).reshape(23, M)
        u32 = self.cells.download(np.uint32, 5 * M, lib.hc_cells_field_offset(23, M)).reshape(5, M)
        vol = self.vol.download(np.float64, 15 * M, lib.hc_volume_field_offset(0, M)).reshape(15, M)
        rates = self.rates.download(np.float64, 4 * M * ns).reshape(4, M * ns)
        return f64, u32, vol, rates

    def free(self):
        for b in (self.cells, self.rates, self.vol, self.mot):
            if b is not None:
                b.free()


def test_volume_update_matches_oracle(ctx):
    lib = H.load()
    n, M, ns = 5000, 5100, 2
    state = make_cells(n, M, ns, 31)
    mask = (np.random.default_rng(32).uniform(size=n) < 0.7).astype(np.uint32)
    dev = Dev(ctx, state, M, ns)
    dmask = ctx.alloc(n * 4)
    dmask.upload(mask)
    c = O.PhenoCells(*(np.ascontiguousarray(a.copy()) for a in (state[0], state[1], state[3], state[2])), ns)
    for rep, (dt, use_ratio) in enumerate([(6.0, 1), (6.0, 0), (0.1, 1)]):
        print(f"volume update {rep}: dt={dt} use_ratio={use_ratio}", flush=True)
        ctx._check(lib.hc_volume_update(ctx.h, dev.cells.h, dev.vol.h, dmask.h, n, M, dt, use_ratio),
                   "hc_volume_update")
        O.volume_update(c, mask, n, dt, use_ratio)
        f64, u32, vol, rates = dev.state()
        assert np.array_equal(vol, c.vol)
        for f in range(23):
            if f in (H.F_RADIUS, H.F_NUCLEAR_RAD):
                np.testing.assert_allclose(f64[f], c.f64[f], rtol=1e-15, atol=0)
            else:
                assert np.array_equal(f64[f], c.f64[f]), f
    dmask.free()
    dev.free()


def test_divide_and_remove_match_oracle(ctx):
    lib = H.load()
    n, M, ns = 3000, 3300, 2
    state = make_cells(n, M, ns, 41)
    from test_gpu_motility import motility_state
    mot = motility_state(M, ns, 43, 0.1)
    dev = Dev(ctx, state, M, ns, mot)
    c = O.PhenoCells(*(np.ascontiguousarray(a.copy()) for a in (state[0], state[1], state[3], state[2])), ns,
                     mot={k: (v.copy() if k != "vec" else [x.copy() for x in v]) for k, v in mot.items()})
    r = np.random.default_rng(44)
    cur = n
    for call in range(3):
        parents = r.choice(cur, 120, replace=False).astype(np.uint32)
        print(f"divide call {call}: n={cur}, {len(parents)} parents", flush=True)
        out = ctypes.c_uint32()
        ctx._check(lib.hc_divide(ctx.h, dev.cells.h, dev.rates.h, dev.vol.h, dev.mot.h, ns, p(parents),
                                 len(parents), cur, M, ctypes.byref(out)), "hc_divide")
        want = O.divide(c, parents, cur, M, seed=99, counter=call)
        assert out.value == want == min(cur + 120, M)
        # live cells only: past the count, a removal leaves unspecified
        # contents (the sequential loop shuffles the tail, the device does not)
        cur = want
        f64, u32, vol, rates = dev.state()
        drawn = (H.F_POS_X, H.F_POS_Y, H.F_ONCOPROTEIN)
        for f in range(23):
            if f in (H.F_RADIUS, H.F_NUCLEAR_RAD):
                np.testing.assert_allclose(f64[f, :cur], c.f64[f, :cur], rtol=1e-15, atol=0)
            elif f in drawn:
                np.testing.assert_allclose(f64[f, :cur], c.f64[f, :cur], rtol=0, atol=1e-12)
            else:
                assert np.array_equal(f64[f, :cur], c.f64[f, :cur]), f
        assert np.array_equal(u32[:, :cur], c.u32[:, :cur]) and np.array_equal(vol[:, :cur], c.vol[:, :cur])
        assert np.array_equal(rates[:, :cur * ns], c.rates[:, :cur * ns])
        victims = r.choice(cur, 150, replace=False).astype(np.uint32)
        print(f"remove call {call}: n={cur}, {len(victims)} removed", flush=True)
        ctx._check(lib.hc_remove(ctx.h, dev.cells.h, dev.rates.h, dev.vol.h, dev.mot.h, ns, p(victims),
                                 len(victims), cur, M, ctypes.byref(out)), "hc_remove")
        want = O.remove(c, victims, cur)
        assert out.value == want == cur - 150
        cur = want
        f64, u32, vol, rates = dev.state()
        for f in range(23):
            if f in (H.F_RADIUS, H.F_NUCLEAR_RAD):
                np.testing.assert_allclose(f64[f, :cur], c.f64[f, :cur], rtol=1e-15, atol=0)
            elif f in drawn:
                np.testing.assert_allclose(f64[f, :cur], c.f64[f, :cur], rtol=0, atol=1e-12)
            else:
                assert np.array_equal(f64[f, :cur], c.f64[f, :cur]), f
        assert np.array_equal(u32[:, :cur], c.u32[:, :cur]) and np.array_equal(vol[:, :cur], c.vol[:, :cur])
        assert np.array_equal(rates[:, :cur * ns], c.rates[:, :cur * ns])
        from test_gpu_motility import read
        for fld, key in ((H.M_PERSISTENCE, "persistence"), (H.M_ELAPSED, "elapsed"), (H.M_BIAS, "bias")):
            assert np.array_equal(read(lib, dev.mot, fld, cur, M, ns), c.mot[key][:cur])
        assert np.array_equal(read(lib, dev.mot, H.M_MOTILE, cur, M, ns, np.uint32), c.mot["motile"][:cur])
        assert np.array_equal(read(lib, dev.mot, H.M_CHEM_DIR, cur, M, ns, np.int32), c.mot["chem_dir"][:cur])
    hdr = dev.vol.download(np.uint64, 2)
    assert hdr[0] == 3 and hdr[1] == 99
    dev.free()


def test_errors(ctx):
    lib = H.load()
    n, M, ns = 100, 120, 1
    dev = Dev(ctx, make_cells(n, M, ns, 51), M, ns)
    out = ctypes.c_uint32()
    dup = np.array([3, 5, 3], np.uint32)
    with pytest.raises(H.HipError, match="distinct"):
        ctx._check(lib.hc_divide(ctx.h, dev.cells.h, dev.rates.h, dev.vol.h, None, ns, p(dup), 3, n, M,
                                 ctypes.byref(out)), "hc_divide")
    big = np.array([n], np.uint32)
    with pytest.raises(H.HipError, match="distinct"):
        ctx._check(lib.hc_remove(ctx.h, dev.cells.h, dev.rates.h, dev.vol.h, None, ns, p(big), 1, n, M,
                                 ctypes.byref(out)), "hc_remove")
    dev.free()


@pytest.mark.parametrize("mixed", [False, True])
def test_simulation_update_phenotype(mixed):
    """pc_sim_update_phenotype between periods: the cell SoA afterwards equals
    the oracle's volume update, divisions and removals applied to the SoA
    before (the volume state given to set_volume), and the simulation keeps
    stepping with the new cell count."""
    from physicell_amd import scenarios as S
    spec, cells = S.small3d(24, 20, 16, n_cells=1500, seed=61)
    n = cells.n
    sim = H.Simulation(spec, cells, max_cells=n + 400, mixed=mixed)
    M = n + 400
    r = np.random.default_rng(62)
    vol = {H.V_FLUID: 0.75 * 2494 * r.uniform(0.9, 1.1, n), H.V_TARGET_SOLID_NUC: r.uniform(100, 200, n),
           H.V_NUC_RATE: r.uniform(0, 0.01, n), H.V_CYTO_RATE: r.uniform(0, 0.01, n)}
    sim.set_volume(seed=7, fields=vol)
    sim.run_periods(2)
    sim.wait()
    before_f = np.zeros((23, M))
    before_u = np.zeros((5, M), np.uint32)
    for f in range(23):
        before_f[f, :n] = sim.field(f)
    for f in range(5):
        before_u[f, :n] = sim.field(23 + f)
    # the volume state set_volume made (fresh values where none given)
    V, Vn, ff = 2494.0, 540.0, 0.75
    fresh = [ff * V, Vn - ff * Vn, (V - Vn) - ff * (V - Vn), ff, 0.0, (V - Vn) - ff * (V - Vn), Vn - ff * Vn, ff,
             3.62, 3.0 / 60.0, 0.33 / 60.0, 0.27 / 60.0, 0.0, 2.0 * V, 2.0]
    vstate = np.array([np.full(M, v) for v in fresh])
    for f, a in vol.items():
        vstate[f, :n] = a
    ns = len(spec.substrates)
    c = O.PhenoCells(before_f, before_u, np.zeros((4, M * ns)), vstate, ns)
    mask = (r.uniform(size=n) < 0.9).astype(np.uint32)
    parents = r.choice(n, 100, replace=False).astype(np.uint32)
    removed = r.choice(n + 100, 80, replace=False).astype(np.uint32)
    print(f"update_phenotype mixed={mixed}: n={n}, 100 divisions, 80 removals", flush=True)
    n2 = sim.update_phenotype(6.0, mask=mask, parents=parents, removed=removed)
    O.volume_update(c, mask, n, 6.0, 1)
    n1 = O.divide(c, parents, n, M, seed=7, counter=0)
    want = O.remove(c, removed, n1)
    assert n2 == want == n + 100 - 80
    for f in range(23):
        got = sim.field(f)
        if f in (H.F_RADIUS, H.F_NUCLEAR_RAD):
            np.testing.assert_allclose(got, c.f64[f, :n2], rtol=1e-15, atol=0)
        elif f in (H.F_POS_X, H.F_POS_Y, H.F_ONCOPROTEIN):
            np.testing.assert_allclose(got, c.f64[f, :n2], rtol=0, atol=1e-12)
        else:
            assert np.array_equal(got, c.f64[f, :n2]), f
    for f in range(5):
        got = sim.field(23 + f)
        if 23 + f == H.F_VOXEL_IDX:
            continue                       # refreshed from the positions afterwards
        assert np.array_equal(got, c.u32[f, :n2]), 23 + f
    sim.run_periods(3)
    sim.wait()
    assert np.all(np.isfinite(sim.density()))
    assert np.all(np.isfinite(sim.field(H.F_POS_X)))
    sim.close()
