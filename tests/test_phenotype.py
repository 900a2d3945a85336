"""Phenotype-side SoA mutation: the oracle's restatement (or_volume_update,
or_divide, or_remove) pinned against the reference's own updateVolume,
divideCell and removeCell (src/cell_phenotype.cpp, src/cell_data.cpp,
compiled from the reference sources into oracle/_ref), and the library's
host-side removal plan (hc_remove_plan) against the oracle.  Host only."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H

REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_path.so")


def ref_lib():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    lib = ctypes.CDLL(REF)
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    lib.ref_phenotype_op.argtypes = [ctypes.c_int, u32, u32, u32, vp, vp, vp, vp, vp, u32, ctypes.c_double,
                                     ctypes.c_int, vp]
    return lib


def p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def make_cells(n, M, ns, seed):
    """Random per-cell state in the library's layout; FP32-representable
    values where the reference stores FP32."""
    r = np.random.default_rng(seed)
    f64 = np.zeros((23, M))
    f64[0:3, :n] = f32(r.uniform(-400, 400, (3, n)))
    f64[3:9, :n] = f32(r.uniform(-1, 1, (6, n)))
    f64[11, :n] = f32(r.uniform(1500, 4000, n))                 # total volume
    f64[12, :n] = f32(f64[11, :n] * r.uniform(0.15, 0.3, n))    # nuclear volume
    f64[9, :n] = f32(np.cbrt(f64[11, :n] / 4.188790204786391))
    f64[10, :n] = f32(np.cbrt(f64[12, :n] / 4.188790204786391))
    f64[13:16, :n] = f32(r.uniform(0.2, 12, (3, n)))
    f64[16:21, :n] = f32(r.uniform(0, 1, (5, n)))
    f64[21, :n] = f32(np.where(r.uniform(size=n) < 0.1, 0.0, r.uniform(0.5, 2, n)))
    f64[22, :n] = f32(r.uniform(0, 3, n))
    u32 = np.zeros((5, M), np.uint32)
    u32[0, :n] = r.integers(0, 3, n)
    u32[1, :n] = r.integers(0, 4, n)
    u32[2, :n] = (r.uniform(size=n) < 0.9)
    u32[3, :n] = r.integers(0, 5000, n)
    u32[4, :n] = r.integers(0, 900, n)
    vol = np.zeros((15, M))
    V, Vn = f64[11, :n], f64[12, :n]
    vol[H.V_FLUID, :n] = 0.75 * V * r.uniform(0.8, 1.2, n)
    vol[H.V_SOLID_NUC, :n] = 0.25 * Vn * r.uniform(0.8, 1.2, n)
    vol[H.V_SOLID_CYTO, :n] = 0.25 * (V - Vn) * r.uniform(0.8, 1.2, n)
    vol[H.V_FLUID_FRAC, :n] = r.uniform(0.6, 0.8, n)
    vol[H.V_CALC_FRAC, :n] = r.uniform(0, 0.1, n)
    vol[H.V_TARGET_SOLID_CYTO, :n] = r.uniform(300, 700, n)
    vol[H.V_TARGET_SOLID_NUC, :n] = r.uniform(100, 200, n)
    vol[H.V_TARGET_FLUID_FRAC, :n] = r.uniform(0.6, 0.8, n)
    vol[H.V_TARGET_CN_RATIO, :n] = r.uniform(2, 5, n)
    vol[H.V_FLUID_RATE, :n] = r.uniform(0, 0.1, n)
    vol[H.V_NUC_RATE, :n] = r.uniform(0, 0.01, n)
    vol[H.V_CYTO_RATE, :n] = r.uniform(0, 0.01, n)
    vol[H.V_CALC_RATE, :n] = r.uniform(0, 0.001, n)
    vol[H.V_RUPTURE_VOL, :n] = 2 * V
    vol[H.V_REL_RUPTURE_VOL, :n] = 2.0
    rates = np.zeros((4, M * ns))
    rates[:, :n * ns] = r.uniform(0, 10, (4, n * ns))
    return f64, u32, vol, rates


def run_ref(op, n, M, ns, state, lst=None, dt=0.0, use_ratio=1):
    f64, u32, vol, rates = (np.ascontiguousarray(a.copy()) for a in state)
    out = ctypes.c_uint32()
    lst_arr = None if lst is None else np.ascontiguousarray(lst, np.uint32)
    count = 0 if lst_arr is None else len(lst_arr)
    ref_lib().ref_phenotype_op(op, n, M, ns, p(f64), p(u32), p(vol), p(rates), p(lst_arr), count, dt, use_ratio,
                               ctypes.byref(out))
    return (f64, u32, vol, rates), out.value


def oracle_cells(state, ns):
    f64, u32, vol, rates = (np.ascontiguousarray(a.copy()) for a in state)
    return O.PhenoCells(f64, u32, rates, vol, ns)


@pytest.mark.parametrize("use_ratio", [0, 1])
def test_volume_update_matches_reference(use_ratio):
    n, M, ns = 700, 800, 2
    state = make_cells(n, M, ns, 3 + use_ratio)
    mask = (np.random.default_rng(9).uniform(size=n) < 0.8).astype(np.uint32)
    for dt in (0.1, 6.0):
        (rf, ru, rv, rr), _ = run_ref(0, n, M, ns, state, mask, dt, use_ratio)
        c = oracle_cells(state, ns)
        O.volume_update(c, mask, n, dt, use_ratio)
        # compartments, fractions and targets: FP64 in both, the same operations
        assert np.array_equal(c.vol, rv)
        # volumes and radii: the reference stores FP32
        for f in (H.F_TOTAL_VOL, H.F_NUCLEAR_VOL, H.F_RADIUS, H.F_NUCLEAR_RAD):
            assert np.array_equal(c.f64[f].astype(np.float32).astype(np.float64), rf[f]), f
        untouched = np.r_[np.flatnonzero(mask == 0), np.arange(n, M)]
        assert np.array_equal(c.f64[:, untouched], state[0][:, untouched])
        moved = np.flatnonzero(mask)
        assert not np.array_equal(c.vol[H.V_FLUID, moved], state[2][H.V_FLUID, moved])


def test_remove_matches_reference():
    n, M, ns = 500, 520, 3
    state = make_cells(n, M, ns, 11)
    r = np.random.default_rng(12)
    for idx in (r.choice(n, 60, replace=False), np.arange(n - 10, n), np.array([0, 1, 2, n - 1]),
                np.array([], np.uint32), r.choice(n, n, replace=False)):
        (rf, ru, rv, rr), rn = run_ref(2, n, M, ns, state, idx)
        c = oracle_cells(state, ns)
        m = O.remove(c, idx, n)
        assert m == rn == n - len(idx)
        assert np.array_equal(c.f64[:, :m], rf[:, :m])
        assert np.array_equal(c.u32[:, :m], ru[:, :m])
        assert np.array_equal(c.vol[:, :m], rv[:, :m])
        assert np.array_equal(c.rates[:, :m * ns], rr[:, :m * ns])


def test_remove_plan_is_the_removal_loop():
    """hc_remove_plan (library, host) gives exactly the moves of the
    sequential loop: label each cell with its index, remove with the oracle,
    and read the moves off the labels."""
    lib = H.load()
    r = np.random.default_rng(13)
    for trial in range(300):
        n = int(r.integers(1, 200))
        k = int(r.integers(0, n + 1))
        idx = r.choice(n, k, replace=False).astype(np.uint32)
        f64 = np.zeros((23, n))
        f64[0] = np.arange(n)
        c = O.PhenoCells(f64, np.zeros((5, n), np.uint32), np.zeros((4, n)), np.zeros((15, n)), 1)
        m = O.remove(c, idx, n)
        want = {(d, int(c.f64[0, d])) for d in range(m) if c.f64[0, d] != d}
        dst, src = np.zeros(max(k, 1), np.uint32), np.zeros(max(k, 1), np.uint32)
        npairs = ctypes.c_uint32()
        assert lib.hc_remove_plan(p(idx), k, n, p(dst), p(src), ctypes.byref(npairs)) == 0
        got = set(zip(dst[:npairs.value].tolist(), src[:npairs.value].tolist()))
        assert got == want, (n, idx)
        assert all(s >= m > d for d, s in got)
    bad = np.array([3, 3], np.uint32)
    assert lib.hc_remove_plan(p(bad), 2, 10, None, None, ctypes.byref(npairs)) != 0
    bad = np.array([10], np.uint32)
    assert lib.hc_remove_plan(p(bad), 1, 10, None, None, ctypes.byref(npairs)) != 0


def test_divide_matches_reference():
    """Everything divideCell does not draw at random agrees with the
    reference (FP32 fields to FP32 rounding); the drawn position lies half a
    radius from the parent in the x-y plane, and the oncoprotein noise is
    +x for the daughter and -x for the parent."""
    n, ns = 400, 2
    r = np.random.default_rng(21)
    parents = r.choice(n, 50, replace=False).astype(np.uint32)
    for M in (n + 80, n + 30):                # room for all; the last 20 skipped
        state = make_cells(n, M, ns, 22)
        (rf, ru, rv, rr), rn = run_ref(1, n, M, ns, state, parents)
        c = oracle_cells(state, ns)
        nn = O.divide(c, parents, n, M, seed=5, counter=0)
        assert nn == rn == min(n + len(parents), M)
        done = nn - n
        pa, da = parents[:done], np.arange(n, nn)
        rows = [f for f in range(23) if f not in (H.F_POS_X, H.F_POS_Y, H.F_ONCOPROTEIN)]
        for f in rows:
            assert np.array_equal(c.f64[f, :nn].astype(np.float32).astype(np.float64), rf[f, :nn]), f
        assert np.array_equal(c.u32[:, :nn], ru[:, :nn])
        assert np.array_equal(c.vol[:, :nn], rv[:, :nn])
        assert np.array_equal(c.rates[:, :nn * ns], rr[:, :nn * ns])
        # untouched cells
        rest = np.setdiff1d(np.arange(n), pa)
        assert np.array_equal(c.f64[:, rest], state[0][:, rest])
        for fx, fy in ((c.f64[0], c.f64[1]), (rf[0], rf[1])):
            dist = np.hypot(fx[da] - fx[pa], fy[da] - fy[pa])
            np.testing.assert_allclose(dist, 0.5 * c.f64[H.F_RADIUS, pa], rtol=2e-5)
        onco0 = state[0][H.F_ONCOPROTEIN, pa]
        od, op = c.f64[H.F_ONCOPROTEIN, da], c.f64[H.F_ONCOPROTEIN, pa]
        both = (od > 0) & (op > 0)
        np.testing.assert_allclose(od[both] + op[both], 2 * onco0[both], rtol=1e-14)
        assert np.all(od[onco0 == 0] == 0) and np.all(op[onco0 == 0] == 0)


def test_division_draws():
    """The angle is uniform and the oncoprotein noise is N(0, 0.02 level):
    many divisions of level-1 cells."""
    n = 20000
    f64 = np.zeros((23, 2 * n))
    f64[H.F_TOTAL_VOL, :n] = 2494.0
    f64[H.F_NUCLEAR_VOL, :n] = 540.0
    f64[H.F_ONCOPROTEIN, :n] = 1.0
    c = O.PhenoCells(f64, np.zeros((5, 2 * n), np.uint32), np.zeros((4, 2 * n)), np.zeros((15, 2 * n)), 1)
    O.divide(c, np.arange(n, dtype=np.uint32), n, 2 * n, seed=77, counter=3)
    noise = c.f64[H.F_ONCOPROTEIN, n:] - 1.0
    assert abs(noise.mean()) < 0.02 * 0.03 and abs(noise.std() / 0.02 - 1) < 0.03
    ang = np.arctan2(c.f64[H.F_POS_Y, n:], c.f64[H.F_POS_X, n:])
    h, _ = np.histogram(ang, bins=8, range=(-np.pi, np.pi))
    assert np.all(np.abs(h / n - 1 / 8) < 0.015)
    # another counter draws other numbers
    c2 = O.PhenoCells(f64.copy(), np.zeros((5, 2 * n), np.uint32), np.zeros((4, 2 * n)), np.zeros((15, 2 * n)), 1)
    c2.f64[:, n:] = 0
    c2.f64[H.F_ONCOPROTEIN, :n] = 1.0
    c2.f64[H.F_TOTAL_VOL, :n] = 2494.0
    c2.f64[H.F_NUCLEAR_VOL, :n] = 540.0
    O.divide(c2, np.arange(n, dtype=np.uint32), n, 2 * n, seed=77, counter=4)
    assert not np.array_equal(c2.f64[H.F_POS_X, n:], c.f64[H.F_POS_X, n:])
