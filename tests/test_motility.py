"""Motility: the oracle's restatement (or_motility, or_uniform) pinned against
the reference's own updateMotility (src/motility.cpp:179-293, compiled from
the reference sources into oracle/_ref) on every case its random generator
does not decide (forced resample with bias 1, non-motile and dead cells, no
gradient, no density), and against closed forms and statistics for the cases
it does (unit speed, the 2-D circle, the resample rate).  Host only."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_path.so")


def ref_lib():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    lib = ctypes.CDLL(REF)
    vp, u32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float
    lib.ref_motility.argtypes = ([u32] + [vp] * 8 + [u32] * 4 + [f32] * 6 + [vp, ctypes.c_double] + [vp] * 11 +
                                 [ctypes.c_int])
    return lib


def p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


class Case:
    """n cells on an nx*ny*nz grid (dx 20, origin -100,-80,-60) with ns substrates;
    every value exactly representable in FP32 (the reference's cell store)."""

    def __init__(self, n, nx, ny, nz, ns, seed, dt=6.0):
        r = np.random.default_rng(seed)
        self.n, self.nx, self.ny, self.nz, self.ns, self.dt = n, nx, ny, nz, ns, dt
        self.dx = 20.0
        self.org = (-100.0, -80.0, -60.0)
        # positions: voxel centres (edge voxels included) plus offsets well inside the voxel
        self.x = f32(self.org[0] + (r.integers(0, nx, n) + 0.5 + r.uniform(-0.3, 0.3, n)) * self.dx)
        self.y = f32(self.org[1] + (r.integers(0, ny, n) + 0.5 + r.uniform(-0.3, 0.3, n)) * self.dx)
        self.z = f32(self.org[2] + (r.integers(0, nz, n) + 0.5 + r.uniform(-0.3, 0.3, n)) * self.dx)
        self.vx, self.vy, self.vz = (f32(r.uniform(-1, 1, n)) for _ in range(3))
        self.alive = (r.uniform(size=n) < 0.9).astype(np.uint32)
        self.speed = f32(r.uniform(0.5, 2.0, n))
        self.persistence = np.full(n, 0.5 * dt)           # below dt: always resample
        self.elapsed = r.uniform(0, 5, n)
        self.bias = np.ones(n)
        self.mv = [r.uniform(-1, 1, n) for _ in range(3)]
        self.motile = (r.uniform(size=n) < 0.8).astype(np.uint32)
        self.r2d = (r.uniform(size=n) < 0.5).astype(np.uint32)
        self.chem_index = r.integers(0, ns, n).astype(np.int32)
        self.chem_dir = np.where(r.uniform(size=n) < 0.5, 1, -1).astype(np.int32)
        self.sens = f32(r.uniform(-2, 2, n * ns))
        # substrate-major field [s][k][j][i], FP32 values
        self.rho = f32(r.uniform(0, 50, ns * nx * ny * nz))

    def grid(self):
        return O.OrGrid(self.nx, self.ny, self.nz, self.ns, self.dx, self.dx, self.dx, 0.01, *self.org)

    def run_oracle(self, sensitivity, rho="field", seed=1, step=0):
        out = {k: np.array(getattr(self, k), copy=True) for k in ("vx", "vy", "vz", "elapsed")}
        mv = [m.copy() for m in self.mv]
        field = self.rho if rho == "field" else rho
        O.motility(self.x, self.y, self.z, out["vx"], out["vy"], out["vz"], self.alive, self.speed, self.grid(),
                   field, self.dt, seed, step, self.persistence.copy(), out["elapsed"], self.bias, *mv,
                   self.motile, self.r2d, self.chem_index, self.chem_dir, self.sens if sensitivity else None)
        out["mv"] = mv
        return out

    def run_ref(self, sensitivity, rho="field"):
        V = self.nx * self.ny * self.nz
        field = self.rho if rho == "field" else rho
        # the reference indexes density[voxel*ns + s]
        dens = None if field is None else np.ascontiguousarray(
            field.reshape(self.ns, V).T.reshape(-1).astype(np.float32))
        v = [a.astype(np.float32) for a in (self.vx, self.vy, self.vz)]
        el = self.elapsed.copy()
        mv = [m.copy() for m in self.mv]
        c32 = [a.astype(np.float32) for a in (self.x, self.y, self.z)]
        ref_lib().ref_motility(self.n, *[p(a) for a in c32], *[p(a) for a in v], p(self.alive),
                               p(self.speed.astype(np.float32)), self.nx, self.ny, self.nz, self.ns,
                               self.dx, self.dx, self.dx, *self.org, p(dens), self.dt, p(self.persistence.copy()),
                               p(el), p(self.bias), *[p(m) for m in mv], p(self.motile), p(self.r2d),
                               p(self.chem_index), p(self.chem_dir), p(self.sens), 0 if sensitivity else 1)
        return {"vx": v[0], "vy": v[1], "vz": v[2], "elapsed": el, "mv": mv}


def compare(a, b):
    for k in range(3):
        np.testing.assert_allclose(a["mv"][k], b["mv"][k], rtol=1e-14, atol=1e-15)
    np.testing.assert_array_equal(a["elapsed"], b["elapsed"])
    for k in ("vx", "vy", "vz"):
        # the reference adds float(vector) to its FP32 velocity
        np.testing.assert_allclose(a[k], b[k].astype(np.float64), rtol=0, atol=4e-7 * 3)


@pytest.mark.parametrize("sensitivity", [False, True])
@pytest.mark.parametrize("shape", [(9, 7, 5, 1), (9, 7, 5, 3), (11, 6, 1, 2)])
def test_forced_resample_bias1_matches_reference(sensitivity, shape):
    """bias 1 and persistence below dt: the vector is speed * the normalised
    chemotactic gradient, whatever the random draws."""
    nx, ny, nz, ns = shape
    c = Case(400, nx, ny, nz, ns, seed=nx * 100 + ns)
    a, b = c.run_oracle(sensitivity), c.run_ref(sensitivity)
    compare(a, b)
    live_motile = (c.alive == 1) & (c.motile == 1)
    mag = np.sqrt(sum(m ** 2 for m in a["mv"]))
    assert np.all(np.abs(mag[live_motile] - c.speed[live_motile]) < 1e-12)
    assert np.all(a["elapsed"][live_motile] == 0.0)
    assert np.all(np.array(a["mv"])[:, (c.alive == 1) & (c.motile == 0)] == 0.0)
    dead = c.alive == 0
    assert all(np.array_equal(a["mv"][k][dead], c.mv[k][dead]) for k in range(3))
    if nz == 1:
        assert np.all(a["mv"][2][live_motile] == 0.0)


def test_no_gradient_and_no_density_match_reference():
    """A uniform field (zero gradient) or no field at all gives a zero bias;
    with bias 1 the blended vector is zero and stays zero after normalising."""
    c = Case(200, 8, 6, 4, 2, seed=3)
    c.rho = np.full_like(c.rho, 7.0)
    for sensitivity in (False, True):
        a, b = c.run_oracle(sensitivity), c.run_ref(sensitivity)
        compare(a, b)
        lm = (c.alive == 1) & (c.motile == 1)
        assert all(np.all(a["mv"][k][lm] == 0.0) for k in range(3))
    a, b = c.run_oracle(False, rho=None), c.run_ref(False, rho=None)
    compare(a, b)


def test_linear_field_points_along_axis():
    """rho = a*i along x (and nothing along y, z): the bias is +x (or -x for
    direction -1), also in the edge voxels where the difference is one-sided."""
    c = Case(300, 10, 6, 4, 1, seed=4)
    i = np.arange(c.nx)
    c.rho = np.broadcast_to(3.0 * i, (c.nz, c.ny, c.nx)).reshape(-1).astype(np.float64)
    a = c.run_oracle(False)
    lm = (c.alive == 1) & (c.motile == 1)
    np.testing.assert_allclose(a["mv"][0][lm], c.chem_dir[lm] * c.speed[lm], rtol=1e-15)
    assert np.all(a["mv"][1][lm] == 0.0) and np.all(a["mv"][2][lm] == 0.0)


def test_unbiased_direction_is_unit_times_speed():
    c = Case(20000, 6, 6, 6, 1, seed=5)
    c.bias = np.zeros(c.n)
    a = c.run_oracle(False, seed=99, step=3)
    lm = (c.alive == 1) & (c.motile == 1)
    mv = np.array(a["mv"])
    mag = np.sqrt((mv ** 2).sum(0))
    np.testing.assert_allclose(mag[lm], c.speed[lm], rtol=1e-14)
    flat = lm & (c.r2d == 1)
    assert np.all(mv[2][flat] == 0.0)
    # theta uniform on [0, 2pi): the mean direction vanishes; phi uniform on [0, pi): E[z] = 0
    unit = mv[:, lm] / c.speed[lm]
    assert np.all(np.abs(unit.mean(1)) < 0.03)
    # velocity = old velocity + vector
    np.testing.assert_allclose(a["vx"][lm], c.vx[lm] + mv[0][lm], rtol=0, atol=1e-15)


def test_resample_rate_and_persistence():
    """persistence 10*dt: a direction is resampled with probability dt/pt =
    0.1; the others keep their vector and accumulate elapsed time."""
    c = Case(40000, 6, 6, 6, 1, seed=6)
    c.persistence = np.full(c.n, 10 * c.dt)
    c.bias = np.zeros(c.n)
    a = c.run_oracle(False, seed=7, step=11)
    lm = (c.alive == 1) & (c.motile == 1)
    res = lm & (a["elapsed"] == 0.0)
    frac = res.sum() / lm.sum()
    assert abs(frac - 0.1) < 0.01, frac
    keep = lm & ~res
    np.testing.assert_array_equal(a["elapsed"][keep], c.elapsed[keep] + c.dt)
    assert all(np.array_equal(a["mv"][k][keep], c.mv[k][keep]) for k in range(3))


def test_uniform_stream_properties():
    """or_uniform: in [0, 1), deterministic, and a different number for every
    (seed, step, cell, draw) coordinate; mean 1/2, variance 1/12, flat
    histogram, no correlation between neighbouring cells."""
    u = np.array([O.uniform(5, 2, i, 1) for i in range(40000)])
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    h, _ = np.histogram(u, bins=20, range=(0, 1))
    assert np.all(np.abs(h / len(u) - 0.05) < 0.006)
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.02
    base = O.uniform(5, 2, 17, 1)
    assert base == O.uniform(5, 2, 17, 1)
    others = {O.uniform(6, 2, 17, 1), O.uniform(5, 3, 17, 1), O.uniform(5, 2, 18, 1), O.uniform(5, 2, 17, 0),
              O.uniform(5, 2, 17, 2)}
    assert base not in others and len(others) == 5
