"""Seeded synthetic inputs for the step hot path.

Each builder returns a ``Scenario`` (grid, substrates, schedule, mechanics
grid) and a ``Cells`` population.  Grid sizing follows the reference:
``n = (uint)((max - min) / d)`` (src/microenvironment.mm:33-35) and the 30 um
mechanics grid ``max(1, (uint)(L / 30))`` with ``nz = 1`` in 2-D
(src/cell_mechanics.mm:27-39).

Random draws use ``numpy.random.default_rng(seed)`` (PCG64); the draw order of
every builder is written next to it, so the same seed always gives the same
population.  Positions can be written/read as ``x,y,z,type`` CSV, the format of
the reference's loadCellsCSV (src/geometry.cpp:184-237).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

MECH_VOXEL = 30.0                      # src/cell_mechanics.mm:27
DEFAULT_RADIUS = 8.41                  # config text: radius 8.41 um
DEFAULT_VOLUME = 4.0 / 3.0 * math.pi * DEFAULT_RADIUS ** 3


@dataclass
class Substrate:
    name: str
    D: float
    decay: float
    initial: float
    dirichlet: bool
    dirichlet_value: float


@dataclass
class Scenario:
    name: str
    x_min: float
    x_max: float
    y_min: float
    y_max: float
    z_min: float
    z_max: float
    dx: float
    dy: float
    dz: float
    use_2d: bool
    substrates: List[Substrate]
    dt_diff: float = 0.01
    dt_mech: float = 0.1
    dt_pheno: float = 6.0
    max_time: float = 60.0
    mech_voxel: float = MECH_VOXEL

    @property
    def nx(self):
        return int((self.x_max - self.x_min) / self.dx)

    @property
    def ny(self):
        return int((self.y_max - self.y_min) / self.dy)

    @property
    def nz(self):
        return 1 if self.use_2d else int((self.z_max - self.z_min) / self.dz)

    @property
    def n_voxels(self):
        return self.nx * self.ny * self.nz

    @property
    def mnx(self):
        return max(1, int((self.x_max - self.x_min) / self.mech_voxel))

    @property
    def mny(self):
        return max(1, int((self.y_max - self.y_min) / self.mech_voxel))

    @property
    def mnz(self):
        return 1 if self.use_2d else max(1, int((self.z_max - self.z_min) / self.mech_voxel))

    @property
    def n_mech_voxels(self):
        return self.mnx * self.mny * self.mnz

    def initial_density(self):
        ns = len(self.substrates)
        rho = np.empty(ns * self.n_voxels, np.float64)
        for s, sub in enumerate(self.substrates):
            rho[s * self.n_voxels:(s + 1) * self.n_voxels] = sub.initial
        return rho


@dataclass
class Cells:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    radius: np.ndarray
    volume: np.ndarray
    rep: np.ndarray
    adh: np.ndarray
    rel_adh: np.ndarray
    sec: np.ndarray        # [n, ns]
    upt: np.ndarray
    sat: np.ndarray
    net: np.ndarray
    alive: np.ndarray
    cell_type: np.ndarray = field(default=None)

    @property
    def n(self):
        return len(self.x)


def uniform_cells(x, y, z, ns, radius=DEFAULT_RADIUS, volume=DEFAULT_VOLUME,
                  rep=10.0, adh=0.4, rel_adh=1.25):
    n = len(x)
    f = lambda v: np.full(n, v, np.float64)
    c64 = lambda a: np.ascontiguousarray(a, np.float64)
    return Cells(x=c64(x), y=c64(y), z=c64(z), radius=f(radius), volume=f(volume),
                 rep=f(rep), adh=f(adh), rel_adh=f(rel_adh),
                 sec=np.zeros((n, ns)), upt=np.zeros((n, ns)), sat=np.zeros((n, ns)),
                 net=np.zeros((n, ns)), alive=np.ones(n, np.uint32),
                 cell_type=np.zeros(n, np.uint32))


def hcp_sites_in_sphere(n, sphere_r, spacing, center=(0.0, 0.0, 0.0)):
    """The n hexagonal-close-packed lattice sites nearest ``center`` inside a
    sphere of radius ``sphere_r``.  Lattice: triangular rows in x-y (row pitch
    spacing*sqrt(3)/2, odd rows shifted by spacing/2), A-B layer stacking along
    z with pitch spacing*sqrt(2/3) (B layers shifted by (spacing/2,
    spacing/(2*sqrt(3)))).  Sites are ordered by squared distance, ties by
    lattice enumeration order (k, j, i)."""
    a = float(spacing)
    ry = a * math.sqrt(3.0) / 2.0
    rz = a * math.sqrt(2.0 / 3.0)
    ni = int(math.ceil(sphere_r / a)) + 2
    nj = int(math.ceil(sphere_r / ry)) + 2
    nk = int(math.ceil(sphere_r / rz)) + 2
    k, j, i = np.meshgrid(np.arange(-nk, nk + 1), np.arange(-nj, nj + 1),
                          np.arange(-ni, ni + 1), indexing="ij")
    k = k.ravel(); j = j.ravel(); i = i.ravel()
    b = (k & 1).astype(np.float64)
    x = a * (i + 0.5 * (j & 1)) + b * 0.5 * a
    y = ry * j + b * a / (2.0 * math.sqrt(3.0))
    z = rz * k
    d2 = x * x + y * y + z * z
    inside = d2 <= sphere_r * sphere_r
    x, y, z, d2 = x[inside], y[inside], z[inside], d2[inside]
    if len(x) < n:
        raise ValueError(f"sphere r={sphere_r} holds {len(x)} sites < {n}")
    order = np.argsort(d2, kind="stable")[:n]
    return (x[order] + center[0], y[order] + center[1], z[order] + center[2])


def config0():
    """2-D, 1000x1000x20 um, 20 um voxels (50x50x1), O2 with Dirichlet 38 on
    all edges, 400 cells uniform in a disc of radius 250 um (seed 0).
    Draw order: u = rng.random((400, 2)); r = 250*sqrt(u[:,0]); theta = 2*pi*u[:,1]."""
    sc = Scenario("config0_2d_400", -500, 500, -500, 500, -10, 10, 20, 20, 20, True,
                  [Substrate("oxygen", 1e5, 0.1, 38.0, True, 38.0)], max_time=60.0)
    rng = np.random.default_rng(0)
    u = rng.random((400, 2))
    r = 250.0 * np.sqrt(u[:, 0])
    th = 2.0 * math.pi * u[:, 1]
    c = uniform_cells(r * np.cos(th), r * np.sin(th), np.zeros(400), 1)
    c.upt[:, 0] = 10.0
    return sc, c


def config1(n_cells=20000):
    """3-D, 1000^3 um, 20 um voxels (50^3).  O2 as config 0 (Dirichlet on all
    faces); factor D=1000, decay 0.01, secreted at 1/min to saturation 1,
    zero-flux (no Dirichlet).  n_cells HCP sites (spacing 0.95*2*radius)
    nearest the centre of a sphere r=270 um, +-1 um uniform jitter (seed 1;
    draw order: jitter = rng.uniform(-1, 1, (n, 3)))."""
    sc = Scenario("config1_3d_20k", -500, 500, -500, 500, -500, 500, 20, 20, 20, False,
                  [Substrate("oxygen", 1e5, 0.1, 38.0, True, 38.0),
                   Substrate("factor", 1000.0, 0.01, 0.0, False, 0.0)], max_time=60.0)
    x, y, z = hcp_sites_in_sphere(n_cells, 270.0, 0.95 * 2 * DEFAULT_RADIUS)
    rng = np.random.default_rng(1)
    j = rng.uniform(-1.0, 1.0, (n_cells, 3))
    c = uniform_cells(x + j[:, 0], y + j[:, 1], z + j[:, 2], 2)
    c.upt[:, 0] = 10.0
    c.sec[:, 1] = 1.0
    c.sat[:, 1] = 1.0
    return sc, c


def config2(n_cells=1_000_000, sphere_r=950.0):
    """3-D, 2000^3 um, 10 um voxels (200^3 = 8M voxels), 4 substrates with
    D = 1e5, 1e4, 1e3, 100 and decay 0.1, 0.05, 0.01, 0.001; Dirichlet
    38/0/0/0 on all faces, initial value = Dirichlet value.  n_cells jittered
    HCP sites nearest the centre of a sphere r=950 um.  Seed 1; draw order:
    jitter (n,3) in [-1,1], then uptake (n,4) in [0,10], then secretion (n,4)
    in [0,10].  Saturation: 38 for substrate 0, 1 for the others.
    200 mechanics steps = 20 min."""
    T = [38.0, 0.0, 0.0, 0.0]
    subs = [Substrate(f"s{k}", D, lam, T[k], True, T[k])
            for k, (D, lam) in enumerate(zip([1e5, 1e4, 1e3, 100.0], [0.1, 0.05, 0.01, 0.001]))]
    sc = Scenario("config2_200cube_1M", -1000, 1000, -1000, 1000, -1000, 1000, 10, 10, 10,
                  False, subs, max_time=20.0)
    x, y, z = hcp_sites_in_sphere(n_cells, sphere_r, 0.95 * 2 * DEFAULT_RADIUS)
    rng = np.random.default_rng(1)
    j = rng.uniform(-1.0, 1.0, (n_cells, 3))
    c = uniform_cells(x + j[:, 0], y + j[:, 1], z + j[:, 2], 4)
    c.upt[:, :] = rng.uniform(0.0, 10.0, (n_cells, 4))
    c.sec[:, :] = rng.uniform(0.0, 10.0, (n_cells, 4))
    c.sat[:, 0] = 38.0
    c.sat[:, 1:] = 1.0
    return sc, c


def small3d(nx=12, ny=10, nz=8, n_cells=300, seed=5, ns=2, dirichlet=(True, False)):
    """Small 3-D case for fast parity tests (ragged, non-cubic grid)."""
    dx = 20.0
    subs = [Substrate("oxygen", 1e5, 0.1, 38.0, dirichlet[0], 38.0),
            Substrate("factor", 1000.0, 0.01, 0.0, dirichlet[1] if len(dirichlet) > 1 else False, 0.0)][:ns]
    sc = Scenario("small3d", 0, nx * dx, 0, ny * dx, 0, nz * dx, dx, dx, dx, False, subs,
                  max_time=1.0)
    rng = np.random.default_rng(seed)
    p = rng.random((n_cells, 3)) * np.array([nx * dx, ny * dx, nz * dx])
    c = uniform_cells(p[:, 0], p[:, 1], p[:, 2], ns)
    c.upt[:, 0] = 10.0
    if ns > 1:
        c.sec[:, 1] = 1.0
        c.sat[:, 1] = 1.0
    return sc, c


HCP_DENSITY = math.sqrt(2.0)       # sites per spacing^3 of a close-packed lattice


def hcp_sites_in_spheroid(n, semi_axes, spacing):
    """The n close-packed lattice sites (lattice of ``hcp_sites_in_sphere``)
    with the smallest normalised radius (x/a)^2 + (y/b)^2 + (z/c)^2 inside the
    spheroid of semi-axes (a, b, c) centred at the origin; ties by lattice
    enumeration order."""
    a, b, c = (float(v) for v in semi_axes)
    r = max(a, b, c)
    x, y, z = hcp_sites_in_sphere(_sites_in_sphere_count(r, spacing), r, spacing)
    q = (x / a) ** 2 + (y / b) ** 2 + (z / c) ** 2
    inside = q <= 1.0
    x, y, z, q = x[inside], y[inside], z[inside], q[inside]
    if len(x) < n:
        raise ValueError(f"spheroid {semi_axes} holds {len(x)} sites < {n}")
    order = np.argsort(q, kind="stable")[:n]
    return x[order], y[order], z[order]


def _sites_in_sphere_count(r, spacing):
    """Number of lattice sites inside the sphere of radius r (all of them)."""
    a = float(spacing)
    ry = a * math.sqrt(3.0) / 2.0
    rz = a * math.sqrt(2.0 / 3.0)
    ni = int(math.ceil(r / a)) + 2
    nj = int(math.ceil(r / ry)) + 2
    nk = int(math.ceil(r / rz)) + 2
    k, j, i = np.meshgrid(np.arange(-nk, nk + 1), np.arange(-nj, nj + 1), np.arange(-ni, ni + 1), indexing="ij")
    b = (k & 1).astype(np.float64)
    x = a * (i + 0.5 * (j & 1)) + b * 0.5 * a
    y = ry * j + b * a / (2.0 * math.sqrt(3.0))
    zz = rz * k
    return int(np.count_nonzero(x * x + y * y + zz * zz <= r * r))


def ensemble_member(seed, edges=(32, 128), cells=(1000, 200_000), dx=20.0, fill=0.8):
    """One simulation of the batched ensemble (config-1 type, 3-D).

    Draw order with ``default_rng(seed)``:
      1. edge[3]  = floor(exp(U(ln edges[0], ln edges[1])))  voxels along x, y, z;
      2. n        = floor(exp(U(ln cells[0], ln cells[1])));
      3. f[7]     = U(0.5, 1.5): factors on O2 D, O2 decay, factor D, factor
                    decay, O2 uptake, repulsion, adhesion (config-1 values 1e5,
                    0.1, 1000, 0.01, 10, 10, 0.4);
      4. jitter   = U(-1, 1, (n', 3)) um.
    Cells sit on the close-packed lattice (spacing 0.95 * 2 * radius) inside a
    spheroid with semi-axes 0.4 * L per axis, centred in the domain
    [-L/2, L/2]; n' = min(n, fill * lattice sites in that spheroid), since a
    small grid cannot hold 200k cells.  The factor is secreted at 1/min to
    saturation 1; O2 has Dirichlet 38 on all faces."""
    rng = np.random.default_rng(seed)
    e = np.floor(np.exp(rng.uniform(math.log(edges[0]), math.log(edges[1]), 3))).astype(int)
    n_draw = int(math.floor(math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1])))))
    f = rng.uniform(0.5, 1.5, 7)
    L = e * dx
    subs = [Substrate("oxygen", 1e5 * f[0], 0.1 * f[1], 38.0, True, 38.0),
            Substrate("factor", 1000.0 * f[2], 0.01 * f[3], 0.0, False, 0.0)]
    sc = Scenario(f"member_{seed}", -L[0] / 2, L[0] / 2, -L[1] / 2, L[1] / 2, -L[2] / 2, L[2] / 2,
                  dx, dx, dx, False, subs, max_time=100.0)
    spacing = 0.95 * 2 * DEFAULT_RADIUS
    semi = 0.4 * L
    cap = int(fill * 4.0 / 3.0 * math.pi * semi[0] * semi[1] * semi[2] * HCP_DENSITY / spacing ** 3)
    n = max(1, min(n_draw, cap))
    x, y, z = hcp_sites_in_spheroid(n, semi, spacing)
    j = rng.uniform(-1.0, 1.0, (n, 3))
    c = uniform_cells(x + j[:, 0], y + j[:, 1], z + j[:, 2], 2, rep=10.0 * f[5], adh=0.4 * f[6])
    c.upt[:, 0] = 10.0 * f[4]
    c.sec[:, 1] = 1.0
    c.sat[:, 1] = 1.0
    return sc, c


def config3(n_sims=64, seed0=100, **kw):
    """The 64-simulation ensemble: members ensemble_member(100..163)."""
    return [ensemble_member(seed0 + i, **kw) for i in range(n_sims)]


BUILDERS = {"config0": config0, "config1": config1, "config2": config2, "small3d": small3d}


def write_csv(path, cells):
    """x,y,z,type rows (src/geometry.cpp:184-237 reads this format)."""
    with open(path, "w") as f:
        f.write("x,y,z,type\n")
        t = cells.cell_type if cells.cell_type is not None else np.zeros(cells.n, np.uint32)
        for a, b, c, d in zip(cells.x, cells.y, cells.z, t):
            f.write(f"{a!r},{b!r},{c!r},{int(d)}\n")


def read_csv(path):
    """Rows that start with a digit, sign or '.' are parsed as x,y,z,type;
    anything else (header, comments) is skipped, as the reference does."""
    xs, ys, zs, ts = [], [], [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or not (line[0].isdigit() or line[0] in "+-."):
                continue
            p = line.split(",")
            if len(p) < 4:
                continue
            xs.append(float(p[0])); ys.append(float(p[1])); zs.append(float(p[2])); ts.append(int(p[3]))
    return (np.array(xs), np.array(ys), np.array(zs), np.array(ts, np.uint32))
