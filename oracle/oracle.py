"""ctypes binding of the FP64 oracle (liboracle.so) and of the reference's own
processSecretion build (oracle/_ref/libref_path.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker / the timed CPU
baseline.  The product path never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_path.so")

u32p = ctypes.POINTER(ctypes.c_uint32)
f64p = ctypes.POINTER(ctypes.c_double)
f32p = ctypes.POINTER(ctypes.c_float)
i32p = ctypes.POINTER(ctypes.c_int32)


class OrGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
                ("ns", ctypes.c_uint32), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("dz", ctypes.c_double), ("dt", ctypes.c_double), ("x_min", ctypes.c_double),
                ("y_min", ctypes.c_double), ("z_min", ctypes.c_double)]


class OrSub(ctypes.Structure):
    _fields_ = [("D", ctypes.c_double), ("decay", ctypes.c_double), ("target", ctypes.c_double),
                ("dirichlet", ctypes.c_int32), ("_pad", ctypes.c_int32)]


class OrMech(ctypes.Structure):
    _fields_ = [("mnx", ctypes.c_uint32), ("mny", ctypes.c_uint32), ("mnz", ctypes.c_uint32),
                ("_pad", ctypes.c_uint32), ("voxel", ctypes.c_double), ("max_dist", ctypes.c_double),
                ("x_min", ctypes.c_double), ("y_min", ctypes.c_double), ("z_min", ctypes.c_double)]


class OrState(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32)] + [
        (k, f64p) for k in ("x", "y", "z", "vx", "vy", "vz", "pvx", "pvy", "pvz", "radius",
                            "volume", "rep", "adh", "rel_adh", "pressure", "sec", "upt",
                            "sat", "net")] + [
        ("alive", u32p), ("voxel", u32p), ("mech_key", u32p), ("rho", f64p)]


class OrSchedule(ctypes.Structure):
    _fields_ = [("dt_diff", ctypes.c_double), ("dt_mech", ctypes.c_double),
                ("dt_pheno", ctypes.c_double), ("max_time", ctypes.c_double)]


class OrCounts(ctypes.Structure):
    _fields_ = [("substeps", ctypes.c_int64), ("mech_steps", ctypes.c_int64),
                ("pheno_steps", ctypes.c_int64)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def ref_lib():
    """The reference's own processSecretion, or None when it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        _ref = ctypes.CDLL(REF_LIB_PATH)
    return _ref


def P(a, t=f64p):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("oracle arrays must be C-contiguous")
    want = {u32p: np.uint32, f32p: np.float32, i32p: np.int32}.get(t, np.float64)
    if a.dtype != want:
        raise TypeError(f"oracle array dtype {a.dtype}, expected {np.dtype(want)}")
    return a.ctypes.data_as(t)


def grid(spec):
    """spec: physicell_amd.scenarios.Scenario (or anything with the same fields)."""
    return OrGrid(spec.nx, spec.ny, spec.nz, len(spec.substrates), spec.dx, spec.dy, spec.dz,
                  spec.dt_diff, spec.x_min, spec.y_min, spec.z_min)


def subs(spec):
    arr = (OrSub * max(1, len(spec.substrates)))()
    for i, s in enumerate(spec.substrates):
        arr[i] = OrSub(s.D, s.decay, s.dirichlet_value, 1 if s.dirichlet else 0, 0)
    return arr


def mech(spec):
    return OrMech(spec.mnx, spec.mny, spec.mnz, 0, spec.mech_voxel, spec.mech_voxel,
                  spec.x_min, spec.y_min, spec.z_min)


def sweep(axis, rho, g, s, nthreads=1):
    lib().or_sweep(ctypes.c_int(axis), P(rho), ctypes.byref(g), s, ctypes.c_int(nthreads))


def dirichlet(rho, g, s):
    lib().or_dirichlet(P(rho), ctypes.byref(g), s)


def diffusion_substep(rho, g, s, nthreads=1):
    lib().or_diffusion_substep(P(rho), ctypes.byref(g), s, ctypes.c_int(nthreads))


def voxel_keys(x, y, z, g):
    out = np.zeros(len(x), np.uint32)
    lib().or_voxel_keys(P(x), P(y), P(z), ctypes.c_uint32(len(x)), ctypes.byref(g), P(out, u32p))
    return out


def secretion(rho, g, dt, voxel, volume, sec, upt, sat, net, alive=None, grouped=False, nthreads=1):
    n = len(voxel)
    al = P(alive, u32p) if alive is not None else None
    if grouped:
        lib().or_secretion_grouped(P(rho), ctypes.byref(g), ctypes.c_double(dt), ctypes.c_uint32(n),
                                   P(voxel, u32p), P(volume), P(sec), P(upt), P(sat), P(net), al,
                                   ctypes.c_int(nthreads))
    else:
        lib().or_secretion(P(rho), ctypes.byref(g), ctypes.c_double(dt), ctypes.c_uint32(n),
                           P(voxel, u32p), P(volume), P(sec), P(upt), P(sat), P(net), al)


def mech_keys(x, y, z, m):
    out = np.zeros(len(x), np.uint32)
    lib().or_mech_keys(P(x), P(y), P(z), ctypes.c_uint32(len(x)), ctypes.byref(m), P(out, u32p))
    return out


def hash_sort(keys, nbins):
    off = np.zeros(nbins + 1, np.uint32)
    srt = np.zeros(len(keys), np.uint32)
    lib().or_hash(P(keys, u32p), ctypes.c_uint32(len(keys)), ctypes.c_uint32(nbins),
                  P(off, u32p), P(srt, u32p))
    return off, srt


def forces(x, y, z, radius, rep, adh, rel, keys, off, srt, m, nthreads=1):
    n = len(x)
    vx, vy, vz, pr = (np.zeros(n) for _ in range(4))
    lib().or_forces(P(x), P(y), P(z), P(radius), P(rep), P(adh), P(rel), ctypes.c_uint32(n),
                    P(keys, u32p), P(off, u32p), P(srt, u32p), ctypes.byref(m),
                    P(vx), P(vy), P(vz), P(pr), ctypes.c_int(nthreads))
    return vx, vy, vz, pr


def neighbors(x, y, z, keys, off, srt, m, cap=128):
    n = len(x)
    counts = np.zeros(n, np.uint32)
    idx = np.zeros(n * cap, np.uint32)
    lib().or_neighbors(P(x), P(y), P(z), ctypes.c_uint32(n), P(keys, u32p), P(off, u32p),
                       P(srt, u32p), ctypes.byref(m), ctypes.c_uint32(cap), P(counts, u32p),
                       P(idx, u32p))
    return counts, idx.reshape(n, cap)


def integrate(x, y, z, vx, vy, vz, pvx, pvy, pvz, m, dt, nthreads=1):
    lib().or_integrate(P(x), P(y), P(z), P(vx), P(vy), P(vz), P(pvx), P(pvy), P(pvz),
                       ctypes.c_uint32(len(x)), ctypes.byref(m), ctypes.c_double(dt),
                       ctypes.c_int(nthreads))


class OracleSim:
    """Whole-simulation state driven by or_run / or_mech_step (numpy-owned arrays)."""

    def __init__(self, spec, cells):
        self.spec = spec
        self.g = grid(spec)
        self.s = subs(spec)
        self.m = mech(spec)
        n = cells.n
        ns = len(spec.substrates)
        self.arr = {}
        for k in ("x", "y", "z"):
            self.arr[k] = np.array(getattr(cells, k), dtype=np.float64, copy=True)
        for k in ("vx", "vy", "vz", "pvx", "pvy", "pvz", "pressure"):
            self.arr[k] = np.zeros(n)
        for k in ("radius", "volume", "rep", "adh", "rel_adh"):
            self.arr[k] = np.array(getattr(cells, k), dtype=np.float64, copy=True)
        for k in ("sec", "upt", "sat", "net"):
            self.arr[k] = np.ascontiguousarray(np.array(getattr(cells, k), dtype=np.float64).reshape(n * ns if ns else 0))
        self.arr["alive"] = np.array(cells.alive, dtype=np.uint32, copy=True)
        self.arr["mech_key"] = np.zeros(n, np.uint32)
        self.arr["rho"] = spec.initial_density().copy()
        self.arr["voxel"] = voxel_keys(self.arr["x"], self.arr["y"], self.arr["z"], self.g)
        self.off = np.zeros(spec.mnx * spec.mny * spec.mnz + 1, np.uint32)
        self.srt = np.zeros(max(1, n), np.uint32)
        st = OrState()
        st.n = n
        for k, _t in OrState._fields_[1:]:
            a = self.arr[k]
            setattr(st, k, P(a, u32p) if a.dtype == np.uint32 else P(a))
        self.st = st

    def run(self, max_time, nthreads=1):
        sch = OrSchedule(self.spec.dt_diff, self.spec.dt_mech, self.spec.dt_pheno, max_time)
        c = OrCounts()
        lib().or_run(ctypes.byref(self.st), ctypes.byref(self.g), self.s, ctypes.byref(self.m),
                     ctypes.byref(sch), P(self.off, u32p), P(self.srt, u32p),
                     ctypes.c_int(nthreads), ctypes.byref(c))
        return c.substeps, c.mech_steps, c.pheno_steps

    def substep(self, nthreads=1):
        diffusion_substep(self.arr["rho"], self.g, self.s, nthreads)
        a = self.arr
        secretion(a["rho"], self.g, self.spec.dt_diff, a["voxel"], a["volume"], a["sec"], a["upt"],
                  a["sat"], a["net"], a["alive"], grouped=nthreads > 1, nthreads=nthreads)

    def mech_step(self, nthreads=1):
        lib().or_mech_step(ctypes.byref(self.st), ctypes.byref(self.m),
                           ctypes.c_double(self.spec.dt_mech), P(self.off, u32p),
                           P(self.srt, u32p), ctypes.c_int(nthreads))

    def refresh_voxels(self):
        a = self.arr
        a["voxel"][:] = voxel_keys(a["x"], a["y"], a["z"], self.g)


def ref_secretion(n, ns, nx, ny, nz, dx, dy, dz, dt, voxel, volume_f32, sec, upt, sat, net, density_f32):
    """Run the reference's own processSecretion (FP32 density and volume)."""
    r = ref_lib()
    if r is None:
        return None
    r.ref_secretion(ctypes.c_uint32(n), ctypes.c_uint32(ns), ctypes.c_uint32(nx), ctypes.c_uint32(ny),
                    ctypes.c_uint32(nz), ctypes.c_float(dx), ctypes.c_float(dy), ctypes.c_float(dz),
                    ctypes.c_double(dt), P(voxel, u32p), P(volume_f32, f32p), P(sec), P(upt),
                    P(sat), P(net), P(density_f32, f32p))
    return density_f32


def uniform(seed, step, cell, draw):
    """or_uniform: the counter-based uniform stream of the motility step."""
    f = lib().or_uniform
    f.restype = ctypes.c_double
    return f(ctypes.c_uint64(seed), ctypes.c_uint64(step), ctypes.c_uint32(cell), ctypes.c_uint32(draw))


def motility(x, y, z, vx, vy, vz, alive, speed, g, rho, dt, seed, step, persistence, elapsed, bias,
             mvx, mvy, mvz, motile, restrict_2d, chem_index, chem_dir, sensitivity=None):
    """or_motility (src/motility.cpp:179-293); updates vx, vy, vz, elapsed and
    the motility vector in place.  g: OrGrid; rho: substrate-major density or
    None."""
    lib().or_motility(P(x), P(y), P(z), P(vx), P(vy), P(vz), P(alive, u32p), P(speed), ctypes.c_uint32(len(x)),
                      ctypes.byref(g), None if rho is None else P(rho), ctypes.c_double(dt),
                      ctypes.c_uint64(seed), ctypes.c_uint64(step), P(persistence), P(elapsed), P(bias),
                      P(mvx), P(mvy), P(mvz), P(motile, u32p), P(restrict_2d, u32p), P(chem_index, i32p),
                      P(chem_dir, i32p), None if sensitivity is None else P(sensitivity))


class OrCells(ctypes.Structure):
    _fields_ = [("f64", f64p * 23), ("u32", u32p * 5), ("rates", f64p * 4), ("vol", f64p * 15),
                ("persistence", f64p), ("elapsed", f64p), ("bias", f64p), ("vec", f64p * 3),
                ("motile", u32p), ("restrict_2d", u32p), ("chem_index", i32p), ("chem_dir", i32p),
                ("sensitivity", f64p), ("ns", ctypes.c_uint32), ("_pad", ctypes.c_uint32)]


class PhenoCells:
    """Host copy of the library's per-cell buffers for the phenotype oracle:
    f64 (23, M), u32 (5, M), rates (4, M*ns), vol (15, M) and, when
    with_motility, the motility fields (numpy arrays, modified in place)."""

    def __init__(self, f64, u32, rates, vol, ns, mot=None):
        self.f64, self.u32, self.rates, self.vol, self.ns, self.mot = f64, u32, rates, vol, ns, mot
        for a in (f64, rates, vol):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        assert u32.dtype == np.uint32 and u32.flags["C_CONTIGUOUS"]

    def struct(self):
        c = OrCells()
        for f in range(23):
            c.f64[f] = self.f64[f].ctypes.data_as(f64p)
        for f in range(5):
            c.u32[f] = self.u32[f].ctypes.data_as(u32p)
        for a in range(4):
            c.rates[a] = self.rates[a].ctypes.data_as(f64p)
        for f in range(15):
            c.vol[f] = self.vol[f].ctypes.data_as(f64p)
        if self.mot is not None:
            m = self.mot
            c.persistence, c.elapsed, c.bias = P(m["persistence"]), P(m["elapsed"]), P(m["bias"])
            for q in range(3):
                c.vec[q] = P(m["vec"][q])
            c.motile, c.restrict_2d = P(m["motile"], u32p), P(m["r2d"], u32p)
            c.chem_index, c.chem_dir = P(m["chem_index"], i32p), P(m["chem_dir"], i32p)
            c.sensitivity = P(m["sens"])
        c.ns = self.ns
        return c


def volume_update(cells, mask, n, dt, use_ratio):
    """or_volume_update (src/cell_phenotype.cpp:52-113)."""
    c = cells.struct()
    lib().or_volume_update(ctypes.byref(c), None if mask is None else P(mask, u32p), ctypes.c_uint32(n),
                           ctypes.c_double(dt), ctypes.c_int(use_ratio))


def divide(cells, parents, n, max_cells, seed, counter):
    """or_divide (src/cell_phenotype.cpp:571-720); returns the new n."""
    c = cells.struct()
    parents = np.ascontiguousarray(parents, np.uint32)
    f = lib().or_divide
    f.restype = ctypes.c_uint32
    return f(ctypes.byref(c), P(parents, u32p), ctypes.c_uint32(len(parents)), ctypes.c_uint32(n),
             ctypes.c_uint32(max_cells), ctypes.c_uint64(seed), ctypes.c_uint64(counter))


def remove(cells, indices, n):
    """or_remove (src/cell_phenotype.cpp:833-844, src/cell_data.cpp:408-542); returns n - count."""
    c = cells.struct()
    indices = np.ascontiguousarray(indices, np.uint32)
    f = lib().or_remove
    f.restype = ctypes.c_uint32
    return f(ctypes.byref(c), P(indices, u32p), ctypes.c_uint32(len(indices)), ctypes.c_uint32(n))


def contact_graph(x, y, z, radius, volume, rel, alive):
    """or_contact_graph (src/output_writer.cpp:457-490): (offsets, idx)."""
    n = len(x)
    off = np.zeros(n + 1, np.uint32)
    f = lib().or_contact_graph
    f.restype = ctypes.c_uint64
    tot = f(P(x), P(y), P(z), P(radius), P(volume), P(rel), P(alive, u32p), ctypes.c_uint32(n), P(off, u32p), None)
    idx = np.zeros(max(1, tot), np.uint32)
    f(P(x), P(y), P(z), P(radius), P(volume), P(rel), P(alive, u32p), ctypes.c_uint32(n), P(off, u32p),
      P(idx, u32p))
    return off, idx[:tot]


MAX_ATTACHMENTS = 12


def form_attachments(x, y, z, radius, rel, alive, rate, dt, seed, count=None, targets=None):
    """or_form_attachments (src/cell_interactions.cpp:411-478): (count,
    targets, formed), updated from the given state (default: none)."""
    n = len(x)
    count = np.zeros(n, np.uint32) if count is None else np.array(count, np.uint32)
    targets = (np.zeros(max(1, n) * MAX_ATTACHMENTS, np.uint32) if targets is None
               else np.array(targets, np.uint32).reshape(-1))
    f = lib().or_form_attachments
    f.restype = ctypes.c_uint32
    formed = f(P(x), P(y), P(z), P(radius), P(rel), P(alive, u32p), P(np.asarray(rate, np.float64)),
               ctypes.c_uint32(n), ctypes.c_double(dt), ctypes.c_uint64(seed), P(count, u32p), P(targets, u32p))
    return count, targets[:n * MAX_ATTACHMENTS].reshape(n, MAX_ATTACHMENTS), formed


def interaction_candidates(x, y, radius, volume, keys, off, srt, m):
    """or_interaction_candidates (src/cell_phenotype.cpp:893-967): (offsets, idx)."""
    n = len(x)
    offs = np.zeros(n + 1, np.uint32)
    f = lib().or_interaction_candidates
    f.restype = ctypes.c_uint64
    args = (P(x), P(y), P(radius), P(volume), ctypes.c_uint32(n), P(keys, u32p), P(off, u32p), P(srt, u32p),
            ctypes.byref(m), P(offs, u32p))
    tot = f(*args, None)
    idx = np.zeros(max(1, tot), np.uint32)
    f(*args, P(idx, u32p))
    return offs, idx[:tot]
