"""ctypes binding of libphysicell_hip.so (the C-ABI of include/physicell_hip.h).

Two levels, as in the header:
  Context / Buffer  -> hc_* (hip_context: device, stream, buffers, dispatches)
  Simulation        -> pc_* (Microenvironment + CellMechanics + CellData + loop)

The product path is the HIP library; there is no CPU fallback.  Loading
fails loudly when the library is missing, and every call checks its return
code and raises HipError with the library's message.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphysicell_hip.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "physicell_hip.h")

c_u32, c_i32, c_f64, c_sz, c_vp = ctypes.c_uint32, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p

# field numbers (include/physicell_hip.h; the reference's CELL_FIELD_*)
F_POS_X, F_POS_Y, F_POS_Z = 0, 1, 2
F_VEL_X, F_VEL_Y, F_VEL_Z = 3, 4, 5
F_PREV_VX, F_PREV_VY, F_PREV_VZ = 6, 7, 8
F_RADIUS, F_TOTAL_VOL = 9, 11
F_REPULSION, F_ADHESION, F_REL_MAX_ADH = 13, 14, 15
F_PRESSURE = 22
F_CELL_TYPE, F_CUR_PHASE, F_IS_ALIVE, F_VOXEL_IDX, F_MECH_VOXEL = 23, 24, 25, 26, 27
F_NUCLEAR_RAD, F_NUCLEAR_VOL, F_MOT_SPEED, F_MOT_BIAS, F_ONCOPROTEIN = 10, 12, 16, 17, 21
# motility buffer fields (include/physicell_hip.h, HC_M_*)
M_PERSISTENCE, M_ELAPSED, M_BIAS, M_VEC_X, M_VEC_Y, M_VEC_Z = 0, 1, 2, 3, 4, 5
M_MOTILE, M_RESTRICT_2D, M_CHEM_INDEX, M_CHEM_DIR, M_SENSITIVITY = 6, 7, 8, 9, 10
NUM_F64_FIELDS = 23

K_SWEEP_X, K_SWEEP_Y, K_SWEEP_Z, K_SWEEP_SMALL, K_DIRICHLET, K_VOXEL_KEYS = 0, 1, 2, 3, 4, 5
K_SEC_PLAN, K_SEC_APPLY, K_HASH, K_FORCES, K_INTEGRATE = 6, 7, 8, 9, 10
K_MOTILITY, K_VOLUME, K_DIVIDE, K_REMOVE = 11, 12, 13, 14
KERNEL_NAMES = ["sweep_x", "sweep_y", "sweep_z", "sweep_small", "dirichlet", "voxel_keys",
                "secretion_plan", "secretion_apply", "hash", "forces", "integrate", "motility",
                "volume", "divide", "remove"]
# volume state fields (include/physicell_hip.h, HC_V_*)
(V_FLUID, V_SOLID_NUC, V_SOLID_CYTO, V_FLUID_FRAC, V_CALC_FRAC, V_TARGET_SOLID_CYTO, V_TARGET_SOLID_NUC,
 V_TARGET_FLUID_FRAC, V_TARGET_CN_RATIO, V_FLUID_RATE, V_NUC_RATE, V_CYTO_RATE, V_CALC_RATE, V_RUPTURE_VOL,
 V_REL_RUPTURE_VOL) = range(15)
NUM_V_FIELDS = 15
QUERY_CONTACT, QUERY_INTERACT, QUERY_ATTACH = 0, 1, 2
MAX_ATTACHMENTS = 12


class HipError(RuntimeError):
    pass


class Grid(ctypes.Structure):
    _fields_ = [("nx", c_u32), ("ny", c_u32), ("nz", c_u32), ("n_substrates", c_u32),
                ("field_fp32", c_u32), ("_pad", c_u32),
                ("dx", c_f64), ("dy", c_f64), ("dz", c_f64), ("dt", c_f64),
                ("x_min", c_f64), ("y_min", c_f64), ("z_min", c_f64)]


class Substrate(ctypes.Structure):
    _fields_ = [("diffusion_coefficient", c_f64), ("decay_rate", c_f64), ("supply_rate", c_f64),
                ("target_value", c_f64), ("dirichlet_enabled", c_u32), ("_pad", c_u32)]


class Mech(ctypes.Structure):
    _fields_ = [("dt", c_f64), ("cell_cell_repulsion", c_f64), ("cell_cell_adhesion", c_f64),
                ("max_interaction_distance", c_f64), ("mechanics_voxel_size", c_f64),
                ("mech_grid_nx", c_u32), ("mech_grid_ny", c_u32), ("mech_grid_nz", c_u32),
                ("forces_fp32", c_u32)]


class SubstrateCfg(ctypes.Structure):
    _fields_ = [("diffusion_coeff", c_f64), ("decay_rate", c_f64), ("initial_value", c_f64),
                ("dirichlet_value", c_f64), ("dirichlet_enabled", c_u32), ("_pad", c_u32)]


class Config(ctypes.Structure):
    _fields_ = [("x_min", c_f64), ("x_max", c_f64), ("y_min", c_f64), ("y_max", c_f64),
                ("z_min", c_f64), ("z_max", c_f64), ("dx", c_f64), ("dy", c_f64), ("dz", c_f64),
                ("use_2D", c_u32), ("n_substrates", c_u32), ("dt_diffusion", c_f64),
                ("dt_mechanics", c_f64), ("dt_phenotype", c_f64), ("max_time", c_f64),
                ("max_cells", c_u32), ("mixed_precision", c_u32), ("substrates", ctypes.POINTER(SubstrateCfg))]


_lib = None


class Snapshot(ctypes.Structure):
    _fields_ = [("n_cells", c_u32), ("n_substrates", c_u32)] + \
               [(f, c_vp) for f in ("x", "y", "z", "total_volume", "nuclear_volume", "radius", "nuclear_radius",
                                    "vx", "vy", "vz", "pressure", "adhesion", "repulsion", "rel_max_adhesion",
                                    "cell_type", "current_phase", "is_alive", "secretion", "uptake", "saturation",
                                    "net_export")] + \
               [("nx", c_u32), ("ny", c_u32), ("nz", c_u32), ("_pad", c_u32)] + \
               [(f, c_f64) for f in ("dx", "dy", "dz", "x_min", "y_min", "z_min")] + [("density", c_vp)]


def declared_functions(header=HEADER):
    """Every function name declared in include/physicell_hip.h."""
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(hc_[a-z0-9_]+|pc_[a-z0-9_]+|pe_[a-z0-9_]+)\s*\(",
                       txt, flags=re.M)
    return sorted(set(names))


def load(path=LIB_PATH):
    """Load the library (no GPU work happens at load time)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HipError(f"HIP library missing: {path} -- run `python physicell-mac-metal_amd/build.py`")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    sig = {
        "hc_create": (c_i32, [P(c_vp), c_i32]), "hc_destroy": (None, [c_vp]),
        "hc_load_kernels": (c_i32, [c_vp, ctypes.c_char_p]),
        "hc_alloc": (c_i32, [c_vp, c_sz, P(c_vp)]), "hc_free": (None, [c_vp]),
        "hc_host_mirror": (c_vp, [c_vp]), "hc_device_ptr": (c_vp, [c_vp]), "hc_buf_size": (c_sz, [c_vp]),
        "hc_upload": (c_i32, [c_vp, c_vp, c_sz, c_sz]), "hc_download": (c_i32, [c_vp, c_vp, c_sz, c_sz]),
        "hc_wait": (c_i32, [c_vp]), "hc_last_error": (ctypes.c_char_p, [c_vp]),
        "hc_cells_bytes": (c_sz, [c_u32]), "hc_cells_field_offset": (c_sz, [c_u32, c_u32]),
        "hc_rates_bytes": (c_sz, [c_u32, c_u32]),
        "hc_thomas_table_bytes": (c_sz, [P(Grid)]),
        "hc_build_thomas_tables": (c_i32, [c_vp, P(Grid), P(Substrate), c_vp]),
        "hc_diffusion_sweep": (c_i32, [c_vp, c_i32, c_vp, P(Grid), P(Substrate), c_vp, c_i32]),
        "hc_dirichlet": (c_i32, [c_vp, c_vp, P(Grid), P(Substrate)]),
        "hc_diffusion_substep": (c_i32, [c_vp, c_vp, P(Grid), P(Substrate), c_vp]),
        "hc_voxel_keys": (c_i32, [c_vp, c_vp, P(Grid), c_u32, c_u32]),
        "hc_secretion_plan_bytes": (c_sz, [P(Grid), c_u32]),
        "hc_secretion_plan": (c_i32, [c_vp, c_vp, c_vp, P(Grid), c_f64, c_u32, c_u32, c_vp]),
        "hc_secretion_apply": (c_i32, [c_vp, c_vp, c_vp, P(Grid)]),
        "hc_secretion": (c_i32, [c_vp, c_vp, c_vp, c_vp, P(Grid), c_f64, c_u32, c_u32, c_vp]),
        "hc_hash_bytes": (c_sz, [c_u32, c_u32]),
        "hc_build_hash": (c_i32, [c_vp, c_vp, c_vp, P(Mech), P(Grid), c_u32, c_u32, c_u32]),
        "hc_forces": (c_i32, [c_vp, c_vp, c_vp, P(Mech), c_u32, c_u32, c_u32]),
        "hc_forces_only_pipeline": (c_i32, [c_vp, c_vp, c_vp, P(Mech), P(Grid), c_u32, c_u32, c_u32]),
        "hc_integrate": (c_i32, [c_vp, c_vp, P(Grid), c_f64, c_u32, c_u32]),
        "hc_mechanics_pipeline": (c_i32, [c_vp, c_vp, c_vp, P(Mech), P(Grid), c_f64, c_u32, c_u32, c_u32]),
        "hc_hash_download": (c_i32, [c_vp, c_vp, c_u32, c_u32, c_u32, c_vp, c_vp]),
        "hc_hash_max_bin": (c_i32, [c_vp, c_vp, c_u32, P(c_u32)]),
        "hc_neighbor_lists": (c_i32, [c_vp, c_vp, c_vp, P(Mech), c_u32, c_u32, c_u32, c_u32, c_vp, c_vp]),
        "hc_set_sweep_mode": (c_i32, [c_vp, c_i32]), "hc_set_substep_order": (c_i32, [c_vp, c_i32]),
        "hc_substrate_totals": (c_i32, [c_vp, c_vp, P(Grid), c_vp]),
        "hc_substrate_totals_noalloc": (c_i32, [c_vp, c_vp, P(Grid), c_vp]),
        "hc_timing_enable": (c_i32, [c_vp, c_u32]),
        "hc_timing_get": (c_i32, [c_vp, c_i32, P(c_f64), P(ctypes.c_int64)]),
        "hc_timing_reset": (c_i32, [c_vp]),
        "pc_sim_create": (c_i32, [P(c_vp), P(Config), c_i32]), "pc_sim_destroy": (None, [c_vp]),
        "pc_last_error": (ctypes.c_char_p, [c_vp]),
        "pc_sim_set_cells": (c_i32, [c_vp, c_u32] + [c_vp] * 13),
        "pc_sim_diffusion_substep": (c_i32, [c_vp]), "pc_sim_mechanics_step": (c_i32, [c_vp]),
        "pc_sim_phenotype_step": (c_i32, [c_vp]),
        "pc_sim_run": (c_i32, [c_vp, c_f64, c_vp]), "pc_sim_run_periods": (c_i32, [c_vp, c_i32, c_i32]),
        "pc_sim_wait": (c_i32, [c_vp]), "pc_sim_get_density": (c_i32, [c_vp, c_vp]),
        "pc_sim_set_overlap": (c_i32, [c_vp, c_i32]), "pc_sim_set_mechanics_mirror": (c_i32, [c_vp, c_i32]), "pe_set_overlap": (c_i32, [c_vp, c_i32]),
        "pc_sim_set_density": (c_i32, [c_vp, c_vp]), "pc_sim_get_field": (c_i32, [c_vp, c_u32, c_vp]),
        "pc_sim_get_hash": (c_i32, [c_vp, c_vp, c_vp]),
        "pc_sim_substrate_totals": (c_i32, [c_vp, c_vp]),
        "pc_sim_info": (c_i32, [c_vp, P(Grid), P(Mech), P(c_u32)]),
        "pc_sim_context": (c_vp, [c_vp]),
        "hc_motility_bytes": (c_sz, [c_u32, c_u32]), "hc_motility_field_offset": (c_sz, [c_u32, c_u32, c_u32]),
        "hc_volume_bytes": (c_sz, [c_u32]), "hc_volume_field_offset": (c_sz, [c_u32, c_u32]),
        "hc_volume_update": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_f64, c_i32]),
        "hc_divide": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_u32, c_vp]),
        "hc_remove": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_u32, c_vp]),
        "hc_remove_plan": (c_i32, [c_vp, c_u32, c_u32, c_vp, c_vp, c_vp]),
        "hc_motility": (c_i32, [c_vp, c_vp, c_vp, c_vp, P(Grid), c_f64, c_u32, c_u32, c_i32]),
        "pc_sim_set_motility": (c_i32, [c_vp, ctypes.c_uint64] + [c_vp] * 8),
        "pc_sim_set_volume": (c_i32, [c_vp, ctypes.c_uint64, c_vp]),
        "pc_sim_neighbor_query": (c_i32, [c_vp, c_i32, c_vp, c_vp, ctypes.c_uint64, c_vp]),
        "pc_sim_form_attachments": (c_i32, [c_vp, ctypes.c_double, c_vp, c_vp, c_vp, ctypes.c_uint64, c_vp]),
        "pe_neighbor_query": (c_i32, [c_vp, c_u32, c_i32, c_vp, c_vp, ctypes.c_uint64, c_vp]),
        "pe_form_attachments": (c_i32, [c_vp, c_u32, ctypes.c_double, c_vp, c_vp, c_vp, ctypes.c_uint64, c_vp]),
        "hc_neighbor_query": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_u32, c_i32, c_vp, c_vp, c_vp]),
        "pc_write_neighbor_graph": (c_i32, [ctypes.c_char_p, c_u32, c_vp, c_vp]),
        "pc_sim_update_phenotype": (c_i32, [c_vp, c_f64, c_i32, c_vp, c_vp, c_u32, c_vp, c_u32, c_vp]),
        "pc_write_matlab4": (c_i32, [ctypes.c_char_p, ctypes.c_char_p, c_u32, c_u32, c_vp]),
        "pc_write_snapshot": (c_i32, [ctypes.c_char_p, c_i32, P(Snapshot)]),
        "pc_sim_write_snapshot": (c_i32, [c_vp, ctypes.c_char_p, c_i32]),
        "pe_create": (c_i32, [P(c_vp), c_i32, c_u32]), "pe_destroy": (None, [c_vp]),
        "pe_last_error": (ctypes.c_char_p, [c_vp]),
        "pe_add": (c_i32, [c_vp, P(Config), c_u32] + [c_vp] * 13 + [P(c_u32)]),
        "pe_remove": (c_i32, [c_vp, c_u32]),
        "pe_set_motility": (c_i32, [c_vp, c_u32, ctypes.c_uint64] + [c_vp] * 8),
        "pe_set_volume": (c_i32, [c_vp, c_u32, ctypes.c_uint64, c_vp, c_u32]),
        "pe_update_phenotype": (c_i32, [c_vp, c_u32, c_f64, c_i32, c_vp, c_vp, c_u32, c_vp, c_u32, c_vp]), "pe_run_periods": (c_i32, [c_vp, c_i32, c_i32]),
        "pe_wait": (c_i32, [c_vp]), "pe_get_density": (c_i32, [c_vp, c_u32, c_vp]),
        "pe_get_field": (c_i32, [c_vp, c_u32, c_u32, c_vp]),
        "pe_member_info": (c_i32, [c_vp, c_u32, P(ctypes.c_uint64), c_vp]),
        "pe_info": (c_i32, [c_vp, P(c_u32), P(ctypes.c_uint64), P(ctypes.c_uint64), P(ctypes.c_uint64)]),
        "pe_set_grid_blocks": (c_i32, [c_vp, c_i32, c_u32]),
        "pe_autotune": (c_i32, [c_vp, c_i32, c_vp, c_vp]),
        "pe_context": (c_vp, [c_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays passed to the C-ABI must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def grid_of(spec, fp32=False):
    return Grid(spec.nx, spec.ny, spec.nz, len(spec.substrates), 1 if fp32 else 0, 0, spec.dx, spec.dy, spec.dz,
                spec.dt_diff, spec.x_min, spec.y_min, spec.z_min)


def mech_grid_of(spec):
    v = spec.mech_voxel
    return Grid(spec.mnx, spec.mny, spec.mnz, 0, 0, 0, v, v, v, 0.0, spec.x_min, spec.y_min, spec.z_min)


def mech_of(spec, fp32=False):
    v = spec.mech_voxel
    return Mech(spec.dt_mech, 10.0, 0.4, v, v, spec.mnx, spec.mny, spec.mnz, 1 if fp32 else 0)


def substrates_of(spec):
    arr = (Substrate * max(1, len(spec.substrates)))()
    for i, s in enumerate(spec.substrates):
        arr[i] = Substrate(s.D, s.decay, 0.0, s.dirichlet_value, 1 if s.dirichlet else 0, 0)
    return arr


class Buffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = c_vp()
        ctx._check(load().hc_alloc(ctx.h, self.nbytes, ctypes.byref(h)), "hc_alloc")
        self.h = h

    def _check_range(self, offset, nbytes, what):
        """The pinned host mirror is exactly self.nbytes long: a copy into or out of
        it must be range-checked here, before any memmove touches it."""
        offset, nbytes = int(offset), int(nbytes)
        if offset < 0 or nbytes < 0 or offset + nbytes > self.nbytes:
            raise HipError(f"{what}: range [{offset}, {offset + nbytes}) outside the {self.nbytes}-byte buffer")

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        self._check_range(offset, arr.nbytes, "upload")
        lib = load()
        host = lib.hc_host_mirror(self.h)
        if not host:
            raise HipError("no host mirror")
        ctypes.memmove(host + offset, arr.ctypes.data, arr.nbytes)
        self.ctx._check(lib.hc_upload(self.ctx.h, self.h, offset, arr.nbytes), "hc_upload")
        self.ctx.wait()

    def download(self, dtype, count, offset=0):
        nb = np.dtype(dtype).itemsize * int(count)
        self._check_range(offset, nb, "download")
        lib = load()
        self.ctx._check(lib.hc_download(self.ctx.h, self.h, offset, nb), "hc_download")
        self.ctx.wait()
        host = lib.hc_host_mirror(self.h)
        out = np.empty(count, dtype)
        ctypes.memmove(out.ctypes.data, host + offset, nb)
        return out

    def free(self):
        if self.h:
            load().hc_free(self.h)
            self.h = None


class Context:
    """hip_context on one device (hc_create + hc_load_kernels)."""

    def __init__(self, device=0):
        lib = load()
        h = c_vp()
        rc = lib.hc_create(ctypes.byref(h), device)
        if rc:
            raise HipError(f"hc_create failed ({rc}): no usable HIP device")
        self.h = h
        self._check(lib.hc_load_kernels(self.h, None), "hc_load_kernels")

    def _check(self, rc, what):
        if rc:
            raise HipError(f"{what} failed ({rc}): {load().hc_last_error(self.h).decode()}")

    def alloc(self, nbytes):
        return Buffer(self, nbytes)

    def wait(self):
        self._check(load().hc_wait(self.h), "hc_wait")

    def close(self):
        if self.h:
            load().hc_destroy(self.h)
            self.h = None

    # timing
    def timing(self, mask):
        self._check(load().hc_timing_enable(self.h, mask), "hc_timing_enable")

    def timing_get(self, k):
        ms, n = c_f64(), ctypes.c_int64()
        self._check(load().hc_timing_get(self.h, k, ctypes.byref(ms), ctypes.byref(n)), "hc_timing_get")
        return ms.value, n.value

    def timing_reset(self):
        self._check(load().hc_timing_reset(self.h), "hc_timing_reset")


class CellBuffer:
    """Device SoA (hc_cells_bytes) plus rate buffer, filled from a Cells population."""

    def __init__(self, ctx, cells, ns, max_cells=None):
        lib = load()
        self.ctx = ctx
        self.n = cells.n
        self.max_cells = max_cells or max(1, cells.n)
        self.ns = ns
        self.soa = ctx.alloc(lib.hc_cells_bytes(self.max_cells))
        self.rates = ctx.alloc(max(8, lib.hc_rates_bytes(self.max_cells, ns)))
        n, M = self.n, self.max_cells
        f64 = np.zeros((NUM_F64_FIELDS, M))
        f64[F_POS_X, :n], f64[F_POS_Y, :n], f64[F_POS_Z, :n] = cells.x, cells.y, cells.z
        f64[F_RADIUS, :n] = cells.radius
        f64[F_TOTAL_VOL, :n] = cells.volume
        f64[F_REPULSION, :n], f64[F_ADHESION, :n], f64[F_REL_MAX_ADH, :n] = cells.rep, cells.adh, cells.rel_adh
        u32 = np.zeros((5, M), np.uint32)
        u32[F_IS_ALIVE - NUM_F64_FIELDS, :n] = cells.alive
        self.soa.upload(f64.ravel())
        self.soa.upload(u32.ravel(), offset=lib.hc_cells_field_offset(23, M))
        if ns:
            r = np.zeros((4, M * ns))
            for k, a in enumerate((cells.sec, cells.upt, cells.sat, cells.net)):
                r[k, :n * ns] = np.asarray(a, np.float64).reshape(-1)
            self.rates.upload(r.ravel())

    def field(self, f, count=None):
        lib = load()
        count = self.n if count is None else count
        dt = np.float64 if f < NUM_F64_FIELDS else np.uint32
        return self.soa.download(dt, count, lib.hc_cells_field_offset(f, self.max_cells))

    def set_field(self, f, values):
        lib = load()
        dt = np.float64 if f < NUM_F64_FIELDS else np.uint32
        self.soa.upload(np.asarray(values, dt), lib.hc_cells_field_offset(f, self.max_cells))


def config_of(spec, max_cells, mixed=False):
    """pc_config of a scenario; returns (config, substrate array) -- keep both alive."""
    ns = len(spec.substrates)
    subs = (SubstrateCfg * max(1, ns))()
    for i, s in enumerate(spec.substrates):
        subs[i] = SubstrateCfg(s.D, s.decay, s.initial, s.dirichlet_value, 1 if s.dirichlet else 0, 0)
    cfg = Config(spec.x_min, spec.x_max, spec.y_min, spec.y_max, spec.z_min, spec.z_max,
                 spec.dx, spec.dy, spec.dz, 1 if spec.use_2d else 0, ns, spec.dt_diff,
                 spec.dt_mech, spec.dt_pheno, spec.max_time, max_cells, 1 if mixed else 0, subs)
    return cfg, subs


def cell_arrays(c):
    """The 13 per-cell arrays of pc_sim_set_cells / pe_add, contiguous."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (c.x, c.y, c.z, c.radius, c.volume, c.rep, c.adh, c.rel_adh)]
    rates = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (c.sec, c.upt, c.sat, c.net)]
    return arrs + rates + [np.ascontiguousarray(c.alive, np.uint32)]


class Simulation:
    """pc_sim: the C++ step loop on one device."""

    def __init__(self, spec, cells, device=0, max_cells=None, mixed=False):
        lib = load()
        self.spec = spec
        ns = len(spec.substrates)
        cfg, self._subs = config_of(spec, max_cells or max(1, cells.n), mixed)
        h = c_vp()
        rc = lib.pc_sim_create(ctypes.byref(h), ctypes.byref(cfg), device)
        if rc:
            raise HipError(f"pc_sim_create failed ({rc})")
        self.h = h
        self.n = cells.n
        self.ns = ns
        self.set_cells(cells)

    def _check(self, rc, what):
        if rc:
            raise HipError(f"{what} failed ({rc}): {load().pc_last_error(self.h).decode()}")

    def set_cells(self, c):
        n = c.n
        self.n = n
        arrs = [np.ascontiguousarray(a, np.float64) for a in (c.x, c.y, c.z, c.radius, c.volume, c.rep, c.adh, c.rel_adh)]
        rates = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (c.sec, c.upt, c.sat, c.net)]
        alive = np.ascontiguousarray(c.alive, np.uint32)
        self._keep = arrs + rates + [alive]
        self._check(load().pc_sim_set_cells(self.h, n, *[_ptr(a) for a in arrs], *[_ptr(a) for a in rates],
                                            _ptr(alive)), "pc_sim_set_cells")

    def substep(self):
        self._check(load().pc_sim_diffusion_substep(self.h), "pc_sim_diffusion_substep")

    def mech_step(self):
        self._check(load().pc_sim_mechanics_step(self.h), "pc_sim_mechanics_step")

    def pheno_step(self):
        self._check(load().pc_sim_phenotype_step(self.h), "pc_sim_phenotype_step")

    def run(self, max_time):
        counts = np.zeros(3, np.int64)
        self._check(load().pc_sim_run(self.h, max_time, _ptr(counts)), "pc_sim_run")
        return tuple(int(v) for v in counts)

    def run_periods(self, periods, use_graph=False):
        self._check(load().pc_sim_run_periods(self.h, periods, 1 if use_graph else 0), "pc_sim_run_periods")

    def wait(self):
        self._check(load().pc_sim_wait(self.h), "pc_sim_wait")

    def set_overlap(self, on):
        self._check(load().pc_sim_set_overlap(self.h, 1 if on else 0), "pc_sim_set_overlap")

    def set_mechanics_mirror(self, on):
        """FP64 mechanics on the hash-ordered mirror (default on; same results)."""
        self._check(load().pc_sim_set_mechanics_mirror(self.h, 1 if on else 0), "pc_sim_set_mechanics_mirror")

    def set_motility(self, seed, persistence=None, bias=None, speed=None, is_motile=None, restrict_2d=None,
                     chem_index=None, chem_dir=None, sensitivity=None):
        conv = [(persistence, np.float64), (bias, np.float64), (speed, np.float64), (is_motile, np.uint32),
                (restrict_2d, np.uint32), (chem_index, np.int32), (chem_dir, np.int32), (sensitivity, np.float64)]
        arrs = [None if a is None else np.ascontiguousarray(np.asarray(a, t).reshape(-1)) for a, t in conv]
        self._keep_mot = arrs
        self._check(load().pc_sim_set_motility(self.h, seed, *[_ptr(a) for a in arrs]), "pc_sim_set_motility")

    def set_volume(self, seed, fields=None):
        """fields: {HC_V_* number: array of n values}; missing ones take
        CellData's fresh values."""
        fields = fields or {}
        arrs = [None if f not in fields else np.ascontiguousarray(fields[f], np.float64) for f in range(NUM_V_FIELDS)]
        table = (c_vp * NUM_V_FIELDS)(*[None if a is None else a.ctypes.data for a in arrs])
        self._check(load().pc_sim_set_volume(self.h, seed, table), "pc_sim_set_volume")

    def update_phenotype(self, dt, mask=None, parents=(), removed=(), use_ratio=True):
        """Volume model (cells with mask != 0), divisions, removals on the
        device; returns the new cell count."""
        m = None if mask is None else np.ascontiguousarray(mask, np.uint32)
        pa = np.ascontiguousarray(parents, np.uint32)
        rm = np.ascontiguousarray(removed, np.uint32)
        out = ctypes.c_uint32()
        self._check(load().pc_sim_update_phenotype(self.h, dt, 1 if use_ratio else 0, _ptr(m), _ptr(pa), len(pa),
                                                   _ptr(rm), len(rm), ctypes.byref(out)), "pc_sim_update_phenotype")
        self.n = out.value
        return self.n

    def neighbor_query(self, kind):
        """pc_sim_neighbor_query: (offsets[n + 1], idx) of QUERY_CONTACT,
        QUERY_INTERACT or QUERY_ATTACH lists over the current positions."""
        off = np.zeros(self.n + 1, np.uint32)
        total = ctypes.c_uint64()
        self._check(load().pc_sim_neighbor_query(self.h, kind, _ptr(off), None, 0, ctypes.byref(total)),
                    "pc_sim_neighbor_query")
        idx = np.zeros(max(1, total.value), np.uint32)
        self._check(load().pc_sim_neighbor_query(self.h, kind, _ptr(off), _ptr(idx), total.value,
                                                 ctypes.byref(total)), "pc_sim_neighbor_query")
        return off, idx[:total.value]

    def form_attachments(self, dt, rate, seed, count=None, targets=None):
        """pc_sim_form_attachments (tryFormAllAttachments): (count[n],
        targets[n, 12], formed), updated from the given lists (default:
        none)."""
        n = self.n
        rate = np.ascontiguousarray(rate, np.float64)
        if len(rate) != n:
            raise ValueError("form_attachments: rate needs one value per cell")
        count = np.zeros(n, np.uint32) if count is None else np.array(count, np.uint32)
        targets = (np.zeros((n, MAX_ATTACHMENTS), np.uint32) if targets is None
                   else np.array(targets, np.uint32).reshape(n, MAX_ATTACHMENTS))
        formed = c_u32()
        self._check(load().pc_sim_form_attachments(self.h, float(dt), _ptr(rate), _ptr(count), _ptr(targets),
                                                   int(seed), ctypes.byref(formed)), "pc_sim_form_attachments")
        return count, targets, formed.value

    def density(self):
        out = np.empty(self.ns * self.spec.n_voxels)
        self._check(load().pc_sim_get_density(self.h, _ptr(out)), "pc_sim_get_density")
        return out

    def set_density(self, rho):
        rho = np.ascontiguousarray(rho, np.float64)
        self._check(load().pc_sim_set_density(self.h, _ptr(rho)), "pc_sim_set_density")

    def field(self, f):
        out = np.empty(self.n, np.float64 if f < NUM_F64_FIELDS else np.uint32)
        self._check(load().pc_sim_get_field(self.h, f, _ptr(out)), "pc_sim_get_field")
        return out

    def hash(self):
        off = np.empty(self.spec.n_mech_voxels + 1, np.uint32)
        cells = np.empty(max(1, self.n), np.uint32)
        self._check(load().pc_sim_get_hash(self.h, _ptr(off), _ptr(cells)), "pc_sim_get_hash")
        return off, cells[:self.n]

    def totals(self):
        out = np.empty(max(1, self.ns))
        self._check(load().pc_sim_substrate_totals(self.h, _ptr(out)), "pc_sim_substrate_totals")
        return out[:self.ns]

    def context(self):
        return load().pc_sim_context(self.h)

    def write_snapshot(self, folder, frame):
        self._check(load().pc_sim_write_snapshot(self.h, str(folder).encode(), frame), "pc_sim_write_snapshot")

    def timing(self, mask):
        rc = load().hc_timing_enable(self.context(), mask)
        if rc:
            raise HipError("hc_timing_enable failed")

    def timing_get(self, k):
        ms, n = c_f64(), ctypes.c_int64()
        rc = load().hc_timing_get(self.context(), k, ctypes.byref(ms), ctypes.byref(n))
        if rc:
            raise HipError("hc_timing_get failed")
        return ms.value, n.value

    def timing_reset(self):
        if load().hc_timing_reset(self.context()):
            raise HipError("hc_timing_reset failed")

    def close(self):
        if self.h:
            load().pc_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """pe_ensemble: members stepped together in packed launches."""

    def __init__(self, device=0, mixed=False):
        lib = load()
        h = c_vp()
        rc = lib.pe_create(ctypes.byref(h), device, 1 if mixed else 0)
        if rc:
            raise HipError(f"pe_create failed ({rc})")
        self.h = h
        self.mixed = mixed
        self.specs = {}
        self.ncells = {}

    def _check(self, rc, what):
        if rc:
            raise HipError(f"{what} failed ({rc}): {load().pe_last_error(self.h).decode()}")

    def add(self, spec, cells):
        cfg, subs = config_of(spec, max(1, cells.n), self.mixed)
        a = cell_arrays(cells)
        mid = c_u32()
        self._check(load().pe_add(self.h, ctypes.byref(cfg), cells.n, *[_ptr(x) for x in a], ctypes.byref(mid)),
                    "pe_add")
        self.specs[mid.value] = spec
        self.ncells[mid.value] = cells.n
        return mid.value

    def remove(self, mid):
        self._check(load().pe_remove(self.h, mid), "pe_remove")
        self.specs.pop(mid, None)
        self.ncells.pop(mid, None)

    def set_motility(self, mid, seed, persistence=None, bias=None, speed=None, is_motile=None, restrict_2d=None,
                     chem_index=None, chem_dir=None, sensitivity=None):
        """pe_set_motility: Simulation.set_motility for one member."""
        conv = [(persistence, np.float64), (bias, np.float64), (speed, np.float64), (is_motile, np.uint32),
                (restrict_2d, np.uint32), (chem_index, np.int32), (chem_dir, np.int32), (sensitivity, np.float64)]
        arrs = [None if a is None else np.ascontiguousarray(np.asarray(a, t).reshape(-1)) for a, t in conv]
        self._keep_mot = getattr(self, "_keep_mot", []) + [arrs]
        self._check(load().pe_set_motility(self.h, mid, seed, *[_ptr(a) for a in arrs]), "pe_set_motility")

    def set_volume(self, mid, seed, fields=None, max_cells=0):
        """pe_set_volume: Simulation.set_volume for one member, with room for
        max_cells cells."""
        fields = fields or {}
        arrs = [None if f not in fields else np.ascontiguousarray(fields[f], np.float64) for f in range(NUM_V_FIELDS)]
        table = (c_vp * NUM_V_FIELDS)(*[None if a is None else a.ctypes.data for a in arrs])
        self._check(load().pe_set_volume(self.h, mid, seed, table, max_cells), "pe_set_volume")

    def update_phenotype(self, mid, dt, mask=None, parents=(), removed=(), use_ratio=True):
        """pe_update_phenotype: Simulation.update_phenotype for one member;
        returns its new cell count."""
        m = None if mask is None else np.ascontiguousarray(mask, np.uint32)
        pa = np.ascontiguousarray(parents, np.uint32)
        rm = np.ascontiguousarray(removed, np.uint32)
        out = ctypes.c_uint32()
        self._check(load().pe_update_phenotype(self.h, mid, dt, 1 if use_ratio else 0, _ptr(m), _ptr(pa), len(pa),
                                               _ptr(rm), len(rm), ctypes.byref(out)), "pe_update_phenotype")
        self.ncells[mid] = out.value
        return out.value

    def run_periods(self, periods, use_graph=True):
        self._check(load().pe_run_periods(self.h, periods, 1 if use_graph else 0), "pe_run_periods")

    def wait(self):
        self._check(load().pe_wait(self.h), "pe_wait")

    def density(self, mid):
        spec = self.specs[mid]
        out = np.empty(len(spec.substrates) * spec.n_voxels)
        self._check(load().pe_get_density(self.h, mid, _ptr(out)), "pe_get_density")
        return out

    def field(self, mid, f):
        out = np.empty(max(1, self.ncells[mid]), np.float64 if f < NUM_F64_FIELDS else np.uint32)
        self._check(load().pe_get_field(self.h, mid, f, _ptr(out)), "pe_get_field")
        return out[:self.ncells[mid]]

    def neighbor_query(self, mid, kind):
        """pe_neighbor_query: (offsets[n + 1], idx) for one member."""
        off = np.zeros(self.ncells[mid] + 1, np.uint32)
        total = ctypes.c_uint64()
        self._check(load().pe_neighbor_query(self.h, mid, kind, _ptr(off), None, 0, ctypes.byref(total)),
                    "pe_neighbor_query")
        idx = np.zeros(max(1, total.value), np.uint32)
        self._check(load().pe_neighbor_query(self.h, mid, kind, _ptr(off), _ptr(idx), total.value,
                                             ctypes.byref(total)), "pe_neighbor_query")
        return off, idx[:total.value]

    def form_attachments(self, mid, dt, rate, seed, count=None, targets=None):
        """pe_form_attachments: (count[n], targets[n, 12], formed) for one member."""
        n = self.ncells[mid]
        rate = np.ascontiguousarray(rate, np.float64)
        if len(rate) != n:
            raise ValueError("form_attachments: rate needs one value per cell")
        count = np.zeros(n, np.uint32) if count is None else np.array(count, np.uint32)
        targets = (np.zeros((n, MAX_ATTACHMENTS), np.uint32) if targets is None
                   else np.array(targets, np.uint32).reshape(n, MAX_ATTACHMENTS))
        formed = c_u32()
        self._check(load().pe_form_attachments(self.h, mid, float(dt), _ptr(rate), _ptr(count), _ptr(targets),
                                               int(seed), ctypes.byref(formed)), "pe_form_attachments")
        return count, targets, formed.value

    def member_info(self, mid):
        b = ctypes.c_uint64()
        counts = np.zeros(3, np.int64)
        self._check(load().pe_member_info(self.h, mid, ctypes.byref(b), _ptr(counts)), "pe_member_info")
        return b.value, tuple(int(v) for v in counts)

    def info(self):
        n, cells, vox, ws = c_u32(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(load().pe_info(self.h, ctypes.byref(n), ctypes.byref(cells), ctypes.byref(vox), ctypes.byref(ws)),
                    "pe_info")
        return {"members": n.value, "cells": cells.value, "voxel_values": vox.value, "workspace_bytes": ws.value}

    def set_overlap(self, on):
        self._check(load().pe_set_overlap(self.h, 1 if on else 0), "pe_set_overlap")

    def set_grid_blocks(self, axis, blocks):
        self._check(load().pe_set_grid_blocks(self.h, axis, blocks), "pe_set_grid_blocks")

    def autotune(self, reps=5):
        blocks = np.zeros(3, np.uint32)
        us = np.zeros(3, np.float64)
        self._check(load().pe_autotune(self.h, reps, _ptr(blocks), _ptr(us)), "pe_autotune")
        return [int(b) for b in blocks], [float(u) for u in us]

    def context(self):
        return load().pe_context(self.h)

    def timing(self, mask):
        if load().hc_timing_enable(self.context(), mask):
            raise HipError("hc_timing_enable failed")

    def timing_get(self, k):
        ms, n = c_f64(), ctypes.c_int64()
        if load().hc_timing_get(self.context(), k, ctypes.byref(ms), ctypes.byref(n)):
            raise HipError("hc_timing_get failed")
        return ms.value, n.value

    def close(self):
        if self.h:
            load().pe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
