"""Writes tests/golden/secretion_ref.npz: seeded crowded secretion inputs and
the output of the reference's own processSecretion
(src/cell_phenotype.cpp:526-564), compiled from the reference sources by
oracle/build_ref.sh.  Cells are applied one after another in ascending index
(a single-thread run of src/main.mm:358-362).  Inputs only; no reference
source is stored.

    sh oracle/build_ref.sh && python tests/golden/make_secretion_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

rng = np.random.default_rng(20261203)
nx, ny, nz, ns, n = 5, 4, 3, 3, 400
V = nx * ny * nz
voxel = rng.integers(0, V + 3, n).astype(np.uint32)          # a few out-of-range voxels
volume = rng.uniform(500, 5000, n).astype(np.float32)
sec = rng.uniform(0, 10, n * ns)
upt = rng.uniform(0, 10, n * ns)
sat = rng.uniform(0, 40, n * ns)
net = rng.uniform(-3000, 1000, n * ns)
rho0 = rng.uniform(0, 50, ns * V).astype(np.float32)
dt = 0.01
out = rho0.copy()
if O.ref_secretion(n, ns, nx, ny, nz, 20.0, 20.0, 20.0, dt, voxel, volume, sec, upt, sat, net, out) is None:
    raise SystemExit("reference build missing: run oracle/build_ref.sh")
np.savez_compressed(os.path.join(HERE, "secretion_ref.npz"), ns=ns, dt=dt, voxel=voxel, volume=volume,
                    sec=sec, upt=upt, sat=sat, net=net, rho0=rho0, rho_ref=out)
print("wrote secretion_ref.npz", float(np.abs(out - rho0).max()))
