"""The packed ensemble and the serving loop against the FP64 oracle.

Every member of an ensemble, and every served job, must equal one run of the
reference's step loop (src/main.mm:275-415) on its own inputs: here the
FP64 oracle's run over the same substeps.

* The full 64-member ensemble of the batched workload (members
  ensemble_member(100..163), 2.05M agents), with the period graph on, over 61
  periods: 610 substeps, 61 mechanics steps and, at the end of period 60,
  every member's phenotype-clock refresh (voxel keys + secretion re-plan).
  The member with the smallest grid, the one with the largest grid and the
  one with the most agents are compared with the oracle: FP64 at 1e-10
  relative with voxel and mechanics keys exact; FP32-mixed at 1e-4 relative
  (max |mixed - oracle| / max |oracle| per substrate) and 1e-3 um.
* Served jobs (continuous batching, FP64 and mixed) against the oracle at the
  same bars.
"""
import os

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H
from physicell_amd import scenarios as S
from physicell_amd.serving import Server, estimate_bytes, poisson_jobs

pytestmark = pytest.mark.gpu
PERIODS = 61
POS = ((H.F_POS_X, "x"), (H.F_POS_Y, "y"), (H.F_POS_Z, "z"))


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    floor = max(float(np.max(np.abs(b))), 1e-300) * 1e-14
    return float(np.max(np.where(d == 0, 0.0, d / np.maximum(np.abs(b), floor)))) if a.size else 0.0


def oracle_after(spec, cells, periods):
    """The oracle's step loop over periods x (10 substeps + 1 mechanics step)."""
    ref = O.OracleSim(spec, cells)
    counts = ref.run((periods * 10 - 1) * spec.dt_diff, nthreads=threads())
    assert counts[0] == periods * 10 and counts[1] == periods, counts
    return ref


def check_fp64(what, spec, rho, pos, keys, ref):
    V = spec.n_voxels
    for s in range(len(spec.substrates)):
        e = rel_err(rho[s * V:(s + 1) * V], ref.arr["rho"][s * V:(s + 1) * V])
        print(f"  {what} substrate {s}: max rel err {e:.3e}", flush=True)
        assert e <= 1e-10, (what, s, e)
    for (f, k), p in zip(POS, pos):
        e = rel_err(p, ref.arr[k])
        print(f"  {what} {k}: max rel err {e:.3e}, bit-identical={np.array_equal(p, ref.arr[k])}", flush=True)
        assert e <= 1e-10, (what, k, e)
    if keys is not None:
        assert np.array_equal(keys[0], ref.arr["voxel"]), f"{what}: voxel keys"
        assert np.array_equal(keys[1], ref.arr["mech_key"]), f"{what}: mechanics keys"


def check_mixed(what, spec, rho, pos, ref):
    V = spec.n_voxels
    for s in range(len(spec.substrates)):
        a, b = rho[s * V:(s + 1) * V], ref.arr["rho"][s * V:(s + 1) * V]
        d = float(np.max(np.abs(a - b)))
        scale = max(float(np.max(np.abs(b))), 1e-300)
        print(f"  {what} substrate {s}: max abs {d:.3e}, rel {d / scale:.3e}", flush=True)
        assert d / scale <= 1e-4, (what, s, d / scale)
    pe = max(float(np.max(np.abs(p - ref.arr[k]))) for (f, k), p in zip(POS, pos))
    print(f"  {what} positions: max abs {pe:.3e} um", flush=True)
    assert pe <= 1e-3, (what, pe)


@pytest.fixture(scope="module")
def members():
    return S.config3()


def picked(members):
    v = [sc.n_voxels for sc, _ in members]
    n = [c.n for _, c in members]
    return sorted({int(np.argmin(v)), int(np.argmax(v)), int(np.argmax(n))})


@pytest.mark.parametrize("mixed", [False, True])
def test_config3_ensemble_members_vs_oracle(members, mixed):
    pick = picked(members)
    print(f"config3 ensemble mixed={mixed}: {len(members)} members, "
          f"{sum(c.n for _, c in members)} agents, {PERIODS} periods (graph); oracle on members {pick}", flush=True)
    ens = H.Ensemble(mixed=mixed)
    ids = [ens.add(sc, c) for sc, c in members]
    ens.run_periods(PERIODS, use_graph=True)
    ens.wait()
    for i in pick:
        sc, c = members[i]
        mid = ids[i]
        _, counts = ens.member_info(mid)
        assert counts[0] == PERIODS * 10 and counts[1] == PERIODS and counts[2] == 1, counts
        rho = ens.density(mid)
        pos = [ens.field(mid, f) for f, _ in POS]
        keys = (ens.field(mid, H.F_VOXEL_IDX), ens.field(mid, H.F_MECH_VOXEL))
        print(f" member {i}: grid {sc.nx}x{sc.ny}x{sc.nz}, {c.n} agents", flush=True)
        ref = oracle_after(sc, c, PERIODS)
        if mixed:
            check_mixed(f"member {i}", sc, rho, pos, ref)
        else:
            check_fp64(f"member {i}", sc, rho, pos, keys, ref)
    ens.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_served_jobs_vs_oracle(mixed):
    jobs = poisson_jobs(6, rate=400.0, seed=3100, steps=(3, 12),
                        member_kw={"edges": (18, 36), "cells": (100, 1500)})
    srv = Server(lambda: H.Ensemble(mixed=mixed), budget_bytes=50e9, max_members=3, mixed=mixed,
                 bytes_of=lambda j: estimate_bytes(j.spec, j.cells.n, mixed, H))
    res, st = srv.run(jobs)
    print(f"serving mixed={mixed}: max live {st.max_live}, admitted while others ran {st.admitted_while_busy}",
          flush=True)
    assert sorted(res) == [j.id for j in jobs]
    assert st.max_live >= 2 and st.admitted_while_busy >= 1
    for j in jobs:
        ref = oracle_after(j.spec, j.cells, j.periods)
        r = res[j.id]
        if mixed:
            check_mixed(f"job {j.id}", j.spec, r.density, list(r.positions), ref)
        else:
            check_fp64(f"job {j.id}", j.spec, r.density, list(r.positions), None, ref)
