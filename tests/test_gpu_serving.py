"""Serving loop on the device: jobs that arrive while others run are batched
with them, and every job's result equals the result of running it alone, bit
for bit (FP32-mixed form)."""
import numpy as np
import pytest

from physicell_amd import hip
from physicell_amd.serving import Server, estimate_bytes, poisson_jobs

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    jobs = poisson_jobs(8, rate=400.0, seed=3000, steps=(3, 12),
                        member_kw={"edges": (18, 36), "cells": (100, 1500)})
    srv = Server(lambda: hip.Ensemble(mixed=True), budget_bytes=50e9, max_members=3, mixed=True,
                 bytes_of=lambda j: estimate_bytes(j.spec, j.cells.n, True, hip))
    res, st = srv.run(jobs)
    s = st.summary()
    print("serving:", {k: s[k] for k in ("batched_steps", "max_live_jobs", "jobs_admitted_while_others_ran",
                                          "p50_step_ms", "p99_step_ms")}, flush=True)
    assert sorted(res) == [j.id for j in jobs]
    assert st.max_live >= 2 and st.admitted_while_busy >= 1
    for j in jobs:
        e = hip.Ensemble(mixed=True)
        mid = e.add(j.spec, j.cells)
        e.run_periods(j.periods)
        e.wait()
        rho = e.density(mid)
        pos = np.stack([e.field(mid, f) for f in (hip.F_POS_X, hip.F_POS_Y, hip.F_POS_Z)])
        e.close()
        assert np.array_equal(rho, res[j.id].density), f"job {j.id}: density differs"
        assert np.array_equal(pos, res[j.id].positions), f"job {j.id}: positions differ"


def test_budget_bounds_admission():
    """A budget that holds two of the jobs at a time: arrivals wait for memory
    to free up (not for slots), the admitted jobs' estimated bytes never exceed
    the budget, each estimate covers the member's own device bytes, and every
    result still equals the job run alone."""
    jobs = poisson_jobs(6, rate=1000.0, seed=3100, steps=(2, 6),
                        member_kw={"edges": (20, 30), "cells": (200, 800)})
    est = {j.id: estimate_bytes(j.spec, j.cells.n, True, hip) for j in jobs}
    budget = sorted(est.values())[-1] + sorted(est.values())[-2] + 1
    srv = Server(lambda: hip.Ensemble(mixed=True), budget_bytes=budget, max_members=64, mixed=True,
                 bytes_of=lambda j: est[j.id])
    res, st = srv.run(jobs)
    s = st.summary()
    print("budget-bound serving:", {k: s[k] for k in ("batched_steps", "max_live_jobs", "max_device_bytes",
                                                       "p99_queue_wait_s")}, f"budget {budget}", flush=True)
    assert sorted(res) == [j.id for j in jobs]
    assert st.max_bytes <= budget
    assert 2 <= st.max_live < len(jobs), "the budget, not the 64 slots, bounds the live jobs"
    assert max(st.queue_wait_s) > 0.0, "some arrival waited for memory"
    for j in jobs:
        e = hip.Ensemble(mixed=True)
        mid = e.add(j.spec, j.cells)
        own, _ = e.member_info(mid)
        assert own <= est[j.id], f"job {j.id}: member holds {own} bytes, estimate {est[j.id]}"
        e.run_periods(j.periods)
        e.wait()
        rho = e.density(mid)
        pos = np.stack([e.field(mid, f) for f in (hip.F_POS_X, hip.F_POS_Y, hip.F_POS_Z)])
        e.close()
        assert np.array_equal(rho, res[j.id].density), f"job {j.id}: density differs"
        assert np.array_equal(pos, res[j.id].positions), f"job {j.id}: positions differ"
