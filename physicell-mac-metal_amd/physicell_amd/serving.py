"""Request loop over a packed ensemble: continuous batching of simulation jobs.

Jobs arrive over time, are admitted while the device-memory budget and the
member slots allow, and step together: every loop iteration runs one period
(10 diffusion substeps + 1 mechanics step) of every live job as one batched
step (one hipGraph replay of the ensemble).  A job leaves when its own number
of mechanics steps is done; its final state (density and positions) is copied
out then and its device memory freed, which lets waiting jobs in.  Jobs that
arrive while others are on the GPU join at the next step boundary, so their
work overlaps on the device with the jobs already running.

A member's arithmetic does not depend on the other members, so each job's
result equals the result of running it alone (tests check this bit for bit).

    python -m physicell_amd.serving --jobs 200 --rate 4 --budget-gb 200
"""
from __future__ import annotations

import argparse
import json
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import scenarios as S


@dataclass
class Job:
    id: int
    spec: S.Scenario
    cells: S.Cells
    periods: int            # mechanics steps to run
    arrival: float          # seconds after the stream starts


@dataclass
class Result:
    job_id: int
    periods: int
    arrival: float
    admitted: float
    finished: float
    density: Optional[np.ndarray] = None
    positions: Optional[np.ndarray] = None    # (3, n)


@dataclass
class Stats:
    jobs: int = 0
    periods_run: int = 0
    agent_steps: int = 0
    voxel_substeps: int = 0
    makespan_s: float = 0.0
    step_ms: List[float] = field(default_factory=list)
    turnaround_s: List[float] = field(default_factory=list)
    queue_wait_s: List[float] = field(default_factory=list)
    max_live: int = 0
    max_bytes: int = 0
    admitted_while_busy: int = 0
    rebuild_steps: int = 0

    def summary(self):
        pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None
        return {
            "jobs": self.jobs, "batched_steps": self.periods_run, "makespan_s": self.makespan_s,
            "agent_steps_per_s": self.agent_steps / self.makespan_s if self.makespan_s else None,
            "voxel_substeps_per_s": self.voxel_substeps / self.makespan_s if self.makespan_s else None,
            "p50_step_ms": pct(self.step_ms, 50), "p99_step_ms": pct(self.step_ms, 99),
            "p50_turnaround_s": pct(self.turnaround_s, 50), "p99_turnaround_s": pct(self.turnaround_s, 99),
            "p50_queue_wait_s": pct(self.queue_wait_s, 50), "p99_queue_wait_s": pct(self.queue_wait_s, 99),
            "max_live_jobs": self.max_live, "max_device_bytes": self.max_bytes,
            "jobs_admitted_while_others_ran": self.admitted_while_busy,
            "steps_after_membership_change": self.rebuild_steps,
        }


def poisson_jobs(n_jobs, rate, seed=1000, steps=(500, 5000), member_kw=None, steps_scale=1.0):
    """The serving stream: job i is ensemble_member(seed + i) with a number of
    mechanics steps drawn U[steps] (integer, inclusive) by default_rng(seed + i)
    after the member's own draws; arrival gaps are Exp(rate) drawn by
    default_rng(seed) in job order.  steps_scale shortens every job (for quick
    runs; 1.0 = the stated stream)."""
    member_kw = member_kw or {}
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, n_jobs)
    t = np.cumsum(gaps)
    jobs = []
    for i in range(n_jobs):
        sc, c = S.ensemble_member(seed + i, **member_kw)
        r = np.random.default_rng([seed + i, 1])
        p = int(r.integers(steps[0], steps[1] + 1))
        jobs.append(Job(i, sc, c, max(1, int(round(p * steps_scale))), float(t[i])))
    return jobs


def estimate_bytes(spec, n_cells, mixed, hip=None):
    """Device bytes a job will hold: density, tables, cell SoA, rates, plus its
    share of the packed workspaces (per agent: keys, slots, permutation, hash
    record, plan arrays; per voxel and mechanics bin: counts and offsets)."""
    ns = len(spec.substrates)
    V = spec.n_voxels
    elem = 4 if mixed else 8
    if hip is not None:
        lib = hip.load()
        cells = lib.hc_cells_bytes(max(1, n_cells))
        rates = lib.hc_rates_bytes(max(1, n_cells), ns)
        g = hip.grid_of(spec, mixed)
        tables = lib.hc_thomas_table_bytes(g)
    else:
        cells = max(1, n_cells) * (23 * 8 + 5 * 4)
        rates = max(1, n_cells) * ns * 4 * 8
        tables = 64 * 1024
    rec = 32 if mixed else 64
    ws = n_cells * (10 * 4 + rec + 3 * ns * 8) + (V + spec.n_mech_voxels) * 12
    return int(V * ns * elem + cells + rates + tables + ws)


class Server:
    """Continuous batching over an ensemble object with the interface of
    hip.Ensemble (add, remove, run_periods, wait, density, field,
    member_info)."""

    def __init__(self, make_ensemble: Callable[[], object], budget_bytes=200e9, max_members=64, mixed=True,
                 keep_results=True, clock=time.perf_counter, sleep=time.sleep, bytes_of=None, log=None,
                 progress_s=10.0):
        self.ens = make_ensemble()
        self.budget = float(budget_bytes)
        self.max_members = int(max_members)
        self.mixed = mixed
        self.keep = keep_results
        self.clock = clock
        self.sleep = sleep
        self.bytes_of = bytes_of or (lambda job: estimate_bytes(job.spec, job.cells.n, mixed))
        self.log = log
        self.progress_s = progress_s

    def run(self, jobs: List[Job]):
        pending = sorted(jobs, key=lambda j: (j.arrival, j.id))
        live: Dict[int, dict] = {}
        results: Dict[int, Result] = {}
        st = Stats(jobs=len(jobs))
        used = 0
        changed = True
        t0 = self.clock()
        next_note = self.progress_s
        while pending or live:
            now = self.clock() - t0
            if self.log and now >= next_note:
                self.log(f"t={now:.1f} s: {len(results)} done, {len(live)} live, {len(pending)} waiting, "
                         f"{st.periods_run} batched steps")
                next_note += self.progress_s
            # admission: arrived jobs, in arrival order, while slots and budget allow
            while pending and pending[0].arrival <= now and len(live) < self.max_members:
                need = self.bytes_of(pending[0])
                if need > self.budget:
                    raise ValueError(f"job {pending[0].id} needs {need} bytes, more than the whole budget")
                if used + need > self.budget:
                    break
                job = pending.pop(0)
                if live:
                    st.admitted_while_busy += 1
                mid = self.ens.add(job.spec, job.cells)
                live[mid] = {"job": job, "done": 0, "admitted": now, "bytes": need}
                used += need
                changed = True
            st.max_live = max(st.max_live, len(live))
            st.max_bytes = max(st.max_bytes, used)
            if not live:
                self.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
                continue
            # one batched step of every live job
            t1 = self.clock()
            self.ens.run_periods(1)
            self.ens.wait()
            st.step_ms.append(1e3 * (self.clock() - t1))
            if changed:
                st.rebuild_steps += 1
                changed = False
            st.periods_run += 1
            for mid, rec in live.items():
                rec["done"] += 1
                job = rec["job"]
                st.agent_steps += job.cells.n
                st.voxel_substeps += job.spec.n_voxels * len(job.spec.substrates) * 10
            # completions
            for mid in [m for m, r in live.items() if r["done"] >= r["job"].periods]:
                rec = live.pop(mid)
                job = rec["job"]
                fin = self.clock() - t0
                res = Result(job.id, job.periods, job.arrival, rec["admitted"], fin)
                if self.keep:
                    res.density = self.ens.density(mid)
                    res.positions = np.stack([self.ens.field(mid, f) for f in (0, 1, 2)])
                self.ens.remove(mid)
                used -= rec["bytes"]
                changed = True
                results[job.id] = res
                st.turnaround_s.append(fin - job.arrival)
                st.queue_wait_s.append(rec["admitted"] - job.arrival)
        st.makespan_s = self.clock() - t0
        return results, st


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--jobs", type=int, default=200)
    ap.add_argument("--rate", type=float, default=4.0, help="mean arrivals per second (Poisson)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-members", type=int, default=64)
    ap.add_argument("--steps-scale", type=float, default=1.0, help="multiply every job's mechanics steps")
    ap.add_argument("--fp64", action="store_true", help="plain FP64 members instead of FP32-mixed")
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    from . import hip
    mixed = not args.fp64
    t = time.perf_counter()
    jobs = poisson_jobs(args.jobs, args.rate, seed=args.seed, steps_scale=args.steps_scale)
    print(f"[serve] {len(jobs)} jobs generated in {time.perf_counter() - t:.1f} s; arrivals over "
          f"{jobs[-1].arrival:.1f} s; {sum(j.cells.n * j.periods for j in jobs)} agent-steps requested", flush=True)
    srv = Server(lambda: hip.Ensemble(mixed=mixed), budget_bytes=args.budget_gb * 1e9,
                 max_members=args.max_members, mixed=mixed, keep_results=False,
                 bytes_of=lambda j: estimate_bytes(j.spec, j.cells.n, mixed, hip),
                 log=lambda m: print("[serve]", m, flush=True))
    results, st = srv.run(jobs)
    out = {"workload": f"serving stream: {args.jobs} jobs, Poisson {args.rate}/s, steps U[500,5000] x "
                       f"{args.steps_scale}, budget {args.budget_gb} GB, {args.max_members} slots, "
                       f"{'FP32-mixed' if mixed else 'plain FP64'}", **st.summary()}
    print(json.dumps(out))
    if args.out:
        json.dump(out, open(args.out, "w"), indent=1)


if __name__ == "__main__":
    main()
