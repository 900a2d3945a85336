"""Benchmark of the step hot path on one MI355X.

Workload (default): 200^3 voxels x 4 substrates, 1,000,000 agents, plain FP64
(the large single-simulation configuration, on which the HBM-roofline target
is stated).  One "step" = one mechanics period: 10 diffusion substeps (x, y, z
Thomas sweeps with the Dirichlet re-apply fused into the z sweep, then
secretion/uptake) followed by one mechanics step (hash, forces, AB2).  By
default the periods are launched eagerly and serially (measured as fast as
hipGraph replay, and overlapping the mechanics step with the substeps on a
second stream measured no faster); --graph and --overlap select those forms.
Inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0) with agent-steps/s as `value`, voxel-substeps/s,
p50/p99 step latency, the roofline of the dominant kernel (its algorithmic
bytes over its average launch time, from HIP events on the context stream
around each launch in an eager serial pass of the same periods right after
the timed region), the whole step's roofline (`step_frac`: the step's
algorithmic bytes over ms_per_step), the CPU baseline (the FP64 oracle on
this host's cores, bounded sample) and, under `ensemble`, the 64-member
ensemble timed in its three forms (batched FP32-mixed, batched FP64, one
plain FP64 simulation per member) with the batched forms' speedups.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "physicell-mac-metal_amd"))

from physicell_amd import hip as H  # noqa: E402
from physicell_amd import scenarios as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "agent-steps/s + voxel-substeps/s at 1 GPU, % HBM roofline; p99 step latency"
SWEEP_KERNELS = (H.K_SWEEP_X, H.K_SWEEP_Y, H.K_SWEEP_Z)


def build_workload(name, n_cells):
    if name == "config2":
        return S.config2(n_cells=n_cells)
    if name == "config1":
        return S.config1()
    if name == "config0":
        return S.config0()
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes(kernel, spec, n_cells, mixed=False):
    """Compulsory HBM bytes of one launch (FP64 plain form).
    Sweeps: 1 read + 1 write of every field value = 16 B per voxel per
    substrate.  Forces: 64-byte sorted record read + bin offsets (amortised
    4 B) + velocity/pressure write 32 B per agent.  Integrate: read pos, v,
    v_prev (72 B) + write pos, v_prev (48 B) per agent."""
    V, ns = spec.n_voxels, len(spec.substrates)
    if kernel in SWEEP_KERNELS:
        return (8.0 if mixed else 16.0) * V * ns
    if kernel == H.K_FORCES:
        return ((32.0 + 8.0) if mixed else 64.0) * n_cells + (4.0 + 32.0) * n_cells
    if kernel == H.K_INTEGRATE:
        return 120.0 * n_cells
    return None


def step_bytes(spec, n_cells, substeps, mixed=False):
    """Algorithmic HBM bytes of one step (substeps diffusion substeps + one
    mechanics step), by unit of work: a voxel-substep (VS) moves 16 B per
    sweep in FP64 (8 B in FP32); an agent-substep of secretion (ASS) reads a
    4-byte key and per substrate two folded coefficients and the density RMW
    (4 + 32 S B; 4 + 24 S B mixed); an agent-step (AS) of mechanics moves
    248 B (180 B mixed): key, sort, forces and AB2 traffic."""
    ns = len(spec.substrates)
    sweeps = 2 if spec.nz == 1 else 3
    vs = spec.n_voxels * ns * substeps
    fields = vs * (8.0 if mixed else 16.0) * sweeps
    secretion = n_cells * substeps * (4.0 + (24.0 if mixed else 32.0) * ns)
    mechanics = n_cells * (180.0 if mixed else 248.0)
    return {"fields": fields, "secretion": secretion, "mechanics": mechanics,
            "total": fields + secretion + mechanics,
            "basis": f"VS {vs} x {(8 if mixed else 16) * sweeps} B, ASS {n_cells * substeps} x "
                     f"{4 + (24 if mixed else 32) * ns} B, AS {n_cells} x {180 if mixed else 248} B"}


def load_oracle():
    """The FP64 oracle module, bound to a copy of liboracle compiled for this
    host's CPU (-O3 -march=native, otherwise the flags of oracle/Makefile) in
    a temporary directory when g++ is available; else to the shipped build.
    Returns (module, description of the build)."""
    import shutil
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    build = "shipped oracle/liboracle.so (-O3, generic x86-64)"
    cxx = shutil.which("g++")
    if cxx and O._lib is None:
        tmp = tempfile.mkdtemp(prefix="oracle_native_")
        import atexit
        atexit.register(shutil.rmtree, tmp, True)
        out = os.path.join(tmp, "liboracle.so")
        cmd = [cxx, "-std=c++17", "-O3", "-march=native", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
               "-shared", os.path.join(ROOT, "oracle", "oracle.cpp"), "-o", out]
        try:
            subprocess.run(cmd, check=True, capture_output=True, timeout=180)
            O.LIB_PATH = out
            build = "oracle.cpp compiled on this host: g++ -O3 -march=native -fopenmp -ffp-contract=off"
        except (OSError, subprocess.SubprocessError):
            pass
    return O, build


def cpu_baseline(spec, cells, periods, substeps):
    """FP64 oracle (the reference's algorithm restated in C++, OpenMP) on this
    host: `periods` x (substeps diffusion substeps + 1 mechanics step)."""
    O, build = load_oracle()
    threads = max(1, min(16, os.cpu_count() or 1))

    def timed(nper, nthreads):
        sim = O.OracleSim(spec, cells)
        t0 = time.perf_counter()
        for _ in range(nper):
            for _ in range(substeps):
                sim.substep(nthreads=nthreads)
            sim.mech_step(nthreads=nthreads)
        return time.perf_counter() - t0

    dt = timed(periods, threads)
    dt1 = timed(1, 1)                        # one period on one host thread
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": cells.n * periods / dt, "unit": "agent-steps/s", "cores": threads, "kind": "port",
            "voxel_substeps_per_s": spec.n_voxels * len(spec.substrates) * substeps * periods / dt,
            "one_core": {"value": cells.n / dt1,
                         "voxel_substeps_per_s": spec.n_voxels * len(spec.substrates) * substeps / dt1,
                         "seconds": dt1},
            "host": {"cpu_model": model, "cpu_count": os.cpu_count()}, "build": build,
            "sample": f"{periods} period(s) of {substeps} diffusion substeps + 1 mechanics step of the "
                      f"same workload, FP64 oracle, OpenMP {threads} threads, {dt:.1f} s; and 1 period on "
                      f"1 thread, {dt1:.1f} s"}


def ensemble_forms(members, steps, warmup, overlap=False):
    """The same ensemble members in three forms, each timed over `steps`
    periods after `warmup`: the batched FP32-mixed ensemble, the batched FP64
    ensemble, and the plain FP64 single-simulation path (one pc_sim per
    member, each replaying its own period graph on its own stream).  One step
    = one period of every member (10 diffusion substeps + 1 mechanics step)."""
    n_cells = sum(c.n for _, c in members)
    n_vals = sum(sc.n_voxels * len(sc.substrates) for sc, _ in members)
    n_sub = 10

    def time_ensemble(mixed):
        ens = H.Ensemble(mixed=mixed)
        for sc, c in members:
            ens.add(sc, c)
        ens.set_overlap(overlap)
        blocks, us = ens.autotune(reps=3)
        ens.run_periods(warmup)
        ens.wait()
        t0 = time.perf_counter()
        ens.run_periods(steps)
        ens.wait()
        T = time.perf_counter() - t0
        lat = []
        for _ in range(max(10, min(steps, 50))):
            t1 = time.perf_counter()
            ens.run_periods(1)
            ens.wait()
            lat.append(1e3 * (time.perf_counter() - t1))
        ens.set_overlap(False)
        ens.timing(0xFFFFFFFF)
        ens.run_periods(2, use_graph=False)
        ens.wait()
        kern = {}
        for k, name in enumerate(H.KERNEL_NAMES):
            ms, n = ens.timing_get(k)
            if n:
                kern[name] = {"total_ms": ms, "launches": n, "avg_us": 1e3 * ms / n}
        ens.timing(0)
        ens.close()
        return {"ms_per_step": 1e3 * T / steps, "agent_steps_per_s": n_cells * steps / T,
                "voxel_substeps_per_s": n_vals * n_sub * steps / T,
                "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
                "autotune": {"grid_blocks": blocks, "sweep_us": us}, "kernels": kern}

    res = {"batched_mixed": time_ensemble(True)}
    print(f"[bench] batched mixed: {res['batched_mixed']['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    res["batched_fp64"] = time_ensemble(False)
    print(f"[bench] batched FP64: {res['batched_fp64']['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    # the per-member path at its faster setting: with 64 members on 64
    # streams the device is already busy, and a second stream per member
    # for the mechanics step (overlap) measured slower (15.0 vs 11.3 ms/step)
    sims = [H.Simulation(sc, c) for sc, c in members]
    for sim in sims:
        sim.set_overlap(False)
        sim.run_periods(warmup, use_graph=True)
    for sim in sims:
        sim.wait()
    t0 = time.perf_counter()
    for _ in range(steps):
        for sim in sims:
            sim.run_periods(1, use_graph=True)
    for sim in sims:
        sim.wait()
    T = time.perf_counter() - t0
    res["plain_fp64_per_sim"] = {"ms_per_step": 1e3 * T / steps, "agent_steps_per_s": n_cells * steps / T,
                                 "voxel_substeps_per_s": n_vals * n_sub * steps / T,
                                 "setting": "one pc_sim per member, own stream, period graph, no overlap"}
    for sim in sims:
        sim.close()
    print(f"[bench] plain FP64 per-sim: {res['plain_fp64_per_sim']['ms_per_step']:.2f} ms/step",
          file=sys.stderr, flush=True)
    return res


def ensemble_summary(res):
    """ms/step of the three forms and the ratios of the batched forms to the
    plain FP64 single-simulation path (agent-steps/s and voxel-substeps/s
    ratios are equal: every form runs the same members and periods)."""
    pl = res["plain_fp64_per_sim"]["ms_per_step"]
    return {"batched_mixed_ms_per_step": res["batched_mixed"]["ms_per_step"],
            "batched_fp64_ms_per_step": res["batched_fp64"]["ms_per_step"],
            "plain_fp64_per_sim_ms_per_step": pl,
            "speedup_batched_mixed_vs_plain_fp64": pl / res["batched_mixed"]["ms_per_step"],
            "speedup_batched_fp64_vs_plain_fp64": pl / res["batched_fp64"]["ms_per_step"]}


def bench_ensemble(args):
    """64-simulation ensemble (members ensemble_member(100..163)): the three
    forms of ensemble_forms, with the batched FP32-mixed form as the line."""
    t_setup = time.perf_counter()
    members = S.config3(n_sims=args.members)
    n_cells = sum(c.n for _, c in members)
    n_vals = sum(sc.n_voxels * len(sc.substrates) for sc, _ in members)
    n_sub = 10
    print(f"[bench] ensemble: {len(members)} members, {n_cells} agents, {n_vals} voxel values; "
          f"setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)
    res = ensemble_forms(members, args.steps, args.warmup, overlap=args.overlap)
    bm, pl = res["batched_mixed"], res["plain_fp64_per_sim"]
    kern = bm["kernels"]
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    ab = None
    if dom in ("sweep_x", "sweep_y", "sweep_z"):
        ab = 8.0 * n_vals
    elif dom == "forces":
        ab = (32.0 + 8.0 + 4.0 + 32.0) * n_cells
    roof = {"bound": "hbm", "kernel": dom, "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
            "avg_launch_us": kern[dom]["avg_us"], "algorithmic_bytes_per_launch": ab,
            "achieved": ab / (kern[dom]["avg_us"] * 1e-6) / 1e9 if ab else None}
    roof["frac"] = roof["achieved"] / HBM_PEAK_GBS if ab else None
    # the whole step: every member's algorithmic bytes (mixed form) over the step time
    sb = [step_bytes(sc, c.n, n_sub, mixed=True) for sc, c in members]
    tot = {k: sum(b[k] for b in sb) for k in ("fields", "secretion", "mechanics", "total")}
    roof["step_achieved"] = tot["total"] / (bm["ms_per_step"] * 1e-3) / 1e9
    roof["step_frac"] = roof["step_achieved"] / HBM_PEAK_GBS
    roof["step_bytes"] = tot
    cpu = None
    if not args.no_cpu_baseline:
        O, build = load_oracle()
        threads = max(1, min(16, os.cpu_count() or 1))
        sample = members[:args.cpu_members]
        dt = 0.0
        for sc, c in sample:
            o = O.OracleSim(sc, c)
            t1 = time.perf_counter()
            for _ in range(args.cpu_periods):
                for _ in range(n_sub):
                    o.substep(nthreads=threads)
                o.mech_step(nthreads=threads)
            dt += time.perf_counter() - t1
        P = args.cpu_periods
        cpu = {"value": sum(c.n for _, c in sample) * P / dt, "unit": "agent-steps/s", "cores": threads,
               "kind": "port", "build": build,
               "voxel_substeps_per_s": sum(sc.n_voxels * len(sc.substrates) for sc, _ in sample) * n_sub * P / dt,
               "sample": f"{P} period(s) of members 0..{len(sample) - 1}, FP64 oracle, OpenMP {threads} threads, "
                         f"{dt:.1f} s"}
    out = {
        "metric": METRIC, "value": bm["agent_steps_per_s"], "unit": "agent-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": bm["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"config3 ensemble: {len(members)} members (seeds 100..{99 + len(members)}), "
                               f"{n_cells} agents, {n_vals} voxel values, batched FP32-mixed; step = one period "
                               f"({n_sub} substeps + 1 mechanics step) of every member", "graph": True},
        "voxel_substeps_per_s": bm["voxel_substeps_per_s"], "p50_step_ms": bm["p50_step_ms"],
        "p99_step_ms": bm["p99_step_ms"], "roofline": roof, "cpu_baseline": cpu,
        "ensemble": ensemble_summary(res),
        "forms": res,
    }
    print(json.dumps(out))


def bench_rows(args):
    """The rows beside the step, on the config-2 population (1M agents, 200^3
    x 4 substrates, plain FP64): motility in every mechanics step
    (hc_motility, every agent motile), and the device phenotype step
    (hc_volume_update on every agent, hc_divide of 1%, hc_remove of 1%).
    Kernel times are HIP events on the context stream; algorithmic bytes per
    launch are stated in the output; the oracle runs the same operations on
    this host (one thread) beside them."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    spec, cells = build_workload("config2", args.cells)
    n = cells.n
    M = n + n // 50
    sim = H.Simulation(spec, cells, max_cells=M)
    r = np.random.default_rng(2024)
    pers = r.uniform(1.0, 10.0, n)
    sim.set_motility(seed=11, persistence=pers, bias=np.full(n, 0.5), speed=np.ones(n),
                     is_motile=np.ones(n, np.uint32), restrict_2d=np.zeros(n, np.uint32),
                     chem_index=np.zeros(n, np.int32), chem_dir=np.ones(n, np.int32))
    sim.set_volume(seed=12)
    ctx = sim.context()
    lib = H.load()
    print(f"[bench] rows: {n} agents, motility on, warmup {args.warmup} periods", file=sys.stderr, flush=True)
    sim.run_periods(max(1, args.warmup), use_graph=True)
    sim.wait()
    t0 = time.perf_counter()
    sim.run_periods(args.steps, use_graph=True)
    sim.wait()
    ms_step = 1e3 * (time.perf_counter() - t0) / args.steps
    # kernel times from eager periods (graph replays record no events)
    lib.hc_timing_reset(ctx)
    lib.hc_timing_enable(ctx, 1 << H.K_MOTILITY)
    sim.run_periods(args.steps, use_graph=False)
    sim.wait()
    out = {}

    def kern(k):
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        lib.hc_timing_get(ctx, k, ctypes.byref(ms), ctypes.byref(cnt))
        return 1e3 * ms.value / max(1, cnt.value), cnt.value

    us, cnt = kern(H.K_MOTILITY)
    resample = float(np.minimum(1.0, spec.dt_mech / pers).sum())
    mot_bytes = 104.0 * n + 100.0 * resample
    out["motility"] = {"avg_us": us, "launches": cnt, "bytes_per_launch": mot_bytes,
                       "achieved_gbs": mot_bytes / us * 1e-3, "frac": mot_bytes / us * 1e-3 / HBM_PEAK_GBS,
                       "bytes_basis": "104 B per motile agent (alive, motile, elapsed RMW, persistence, velocity "
                                      "RMW, vector) + 100 B per resample (bias, flags, speed, position, gradient)"}
    # device phenotype step, 1% divisions and 1% removals per call
    lib.hc_timing_reset(ctx)
    lib.hc_timing_enable(ctx, (1 << H.K_VOLUME) | (1 << H.K_DIVIDE) | (1 << H.K_REMOVE))
    calls, walls = 3, []
    cur = n
    for _ in range(calls):
        parents = r.choice(cur, n // 100, replace=False).astype(np.uint32)
        removed = r.choice(cur + len(parents), n // 100, replace=False).astype(np.uint32)
        t1 = time.perf_counter()
        cur = sim.update_phenotype(6.0, parents=parents, removed=removed)
        walls.append(1e3 * (time.perf_counter() - t1))
    ns = len(spec.substrates)
    per = {H.K_VOLUME: ("volume", 192.0 * n, "192 B per agent (2 volumes + 12 state reads; 4 geometry + 6 state writes)"),
           H.K_DIVIDE: ("divide", (2 * 8 * 23 + 2 * 4 * 5 + 2 * 8 * 15 + 4 * 8 * ns * 2) * (n // 100),
                        "per division: parent read + parent and daughter writes of the cell SoA, volume state, rates"),
           H.K_REMOVE: ("remove", 2 * (8 * 23 + 4 * 5 + 8 * 15 + 4 * 8 * ns) * (n // 100),
                        "per removal at most one move: read + write of the cell SoA, volume state, rates")}
    for k, (name, b, basis) in per.items():
        us, cnt = kern(k)
        out[name] = {"avg_us": us, "launches": cnt, "bytes_per_launch": b, "achieved_gbs": b / us * 1e-3,
                     "frac": b / us * 1e-3 / HBM_PEAK_GBS, "bytes_basis": basis}
    # neighbour queries over all agents (hash from the current positions,
    # count, scan, fill, copy-out); the second call of each is timed
    queries = {}
    for kind, name in ((H.QUERY_CONTACT, "contact_graph"), (H.QUERY_INTERACT, "interaction_candidates")):
        for _ in range(2):
            t1 = time.perf_counter()
            qoff, qidx = sim.neighbor_query(kind)
            ms = 1e3 * (time.perf_counter() - t1)
        queries[name] = {"ms": ms, "entries": int(len(qidx)), "cells": int(len(qoff) - 1)}
    # attachments (tryFormAllAttachments): device pair scan + host greedy pass,
    # rates U[0, 0.2] per minute over dt 1 min; the second call is timed
    arate = r.uniform(0.0, 0.2, sim.n)
    for k in range(2):
        t1 = time.perf_counter()
        acnt, _, aformed = sim.form_attachments(1.0, arate, seed=20 + k)
        ms = 1e3 * (time.perf_counter() - t1)
    queries["form_attachments"] = {"ms": ms, "formed": int(aformed), "cells": int(len(acnt))}
    # the oracle beside it (one host thread), same operations and sizes
    cpu = {}
    sample = 20000
    a = {k: np.ascontiguousarray(sim.field(f)[:sample]) for k, f in
         (("x", H.F_POS_X), ("y", H.F_POS_Y), ("z", H.F_POS_Z), ("r", H.F_RADIUS), ("v", H.F_TOTAL_VOL),
          ("rel", H.F_REL_MAX_ADH))}
    t1 = time.perf_counter()
    O.contact_graph(a["x"], a["y"], a["z"], a["r"], a["v"], a["rel"], np.ascontiguousarray(sim.field(H.F_IS_ALIVE)[:sample]))
    cpu["contact_graph_ms_20k_cells"] = 1e3 * (time.perf_counter() - t1)
    t1 = time.perf_counter()
    O.form_attachments(a["x"], a["y"], a["z"], a["r"], a["rel"], np.ascontiguousarray(sim.field(H.F_IS_ALIVE)[:sample]),
                       np.ascontiguousarray(arate[:sample]), 1.0, 21)
    cpu["form_attachments_ms_20k_cells"] = 1e3 * (time.perf_counter() - t1)
    x, y, z = (np.ascontiguousarray(np.asarray(a, np.float64)) for a in (cells.x, cells.y, cells.z))
    vel = [np.zeros(n) for _ in range(3)]
    mv = [np.zeros(n) for _ in range(3)]
    rho = np.full(spec.n_voxels * ns, 1.0)
    t1 = time.perf_counter()
    O.motility(x, y, z, *vel, np.ones(n, np.uint32), np.ones(n), O.grid(spec), rho, spec.dt_mech, 11, 0, pers.copy(),
               np.zeros(n), np.full(n, 0.5), *mv, np.ones(n, np.uint32), np.zeros(n, np.uint32),
               np.zeros(n, np.int32), np.ones(n, np.int32))
    cpu["motility_ms"] = 1e3 * (time.perf_counter() - t1)
    f64 = np.zeros((23, M))
    f64[H.F_TOTAL_VOL, :n], f64[H.F_NUCLEAR_VOL, :n], f64[H.F_ONCOPROTEIN, :n] = 2494.0, 540.0, 1.0
    vol = np.zeros((15, M))
    vol[:, :] = np.array([0.75 * 2494, 135.0, 488.5, 0.75, 0, 488.5, 135.0, 0.75, 3.62, 0.05, 0.0055, 0.0045, 0,
                          4988.0, 2.0])[:, None]
    pc = O.PhenoCells(f64, np.zeros((5, M), np.uint32), np.zeros((4, M * ns)), vol, ns)
    t1 = time.perf_counter()
    O.volume_update(pc, None, n, 6.0, 1)
    cpu["volume_ms"] = 1e3 * (time.perf_counter() - t1)
    parents = r.choice(n, n // 100, replace=False).astype(np.uint32)
    t1 = time.perf_counter()
    n1 = O.divide(pc, parents, n, M, 12, 0)
    cpu["divide_ms"] = 1e3 * (time.perf_counter() - t1)
    t1 = time.perf_counter()
    O.remove(pc, r.choice(n1, n // 100, replace=False).astype(np.uint32), n1)
    cpu["remove_ms"] = 1e3 * (time.perf_counter() - t1)
    line = {
        "metric": "kernel time and HBM roofline fraction of the rows beside the step (motility every mechanics "
                  "step; device phenotype step: volume model, 1% divisions, 1% removals)",
        "value": n / (out["motility"]["avg_us"] * 1e-6), "unit": "motile agent-steps/s (motility kernel)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"config2 population ({n} agents, 200^3 x {ns} substrates), every agent motile "
                               f"(persistence U[1,10] min, bias 0.5, speed 1); phenotype calls: {n // 100} "
                               f"divisions + {n // 100} removals each", "graph": True},
        "kernels": out,
        "phenotype_call_ms": walls,
        "neighbor_queries": queries,
        "cpu_baseline": {"kind": "port", "cores": 1, **cpu,
                         "sample": f"the FP64 oracle on this host, one thread: motility over {n} agents, volume "
                                   f"model over {n}, {n // 100} divisions, {n // 100} removals; the contact "
                                   f"graph's and the attachments' O(N^2) loops (the reference's) over the first {sample} "
                                   f"agents only"},
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="config2")
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--graph", action="store_true",
                    help="replay each period as a hipGraph in the timed region (kernel times then come from a "
                         "separate eager pass)")
    ap.add_argument("--overlap", action="store_true",
                    help="run each period's mechanics step concurrently with its substeps (kernel times then come "
                         "from a separate serial pass)")
    ap.add_argument("--no-graph", action="store_true", help="(default) eager launches in the timed region")
    ap.add_argument("--no-overlap", action="store_true", help="(default) serial periods")
    ap.add_argument("--mixed", action="store_true", help="FP32-mixed form instead of plain FP64")
    ap.add_argument("--sweep-mode", type=int, default=0,
                    help="0 partitioned sweeps (default), 1 serial elimination, 2 partitioned with LDS-staged x")
    ap.add_argument("--substep-order", type=int, default=0,
                    help="0 default (axis-major), 1 axis-major, 2 substrate-major sweeps inside a substep")
    ap.add_argument("--cpu-periods", type=int, default=12, help="periods in the CPU sample (about 9 s on 16 host threads)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_config2.json"))
    ap.add_argument("--members", type=int, default=64, help="ensemble size (workload config3)")
    ap.add_argument("--cpu-members", type=int, default=64, help="members in the CPU sample (workload config3)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="leave out the ensemble leg of the default line (config3 members in the batched mixed, "
                         "batched FP64 and per-member FP64 forms)")
    ap.add_argument("--ensemble-steps", type=int, default=20, help="timed steps per form in the ensemble leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain FP64 device copy of 2^27 elements (1 GiB read + 1 GiB "
                         "written) five times, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); every box here has one GPU "
              f"and the multi-GPU decomposition is design-only.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "config3":
        return bench_ensemble(args)
    if args.workload == "rows":
        return bench_rows(args)

    t_setup = time.perf_counter()
    spec, cells = build_workload(args.workload, args.cells)
    sim = H.Simulation(spec, cells, mixed=args.mixed)
    sim.set_overlap(args.overlap)
    if args.sweep_mode:
        rc = H.load().hc_set_sweep_mode(sim.context(), args.sweep_mode)
        if rc:
            raise SystemExit("hc_set_sweep_mode failed")
    if args.substep_order:
        if H.load().hc_set_substep_order(sim.context(), args.substep_order):
            raise SystemExit("hc_set_substep_order failed")
    k_sub = 10 if abs(spec.dt_mech / spec.dt_diff - 10.0) < 1e-9 else None
    use_graph = args.graph
    print(f"[bench] {spec.name}: grid {spec.nx}x{spec.ny}x{spec.nz} x {len(spec.substrates)} substrates, "
          f"{cells.n} agents; setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    # warmup (also builds / instantiates the graph)
    sim.run_periods(args.warmup, use_graph=use_graph)
    sim.wait()

    # timed region: K periods
    sim.wait()
    t0 = time.perf_counter()
    sim.run_periods(args.steps, use_graph=use_graph)
    sim.wait()
    T = time.perf_counter() - t0
    ms_per_step = 1e3 * T / args.steps
    print(f"[bench] timed {args.steps} steps: {T:.3f} s", file=sys.stderr, flush=True)

    # per-step latency: each period launched and waited for on its own
    lat = []
    for _ in range(max(10, min(args.steps, 100))):
        t1 = time.perf_counter()
        sim.run_periods(1, use_graph=use_graph)
        sim.wait()
        lat.append(1e3 * (time.perf_counter() - t1))
    lat = np.array(lat)

    # kernel timing: HIP events on the context stream around each launch, in
    # an eager serial pass right after the timed region: by default the same
    # launches the timed region makes, without events in the timed region
    # itself (events around every launch cost 9% of the step, around the
    # sweeps 5%; graph replays record none, and overlapped kernels contend)
    sim.set_overlap(False)
    sim.timing_reset()
    sim.timing(0xFFFFFFFF)
    kp = max(2, min(args.steps, 10))
    sim.run_periods(kp, use_graph=False)
    sim.wait()
    kern = {}
    for k, name in enumerate(H.KERNEL_NAMES):
        ms, n = sim.timing_get(k)
        if n:
            kern[name] = {"total_ms": ms, "launches": n, "avg_us": 1e3 * ms / n}
    sim.timing(0)
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    dom_id = H.KERNEL_NAMES.index(dom)
    ab = algorithmic_bytes(dom_id, spec, cells.n, args.mixed)
    avg_s = kern[dom]["avg_us"] * 1e-6
    if ab and dom_id in SWEEP_KERNELS:
        # a substep may run one launch per substrate (substrate-major order)
        ab /= kern[dom]["launches"] / (kp * (k_sub or 10))
    traffic = None
    if os.path.exists(args.traffic):
        try:
            tr = json.load(open(args.traffic))
            traffic = tr.get(dom, {}).get("bytes_per_launch")
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "kernel": dom, "achieved": ab / avg_s / 1e9 if ab else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (ab / avg_s / 1e9) / HBM_PEAK_GBS if ab else None, "traffic": traffic,
            "algorithmic_bytes_per_launch": ab, "avg_launch_us": kern[dom]["avg_us"]}
    # the three sweeps together (the diffusion substep's roofline)
    sw = [kern[n] for n in ("sweep_x", "sweep_y", "sweep_z") if n in kern]
    if sw:
        sweep_bytes = (8.0 if args.mixed else 16.0) * spec.n_voxels * len(spec.substrates) * len(sw)
        sweep_s = sum(x["total_ms"] for x in sw) * 1e-3 / (kp * (k_sub or 10))
        roof["sweeps_combined_frac"] = sweep_bytes / sweep_s / 1e9 / HBM_PEAK_GBS

    n_sub = k_sub or 10
    value = cells.n * args.steps / T
    vss = spec.n_voxels * len(spec.substrates) * n_sub * args.steps / T
    # the whole step's roofline: algorithmic bytes of one step over its time
    sb = step_bytes(spec, cells.n, n_sub, args.mixed)
    roof["step_achieved"] = sb["total"] / (ms_per_step * 1e-3) / 1e9
    roof["step_frac"] = roof["step_achieved"] / HBM_PEAK_GBS
    roof["step_bytes"] = sb
    sim.close()
    ens = None
    if not args.no_ensemble and args.workload == "config2":
        t1 = time.perf_counter()
        members = S.config3(n_sims=args.members)
        print(f"[bench] ensemble leg: {len(members)} members, {sum(c.n for _, c in members)} agents; setup "
              f"{time.perf_counter() - t1:.1f} s", file=sys.stderr, flush=True)
        ens = ensemble_summary(ensemble_forms(members, args.ensemble_steps, 3))
        ens["workload"] = (f"config3 ensemble: {len(members)} members (seeds 100..{99 + len(members)}), "
                           f"{sum(c.n for _, c in members)} agents, "
                           f"{sum(sc.n_voxels * len(sc.substrates) for sc, _ in members)} voxel values; "
                           f"step = one period of every member; {args.ensemble_steps} timed steps per form")
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(spec, cells, args.cpu_periods, n_sub)
    out = {
        "metric": METRIC, "value": value, "unit": "agent-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if args.mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"{spec.name}: {spec.nx}x{spec.ny}x{spec.nz} voxels x {len(spec.substrates)} "
                               f"substrates, {cells.n} agents, {'FP32-mixed' if args.mixed else 'plain FP64'}; step = {n_sub} diffusion substeps "
                               f"+ 1 mechanics step", "graph": use_graph, "overlap": args.overlap,
                   "kernel_times": f"HIP events over {kp} eager serial periods right after the timed region",
                   "sweep_mode": args.sweep_mode,
                   "substep_order": args.substep_order},
        "voxel_substeps_per_s": vss,
        "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
        "roofline": roof, "kernels": kern, "cpu_baseline": cpu, "ensemble": ens,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
