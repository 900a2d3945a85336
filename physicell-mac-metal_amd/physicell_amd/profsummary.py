"""Summaries of rocprofv3 result databases (rocpd SQLite) for this library.

    python -m physicell_amd.profsummary trace  <results.db>
    python -m physicell_amd.profsummary traffic <fetch.db> <write.db> <copy_fetch.db> <copy_write.db> [--json out]
    python -m physicell_amd.profsummary counters <pmc.db> [<pmc.db> ...] [--kernels a,b,...]

`trace` prints per-kernel calls / total / average duration (the
`--kernel-trace --stats` view).  `counters` prints the per-dispatch average
of every counter found in the databases, per kernel (optionally only the
kernels whose names contain one of the given substrings).  `traffic` turns per-dispatch FETCH_SIZE and
WRITE_SIZE readings (KiB) into bytes per launch, scaled by what the same
counters read for a plain device copy of known size (1 GiB read + 1 GiB
written, `bench.py --calibrate-copy`): scale = known bytes / counter bytes.
"""
import json
import re
import sqlite3
import sys

KERNEL_ALIASES = [
    ("k_sweep_x", "sweep_x"), ("k_sweep_lane<0", "sweep_x"), ("k_sweep_lane<1", "sweep_y"),
    ("k_sweep_lane<2", "sweep_z"), ("k_sweep_part<0", "sweep_x"), ("k_sweep_part<1", "sweep_y"),
    ("k_sweep_part<2", "sweep_z"), ("k_forces", "forces"), ("k_sec_apply", "secretion_apply"),
    ("k_integrate", "integrate"), ("k_stable_gather", "hash_gather"), ("k_mech_keys", "hash_keys"),
]


def short(name):
    for key, alias in KERNEL_ALIASES:
        if key in name:
            return alias
    return re.sub(r"\(.*", "", name)


def trace(db):
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    # the top_kernels view reports durations in microseconds
    out = [f"{'kernel':<60s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>10s} {'pct':>6s}"]
    for name, calls, tot, avg, pct in rows:
        out.append(f"{re.sub(r'[(].*', '', name)[:60]:<60s} {calls:>7d} {tot / 1e3:>10.2f} {avg:>10.1f} {pct:>6.2f}")
    return "\n".join(out)


def per_kernel_counter(db, counter):
    c = sqlite3.connect(db)
    acc = {}
    for name, val in c.execute("select kernel_name, value from counters_collection where counter_name=?", (counter,)):
        acc.setdefault(name, []).append(val)
    return acc


def copy_scale(db, counter, known_bytes):
    acc = per_kernel_counter(db, counter)
    # the copy is the largest-reading dispatch
    best = max((max(v) for v in acc.values()), default=0.0)
    return known_bytes / (best * 1024.0) if best else None


def traffic(fetch_db, write_db, cfetch_db, cwrite_db):
    known = float(1 << 30)
    sf = copy_scale(cfetch_db, "FETCH_SIZE", known)
    sw = copy_scale(cwrite_db, "WRITE_SIZE", known)
    f = per_kernel_counter(fetch_db, "FETCH_SIZE")
    w = per_kernel_counter(write_db, "WRITE_SIZE")
    res = {"scale_fetch": sf, "scale_write": sw}
    for name in sorted(set(f) | set(w)):
        fk = sum(f.get(name, [0])) / max(1, len(f.get(name, [1])))
        wk = sum(w.get(name, [0])) / max(1, len(w.get(name, [1])))
        fetch_b = fk * 1024.0 * (sf or 1.0)
        write_b = wk * 1024.0 * (sw or 1.0)
        key = short(name)
        if key in res:                        # two kernels with one short name: keep both apart
            key = re.sub(r"\(.*", "", name)
        res[key] = {"fetch_bytes": fetch_b, "write_bytes": write_b,
                    "bytes_per_launch": fetch_b + write_b, "dispatches": len(f.get(name, []))}
    return res


def counters(dbs, kernels=None):
    """{kernel: {counter: per-dispatch average}} over the given databases."""
    out = {}
    for db in dbs:
        c = sqlite3.connect(db)
        rows = c.execute("select kernel_name, counter_name, value from counters_collection")
        acc = {}
        for name, cn, val in rows:
            acc.setdefault((name, cn), []).append(val)
        for (name, cn), vals in acc.items():
            if kernels and not any(k in name for k in kernels):
                continue
            out.setdefault(re.sub(r"[(].*", "", name), {})[cn] = sum(vals) / len(vals)
    return out


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        print(trace(sys.argv[2]))
    elif sys.argv[1] == "counters":
        args = sys.argv[2:]
        ks = None
        if "--kernels" in args:
            i = args.index("--kernels")
            ks = args[i + 1].split(",")
            args = args[:i] + args[i + 2:]
        for k, v in sorted(counters(args, ks).items()):
            print(k)
            for cn, val in sorted(v.items()):
                print(f"    {cn:<28s} {val:16.1f}")
    elif sys.argv[1] == "traffic":
        r = traffic(*sys.argv[2:6])
        if "--json" in sys.argv:
            json.dump(r, open(sys.argv[sys.argv.index("--json") + 1], "w"), indent=1)
        for k, v in r.items():
            if isinstance(v, dict):
                print(f"{k:<20s} fetch {v['fetch_bytes'] / 1e6:10.1f} MB  write {v['write_bytes'] / 1e6:10.1f} MB  "
                      f"dispatches {v['dispatches']}")
            else:
                print(k, v)
