"""Snapshot-pair comparison with the semantics of the reference's parity
harness (parity/diff.py:150-300), so that a reference run and a run of this
library are judged the same way:

- cells (output%08d_cells.mat): columns matched by cell ID (row 0, the last
  column wins for a repeated ID, parity/mat_reader.py:86-92); the largest
  absolute error of rows 1-3 (position, tolerance `position`, default 1e-6)
  and row 4 (total volume, `volume`, 1e-3) must not exceed its tolerance,
  nor that of row 7 (phase index, `phase`, 0) when both files have it.  A
  cell-count difference is reported and fails only when `cell_count_abs` is
  given and exceeded; no shared ID fails.
- microenvironment (output%08d_microenvironment0.mat): the voxel counts
  must agree and every voxel centre (rows 0-2) within `voxel_coord` (1e-4);
  then every substrate row present in both must keep its largest error
  within `substrate_linf` (1e-3) and its RMSE within `substrate_rmse` (1e-4).

Every field also reports its normwise relative error (largest absolute
error over the largest reference magnitude), the measure of the FP64 bar
(1e-10).  The arithmetic is numpy's; the reference walks Python lists.

    python -m physicell_amd.snapdiff REF_DIR OURS_DIR [--frames 0 3] [--tol position=1e-9 ...]
"""
import argparse
import glob
import os
import re
import sys
from dataclasses import dataclass, field

import numpy as np

from . import matio

CELL_FIELDS = (("pos_x", matio.CELL_ROW_X, "position", 1e-6), ("pos_y", matio.CELL_ROW_Y, "position", 1e-6),
               ("pos_z", matio.CELL_ROW_Z, "position", 1e-6),
               ("volume", matio.CELL_ROW_TOTAL_VOLUME, "volume", 1e-3))


@dataclass
class FieldStat:
    name: str
    n: int = 0
    max_abs: float = 0.0
    mean_abs: float = 0.0
    rmse: float = 0.0
    rel: float = 0.0
    ref_min: float = 0.0
    ref_max: float = 0.0
    tol: float = 0.0
    ok: bool = True

    def line(self):
        return (f"  {self.name:<14s} n={self.n:<8d} Linf={self.max_abs:.6g} RMSE={self.rmse:.6g} "
                f"rel={self.rel:.3g} ref in [{self.ref_min:.4g}, {self.ref_max:.4g}] tol={self.tol:.4g} "
                f"{'PASS' if self.ok else 'FAIL'}")


def field_stat(name, ref, ours, tol):
    ref = np.asarray(ref, np.float64)
    ours = np.asarray(ours, np.float64)
    if ref.size == 0:
        return FieldStat(name, tol=tol)
    e = np.abs(ref - ours)
    mx = float(e.max())
    scale = float(np.abs(ref).max())
    return FieldStat(name, n=int(ref.size), max_abs=mx, mean_abs=float(e.mean()), rmse=float(np.sqrt(np.mean(e * e))),
                     rel=mx / scale if scale > 0 else mx, ref_min=float(ref.min()), ref_max=float(ref.max()),
                     tol=tol, ok=bool(mx <= tol))


@dataclass
class Part:
    ok: bool = True
    counts: tuple = (0, 0)
    rows: tuple = (0, 0)
    matched: int = 0
    only_ref: int = 0
    only_ours: int = 0
    fields: list = field(default_factory=list)
    notes: list = field(default_factory=list)


def _read(path, part, who):
    try:
        return matio.read_mat(path)[1]
    except Exception as e:          # reported, as the harness does, never raised
        part.ok = False
        part.notes.append(f"cannot read {who} {path}: {e}")
        return None


def _by_id(m):
    ids = m[matio.CELL_ROW_ID].astype(np.int64)
    return dict(zip(ids.tolist(), range(len(ids))))


def diff_cells(ref_path, ours_path, tol=None):
    tol = tol or {}
    out = Part()
    a, b = _read(ref_path, out, "ref"), _read(ours_path, out, "ours")
    if a is None or b is None:
        return out
    out.counts, out.rows = (a.shape[1], b.shape[1]), (a.shape[0], b.shape[0])
    if a.shape[0] != b.shape[0]:
        out.notes.append(f"row schema differs: ref {a.shape[0]} rows, ours {b.shape[0]}; leading rows compared")
    ia, ib = _by_id(a), _by_id(b)
    common = sorted(set(ia) & set(ib))
    out.matched, out.only_ref, out.only_ours = len(common), len(set(ia) - set(ib)), len(set(ib) - set(ia))
    if not common:
        out.ok = False
        out.notes.append("no shared cell IDs")
        return out
    ca = np.array([ia[i] for i in common])
    cb = np.array([ib[i] for i in common])
    rows = list(CELL_FIELDS)
    if min(a.shape[0], b.shape[0]) > matio.CELL_ROW_PHASE:
        rows.append(("cycle_phase", matio.CELL_ROW_PHASE, "phase", 0.0))
    for name, row, key, default in rows:
        out.fields.append(field_stat(name, a[row, ca], b[row, cb], tol.get(key, default)))
    if out.counts[0] != out.counts[1]:
        out.notes.append(f"cell count differs: ref {out.counts[0]}, ours {out.counts[1]} (matched {out.matched}, "
                         f"only ref {out.only_ref}, only ours {out.only_ours})")
        lim = tol.get("cell_count_abs")
        if lim is not None and abs(out.counts[0] - out.counts[1]) > lim:
            out.ok = False
    out.ok = out.ok and all(f.ok for f in out.fields)
    return out


def diff_microenvironment(ref_path, ours_path, tol=None):
    tol = tol or {}
    out = Part()
    a, b = _read(ref_path, out, "ref"), _read(ours_path, out, "ours")
    if a is None or b is None:
        return out
    out.counts, out.rows = (a.shape[1], b.shape[1]), (a.shape[0], b.shape[0])
    if a.shape[1] != b.shape[1]:
        out.ok = False
        out.notes.append(f"voxel count differs: ref {a.shape[1]}, ours {b.shape[1]}; substrates not compared")
        return out
    if a.shape[0] != b.shape[0]:
        out.notes.append(f"substrate count differs: ref {a.shape[0]} rows, ours {b.shape[0]}")
    ct = tol.get("voxel_coord", 1e-4)
    bad = int(np.count_nonzero(np.any(np.abs(a[:3] - b[:3]) > ct, axis=0)))
    if bad:
        out.ok = False
        out.notes.append(f"{bad}/{a.shape[1]} voxel centres differ by more than {ct}")
        return out
    ns = min(a.shape[0], b.shape[0]) - matio.ME_FIRST_SUBSTRATE
    if ns <= 0:
        out.notes.append("no substrate rows")
        return out
    linf, rmse = tol.get("substrate_linf", 1e-3), tol.get("substrate_rmse", 1e-4)
    for s in range(ns):
        r = matio.ME_FIRST_SUBSTRATE + s
        st = field_stat(f"substrate_{s}", a[r], b[r], linf)
        st.ok = st.max_abs <= linf and st.rmse <= rmse
        out.fields.append(st)
    out.ok = all(f.ok for f in out.fields)
    return out


def diff_frame(ref_dir, ours_dir, frame, tol=None):
    """(cells Part or None, microenvironment Part or None) for one frame; a
    file missing on either side leaves its part None, as in the harness."""
    label = f"output{frame:08d}"
    parts = []
    for suffix, fn in (("_cells.mat", diff_cells), ("_microenvironment0.mat", diff_microenvironment)):
        pa, pb = os.path.join(ref_dir, label + suffix), os.path.join(ours_dir, label + suffix)
        parts.append(fn(pa, pb, tol) if os.path.isfile(pa) and os.path.isfile(pb) else None)
    return tuple(parts)


def frames_in(folder):
    got = set()
    for p in glob.glob(os.path.join(folder, "output*_cells.mat")) + \
            glob.glob(os.path.join(folder, "output*_microenvironment0.mat")):
        m = re.match(r"output(\d{8})_", os.path.basename(p))
        if m:
            got.add(int(m.group(1)))
    return sorted(got)


def report(frame, cells, micro):
    ok = (cells is None or cells.ok) and (micro is None or micro.ok)
    lines = [f"[{'PASS' if ok else 'FAIL'}] output{frame:08d}"]
    for what, p in (("cells", cells), ("microenvironment", micro)):
        if p is None:
            continue
        lines.append(f" {what}: ref {p.counts[0]}, ours {p.counts[1]} (rows {p.rows[0]} / {p.rows[1]})"
                     + (f", matched {p.matched}" if what == "cells" else ""))
        lines += [f"   note: {n}" for n in p.notes]
        lines += [f.line() for f in p.fields]
    return ok, "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("ref_dir")
    ap.add_argument("ours_dir")
    ap.add_argument("--frames", type=int, nargs="*", help="default: every frame present in both folders")
    ap.add_argument("--tol", nargs="*", default=[], help="key=value (position, volume, phase, cell_count_abs, "
                                                          "voxel_coord, substrate_linf, substrate_rmse)")
    a = ap.parse_args(argv)
    tol = {k: float(v) for k, v in (t.split("=", 1) for t in a.tol)}
    frames = a.frames if a.frames else sorted(set(frames_in(a.ref_dir)) & set(frames_in(a.ours_dir)))
    if not frames:
        print("no frame present in both folders")
        return 1
    all_ok = True
    for f in frames:
        ok, text = report(f, *diff_frame(a.ref_dir, a.ours_dir, f, tol))
        print(text)
        all_ok = all_ok and ok
    return 0 if all_ok else 1


if __name__ == "__main__":
    sys.exit(main())
