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

Exact mode (--exact): the integer outputs of the path, written beside the
snapshot by write_exact as Level-4 matrices (integers below 2^53 are exact
in them), must be equal, value for value:
- output%08d_keys.mat: per cell (one column each) its ID, cell->voxel key
  and mechanics-voxel key;
- output%08d_bins.mat: the hash's bin offsets (one row, bins + 1 values) and
  output%08d_sorted.mat: the cells in hash order;
- output%08d_neighbors.mat: the neighbour lists of the force predicates
  (same or adjacent mechanics voxel, d^2 < cutoff^2, d^2 >= 1e-12), two rows
  (cell, neighbour) in visiting order; compared as sorted sets per cell and
  in visiting order.
A file missing on both sides is skipped; on one side it fails.

    python -m physicell_amd.snapdiff REF_DIR OURS_DIR [--frames 0 3] [--tol position=1e-9 ...] [--exact]
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


EXACT_FILES = ("keys", "bins", "sorted", "neighbors")


def write_exact(folder, frame, cell_ids, voxel_keys, mech_keys, offsets, sorted_cells, nbr_offsets, nbr_idx):
    """The exact-mode files of one frame: keys, bins, sorted list and the
    neighbour lists (CSR: nbr_offsets[n + 1], nbr_idx in visiting order)."""
    label = os.path.join(str(folder), f"output{frame:08d}")
    ids = np.asarray(cell_ids, np.float64)
    matio.write_mat(label + "_keys.mat", "keys",
                    np.vstack([ids, np.asarray(voxel_keys, np.float64), np.asarray(mech_keys, np.float64)]))
    matio.write_mat(label + "_bins.mat", "bins", np.asarray(offsets, np.float64)[None, :])
    matio.write_mat(label + "_sorted.mat", "sorted", np.asarray(sorted_cells, np.float64)[None, :])
    nbr_offsets = np.asarray(nbr_offsets, np.int64)
    owner = np.repeat(np.arange(len(nbr_offsets) - 1), np.diff(nbr_offsets))
    matio.write_mat(label + "_neighbors.mat", "neighbors",
                    np.vstack([owner.astype(np.float64), np.asarray(nbr_idx, np.float64)]))


@dataclass
class ExactStat:
    name: str
    ok: bool = True
    n: int = 0
    note: str = ""

    def line(self):
        return f"  {self.name:<14s} n={self.n:<10d} {'EXACT' if self.ok else 'DIFFER'} {self.note}"


def _first_diff(a, b):
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    idx = np.argwhere(a != b)
    return f"{len(idx)} values differ, first at {tuple(int(v) for v in idx[0])}: {a[tuple(idx[0])]!r} vs " \
           f"{b[tuple(idx[0])]!r}" if len(idx) else ""


def diff_exact(ref_dir, ours_dir, frame):
    """[ExactStat] for the exact-mode files of one frame."""
    label = f"output{frame:08d}"
    out = []
    for what in EXACT_FILES:
        pa = os.path.join(ref_dir, f"{label}_{what}.mat")
        pb = os.path.join(ours_dir, f"{label}_{what}.mat")
        ha, hb = os.path.isfile(pa), os.path.isfile(pb)
        if not ha and not hb:
            continue
        if ha != hb:
            out.append(ExactStat(what, False, 0, f"present only in {'ref' if ha else 'ours'}"))
            continue
        a, b = matio.read_mat(pa)[1], matio.read_mat(pb)[1]
        if what == "neighbors":
            st = ExactStat("neighbors", True, a.shape[1])
            if a.shape != b.shape or not np.array_equal(a[0], b[0]):
                st.ok, st.note = False, "neighbour counts differ: " + _first_diff(
                    np.bincount(a[0].astype(np.int64)), np.bincount(b[0].astype(np.int64)))
            else:
                # sorted sets per cell, then the visiting order
                ka = np.lexsort((a[1], a[0]))
                kb = np.lexsort((b[1], b[0]))
                if not np.array_equal(a[1][ka], b[1][kb]):
                    st.ok, st.note = False, "sets differ: " + _first_diff(a[:, ka], b[:, kb])
                elif not np.array_equal(a[1], b[1]):
                    st.ok, st.note = False, "same sets, visiting order differs: " + _first_diff(a, b)
                else:
                    st.note = "(sets and visiting order)"
            out.append(st)
        else:
            d = _first_diff(a, b)
            out.append(ExactStat(what, not d, a.size, d))
    return out


def compare(ref_dir, ours_dir, frames=None, tol=None, exact=False):
    """(all passed, report text) over the frames present in both folders (or
    the given ones): the snapshot rules, plus the exact-mode files."""
    frames = frames if frames else sorted(set(frames_in(ref_dir)) & set(frames_in(ours_dir)))
    if not frames:
        return False, "no frame present in both folders"
    all_ok, texts = True, []
    for f in frames:
        ok, text = report(f, *diff_frame(ref_dir, ours_dir, f, tol))
        if exact:
            ex = diff_exact(ref_dir, ours_dir, f)
            if not ex:
                ok, text = False, text + "\n exact: no exact-mode files"
            else:
                ok = ok and all(e.ok for e in ex)
                text += "\n exact:\n" + "\n".join(e.line() for e in ex)
            text = text.replace("[PASS]", "[FAIL]", 1) if not ok else text
        texts.append(text)
        all_ok = all_ok and ok
    return all_ok, "\n".join(texts)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("ref_dir")
    ap.add_argument("ours_dir")
    ap.add_argument("--frames", type=int, nargs="*", help="default: every frame present in both folders")
    ap.add_argument("--tol", nargs="*", default=[], help="key=value (position, volume, phase, cell_count_abs, "
                                                          "voxel_coord, substrate_linf, substrate_rmse)")
    ap.add_argument("--exact", action="store_true",
                    help="also require the exact-mode files (keys, bins, sorted list, neighbour lists) to be equal")
    a = ap.parse_args(argv)
    tol = {k: float(v) for k, v in (t.split("=", 1) for t in a.tol)}
    ok, text = compare(a.ref_dir, a.ours_dir, a.frames, tol, a.exact)
    print(text)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
