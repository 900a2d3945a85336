"""physicell_amd.snapdiff: the parity harness's comparison semantics
(parity/diff.py:150-300) on hand-built snapshot pairs whose answers are known:
ID matching under reordering, the absolute tolerances and their defaults,
the integer phase, the count rule, the voxel-alignment gate and the
substrate L-infinity / RMSE pair.  Host only."""
import numpy as np

from physicell_amd import matio
from physicell_amd import snapdiff as D


def cells_matrix(n, rows=40, seed=0):
    r = np.random.default_rng(seed)
    m = r.uniform(-100, 100, (rows, n))
    m[matio.CELL_ROW_ID] = np.arange(n)
    m[matio.CELL_ROW_TOTAL_VOLUME] = r.uniform(2000, 3000, n)
    m[matio.CELL_ROW_PHASE] = r.integers(0, 4, n)
    return m


def micro_matrix(ns=2, n=60, seed=1):
    r = np.random.default_rng(seed)
    m = np.zeros((4 + ns, n))
    m[0], m[1], m[2] = np.arange(n) * 20.0 - 590, 10.0, 0.0
    m[3] = 8000.0
    m[4:] = r.uniform(0, 40, (ns, n))
    return m


def write(tmp_path, side, frame, cells=None, micro=None):
    d = tmp_path / side
    d.mkdir(exist_ok=True)
    if cells is not None:
        matio.write_mat(str(d / f"output{frame:08d}_cells.mat"), "cells", cells)
    if micro is not None:
        matio.write_mat(str(d / f"output{frame:08d}_microenvironment0.mat"), "multiscale_microenvironment", micro)
    return str(d)


def test_identical_and_reordered(tmp_path):
    c, me = cells_matrix(50), micro_matrix()
    a = write(tmp_path, "ref", 0, c, me)
    perm = np.random.default_rng(3).permutation(50)
    b = write(tmp_path, "ours", 0, c[:, perm], me)
    cells, micro = D.diff_frame(a, b, 0)
    assert cells.ok and micro.ok and cells.matched == 50
    assert all(f.max_abs == 0 and f.rel == 0 for f in cells.fields + micro.fields)
    assert [f.name for f in cells.fields] == ["pos_x", "pos_y", "pos_z", "volume", "cycle_phase"]
    assert D.main([a, b]) == 0


def test_position_tolerance_and_stats(tmp_path):
    c = cells_matrix(40)
    o = c.copy()
    o[matio.CELL_ROW_Y, 7] += 2e-6                 # default position tolerance 1e-6
    a, b = write(tmp_path, "ref", 1, c), write(tmp_path, "ours", 1, o)
    cells, micro = D.diff_frame(a, b, 1)
    assert micro is None and not cells.ok
    y = cells.fields[1]
    assert not y.ok and np.isclose(y.max_abs, 2e-6, rtol=1e-6)
    assert np.isclose(y.rmse, 2e-6 / np.sqrt(40), rtol=1e-6) and np.isclose(y.mean_abs, 2e-6 / 40, rtol=1e-6)
    assert cells.fields[0].ok and cells.fields[2].ok
    assert D.diff_frame(a, b, 1, {"position": 1e-5})[0].ok
    assert D.main([a, b]) == 1


def test_phase_is_exact_and_volume_tolerance(tmp_path):
    c = cells_matrix(30)
    o = c.copy()
    o[matio.CELL_ROW_TOTAL_VOLUME, 3] += 9e-4       # within the default 1e-3
    a, b = write(tmp_path, "ref", 2, c), write(tmp_path, "ours", 2, o)
    assert D.diff_frame(a, b, 2)[0].ok
    o[matio.CELL_ROW_PHASE, 5] += 1
    b = write(tmp_path, "ours", 2, o)
    cells = D.diff_frame(a, b, 2)[0]
    assert not cells.ok and not cells.fields[-1].ok and cells.fields[-1].name == "cycle_phase"
    # without row 7 on one side, the phase is not compared
    b = write(tmp_path, "ours", 2, o[:7])
    cells = D.diff_frame(a, b, 2)[0]
    assert cells.ok and len(cells.fields) == 4 and any("row schema" in n for n in cells.notes)


def test_cell_count_rule(tmp_path):
    c = cells_matrix(30)
    a, b = write(tmp_path, "ref", 3, c), write(tmp_path, "ours", 3, c[:, :27])
    cells = D.diff_frame(a, b, 3)[0]
    assert cells.ok and cells.only_ref == 3 and cells.matched == 27
    assert not D.diff_frame(a, b, 3, {"cell_count_abs": 2})[0].ok
    assert D.diff_frame(a, b, 3, {"cell_count_abs": 3})[0].ok
    o = c.copy()
    o[matio.CELL_ROW_ID] += 1000                   # no shared ID
    b = write(tmp_path, "ours", 3, o)
    assert not D.diff_frame(a, b, 3)[0].ok


def test_microenvironment_gates(tmp_path):
    me = micro_matrix()
    o = me.copy()
    o[5] += 5e-5                                     # constant offset: Linf = RMSE = 5e-5
    a, b = write(tmp_path, "ref", 4, micro=me), write(tmp_path, "ours", 4, micro=o)
    micro = D.diff_frame(a, b, 4)[1]
    s1 = micro.fields[1]
    assert micro.ok and np.isclose(s1.max_abs, 5e-5) and np.isclose(s1.rmse, 5e-5)
    assert np.isclose(s1.rel, 5e-5 / np.abs(me[5]).max())
    o[5, 0] += 5e-4                                  # Linf 5.5e-4 < 1e-3 and RMSE 8.7e-5 < 1e-4
    micro = D.diff_frame(a, write(tmp_path, "ours", 4, micro=o), 4)[1]
    assert micro.ok
    o[5] += 2e-4                                     # RMSE > 1e-4 with Linf < 1e-3
    micro = D.diff_frame(a, write(tmp_path, "ours", 4, micro=o), 4)[1]
    assert not micro.ok and micro.fields[1].max_abs < 1e-3
    o = me.copy()
    o[0, 10] += 1e-3                                 # a voxel centre moved: no substrate comparison
    micro = D.diff_frame(a, write(tmp_path, "ours", 4, micro=o), 4)[1]
    assert not micro.ok and not micro.fields
    micro = D.diff_frame(a, write(tmp_path, "ours", 4, micro=me[:, :59]), 4)[1]
    assert not micro.ok and "voxel count" in micro.notes[0]


def exact_state(n=50, nb=12, seed=3):
    """Hand-built integer outputs: keys, bins of a counting sort, and CSR
    neighbour lists with 0-4 entries per cell."""
    r = np.random.default_rng(seed)
    mech = r.integers(0, nb, n)
    off = np.concatenate([[0], np.cumsum(np.bincount(mech, minlength=nb))])
    srt = np.argsort(mech, kind="stable")
    cnt = r.integers(0, 5, n)
    noff = np.concatenate([[0], np.cumsum(cnt)])
    nidx = r.integers(0, n, int(noff[-1]))
    return dict(cell_ids=np.arange(n), voxel_keys=r.integers(0, 1000, n), mech_keys=mech, offsets=off,
                sorted_cells=srt, nbr_offsets=noff, nbr_idx=nidx)


def test_exact_mode(tmp_path):
    """The exact-mode files: equal outputs pass; one changed key, bin, sorted
    entry or neighbour fails and is named; the same neighbour sets in another
    visiting order fail as an order difference; a file on one side only
    fails.  The command line (parity/diff.py) returns the status."""
    import os
    import subprocess
    import sys
    st = exact_state()
    ref = write(tmp_path, "ref", 2, cells_matrix(50), micro_matrix())
    D.write_exact(ref, 2, **st)
    for what, change in (("keys", lambda s: s["voxel_keys"].__setitem__(7, s["voxel_keys"][7] + 1)),
                         ("bins", lambda s: s["offsets"].__setitem__(3, s["offsets"][3] + 1)),
                         ("sorted", lambda s: s["sorted_cells"].__setitem__(slice(0, 2), s["sorted_cells"][1::-1])),
                         ("neighbors", lambda s: s["nbr_idx"].__setitem__(0, (s["nbr_idx"][0] + 1) % 50)),
                         (None, None)):
        ours = write(tmp_path, f"ours_{what}", 2, cells_matrix(50), micro_matrix())
        s2 = {k: np.array(v) for k, v in st.items()}
        if change:
            change(s2)
        D.write_exact(ours, 2, **s2)
        ok, text = D.compare(ref, ours, exact=True)
        stats = {e.name: e for e in D.diff_exact(ref, ours, 2)}
        if what is None:
            assert ok and all(e.ok for e in stats.values()), text
        else:
            assert not ok and not stats[what].ok, text
            assert all(e.ok for k, e in stats.items() if k != what), text
    # same sets, another visiting order
    ours = write(tmp_path, "ours_order", 2, cells_matrix(50), micro_matrix())
    s2 = {k: np.array(v) for k, v in st.items()}
    i = int(np.argmax(np.diff(s2["nbr_offsets"]) >= 2))
    a = s2["nbr_offsets"][i]
    s2["nbr_idx"][a], s2["nbr_idx"][a + 1] = s2["nbr_idx"][a + 1], s2["nbr_idx"][a]
    if s2["nbr_idx"][a] != s2["nbr_idx"][a + 1]:
        D.write_exact(ours, 2, **s2)
        e = {x.name: x for x in D.diff_exact(ref, ours, 2)}["neighbors"]
        assert not e.ok and "visiting order" in e.note
    # a file on one side only
    ours = write(tmp_path, "ours_missing", 2, cells_matrix(50), micro_matrix())
    D.write_exact(ours, 2, **st)
    os.remove(os.path.join(ours, "output00000002_bins.mat"))
    ok, text = D.compare(ref, ours, exact=True)
    assert not ok and "present only in ref" in text
    # the command line
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    good = write(tmp_path, "ours_None", 2)
    cli = [sys.executable, os.path.join(root, "parity", "diff.py"), ref]
    assert subprocess.run(cli + [good, "--exact"], capture_output=True).returncode == 0
    assert subprocess.run(cli + [ours, "--exact"], capture_output=True).returncode == 1
