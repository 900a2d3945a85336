"""Neighbour queries for the phenotype side: the oracle's contact graph
(or_contact_graph) against the reference's own writeCellNeighborGraph
(src/output_writer.cpp:457-490, compiled from the reference sources into
oracle/_ref) byte for byte through the library's writer, and against a
numpy all-pairs computation; the oracle's interaction candidates
(or_interaction_candidates) against an independent numpy formulation of the
reference's scan (bins from the positions, not from a hash).  Host only."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H

REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_path.so")


def ref_lib():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    lib = ctypes.CDLL(REF)
    vp = ctypes.c_void_p
    lib.ref_write_neighbor_graph.argtypes = [ctypes.c_char_p, ctypes.c_uint32] + [vp] * 7
    return lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def population(n, seed, box=200.0):
    r = np.random.default_rng(seed)
    c = {"x": f32(r.uniform(-box, box, n)), "y": f32(r.uniform(-box, box, n)), "z": f32(r.uniform(-box / 2, box / 2, n)),
         "radius": f32(r.uniform(5, 10, n)), "volume": f32(r.uniform(1000, 4000, n)),
         "rel": f32(r.uniform(1.0, 1.4, n)), "alive": (r.uniform(size=n) < 0.9).astype(np.uint32)}
    c["volume"][r.uniform(size=n) < 0.03] = 0.0
    return c


def numpy_contacts(c):
    x, y, z = c["x"], c["y"], c["z"]
    d2 = (x[:, None] - x[None]) ** 2 + (y[:, None] - y[None]) ** 2 + (z[:, None] - z[None]) ** 2
    md = np.maximum(c["rel"][:, None], c["rel"][None]) * (c["radius"][:, None] + c["radius"][None])
    ok = (c["alive"] == 1) & (c["volume"] > 1e-15)
    m = (d2 <= md * md) & ok[:, None] & ok[None] & np.triu(np.ones_like(d2, bool), 1)
    return [tuple(v) for v in np.argwhere(m)]


def test_contact_graph_matches_reference_file(tmp_path):
    lib = H.load()
    ref = ref_lib()
    for seed, n, box in ((1, 1500, 200.0), (2, 300, 70.0)):
        c = population(n, seed, box)
        off, idx = O.contact_graph(c["x"], c["y"], c["z"], c["radius"], c["volume"], c["rel"], c["alive"])
        pairs = [(i, int(j)) for i in range(n) for j in idx[off[i]:off[i + 1]]]
        assert pairs == numpy_contacts(c)
        assert len(pairs) > n // 2
        ours = tmp_path / f"ours{seed}.txt"
        assert lib.pc_write_neighbor_graph(str(ours).encode(), n, p(off), p(idx)) == 0
        d = tmp_path / f"ref{seed}"
        d.mkdir()
        a32 = [np.ascontiguousarray(c[k], np.float32) for k in ("x", "y", "z", "radius", "volume", "rel")]
        ref.ref_write_neighbor_graph(str(d).encode(), n, *[p(a) for a in a32], p(c["alive"]))
        assert ours.read_bytes() == (d / "final_cell_neighbor_graph.txt").read_bytes()


def test_interaction_candidates_independent_formulation():
    n = 2000
    c = population(n, 3, box=150.0)
    vox = 30.0
    m = O.OrMech(10, 10, 5, 0, vox, vox, -150.0, -150.0, -75.0)
    keys = O.mech_keys(c["x"], c["y"], c["z"], m)
    off, srt = O.hash_sort(keys, 10 * 10 * 5)
    offs, idx = O.interaction_candidates(c["x"], c["y"], c["radius"], c["volume"], keys, off, srt, m)
    # independent: bins from the positions, the 3x3x3 neighbourhood test per
    # pair, then the reference's filters, ordered by (dk, dj, di, j)
    b = np.stack([np.clip(((c[a] - lo) / vox).astype(int), 0, hi - 1)
                  for a, lo, hi in (("x", -150.0, 10), ("y", -150.0, 10), ("z", -75.0, 5))])
    for i in range(0, n, 7):
        d = b - b[:, i:i + 1]
        near = np.all(np.abs(d) <= 1, axis=0)
        near[i] = False
        cand = np.flatnonzero(near)
        order = np.lexsort((cand, d[0, cand], d[1, cand], d[2, cand]))
        cand = cand[order][:512]
        dx, dy = c["x"][cand] - c["x"][i], c["y"][cand] - c["y"][i]
        d2 = dx * dx + dy * dy
        rj = c["radius"][cand]
        keep = (c["volume"][cand] >= 1e-15) & ~(d2 > (2.5 * c["radius"][i] + rj) ** 2) & \
               ~(np.sqrt(d2) > (c["radius"][i] + rj) * 1.25)
        assert idx[offs[i]:offs[i + 1]].tolist() == cand[keep].tolist(), i


def test_interaction_candidates_cap():
    """512 candidates at most are visited (before the distance filters)."""
    n = 800
    r = np.random.default_rng(4)
    x, y, z = (r.uniform(0, 10, n) for _ in range(3))
    m = O.OrMech(4, 4, 4, 0, 30.0, 30.0, 0.0, 0.0, 0.0)
    keys = O.mech_keys(x, y, z, m)
    off, srt = O.hash_sort(keys, 64)
    offs, idx = O.interaction_candidates(x, y, np.full(n, 8.0), np.full(n, 2000.0), keys, off, srt, m)
    counts = np.diff(offs)
    assert counts.max() == 512 and counts.min() == 512
    assert idx[offs[5]:offs[6]].tolist() == [j for j in range(n) if j != 5][:512]
