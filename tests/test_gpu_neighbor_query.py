"""Neighbour queries on the device hash (pc_sim_neighbor_query /
hc_neighbor_query) against the oracle on the same positions: the contact
graph (or_contact_graph, the reference's O(N^2) loop restated) and the
interaction candidates (or_interaction_candidates), both bit for bit as CSR
lists, FP64 and mixed hashes, and the reach check."""
import numpy as np
import pytest

import oracle as O
from physicell_amd import hip as H, scenarios as S

pytestmark = pytest.mark.gpu


def sim_arrays(sim):
    g = {k: sim.field(f) for k, f in (("x", H.F_POS_X), ("y", H.F_POS_Y), ("z", H.F_POS_Z), ("radius", H.F_RADIUS),
                                       ("volume", H.F_TOTAL_VOL), ("rel", H.F_REL_MAX_ADH))}
    g["alive"] = sim.field(H.F_IS_ALIVE)
    return g


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("case", ["small3d", "config1", "config0"])
def test_queries_match_oracle(case, mixed):
    if case == "small3d":
        spec, cells = S.small3d(30, 24, 20, n_cells=3000, seed=71)
    elif case == "config0":
        spec, cells = S.config0()            # 2-D: 9 bins around each cell
    else:
        spec, cells = S.config1()
    r = np.random.default_rng(72)
    cells.alive = (r.uniform(size=cells.n) < 0.95).astype(np.uint32)
    sim = H.Simulation(spec, cells, mixed=mixed)
    sim.run_periods(2)
    sim.wait()
    print(f"neighbour queries: {case} n={cells.n} mixed={mixed}", flush=True)
    off, idx = sim.neighbor_query(H.QUERY_CONTACT)
    a = sim_arrays(sim)
    woff, widx = O.contact_graph(a["x"], a["y"], a["z"], a["radius"], a["volume"], a["rel"], a["alive"])
    assert np.array_equal(off, woff) and np.array_equal(idx, widx)
    assert len(idx) > cells.n // 4 or case == "config0"
    off, idx = sim.neighbor_query(H.QUERY_INTERACT)
    keys = sim.field(H.F_MECH_VOXEL)
    v = spec.mech_voxel
    m = O.OrMech(spec.mnx, spec.mny, spec.mnz, 0, v, v, spec.x_min, spec.y_min, spec.z_min)
    assert np.array_equal(keys, O.mech_keys(a["x"], a["y"], a["z"], m))
    boff, bsrt = O.hash_sort(keys, spec.mnx * spec.mny * spec.mnz)
    woff, widx = O.interaction_candidates(a["x"], a["y"], a["radius"], a["volume"], keys, boff, bsrt, m)
    assert np.array_equal(off, woff) and np.array_equal(idx, widx)
    sim.close()


def test_contact_reach_check():
    spec, cells = S.small3d(12, 10, 8, n_cells=200, seed=73)
    cells.radius = np.full(cells.n, 14.0)           # 1.25 x 28 um > one 30 um voxel
    sim = H.Simulation(spec, cells)
    with pytest.raises(H.HipError, match="exceeds the mechanics voxel"):
        sim.neighbor_query(H.QUERY_CONTACT)
    off, idx = sim.neighbor_query(H.QUERY_INTERACT)    # the reference's own scan has no such limit
    assert off[-1] == len(idx)
    sim.close()


def test_empty_population():
    """No cells: empty lists, and the phenotype step and queries accept it."""
    spec, cells = S.small3d(12, 10, 8, n_cells=0, seed=74)
    sim = H.Simulation(spec, cells, max_cells=16)
    for kind in (H.QUERY_CONTACT, H.QUERY_INTERACT, H.QUERY_ATTACH):
        off, idx = sim.neighbor_query(kind)
        assert off.tolist() == [0] and len(idx) == 0
    cnt, tg, formed = sim.form_attachments(1.0, np.zeros(0), seed=1)
    assert formed == 0 and len(cnt) == 0
    sim.set_volume(seed=1)
    assert sim.update_phenotype(6.0) == 0
    sim.run_periods(1)
    sim.wait()
    sim.close()


def test_attachment_input_checks():
    """Lists longer than 12 partners and rates of the wrong length are
    rejected before any work."""
    spec, cells = S.small3d(12, 10, 8, n_cells=50, seed=75)
    sim = H.Simulation(spec, cells)
    bad = np.zeros(cells.n, np.uint32)
    bad[3] = 13
    with pytest.raises(H.HipError, match="exceeds 12"):
        sim.form_attachments(1.0, np.full(cells.n, 0.1), seed=1, count=bad)
    with pytest.raises(ValueError):
        sim.form_attachments(1.0, np.full(cells.n - 1, 0.1), seed=1)
    cnt, tg, formed = sim.form_attachments(1.0, np.full(cells.n, 0.1), seed=1)
    assert cnt.max(initial=0) <= H.MAX_ATTACHMENTS
    sim.close()


def attach_candidates_bruteforce(a):
    """All pairs j > i, both alive, sqrt(|p_j - p_i|^2) <= max(rel) (r_i + r_j)
    (tryFormAttachment's test, src/cell_interactions.cpp:429-440)."""
    x, y, z = a["x"], a["y"], a["z"]
    n = len(x)
    off = np.zeros(n + 1, np.uint32)
    idx = []
    for i in range(n):
        off[i] = len(idx)
        if not a["alive"][i]:
            continue
        j = np.arange(i + 1, n)
        d = np.sqrt((x[j] - x[i]) ** 2 + (y[j] - y[i]) ** 2 + (z[j] - z[i]) ** 2)
        md = np.maximum(a["rel"][i], a["rel"][j]) * (a["radius"][i] + a["radius"][j])
        idx.extend(j[(a["alive"][j] == 1) & ~(d > md)].tolist())
    off[n] = len(idx)
    return off, np.array(idx, np.uint32)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("case", ["small3d", "config0", "dense"])
def test_attachments_match_oracle(case, mixed):
    """HC_QUERY_ATTACH against the all-pairs test, and pc_sim_form_attachments
    (device candidates + host greedy pass) against or_form_attachments (the
    reference's O(N^2) loop restated, same std::mt19937 draws), bit for bit,
    over two calls (the second starts from the first's lists)."""
    if case == "small3d":
        spec, cells = S.small3d(30, 24, 20, n_cells=2500, seed=81)
    elif case == "config0":
        spec, cells = S.config0()
    else:
        spec, cells = S.small3d(12, 10, 8, n_cells=1500, seed=82)    # lists fill to 12
    r = np.random.default_rng(83)
    cells.alive = (r.uniform(size=cells.n) < 0.93).astype(np.uint32)
    sim = H.Simulation(spec, cells, mixed=mixed)
    sim.run_periods(2)
    sim.wait()
    a = sim_arrays(sim)
    print(f"attachments: {case} n={cells.n} mixed={mixed}", flush=True)
    off, idx = sim.neighbor_query(H.QUERY_ATTACH)
    woff, widx = attach_candidates_bruteforce(a)
    assert np.array_equal(off, woff) and np.array_equal(idx, widx)
    rate = r.uniform(0.0, 0.2, cells.n)
    rate[r.uniform(size=cells.n) < 0.1] = 0.0
    if case == "dense":
        rate[:] = np.where(rate > 0, 1.0, 0.0)          # every draw accepts
    cnt, tg, formed = sim.form_attachments(6.0 if case == "dense" else 1.0, rate, seed=5)
    wc, wt, wf = O.form_attachments(a["x"], a["y"], a["z"], a["radius"], a["rel"], a["alive"], rate,
                                    6.0 if case == "dense" else 1.0, 5)
    assert formed == wf > 0 and np.array_equal(cnt, wc) and np.array_equal(tg, wt)
    if case == "dense":
        assert cnt.max() == H.MAX_ATTACHMENTS
    cnt2, tg2, f2 = sim.form_attachments(1.0, rate, seed=6, count=cnt, targets=tg)
    wc2, wt2, wf2 = O.form_attachments(a["x"], a["y"], a["z"], a["radius"], a["rel"], a["alive"], rate, 1.0, 6,
                                       count=wc, targets=wt)
    assert f2 == wf2 and np.array_equal(cnt2, wc2) and np.array_equal(tg2, wt2)
    sim.close()
