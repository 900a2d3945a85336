"""Packed ensembles (pe_*): each member's results equal the single-simulation
path's (pc_sim) on the same inputs, bit for bit, alone and in any mix, across
membership changes, graph replay and autotuning."""
import math

import numpy as np
import pytest

from physicell_amd import hip, scenarios as S

pytestmark = pytest.mark.gpu

POS = (hip.F_POS_X, hip.F_POS_Y, hip.F_POS_Z)
VEL = (hip.F_VEL_X, hip.F_VEL_Y, hip.F_VEL_Z)
VOLUMES = (hip.F_RADIUS, hip.F_NUCLEAR_RAD, hip.F_TOTAL_VOL, hip.F_NUCLEAR_VOL)


def small_member(seed, dt_pheno=0.5, **kw):
    kw.setdefault("edges", (18, 40))
    kw.setdefault("cells", (100, 3000))
    sc, c = S.ensemble_member(seed, **kw)
    sc.dt_pheno = dt_pheno
    return sc, c


def substeps_until(dtd, dt):
    acc = 0.0
    for k in range(1, 1000000):
        acc += dtd
        if acc >= dt - 0.5 * dtd:
            return k
    return 1000000


def run_single(sc, c, periods, mixed, start=0, motility=None):
    """pc_sim driven through the same schedule: per period k substeps + one
    mechanics step, the phenotype step where the member's own clock fires.
    motility: (seed, keyword arrays) for Simulation.set_motility."""
    sim = hip.Simulation(sc, c, mixed=mixed)
    if motility is not None:
        sim.set_motility(motility[0], **motility[1])
    k = substeps_until(sc.dt_diff, sc.dt_mech)
    kp = substeps_until(sc.dt_diff, sc.dt_pheno)
    for p in range(periods):
        for _ in range(k):
            sim.substep()
        sim.mech_step()
        if ((p + 1) * k) % kp == 0:
            sim.pheno_step()
    sim.wait()
    out = {"rho": sim.density()}
    for f in POS + VEL:
        out[f] = sim.field(f)
    sim.close()
    return out


def member_state(ens, mid):
    out = {"rho": ens.density(mid)}
    for f in POS + VEL:
        out[f] = ens.field(mid, f)
    return out


def assert_same(a, b, what):
    bad = []
    for key in a:
        if not np.array_equal(a[key], b[key]):
            diff = np.abs(a[key] - b[key])
            bad.append(f"{key} (max abs {np.max(diff):.3e} at {int(np.argmax(diff))}, {int(np.count_nonzero(diff))} values)")
    if bad:
        raise AssertionError(f"{what}: " + "; ".join(bad) + " differ")


@pytest.mark.parametrize("mixed", [False, True])
def test_member_equals_single(mixed):
    members = [small_member(200), small_member(201, dt_pheno=1.0), small_member(202)]
    P = 12
    ens = hip.Ensemble(mixed=mixed)
    ids = [ens.add(sc, c) for sc, c in members]
    ens.run_periods(P, use_graph=True)
    ens.wait()
    for mid, (sc, c) in zip(ids, members):
        print(f"member {mid}: grid {sc.nx}x{sc.ny}x{sc.nz} cells {c.n} mixed={mixed}", flush=True)
        assert_same(member_state(ens, mid), run_single(sc, c, P, mixed), f"member {mid}")
        b, counts = ens.member_info(mid)
        k = substeps_until(sc.dt_diff, sc.dt_mech)
        kp = substeps_until(sc.dt_diff, sc.dt_pheno)
        assert counts == (P * k, P, (P * k) // kp)
        assert b > 0
    ens.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_crowded_member_equals_single(mixed):
    """A member with bins of more than 16 agents (the hash ranks their agents
    one by one) beside ordinary members."""
    crowded = S.small3d(12, 10, 8, n_cells=4000, seed=11)
    members = [small_member(230), crowded, small_member(231)]
    P = 4
    ens = hip.Ensemble(mixed=mixed)
    ids = [ens.add(sc, c) for sc, c in members]
    ens.run_periods(P, use_graph=False)
    ens.wait()
    for mid, (sc, c) in zip(ids, members):
        assert_same(member_state(ens, mid), run_single(sc, c, P, mixed), f"member {mid}")
    ens.close()


def motility_arrays(n, ns, seed, sensitivities):
    r = np.random.default_rng(seed)
    kw = dict(persistence=r.uniform(0.05, 2.0, n), bias=r.uniform(0.0, 1.0, n), speed=r.uniform(0.5, 2.0, n),
              is_motile=(r.uniform(size=n) < 0.8).astype(np.uint32), chem_index=r.integers(0, ns, n),
              chem_dir=np.where(r.uniform(size=n) < 0.5, 1, -1))
    if sensitivities:
        kw["sensitivity"] = r.uniform(-1.0, 1.0, n * ns)
    return kw


@pytest.mark.parametrize("mixed", [False, True])
def test_member_motility_equals_single(mixed):
    """pe_set_motility: members with motility (chemotaxis by substrate index on
    one, by sensitivities on another) beside a member without, over graph
    replays (fresh draws every mechanics step), bit for bit against pc_sim
    with the same motility."""
    members = [small_member(240), small_member(241), small_member(242)]
    P = 6
    ens = hip.Ensemble(mixed=mixed)
    ids = [ens.add(sc, c) for sc, c in members]
    mot = {}
    for j in (0, 2):
        sc, c = members[j]
        mot[j] = (1000 + j, motility_arrays(c.n, len(sc.substrates), 300 + j, sensitivities=(j == 2)))
        ens.set_motility(ids[j], mot[j][0], **mot[j][1])
    ens.run_periods(P, use_graph=True)
    ens.wait()
    for j, (mid, (sc, c)) in enumerate(zip(ids, members)):
        got = member_state(ens, mid)
        want = run_single(sc, c, P, mixed, motility=mot.get(j))
        print(f"member {mid}: cells {c.n} motility={j in mot} mixed={mixed}", flush=True)
        assert_same(got, want, f"member {mid}")
    # motility moves the cells: the member without it differs from one with it
    assert not np.array_equal(member_state(ens, ids[0])[hip.F_POS_X],
                              run_single(*members[0], P, mixed)[hip.F_POS_X])
    ens.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_member_phenotype_equals_single(mixed):
    """pe_set_volume / pe_update_phenotype: the volume model, divisions (into
    room made by pe_set_volume) and removals of one member, with motility on
    it, between periods; bit for bit against pc_sim driven the same way."""
    members = [small_member(250), small_member(251), small_member(252)]
    sc, c = members[1]
    n, ns = c.n, len(sc.substrates)
    M = n + 64
    mot = (77, motility_arrays(n, ns, 350, sensitivities=False))
    r = np.random.default_rng(450)
    fields = {hip.V_TARGET_SOLID_NUC: r.uniform(100, 200, n), hip.V_FLUID: r.uniform(1500, 2500, n)}
    mask = (r.uniform(size=n) < 0.7).astype(np.uint32)
    parents = r.choice(n, 40, replace=False)
    removed = r.choice(n + 40, 25, replace=False)
    k = substeps_until(sc.dt_diff, sc.dt_mech)
    kp = substeps_until(sc.dt_diff, sc.dt_pheno)
    P1, P2 = 3, 3

    ens = hip.Ensemble(mixed=mixed)
    ids = [ens.add(s_, c_) for s_, c_ in members]
    ens.set_motility(ids[1], mot[0], **mot[1])
    ens.set_volume(ids[1], 9, fields, max_cells=M)
    ens.run_periods(P1, use_graph=True)
    n_ens = ens.update_phenotype(ids[1], 0.1, mask=mask, parents=parents, removed=removed)
    ens.run_periods(P2, use_graph=True)
    ens.wait()

    sim = hip.Simulation(sc, c, max_cells=M, mixed=mixed)
    sim.set_motility(mot[0], **mot[1])
    sim.set_volume(9, fields)
    for p in range(P1 + P2):
        if p == P1:
            n_single = sim.update_phenotype(0.1, mask=mask, parents=parents, removed=removed)
        for _ in range(k):
            sim.substep()
        sim.mech_step()
        if ((p + 1) * k) % kp == 0:
            sim.pheno_step()
    sim.wait()
    want = {"rho": sim.density()}
    for f in POS + VEL + VOLUMES:
        want[f] = sim.field(f)
    sim.close()
    print(f"phenotype member: n {n} -> {n_ens} (single {n_single}), mixed={mixed}", flush=True)
    assert n_ens == n_single == n + 40 - 25
    got = member_state(ens, ids[1])
    for f in VOLUMES:
        got[f] = ens.field(ids[1], f)
    assert_same(got, want, "phenotype member")
    # the other members are untouched by the resize and the phenotype step
    for j in (0, 2):
        assert_same(member_state(ens, ids[j]), run_single(*members[j], P1 + P2, mixed), f"member {ids[j]}")
    ens.close()


def test_alone_and_in_mix_mixed():
    x = small_member(210)
    others = [small_member(s) for s in (211, 212, 213)]
    P = 10
    alone = hip.Ensemble(mixed=True)
    a = alone.add(*x)
    alone.run_periods(P)
    ra = member_state(alone, a)
    alone.close()
    mix = hip.Ensemble(mixed=True)
    mix.add(*others[0])
    b = mix.add(*x)
    for o in others[1:]:
        mix.add(*o)
    mix.run_periods(P)
    rb = member_state(mix, b)
    mix.close()
    assert_same(ra, rb, "alone vs mix")


def test_membership_changes():
    """Continuous batching: members join and leave between periods."""
    A, B, C = small_member(220), small_member(221), small_member(222)
    ens = hip.Ensemble(mixed=False)
    ia = ens.add(*A)
    ib = ens.add(*B)
    ens.run_periods(3)
    ic = ens.add(*C)
    ens.run_periods(3, use_graph=False)
    ens.remove(ia)
    ens.run_periods(2)
    ens.wait()
    assert ens.info()["members"] == 2
    assert_same(member_state(ens, ib), run_single(*B, 8, False), "B (8 periods)")
    assert_same(member_state(ens, ic), run_single(*C, 5, False), "C (5 periods)")
    with pytest.raises(hip.HipError, match="unknown member"):
        ens.remove(ia)
    ens.close()


def test_autotune_keeps_results():
    members = [small_member(s) for s in (230, 231)]
    ref = hip.Ensemble(mixed=True)
    ids = [ref.add(*m) for m in members]
    ref.run_periods(4)
    r0 = [member_state(ref, i) for i in ids]
    ref.close()
    ens = hip.Ensemble(mixed=True)
    ids = [ens.add(*m) for m in members]
    ens.run_periods(2)
    blocks, us = ens.autotune(reps=2)
    print("autotune", blocks, us, flush=True)
    assert all(b == 0 or b >= 256 for b in blocks)    # 0: one workgroup per work item
    ens.run_periods(2)
    for i, r in zip(ids, r0):
        assert_same(member_state(ens, i), r, "after autotune")
    ens.close()


def test_schedule_and_grid_errors():
    ens = hip.Ensemble(mixed=False)
    sc, c = small_member(240)
    ens.add(sc, c)
    sc2, c2 = small_member(241)
    sc2.dt_mech = 0.2
    with pytest.raises(hip.HipError, match="share dt_diffusion"):
        ens.add(sc2, c2)
    sc3, c3 = S.small3d(nx=12, ny=10, nz=4)
    with pytest.raises(hip.HipError, match="grid edges"):
        ens.add(sc3, c3)
    sc4, c4 = small_member(242)
    sc4.dt_pheno = 0.15
    with pytest.raises(hip.HipError, match="whole number"):
        ens.add(sc4, c4)
    ens.close()


def test_mixed_ensemble_error_vs_fp64():
    members = [small_member(s, edges=(24, 48), cells=(500, 4000)) for s in (250, 251)]
    P = 20
    r = {}
    for mixed in (False, True):
        ens = hip.Ensemble(mixed=mixed)
        ids = [ens.add(*m) for m in members]
        ens.run_periods(P)
        r[mixed] = [member_state(ens, i) for i in ids]
        ens.close()
    for a, b in zip(r[False], r[True]):
        rel = np.max(np.abs(a["rho"] - b["rho"]) / np.maximum(np.abs(a["rho"]), 1e-3))
        pos = max(np.max(np.abs(a[f] - b[f])) for f in POS)
        print(f"mixed vs FP64: density max rel {rel:.3e}, positions max abs {pos:.3e} um", flush=True)
        assert rel <= 1e-4
        assert pos <= 1e-3


def single_after(sc, c, periods, mixed):
    """A pc_sim driven through the member schedule, left open for queries."""
    sim = hip.Simulation(sc, c, mixed=mixed)
    k = substeps_until(sc.dt_diff, sc.dt_mech)
    kp = substeps_until(sc.dt_diff, sc.dt_pheno)
    for p in range(periods):
        for _ in range(k):
            sim.substep()
        sim.mech_step()
        if ((p + 1) * k) % kp == 0:
            sim.pheno_step()
    sim.wait()
    return sim


@pytest.mark.parametrize("mixed", [False, True])
def test_member_queries_equal_single(mixed):
    """pe_neighbor_query (contact, interaction, attachment lists) and
    pe_form_attachments on one member equal pc_sim's on the same schedule,
    bit for bit, beside other members."""
    members = [small_member(270, cells=(1500, 3000)), small_member(271), small_member(272, cells=(1500, 3000))]
    P = 4
    ens = hip.Ensemble(mixed=mixed)
    ids = [ens.add(s_, c_) for s_, c_ in members]
    ens.run_periods(P, use_graph=True)
    ens.wait()
    for j in (0, 2):
        sc, c = members[j]
        sim = single_after(sc, c, P, mixed)
        print(f"member queries: member {j} n={c.n} mixed={mixed}", flush=True)
        for kind in (hip.QUERY_CONTACT, hip.QUERY_INTERACT, hip.QUERY_ATTACH):
            mo, mi = ens.neighbor_query(ids[j], kind)
            so, si = sim.neighbor_query(kind)
            assert np.array_equal(mo, so) and np.array_equal(mi, si), kind
            assert len(mi) > 0
        rate = np.random.default_rng(273 + j).uniform(0.0, 0.3, c.n)
        got = ens.form_attachments(ids[j], 1.0, rate, seed=9)
        want = sim.form_attachments(1.0, rate, seed=9)
        assert got[2] == want[2] > 0 and np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        sim.close()
    # the queries leave the members' steps unchanged
    ens.run_periods(1, use_graph=True)
    ens.wait()
    for j in (0, 1):
        assert_same(member_state(ens, ids[j]), run_single(*members[j], P + 1, mixed), f"member {j} after queries")
    ens.close()


def box_member(nx, ny, nz, n_cells, seed):
    """A member on an nx x ny x nz grid of 20 um voxels with uniform random agents."""
    dx = 20.0
    subs = [S.Substrate("oxygen", 1e5, 0.1, 38.0, True, 38.0), S.Substrate("factor", 1000.0, 0.01, 0.0, False, 0.0)]
    sc = S.Scenario(f"box{nx}x{ny}x{nz}", 0.0, nx * dx, 0.0, ny * dx, 0.0, nz * dx, dx, dx, dx, False, subs,
                    max_time=10.0)
    sc.dt_pheno = 0.5
    p = np.random.default_rng(seed).random((n_cells, 3)) * np.array([nx * dx, ny * dx, nz * dx])
    c = S.uniform_cells(p[:, 0], p[:, 1], p[:, 2], 2)
    c.upt[:, 0] = 10.0
    c.sec[:, 1] = 1.0
    c.sat[:, 1] = 1.0
    return sc, c


@pytest.mark.parametrize("mixed", [False, True])
def test_long_x_lines_members_equal_single(mixed):
    """x edges of 131..258 voxels: 8 segments per line in the FP32-mixed form
    (16 in FP64) in their own work list, beside a member of short lines."""
    members = [box_member(150, 12, 10, 1500, 61), box_member(258, 8, 7, 1500, 62), small_member(263)]
    P = 7
    ens = hip.Ensemble(mixed=mixed)
    ids = [ens.add(sc, c) for sc, c in members]
    ens.run_periods(P, use_graph=True)
    for mid, (sc, c) in zip(ids, members):
        assert_same(member_state(ens, mid), run_single(sc, c, P, mixed), f"member {sc.name} mixed={mixed}")
    ens.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_overlapped_periods_equal_serial(mixed):
    """pe_set_overlap(1): the mechanics step on a second stream beside the
    substeps gives the serial periods' results bit for bit (graph and eager)."""
    members = [small_member(280), small_member(281), small_member(282)]
    out = []
    for overlap in (False, True):
        ens = hip.Ensemble(mixed=mixed)
        ens.set_overlap(overlap)
        ids = [ens.add(sc, c) for sc, c in members]
        ens.run_periods(4, use_graph=True)
        ens.run_periods(3, use_graph=False)
        out.append([member_state(ens, i) for i in ids])
        ens.close()
    for a, b, (sc, _) in zip(out[0], out[1], members):
        assert_same(b, a, f"member {sc.name} overlapped, mixed={mixed}")


def test_mixed_mirror_resort_and_writeback_equal_single():
    """The mixed ensemble keeps its mechanics state in an agent mirror in
    record order, re-sorted every 64 mechanics steps, and writes the cell
    fields back only when they are read.  Reads in the middle of the run
    (fields after 50 periods, a neighbour query after 90), phenotype
    refreshes and two re-sorts must leave every member equal to its single
    simulation at each point, bit for bit, previous velocities, pressure and
    the mechanics voxel included."""
    extra = (hip.F_PREV_VX, hip.F_PREV_VY, hip.F_PREV_VZ, hip.F_PRESSURE, hip.F_MECH_VOXEL)
    members = [small_member(260, dt_pheno=1.0), small_member(261, dt_pheno=2.0)]
    ens = hip.Ensemble(mixed=True)
    ids = [ens.add(sc, c) for sc, c in members]
    sims = [hip.Simulation(sc, c, mixed=True) for sc, c in members]

    def advance(periods, use_graph):
        ens.run_periods(periods, use_graph=use_graph)
        ens.wait()
        for sim, (sc, c) in zip(sims, members):
            k = substeps_until(sc.dt_diff, sc.dt_mech)
            kp = substeps_until(sc.dt_diff, sc.dt_pheno)
            for _ in range(periods):
                for _ in range(k):
                    sim.substep()
                sim.mech_step()
                sim.mech_periods = getattr(sim, "mech_periods", 0) + 1
                if (sim.mech_periods * k) % kp == 0:
                    sim.pheno_step()
            sim.wait()

    def compare(when):
        for mid, sim in zip(ids, sims):
            a = {"rho": ens.density(mid)}
            b = {"rho": sim.density()}
            for f in POS + VEL + extra:
                a[f] = ens.field(mid, f)
                b[f] = sim.field(f)
            assert_same(a, b, f"member {mid} {when}")

    advance(50, True)
    compare("after 50 periods")
    advance(40, True)
    for mid, sim in zip(ids, sims):
        qa = ens.neighbor_query(mid, hip.QUERY_CONTACT)
        qb = sim.neighbor_query(hip.QUERY_CONTACT)
        assert all(np.array_equal(x, y) for x, y in zip(qa, qb)), f"member {mid}: contact lists after 90 periods"
    advance(50, False)
    compare("after 140 periods")
    for sim in sims:
        sim.close()
    ens.close()
