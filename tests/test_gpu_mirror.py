"""The mechanics mirror of the single simulation (pc_sim_set_mechanics_mirror),
plain FP64 and FP32-mixed: a simulation stepping on the mirror equals the
same simulation stepping on the cell fields directly, bit for bit, in every
cell field, the density and the hash, through eager and graph periods, the
phenotype clock (which writes the mirror back before the voxel refresh),
field reads in the middle of a run, bins above and below 16 members, 2-D and
3-D grids, and the step loop of pc_sim_run."""
import numpy as np
import pytest

from physicell_amd import hip as H
from physicell_amd import scenarios as S

pytestmark = pytest.mark.gpu

FIELDS = (H.F_POS_X, H.F_POS_Y, H.F_POS_Z, H.F_VEL_X, H.F_VEL_Y, H.F_VEL_Z, H.F_PREV_VX, H.F_PREV_VY,
          H.F_PREV_VZ, H.F_PRESSURE, H.F_MECH_VOXEL, H.F_VOXEL_IDX)


def crowded():
    """About 12 agents per 30 um bin: bins of both sort paths (<= 16 and > 16)."""
    return S.small3d(12, 10, 8, n_cells=4000, seed=11)


def packed():
    """About 33 agents per 30 um bin: rows of three bins above 64 records (the
    x-trimmed mark pass walks them 64 records at a time) and more marked
    candidates per agent than its LDS list holds (the list is flushed)."""
    return S.small3d(12, 10, 8, n_cells=8000, seed=13)


BUILDERS = dict(S.BUILDERS, crowded=crowded, packed=packed)


def state(sim):
    out = {"rho": sim.density()}
    for f in FIELDS:
        out[f] = sim.field(f)
    off, cells = sim.hash()
    out["offsets"], out["sorted"] = off, cells
    return out


def assert_same(a, b, what):
    for k in a:
        if not np.array_equal(a[k], b[k]):
            d = np.abs(np.asarray(a[k], np.float64) - np.asarray(b[k], np.float64))
            raise AssertionError(f"{what} {k}: {int(np.count_nonzero(d))} values differ, max abs {d.max():.3e}")


def make(builder, mirror, dt_pheno=None, mixed=False):
    spec, cells = BUILDERS[builder]()
    if dt_pheno is not None:
        spec.dt_pheno = dt_pheno
    sim = H.Simulation(spec, cells, mixed=mixed)
    sim.set_mechanics_mirror(mirror)
    return sim


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("builder", ["config0", "small3d", "config1", "crowded", "packed"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_mirror_equals_direct_periods(builder, use_graph, mixed):
    print(f"mirror {builder} graph={use_graph} mixed={mixed}", flush=True)
    out = []
    for mirror in (True, False):
        # phenotype clock every 3 periods: the mirror is written back before
        # each voxel refresh and the run continues on it
        sim = make(builder, mirror, dt_pheno=0.3, mixed=mixed)
        sim.run_periods(4, use_graph=use_graph)
        mid = {f: sim.field(f) for f in FIELDS}        # a read in the middle of the run
        sim.run_periods(5, use_graph=use_graph)
        sim.wait()
        st = state(sim)
        st.update({f"mid{f}": v for f, v in mid.items()})
        out.append(st)
        sim.close()
    assert_same(out[0], out[1], f"{builder} graph={use_graph} mixed={mixed}")
    off = out[0]["offsets"].astype(np.int64)
    if builder == "crowded":
        sizes = np.diff(off)
        assert sizes.max() > 16 and ((sizes > 0) & (sizes <= 16)).any(), "both bin paths"
    if builder == "packed":
        sizes = np.diff(off)
        assert (sizes[:-2] + sizes[1:-1] + sizes[2:]).max() > 64, "rows above 64 records"


@pytest.mark.parametrize("builder", ["config0", "config1"])
def test_mirror_equals_direct_run(builder):
    out = []
    for mirror in (True, False):
        sim = make(builder, mirror)
        counts = sim.run(3.0)
        out.append((counts, state(sim)))
        sim.close()
    assert out[0][0] == out[1][0]
    assert_same(out[0][1], out[1][1], f"{builder} run")


def test_mirror_switch_mid_run():
    """Turning the mirror off and on between periods keeps the trajectory."""
    ref = make("crowded", False)
    ref.run_periods(6)
    ref.wait()
    a = state(ref)
    ref.close()
    sim = make("crowded", True)
    sim.run_periods(2)
    sim.set_mechanics_mirror(False)
    sim.run_periods(2)
    sim.set_mechanics_mirror(True)
    sim.run_periods(2, use_graph=True)
    sim.wait()
    assert_same(state(sim), a, "switching")
    sim.close()


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("builder", ["config1", "crowded"])
def test_mirror_resort_long_run(builder, mixed):
    """The mirror's slots keep their order between steps and are put back in
    record order every 64 mechanics steps; over 140 periods (two re-sorts,
    graph replays and eager periods, a field read and a neighbour query in
    between) the run equals the direct one bit for bit."""
    out = []
    for mirror in (True, False):
        sim = make(builder, mirror, mixed=mixed)
        sim.run_periods(70, use_graph=True)
        mid = {f: sim.field(f) for f in FIELDS}
        sim.run_periods(30, use_graph=False)
        q = sim.neighbor_query(H.QUERY_CONTACT)
        sim.run_periods(40, use_graph=True)
        sim.wait()
        st = state(sim)
        st.update({f"mid{f}": v for f, v in mid.items()})
        st["q_off"], st["q_idx"] = q
        out.append(st)
        sim.close()
    assert_same(out[0], out[1], f"{builder} 140 periods mixed={mixed}")


def test_mirror_giant_bin_takes_exact_pass():
    """One 30 um bin holding 48000 agents: their record and mark positions lie
    more than 32767 apart, beyond the 16-bit field of the plain form's 8-byte
    mark entries, so the gather flags the step and every agent takes the
    exact FP64 pass; the run still equals the direct one bit for bit."""
    spec, cells = S.small3d(2, 2, 1, n_cells=48000, seed=3)
    out = []
    for mirror in (True, False):
        sim = H.Simulation(spec, cells)
        sim.set_mechanics_mirror(mirror)
        sim.run_periods(2)
        sim.wait()
        out.append(state(sim))
        sim.close()
    sizes = np.diff(out[0]["offsets"].astype(np.int64))
    assert sizes.max() > 43690, "one bin large enough to overflow the entry"
    assert_same(out[0], out[1], "giant bin")
