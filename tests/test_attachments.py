"""Spring attachments (tryFormAllAttachments, src/cell_interactions.cpp:
411-478): the oracle's O(N^2) restatement (or_form_attachments) against the
reference's own compiled function (oracle/_ref) where its unseeded draws
cannot matter (rate * dt >= 1 accepts every draw), including lists that fill
to the 12-partner cap, dead cells and cells with rate 0; the draw itself
(u < max(rate_i, rate_j) dt) against its acceptance probability on isolated
pairs; and the invariants every result must keep.  Host only."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_path.so")
M = O.MAX_ATTACHMENTS


def ref_lib():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    lib = ctypes.CDLL(REF)
    vp = ctypes.c_void_p
    lib.ref_form_attachments.argtypes = [ctypes.c_uint32] + [vp] * 7 + [ctypes.c_double, vp, vp]
    return lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def population(n, seed, box, rate_hi):
    """FP32-representable values (the reference stores FP32), so both sides
    see the same numbers."""
    r = np.random.default_rng(seed)
    c = {"x": f32(r.uniform(-box, box, n)), "y": f32(r.uniform(-box, box, n)),
         "z": f32(r.uniform(-box / 2, box / 2, n)), "radius": f32(r.uniform(5, 10, n)),
         "rel": f32(r.uniform(1.0, 1.4, n)), "alive": (r.uniform(size=n) < 0.9).astype(np.uint32)}
    rate = f32(r.uniform(0.5, 1.0, n) * rate_hi)
    rate[r.uniform(size=n) < 0.1] = 0.0
    c["rate"] = rate
    return c


def run_ref(c, dt):
    lib = ref_lib()
    n = len(c["x"])
    cnt = np.zeros(n, np.uint32)
    tg = np.zeros(n * M, np.uint32)
    keep = [np.ascontiguousarray(c[k], np.float32) for k in ("x", "y", "z", "radius", "rel")]
    alive = np.ascontiguousarray(c["alive"], np.uint32)
    rate = np.ascontiguousarray(c["rate"], np.float32)
    assert lib.ref_form_attachments(n, *(p(a) for a in keep), p(alive), p(rate), dt, p(cnt), p(tg)) == 0
    return cnt, tg.reshape(n, M)


def run_oracle(c, dt, seed):
    return O.form_attachments(c["x"], c["y"], c["z"], c["radius"], c["rel"], c["alive"], c["rate"], dt, seed)


def check_invariants(c, cnt, tg):
    n = len(cnt)
    assert cnt.max(initial=0) <= M
    pairs = set()
    for i in range(n):
        row = tg[i, :cnt[i]].tolist()
        assert len(set(row)) == len(row) and i not in row
        for j in row:
            assert c["alive"][i] and c["alive"][j] and c["rate"][i] > 0 and c["rate"][j] > 0
            d = np.sqrt((c["x"][j] - c["x"][i]) ** 2 + (c["y"][j] - c["y"][i]) ** 2 + (c["z"][j] - c["z"][i]) ** 2)
            assert d <= max(c["rel"][i], c["rel"][j]) * (c["radius"][i] + c["radius"][j])
            pairs.add((min(i, j), max(i, j)))
    # symmetric: every attachment sits in both lists
    assert sum(int(v) for v in cnt) == 2 * len(pairs)


@pytest.mark.parametrize("n,box,seed", [(600, 150.0, 1), (400, 40.0, 2), (250, 18.0, 3), (1, 10.0, 4)])
def test_oracle_equals_reference_when_every_draw_accepts(n, box, seed):
    """rate * dt >= 1: the reference's result is fixed by its order, cap and
    distance test alone.  The dense boxes fill lists to 12."""
    dt = 6.0
    c = population(n, seed, box, rate_hi=2.0 / dt)
    c["rate"][c["rate"] > 0] = np.maximum(c["rate"][c["rate"] > 0], np.float32(1.0 / dt * 1.0001))
    c["rate"] = f32(c["rate"])
    rc, rt = run_ref(c, dt)
    oc, ot, formed = run_oracle(c, dt, seed=99)
    assert np.array_equal(oc, rc)
    for i in range(n):
        assert np.array_equal(ot[i, :oc[i]], rt[i, :rc[i]]), i
    assert formed == int(oc.sum()) // 2
    check_invariants(c, oc, ot)
    if box <= 40.0 and n > M:
        assert oc.max() == M                      # the cap is exercised


def test_no_rate_no_attachments():
    c = population(300, 5, 40.0, rate_hi=1.0)
    c["rate"][:] = 0.0
    rc, _ = run_ref(c, 6.0)
    oc, _, formed = run_oracle(c, 6.0, seed=1)
    assert formed == 0 and not oc.any() and not rc.any()


def test_draw_probability_isolated_pairs():
    """One pair in reach per call: attached with probability max(rate_i,
    rate_j) dt (the larger rate counts; u < p, not u <= p)."""
    dt = 1.0
    c = {"x": np.array([0.0, 10.0]), "y": np.zeros(2), "z": np.zeros(2), "radius": np.full(2, 6.0),
         "rel": np.full(2, 1.25), "alive": np.ones(2, np.uint32), "rate": np.array([0.05, 0.3])}
    hits = sum(run_oracle(c, dt, seed=s)[2] for s in range(4000))
    assert abs(hits / 4000 - 0.3) < 4 * np.sqrt(0.3 * 0.7 / 4000)
    c["x"][1] = 15.1                               # beyond 1.25 x 12 um: never
    assert sum(run_oracle(c, dt, seed=s)[2] for s in range(200)) == 0


def test_partial_rates_invariants_and_existing_lists():
    """rate * dt < 1 and lists that already hold partners: still symmetric,
    capped, in reach; existing partners are not attached twice."""
    c = population(500, 6, 60.0, rate_hi=0.1)
    cnt, tg, f1 = run_oracle(c, 1.0, seed=7)
    check_invariants(c, cnt, tg)
    cnt2, tg2, f2 = O.form_attachments(c["x"], c["y"], c["z"], c["radius"], c["rel"], c["alive"], c["rate"], 1.0,
                                       8, count=cnt, targets=tg)
    check_invariants(c, cnt2, tg2)
    assert f1 > 0 and f2 > 0
    for i in range(len(cnt)):
        assert np.array_equal(tg2[i, :cnt[i]], tg[i, :cnt[i]])   # old partners stay, new ones append
