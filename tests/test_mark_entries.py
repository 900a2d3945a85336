"""Error bounds of the 8-byte mark entries (cells_common.h: mark_entry16,
force_agent64t, force_agent32x<true>).

A mark entry holds x - x_min in 1/16 um and y, z relative to the bin's lower
corner plus 64 um in 1/512 um.  The mark pass forms each candidate's d^2 in
FP32 from the entry and the agent's own FP32 coordinates (one FMA per axis:
the products with -1/16 and -1/512 are exact) and compares it with a
threshold that carries MARK16_SLACK = 2.25 um^2 above the force law's bound.
The kernels rely on two bounds, checked here on sampled pairs within 32 um
(the largest cutoff the mark pass serves) over the whole coordinate range
(x - x_min up to 4096 um, bin-relative y and z in [0, 60) um):
  - |d^2_mark - d^2| < 2.14 um^2 (below the slack: no pair the law keeps is
    missed);
  - |dx_mark - dx| < 0.032 um (below XTRIM_MARGIN = 0.05 um: the x-trimmed
    walk covers every marked candidate).
A mistake in the encoding (a dropped offset, a wrong scale, truncation
instead of rounding) breaks one of them."""
import numpy as np

SLACK = 2.25
XTRIM_MARGIN = 0.05
F32 = np.float32


def encode(xg, yb, zb):
    qx = np.rint(xg * 16.0)
    qy = np.rint((yb + 64.0) * 512.0)
    qz = np.rint((zb + 64.0) * 512.0)
    assert qx.min() >= 0 and qx.max() <= 65535 and qy.min() >= 0 and qy.max() <= 65535
    assert qz.min() >= 0 and qz.max() <= 65535
    return qx, qy, qz


def fma32(a, b, c):
    # a * b is exact in FP64 for FP32 operands; one rounding to FP32
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def mark_d2(agent, cand):
    """The kernel's FP32 arithmetic: agent (xg, yb, zb) as FP32 of the FP64
    values (y, z plus 64), candidate from its entry; dk, dj row offsets are
    folded into the candidate's bin-relative coordinates here."""
    axg, ayb, azb = agent
    qx, qy, qz = encode(*cand)
    mxg = axg.astype(F32)
    my64 = (ayb + 64.0).astype(F32)
    mz64 = (azb + 64.0).astype(F32)
    ddx = fma32(qx.astype(F32), F32(-0.0625) * np.ones_like(mxg), mxg)
    ddy = fma32(qy.astype(F32), F32(-1.0 / 512.0) * np.ones_like(my64), my64)
    ddz = fma32(qz.astype(F32), F32(-1.0 / 512.0) * np.ones_like(mz64), mz64)
    d2 = fma32(ddz, ddz, fma32(ddy, ddy, (ddx * ddx).astype(F32)))
    return d2.astype(np.float64), ddx.astype(np.float64)


def test_mark_entry_error_within_slack():
    rng = np.random.default_rng(7)
    n = 2_000_000
    # candidates anywhere in the coordinate range, the agent within 32 um
    cxg = rng.uniform(40.0, 4050.0, n)
    cyb = rng.uniform(0.0, 60.0, n)
    czb = rng.uniform(0.0, 60.0, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = 32.0 * rng.random(n) ** (1.0 / 3.0)
    d = u * r[:, None]
    # the agent's coordinates in the candidate's row frame (bin-relative y, z
    # may leave [0, 60) by up to the pair distance)
    axg, ayb, azb = cxg + d[:, 0], cyb + d[:, 1], czb + d[:, 2]
    d2_exact = (axg - cxg) ** 2 + (ayb - cyb) ** 2 + (azb - czb) ** 2
    d2, ddx = mark_d2((axg, ayb, azb), (cxg, cyb, czb))
    err = np.abs(d2 - d2_exact)
    assert err.max() < 2.14 < SLACK, err.max()
    assert np.abs(ddx - (axg - cxg)).max() < 0.032 < XTRIM_MARGIN


def test_mark_entry_rounds_to_nearest():
    # the quantisation step is 1/16 um in x and 1/512 um in y, z, rounded to
    # nearest: at most half a step from the value, and exact on the grid
    x = np.array([0.0, 1.0 / 16.0, 0.03125 - 1e-9, 4095.9375])
    y = np.array([0.0, 59.998046875, 1.0 / 1024.0 - 1e-9, -63.998046875])
    qx, qy, _ = encode(x, y, y)
    assert np.array_equal(qx, [0.0, 1.0, 0.0, 65535.0])
    assert np.array_equal(qy / 512.0 - 64.0, [0.0, 59.998046875, 0.0, -63.998046875])
