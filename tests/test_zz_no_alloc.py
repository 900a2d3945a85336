"""hc_substrate_totals_noalloc (static device scratch, no allocation) against
hc_substrate_totals (per-call scratch): the same kernels over the same
values, so the totals are equal bit for bit.  Two calls in a row, the second
on a smaller input (fewer substrates, FP32 field), then a third that exceeds
the static maximum of substrates and is rejected by the host check."""
import ctypes

import numpy as np
import pytest

from physicell_amd import hip as H
from physicell_amd import scenarios as S

pytestmark = pytest.mark.gpu
MAX_SUBSTRATES = 64


def _totals(ctx, fn, buf, g, ns):
    out = np.full(ns, np.nan)
    ctx._check(getattr(H.load(), fn)(ctx.h, buf.h, g, H._ptr(out)), fn)
    return out


def _case(ctx, nx, ny, nz, ns, fp32, seed):
    spec, _ = S.small3d(nx, ny, nz, n_cells=1)
    g = H.grid_of(spec, fp32=fp32)
    g.n_substrates = ns
    rng = np.random.default_rng(seed)
    rho = rng.uniform(-1.0, 50.0, ns * nx * ny * nz)
    rho = rho.astype(np.float32) if fp32 else rho
    buf = ctx.alloc(rho.nbytes)
    buf.upload(rho)
    return g, buf, rho


def test_noalloc_totals_equal_allocating_totals():
    ctx = H.Context(0)
    for nx, ny, nz, ns, fp32, seed in ((120, 90, 70, 6, False, 1), (17, 9, 5, 2, True, 2)):
        print(f"noalloc totals: grid {nx}x{ny}x{nz}, {ns} substrates, fp32={fp32}", flush=True)
        g, buf, rho = _case(ctx, nx, ny, nz, ns, fp32, seed)
        ref = _totals(ctx, "hc_substrate_totals", buf, g, ns)
        got = _totals(ctx, "hc_substrate_totals_noalloc", buf, g, ns)
        assert np.array_equal(got, ref), (got, ref)
        V = nx * ny * nz
        exact = np.array([np.sum(rho[s * V:(s + 1) * V].astype(np.float64)) for s in range(ns)])
        assert np.allclose(got, exact, rtol=1e-12, atol=0)
        buf.free()
    # beyond the static maximum: rejected before any launch
    g, buf, _ = _case(ctx, 4, 3, 2, MAX_SUBSTRATES + 1, False, 3)
    out = np.zeros(MAX_SUBSTRATES + 1)
    rc = H.load().hc_substrate_totals_noalloc(ctx.h, buf.h, g, H._ptr(out))
    assert rc == -5, rc
    assert b"exceed the static scratch" in H.load().hc_last_error(ctx.h)
    buf.free()
    ctx.close()
