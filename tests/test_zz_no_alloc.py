"""adg_conserved_totals_noalloc (static device scratch, no allocation)
against adg_conserved_totals on the same state: bitwise equal, on two calls
in a row, the second with a smaller input, and a clear error for an input
beyond the scratch."""
import numpy as np
import pytest

from devbuf import DeviceArray
from layouts import kh_bands, single_level

pytestmark = pytest.mark.gpu

VORTEX = [5.0, 5.0, 5.0, 1.0, 0.0]


def _ctx(adg, D, P, levels, dom, phi, ic, prm):
    ctx = adg.Context(D, P, 1.4, 0)
    ctx.set_geometry(dom, [0] * D, phi)
    ctx.set_levels(levels)
    for l in range(len(levels)):
        ctx.project_ic(l, ic, prm)
    ctx.average_down()
    ctx.fill_ghosts()
    ctx.set_dt(0.5, 1.0, 0.0)
    return ctx


def _noalloc(ctx, Q):
    out = DeviceArray(Q)
    ctx.totals_noalloc(out.ptr)
    ctx.sync()
    got = out.read()
    out.close()
    return got


def test_noalloc_totals_two_calls(adg):
    print("case noalloc totals: KH p=3 3 levels, then vortex p=2 16^2", flush=True)
    big = _ctx(adg, 2, 3, kh_bands(64, 3, 32), [64, 64], [1, 1], 1, [0.01, 1e-3, 7])
    big.step(0.5, 1.0)
    ref1 = big.totals()
    got1 = _noalloc(big, 4)
    assert np.array_equal(got1, ref1), (got1, ref1)
    small = _ctx(adg, 2, 2, single_level(16, 8), [16, 16], [10, 10], 0, VORTEX)
    ref2 = small.totals()
    got2 = _noalloc(small, 4)
    assert np.array_equal(got2, ref2), (got2, ref2)
    # the first context again, after the smaller call left its partials behind
    assert np.array_equal(_noalloc(big, 4), ref1)


def test_noalloc_totals_rejects_oversize(adg):
    """1664^2 = 2 768 896 uncovered cells, above the 2 621 440 of 2D."""
    print("case noalloc totals oversize: 1664^2 single level p=1", flush=True)
    ctx = adg.Context(2, 1, 1.4, 0)
    ctx.set_geometry([1664, 1664], [0, 0], [1, 1])
    ctx.set_levels(single_level(1664, 416))
    out = DeviceArray(4, fill=0.0)
    with pytest.raises(adg.AdgError, match="static scratch"):
        ctx.totals_noalloc(out.ptr)
    ctx.sync()
    assert not out.read().any()
    out.close()
