"""The TVB limiter on the GPU (Limiter_w / Limiter_linear_tvb,
AmrDG.h:1657-1855) through the C-ABI against the oracle: the limiter alone
on a discontinuous state (same cells limited, U_w per (variable, mode) at
1e-12), steps with the limiter inside the step (2D KH on 3 levels with the
limiting M = 200, AMR_TVB_C = 1; the 3D spherical shock tube),
and the 3D shock tube with the limiter conserving mass and energy over 20
steps."""
import numpy as np
import pytest

from bars import mode_err
from layouts import kh_bands, single_level

pytestmark = pytest.mark.gpu

TOL = 1e-12
KH = [0.01, 1e-3, 7]
SHOCK = [0.4, 1.0, 1.0, 0.125, 0.1, 0.5, 0.5, 0.5]


def _pair(adg, orc, D, P, levels, dom, ic, prm, precision=0):
    T = orc.Tables(P, D)
    H = orc.Hier(D, levels, dom, [0] * D, [1] * D)
    Uw = orc.project_ic(T, H, ic, prm)
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    ctx = adg.Context(D, P, 1.4, precision)
    ctx.set_geometry(dom, [0] * D, [1] * D)
    ctx.set_levels(levels)
    for l in range(len(levels)):
        ctx.upload_Uw(l, Uw[l])
    return T, H, Uw, ctx


@pytest.mark.parametrize("M", [0.0, 200.0])
def test_limiter_alone_kh(adg, orc, M):
    """Two unlimited steps, then the limiter alone on the finest level."""
    print(f"case limiter alone KH P=3 3 levels 16^2 base, M={M}", flush=True)
    levels = kh_bands(16, 3, 16)
    T, H, Uw, ctx = _pair(adg, orc, 2, 3, levels, [16, 16], 1, KH)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    for _ in range(2):
        ctx.step(0.5, 1.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=2)
    tvb = [1.0, 1.0, 1.0]
    ctx.set_limiter(False, M, tvb)
    ctx.set_limiter(True, M, tvb)
    n_gpu = ctx.limit()
    n_orc = orc.limiter_tvb(T, H, Uw, 2, M, tvb)
    print(f"  cells limited: GPU {n_gpu}, oracle {n_orc}", flush=True)
    assert n_gpu == n_orc and n_orc > 0
    valid = H.valid_mask(2)
    e, at = mode_err(ctx.download_Uw(2)[:, :, valid], Uw[2][:, :, valid])
    print(f"  finest level per-mode error {e:.2e} at {at}", flush=True)
    assert e < TOL
    assert ctx.limiter_stats() == n_gpu


@pytest.mark.parametrize("kind", [0, 1])
def test_limited_steps_kh(adg, orc, kind):
    """3 steps of evolve with the limiter in the step (M = 200: the
    reference's M = 5000 limits nothing at these coarse sizes)."""
    print(f"case limited steps KH P=3 3 levels kind={kind}", flush=True)
    levels = kh_bands(16, 3, 16)
    T, H, Uw, ctx = _pair(adg, orc, 2, 3, levels, [16, 16], 1, KH)
    if kind:
        ctx.set_predictor_kernel(1)
    tvb = [1.0, 1.0, 1.0]
    ctx.set_limiter(True, 200.0, tvb, 1)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    m0 = ctx.totals()
    for _ in range(3):
        ctx.step(0.5, 1.0)
    n, _, _, nl = orc.evolve_lim(T, H, Uw, 0.0, Dt, 1.0, 0.5, 200.0, tvb, 1, nsteps_max=3)
    print(f"  cells limited: GPU {ctx.limiter_stats()}, oracle {nl}", flush=True)
    assert ctx.limiter_stats() == nl and nl > 0
    for l in range(3):
        e, at = mode_err(ctx.download_Uw(l), Uw[l])
        print(f"  level {l}: per-mode error {e:.2e} at {at}", flush=True)
        assert e < TOL, f"level {l}"
    m1 = ctx.totals()
    assert abs(m1[0] - m0[0]) <= 1e-14 * abs(m0[0])


def test_limited_graph_steps_match_eager(adg):
    """The limiter inside the captured step graph gives the eager steps'
    bits on every replay (t_limit = 2: the device step counter)."""
    print("case limited graph replays KH P=3 3 levels, t_limit 2", flush=True)
    levels = kh_bands(16, 3, 16)
    out = []
    for graph in (False, True):
        ctx = adg.Context(2, 3, 1.4, 0)
        ctx.set_geometry([16, 16], [0, 0], [1, 1])
        ctx.set_levels(levels)
        for l in range(3):
            ctx.project_ic(l, 1, KH)
        ctx.average_down()
        ctx.fill_ghosts()
        ctx.set_limiter(True, 200.0, [1.0] * 3, 2)
        ctx.set_dt(0.5, 1.0, 0.0)
        snaps = []
        for _ in range(4):
            if graph:
                ctx.step_graph(0.5, 1.0, 1)
            else:
                ctx.step(0.5, 1.0)
            snaps.append([ctx.download_Uw(l) for l in range(3)])
        out.append((snaps, ctx.limiter_stats()))
        ctx.close()
    assert out[0][1] == out[1][1] and out[0][1] > 0
    for k in range(4):
        for l in range(3):
            assert np.array_equal(out[0][0][k][l], out[1][0][k][l]), f"step {k} level {l}"


def _limit_ic_oracle(orc, T, H, Uw, M):
    orc.limiter_tvb(T, H, Uw, len(Uw) - 1, M, [1.0] * len(Uw))
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)


def test_limited_shock_tube_3d(adg, orc):
    """The spherical shock tube, its projected initial state limited once
    (unlimited, its p = 2 projection has negative nodal density and
    pressure and the first step diverges), then 3 limited steps."""
    print("case limited 3D shock tube P=2 12^3, 3 steps vs oracle", flush=True)
    levels = single_level(12, 6, D=3)
    T, H, Uw, ctx = _pair(adg, orc, 3, 2, levels, [12] * 3, 3, SHOCK)
    ctx.set_limiter(True, 0.0, [1.0], 1)
    n0 = ctx.limit()
    ctx.fill_ghosts()
    _limit_ic_oracle(orc, T, H, Uw, 0.0)
    assert n0 == 12 ** 3
    ctx.limiter_stats(reset=True)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    for _ in range(3):
        ctx.step(0.5, 1.0)
    _, _, _, nl = orc.evolve_lim(T, H, Uw, 0.0, Dt, 1.0, 0.5, 0.0, [1.0], 1, nsteps_max=3)
    assert ctx.limiter_stats() == nl and nl > 0
    e, at = mode_err(ctx.download_Uw(0), Uw[0])
    print(f"  per-mode error {e:.2e} at {at}; {nl} cell limitings", flush=True)
    assert e < TOL


@pytest.mark.parametrize("precision", [0, 1])
def test_limited_shock_tube_3d_conserves(adg, precision):
    """20 steps of the periodic 3D spherical shock tube with the limiter
    (M = 0; the initial state limited once) on one 16^3 level: mass and
    energy drift below 1e-13, cell-mean density within its initial range."""
    print(f"case limited 3D shock tube 16^3 20 steps precision {precision}", flush=True)
    ctx = adg.Context(3, 2, 1.4, precision)
    ctx.set_geometry([16] * 3, [0] * 3, [1] * 3)
    ctx.set_levels(single_level(16, 8, D=3))
    ctx.project_ic(0, 3, SHOCK)
    ctx.fill_ghosts()
    ctx.set_limiter(True, 0.0, [1.0], 1)
    ctx.limit()
    ctx.fill_ghosts()
    ctx.set_dt(0.5, 1.0, 0.0)
    m0 = ctx.totals()
    for _ in range(20):
        ctx.step(0.5, 1.0)
    m1 = ctx.totals()
    U = ctx.download_Uw(0)
    dm, de = abs(m1[0] - m0[0]) / m0[0], abs(m1[4] - m0[4]) / m0[4]
    print(f"  mass drift {dm:.2e}, energy drift {de:.2e}, min cell density {U[0, 0].min():.3f}, "
          f"{ctx.limiter_stats()} cell limitings", flush=True)
    assert np.isfinite(U).all()
    assert dm < 1e-13 and de < 1e-13
    assert 0.12 < U[0, 0].min() and U[0, 0].max() < 1.01
