"""Parity at the benchmark's own sizes: the FP64 form against the oracle on
the 3-level Kelvin-Helmholtz hierarchy of the bench's kernel measurement
(64^2 -> 128^2 -> 256^2 bands, 28 672 elements, p = 3) and on a 3-level 3D
hierarchy, per (variable, mode) at 1e-12 (tests/bars.py); and the committed
warm-start network (weights/kh_p3_2d, trained on the bench's tag-driven
jobs) on the bench's state: fallback fraction, density L1 difference to
the FP64 form, mass and energy drift, and what it gains over the plain
start at equal Picard iterations, early on the band hierarchy and late in
the tag-driven job."""
import os

import numpy as np
import pytest

from bars import density_l1_diff, mode_err
from layouts import kh_bands, tiles, two_level_centre

pytestmark = pytest.mark.gpu

TOL = 1e-12
KH_BENCH = [0.01, 1e-3, 20240611]   # the bench's KH state: amplitude, noise, seed
WAVE = [0.2, 1.0]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "ader-dg-amr_amd", "weights", "kh_p3_2d.f32")


def _pair(adg, orc, D, P, levels, dom, phi, ic, prm, precision=0):
    T = orc.Tables(P, D)
    H = orc.Hier(D, levels, dom, [0] * D, phi)
    Uw = orc.project_ic(T, H, ic, prm)
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    ctx = adg.Context(D, P, 1.4, precision)
    ctx.set_geometry(dom, [0] * D, phi)
    ctx.set_levels(levels)
    for l in range(len(levels)):
        ctx.upload_Uw(l, Uw[l])
    return T, H, Uw, ctx


def test_kh_bench_hierarchy_fp64(adg, orc):
    print("case KH p=3 bench hierarchy 64^2/128^2/256^2, 2 steps FP64", flush=True)
    levels = kh_bands(64, 3, 32)
    T, H, Uw, ctx = _pair(adg, orc, 2, 3, levels, [64, 64], [1, 1], 1, KH_BENCH)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    m0 = ctx.totals()
    for _ in range(2):
        ctx.step(0.5, 1.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=2)
    for l in range(3):
        e, at = mode_err(ctx.download_Uw(l), Uw[l])
        print(f"  level {l}: per-mode error {e:.2e} at (q, n) = {at}", flush=True)
        assert e < TOL, f"level {l}"
    m1 = ctx.totals()
    assert abs(m1[0] - m0[0]) <= 1e-14 * abs(m0[0])


def _three_level_3d():
    # 8^3, the centre half at level 1, the centre quarter of level 1 at level 2
    return two_level_centre(8, 4, D=3) + [tiles(32, 8, 3, [12] * 3, [19] * 3)]


@pytest.mark.parametrize("P", [1, 2])
def test_3d_three_levels_fp64(adg, orc, P):
    print(f"case 3D density wave P={P} 3 levels 8^3/16^3/32^3, 2 steps FP64", flush=True)
    levels = _three_level_3d()
    T, H, Uw, ctx = _pair(adg, orc, 3, P, levels, [8] * 3, [1] * 3, 5, WAVE)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    m0 = ctx.totals()
    for _ in range(2):
        ctx.step(0.5, 1.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=2)
    for l in range(3):
        e, at = mode_err(ctx.download_Uw(l), Uw[l])
        print(f"  level {l}: per-mode error {e:.2e} at (q, n) = {at}", flush=True)
        assert e < TOL, f"level {l}"
    m1 = ctx.totals()
    assert abs(m1[0] - m0[0]) <= 1e-14 * abs(m0[0])
    assert abs(m1[4] - m0[4]) <= 1e-14 * abs(m0[4])


def _kh_ctx(adg, levels, precision, weights=None, tol=1e-4, max_iters=2):
    ctx = adg.Context(2, 3, 1.4, precision)
    ctx.set_geometry([64, 64], [0, 0], [1, 1])
    ctx.set_levels(levels)
    for l in range(len(levels)):
        ctx.project_ic(l, 7, KH_BENCH + [64])   # the bench's state: noise on the 64^2 grid
    ctx.average_down()
    ctx.fill_ghosts()
    ctx.set_dt(0.5, 1.0, 0.0)
    if weights is not None:
        ctx.mlp_init(1)
        ctx.mlp_set_params(weights)
        ctx.set_learned(True, tol=tol, max_iters=max_iters)
        ctx.learned_stats(reset=True)
    return ctx


def test_committed_network_on_bench_hierarchy(adg):
    """The network weights/kh_p3_2d on the band hierarchy of the bench's
    kernel measurement (the bench's state, held out of training), 50 steps
    from t = 0, learned + FP32 at the bench's settings (tol 1e-4, at most 2
    iterations) against the FP64 form.  The network is trained on the
    tag-driven jobs, where fewer than 1e-5 of the elements fall back; on
    these early band states 1.4 % do (elements whose start is far off: the
    acceptance test fails only for a start error above about 0.3 rho0 at
    this Picard contraction), within the 2 % bar; every one of them takes
    the plain predictor, so the accuracy bar below is unchanged."""
    print("case committed network kh_p3_2d on the bench hierarchy, 50 steps", flush=True)
    levels = kh_bands(64, 3, 32)
    w = np.fromfile(WEIGHTS, dtype=np.float32)
    nsteps = 50
    runs = {}
    for form, prec, wt in (("fp64", 0, None), ("fp32", 1, None), ("learned", 1, w)):
        ctx = _kh_ctx(adg, levels, prec, wt)
        t0 = ctx.totals()
        for _ in range(nsteps):
            ctx.step(0.5, 1.0)
        t1 = ctx.totals()
        stats = ctx.learned_stats() if wt is not None else None
        runs[form] = ([ctx.download_Uw(l) for l in range(3)], t0, t1, stats)
        ctx.close()
    ref = runs["fp64"][0]
    for form in ("fp32", "learned"):
        U, t0, t1, stats = runs[form]
        l1 = density_l1_diff(U, ref, levels)
        dm = abs(t1[0] - t0[0]) / abs(t0[0])
        de = abs(t1[3] - t0[3]) / abs(t0[3])
        msg = f"  {form}: density L1 difference to FP64 {l1:.2e}, mass drift {dm:.1e}, energy drift {de:.1e}"
        if stats is not None:
            ne, nf = stats
            msg += f", {nf} of {ne} elements fell back ({nf / max(ne, 1):.3%})"
            assert nf <= 2e-2 * ne
        print(msg, flush=True)
        assert l1 < 1e-6
        assert dm < 1e-12 and de < 1e-12
    # the learned form is as close to FP64 as plain FP32 is (within 2x)
    l1_plain = density_l1_diff(runs["fp32"][0], ref, levels)
    l1_learn = density_l1_diff(runs["learned"][0], ref, levels)
    assert l1_learn < 2.0 * l1_plain + 1e-12


def test_committed_network_gain_on_bench_hierarchy(adg):
    """At one Picard iteration the learned start ends closer to the
    converged predictor (2P + 2 iterations) than the reference's start on
    every level of the bench hierarchy; at two, both are below the FP32
    round-off of the predictor (1e-7 of the scale), where the early exit
    accepts (on the finest level, where the CFL number is largest, the
    plain start's second iteration is the closer one)."""
    print("case committed network gain at equal iterations, bench hierarchy", flush=True)
    levels = kh_bands(64, 3, 32)
    w = np.fromfile(WEIGHTS, dtype=np.float32)
    g = _kh_ctx(adg, levels, 0)
    g.mlp_init(1)
    g.mlp_set_params(w)
    g.keep_Hw(True)
    g.mlp_collect(8)
    conv = [g.download_Hw(l) for l in range(3)]
    res = {}
    for it in (1, 2):
        g.mlp_collect(it)
        res[f"plain + {it}"] = [g.download_Hw(l) for l in range(3)]
        g.set_learned(2, max_iters=it)   # exactly `it` iterations from the learned start
        g.predictor(-1, 1)
        res[f"learned + {it}"] = [g.download_Hw(l) for l in range(3)]
    for l in range(3):
        sc = np.abs(conv[l]).max()
        d = {k: np.abs(v[l] - conv[l]).max() / sc for k, v in res.items()}
        print(f"  level {l}: " + ", ".join(f"{k} it {v:.2e}" for k, v in d.items()), flush=True)
        assert d["learned + 1"] < 0.3 * d["plain + 1"]
        assert d["learned + 2"] < 1e-7 and d["plain + 2"] < 1e-6


def _tagged_job(adg, precision, weights=None):
    """The bench's job: three levels from density-gradient tags, the TVB
    limiter (M = 5000), both composite rules, regrid every 0.05."""
    ctx = adg.Context(2, 3, 1.4, precision)
    ctx.set_geometry([64, 64], [0, 0], [1, 1])
    ctx.set_conservative_reflux(True)
    ctx.set_cfl_uncovered(True)
    ctx.set_levels([tiles(64, 32, 2)])
    prm = KH_BENCH + [64]
    ctx.project_ic(0, 7, prm)
    ctx.fill_ghosts()
    ctx.set_amr(2, max_grid_size=32)
    ctx.set_tagging(2, [0.05])
    ctx.set_limiter(True, 5000.0, [1.0] * 3, 1)
    for _ in range(2):
        ctx.regrid(0.0)
        for l in range(len(ctx.get_levels())):
            ctx.project_ic(l, 7, prm)
        ctx.average_down()
        ctx.fill_ghosts()
    if weights is not None:
        ctx.mlp_init(1)
        ctx.mlp_set_params(weights)
    ctx.set_dt(0.5, 1.0, 0.0)
    return ctx


def _advance(ctx, t_stop):
    t, Dt = ctx.get_time()
    next_rg = 0.05
    while t < t_stop - 1e-12:
        n = int(min(64, 0.5 * min(t_stop - t, next_rg - t) / Dt))
        if n >= 2:
            ctx.step_graph(0.5, 1.0, n)
        else:
            ctx.step(0.5, 1.0)
        t, Dt = ctx.get_time()
        if t >= next_rg - 1e-12:
            ctx.regrid(t)
            next_rg = t + 0.05


def test_committed_network_on_a_late_tagged_state(adg):
    """The network on states it meets late in the bench's job (the bench's
    seed is held out of training): at t = 0.6, on the tag-driven hierarchy
    of that time, one Picard iteration from the learned start ends at least
    20x closer to the converged predictor than one from the reference's
    start, and two end closer than the reference's three; then 2 000 steps
    of the learned form fall back for fewer than 1e-4 of the elements and
    conserve mass."""
    print("case committed network at t=0.6 of the tagged KH job", flush=True)
    w = np.fromfile(WEIGHTS, dtype=np.float32)
    ctx = _tagged_job(adg, 0, w)
    _advance(ctx, 0.6)
    levels = ctx.get_levels()
    ctx.keep_Hw(True)
    ctx.mlp_collect(8)
    conv = [ctx.download_Hw(l) for l in range(len(levels))]
    res = {}
    for it in (1, 2, 3):
        ctx.mlp_collect(it)
        res[f"plain + {it}"] = [ctx.download_Hw(l) for l in range(len(levels))]
    for it in (1, 2):
        ctx.set_learned(2, max_iters=it)
        ctx.predictor(-1, 1)
        res[f"learned + {it}"] = [ctx.download_Hw(l) for l in range(len(levels))]
    sc = max(np.abs(c).max() for c in conv)
    err = {k: max(np.abs(v[l] - conv[l]).max() for l in range(len(levels))) / sc for k, v in res.items()}
    print("  distance to converged H_w: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["learned + 1"] < err["plain + 1"] / 20
    assert err["learned + 2"] < err["plain + 3"]
    ctx.close()

    ctx = _tagged_job(adg, 1, w)
    _advance(ctx, 0.6)
    ctx.set_learned(True, tol=1e-4, max_iters=2)
    ctx.learned_stats(reset=True)
    m0 = ctx.totals()
    ctx.step_graph(0.5, 1.0, 2000)
    m1 = ctx.totals()
    ne, nf = ctx.learned_stats()
    print(f"  2000 learned steps from t=0.6: {nf} of {ne} elements fell back, "
          f"mass drift {abs(m1[0] - m0[0]) / m0[0]:.1e}", flush=True)
    assert ne > 0 and nf < 1e-4 * ne
    assert abs(m1[0] - m0[0]) / m0[0] < 1e-13
    ctx.close()
