"""Learned warm start in 3D (Q = 5, p = 2), run by the quad-gather and the
node-lane MFMA predictors, with the network trained on seeded 3D density waves
(weights/wave_p2_3d): the inference kernel against the oracle's forward
pass, the gain of the learned start at equal Picard iterations on a held-out
wave, the early exit with its per-element fallback (a tolerance nothing
passes, or a NaN network output, reproduces the plain predictor bit for
bit), conservation and the error against FP64 over steps, and finite steps
on the spherical shock tube with outflow boundaries."""
import os

import numpy as np
import pytest

from layouts import two_level_centre

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "ader-dg-amr_amd", "weights", "wave_p2_3d.f32")
SHOCK = [0.4, 1.0, 1.0, 0.125, 0.1]
WAVE = [0.2, 1.0]


def _ctx3(adg, precision=1, ic=(3, SHOCK), periodic=False, n=16):
    ctx = adg.Context(3, 2, 1.4, precision)
    ctx.set_geometry([n] * 3, [0] * 3, [1] * 3, periodic=[1, 1, 1] if periodic else [0, 0, 0])
    levels = two_level_centre(n, 16, D=3)
    ctx.set_levels(levels)
    for l in range(len(levels)):
        ctx.project_ic(l, ic[0], ic[1])
    ctx.average_down()
    ctx.fill_ghosts()
    ctx.set_dt(0.5, 10.0, 0.0)
    return ctx, levels


def _weights():
    return np.fromfile(WEIGHTS, dtype=np.float32)


def test_inference_3d_matches_oracle(adg, orc):
    """Warm-start increments of the 3D network on a density-wave state
    against the oracle forward pass (BF16 operands, FP32 accumulation)."""
    print("case 3D warm start vs oracle, density wave p=2 16^3 single level", flush=True)
    ctx = adg.Context(3, 2, 1.4, 1)
    ctx.set_geometry([16] * 3, [0] * 3, [1] * 3)
    from layouts import single_level
    ctx.set_levels(single_level(16, 16, D=3))
    ctx.project_ic(0, 5, [0.3, 1.2])
    ctx.fill_ghosts()
    Dt = ctx.set_dt(0.5, 10.0, 0.0)
    ctx.mlp_init(1)
    ctx.mlp_set_params(_weights())
    sz = ctx.mlp_sizes()
    nin, nin_p, nh, nout, nout_p, npar = (sz[k] for k in ("nin", "nin_p", "nh", "nout", "nout_p", "npar"))
    assert (nin, nout) == (5 * 10 + 1 + 3, 5 * 15)
    par = ctx.mlp_params()
    norm = par[npar:]
    mux, sdx = norm[:nin_p], norm[nin_p:2 * nin_p]
    muy, sdy = norm[2 * nin_p:2 * nin_p + nout_p], norm[2 * nin_p + nout_p:]
    U = ctx.download_Uw(0)
    T = orc.Tables(2, 3)
    delta = ctx.mlp_warm_start(0)
    tau = np.array([Dt * 16] * 3)
    ns = U.shape[2]
    xs = np.zeros((ns, nin_p), np.float32)
    for s in range(ns):
        xs[s, :nin] = (orc.mlp_features(T, U[:, :, s], tau).astype(np.float32) - mux[:nin]) / sdx[:nin]
    _, _, yp = orc.mlp_loss_grad((nin_p, nh, nout_p, nout), par[:npar], xs, np.zeros((ns, nout_p), np.float32),
                                 bf16=True, want_grad=False)
    ref = (U[0, 0] * tau.max())[None, :] * (muy[:nout, None] + sdy[:nout, None] * yp[:, :nout].T)
    e = np.linalg.norm(delta - ref) / np.linalg.norm(ref)
    print(f"  warm start normwise relative difference {e:.2e}", flush=True)
    assert e < 1e-2


@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("precision", [0, 1])
def test_learned_3d_gain_and_fallback(adg, precision, kind):
    """Held-out density wave (amplitude and pressure of no training run).
    At equal iterations the learned start is closer to the converged
    predictor; with a tolerance no element meets, every element falls back
    and the result equals the plain predictor bitwise."""
    print(f"case 3D learned predictor density wave two levels, precision {precision} kind {kind}", flush=True)
    g, _ = _ctx3(adg, precision=precision, ic=(5, [0.27, 0.8]), periodic=True)
    g.set_predictor_kernel(kind)
    g.mlp_init(1)
    g.mlp_set_params(_weights())
    g.keep_Hw(True)
    nl = 2
    g.mlp_collect(6)
    Hs = [g.download_Hw(l) for l in range(nl)]
    res = {}
    for it in (1, 2):
        g.mlp_collect(it)
        res[f"plain {it} it"] = [g.download_Hw(l) for l in range(nl)]
    for it in (1, 2):
        g.set_learned(2, max_iters=it)
        g.predictor(-1, 1)
        res[f"learned + {it} it"] = [g.download_Hw(l) for l in range(nl)]
    sc = max(np.abs(h).max() for h in Hs)
    err = {k: max(np.abs(v[l] - Hs[l]).max() for l in range(nl)) / sc for k, v in res.items()}
    print("  distance to converged H_w: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["learned + 1 it"] < err["plain 1 it"]
    # fallback: nothing passes 1e-30, every element recomputed by the plain path
    g.set_learned(True, tol=1e-30, max_iters=1)
    nf = g.predictor(-1, 1)
    learned = [g.download_Hw(l) for l in range(nl)]
    g.predictor(-1, 0)
    plain = [g.download_Hw(l) for l in range(nl)]
    print(f"  tol 1e-30: {nf} elements fell back", flush=True)
    assert nf > 0
    for l in range(nl):
        assert np.array_equal(learned[l], plain[l]), f"level {l}"


def test_learned_3d_conservation_and_accuracy(adg):
    """FP32 + learned start (tol 1e-4, 2 iterations) on two levels of the
    periodic density wave: mass and energy conserved to 1e-13 over 4 steps,
    within 1e-5 of FP64.  On the spherical shock tube with outflow (a
    discontinuity with no limiter, which FP64 and FP32 themselves resolve
    differently after a step) two learned steps stay finite."""
    print("case 3D learned FP32 steps: density wave vs FP64, shock tube finite", flush=True)
    out = {}
    for name, prec, learned in (("fp64", 0, False), ("fp32", 1, False), ("learned", 1, True)):
        c, levels = _ctx3(adg, precision=prec, ic=(5, WAVE), periodic=True)
        if learned:
            c.mlp_init(1)
            c.mlp_set_params(_weights())
            c.set_learned(True, tol=1e-4, max_iters=2)
            c.learned_stats(reset=True)
        m0 = c.totals()
        for _ in range(4):
            c.step(0.5, 10.0)
        m1 = c.totals()
        dm, de = abs(m1[0] - m0[0]) / abs(m0[0]), abs(m1[4] - m0[4]) / abs(m0[4])
        extra = ""
        if learned:
            ne, nf = c.learned_stats()
            extra = f", {ne} predictions, {nf} fell back"
        print(f"  {name}: mass drift {dm:.1e}, energy drift {de:.1e}{extra}", flush=True)
        assert dm <= 1e-13 and de <= 1e-13, name
        out[name] = [c.download_Uw(l) for l in range(len(levels))]
    sc = max(np.abs(u).max() for u in out["fp64"])
    ew = max(np.abs(a - b).max() for a, b in zip(out["learned"], out["fp64"])) / sc
    e32 = max(np.abs(a - b).max() for a, b in zip(out["fp32"], out["fp64"])) / sc
    print(f"  after 4 steps vs FP64: plain FP32 {e32:.2e}, learned FP32 {ew:.2e} of the state scale", flush=True)
    assert ew < 1e-5
    c, levels = _ctx3(adg, precision=1, ic=(3, SHOCK))
    c.mlp_init(1)
    c.mlp_set_params(_weights())
    c.set_learned(True, tol=1e-4, max_iters=2)
    for _ in range(2):
        c.step(0.5, 10.0)
    assert all(np.isfinite(c.download_Uw(l)).all() for l in range(len(levels)))


@pytest.mark.parametrize("kind", [2, 3])
def test_learned_3d_nan_start_falls_back(adg, kind):
    """A network whose output is NaN (NaN output biases): every element
    fails the acceptance test and the result is the plain predictor's."""
    print(f"case 3D learned start with NaN network output, kind {kind}", flush=True)
    g, levels = _ctx3(adg, precision=1, ic=(5, WAVE), periodic=True)
    g.set_predictor_kernel(kind)
    g.mlp_init(1)
    par = _weights().copy()
    sz = g.mlp_sizes()
    nin_p, nh, nout_p, npar = (sz[k] for k in ("nin_p", "nh", "nout_p", "npar"))
    par[npar - nout_p:npar] = np.nan   # b3
    g.mlp_set_params(par)
    g.keep_Hw(True)
    g.set_learned(True, tol=1e-4, max_iters=2)
    nf = g.predictor(-1, 1)
    learned = [g.download_Hw(l) for l in range(len(levels))]
    g.predictor(-1, 0)
    plain = [g.download_Hw(l) for l in range(len(levels))]
    npred = sum(g.npred(l) for l in range(len(levels)))
    print(f"  {nf} of {npred} elements fell back", flush=True)
    assert nf == npred
    for l in range(len(levels)):
        assert np.array_equal(learned[l], plain[l]), f"level {l}"
