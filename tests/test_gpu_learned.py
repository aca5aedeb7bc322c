"""Learned warm start on the GPU, through the C-ABI, against the CPU oracle:
the multi-mode KH initial condition, the (features, target) pairs the
collection kernel writes, the MLP loss and gradient of the BF16-MFMA
training kernels, the fused inference kernel, and the early-exit predictor
(accuracy against the plain predictor and what the network gains)."""
import numpy as np
import pytest

from layouts import single_level

pytestmark = pytest.mark.gpu

VORTEX = [5.0, 5.0, 5.0, 1.0, 0.0]


def _vortex_ctx(adg, P=3, n=16, precision=0):
    ctx = adg.Context(2, P, 1.4, precision)
    ctx.set_geometry([n, n], [0, 0], [10, 10])
    ctx.set_levels(single_level(n))
    ctx.project_ic(0, 0, VORTEX)
    ctx.fill_ghosts()
    Dt = ctx.set_dt(0.5, 1.0, 0.0)
    return ctx, Dt


def _kh_ctx(adg, seed, n=32, precision=1, P=3):
    ctx = adg.Context(2, P, 1.4, precision)
    ctx.set_geometry([n, n], [0, 0], [1, 1])
    ctx.set_levels(single_level(n, 16))
    ctx.project_ic(0, 4, [seed, 0.02])
    ctx.fill_ghosts()
    ctx.set_dt(0.5, 10.0, 0.0)
    return ctx


def test_kh_multimode_ic_matches_oracle(adg, orc):
    print("case IC multi-mode KH P=3 16^2", flush=True)
    T = orc.Tables(3, 2)
    levels = single_level(16)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [1, 1])
    ref = orc.project_ic(T, H, 4, [11, 0.02])[0]
    ctx = adg.Context(2, 3, 1.4, 0)
    ctx.set_geometry([16, 16], [0, 0], [1, 1])
    ctx.set_levels(levels)
    ctx.project_ic(0, 4, [11, 0.02])
    got = ctx.download_Uw(0)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # the perturbation is there: v differs from zero with amplitude <= 8 * 0.02
    v = ref[2, 0] / ref[0, 0]
    assert 1e-3 < np.abs(v).max() < 0.16


def test_mlp_gradient_matches_oracle(adg, orc):
    """Loss and gradient of the device forward/backward kernels (BF16 MFMA
    operands, FP32 accumulation) against the oracle in FP64 with the same
    BF16 operand rounding; normwise per parameter block."""
    print("case MLP gradient vs oracle, 64 rows", flush=True)
    ctx = adg.Context(2, 3, 1.4, 1)
    ctx.mlp_init(5)
    sz = ctx.mlp_sizes()
    nin, nin_p, nh, nout, nout_p, npar = (sz[k] for k in ("nin", "nin_p", "nh", "nout", "nout_p", "npar"))
    par = ctx.mlp_params()[:npar]
    rng = np.random.default_rng(9)
    x = np.zeros((64, nin_p), np.float32)
    x[:, :nin] = rng.standard_normal((64, nin))
    y = np.zeros((64, nout_p), np.float32)
    y[:, :nout] = rng.standard_normal((64, nout))
    loss, g = ctx.mlp_loss_grad(x, y)
    lo, go, _ = orc.mlp_loss_grad((nin_p, nh, nout_p, nout), par, x, y, bf16=True)
    print(f"  loss device {loss:.8e} oracle {lo:.8e}", flush=True)
    assert abs(loss - lo) <= 2e-4 * lo
    offs = np.cumsum([0, nh * nin_p, nh, nh * nh, nh, nout_p * nh, nout_p])
    for name, a, b in zip(("W1", "b1", "W2", "b2", "W3", "b3"), offs[:-1], offs[1:]):
        e = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        print(f"  {name}: normwise relative gradient error {e:.2e}", flush=True)
        assert e < 5e-3, name
    # padded input columns and output rows carry no gradient
    W1g = g[: nh * nin_p].reshape(nh, nin_p)
    assert np.all(W1g[:, nin:] == 0)
    W3g = g[offs[4]:offs[5]].reshape(nout_p, nh)
    assert np.all(W3g[nout:] == 0)


def test_collect_matches_oracle(adg, orc):
    """Features and targets of every valid element after a converged
    (8-iteration) FP64 predictor, against the oracle's definitions."""
    print("case MLP collect vortex P=3 16^2", flush=True)
    ctx, Dt = _vortex_ctx(adg)
    ctx.mlp_init(1)
    n = ctx.mlp_collect(8)
    assert n == 256
    x, y = ctx.mlp_dataset(0, n)
    T = orc.Tables(3, 2)
    H = orc.Hier(2, single_level(16), [16, 16], [0, 0], [10, 10])
    Uw = orc.project_ic(T, H, 0, VORTEX)
    orc.fill_patch(T, H, Uw)
    Dto = orc.set_dt(T, H, Uw, 0.5)
    assert abs(Dto - Dt) <= 1e-14 * Dt
    Hs = orc.predictor_iters(T, H, 0, Uw[0], Dt, 8)
    valid = np.nonzero(H.valid_mask(0))[0]
    sz = ctx.mlp_sizes()
    tau = np.array([Dt / (10 / 16)] * 2)
    ymax = 0.0
    for i, s in enumerate(valid):
        xo = orc.mlp_features(T, Uw[0][:, :, s], tau)
        yo = orc.mlp_target(T, Uw[0][:, :, s], Hs[:, :, s], tau)
        assert np.allclose(x[i, :sz["nin"]], xo, rtol=2e-7, atol=1e-7), i
        assert np.all(x[i, sz["nin"]:] == 0)
        assert np.abs(y[i, :sz["nout"]] - yo).max() <= 2e-7 * max(1.0, np.abs(yo).max()) + 1e-6, i
        ymax = max(ymax, np.abs(yo).max())
    assert ymax > 1e-2


def _collect_train(adg, seeds, steps_per_seed, nsteps, n=32, batch=512, lr=2e-3, schedule=None):
    ctx = _kh_ctx(adg, seeds[0], n=n)
    ctx.mlp_init(3)
    for j, sd in enumerate(seeds):
        if j:
            ctx.project_ic(0, 4, [sd, 0.02])
            ctx.fill_ghosts()
            ctx.set_dt(0.5, 10.0, 0.0)
        for _ in range(steps_per_seed):
            ctx.mlp_collect(2 * 3 + 2)
            ctx.step(0.5, 10.0)
    ctx.mlp_fit_normalization()
    loss0 = ctx.mlp_train(1, batch, 0.0)
    for k, (ns, lr_k) in enumerate(schedule or [(nsteps, lr)]):
        loss = ctx.mlp_train(ns, batch, lr_k, seed=11 + k)
    return ctx, loss0, loss


def test_training_reduces_loss_and_inference_matches_oracle(adg, orc):
    print("case MLP train KH seeds 1-2, 32^2, then warm start vs oracle", flush=True)
    ctx, loss0, loss = _collect_train(adg, [1, 2], 4, 400)
    print(f"  loss {loss0:.4e} -> {loss:.4e}", flush=True)
    assert loss < 0.1 * loss0
    # fused inference kernel vs oracle forward on the same state
    sz = ctx.mlp_sizes()
    nin, nin_p, nh, nout, nout_p, npar = (sz[k] for k in ("nin", "nin_p", "nh", "nout", "nout_p", "npar"))
    par = ctx.mlp_params()
    norm = par[npar:]
    mux, sdx = norm[:nin_p], norm[nin_p:2 * nin_p]
    muy, sdy = norm[2 * nin_p:2 * nin_p + nout_p], norm[2 * nin_p + nout_p:]
    U = ctx.download_Uw(0)
    T = orc.Tables(3, 2)
    Dt = ctx.set_dt(0.5, 10.0, 0.0)
    delta = ctx.mlp_warm_start(0)
    tau = np.array([Dt * 32] * 2)
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
    # the other launch shapes of the inference kernel give the same increments
    for tiles in (2, 0, 1):
        ctx.set_warm_tiles(tiles)
        d2 = ctx.mlp_warm_start(0)
        e2 = np.abs(d2 - delta).max() / np.abs(delta).max()
        print(f"  warm tiles {tiles}: max difference to the default shape {e2:.2e}", flush=True)
        assert e2 <= 1e-6


def test_learned_predictor_accuracy_and_gain(adg, orc):
    """Held-out seed.  (1) The gain: at equal Picard iterations the learned
    start ends closer to the converged predictor than the plain start.
    (2) Early exit and fallback: a tight tolerance falls back and reproduces
    plain FP32; a loose one accepts most starts and stays close to FP64; the
    error against FP64 grows with the tolerance."""
    print("case learned predictor KH held-out seed 9", flush=True)
    ctx, _, loss = _collect_train(adg, [1, 2, 3, 4], 8, 0, batch=1024,
                                  schedule=[(1500, 2e-3), (1500, 5e-4), (1000, 1e-4)])
    print(f"  training loss {loss:.3e}", flush=True)
    par = ctx.mlp_params()
    g = _kh_ctx(adg, 9, precision=0)
    g.mlp_set_params(par)
    g.keep_Hw(True)
    g.mlp_collect(8)
    Hs = g.download_Hw(0)
    res = {}
    for it in (1, 2, 3):
        g.mlp_collect(it)
        res[f"plain {it} it"] = g.download_Hw(0)
    for it in (1, 2):
        g.set_learned(2, max_iters=it)
        g.predictor(-1, 1)
        res[f"learned + {it} it"] = g.download_Hw(0)
    Hh = orc.Hier(2, single_level(32, 16), [32, 32], [0, 0], [1, 1])
    valid = np.nonzero(Hh.valid_mask(0))[0]
    sc = np.abs(Hs[:, :, valid]).max()
    err = {k: np.abs(v[:, :, valid] - Hs[:, :, valid]).max() / sc for k, v in res.items()}
    print("  distance to converged H_w: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    # at equal Picard iterations the learned start is closer to converged
    assert err["learned + 1 it"] < 0.3 * err["plain 1 it"]
    assert err["learned + 2 it"] < err["plain 2 it"]

    ref64 = _kh_ctx(adg, 9, precision=0)
    plain = _kh_ctx(adg, 9)
    for _ in range(6):
        ref64.step(0.5, 10.0)
        plain.step(0.5, 10.0)
    r = ref64.download_Uw(0)
    scale = np.abs(r).max()
    e_fp32 = np.abs(plain.download_Uw(0) - r).max() / scale
    print(f"  plain FP32 max error vs FP64 {e_fp32:.2e}", flush=True)
    rows = {}
    for tol in (1e-6, 1e-5, 1e-4, 1e-3):
        for mi in (1, 2):
            learn = _kh_ctx(adg, 9)
            learn.mlp_set_params(par)
            learn.set_learned(True, tol=tol, max_iters=mi)
            learn.learned_stats(reset=True)
            m0 = learn.totals()
            for _ in range(6):
                learn.step(0.5, 10.0)
            m1 = learn.totals()
            ne, nf = learn.learned_stats()
            e = np.abs(learn.download_Uw(0) - r).max() / scale
            dm, de = abs(m1[0] - m0[0]) / abs(m0[0]), abs(m1[3] - m0[3]) / abs(m0[3])
            print(f"  tol {tol:g} max_iters {mi}: {ne} elements, {nf / max(ne, 1):.1%} fell back, "
                  f"max error vs FP64 {e:.2e}, mass drift {dm:.1e}, energy drift {de:.1e}", flush=True)
            # conservation holds whatever each element decided (4 patches:
            # ghost copies and their source cells must take the same path)
            assert dm <= 1e-14 and de <= 1e-14
            assert 0 < ne and nf <= ne
            rows[(tol, mi)] = (ne, nf, e)
    # a tight tolerance sends (almost) every element to the plain fallback,
    # which then reproduces the plain FP32 result
    ne, nf, e = rows[(1e-6, 1)]
    assert nf > 0.9 * ne and e < 2 * e_fp32 + 1e-12
    # a loose tolerance accepts most starts and the steps stay close to FP64
    ne, nf, e = rows[(1e-3, 2)]
    assert nf < 0.2 * ne and e < 1e-3
    # the error grows monotonically as the tolerance is relaxed
    errs = [rows[(t, 2)][2] for t in (1e-6, 1e-5, 1e-4, 1e-3)]
    assert all(x <= 1.5 * y for x, y in zip(errs, errs[1:]))


def test_nan_start_falls_back(adg):
    """A network whose output is NaN (NaN output biases): every element
    fails the acceptance test (a NaN residual must not pass through a max
    that drops NaN) and the result is the plain predictor's, bit for bit."""
    import os
    print("case learned start with NaN network output, KH 32^2 FP32 node-lane", flush=True)
    w = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ader-dg-amr_amd", "weights",
                     "kh_p3_2d.f32")
    g = _kh_ctx(adg, 5)
    g.mlp_init(1)
    par = np.fromfile(w, dtype=np.float32).copy()
    sz = g.mlp_sizes()
    par[sz["npar"] - sz["nout_p"]:sz["npar"]] = np.nan   # b3
    g.mlp_set_params(par)
    g.keep_Hw(True)
    g.set_learned(True, tol=1e-4, max_iters=2)
    nf = g.predictor(-1, 1)
    learned = g.download_Hw(0)
    g.predictor(-1, 0)
    plain = g.download_Hw(0)
    print(f"  {nf} of {g.npred(0)} elements fell back", flush=True)
    assert nf == g.npred(0)
    assert np.array_equal(learned, plain)
