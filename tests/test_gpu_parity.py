"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same seeded inputs.  Bar: FP64 form within 1e-12 relative per
(conserved variable, mode), each mode against its own size (floored at 1 %
of its variable group's scale, tests/bars.py)."""
import numpy as np
import pytest

from bars import mode_err

from layouts import kh_bands, single_level, tiles, two_level_centre

pytestmark = pytest.mark.gpu

TOL = 1e-12
VORTEX = [5.0, 5.0, 5.0, 1.0, 0.0]


def rel_err(a, b):
    """Per-(variable, mode) relative error (tests/bars.py: each mode against
    its own size, floored at 1 % of its variable group's scale)."""
    return mode_err(a, b)[0]


def setup(adg, orc, D, P, levels, dom, plo, phi, ic_case, params, gamma=1.4, kind=0, precision=0):
    T = orc.Tables(P, D)
    H = orc.Hier(D, levels, dom, plo, phi)
    Uw = orc.project_ic(T, H, ic_case, params, gamma)
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    ctx = adg.Context(D, P, gamma, precision)
    if kind:
        ctx.set_predictor_kernel(kind)
    ctx.set_geometry(dom, plo, phi)
    ctx.set_levels(levels)
    for l in range(len(levels)):
        assert ctx.nslot(l) == H.nslot(l)
        ctx.upload_Uw(l, Uw[l])
    return T, H, Uw, ctx


KINDS = [1, 2, 3]  # VALU, MFMA quad-gather, MFMA node-lane predictor


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("P", [1, 2, 3])
def test_predictor_single_patch(adg, orc, P, kind):
    print(f"case predictor vortex P={P} 16^2 kind={kind}", flush=True)
    levels = single_level(16)
    T, H, Uw, ctx = setup(adg, orc, 2, P, levels, [16, 16], [0, 0], [10, 10], 0, VORTEX, kind=kind)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.keep_Hw(True)
    ctx.set_dt_value(Dt)
    ctx.predictor()
    Hw = ctx.download_Hw(0)
    ref = orc.predictor(T, H, 0, Uw[0], Dt)
    assert rel_err(Hw, ref) < TOL


def test_set_dt_matches(adg, orc):
    print("case set_dt vortex P=2 32^2 4 patches", flush=True)
    levels = single_level(32, 16)
    T, H, Uw, ctx = setup(adg, orc, 2, 2, levels, [32, 32], [0, 0], [10, 10], 0, VORTEX)
    dt_ref = orc.set_dt(T, H, Uw, 0.5)
    dt = ctx.set_dt(0.5, 10.0, 0.0)
    assert abs(dt - dt_ref) <= 1e-14 * dt_ref


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("P,mg", [(2, 32), (2, 16), (3, 8)])
def test_ader_step_single_level(adg, orc, P, mg, kind):
    print(f"case ader vortex P={P} 32^2 patches {mg} kind={kind}", flush=True)
    levels = single_level(32, mg)
    T, H, Uw, ctx = setup(adg, orc, 2, P, levels, [32, 32], [0, 0], [10, 10], 0, VORTEX, kind=kind)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt_value(Dt)
    ctx.ader()
    orc.ader(T, H, Uw, Dt)
    U = ctx.download_Uw(0)
    valid = H.valid_mask(0)
    assert rel_err(U[:, :, valid], Uw[0][:, :, valid]) < TOL


def test_face_flux_matches(adg, orc):
    print("case Fnum vortex P=2 16^2 2x2 patches", flush=True)
    levels = single_level(16, 8)
    T, H, Uw, ctx = setup(adg, orc, 2, 2, levels, [16, 16], [0, 0], [10, 10], 0, VORTEX)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt_value(Dt)
    ctx.predictor()
    ctx.face_flux()
    fn = orc.ader(T, H, [u.copy() for u in Uw], Dt, want_fnum=True)
    at = 0
    for d in range(2):
        nf = H.nfslot(0, d)
        ref = fn[at:at + T.Q * T.qMp_st_bd * nf].reshape(T.Q, T.qMp_st_bd, nf)
        at += ref.size
        assert rel_err(ctx.download_Fnum(0, d), ref) < TOL


@pytest.mark.parametrize("P", [2, 3])
def test_evolve_steps_single_level(adg, orc, P):
    print(f"case evolve 10 steps vortex P={P} 32^2", flush=True)
    levels = single_level(32, 16)
    T, H, Uw, ctx = setup(adg, orc, 2, P, levels, [32, 32], [0, 0], [10, 10], 0, VORTEX)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    for _ in range(10):
        ctx.step(0.5, 1.0)
    n, t, _ = orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=10)
    assert n == 10
    valid = H.valid_mask(0)
    assert rel_err(ctx.download_Uw(0)[:, :, valid], Uw[0][:, :, valid]) < TOL


@pytest.mark.parametrize("P", [2, 3])
def test_two_level_step(adg, orc, P):
    print(f"case two-level vortex P={P} 16^2+centre", flush=True)
    levels = two_level_centre(16, 8)
    T, H, Uw, ctx = setup(adg, orc, 2, P, levels, [16, 16], [0, 0], [10, 10], 0, VORTEX)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    for _ in range(3):
        ctx.step(0.5, 1.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=3)
    for l in range(2):
        U = ctx.download_Uw(l)
        assert rel_err(U, Uw[l]) < TOL, f"level {l}"


@pytest.mark.parametrize("kind", KINDS)
def test_kh_three_levels(adg, orc, kind):
    print(f"case KH P=3 16^2 3 levels kind={kind}", flush=True)
    levels = kh_bands(16, 3, 16)
    T, H, Uw, ctx = setup(adg, orc, 2, 3, levels, [16, 16], [0, 0], [1, 1], 1, [0.01, 1e-3, 7], kind=kind)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    m0 = ctx.totals()
    for _ in range(3):
        ctx.step(0.5, 1.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=3)
    for l in range(3):
        assert rel_err(ctx.download_Uw(l), Uw[l]) < TOL, f"level {l}"
    m1 = ctx.totals()
    assert abs(m1[0] - m0[0]) / abs(m0[0]) < 1e-13


def test_gpu_ic_matches_oracle(adg, orc):
    print("case IC projection vortex + KH", flush=True)
    levels = single_level(16, 8)
    T = orc.Tables(3, 2)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [10, 10])
    ctx = adg.Context(2, 3)
    ctx.set_geometry([16, 16], [0, 0], [10, 10])
    ctx.set_levels(levels)
    ctx.project_ic(0, 0, VORTEX)
    ref = orc.project_ic(T, H, 0, VORTEX)[0]
    assert rel_err(ctx.download_Uw(0), ref) < 1e-13
    ctx.project_ic(0, 1, [0.01, 1e-3, 7])
    ref = orc.project_ic(T, H, 1, [0.01, 1e-3, 7])[0]
    assert rel_err(ctx.download_Uw(0), ref) < 1e-13
    # the fixed-grid noise variant on the unit square, two levels
    levels = two_level_centre(16, 8)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [1, 1])
    ctx = adg.Context(2, 3)
    ctx.set_geometry([16, 16], [0, 0], [1, 1])
    ctx.set_levels(levels)
    prm = [0.01, 1e-3, 7, 8]
    ref = orc.project_ic(T, H, 7, prm)
    for l in range(2):
        ctx.project_ic(l, 7, prm)
        assert rel_err(ctx.download_Uw(l), ref[l]) < 1e-13, f"level {l}"


@pytest.mark.parametrize("kind,waves", [(2, 8), (3, 8), (3, 16)])
@pytest.mark.parametrize("P", [2, 3])
def test_fp32_form_kh(adg, orc, P, kind, waves):
    """Precision form: FP32 MFMA contractions and FP32 face fluxes, FP64 U_w
    accumulation, reflux and reductions.  Stated error vs the FP64 oracle
    after 10 steps on a 3-level KH hierarchy: per variable, max|dU| over the
    largest |U| of all variables (the FP32 round-off of 10 increments) below
    1e-6; composite mass and energy conserved to 1e-13."""
    print(f"case FP32 form KH P={P} 16^2 3 levels kind={kind} waves={waves}", flush=True)
    levels = kh_bands(16, 3, 16)
    T, H, Uw, ctx = setup(adg, orc, 2, P, levels, [16, 16], [0, 0], [1, 1], 1, [0.01, 1e-3, 7], precision=1,
                          kind=kind)
    ctx.set_predictor_waves(waves)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    m0 = ctx.totals()
    for _ in range(10):
        ctx.step(0.5, 1.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=10)
    for l in range(3):
        got = ctx.download_Uw(l)
        scale = np.abs(Uw[l]).max()
        errs = [np.abs(got[q] - Uw[l][q]).max() / scale for q in range(T.Q)]
        print(f"  level {l} FP32-vs-FP64 error per variable: {['%.2e' % e for e in errs]}", flush=True)
        assert max(errs) < 1e-6, f"level {l}"
    m1 = ctx.totals()
    assert abs(m1[0] - m0[0]) / abs(m0[0]) < 1e-13
    assert abs(m1[3] - m0[3]) / abs(m0[3]) < 1e-13


def test_mfma_layout(adg):
    """Pins the 16x16x4 MFMA lane layouts the predictor relies on:
    A[i][k] on lane i+16k, B[k][j] on lane j+16k, accumulator register r of
    lane group g holds row 4g+r (FP32) and row 4r+g (FP64)."""
    import ctypes as C
    L = adg.lib()
    L.adg_probe_mfma_layout.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    for prec in (0, 1):
        o = np.zeros(256)
        assert L.adg_probe_mfma_layout(prec, 0, o.ctypes.data_as(C.POINTER(C.c_double))) == 0
        o = o.reshape(64, 4)
        for lane in range(64):
            g = lane // 16
            rows = [4 * r + g if prec == 0 else 4 * g + r for r in range(4)]
            assert list(o[lane]) == [row + 1 for row in rows]
        assert L.adg_probe_mfma_layout(prec, 1, o.ctypes.data_as(C.POINTER(C.c_double))) == 0
        o = o.reshape(64, 4)
        for lane in range(64):
            assert list(o[lane]) == [lane % 16 + 1] * 4


@pytest.mark.parametrize("precision", [0, 1])
def test_step_graph_replays(adg, orc, precision):
    """hipGraph replays of a whole step equal eager steps on every replay,
    U_w and (t, Dt) bitwise (eager work and host syncs interleaved), and the
    FP64 form equals the oracle."""
    print(f"case graph replays KH P=3 16^2 3 levels precision={precision}", flush=True)
    levels = kh_bands(16, 3, 16)
    T, H, Uw, ctx = setup(adg, orc, 2, 3, levels, [16, 16], [0, 0], [1, 1], 1, [0.01, 1e-3, 7],
                          precision=precision)
    eager = adg.Context(2, 3, 1.4, precision)
    eager.set_geometry([16, 16], [0, 0], [1, 1])
    eager.set_levels(levels)
    for l in range(3):
        eager.upload_Uw(l, Uw[l])
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    eager.set_dt(0.5, 1.0, 0.0)
    for rep in range(8):
        ctx.step_graph(0.5, 1.0, 1)
        eager.step(0.5, 1.0)
        ctx.sync()
        eager.totals()
        assert ctx.get_time() == eager.get_time(), f"replay {rep}: (t, Dt)"
        for l in range(3):
            assert np.array_equal(ctx.download_Uw(l), eager.download_Uw(l)), f"replay {rep} level {l}"
    if precision == 0:
        orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=8)
        for l in range(3):
            assert rel_err(ctx.download_Uw(l), Uw[l]) < TOL


@pytest.mark.parametrize("learned", [False, True])
def test_batched_steps_equal_single_steps(adg, learned):
    """adg_steps (a batch of stream launches), adg_step_graph with 4 steps
    per captured graph (a batch of 5: one graph and one stream step) and
    single adg_step calls give bitwise the same U_w, t and Dt: 3 batches of
    5 steps of the 3-level KH state in the FP32 form, plain and with the
    committed warm-start network."""
    print(f"case batched steps vs single steps KH P=3 16^2 3 levels learned={learned}", flush=True)
    import os
    levels = kh_bands(16, 3, 16)
    w = np.fromfile(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                 "ader-dg-amr_amd", "weights", "kh_p3_2d.f32"), dtype=np.float32)
    ctxs = []
    for _ in range(3):
        c = adg.Context(2, 3, 1.4, 1)
        c.set_geometry([16, 16], [0, 0], [1, 1])
        c.set_levels(levels)
        for l in range(3):
            c.project_ic(l, 1, [0.01, 1e-3, 7])
        c.average_down()
        c.fill_ghosts()
        if learned:
            c.mlp_init(1)
            c.mlp_set_params(w)
            c.set_learned(True, tol=1e-4, max_iters=2)
        c.set_dt(0.5, 1.0, 0.0)
        ctxs.append(c)
    ctxs[1].set_graph_unroll(4)
    for b in range(3):
        ctxs[0].steps(0.5, 1.0, 5)
        ctxs[1].step_graph(0.5, 1.0, 5)
        for _ in range(5):
            ctxs[2].step(0.5, 1.0)
        ref = ctxs[2].get_time()
        for k in (0, 1):
            assert ctxs[k].get_time() == ref, f"batch {b} form {k}: (t, Dt)"
            for l in range(3):
                assert np.array_equal(ctxs[k].download_Uw(l), ctxs[2].download_Uw(l)), f"batch {b} form {k} level {l}"
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("nlev,periodic", [(3, [1, 1]), (4, [1, 1]), (3, [1, 0]), (2, [1, 1])])
@pytest.mark.parametrize("precision", [0, 1])
def test_fused_step_equals_unfused(adg, nlev, periodic, precision):
    """The step's fused launches (reflux beside the finest average down, the
    ghost fill beside the CFL reduction and the time advance) give bitwise
    the U_w, t and Dt of the separate calls in the reference's order
    (AmrDG.h:1000-1029): predictor, face flux, corrector, reflux, average
    down from the finest level, ghost fill, CFL and advance."""
    print(f"case fused vs unfused step KH P=3 16^2 {nlev} levels periodic={periodic} precision={precision}",
          flush=True)
    levels = kh_bands(16, nlev, 16)
    ctxs = []
    for _ in range(2):
        ctx = adg.Context(2, 3, 1.4, precision)
        ctx.set_geometry([16, 16], [0, 0], [1, 1], periodic=periodic)
        ctx.set_levels(levels)
        for l in range(nlev):
            ctx.project_ic(l, 1, [0.01, 1e-3, 7])
        ctx.average_down()
        ctx.fill_ghosts()
        ctx.set_dt(0.5, 1.0, 0.0)
        ctxs.append(ctx)
    fused, split = ctxs
    for step in range(3):
        fused.step(0.5, 1.0)
        split.ader()
        split.average_down()
        split.fill_ghosts()
        split.advance(0.5, 1.0)
        assert fused.get_time() == split.get_time(), f"step {step}"
        for l in range(nlev):
            assert np.array_equal(fused.download_Uw(l), split.download_Uw(l)), f"step {step} level {l}"


@pytest.mark.parametrize("periodic", [[1, 1, 1], [0, 0, 0]])
@pytest.mark.parametrize("precision", [0, 1])
def test_fused_step_equals_unfused_3d(adg, periodic, precision):
    """The same in 3D (p=2, three levels, the shock tube limited once):
    the fused epilogue's reflux rows (one thread per row and variable for
    every mode) equal the separate flux-register moments and reflux
    bitwise."""
    print(f"case fused vs unfused step 3D P=2 8^3 3 levels periodic={periodic} precision={precision}", flush=True)
    levels = [tiles(8, 8, 3), tiles(16, 8, 3, [4, 4, 4], [11, 11, 11]),
              tiles(32, 8, 3, [10, 10, 12], [21, 21, 19])]
    ctxs = []
    for _ in range(2):
        ctx = adg.Context(3, 2, 1.4, precision)
        ctx.set_geometry([8, 8, 8], [0, 0, 0], [1, 1, 1], periodic=periodic)
        ctx.set_conservative_reflux(True)
        ctx.set_levels(levels)
        for l in range(3):
            ctx.project_ic(l, 3, [0.3, 1.0, 1.0, 0.125, 0.1])
        ctx.set_limiter(True, 0.0, [1.0] * 3, 1)
        ctx.limit()
        ctx.average_down()
        ctx.fill_ghosts()
        ctx.set_limiter(False, 0.0, [1.0] * 3, 1)
        ctx.set_dt(0.5, 1.0, 0.0)
        ctxs.append(ctx)
    fused, split = ctxs
    for step in range(3):
        fused.step(0.5, 1.0)
        split.ader()
        split.average_down()
        split.fill_ghosts()
        split.advance(0.5, 1.0)
        assert fused.get_time() == split.get_time(), f"step {step}"
        for l in range(3):
            assert np.array_equal(fused.download_Uw(l), split.download_Uw(l)), f"step {step} level {l}"


def test_kh_outflow_y_three_levels(adg, orc):
    """Periodic in x, transmissive (mirror-image) ghosts at y = 0 and 1."""
    print("case KH P=3 16^2 3 levels, outflow in y", flush=True)
    levels = kh_bands(16, 3, 16)
    T = orc.Tables(3, 2)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [1, 1], periodic=[1, 0])
    Uw = orc.project_ic(T, H, 1, [0.01, 1e-3, 7])
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    ctx = adg.Context(2, 3, 1.4, 0)
    ctx.set_geometry([16, 16], [0, 0], [1, 1], periodic=[1, 0])
    ctx.set_levels(levels)
    for l in range(3):
        ctx.upload_Uw(l, Uw[l])
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 1.0, 0.0)
    for _ in range(3):
        ctx.step(0.5, 1.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=3)
    for l in range(3):
        assert rel_err(ctx.download_Uw(l), Uw[l]) < TOL, f"level {l}"


@pytest.mark.parametrize("precision", [0, 1])
def test_nodelane_wave_counts_agree_bitwise(adg, precision):
    """The node-lane predictor with 8 and 16 wavefronts per workgroup gives
    the same bits (the autotuner's choice between them cannot change a
    result); 3 steps of the 3-level KH state."""
    print(f"case node-lane 8 vs 16 waves bitwise, precision {precision}", flush=True)
    levels = kh_bands(16, 3, 16)
    out = []
    for waves in (8, 16):
        ctx = adg.Context(2, 3, 1.4, precision)
        ctx.set_predictor_kernel(3)
        ctx.set_predictor_waves(waves)
        ctx.set_geometry([16, 16], [0, 0], [1, 1])
        ctx.set_levels(levels)
        for l in range(3):
            ctx.project_ic(l, 1, [0.01, 1e-3, 7])
        ctx.average_down()
        ctx.fill_ghosts()
        ctx.set_dt(0.5, 1.0, 0.0)
        for _ in range(3):
            ctx.step(0.5, 1.0)
        out.append([ctx.download_Uw(l) for l in range(3)])
        ctx.close()
    for l in range(3):
        assert np.array_equal(out[0][l], out[1][l]), f"level {l}"


@pytest.mark.parametrize("conservative", [False, True])
def test_reflux_rules_on_a_box_edge(adg, orc, conservative):
    """Both reflux rules against the oracle on a layout where they differ (a
    coarse-fine face on the low edge of coarse boxes): the reference's rule
    drifts, the conservative one keeps the composite totals to rounding.
    The conservative rule is switched on after the levels are set, so the
    register lists are rebuilt on the live layout."""
    print(f"case reflux rule conservative={conservative} vortex P=2 16^2 + two fine boxes", flush=True)
    levels = [tiles(16, 8, 2), [((8, 8), (15, 15)), ((16, 8), (23, 15))]]
    T = orc.Tables(2, 2)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [10, 10], conservative_reflux=conservative)
    Uw = orc.project_ic(T, H, 0, VORTEX)
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    ctx = adg.Context(2, 2, 1.4, 0)
    ctx.set_geometry([16, 16], [0, 0], [10, 10])
    ctx.set_levels(levels)
    for l in range(2):
        ctx.upload_Uw(l, Uw[l])
    if conservative:
        ctx.set_conservative_reflux(True)
    m0 = ctx.totals()
    Dt = orc.set_dt(T, H, Uw, 0.5)
    ctx.set_dt(0.5, 10.0, 0.0)
    for _ in range(3):
        ctx.step(0.5, 10.0)
    orc.evolve(T, H, Uw, 0.0, Dt, 10.0, 0.5, nsteps_max=3)
    for l in range(2):
        v = H.valid_mask(l)
        assert rel_err(ctx.download_Uw(l)[:, :, v], Uw[l][:, :, v]) < TOL, f"level {l}"
    m = ctx.totals()
    drift = max(abs(m[q] - m0[q]) / max(abs(m0[q]), 1.0) for q in range(4))
    assert (drift < 1e-13) if conservative else (drift > 1e-7)
    ctx.close()


@pytest.mark.parametrize("cfl_uncovered", [False, True])
def test_cfl_rules_with_a_bad_covered_cell(adg, orc, cfl_uncovered):
    """set_Dt under both rules against the oracle on a state whose one
    covered coarse cell has negative density at spatial node 0 (as the
    projection of a discontinuity can leave it): the reference's rule
    collapses Dt, the uncovered-cell rule ignores the cell; then steps with
    the step graph, which must pick the rule up."""
    print(f"case CFL rule cfl_uncovered={cfl_uncovered} vortex P=2 16^2 + centre", flush=True)
    levels = two_level_centre(16, 16)
    T = orc.Tables(2, 2)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [10, 10], cfl_uncovered=cfl_uncovered)
    Uw = orc.project_ic(T, H, 0, VORTEX)
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    dt_clean = orc.set_dt(T, H, Uw, 0.5)
    s = int(np.flatnonzero(H.valid_mask(0))[0]) + 18 * 9 + 9   # coarse cell (9, 9), covered
    Uw[0][0, 1, s] = -5.0 * Uw[0][0, 0, s]
    dt_ref = orc.set_dt(T, H, Uw, 0.5)
    assert (dt_ref == dt_clean) if cfl_uncovered else (dt_ref < 0.1 * dt_clean)
    ctx = adg.Context(2, 2, 1.4, 0)
    ctx.set_geometry([16, 16], [0, 0], [10, 10])
    ctx.set_levels(levels)
    for l in range(2):
        ctx.upload_Uw(l, Uw[l])
    ctx.set_dt(0.5, 10.0, 0.0)
    ctx.step_graph(0.5, 10.0, 1)      # a graph captured under the default rule
    for l in range(2):
        ctx.upload_Uw(l, Uw[l])
    ctx.set_cfl_uncovered(cfl_uncovered)
    dt = ctx.set_dt(0.5, 10.0, 0.0)
    assert abs(dt - dt_ref) <= 1e-14 * dt_ref
    ctx.step_graph(0.5, 10.0, 2)
    _, t, Dt = orc.evolve(T, H, Uw, 0.0, dt_ref, 10.0, 0.5, nsteps_max=2)
    t_dev, Dt_dev = ctx.get_time()
    assert abs(t_dev - t) <= 1e-14 * t and abs(Dt_dev - Dt) <= 1e-12 * Dt
    for l in range(2):
        v = H.valid_mask(l)
        assert rel_err(ctx.download_Uw(l)[:, :, v], Uw[l][:, :, v]) < TOL, f"level {l}"
    ctx.close()
