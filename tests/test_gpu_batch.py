"""Batched independent patches: every patch its own periodic domain, all in
one launch per stage.  Each patch of a small uneven batch against the oracle
run of that patch alone (same fixed Dt), and the autotuner."""
import numpy as np
import pytest

from layouts import batch_patches

pytestmark = pytest.mark.gpu

DX = 10.0 / 64


def test_batch_patches_match_single_patch_oracle(adg, orc):
    print("case batch of 3 uneven vortex patches P=3 vs per-patch oracle", flush=True)
    P = 3
    boxes = [((0, 0), (7, 7)), ((8, 0), (19, 15)), ((20, 0), (35, 9))]
    dom = [36, 16]
    rng = np.random.default_rng(5)
    prm = []
    for lo, hi in boxes:
        n = [hi[d] - lo[d] + 1 for d in range(2)]
        prm.append([rng.uniform(1, 5), rng.uniform(0, n[0] * DX), rng.uniform(0, n[1] * DX),
                    rng.uniform(-1, 1), rng.uniform(-1, 1)])
    ctx = adg.Context(2, P, 1.4, 0)
    ctx.set_independent_patches(True)
    ctx.set_geometry(dom, [0, 0], [dom[0] * DX, dom[1] * DX])
    ctx.set_levels([boxes])
    ctx.project_ic_patches(0, 0, prm)
    ctx.fill_ghosts()
    Dt = 0.8 * ctx.set_dt(0.5, 1.0, 0.0)
    ctx.set_dt_value(Dt)
    T = orc.Tables(P, 2)
    refs = []
    for (lo, hi), pp in zip(boxes, prm):
        n = [hi[d] - lo[d] + 1 for d in range(2)]
        H = orc.Hier(2, [[((0, 0), (n[0] - 1, n[1] - 1))]], n, [0, 0], [n[0] * DX, n[1] * DX])
        Uw = orc.project_ic(T, H, 0, pp)
        orc.fill_patch(T, H, Uw)
        refs.append((H, Uw))
    U0 = ctx.download_Uw(0)
    off = 0
    for (lo, hi), (H, Uw) in zip(boxes, refs):
        ns = H.nslot(0)
        assert np.abs(U0[:, :, off:off + ns] - Uw[0]).max() <= 1e-13 * np.abs(Uw[0]).max()
        off += ns
    for _ in range(3):
        ctx.ader()
        ctx.fill_ghosts()
        for H, Uw in refs:
            orc.ader(T, H, Uw, Dt)
            orc.fill_patch(T, H, Uw)
    U = ctx.download_Uw(0)
    off = 0
    for (H, Uw) in refs:
        ns = H.nslot(0)
        v = H.valid_mask(0)
        got, ref = U[:, :, off:off + ns][:, :, v], Uw[0][:, :, v]
        for q in range(4):
            assert np.abs(got[q] - ref[q]).max() <= 1e-12 * np.abs(ref[q]).max(), q
        off += ns


def test_autotuner_picks_a_valid_shape(adg):
    print("case autotune batch of 24 patches 16..48, FP32 and FP64", flush=True)
    levels, dom, _ = batch_patches(24, 16, 48, seed=3, row_width=256)
    rng = np.random.default_rng(3)
    prm = [[rng.uniform(1, 5), 2.0, 2.0, rng.uniform(-1, 1), rng.uniform(-1, 1)] for _ in levels[0]]
    for prec in (0, 1):
        ctx = adg.Context(2, 3, 1.4, prec)
        ctx.set_independent_patches(True)
        ctx.set_geometry(dom, [0, 0], [dom[0] * DX, dom[1] * DX])
        ctx.set_levels(levels)
        ctx.project_ic_patches(0, 0, prm)
        ctx.fill_ghosts()
        ctx.set_dt(0.5, 1.0, 0.0)
        times, choice = ctx.autotune(5)
        print(f"  precision {prec}: {times} -> {choice}", flush=True)
        assert choice[0] in (1, 2, 3) and choice[1] in (8, 16)
        pred = {k: v for k, v in times.items() if not k.startswith("warm")}
        assert min(pred.values()) == pytest.approx(pred[{(1, 8): "valu", (2, 8): "quad8", (3, 8): "nodelane8",
                                                        (3, 16): "nodelane16"}[choice]])
        t0 = ctx.totals()
        ctx.step_graph(0.5, 1.0, 3)
        t1 = ctx.totals()
        assert abs(t1[0] - t0[0]) <= 1e-13 * abs(t0[0])
