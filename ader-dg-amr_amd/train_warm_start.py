#!/usr/bin/env python3
"""Trains the predictor warm-start network on one GPU with the library's own
kernels and writes its parameters (float32, adg_mlp_get_params layout) plus a
JSON record of the data, schedule and held-out evaluation.

Data: (U_w, converged H_w) pairs of every valid element, collected by the
p=3 solver itself (adg_mlp_collect, 2P+2 Picard iterations) along
Kelvin-Helmholtz runs on the 3-level band hierarchy 64^2 -> 128^2 -> 256^2:
odd seeds with the multi-mode perturbation v = sum_{k=1..8} a_k sin(2 pi k x),
a_k ~ U[0, 0.02] (initial condition 4), even seeds with v = 0.01 sin(4 pi x)
plus seeded cell noise of amplitude 1e-3 (initial condition 1).
Evaluation: held-out seeds, distance of the learned start after 1 and 2
Picard iterations to the converged predictor, next to the plain predictor
after 1, 2 and 3 iterations.

With --case kh-job (the committed weights/kh_p3_2d): the pairs come from
whole Kelvin-Helmholtz jobs, t = 0 -> 1, on three levels built and rebuilt
from density-gradient tags (a regrid every 0.05), the TVB limiter (M = 5000)
and the composite rules, one collection after every regrid: odd seeds the
multi-mode state (initial condition 4), even seeds the fixed-grid noise
state (initial condition 7); optionally (--band-pairs) pairs from early
states (a collection every 20 steps from t = 0) on the fixed band hierarchy
of the bench's kernel measurement; held-out seeds are evaluated at t = 0, 0.5 and
0.9 of their own jobs.

With --case shock3d (weights/shock_p2_3d): the 3D p=2 network for shock
tubes, from spherical shock tubes with seeded radius and states on the
outflow [0,1]^3, three levels 16^3 -> 64^3 (--shock-n0) from density-gradient tags
(0.2), a regrid every 16 steps, the TVB limiter (M = 0) on the finest level
and on the initial state, the composite rules, one collection after every
regrid over the first --shock-steps steps.

With --case wave3d: the 3D p=2 network, from runs of the smooth density wave
(initial condition 5) with seeded amplitude and pressure on the periodic
[0,1]^3 with two levels (32^3, the centre half refined).

    python train_warm_start.py --case kh-job --out weights/kh_p3_2d
    python train_warm_start.py --case wave3d --out weights/wave_p2_3d
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import pyadg  # noqa: E402
from layouts import kh_bands, two_level_centre  # noqa: E402

C_DT, T_END = 0.5, 10.0


def wave3d_params(seed):
    """Smooth density wave (IC 5) with seeded amplitude and pressure."""
    rng = np.random.default_rng(9500 + seed)
    return [rng.uniform(0.05, 0.4), rng.uniform(0.5, 2.0)]


def wave3d_project(ctx, levels, seed):
    for l in range(len(levels)):
        ctx.project_ic(l, 5, wave3d_params(seed))
    ctx.average_down()
    ctx.fill_ghosts()
    ctx.set_dt(C_DT, T_END, 0.0)


def wave3d_context(seed, precision=pyadg.ADG_FP32_MFMA, n0=32):
    """3D p=2 on the periodic [0,1]^3, two levels."""
    ctx = pyadg.Context(3, 2, 1.4, precision)
    ctx.set_geometry([n0] * 3, [0] * 3, [1] * 3)
    levels = two_level_centre(n0, 32, D=3)
    ctx.set_levels(levels)
    wave3d_project(ctx, levels, seed)
    return ctx, levels


def kh_context(seed, P=3, n0=64, nlev=3, precision=pyadg.ADG_FP32_MFMA, ic=4):
    ctx = pyadg.Context(2, P, 1.4, precision)
    ctx.set_geometry([n0, n0], [0, 0], [1, 1])
    levels = kh_bands(n0, nlev, 32)
    ctx.set_levels(levels)
    for l in range(len(levels)):
        ctx.project_ic(l, ic, [seed, 0.02] if ic == 4 else [0.01, 1e-3, seed])
    ctx.average_down()
    ctx.fill_ghosts()
    ctx.set_dt(C_DT, T_END, 0.0)
    return ctx, levels


SHOCK_N0 = [16]   # level-0 cells per direction of the shock3d runs


def shock3d_params(seed):
    """Spherical shock tube (IC 3) with seeded radius and states."""
    rng = np.random.default_rng(9700 + seed)
    return [rng.uniform(0.25, 0.4), rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0),
            rng.uniform(0.08, 0.3), rng.uniform(0.05, 0.3)]


def shock3d_reset(ctx, seed, n0=None):
    """Three tagged levels from the seed's initial state (analytic on each,
    limited), in place (the dataset on the device is kept)."""
    from layouts import tiles
    n0 = n0 or SHOCK_N0[0]
    prm = shock3d_params(seed)
    ctx.set_levels([tiles(n0, 32, 3)])
    ctx.project_ic(0, 3, prm)
    ctx.fill_ghosts()
    ctx.set_limiter(True, 0.0, [1.0] * 3, 1)
    for _ in range(2):
        ctx.limit()
        ctx.average_down()
        ctx.fill_ghosts()
        ctx.regrid(0.0)
        for l in range(len(ctx.get_levels())):
            ctx.project_ic(l, 3, prm)
        ctx.average_down()
        ctx.fill_ghosts()
    ctx.limit()
    ctx.average_down()
    ctx.fill_ghosts()
    ctx.set_dt(C_DT, 1.0, 0.0)


def shock3d_context(seed, precision=pyadg.ADG_FP32_MFMA, n0=None):
    n0 = n0 or SHOCK_N0[0]
    ctx = pyadg.Context(3, 2, 1.4, precision)
    ctx.set_geometry([n0] * 3, [0] * 3, [1] * 3, [0, 0, 0])
    ctx.set_conservative_reflux(True)
    ctx.set_cfl_uncovered(True)
    ctx.set_amr(2, max_grid_size=32)
    ctx.set_tagging(2, [0.2])
    shock3d_reset(ctx, seed, n0)
    return ctx


def shock3d_advance(ctx, nsteps, on_regrid=None):
    """nsteps steps with a regrid every 16; on_regrid(ctx) after each."""
    done = 0
    while done < nsteps:
        n = min(16, nsteps - done)
        ctx.step_graph(C_DT, 1.0, n)
        done += n
        t, _ = ctx.get_time()
        ctx.regrid(t)
        if on_regrid:
            on_regrid(ctx)


def kh_job_context(seed, ic, precision=pyadg.ADG_FP32_MFMA, n0=64):
    """Three levels from the initial state's tags (analytic state on each),
    the TVB limiter and both composite rules, as the bench's job."""
    ctx = pyadg.Context(2, 3, 1.4, precision)
    ctx.set_geometry([n0, n0], [0, 0], [1, 1])
    ctx.set_conservative_reflux(True)
    ctx.set_cfl_uncovered(True)
    ctx.set_amr(2, max_grid_size=32)
    ctx.set_tagging(2, [0.05])
    kh_job_reset(ctx, seed, ic, n0)
    return ctx


def kh_job_reset(ctx, seed, ic, n0=64):
    """The job's initial hierarchy and state for another seed, in place (the
    dataset on the device is kept)."""
    from layouts import tiles
    prm = [seed, 0.02] if ic == 4 else [0.01, 1e-3, 7000 + seed, n0]
    ctx.set_levels([tiles(n0, 32, 2)])
    ctx.project_ic(0, ic, prm)
    ctx.fill_ghosts()
    ctx.set_limiter(True, 5000.0, [1.0] * 3, 1)
    for _ in range(2):
        ctx.regrid(0.0)
        for l in range(len(ctx.get_levels())):
            ctx.project_ic(l, ic, prm)
        ctx.average_down()
        ctx.fill_ghosts()
    ctx.set_dt(C_DT, 1.0, 0.0)


def kh_job_advance(ctx, t_stop, on_regrid=None):
    """Steps the job to t_stop with a regrid every 0.05 (graph batches that
    do not cross a regrid); on_regrid(ctx) after every regrid."""
    t, Dt = ctx.get_time()
    next_rg = (int(t / 0.05 + 1e-9) + 1) * 0.05
    while t < t_stop - 1e-12:
        room = 0.5 * min(t_stop - t, next_rg - t) / Dt
        n = int(min(64, room))
        if n >= 2:
            ctx.step_graph(C_DT, 1.0, n)
        else:
            ctx.step(C_DT, 1.0)
        t, Dt = ctx.get_time()
        if t >= next_rg - 1e-12 and t < 1.0 - 1e-12:
            ctx.regrid(t)
            next_rg = t + 0.05
            if on_regrid:
                on_regrid(ctx)


def valid_slots(levels, lev):
    """Slot indices of the valid cells of a level (grown boxes, x fastest)."""
    out, off = [], 0
    for lo, hi in levels[lev]:
        g = [hi[d] - lo[d] + 3 for d in range(len(lo))]
        inner = np.ix_(*[np.arange(1, n - 1) for n in reversed(g)])
        idx = np.arange(int(np.prod(g))).reshape(tuple(reversed(g)))[inner]
        out.extend((off + idx).ravel().tolist())
        off += int(np.prod(g))
    return np.array(out)


def evaluate(par, seed, ic, P=3, case="kh", t_at=0.0):
    if case == "kh-job":
        ctx = kh_job_context(seed, ic, precision=pyadg.ADG_FP64)
        if t_at > 0:
            kh_job_advance(ctx, t_at)
        levels = ctx.get_levels()
    elif case == "shock3d":
        ctx = shock3d_context(seed, precision=pyadg.ADG_FP64)
        if t_at > 0:
            shock3d_advance(ctx, int(t_at))
        levels = ctx.get_levels()
    elif case == "kh":
        ctx, levels = kh_context(seed, precision=pyadg.ADG_FP64, ic=ic)
    else:
        ctx, levels = wave3d_context(seed, precision=pyadg.ADG_FP64)
    ctx.mlp_set_params(par)
    ctx.keep_Hw(True)
    res = {}
    ctx.mlp_collect(2 * P + 2)
    conv = [ctx.download_Hw(l) for l in range(len(levels))]
    for it in (1, 2, 3):
        ctx.mlp_collect(it)
        res[f"plain_{it}"] = [ctx.download_Hw(l) for l in range(len(levels))]
    for it in (1, 2):
        ctx.set_learned(2, max_iters=it)
        ctx.predictor(-1, 1)
        res[f"learned_{it}"] = [ctx.download_Hw(l) for l in range(len(levels))]
    out = {}
    for k, H in res.items():
        num = max(np.abs(H[l][:, :, valid_slots(levels, l)] - conv[l][:, :, valid_slots(levels, l)]).max()
                  for l in range(len(levels)))
        den = max(np.abs(conv[l]).max() for l in range(len(levels)))
        out[k] = num / den
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", choices=["kh", "kh-job", "wave3d", "shock3d"], default="kh-job")
    ap.add_argument("--shock-steps", type=int, default=128, help="shock3d: steps per seed")
    ap.add_argument("--shock-n0", type=int, default=16, help="shock3d: level-0 cells per direction")
    ap.add_argument("--out", default=None)
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--band-pairs", type=int, default=0,
                    help="kh-job: pairs from early states on the band hierarchy, after the job pairs")
    ap.add_argument("--every", type=int, default=20, help="steps between collections")
    ap.add_argument("--per-seed", type=int, default=9, help="collections per seed")
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--schedule", default="20000:2e-3,20000:5e-4,10000:1e-4",
                    help="comma list of steps:lr phases")
    ap.add_argument("--heldout", default="4:1001,4:1002,1:20240611",
                    help="comma list of ic:seed evaluation runs (not used for training)")
    args = ap.parse_args()
    kh = args.case in ("kh", "kh-job")
    P = 3 if kh else 2
    if args.out is None:
        args.out = os.path.join(HERE, "weights", "kh_p3_2d" if kh else
                                "shock_p2_3d" if args.case == "shock3d" else "wave_p2_3d")
    if args.case == "shock3d" and args.heldout == ap.get_default("heldout"):
        args.heldout = "3:1001,3:1002"
    if args.case == "wave3d" and args.heldout == ap.get_default("heldout"):
        args.heldout = "5:1001,5:1002"
    if args.case == "kh-job" and args.heldout == ap.get_default("heldout"):
        args.heldout = "4:1001,7:1002"

    t0 = time.time()
    seed, n = 1, 0
    seeds = []
    if args.case == "shock3d":
        SHOCK_N0[0] = args.shock_n0
        ctx = shock3d_context(1)
        ctx.mlp_init(2024)
        state = {"n": 0}

        def collect3(c):
            if state["n"] < args.pairs:
                state["n"] = c.mlp_collect(2 * P + 2)

        while state["n"] < args.pairs:
            if seed > 1:
                shock3d_reset(ctx, seed)
            # no collection of the initial jump: its Picard iteration has not
            # converged after 2P + 2 iterations, so its targets are not H_w*
            shock3d_advance(ctx, args.shock_steps, on_regrid=collect3)
            seeds.append(seed)
            print(f"seed {seed}: {state['n']} pairs ({time.time() - t0:.1f} s)", flush=True)
            seed += 1
        n = state["n"]
    elif args.case == "kh-job":
        ctx = kh_job_context(1, 4)
        ctx.mlp_init(2024)
        state = {"n": 0}

        def collect(c):
            state["n"] = c.mlp_collect(2 * P + 2)

        while state["n"] < args.pairs:
            if seed > 1:
                kh_job_reset(ctx, seed, 4 if seed % 2 else 7)
            collect(ctx)
            kh_job_advance(ctx, 1.0, on_regrid=lambda c: collect(c) if state["n"] < args.pairs else None)
            seeds.append(seed)
            print(f"seed {seed}: {state['n']} pairs ({time.time() - t0:.1f} s)", flush=True)
            seed += 1
        # early states on the fixed band hierarchy of the bench's kernel
        # measurement, collected every --every steps from t = 0
        levels = kh_bands(64, 3, 32)
        target = state["n"] + args.band_pairs
        while state["n"] < target:
            ic = 4 if seed % 2 else 7
            prm = [seed, 0.02] if ic == 4 else [0.01, 1e-3, 7000 + seed, 64]
            ctx.set_levels(levels)
            for l in range(len(levels)):
                ctx.project_ic(l, ic, prm)
            ctx.average_down()
            ctx.fill_ghosts()
            ctx.set_dt(C_DT, T_END, 0.0)
            for _ in range(args.per_seed):
                collect(ctx)
                ctx.step_graph(C_DT, T_END, args.every)
                if state["n"] >= target:
                    break
            seeds.append(seed)
            print(f"seed {seed} (bands): {state['n']} pairs ({time.time() - t0:.1f} s)", flush=True)
            seed += 1
        n = state["n"]
    else:
        ctx, levels = kh_context(1) if kh else wave3d_context(1)
        ctx.mlp_init(2024)
    while args.case not in ("kh-job", "shock3d") and n < args.pairs:
        if not kh and seed > 1:
            wave3d_project(ctx, levels, seed)
        elif seed > 1:   # same hierarchy, next seed (the dataset stays on the device)
            for l in range(len(levels)):
                # odd seeds: multi-mode perturbation; even seeds: single mode
                # plus seeded cell noise (initial condition 1, other seeds
                # than any evaluation run)
                if seed % 2:
                    ctx.project_ic(l, 4, [seed, 0.02])
                else:
                    ctx.project_ic(l, 1, [0.01, 1e-3, 7000 + seed])
            ctx.average_down()
            ctx.fill_ghosts()
            ctx.set_dt(C_DT, T_END, 0.0)
        for _ in range(args.per_seed):
            n = ctx.mlp_collect(2 * P + 2)
            ctx.step_graph(C_DT, T_END, args.every)
            if n >= args.pairs:
                break
        seeds.append(seed)
        print(f"seed {seed}: {n} pairs ({time.time() - t0:.1f} s)", flush=True)
        seed += 1
    ctx.mlp_fit_normalization()
    loss0 = ctx.mlp_train(1, args.batch, 0.0)
    hist = [("init", loss0)]
    for k, ph in enumerate(args.schedule.split(",")):
        ns, lr = ph.split(":")
        loss = ctx.mlp_train(int(ns), args.batch, float(lr), seed=100 + k)
        hist.append((ph, loss))
        print(f"phase {ph}: loss {loss:.3e} ({time.time() - t0:.1f} s)", flush=True)
    par = ctx.mlp_params()
    evals = {}
    for hs in args.heldout.split(","):
        ic, sd = hs.split(":")
        for t_at in ((0.0, 0.5, 0.9) if args.case == "kh-job" else (0.0, 64, 128) if args.case == "shock3d"
                     else (0.0,)):
            key = hs if t_at == 0.0 else (f"{hs}@step={int(t_at)}" if args.case == "shock3d" else f"{hs}@t={t_at}")
            evals[key] = evaluate(par, int(sd), int(ic), P=P, case=args.case, t_at=t_at)
            print(f"held-out seed {key}: " + ", ".join(f"{k} {v:.2e}" for k, v in evals[key].items()), flush=True)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    par.astype(np.float32).tofile(args.out + ".f32")
    sz = ctx.mlp_sizes()
    meta = {"widths": sz, "pairs": n, "train_seeds": seeds, "collect_every_steps": args.every,
            "collect_iterations": 2 * P + 2, "batch": args.batch, "loss_history": hist, "heldout_distance": evals,
            "workload": ("KH whole jobs t=0..1, odd seeds IC 4 (multi-mode, amax 0.02), even seeds IC 7 "
                         "(single mode + fixed-grid noise), p=3, 3 tagged levels 64^2/128^2/256^2, regrid every "
                         "0.05, TVB M=5000, collection after every regrid"
                         + (f"; then {args.band_pairs} pairs from early states on the 3-level band hierarchy"
                            if args.band_pairs else "") + ", FP32 form") if args.case == "kh-job" else
                        ("KH, odd seeds IC 4 (multi-mode, amax 0.02), even seeds IC 1 (single mode + noise), "
                         "p=3, 3 levels 64^2/128^2/256^2 bands, FP32 form") if kh else
                        ("3D p=2 spherical shock tubes (IC 3) with seeded radius U[0.25,0.4] and states, outflow "
                         f"[0,1]^3, 3 tagged levels from {args.shock_n0}^3, regrid every 16 steps, TVB M=0, collection "
                         f"after every regrid over {args.shock_steps} steps, FP32 form") if args.case == "shock3d" else
                        ("3D p=2, periodic [0,1]^3, 32^3 + centre half refined, density wave (IC 5) with "
                         "seeded amplitude and pressure, FP32 form"),
            "seconds": time.time() - t0}
    with open(args.out + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps({"out": args.out, "pairs": n, "final_loss": hist[-1][1]}), flush=True)


if __name__ == "__main__":
    main()
