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

With --case wave3d: the 3D p=2 network, from runs of the smooth density wave
(initial condition 5) with seeded amplitude and pressure on the periodic
[0,1]^3 with two levels (32^3, the centre half refined).

    python train_warm_start.py --out weights/kh_p3_2d
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


def evaluate(par, seed, ic, P=3, case="kh"):
    if case == "kh":
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
    ap.add_argument("--case", choices=["kh", "wave3d"], default="kh")
    ap.add_argument("--out", default=None)
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--every", type=int, default=20, help="steps between collections")
    ap.add_argument("--per-seed", type=int, default=9, help="collections per seed")
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--schedule", default="20000:2e-3,20000:5e-4,10000:1e-4",
                    help="comma list of steps:lr phases")
    ap.add_argument("--heldout", default="4:1001,4:1002,1:20240611",
                    help="comma list of ic:seed evaluation runs (not used for training)")
    args = ap.parse_args()
    kh = args.case == "kh"
    P = 3 if kh else 2
    if args.out is None:
        args.out = os.path.join(HERE, "weights", "kh_p3_2d" if kh else "wave_p2_3d")
    if not kh and args.heldout == ap.get_default("heldout"):
        args.heldout = "5:1001,5:1002"

    t0 = time.time()
    ctx, levels = kh_context(1) if kh else wave3d_context(1)
    ctx.mlp_init(2024)
    seed, n = 1, 0
    seeds = []
    while n < args.pairs:
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
        evals[hs] = evaluate(par, int(sd), int(ic), P=P, case=args.case)
        print(f"held-out seed {hs}: " + ", ".join(f"{k} {v:.2e}" for k, v in evals[hs].items()), flush=True)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    par.astype(np.float32).tofile(args.out + ".f32")
    sz = ctx.mlp_sizes()
    meta = {"widths": sz, "pairs": n, "train_seeds": seeds, "collect_every_steps": args.every,
            "collect_iterations": 2 * P + 2, "batch": args.batch, "loss_history": hist, "heldout_distance": evals,
            "workload": ("KH, odd seeds IC 4 (multi-mode, amax 0.02), even seeds IC 1 (single mode + noise), "
                         "p=3, 3 levels 64^2/128^2/256^2 bands, FP32 form") if kh else
                        ("3D p=2, periodic [0,1]^3, 32^3 + centre half refined, density wave (IC 5) with "
                         "seeded amplitude and pressure, FP32 form"),
            "seconds": time.time() - t0}
    with open(args.out + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps({"out": args.out, "pairs": n, "final_loss": hist[-1][1]}), flush=True)


if __name__ == "__main__":
    main()
