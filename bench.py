#!/usr/bin/env python3
"""Throughput of the ADER-DG hot path on one MI355X.

Workload (the metric's configuration): 2D compressible Euler (gamma=1.4)
Kelvin-Helmholtz shear layer on periodic [0,1]^2, DG p=3 (Np_s=10,
Np_st=20), three levels 64^2 -> 128^2 -> 256^2 (ratio 2; level 1 covers the
bands |y-1/4|,|y-3/4| < 1/8, level 2 the bands < 1/16), v = 0.01 sin(4 pi x)
plus seeded per-cell uniform noise of amplitude 1e-3.  Inputs are synthetic
and resident in HBM before the timed region.

A step is one iteration of the evolve loop on the device: predictor, face
flux, corrector on every level, flux-register reflux, average down, ghost
fill (same level, periodic, coarse-fine projection), CFL reduction and time
advance.  value = sum over levels of valid elements * Q * Np_st per second.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ader-dg-amr_amd")
sys.path.insert(0, PKG)

METRIC = "space-time DOF-updates/s at p=3, learned+FP32 vs plain FP64, matched L1 error"
PEAK_FP64 = 78.6e12   # MI355X dense FP64 (vector and matrix)
PEAK_FP32 = 157.3e12  # MI355X dense FP32 matrix
PEAK_HBM = 8.0e12
KH_PARAMS = [0.01, 1e-3, 20240611]
# HBM bytes per launch of the (autotuned) node-lane predictor on this workload,
# measured with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE and scaled by
# the counters' reading of a plain 256 MiB device copy (x2.0 reads, x1.0 writes)
PRED_TRAFFIC = {"fp64": 147.2e6, "fp32": 53.1e6, "learned": 53.1e6}


def algorithmic_flops(P, D, Q, Ns, Nst, Mst, Mbd):
    """Per valid element per step (2 flops per FMA): P Picard iterations plus
    the final evaluation, traces, corrector volume and boundary terms."""
    f_iter = 2 * Q * (Mst * Nst + Nst * Ns + D * Nst * Mst + Nst * Nst)
    f_corr = 2 * Q * (Mst * Nst + 2 * D * Mbd * Nst + D * Ns * Mst + 2 * D * Ns * Mbd)
    f_pred_kernel = P * f_iter + f_corr - 2 * Q * 2 * D * Ns * Mbd  # the predictor kernel's share
    return P * f_iter + f_corr, f_pred_kernel


def cpu_baseline(levels, P, nsteps, threads):
    """The oracle (reference-structured FP64 CPU restatement) on the same
    workload, OpenMP over cells, a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    T = O.Tables(P, 2)
    H = O.Hier(2, levels, [64, 64], [0, 0], [1, 1])
    Uw = O.project_ic(T, H, 1, KH_PARAMS)
    O.average_down(T, H, Uw)
    O.fill_patch(T, H, Uw)
    Dt = O.set_dt(T, H, Uw, 0.5)
    t0 = time.perf_counter()
    n, _, _ = O.evolve(T, H, Uw, 0.0, Dt, 1.0, 0.5, nsteps_max=nsteps)
    el = time.perf_counter() - t0
    return n, el


def uncovered_cells(levels):
    """Per level: (slot indices of valid cells not covered by the next level,
    cell volume factor 4^-l), slots of the grown boxes, x fastest."""
    out = []
    for l, lev in enumerate(levels):
        fine = levels[l + 1] if l + 1 < len(levels) else []
        slots, off = [], 0
        for lo, hi in lev:
            g0, g1 = hi[0] - lo[0] + 3, hi[1] - lo[1] + 3
            for j in range(lo[1], hi[1] + 1):
                for i in range(lo[0], hi[0] + 1):
                    cov = any(flo[0] <= 2 * i <= fhi[0] and flo[1] <= 2 * j <= fhi[1] for flo, fhi in fine)
                    if not cov:
                        slots.append(off + (j - lo[1] + 1) * g0 + (i - lo[0] + 1))
            off += g0 * g1
        out.append((slots, 0.25 ** l))
    return out


def accuracy_vs_fp64(make, levels, nsteps, c_dt, Tend):
    """Each form from the same initial state for nsteps steps: composite
    density L1 difference (cell averages, uncovered cells) against the FP64
    form, relative to the FP64 density L1 norm, and relative drift of the
    conserved mass and energy over the run."""
    unc = uncovered_cells(levels)
    fin = {}
    drift = {}
    for form in ("fp64", "fp32", "learned"):
        ctx = make(form)
        t0 = ctx.totals()
        ctx.step_graph(c_dt, Tend, nsteps)
        t1 = ctx.totals()
        fin[form] = [ctx.download_Uw(l)[0, 0] for l in range(len(levels))]
        drift[form] = {"mass": abs(t1[0] - t0[0]) / abs(t0[0]), "energy": abs(t1[-1] - t0[-1]) / abs(t0[-1])}
        ctx.close()
    out = {"steps": nsteps}
    ref_norm = sum(w * sum(abs(fin["fp64"][l][s]) for s in sl) for l, (sl, w) in enumerate(unc))
    for form in ("fp32", "learned"):
        num = sum(w * sum(abs(fin[form][l][s] - fin["fp64"][l][s]) for s in sl) for l, (sl, w) in enumerate(unc))
        out[form] = {"rho_L1_vs_fp64": num / ref_norm, **{f"{k}_drift": v for k, v in drift[form].items()}}
    out["fp64"] = {f"{k}_drift": v for k, v in drift["fp64"].items()}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["fp64", "fp32", "learned"], default="learned",
                    help="form reported as value (all three are measured)")
    ap.add_argument("--tol", type=float, default=1e-4,
                    help="learned form: accept the warm start when max|H_k - H_k-1| <= tol * rho0")
    ap.add_argument("--max-iters", type=int, default=2, help="learned form: Picard iterations from the warm start")
    ap.add_argument("--weights", default=os.path.join(PKG, "weights", "kh_p3_2d.f32"))
    ap.add_argument("--acc-steps", type=int, default=200, help="steps of the accuracy comparison (0 = skip)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-patch measurement")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D density-wave measurement")
    ap.add_argument("--kernel", type=int, default=0, help="predictor kernel: 0 auto, 1 VALU, 2 MFMA")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        from hiptime import copy_probe
        copy_probe(args.copy_probe)
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "reps": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the hot path is single-GPU by design "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    import pyadg
    from hiptime import Event
    from layouts import kh_bands

    P, D = 3, 2
    levels = kh_bands(64, 3, 32)
    n_valid = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) for lev in levels for lo, hi in lev)
    c_dt, Tend = 0.5, 1.0e9   # no clamp inside the timed window

    import numpy as np
    weights = np.fromfile(args.weights, dtype=np.float32)

    def make(form):
        prec = pyadg.ADG_FP64 if form == "fp64" else pyadg.ADG_FP32_MFMA
        ctx = pyadg.Context(D, P, 1.4, prec)
        if args.kernel:
            ctx.set_predictor_kernel(args.kernel)
        ctx.set_geometry([64, 64], [0, 0], [1, 1])
        ctx.set_levels(levels)
        for l in range(len(levels)):
            ctx.project_ic(l, 1, KH_PARAMS)
        ctx.average_down()
        ctx.fill_ghosts()
        ctx.set_dt(c_dt, Tend, 0.0)
        if form == "learned":
            ctx.mlp_set_params(weights)
            ctx.set_learned(True, tol=args.tol, max_iters=args.max_iters)
        return ctx

    def measure(form):
        ctx = make(form)
        tune = ctx.autotune(5)[1] if not args.kernel else None   # launch shapes for this hierarchy
        stream = ctx.stream()
        ctx.step_graph(c_dt, Tend, args.warmup)   # captures the step graph, then warms up
        ctx.sync()
        if form == "learned":
            ctx.learned_stats(reset=True)
        e0, e1 = Event(), Event()
        e0.record(stream)
        ctx.step_graph(c_dt, Tend, args.steps)
        e1.record(stream)
        e1.synchronize()
        ms_step = e0.elapsed_ms(e1) / args.steps
        fell = None
        if form == "learned":
            ne, nf = ctx.learned_stats()
            fell = nf / max(ne, 1)
        # the dominant kernel on its own, same stream, same state
        p0, p1 = Event(), Event()
        kp = max(20, args.steps // 4)
        p0.record(stream)
        for _ in range(kp):
            ctx.predictor()
        p1.record(stream)
        p1.synchronize()
        pred_ms = p0.elapsed_ms(p1) / kp
        dof = n_valid * ctx.Q * ctx.Np_st
        out = {"ms_per_step": ms_step, "value": dof / (ms_step * 1e-3), "pred_ms": pred_ms, "dof": dof,
               "sizes": (ctx.Q, ctx.Np_s, ctx.Np_st, ctx.qMp_st, ctx.qMp_st_bd), "fallback_fraction": fell,
               "autotune": tune}
        ctx.close()
        return out

    def measure_batch():
        """512 independent periodic vortices, edges ~ U{16..128} (seeded),
        beta ~ U[1,5], centres ~ U(patch), background ~ U[-1,1]^2, p=3,
        FP32-MFMA, one launch per stage, autotuned predictor shape."""
        from layouts import batch_patches
        blev, bdom, edges = batch_patches(512, 16, 128, seed=2024)
        dx = 10.0 / 64
        rng = np.random.default_rng(2024)
        prm = [[rng.uniform(1, 5), rng.uniform(0, n * dx), rng.uniform(0, n * dx), rng.uniform(-1, 1),
                rng.uniform(-1, 1)] for n in edges]
        ctx = pyadg.Context(D, P, 1.4, pyadg.ADG_FP32_MFMA)
        ctx.set_independent_patches(True)
        ctx.set_geometry(bdom, [0, 0], [bdom[0] * dx, bdom[1] * dx])
        ctx.set_levels(blev)
        ctx.project_ic_patches(0, 0, prm)
        ctx.fill_ghosts()
        ctx.set_dt(c_dt, Tend, 0.0)
        tune, choice = ctx.autotune(5)
        nb = 20
        ctx.step_graph(c_dt, Tend, 3)
        ctx.sync()
        e0, e1 = Event(), Event()
        e0.record(ctx.stream())
        ctx.step_graph(c_dt, Tend, nb)
        e1.record(ctx.stream())
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / nb
        nel = int(sum(n * n for n in edges))
        ctx.close()
        return {"workload": "512 independent periodic vortices, edges U{16..128}, p=3, FP32-MFMA",
                "elements": nel, "ms_per_step": ms, "value": nel * ctx.Q * ctx.Np_st / (ms * 1e-3),
                "autotune_ms": tune, "autotune_choice": {"kind": choice[0], "waves": choice[1]}}

    def measure_wave3d():
        """3D p=2 density wave (IC 5, amplitude 0.2, pressure 1) on the
        periodic [0,1]^3, 64^3 and the centre half refined (524 288 valid
        elements): FP64, FP32-MFMA and learned + FP32 (quad-gather MFMA
        predictor, network weights/wave_p2_3d trained on other amplitudes
        and pressures)."""
        from layouts import two_level_centre
        lev3 = two_level_centre(64, 32, D=3)
        w3 = np.fromfile(os.path.join(PKG, "weights", "wave_p2_3d.f32"), dtype=np.float32)
        out = {"workload": "3D density wave p=2, periodic [0,1]^3, 64^3 + centre half refined"}
        nel = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lv in lev3 for lo, hi in lv)
        out["elements"] = nel
        for form in ("fp64", "fp32", "learned"):
            prec = pyadg.ADG_FP64 if form == "fp64" else pyadg.ADG_FP32_MFMA
            ctx = pyadg.Context(3, 2, 1.4, prec)
            ctx.set_geometry([64] * 3, [0] * 3, [1] * 3)
            ctx.set_levels(lev3)
            for l in range(len(lev3)):
                ctx.project_ic(l, 5, [0.2, 1.0])
            ctx.average_down()
            ctx.fill_ghosts()
            ctx.set_dt(c_dt, Tend, 0.0)
            if form == "learned":
                ctx.mlp_set_params(w3)
                ctx.set_learned(True, tol=args.tol, max_iters=args.max_iters)
            ctx.autotune(3)
            ctx.step_graph(c_dt, Tend, 3)
            ctx.sync()
            if form == "learned":
                ctx.learned_stats(reset=True)
            nb = 10
            e0, e1 = Event(), Event()
            e0.record(ctx.stream())
            ctx.step_graph(c_dt, Tend, nb)
            e1.record(ctx.stream())
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / nb
            out[f"{form}_ms_per_step"] = ms
            out[f"{form}_value"] = nel * ctx.Q * ctx.Np_st / (ms * 1e-3)
            if form == "learned":
                ne, nf = ctx.learned_stats()
                out["learned_fallback_fraction"] = nf / max(ne, 1)
            ctx.close()
        out["fp32_over_fp64"] = out["fp64_ms_per_step"] / out["fp32_ms_per_step"]
        out["learned_over_fp64"] = out["fp64_ms_per_step"] / out["learned_ms_per_step"]
        return out

    res = {f: measure(f) for f in ("fp64", "fp32", "learned")}
    batch = measure_batch() if not args.no_batch else None
    wave3d = measure_wave3d() if not args.no_3d else None
    head = res[args.form]
    accuracy = accuracy_vs_fp64(make, levels, args.acc_steps, c_dt, Tend) if args.acc_steps > 0 else None
    Q, Ns, Nst, Mst, Mbd = head["sizes"]
    f_step, f_pred = algorithmic_flops(P, D, Q, Ns, Nst, Mst, Mbd)
    achieved = f_pred * n_valid / (head["pred_ms"] * 1e-3)
    peak = PEAK_FP64 if args.form == "fp64" else PEAK_FP32
    roofline = {"bound": "mfma",
                "kernel": f"k_predictor_nl<{'double' if args.form == 'fp64' else 'float'},3,2,16> (P Picard iterations)",
                "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": PRED_TRAFFIC[args.form], "avg_launch_ms": head["pred_ms"],
                "flops_per_launch": f_pred * n_valid}

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        n, el = cpu_baseline(levels, P, args.cpu_steps, threads)
        cpu = {"value": head["dof"] * n / el, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
               "sample": f"{n} full steps of the same 3-level workload ({n_valid} elements) by the FP64 "
                         f"CPU oracle (OpenMP, {threads} threads)"}

    line = {
        "metric": METRIC, "value": head["value"], "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "data": "synthetic",
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "learned": "f32+f64+bf16"}[args.form],
        "config": {"workload": "kelvin-helmholtz 2D p=3, 3 levels 64^2/128^2/256^2 (bands)",
                   "form": args.form, "elements": n_valid, "dof_per_step": head["dof"],
                   "algorithmic_flops_per_step": f_step * n_valid,
                   "fp64_value": res["fp64"]["value"], "fp64_ms_per_step": res["fp64"]["ms_per_step"],
                   "fp32_value": res["fp32"]["value"], "fp32_ms_per_step": res["fp32"]["ms_per_step"],
                   "learned_value": res["learned"]["value"],
                   "learned_ms_per_step": res["learned"]["ms_per_step"],
                   "learned_fallback_fraction": res["learned"]["fallback_fraction"],
                   "learned_tol": args.tol, "learned_max_iters": args.max_iters,
                   "fp32_over_fp64": res["fp32"]["value"] / res["fp64"]["value"],
                   "learned_over_fp64": res["learned"]["value"] / res["fp64"]["value"],
                   "autotune": {f: res[f]["autotune"] for f in res},
                   "accuracy": accuracy,
                   "batch": batch,
                   "wave3d": wave3d,
                   "step": "hipGraph replay of predictor, face flux, corrector, reflux+average down, "
                           "ghost fill+CFL+advance"},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
