#!/usr/bin/env python3
"""Throughput of the ADER-DG hot path on one MI355X.

Workload of the metric (the headline): 2D compressible Euler (gamma=1.4)
Kelvin-Helmholtz shear layer on the periodic [0,1]^2, DG p=3 (Np_s=10,
Np_st=20), three AMR levels 64^2 -> 128^2 -> 256^2 (ratio 2) built and
rebuilt from density-gradient tags (criterion |d rho| > 0.05 between face
neighbours, one buffer cell, blocking factor 4, boxes of at most 32^2, a
regrid every 0.05 time units), the TVB limiter on the finest level every
step (M = 5000, the reference's Kelvin-Helmholtz setting), v = 0.01 sin(4 pi
x) plus seeded uniform noise of amplitude 1e-3 constant on the cells of the
64^2 grid, run from t = 0 to t = 1.  The two composite rules that are not the
reference's are on: the conservative reflux (every coarse cell the flux
registers reach is corrected) and the time step from uncovered cells.
Inputs are synthetic and resident in HBM before a timed region starts.

A step is one iteration of the evolve loop on the device (predictor, face
flux, corrector, reflux, limiter, average down, ghost fill, CFL reduction
and time advance; one hipGraph replay) plus, when one is due, the regrid
between steps (device tags, host grid generation, remake on the device).
value = space-time DOF-updates per second of the learned + FP32 form over
exactly --steps steps after --warmup steps from t = 0: the sum over the
timed steps of valid elements x Q x Np_st, over the time of those steps.
The whole job of every form (t = 0 -> 1) and its density L1 error against a
finer FP64 run of the same initial state are in config.

Prints ONE JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ader-dg-amr_amd")
sys.path.insert(0, PKG)

METRIC = "space-time DOF-updates/s at p=3, learned+FP32 vs plain FP64, matched L1 error"
PEAK_FP64 = 78.6e12   # MI355X dense FP64 (vector and matrix)
PEAK_FP32 = 157.3e12  # MI355X dense FP32 matrix
PEAK_HBM = 8.0e12
KH_SEED = 20240611
# HBM bytes per launch of the node-lane predictor on the 28 672-element band
# hierarchy (the kernel measurement below), from rocprofv3 --pmc FETCH_SIZE
# and --pmc WRITE_SIZE scaled by the counters' reading of a plain 256 MiB
# device copy (x2.0 reads, x1.0 writes)
PRED_TRAFFIC = {"fp64": 98.1e6, "fp32": 56.8e6, "learned": 56.8e6}
SHOCK = [0.4, 1.0, 1.0, 0.125, 0.1]


def algorithmic_flops(P, D, Q, Ns, Nst, Mst, Mbd):
    """Per valid element per step (2 flops per FMA): P Picard iterations plus
    the final evaluation, traces, corrector volume and boundary terms."""
    f_iter = 2 * Q * (Mst * Nst + Nst * Ns + D * Nst * Mst + Nst * Nst)
    f_corr = 2 * Q * (Mst * Nst + 2 * D * Mbd * Nst + D * Ns * Mst + 2 * D * Ns * Mbd)
    f_pred_kernel = P * f_iter + f_corr - 2 * Q * 2 * D * Ns * Mbd  # the predictor kernel's share
    return P * f_iter + f_corr, f_pred_kernel


def log(msg):
    """Progress on stderr (stdout carries only the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def nvalid_of(levels, D):
    n = 0
    for lev in levels:
        for lo, hi in lev:
            k = 1
            for d in range(D):
                k *= hi[d] - lo[d] + 1
            n += k
    return n


class Job:
    """One run of the evolve loop with regrids, driven in batches of steps:
    replays of the captured step graph (16 steps per graph, leftover steps
    as stream launches) or, with graph=False, stream launches; no host
    synchronisation inside a batch.

    Regrids are due every regrid_dt time units (the reference's dt_regrid)
    or every regrid_steps steps (dtn_regrid); a batch never crosses a due
    regrid, the end time or a step budget.  dof counts valid elements x Q x
    Np_st for every step taken."""

    def __init__(self, ctx, D, c_dt, T, regrid_dt=None, regrid_steps=None, max_batch=64, name=None, graph=True):
        self.ctx, self.D, self.c_dt, self.T = ctx, D, c_dt, T
        self.graph = graph
        if graph:
            ctx.set_graph_unroll(16)
        self.name = name
        self.last_log = time.perf_counter()
        self.regrid_dt, self.regrid_steps, self.max_batch = regrid_dt, regrid_steps, max_batch
        self.t, self.Dt = ctx.get_time()
        self.steps = 0
        self.dof = 0.0
        self.regrids = 0
        self.next_rg = self.t + regrid_dt if regrid_dt else None
        self.since_rg = 0
        self.levels = ctx.get_levels()
        self.nvalid = nvalid_of(self.levels, D)
        self.max_nvalid = self.nvalid
        self.per_elem = ctx.Q * ctx.Np_st

    def done(self):
        return self.t >= self.T - 1e-12

    def _maybe_regrid(self):
        due = (self.regrid_dt and self.t >= self.next_rg - 1e-12) or \
              (self.regrid_steps and self.since_rg >= self.regrid_steps)
        if due and not self.done():
            self.ctx.regrid(self.t)
            self.regrids += 1
            self.since_rg = 0
            if self.regrid_dt:
                self.next_rg = self.t + self.regrid_dt
            self.levels = self.ctx.get_levels()
            self.nvalid = nvalid_of(self.levels, self.D)
            self.max_nvalid = max(self.max_nvalid, self.nvalid)

    def run(self, nsteps=None):
        """Advances nsteps steps (or to T); returns the steps taken."""
        taken = 0
        while not self.done() and (nsteps is None or taken < nsteps):
            self._maybe_regrid()
            room = 0.5 * (self.T - self.t) / self.Dt
            if self.regrid_dt:
                room = min(room, 0.5 * (self.next_rg - self.t) / self.Dt)
            batch = int(min(self.max_batch, max(room, 0)))
            if self.regrid_steps:
                batch = min(batch, self.regrid_steps - self.since_rg)
            if nsteps is not None:
                batch = min(batch, nsteps - taken)
            if batch >= 2:
                if self.graph:
                    self.ctx.step_graph(self.c_dt, self.T, batch)
                else:
                    self.ctx.steps(self.c_dt, self.T, batch)
            else:
                batch = 1
                self.ctx.step(self.c_dt, self.T)
            taken += batch
            self.steps += batch
            self.since_rg += batch
            self.dof += batch * self.nvalid * self.per_elem
            self.t, self.Dt = self.ctx.get_time()
            if self.name and time.perf_counter() - self.last_log > 20:
                log(f"{self.name}: t={self.t:.4f} steps={self.steps} elements={self.nvalid} Dt={self.Dt:.2e}")
                self.last_log = time.perf_counter()
        return taken


def build_tagged(pyadg, D, P, prec, n0, ic, prm, periodic, max_grid_size, amr_c, limiter_M, limit_ic):
    """Three levels from the initial state's tags, the analytic state
    projected on each (the reference's InitFromScratch in validation mode),
    both composite rules on, the TVB limiter on the finest level."""
    from layouts import tiles
    ctx = pyadg.Context(D, P, 1.4, prec)
    ctx.set_geometry([n0] * D, [0] * D, [1] * D, periodic)
    ctx.set_conservative_reflux(True)
    ctx.set_cfl_uncovered(True)
    ctx.set_levels([tiles(n0, 32, D)])
    ctx.project_ic(0, ic, prm)
    ctx.fill_ghosts()
    ctx.set_amr(2, max_grid_size=max_grid_size)
    ctx.set_tagging(2, [amr_c])
    ctx.set_limiter(True, limiter_M, [1.0] * 3, 1)
    for _ in range(2):
        if limit_ic:
            ctx.limit()
            ctx.average_down()
            ctx.fill_ghosts()
        ctx.regrid(0.0)
        for l in range(len(ctx.get_levels())):
            ctx.project_ic(l, ic, prm)
        ctx.average_down()
        ctx.fill_ghosts()
    if limit_ic:
        ctx.limit()
        ctx.average_down()
        ctx.fill_ghosts()
    ctx.limiter_stats(reset=True)
    return ctx


def level0_means(ctx, n0):
    """Composite cell means on level 0 (after the average down), [q][y][x]."""
    import numpy as np
    U = ctx.download_Uw(0)
    lv = ctx.get_levels()[0]
    out = np.zeros((U.shape[0], n0, n0))
    off = 0
    for lo, hi in lv:
        g = [hi[d] - lo[d] + 3 for d in range(2)]
        blk = U[:, 0, off:off + g[0] * g[1]].reshape(U.shape[0], g[1], g[0])[:, 1:-1, 1:-1]
        out[:, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = blk
        off += g[0] * g[1]
    return out


def native_oracle():
    """The oracle built for this host's vector ISA (oracle/Makefile native);
    the committed build if that fails.  Returns (path, kind)."""
    out = f"/tmp/liboracle_native_{os.getpid()}.so"
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVE_OUT={out}"],
                       check=True, capture_output=True, timeout=300)
        return out, "-O3 -march=native -fopenmp -ffp-contract=off"
    except Exception:
        return os.path.join(ROOT, "oracle", "liboracle.so"), "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(levels, nsteps):
    """The oracle (reference-structured FP64 CPU restatement, OpenMP over
    cells) on the headline workload's initial hierarchy: the same state, the
    TVB limiter and the two composite rules, steps timed one by one; the
    median step."""
    import numpy as np
    path, build = native_oracle()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    O.lib(path)
    T = O.Tables(3, 2)
    H = O.Hier(2, levels, [64, 64], [0, 0], [1, 1], conservative_reflux=True, cfl_uncovered=True)
    Uw = O.project_ic(T, H, 7, [0.01, 1e-3, KH_SEED, 64])
    O.average_down(T, H, Uw)
    O.fill_patch(T, H, Uw)
    Dt = O.set_dt(T, H, Uw, 0.5)
    t, times = 0.0, []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        _, t, Dt, _ = O.evolve_lim(T, H, Uw, t, Dt, 1.0, 0.5, 5000.0, [1.0] * 3, 1, nsteps_max=1)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), build


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--T", type=float, default=1.0, help="end time of the Kelvin-Helmholtz job")
    ap.add_argument("--form", choices=["fp64", "fp32", "learned"], default="learned",
                    help="form reported as value (all three are measured)")
    ap.add_argument("--tol", type=float, default=1e-4,
                    help="learned form: accept the warm start when max|H_k - H_k-1| <= tol * rho0")
    ap.add_argument("--max-iters", type=int, default=2, help="learned form: Picard iterations from the warm start")
    ap.add_argument("--weights", default=os.path.join(PKG, "weights", "kh_p3_2d.f32"))
    ap.add_argument("--weights3d", default=os.path.join(PKG, "weights", "wave_p2_3d.f32"),
                    help="3D shock tube: the learned form's network")
    ap.add_argument("--no-ref", action="store_true", help="skip the finer FP64 reference run (no L1 error)")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-patch measurement")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D shock-tube measurement")
    ap.add_argument("--no-vortex", action="store_true", help="skip the vortex accuracy case")
    ap.add_argument("--shock-steps", type=int, default=128,
                    help="3D shock tube: steps timed after 16 warmup steps (0 with --shock-T: the whole job)")
    ap.add_argument("--shock-T", type=float, default=0.25, help="3D shock tube: end time of a whole-job run")
    ap.add_argument("--shock-forms", default="fp64,fp32,learned", help="3D shock tube: the forms to run")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the predictor measurement on the band hierarchy (the profiled command)")
    ap.add_argument("--shock-only", action="store_true", help="only the 3D shock-tube measurement")
    ap.add_argument("--pred-shape", default=None, metavar="KIND,WAVES",
                    help="kernel measurement: pin the predictor launch shape (e.g. 3,16) instead of autotuning; "
                         "counter runs use it, since the tuner's timings change under counter collection")
    ap.add_argument("--no-graph", action="store_true",
                    help="step the jobs one step per call (the profiled commands)")
    ap.add_argument("--stream", action="store_true",
                    help="run the jobs' batches of steps as stream launches instead of replays of the captured "
                         "step graph")
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

    import numpy as np
    import pyadg
    from hiptime import Event
    from layouts import kh_bands

    P, D = 3, 2
    c_dt = 0.5
    KH = [0.01, 1e-3, KH_SEED, 64]
    weights = np.fromfile(args.weights, dtype=np.float32)
    prec_of = {"fp64": pyadg.ADG_FP64, "fp32": pyadg.ADG_FP32_MFMA, "learned": pyadg.ADG_FP32_MFMA}

    def kh_ctx(form, n0=64):
        ctx = build_tagged(pyadg, 2, P, prec_of[form], n0, 7, KH, [1, 1], 32, 0.05, 5000.0, False)
        if form == "learned":
            ctx.mlp_set_params(weights)
            ctx.set_learned(True, tol=args.tol, max_iters=args.max_iters)
        ctx.set_dt(c_dt, args.T, 0.0)
        return ctx

    def time_launches(ctx, fn, reps):
        s = ctx.stream()
        e0, e1 = Event(), Event()
        fn()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    def kernel_measurement(form):
        """The predictor and the bandwidth stages on the fixed 3-level band
        hierarchy (28 672 elements), the library stream, HIP events."""
        levels = kh_bands(64, 3, 32)
        ctx = pyadg.Context(D, P, 1.4, prec_of[form])
        ctx.set_geometry([64, 64], [0, 0], [1, 1])
        ctx.set_levels(levels)
        for l in range(len(levels)):
            ctx.project_ic(l, 7, KH)
        ctx.average_down()
        ctx.fill_ghosts()
        ctx.set_dt(c_dt, 1e9, 0.0)
        if form == "learned":
            ctx.mlp_set_params(weights)
            ctx.set_learned(True, tol=args.tol, max_iters=args.max_iters)
        if args.pred_shape:
            tune = [int(x) for x in args.pred_shape.split(",")]
            ctx.set_predictor_kernel(tune[0])
            ctx.set_predictor_waves(tune[1])
        else:
            tune = ctx.autotune(5)[1]
        ctx.step_graph(c_dt, 1e9, 10)
        reps = 50
        out = {"autotune": tune, "elements": nvalid_of(levels, D)}
        out["pred_ms"] = time_launches(ctx, lambda: ctx.predictor(), reps)
        out["faceflux_ms"] = time_launches(ctx, lambda: ctx.face_flux(), reps)
        out["corrector_ms"] = time_launches(ctx, lambda: ctx.corrector(), reps)
        out["step_ms"] = time_launches(ctx, lambda: ctx.steps(c_dt, 1e9, 1), reps)
        out["step_graph1_ms"] = time_launches(ctx, lambda: ctx.step_graph(c_dt, 1e9, 1), reps)   # 1 step per graph
        ctx.set_graph_unroll(16)
        ctx.step_graph(c_dt, 1e9, 16)
        out["step_graph_ms"] = time_launches(ctx, lambda: ctx.step_graph(c_dt, 1e9, 16), 8) / 16
        nslot = sum(ctx.nslot(l) for l in range(len(levels)))
        nface = sum(ctx.nface(l, d) for l in range(len(levels)) for d in range(D))
        nfine = sum(nvalid_of([levels[l]], D) for l in range(1, len(levels)))
        out.update(nslot=nslot, nface=nface, nfine=nfine,
                   sizes=(ctx.Q, ctx.Np_s, ctx.Np_st, ctx.qMp_st, ctx.qMp_st_bd), esz=8 if form == "fp64" else 4)
        ctx.close()
        return out

    def stage_bandwidth(k):
        """Algorithmic bytes over measured time for the bandwidth stages."""
        Q, Ns, Nst, Mst, Mbd = k["sizes"]
        esz, nv = k["esz"], k["elements"]
        ff = k["nface"] * 3 * Q * Mbd * esz                      # two traces read, one flux written
        corr = nv * (2 * Q * Ns * 8 + Q * Ns * esz + 2 * D * Q * Mbd * esz)
        nghost = k["nslot"] - nv
        epi = (nv * Q * Ns * 8                                   # CFL reads U_w
               + nghost * 2 * Q * Ns * 8                         # ghost fill
               + (k["nfine"] // 4) * 5 * Q * Ns * 8)             # average down: 4 children + parent
        epi_ms = k["step_ms"] - k["pred_ms"] - k["faceflux_ms"] - k["corrector_ms"]
        return {"face_flux": {"GB/s": ff / (k["faceflux_ms"] * 1e-3) / 1e9, "bytes": ff, "ms": k["faceflux_ms"]},
                "corrector": {"GB/s": corr / (k["corrector_ms"] * 1e-3) / 1e9, "bytes": corr, "ms": k["corrector_ms"]},
                "epilogue": {"GB/s": epi / (epi_ms * 1e-3) / 1e9, "bytes": epi, "ms": epi_ms,
                             "note": "reflux, limiter, average down, ghost fill, CFL: step minus the other stages"},
                "peak_GB/s": PEAK_HBM / 1e9}

    if args.kernel_only:
        k = kernel_measurement(args.form)
        print(json.dumps({"kernel_only": True, "form": args.form, **{x: k[x] for x in k if x != "sizes"}}), flush=True)
        return

    def kh_job(form):
        ctx = kh_ctx(form)
        tune = ctx.autotune(5)[1]
        ctx.regrid_stats(reset=True)
        s = ctx.stream()
        job = Job(ctx, D, c_dt, args.T, regrid_dt=0.05, max_batch=1 if args.no_graph else 64, graph=not args.stream)
        levels0 = job.levels
        m0 = ctx.totals()
        j0, e0, e1, j1 = Event(), Event(), Event(), Event()
        w0 = time.perf_counter()
        j0.record(s)
        job.run(args.warmup)
        if form == "learned":
            ctx.learned_stats(reset=True)
        dof0, st0 = job.dof, job.steps
        e0.record(s)
        taken = job.run(args.steps)
        e1.record(s)
        e1.synchronize()
        timed_ms = e0.elapsed_ms(e1)
        timed_dof = job.dof - dof0
        fb = ctx.learned_stats() if form == "learned" else None
        job.run()
        j1.record(s)
        j1.synchronize()
        wall = time.perf_counter() - w0
        job_ms = j0.elapsed_ms(j1)
        m1 = ctx.totals()
        out = {"timed_steps": taken, "timed_ms": timed_ms, "value": timed_dof / (timed_ms * 1e-3),
               "ms_per_step": timed_ms / max(taken, 1),
               "job": {"t_end": job.t, "steps": job.steps, "ms": job_ms, "wall_s": wall,
                       "value": job.dof / (job_ms * 1e-3), "regrids": job.regrids,
                       "elements_start": nvalid_of(levels0, D), "elements_end": job.nvalid,
                       "elements_max": job.max_nvalid, "regrid_host_s": ctx.regrid_stats(),
                       "cells_limited": ctx.limiter_stats(),
                       "mass_drift": abs(m1[0] - m0[0]) / abs(m0[0]),
                       "energy_drift": abs(m1[-1] - m0[-1]) / abs(m0[-1])},
               "fallback_fraction_timed": (fb[1] / max(fb[0], 1)) if fb else None,
               "autotune": tune, "levels0": levels0}
        if form == "learned":
            ne, nf = ctx.learned_stats()
            out["job"]["fallback_fraction"] = nf / max(ne, 1)
        out["rho0"] = level0_means(ctx, 64)[0]
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
        ctx.set_dt(c_dt, 1e9, 0.0)
        tune, choice = ctx.autotune(5)
        ctx.steps(c_dt, 1e9, 3)
        ms = time_launches(ctx, lambda: ctx.steps(c_dt, 1e9, 1), 20)
        nel = int(sum(n * n for n in edges))
        Q, Nst = ctx.Q, ctx.Np_st
        ctx.close()
        return {"workload": "512 independent periodic vortices, edges U{16..128}, p=3, FP32-MFMA",
                "elements": nel, "ms_per_step": ms, "value": nel * Q * Nst / (ms * 1e-3),
                "autotune_choice": {"kind": choice[0], "waves": choice[1]}}

    def measure_vortex():
        """The plain method's case: isentropic vortex p=2 on the periodic
        [0,10]^2, 32^2, CFL 0.5, t = 0 -> 1, FP64 and FP32 against the exact
        advected vortex."""
        from layouts import single_level
        V = [5.0, 5.0, 5.0, 1.0, 1.0]
        out = {"workload": "isentropic vortex p=2, 32^2, u_inf=(1,1), t=1"}
        for form in ("fp64", "fp32"):
            ctx = pyadg.Context(2, 2, 1.4, prec_of[form])
            ctx.set_geometry([32, 32], [0, 0], [10, 10])
            ctx.set_levels(single_level(32, 16))
            ctx.project_ic(0, 0, V)
            ctx.fill_ghosts()
            ctx.set_dt(c_dt, 1.0, 0.0)
            m0 = ctx.totals()
            t, n = 0.0, 0
            while t < 1.0 - 1e-12:
                t = ctx.step(c_dt, 1.0, want_t=True)
                n += 1
            m1 = ctx.totals()
            out[form] = {"steps": n, "t": t, "rho_L1": ctx.lp_error(0, 1, 0, V, t),
                         "rho_L2": ctx.lp_error(0, 2, 0, V, t), "mass_drift": abs(m1[0] - m0[0]) / m0[0]}
            ctx.close()
        out["fp32_over_fp64_L1"] = out["fp32"]["rho_L1"] / out["fp64"]["rho_L1"]
        return out

    def measure_shock3d():
        """3D spherical shock tube on [0,1]^3, outflow boundaries, p=2, three
        levels 64^3 -> 128^3 -> 256^3 from density-gradient tags (0.2),
        regrid every 16 steps, TVB limiter (M = 0) on the finest level and
        on the initial state, composite rules on: 16 warmup steps then
        --shock-steps timed steps (or, with --shock-steps 0, the whole job to
        --shock-T), each form."""
        w3 = np.fromfile(args.weights3d, dtype=np.float32)
        f3_pred = algorithmic_flops(2, 3, 5, 10, 15, 81, 27)[1]
        out = {"workload": "3D spherical shock tube p=2, outflow, 3 tagged levels 64^3/128^3/256^3, "
                           "regrid every 16 steps, TVB M=0"}
        for form in args.shock_forms.split(","):
            ctx = build_tagged(pyadg, 3, 2, prec_of[form], 64, 3, SHOCK, [0, 0, 0], 32, 0.2, 0.0, True)
            if form == "learned":
                ctx.mlp_set_params(w3)
                ctx.set_learned(True, tol=args.tol, max_iters=args.max_iters)
            ctx.set_dt(c_dt, args.shock_T, 0.0)
            shapes, choice = ctx.autotune(3)
            n_tune = nvalid_of(ctx.get_levels(), 3)
            ctx.regrid_stats(reset=True)
            job = Job(ctx, 3, c_dt, args.shock_T, regrid_steps=16, max_batch=1 if args.no_graph else 64, graph=not args.stream,
                      name=f"shock tube {form}")
            m0 = ctx.totals()
            s = ctx.stream()
            job.run(16)
            if form == "learned":
                ctx.learned_stats(reset=True)
            e0, e1 = Event(), Event()
            d0, st0 = job.dof, job.steps
            e0.record(s)
            taken = job.run(args.shock_steps if args.shock_steps > 0 else None)
            e1.record(s)
            e1.synchronize()
            ms = e0.elapsed_ms(e1)
            m1 = ctx.totals()
            r = {"steps": taken, "t": job.t, "ms": ms, "ms_per_step": ms / max(taken, 1),
                 "value": (job.dof - d0) / (ms * 1e-3), "elements_end": job.nvalid, "elements_max": job.max_nvalid,
                 "regrids": job.regrids, "regrid_host_s": ctx.regrid_stats(),
                 "mass_drift": abs(m1[0] - m0[0]) / m0[0], "energy_drift": abs(m1[-1] - m0[-1]) / m0[-1],
                 "Dt": job.Dt, "predictor_shapes_ms": shapes, "predictor_choice": choice,
                 "elements_mean": (job.dof - d0) / job.per_elem / max(taken, 1)}
            # the autotuned predictor launch on the initial hierarchy against the
            # matrix peak (algorithmic flops of its valid elements)
            best = min(v for k, v in shapes.items() if not k.startswith("warm"))
            fl = f3_pred * n_tune / (best * 1e-3)
            pk = PEAK_FP64 if form == "fp64" else PEAK_FP32
            r["predictor_roofline"] = {"elements": n_tune, "flops_per_element": f3_pred, "launch_ms": best,
                                       "achieved": fl / 1e12, "peak": pk / 1e12, "unit": "TFLOP/s", "frac": fl / pk}
            if form == "learned":
                ne, nf = ctx.learned_stats()
                r["fallback_fraction"] = nf / max(ne, 1)
            out[form] = r
            log(f"shock tube {form}: {r['ms_per_step']:.1f} ms/step, t={r['t']:.4f}, elements {r['elements_end']}")
            ctx.close()
        if "fp64" in out:
            for f in ("fp32", "learned"):
                if f in out:
                    out[f"{f}_over_fp64"] = out[f]["value"] / out["fp64"]["value"]
        out["note"] = ("mass drift includes the flux through the open boundaries: the shock reaches them at "
                       "t ~ 0.07 and leaves; before that, only the FP32 forms' round-off momentum of the state at "
                       "rest (about 1e-8 of the sound speed) crosses them; periodic runs conserve to 1e-15")
        return out

    if args.shock_only:
        print(json.dumps({"shock_only": True, "shock3d": measure_shock3d()}), flush=True)
        return

    res = {}
    for f in ("fp64", "fp32", "learned"):
        res[f] = kh_job(f)
        log(f"kelvin-helmholtz {f}: {res[f]['ms_per_step']:.3f} ms/step timed, job {res[f]['job']['steps']} steps "
            f"in {res[f]['job']['ms'] / 1e3:.2f} s, mass drift {res[f]['job']['mass_drift']:.1e}")

    # density L1 error at T against the same state on a 2x finer hierarchy
    # (128^2 -> 512^2, FP64), compared on the 64^2 composite cell means
    accuracy = None
    if not args.no_ref:
        ctx = build_tagged(pyadg, 2, P, pyadg.ADG_FP64, 128, 7, KH, [1, 1], 32, 0.05, 5000.0, False)
        ctx.set_dt(c_dt, args.T, 0.0)
        ref_job = Job(ctx, D, c_dt, args.T, regrid_dt=0.05)
        r0 = time.perf_counter()
        while not ref_job.done():
            ref_job.run(2000)
            log(f"fine reference t={ref_job.t:.3f} steps={ref_job.steps}")
        ref = level0_means(ctx, 128)[0]
        ctx.close()
        ref = 0.25 * (ref[0::2, 0::2] + ref[1::2, 0::2] + ref[0::2, 1::2] + ref[1::2, 1::2])
        norm = np.abs(ref).sum()
        err = {f: float(np.abs(res[f]["rho0"] - ref).sum() / norm) for f in res}
        accuracy = {"reference": "FP64, 128^2 -> 256^2 -> 512^2, same state and settings, "
                                 f"{ref_job.steps} steps in {time.perf_counter() - r0:.1f} s",
                    "rho_L1_error": err,
                    "fp32_over_fp64": err["fp32"] / err["fp64"], "learned_over_fp64": err["learned"] / err["fp64"],
                    "rho_L1_diff_to_fp64": {f: float(np.abs(res[f]["rho0"] - res["fp64"]["rho0"]).sum() / norm)
                                            for f in ("fp32", "learned")}}

    kern = {f: kernel_measurement(f) for f in ("fp64", "fp32", "learned")}
    log("kernel measurements done")

    batch = measure_batch() if not args.no_batch else None
    log("batch done")
    vortex = measure_vortex() if not args.no_vortex else None
    log("vortex done")
    shock3d = measure_shock3d() if not args.no_3d else None

    head = res[args.form]
    # the learned form's predictor is the FP32 kernel with an early exit:
    # the roofline is that kernel's plain launch (P iterations)
    k = kern["fp32" if args.form == "learned" else args.form]
    Q, Ns, Nst, Mst, Mbd = k["sizes"]
    f_step, f_pred = algorithmic_flops(P, D, Q, Ns, Nst, Mst, Mbd)
    achieved = f_pred * k["elements"] / (k["pred_ms"] * 1e-3)
    peak = PEAK_FP64 if args.form == "fp64" else PEAK_FP32
    roofline = {"bound": "mfma",
                "kernel": f"k_predictor_nl<{'double' if args.form == 'fp64' else 'float'},3,2,W,false> "
                          f"(plain form, P Picard iterations; W={k['autotune'][1]})",
                "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": PRED_TRAFFIC[args.form], "avg_launch_ms": k["pred_ms"],
                "flops_per_launch": f_pred * k["elements"],
                "measured_on": "kelvin-helmholtz p=3, 3-level band hierarchy 64^2/128^2/256^2, 28 672 elements",
                "stages": stage_bandwidth(k)}

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("cpu baseline")
        med, build = cpu_baseline(res["fp64"]["levels0"], args.cpu_steps)
        n0 = nvalid_of(res["fp64"]["levels0"], D)
        cpu = {"value": n0 * Q * Nst / med, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
               "sample": f"median of {args.cpu_steps} single steps of the headline workload at t=0 "
                         f"({n0} elements, 3 tagged levels, limiter and composite rules) by the FP64 CPU "
                         f"oracle, OpenMP {threads} threads, build {build}",
               "host": {"nproc": os.cpu_count(), "model": cpu_model()}}

    def strip(r):
        return {x: r[x] for x in r if x not in ("rho0", "levels0")}

    line = {
        "metric": METRIC, "value": head["value"], "unit": "DOF-updates/s", "n_gpus": 1, "steps": head["timed_steps"],
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "data": "synthetic",
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "learned": "f32+f64+bf16"}[args.form],
        "config": {"workload": "kelvin-helmholtz 2D p=3, 3 tagged levels 64^2/128^2/256^2, regrid every 0.05, "
                               "TVB M=5000, t=0..1",
                   "form": args.form, "learned_tol": args.tol, "learned_max_iters": args.max_iters,
                   "forms": {f: strip(res[f]) for f in res},
                   "fp32_over_fp64": res["fp32"]["value"] / res["fp64"]["value"],
                   "learned_over_fp64": res["learned"]["value"] / res["fp64"]["value"],
                   "job_fp32_over_fp64": res["fp32"]["job"]["value"] / res["fp64"]["job"]["value"],
                   "job_learned_over_fp64": res["learned"]["job"]["value"] / res["fp64"]["job"]["value"],
                   "accuracy": accuracy,
                   "algorithmic_flops_per_element_step": f_step,
                   "kernels_band_hierarchy": {f: {x: kern[f][x] for x in ("pred_ms", "faceflux_ms", "corrector_ms",
                                                                       "step_ms", "step_graph_ms", "step_graph1_ms",
                                                                       "autotune")}
                                              for f in kern},
                   "batch": batch, "vortex": vortex, "shock3d": shock3d,
                   "step": "predictor, face flux, corrector, limiter, reflux+average down, ghost fill+CFL+advance, "
                           "batches of up to 64 steps as replays of a hipGraph of 16 captured steps (--stream: "
                           "stream launches); regrid between batches when due",
                   "composite_rules": "conservative reflux and CFL over uncovered cells (not the reference's rules)"},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
