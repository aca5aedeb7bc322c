"""The C++ host side end to end on the GPU: the driver in the shape of the
reference's main.cpp builds Simulation<AmrDG<p>, Compressible_Euler>, runs
its evolve loop (single steps and hipGraph batches) through the C-ABI, and
its U_w after N steps matches the oracle's on the same hierarchy to 1e-12
per conserved variable; the norm it reports matches the oracle's Lp error."""
import json
import os
import subprocess
from math import comb

import numpy as np
import pytest

from bars import mode_err

from layouts import two_level_centre

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "ader-dg-amr_amd", "host")
TOL = 1e-12


def rel_err(a, b):
    """Per-(variable, mode) relative error (tests/bars.py: each mode against
    its own size, floored at 1 % of its variable group's scale)."""
    return mode_err(a, b)[0]


def run_driver(dim, args, dump):
    cmd = [os.path.join(HOST, f"adg_main_{dim}d")] + args + ["--dump", dump, "--quiet"]
    print("  running", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    Q, Ns = dim + 2, comb(summary["p"] + dim, dim)
    levels, layout = [], []
    with open(dump, "rb") as f:
        nlev = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        for _ in range(nlev):
            ns = int(np.fromfile(f, dtype=np.int32, count=1)[0])
            levels.append(np.fromfile(f, dtype=np.float64, count=Q * Ns * ns).reshape(Q, Ns, ns))
        for _ in range(nlev):
            nb = int(np.fromfile(f, dtype=np.int32, count=1)[0])
            b = np.fromfile(f, dtype=np.int32, count=2 * dim * nb).reshape(nb, 2 * dim)
            layout.append([(tuple(int(v) for v in r[:dim]), tuple(int(v) for v in r[dim:])) for r in b])
    summary["layout"] = layout
    return summary, levels


def oracle_run(orc, D, P, n, L, ic, prm, c_dt, Tend, nsteps):
    levels = two_level_centre(n, 32, D=D)
    T = orc.Tables(P, D)
    H = orc.Hier(D, levels, [n] * D, [0] * D, [L] * D)
    Uw = orc.project_ic(T, H, ic, prm)
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    Dt = orc.set_dt(T, H, Uw, c_dt)
    steps, t, _ = orc.evolve(T, H, Uw, 0.0, Dt, Tend, c_dt, nsteps_max=nsteps)
    return T, H, Uw, steps, t


@pytest.mark.parametrize("graph_steps", [0, 4])
def test_cpp_driver_vortex_two_levels(orc, tmp_path, graph_steps):
    print(f"case C++ driver isentropic_vortex p=2 16^2 + centre, 6 steps, graph_steps={graph_steps}", flush=True)
    summary, got = run_driver(2, ["--case", "isentropic_vortex", "--p", "2", "--n", "16", "--max_level", "1",
                                  "--c_dt", "0.5", "--max_steps", "6", "--graph_steps", str(graph_steps)],
                              str(tmp_path / "uw.bin"))
    VORTEX = [5.0, 5.0, 5.0, 1.0, 0.0]
    T, H, Uw, steps, t = oracle_run(orc, 2, 2, 16, 10.0, 0, VORTEX, 0.5, 10.0, 6)
    assert summary["steps"] == 6 and steps == 6
    assert abs(summary["t"] - t) <= 1e-12 * t
    for l in range(2):
        e = rel_err(got[l], Uw[l])
        print(f"  level {l} relative error {e:.2e}", flush=True)
        assert e < TOL, f"level {l}"
    for q in range(4):
        ref = orc.lp_error(T, H, Uw, q, 1, 0, VORTEX, t)
        assert abs(summary["L1"][q] - ref) <= 1e-10 * ref, (q, summary["L1"][q], ref)


def test_cpp_driver_density_wave_3d(orc, tmp_path):
    print("case C++ driver density_wave 3D p=2 8^3 + centre, 3 steps", flush=True)
    summary, got = run_driver(3, ["--case", "density_wave", "--p", "2", "--n", "8", "--max_level", "1",
                                  "--c_dt", "0.5", "--max_steps", "3", "--graph_steps", "2"],
                              str(tmp_path / "uw.bin"))
    T, H, Uw, steps, t = oracle_run(orc, 3, 2, 8, 1.0, 5, [0.2, 1.0], 0.5, 1.0, 3)
    assert summary["steps"] == 3
    for l in range(2):
        e = rel_err(got[l], Uw[l])
        print(f"  level {l} relative error {e:.2e}", flush=True)
        assert e < TOL, f"level {l}"


def test_cpp_driver_learned_fp32(tmp_path):
    """The learned warm start and the FP32-MFMA form through the C++ host:
    the run completes and conserves mass against the plain FP64 run of the
    same case to 1e-12."""
    print("case C++ driver kelvin_helmholtz_multimode p=3 32^2 FP32 + learned, 5 steps", flush=True)
    w = os.path.join(ROOT, "ader-dg-amr_amd", "weights", "kh_p3_2d.f32")
    base = ["--case", "kelvin_helmholtz_multimode", "--p", "3", "--n", "32", "--max_level", "0",
            "--max_steps", "5", "--graph_steps", "4"]
    s64, _ = run_driver(2, base, str(tmp_path / "a.bin"))
    s32, _ = run_driver(2, base + ["--precision", "fp32", "--weights", w], str(tmp_path / "b.bin"))
    assert s64["steps"] == s32["steps"] == 5
    assert abs(s32["totals"][0] - s64["totals"][0]) <= 1e-12 * abs(s64["totals"][0])
    assert all(np.isfinite(s32["totals"]))


def test_cpp_driver_reference_kh(orc, tmp_path):
    """The reference's own kelvin_helmolz_instability state
    (Compressible_Euler.h:237-256) through the C++ host, two steps at p=3 on
    one 32^2 level, against the oracle."""
    print("case C++ driver kelvin_helmolz_instability p=3 32^2, 2 steps", flush=True)
    summary, got = run_driver(2, ["--case", "kelvin_helmolz_instability", "--p", "3", "--n", "32",
                                  "--max_level", "0", "--max_steps", "2", "--graph_steps", "0"],
                              str(tmp_path / "uw.bin"))
    from layouts import single_level
    T = orc.Tables(3, 2)
    H = orc.Hier(2, single_level(32, 32), [32, 32], [0, 0], [1, 1])
    Uw = orc.project_ic(T, H, 6, [0.2, 0.05 / np.sqrt(2.0)])
    orc.fill_patch(T, H, Uw)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    orc.evolve(T, H, Uw, 0.0, Dt, 7.0, 0.5, nsteps_max=2)
    assert summary["steps"] == 2
    e = rel_err(got[0], Uw[0])
    print(f"  relative error {e:.2e}", flush=True)
    assert e < TOL


def test_cpp_driver_kh_regrid_limiter(orc, tmp_path):
    """The noisy KH state at p=3 on 16^2 with two tagged levels through the
    C++ host (--regrid: full-domain levels, density-gradient tags every 2
    steps, max_grid_size 32) with the TVB limiter (M = 200) and 2-step graph
    batches between regrid points, 6 steps, against the oracle remade onto
    the layouts the host grid generation makes from the oracle's own tags."""
    import pyadg
    from layouts import tiles
    print("case C++ driver KH p=3 16^2 --regrid (2 levels, every 2 steps) + TVB limiter, 6 steps", flush=True)
    summary, got = run_driver(2, ["--case", "kelvin_helmholtz_noise", "--p", "3", "--n", "16", "--regrid",
                                  "--dtn_regrid", "2", "--amr_c", "0.05", "--limiter", "TVB", "--tvb_m", "200",
                                  "--t_limit", "1", "--max_steps", "6", "--graph_steps", "4", "--T", "1.0"],
                              str(tmp_path / "uw.bin"))
    T = orc.Tables(3, 2)
    cur = [tiles(16 << l, 32, 2) for l in range(3)]
    H = orc.Hier(2, cur, [16, 16], [0, 0], [1, 1])
    Uw = orc.project_ic(T, H, 1, [0.01, 1e-3, 1])
    orc.average_down(T, H, Uw)
    orc.fill_patch(T, H, Uw)
    Dt = orc.set_dt(T, H, Uw, 0.5)
    t, nl_tot = 0.0, 0
    for n in range(6):
        if n % 2 == 0:
            tags = []
            for l in range(min(len(cur) - 1, 1) + 1):
                tg = orc.tag_cells(T, H, Uw, l, 2, 0.05)
                cells = [c for lo, hi in cur[l] for c in
                         ((lo[0] + i, lo[1] + j) for j in range(hi[1] - lo[1] + 1) for i in range(hi[0] - lo[0] + 1))]
                tags.append([cells[i] for i in np.nonzero(tg)[0]])
            new = pyadg.make_grids(2, [16, 16], cur, tags, 2, max_grid_size=32)
            if new != cur:
                Hn = orc.Hier(2, new, [16, 16], [0, 0], [1, 1])
                Uw = orc.remake(T, H, Uw, Hn)
                H, cur = Hn, new
        _, t, Dt, nl = orc.evolve_lim(T, H, Uw, t, Dt, 1.0, 0.5, 200.0, [1.0] * 3, 1, nsteps_max=1)
        nl_tot += nl
    print(f"  driver: {summary['regrids']} regrids, {summary['limited']} cells limited, layout "
          f"{[len(l) for l in summary['layout']]}; oracle {nl_tot} limited, layout {[len(l) for l in cur]}",
          flush=True)
    assert summary["steps"] == 6
    assert summary["layout"] == cur
    assert summary["limited"] == nl_tot
    for l in range(len(cur)):
        e = rel_err(got[l], Uw[l])
        print(f"  level {l}: per-mode error {e:.2e}", flush=True)
        assert e < TOL, f"level {l}"
