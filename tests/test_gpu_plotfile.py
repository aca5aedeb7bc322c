"""Plotfile output and restart through the C++ host (PlotFile / ReadPlotFile,
Core/Solver.h:599-652,922-968): the driver writes one plotfile per
conserved variable, tstep_<n>_q_<q>_plt with components <name>_<mode>, whose
valid-cell data equal the state it dumps; a run restarted from the plotfiles
of step 3 continues to step 6 with the bits of the uninterrupted run, for a
static and for a regridding hierarchy.  (AMReX and yt are not available to
read these files: the format follows AMReX's plotfile layout and is read
back here by an independent reader.)"""
import os

import numpy as np
import pytest

from plotfile_reader import read
from test_gpu_host_cpp import run_driver

pytestmark = pytest.mark.gpu

NAMES = ["mass_density", "momentum_x", "momentum_y", "energy_density"]
BASE = ["--case", "kelvin_helmholtz_noise", "--p", "3", "--n", "16", "--T", "1.0", "--graph_steps", "4"]


def _valid(u, boxes, dim=2):
    """valid-cell values of a dumped level [q][n][slot], box by box."""
    out, off = [], 0
    for lo, hi in boxes:
        g = [hi[d] - lo[d] + 3 for d in range(dim)]
        idx = [off + (x + 1) + g[0] * (y + 1) for y in range(hi[1] - lo[1] + 1) for x in range(hi[0] - lo[0] + 1)]
        out.append(u[:, :, idx])
        off += int(np.prod(g))
    return out


def test_plotfiles_hold_the_state(tmp_path):
    print("case plotfiles of KH p=3 16^2 + centre level, every 2 steps", flush=True)
    d = str(tmp_path / "plt")
    s, got = run_driver(2, BASE + ["--max_level", "1", "--max_steps", "4", "--plot_every", "2",
                                   "--data_dir", d], str(tmp_path / "uw.bin"))
    for n in (0, 2, 4):
        for q in range(4):
            assert os.path.isdir(os.path.join(d, f"tstep_{n}_q_{q}_plt"))
    for q in range(4):
        pf = read(os.path.join(d, f"tstep_4_q_{q}_plt"))
        assert pf["names"] == [f"{NAMES[q]}_{m}" for m in range(10)]
        assert pf["finest"] == 1 and pf["step"] == 4 and pf["time"] == s["t"]
        for l in range(2):
            lev = pf["levels"][l]
            assert lev["boxes"] == s["layout"][l]
            for b, v in zip(_valid(got[l], lev["boxes"]), lev["data"]):
                assert np.array_equal(b[q], v)


@pytest.mark.parametrize("mode", ["static", "regrid"])
def test_restart_continues_bitwise(tmp_path, mode):
    print(f"case restart from step 3 of KH p=3 16^2, {mode} hierarchy", flush=True)
    extra = ["--max_level", "1"] if mode == "static" else ["--regrid", "--dtn_regrid", "1", "--amr_c", "0.05"]
    d = str(tmp_path / "plt")
    sa, a = run_driver(2, BASE + extra + ["--max_steps", "6"], str(tmp_path / "a.bin"))
    run_driver(2, BASE + extra + ["--max_steps", "3", "--plot_every", "3", "--data_dir", d],
               str(tmp_path / "b.bin"))
    sc, c = run_driver(2, BASE + extra + ["--restart", "3", "--max_steps", "3", "--data_dir", d],
                       str(tmp_path / "c.bin"))
    print(f"  uninterrupted: {sa['steps']} steps t={sa['t']!r}; restarted: {sc['steps']} steps t={sc['t']!r}",
          flush=True)
    assert sc["steps"] == sa["steps"] == 6 and sc["t"] == sa["t"]
    assert sc["layout"] == sa["layout"]
    for l in range(len(a)):
        assert np.array_equal(a[l], c[l]), f"level {l}"
