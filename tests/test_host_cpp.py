"""The C++ host side (Simulation / Solver / AmrDG / ModelEquation /
Compressible_Euler / Mesh / BoundaryCondition over the C-ABI) builds and its
driver refuses what the library does not run, before any GPU call."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "ader-dg-amr_amd", "host")


def _bin(dim):
    path = os.path.join(HOST, f"adg_main_{dim}d")
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built; run __graft_entry__.build() or make -C ader-dg-amr_amd")
    return path


@pytest.mark.parametrize("dim", [2, 3])
def test_driver_help(dim):
    r = subprocess.run([_bin(dim), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert f"space dimension: {dim}" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--case", "keplerian_disc"], "not part of this library"),
    (["--case", "no_such_case"], "unknown case"),
    (["--p", "4"], "p must be 1, 2 or 3"),
    (["--max_level", "2"], "at most one fine level"),
    (["--regrid", "--max_level", "5"], "max_level must be at most 4"),
    (["--limiter", "MUSCL"], "unknown limiter"),
    (["--bogus"], "unknown option"),
])
def test_driver_refuses_before_the_device(args, msg):
    r = subprocess.run([_bin(2)] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert msg in r.stderr


def test_vortex_is_two_dimensional():
    r = subprocess.run([_bin(3), "--case", "isentropic_vortex"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "two-dimensional" in r.stderr
