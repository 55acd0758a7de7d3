"""Reference-compatible command-line drivers (emsolvers_amd/bin): argv
handling and parameter checks that end before any device work, so they run
without a GPU.  The runs themselves are checked in test_drivers_gpu.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "emsolvers_amd", "bin")
PROGRAMS = ["ems_fdtd", "ems_fem", "ems_fem_abm", "ems_fem_v2", "ems_symplectic_2", "ems_symplectic_4"]


def run(prog, *args, cwd=None):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.skip("drivers not built (run __graft_entry__.build())")
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_all_drivers_built():
    missing = [p for p in PROGRAMS if not os.path.exists(os.path.join(BIN, p))]
    if len(missing) == len(PROGRAMS):
        pytest.skip("drivers not built")
    assert not missing


@pytest.mark.parametrize("prog", ["ems_fem_v2", "ems_symplectic_2", "ems_symplectic_4"])
def test_hex_usage_on_wrong_argc(prog):
    """src/fem_v2/main.cpp:51-65: argc != 18 prints the usage and returns 1."""
    r = run(prog, 1, 1, 1)
    assert r.returncode == 1
    assert r.stdout.startswith("Usage:") and "<dim_x> <dim_y> <dim_z>" in r.stdout


def test_hex_rejects_a_process_grid():
    """One GPU owns the domain: dims other than 1 1 1 are an error (before
    any device work)."""
    r = run("ems_fem_v2", 1, 1, 1, 1.0, 1.0, 1.0, 8, 2e-9, 0.5, 0.5, 0.5, 0.625, 0.5, 0.5, 2, 1, 1)
    assert r.returncode == 1 and "1 1 1" in r.stderr


def test_hex_invalid_parameters():
    """check_params (src/fem_v2/fem_simulation.cpp:1411-1444): a negative
    duration fails before the solver is created."""
    r = run("ems_symplectic_4", 1, 1, 1, 1.0, 1.0, 1.0, 8, -1e-9, 0.5, 0.5, 0.5, 0.625, 0.5, 0.5, 1, 1, 1)
    assert r.returncode == 1 and "Invalid parameters" in r.stderr


def test_fem_rejects_nonpositive_sizes():
    r = run("ems_fem", 0, 0.3, 0)
    assert r.returncode == 1
    r = run("ems_fdtd", 0)
    assert r.returncode == 1
