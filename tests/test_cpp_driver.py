"""The C++ class surface (include/pargemslr_amd.hpp) drives a full solve the
way the reference's sequential driver does."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "driver_laplacian")


def _driver():
    assert os.path.exists(DRIVER), "tests/cpp/driver_laplacian is missing: run __graft_entry__.build()"
    return DRIVER


def test_driver_is_built_and_links():
    out = subprocess.run([_driver()], capture_output=True, text=True)
    assert out.returncode == 2 and "usage" in out.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["64", "64", "1", "2", "4", "1e-2", "10", "0", "30", "1e-8", "1"],
    ["32", "32", "32", "3", "8", "1e-3", "20", "16", "50", "1e-8", "2"],
    # complexd classes, complex-shifted Laplacian, complex64 apply (the reference's gemslrz driver)
    ["24", "24", "24", "3", "8", "1e-3", "20", "16", "50", "1e-8", "2", "0.2", "0.02", "1"],
    ["24", "24", "24", "3", "8", "1e-3", "20", "16", "50", "1e-8", "1", "0.2", "0.02", "1"],
], ids=["real-c1", "real-mixed", "complex-mixed", "complex-fp64"])
def test_driver_solves(args, gpu):
    out = subprocess.run([_driver()] + args, capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "Number of iterations" in out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["32", "32", "32", "3", "8", "1e-3", "20", "16", "50", "1e-8", "2"],
    ["24", "24", "24", "3", "8", "1e-3", "20", "16", "50", "1e-8", "1", "0.2", "0.02", "1"],
], ids=["real-mixed", "complex-fp64"])
def test_driver_with_device_low_rank_build(args, gpu):
    """GemslrClass::SetLowRankDevice: the same solve with the low-rank
    corrections built on the device takes the same iteration count, within one."""
    its = []
    for env in ({}, {"PGS_LOWRANK_DEVICE": "0"}):
        out = subprocess.run([_driver()] + args, capture_output=True, text=True, timeout=120,
                             env=dict(os.environ, **env))
        print(out.stdout)
        assert out.returncode == 0, out.stdout + out.stderr
        its.append(int(next(l for l in out.stdout.splitlines() if "Number of iterations" in l).split(":")[1]))
    # the two sets of bases differ by rounding only (1e-14 to 1e-12), which can move the count by one
    assert abs(its[0] - its[1]) <= 1, its
