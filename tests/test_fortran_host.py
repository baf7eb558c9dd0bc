"""The Fortran host (phasta_amd/fortran/phg_boxdrv, module phgpu) against
the oracle.

The driver runs the reference's time loop shape (itrdrv.f:398-652) through
the C-ABI in two ways: 'host' keeps the state in Fortran arrays and calls
phg_solgmrs where the reference calls SolGMRs (itrdrv.f:477); 'resident'
keeps it in HBM.  Both must reproduce the oracle's yold after two time
steps of two Newton iterations each, with the oracle's Krylov counts.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle
from phasta_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "phasta_amd", "fortran", "phg_boxdrv")
needs_drv = pytest.mark.skipif(not os.path.exists(DRV), reason="Fortran driver not built")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@needs_drv
def test_driver_rejects_bad_usage():
    """Argument errors end the driver before any library call."""
    p = subprocess.run([DRV], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stdout
    p = subprocess.run([DRV, "4", "0.3", "10", "1", "1", "nomode", "/dev/null"],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "unknown mode" in p.stdout


@pytest.fixture(scope="module")
def oracle_box4():
    case = mesh.box_case(4, mach=0.3, kspace=10)
    orc = pyoracle.Oracle(case)
    acold = np.zeros((case.nshg, 5), order="F")
    yold, _, hist = orc.time_steps(case.yold, acold, steps=2, newton=2)
    return case, yold, [h["iKs"] for h in hist]


@pytest.mark.gpu
@needs_drv
@pytest.mark.parametrize("mode", ["host", "resident"])
def test_fortran_driver_matches_oracle(mode, oracle_box4, tmp_path):
    case, yold_ref, iks_ref = oracle_box4
    out = tmp_path / f"yold_{mode}.bin"
    print(f"[fortran] box4 mode={mode}", flush=True)
    p = subprocess.run([DRV, "4", "0.3", "10", "2", "2", mode, str(out)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    iks = [int(m) for m in re.findall(r"iKs\s+(\d+)", p.stdout)]
    assert iks == iks_ref
    yold = np.fromfile(out, dtype=np.float64).reshape((case.nshg, 5), order="F")
    assert rel(yold, yold_ref) < 1e-12
