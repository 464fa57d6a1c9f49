"""The Fortran drop-in (bipoly_hydro_amd/fortran): the replacement potsetup /
potential_solver with the reference's COMMON blocks, driven by a 2 x 2 MPI
rank grid as the hydro code decomposes the grid, against the reference's own
golden output."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from golden_cases import load_case, rell2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "bipoly_hydro_amd", "fortran")
DEMO = os.path.join(FDIR, "_build", "g42", "bh_fortran_demo")
MPIEXEC = "/opt/conda/bin/mpiexec"
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def build_demo():
    subprocess.run(["make", "-s", "-C", FDIR, "NAME=g42", "NUMR=42", "NUMZ=26", "NUMPHI=16", "NRP=2", "NZP=2"],
                   check=True)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_dropin_builds_and_links():
    build_demo()
    assert os.path.exists(DEMO)
    out = subprocess.run(["ldd", DEMO], capture_output=True, text=True).stdout
    assert "libbh_potential.so" in out


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-8)])
def test_fortran_dropin_under_mpi_matches_reference(golden_dir, tmp_path, precision, tol):
    if not (os.path.exists(DEMO) and os.path.exists(MPIEXEC)):
        pytest.skip("demo driver or mpiexec missing")
    (numr, numz, numphi), rho1, rho2, ref1, ref2 = load_case(golden_dir, "g42x26x16")
    f = {k: str(tmp_path / f"{k}.bin") for k in ("rho1", "pot1", "rho2", "pot2")}
    rho1.tofile(f["rho1"])
    rho2.tofile(f["rho2"])
    env = dict(os.environ, BH_PRECISION=precision)
    r = subprocess.run(["timeout", "-k", "5", "120", MPIEXEC, "-n", "4", DEMO, f["rho1"], f["pot1"], f["rho2"],
                        f["pot2"]], capture_output=True, text=True, env=env)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    pot1 = np.fromfile(f["pot1"]).reshape(rho1.shape)
    pot2 = np.fromfile(f["pot2"]).reshape(rho1.shape)
    e1, e2 = rell2(pot1, ref1), rell2(pot2, ref2)
    print(f"fortran drop-in {precision}: cold {e1:.3e} warm {e2:.3e}", flush=True)
    assert e1 <= tol and e2 <= tol
