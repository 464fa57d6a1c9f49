"""The device source step (bh_source_device, bh_make_source_temp_device)
against the reference's own source.F90 output (tests/golden/source_*.npz,
written by make_golden.py --source from the reference compiled here) and
against the oracle's restatement on larger grids, and fed by the solve's
device potential.  Bar: relative L2 <= 1e-12 per field (FP64 throughout;
the device may contract multiply-adds)."""
import os

import numpy as np
import pytest

from bipoly_hydro_amd import density
from bipoly_hydro_amd import potential as P
from golden_cases import SOURCE_CASES, SOURCE_GRID, rell2, source_inputs

pytestmark = pytest.mark.gpu
TOL = 1e-12
OUT = ("s", "t", "a", "temps", "tempa")


def to_dev(f):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}


def params(sc, tstep, bc):
    return dict(dt=sc["dt"], omega=sc["omega"], tstep=tstep, bc=bc, R_com=sc["R_com"],
                cylin_a_com=sc["cylin_a_com"], cylin_v_com=sc["cylin_v_com"])


def run_device(plan, f, sc, st, tstep, bc, mst_only=False):
    import torch
    d = to_dev(f)
    ptrs = {k: v.data_ptr() for k, v in d.items()}
    s = torch.cuda.current_stream().cuda_stream
    if mst_only:
        plan.make_source_temp_device(ptrs, params(sc, tstep, bc), s)
    else:
        plan.source_device(st, ptrs, params(sc, tstep, bc), s)
    torch.cuda.synchronize()
    return {k: d[k].cpu().numpy() for k in OUT}


@pytest.mark.parametrize("tag,st,tstep,bc", SOURCE_CASES)
def test_source_matches_reference(golden_dir, tag, st, tstep, bc):
    g = np.load(os.path.join(golden_dir, "source_%dx%dx%d.npz" % SOURCE_GRID))
    f, sc = source_inputs()
    with P.PotentialPlan(*SOURCE_GRID) as plan:
        got = run_device(plan, f, sc, st, tstep, bc)
    for k in OUT:
        e = rell2(got[k], g[f"{tag}_{k}"])
        print(f"{tag} {k}: relL2 vs reference {e:.2e}", flush=True)
        assert e <= TOL, (k, e)


@pytest.mark.parametrize("grid,tstep,bc", [((130, 66, 64), 10, (1, 2, 0)), ((258, 130, 128), 4000, (0, 1, 2))])
def test_source_matches_oracle_on_larger_grids(grid, tstep, bc):
    from oracle import oracle as O
    f, sc = source_inputs(grid, seed=99)
    with P.PotentialPlan(*grid) as plan:
        got = run_device(plan, f, sc, 2, tstep, bc)
        mst = run_device(plan, f, sc, 0, tstep, bc, mst_only=True)
    ref = O.source(f, source_time=2, tstep=tstep, bc=bc, dt=sc["dt"], omega=sc["omega"], R_com=sc["R_com"],
                   cylin_a_com=sc["cylin_a_com"], cylin_v_com=sc["cylin_v_com"])
    refm = O.make_source_temp(f, tstep=tstep, bc=bc, dt=sc["dt"], omega=sc["omega"], R_com=sc["R_com"],
                              cylin_a_com=sc["cylin_a_com"], cylin_v_com=sc["cylin_v_com"])
    for k in OUT:
        assert rell2(got[k], ref[k]) <= TOL, k
    for k in ("temps", "tempa"):
        assert rell2(mst[k], refm[k]) <= TOL, k
    np.testing.assert_array_equal(mst["s"], f["s"])  # make_source_temp leaves the momenta alone


def test_source_reads_the_solved_potential_on_the_device():
    """solve -> source with the potential never leaving the device equals the
    oracle's source step on the downloaded potential."""
    import torch
    from oracle import oracle as O
    grid = (66, 66, 64)
    f, sc = source_inputs(grid, seed=5)
    rho = density.config1_density(seed=3)
    f["rho"] = rho
    with P.PotentialPlan(*grid, precision="mixed") as plan:
        d = to_dev(f)
        d_pot = torch.empty_like(d["rho"])
        s = torch.cuda.current_stream().cuda_stream
        plan.potential_solver_device(d["rho"].data_ptr(), d_pot.data_ptr(), 0, s)
        ptrs = {k: v.data_ptr() for k, v in d.items()}
        ptrs["pot"] = d_pot.data_ptr()
        plan.source_device(2, ptrs, params(sc, 20, (1, 1, 1)), s)
        torch.cuda.synchronize()
        got = {k: d[k].cpu().numpy() for k in OUT}
        f["pot"] = d_pot.cpu().numpy()
    ref = O.source(f, source_time=2, tstep=20, bc=(1, 1, 1), dt=sc["dt"], omega=sc["omega"])
    for k in OUT:
        assert rell2(got[k], ref[k]) <= TOL, k
