"""The serving loop (continuous batching of SCF solve sequences): the device
density generator against its host statement, a request's results alone and
in a mix of other requests arriving while it runs, agreement with the same
sequence of single-plan solves, and admission by device-memory budget."""
import ctypes
import time

import numpy as np
import pytest

from bipoly_hydro_amd import density
from bipoly_hydro_amd import potential as P

pytestmark = pytest.mark.gpu

MODEL_A = dict(radius=(8.0, 6.0), separation=20.0, drift=-0.01, noise=1e-4, seed=5)


def host_density(grid, model, i, prof, isym=1):
    r, z, p = grid
    m = dict(model)
    m.setdefault("rho_c", (1.0, 1.0))
    return density.bipolytrope_binary(r, z, p, m, i, prof, isym=isym)


def test_device_density_matches_host_statement():
    import torch
    grid = (66, 34, 32)
    prof, _, _ = P.bipolytrope_profile()
    L = P.load_library()
    L.bhx_density.argtypes = [ctypes.POINTER(P.GridDesc), ctypes.POINTER(P.BinaryModel), ctypes.c_int,
                              ctypes.c_void_p]
    m = P.BinaryModel()
    m.rho_c[:] = [1.0, 0.8]
    m.radius[:] = [8.0, 6.0]
    m.separation, m.drift, m.noise, m.seed = 20.0, -0.01, 1e-4, 5
    d = P.GridDesc(66, 34, 32, 1, 1.0 / 127.0, 30, 20, 5)
    out = torch.zeros((32, 34, 66), dtype=torch.float64, device="cuda:0")
    for i in (0, 3):
        assert L.bhx_density(ctypes.byref(d), ctypes.byref(m), i, out.data_ptr()) == 0
        ref = host_density(grid, dict(MODEL_A, rho_c=(1.0, 0.8)), i, prof)
        got = out.cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.count_nonzero(ref) > 1000


def test_request_alone_and_in_a_mix_give_the_same_results():
    grid_a, n_a = (66, 34, 32), 6
    with P.PotentialServer(precision="mixed") as srv:
        ra = srv.submit(grid_a, n_a, MODEL_A, want_potential=True)
        res_alone, pot_alone, lat_alone = srv.wait(ra)
    assert res_alone["status"] == 0 and res_alone["nsolves"] == n_a
    others = [((42, 26, 16), 8, dict(radius=(5.0, 4.0), separation=14.0, drift=0.005, seed=1)),
              ((66, 66, 64), 5, dict(radius=(10.0, 7.0), separation=24.0, drift=-0.002, noise=1e-5, seed=2)),
              ((34, 18, 16), 10, dict(radius=(4.0, 3.0), separation=9.0, seed=3))]
    with P.PotentialServer(precision="mixed") as srv:
        ids = [srv.submit(*others[0]), srv.submit(*others[1])]
        time.sleep(0.002)  # A arrives while the first two are running
        ra = srv.submit(grid_a, n_a, MODEL_A, want_potential=True)
        time.sleep(0.001)
        ids.append(srv.submit(*others[2]))
        res_mix, pot_mix, lat_mix = srv.wait(ra)
        for i in ids:
            r, _, lat = srv.wait(i)
            assert r["status"] == 0 and np.all(lat > 0)
        st = srv.stats()
    assert res_mix["status"] == 0
    assert np.array_equal(pot_mix, pot_alone)
    assert st["completed"] == 4 and st["solves"] == n_a + 8 + 5 + 10
    assert st["peak_active"] >= 2  # requests overlapped on the device
    assert 0 < st["p50_ms"] <= st["p99_ms"]


def test_served_sequence_matches_single_plan_solves():
    grid, n = (42, 26, 16), 4
    model = dict(radius=(5.0, 4.0), separation=14.0, drift=0.01, noise=1e-5, seed=9)
    with P.PotentialServer(precision="fp64") as srv:
        rid = srv.submit(grid, n, model, want_potential=True)
        res, pot, _ = srv.wait(rid)
    assert res["status"] == 0
    prof, _, _ = P.bipolytrope_profile()
    with P.PotentialPlan(*grid, precision="fp64") as plan:
        for i in range(n):
            ref = plan.potential_solver(host_density(grid, model, i, prof), tstep=0 if i == 0 else 2)
    assert np.linalg.norm(pot - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("isym", [2, 3])
def test_served_symmetric_grid_matches_single_plan_solves(isym):
    # an isym = 2 or 3 request beside an isym = 1 one of the same shape:
    # separate plan pools, the device density on the half-space grid
    grid, n = (42, 26, 16), 3
    model = dict(radius=(5.0, 4.0), separation=14.0, drift=0.01, noise=1e-5, seed=9)
    with P.PotentialServer(precision="fp64") as srv:
        r1 = srv.submit(grid, n, model, want_potential=True)
        rs = srv.submit(grid + (isym,), n, model, want_potential=True)
        res1, pot1, _ = srv.wait(r1)
        res, pot, _ = srv.wait(rs)
    assert res["status"] == 0 and res1["status"] == 0
    prof, _, _ = P.bipolytrope_profile()
    with P.PotentialPlan(*grid, precision="fp64", isym=isym) as plan:
        for i in range(n):
            ref = plan.potential_solver(host_density(grid, model, i, prof, isym), tstep=0 if i == 0 else 2)
    assert np.linalg.norm(pot - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.linalg.norm(pot1 - ref) > 1e-3 * np.linalg.norm(ref)


def test_admission_by_device_memory_budget():
    grid = (66, 66, 64)
    with P.PotentialPlan(*grid, precision="mixed") as plan:
        plan_bytes = plan.device_bytes()
    arrays = 2 * 8 * 66 * 66 * 64
    # room for the shared tables and one request's own buffers (plus 1 MB
    # for the profile table and launch arguments), not for a second request
    budget = plan_bytes + arrays + (1 << 20)
    model = dict(radius=(10.0, 7.0), separation=24.0, seed=4)
    with P.PotentialServer(precision="mixed", hbm_budget=budget) as srv:
        ids = [srv.submit(grid, 3, dict(model, seed=k)) for k in range(4)]
        for i in ids:
            r, _, _ = srv.wait(i)
            assert r["status"] == 0
        st = srv.stats()
        assert st["peak_active"] == 1
        assert st["peak_bytes"] <= budget
        with pytest.raises(P.BHError) as e:
            srv.submit((258, 130, 512), 2, model)
        assert e.value.code == -2


def test_a_request_id_is_waited_for_once():
    with P.PotentialServer(precision="mixed") as srv:
        rid = srv.submit((42, 26, 16), 2, dict(radius=(5.0, 4.0), separation=14.0, seed=1))
        res, _, _ = srv.wait(rid)
        assert res["status"] == 0
        with pytest.raises(P.BHError) as e:
            srv.wait(rid)
        assert e.value.code == -5
        log = srv.iterations()
        assert log.shape[0] >= 2 and np.all(log[:, 3] >= log[:, 2]) and np.all(log[:, 2] >= log[:, 1])


def test_served_sequence_resumes_from_continuation_files(tmp_path):
    """A request resumed from a potential written as continuation files (the
    reference's fort.12 format) gives the single-plan warm solves from that
    potential."""
    grid, n = (42, 26, 16), 3
    model = dict(radius=(5.0, 4.0), separation=14.0, drift=0.01, noise=1e-5, seed=9)
    prof, _, _ = P.bipolytrope_profile()
    with P.PotentialPlan(*grid, precision="fp64") as plan:
        start = plan.potential_solver(host_density(grid, dict(model, seed=3), 0, prof), tstep=0)
    P.cont_write(str(tmp_path), 77, grid, (2, 2), {"pot": start})
    back = P.cont_read(str(tmp_path / "fort.12.   77."), grid, (2, 2), names=("pot",))["pot"]
    np.testing.assert_array_equal(back, start)
    with P.PotentialServer(precision="fp64") as srv:
        rid = srv.submit(grid, n, model, want_potential=True, start_pot=back)
        res, pot, _ = srv.wait(rid)
    assert res["status"] == 0
    with P.PotentialPlan(*grid, precision="fp64") as plan:
        plan.set_potential_from_cont(str(tmp_path / "fort.12.   77."), 2, 2)
        for i in range(n):
            ref = plan.potential_solver(host_density(grid, model, i, prof), tstep=2)
    assert np.linalg.norm(pot - ref) <= 1e-12 * np.linalg.norm(ref)


def test_lanes_by_grid_size_run_side_by_side():
    """A small grid (lane 0, < 2^20 points) and a large one (lane 2, >= 2^22
    points) served together: each lane has its own iterations, and the small
    request's results equal those it gets alone."""
    small, big = (42, 26, 16), (258, 130, 128)
    ms = dict(radius=(5.0, 4.0), separation=14.0, drift=0.005, seed=1)
    mb = dict(radius=(20.0, 15.0), separation=60.0, drift=0.002, seed=2)
    with P.PotentialServer(precision="mixed") as srv:
        r0 = srv.submit(small, 6, ms, want_potential=True)
        alone = srv.wait(r0)[1]
    with P.PotentialServer(precision="mixed") as srv:
        rb = srv.submit(big, 3, mb, want_potential=True)
        rs = srv.submit(small, 6, ms, want_potential=True)
        res_s, pot_s, _ = srv.wait(rs)
        res_b, _, _ = srv.wait(rb)
        log = srv.iterations()
    assert res_s["status"] == 0 and res_b["status"] == 0
    assert np.array_equal(pot_s, alone)
    lanes = set(int(x) for x in log[:, 7])
    assert {0, 2} <= lanes
    # no iteration mixes the two: each solved only its own lane's points
    for row in log:
        pts = row[6]
        assert (pts < (1 << 20) * 6) if row[7] == 0 else (pts >= (1 << 22))
