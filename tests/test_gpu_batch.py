"""Batched solves of uneven grids (grouped persistent launches, hipGraph
replay, autotuned tile shapes) against the same grids solved one plan at a
time.  The grouped kernels run the same per-tile code as the single-grid
launches and every tile shape sums each product in the same order, so the
results must agree bit for bit."""
import numpy as np
import pytest

from bipoly_hydro_amd import density
from bipoly_hydro_amd import potential as P

pytestmark = pytest.mark.gpu

GRIDS = [(34, 18, 16), (42, 26, 16), (66, 34, 32), (50, 22, 64), (34, 34, 8), (66, 66, 64)]


def densities(grids, seed0, scale=1.0):
    return [scale * density.scaled_binary(r, z, p, seed=seed0 + i) for i, (r, z, p) in enumerate(grids)]


def singles(grids, rhos_seq, tsteps_seq, precision):
    """Each grid alone through its own plan, over a sequence of solves."""
    out = []
    for i, (r, z, p) in enumerate(grids):
        with P.PotentialPlan(r, z, p, precision=precision) as plan:
            out.append([plan.potential_solver(rhos[i], tstep=ts[i]) for rhos, ts in zip(rhos_seq, tsteps_seq)])
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_equals_single_plans_bitwise(precision):
    rhos1 = densities(GRIDS, 11)
    rhos2 = densities(GRIDS, 11, scale=1.01)
    # second solve: a mix of warm and cold grids
    ts1 = [0] * len(GRIDS)
    ts2 = [2, 0, 2, 2, 0, 2]
    ref = singles(GRIDS, [rhos1, rhos2], [ts1, ts2], precision)
    with P.PotentialBatch(GRIDS, precision=precision) as b:
        got1 = b.solve(rhos1, ts1)
        got2 = b.solve(rhos2, ts2)
        st = b.stats()
    assert st["ngrids"] == len(GRIDS)
    for i in range(len(GRIDS)):
        assert np.array_equal(got1[i], ref[i][0]), f"grid {GRIDS[i]} cold"
        assert np.array_equal(got2[i], ref[i][1]), f"grid {GRIDS[i]} second solve"


def test_fp64_batch_of_mixed_symmetries_equals_single_plans_bitwise():
    # isym 1, 2 and 3 grids in one batch: the symmetric grids' kernels
    # (side contraction with the image table, equator rows) group beside the
    # plain ones
    grids = [(34, 18, 16, 2), (42, 26, 16, 1), (66, 33, 32, 3), (50, 22, 64, 2)]
    rhos = [[density.lane_emden_n1_binary(r, z, p, radius_cells=(0.2 * r, 0.15 * r), centre_cells=(0.3 * r, -0.3 * r),
                                          noise=1e-3, seed=7 + i, isym=sym) * sc
             for i, (r, z, p, sym) in enumerate(grids)] for sc in (1.0, 1.01)]
    ts = [[0, 0, 0, 0], [2, 0, 2, 2]]
    ref = []
    for i, (r, z, p, sym) in enumerate(grids):
        with P.PotentialPlan(r, z, p, precision="fp64", isym=sym) as plan:
            ref.append([plan.potential_solver(rhos[k][i], tstep=ts[k][i]) for k in range(2)])
    with P.PotentialBatch(grids, precision="fp64") as b:
        got = [b.solve(rhos[k], ts[k]) for k in range(2)]
    for i in range(len(grids)):
        for k in range(2):
            assert np.array_equal(got[k][i], ref[i][k]), f"grid {grids[i]} solve {k}"


def test_batch_graph_replays_match_eager():
    import torch
    dev = torch.device("cuda:0")
    rhos = densities(GRIDS, 23)
    d_rho = [torch.from_numpy(r).to(dev) for r in rhos]
    with P.PotentialBatch(GRIDS, precision="mixed") as eager:
        ref = eager.solve(rhos, 0)
    with P.PotentialBatch(GRIDS, precision="mixed", graph=True) as b:
        d_pot = [torch.empty_like(x) for x in d_rho]
        s = torch.cuda.current_stream()
        for rep in range(3):
            for x in d_pot:
                x.fill_(float("nan"))
            b.solve_device([x.data_ptr() for x in d_rho], [x.data_ptr() for x in d_pot], 0, s.cuda_stream)
            torch.cuda.synchronize()
            assert b.stats()["graph"] == 1
            for i, x in enumerate(d_pot):
                assert np.array_equal(x.cpu().numpy(), ref[i]), f"replay {rep} grid {GRIDS[i]}"
            # unrelated eager work between replays
            a = torch.randn(512, 512, device=dev)
            (a @ a).sum().item()


def test_batch_autotune_keeps_results_and_reports_choices():
    rhos = densities(GRIDS, 31)
    with P.PotentialBatch(GRIDS, precision="mixed") as eager:
        ref = eager.solve(rhos, 0)
    with P.PotentialBatch(GRIDS, precision="mixed", autotune=True, graph=True) as b:
        b.solve(rhos, 0)
        left = b.stats()["tuning_left"]
        assert left > 0
        for _ in range(left + 1):  # the remaining tuning solves, then the tuned sequence
            got = b.solve(rhos, 0)
            for i in range(len(GRIDS)):
                assert np.array_equal(got[i], ref[i])
        st = b.stats()
        assert st["tuning_left"] == 0 and st["graph"] == 1
        tune = b.tuning()
    # the mixed solve's residuals and every launch's grid and order were searched
    for ax, allowed in (("rows", {2, 4, 8}), ("grid", {0, 1, 2, 4}), ("order", {0, 1})):
        chosen = [c for c in tune[ax]["choice"] if c >= 0]
        assert chosen and set(chosen) <= allowed, (ax, tune[ax])


def test_batch_over_budget_is_refused():
    with pytest.raises(P.BHError) as e:
        P.PotentialBatch(GRIDS, precision="mixed", hbm_budget=1 << 20)
    assert e.value.code == -2
