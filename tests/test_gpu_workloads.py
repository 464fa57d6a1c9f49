"""The batched and served workloads at their full sizes.

Batched: the 64 uneven grids of the batched workload (density.batched_grids),
mixed precision, solved as one batch with the hipGraph and the autotuner on.
Every grid equals its own single-plan solve bit for bit, through the tuning
solves and the tuned graph replays; the 258 x 130 x 512 members (given the
reference's own density) hold the reference's output to the mixed bar, and
the 66 x 34 x 64 members the oracle's.

Served: SCF requests of 20-30 dependent solves on grid sizes drawn as for the
batched workload, arriving as a Poisson stream so that they overlap on the
device; every request's final potential equals the one it gives alone."""
import ctypes
import os
import time

import numpy as np
import pytest

from bipoly_hydro_amd import density
from bipoly_hydro_amd import potential as P
from golden_cases import c2_digests, c2_slices, rell2

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-8


def test_batched_64_uneven_grids(golden_dir):
    import torch
    dev = torch.device("cuda:0")
    grids = density.batched_grids(64)
    c2 = np.load(os.path.join(golden_dir, "c2_258x130x512.npz"))
    rhos = []
    for r, z, p, s in grids:
        if (r, z, p) == (258, 130, 512):
            rhos.append(density.config2_density(seed=int(c2["seeds"][0])))
        else:
            rhos.append(density.scaled_binary(r, z, p, seed=s))
    d_rho = [torch.from_numpy(x).to(dev) for x in rhos]
    d_pot = [torch.empty_like(x) for x in d_rho]
    rp, pp = [x.data_ptr() for x in d_rho], [x.data_ptr() for x in d_pot]
    stream = torch.cuda.current_stream().cuda_stream
    print(f"batch of {len(grids)} grids, {sum(x.numel() for x in d_rho) / 1e6:.1f} M points", flush=True)
    with P.PotentialBatch([g[:3] for g in grids], precision="mixed", graph=True, autotune=True) as b:
        outs = []
        k = 0
        while k < 2 or b.stats()["tuning_left"] > 0 or k < 3:  # the tuning solves, then tuned graph replays
            b.solve_device(rp, pp, 0, stream)
            torch.cuda.synchronize()
            outs.append([x.cpu().numpy() for x in d_pot])
            k += 1
        outs.append(None)
        b.solve_device(rp, pp, 0, stream)
        torch.cuda.synchronize()
        outs[-1] = [x.cpu().numpy() for x in d_pot]
        print(f"{len(outs)} batched solves, tuning: "
              f"{ {a: (v['gain_ms'], sorted(set(v['choice']))) for a, v in b.tuning().items()} }", flush=True)
        st = b.stats()
        assert st["tuning_left"] == 0 and st["graph"] == 1
        for k in range(1, len(outs)):
            for i in range(len(grids)):
                assert np.array_equal(outs[k][i], outs[0][i]), f"solve {k} grid {i} {grids[i][:3]}"
        # every grid alone through its plan's own (ungrouped) launches
        L = P.load_library()
        for i in range(len(grids)):
            h = ctypes.c_void_p()
            P._check(L.bh_batch_plan(b._h, i, ctypes.byref(h)))
            P._check(L.bh_solve_device(h, ctypes.c_void_p(rp[i]), ctypes.c_void_p(pp[i]), 0,
                                       ctypes.c_void_p(stream)))
        torch.cuda.synchronize()
        for i in range(len(grids)):
            assert np.array_equal(d_pot[i].cpu().numpy(), outs[0][i]), f"single plan, grid {i} {grids[i][:3]}"
    n_c2 = n_or = 0
    from oracle import oracle as O
    for i, (r, z, p, s) in enumerate(grids):
        if (r, z, p) == (258, 130, 512):
            parts = c2_slices(outs[0][i], "1")
            for k, v in parts.items():
                assert rell2(v, c2[k]) <= TOL_MIXED, (i, k)
            for k, v in c2_digests(outs[0][i], c2["digest_weights"]).items():
                assert rell2(v, c2[k + "1"]) <= TOL_MIXED, (i, k)
            n_c2 += 1
        elif (r, z, p) == (66, 34, 64):
            ref = O.OraclePlan(r, z, p).solve(rhos[i], tstep=0)
            e = rell2(outs[0][i], ref)
            print(f"grid {i} {r}x{z}x{p}: relL2 vs oracle {e:.2e}", flush=True)
            assert e <= TOL_MIXED
            n_or += 1
    assert n_c2 >= 1 and n_or >= 1


def served_requests(n, seed=777):
    grids = density.batched_grids(n, seed=seed)
    u = density.splitmix64_uniform(seed + 1, 4 * n)
    out = []
    for i, (r, z, p, s) in enumerate(grids):
        rg = min(r - 3, z - 2)
        sep = (0.4 + 0.4 * u[4 * i]) * rg
        r1 = max(3.0, min(0.45 * sep, 0.45 * (z - 4), 0.45 * (r - 3)))
        r2 = max(2.0, r1 * (0.3 + 0.7 * u[4 * i + 1]) ** (1.0 / 3.0))
        model = dict(radius=(r1, r2), separation=sep, drift=-0.002 * (0.5 + u[4 * i + 2]), noise=1e-4, seed=s)
        out.append(((r, z, p), 20 + int(11 * u[4 * i + 3]), model))
    return out


def test_served_poisson_stream_alone_equals_in_mix():
    reqs = served_requests(8)
    alone = []
    with P.PotentialServer(precision="mixed") as srv:
        for g, n, m in reqs:
            res, pot, lat = srv.wait(srv.submit(g, n, m, want_potential=True))
            assert res["status"] == 0 and res["nsolves"] == n
            alone.append(pot)
    gaps = -np.log(1.0 - density.splitmix64_uniform(99, len(reqs))) / 400.0  # 400 arrivals per second
    with P.PotentialServer(precision="mixed") as srv:
        ids = []
        t_next = time.perf_counter()
        for (g, n, m), gap in zip(reqs, gaps):
            t_next += gap
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids.append(srv.submit(g, n, m, want_potential=True))
        mix = []
        for rid in ids:
            res, pot, lat = srv.wait(rid)
            assert res["status"] == 0 and np.all(lat > 0)
            mix.append(pot)
        st = srv.stats()
    print(f"served: {st['solves']} solves, peak active {st['peak_active']}, p50 {st['p50_ms']:.3f} ms, "
          f"p99 {st['p99_ms']:.3f} ms", flush=True)
    assert st["peak_active"] >= 2
    for i, (a, b) in enumerate(zip(alone, mix)):
        assert np.array_equal(a, b), f"request {i} {reqs[i][0]}"
