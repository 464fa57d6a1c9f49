"""Benchmark: Poisson (self-gravity) solves per second at 258 x 130 x 512.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid R Z P] [--warm]

A step is one potential_solver call (cold start: 20 ADI steps from zero, the
reference's setting for a solve without a previous potential) on a seeded
binary of n = 1 polytropes already resident in HBM, enqueued through the C
ABI on one stream.  Prints one JSON line.  One GPU only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Poisson solves/s @258×130×512 mixed vs FP64 at relL2≤1e-8; p50/p99 ms"
HBM_PEAK_GBS = 8000.0
# HBM bytes per solve of the boundary contraction (with the projections
# scaled as they are read; its partials are summed by the face kernel) and of one FP64 ADI
# half-step at 258 x 130 x 512, from rocprofv3 FETCH_SIZE (x 2.000, the
# copy-kernel calibration) + WRITE_SIZE;
# profiles/potential_solver_258x130x512_hbm_traffic.md
CONTRACT_TRAFFIC = (546340.0 * 2.0 + 7200.0) * 1024  # FP64 table
# the mixed cold inverse DFT (radix 8, x = faces + delta1 + correction sum
# in, phi out; the warm-start state stays as the two corrections): HBM bytes
# per launch (FETCH x 2.000 + WRITE) and the FP64 MFMA operations it executes
# (SQ_INSTS_VALU_MFMA_MOPS_F64 x 512); profiles/potential_solver_258x130x512_{hbm_traffic,mfma}.md
INV_TRAFFIC = (74626.0 * 2.0 + 138799.0) * 1024
INV_MFMA_FLOP = 2147328.0 * 512
FP64_PEAK_TFLOPS = 78.6
SWEEP_TRAFFIC = ((159728.0 + 158982.0) * 2.0 + (131072.0 + 131512.0)) * 1024 / 2  # mean of the two half-steps


def rell2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(numr, numz, numphi, rho, threads, samples=5):
    """The oracle (C restatement, OpenMP) timed on cold solves of the same
    grid, the median of `samples`.  Its Green's tables are allocated and
    every page is written with a nonzero constant, so that the 5 GB table
    stream reaches DRAM; their generation is one-time setup (potsetup),
    which the reference's timer excludes too."""
    from oracle import oracle as O
    O.set_threads(threads)
    plan = O.OraclePlan(numr, numz, numphi, with_tables="filled")
    ts = []
    for _ in range(samples):
        t0 = time.perf_counter()
        plan.solve(rho, tstep=0)
        ts.append(time.perf_counter() - t0)
    del plan
    dt = float(np.median(ts))
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"median of {samples} cold solves {numr}x{numz}x{numphi} (oracle/bh_oracle.c, OpenMP, "
                      f"{threads} threads on {cpu_model()}; Green's tables filled, every page written): "
                      f"{', '.join(f'{t:.2f}' for t in ts)} s"}


def reference_errors(pot, tag="1"):
    """relL2 of a 258 x 130 x 512 potential against the reference's own
    output kept in tests/golden/c2_258x130x512.npz: over the kept slices,
    and over the whole-field digests (plane / K / J norms and weighted row
    sums); tag "1" is the cold solve, "2" the warm one that follows it."""
    path = os.path.join(ROOT, "tests", "golden", "c2_258x130x512.npz")
    if not os.path.exists(path) or pot.shape != (512, 130, 258):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from golden_cases import c2_digests, c2_slices
    g = np.load(path)
    parts = [(v, g[k]) for k, v in c2_slices(pot, tag).items()]
    num = sum(float(np.sum((a - b) ** 2)) for a, b in parts)
    den = sum(float(np.sum(b ** 2)) for _, b in parts)
    out = {"slices": (num / den) ** 0.5}
    if "row_dot" + tag in g:
        for k, v in c2_digests(pot, g["digest_weights"]).items():
            out[k] = rell2(v, g[k + tag])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, nargs=3, default=[258, 130, 512])
    ap.add_argument("--warm", action="store_true", help="time warm solves (tstep > 1, 5 ADI steps)")
    ap.add_argument("--passes", type=int, default=2, help="refinement passes of the headline mixed solve")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the short batched and served runs of the default line")
    ap.add_argument("--only", choices=["mixed", "fp64", "mixed_t32", "spectral64"],
                    help="time one precision alone (profiling runs); prints a reduced line")
    ap.add_argument("--mmax", type=int, default=30,
                    help="boundary azimuthal modes of the --only solve (the reference's 30 by default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batched", action="store_true",
                    help="instead of the headline: 64 uneven grids solved as one batch (grouped "
                         "persistent launches, hipGraph, autotuned tiles), against the same grids "
                         "solved one plan at a time")
    ap.add_argument("--ngrids", type=int, default=64)
    ap.add_argument("--serve", action="store_true",
                    help="instead of the headline: a served stream of SCF requests (Poisson arrivals, "
                         "mixed grid sizes, 20-50 dependent solves each) through the continuous-"
                         "batching server; reports solves/s and per-solve p50/p99 latency")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--rate", type=float, default=40.0, help="request arrivals per second (--serve)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first copy 256 MiB device-to-device three times (a known-size kernel for "
                         "calibrating the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from bipoly_hydro_amd import density
    from bipoly_hydro_amd import potential as P

    if args.batched:
        return run_batched(args)
    if args.serve:
        return run_serve(args)
    numr, numz, numphi = args.grid
    dev = torch.device("cuda:0")
    if args.calib_copy:
        src = torch.ones(1 << 25, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if (numr, numz, numphi) == (258, 130, 512):
        rho = density.config2_density(seed=1234)
    else:
        rho = density.scaled_binary(numr, numz, numphi, seed=1234)
    d_rho = torch.from_numpy(rho).to(dev)
    d_pot = torch.empty_like(d_rho)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    tstep = 2 if args.warm else 0

    def timed(plan, tstep, steps):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(stream)
            plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), tstep, sp)
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array([a.elapsed_time(b) for a, b in ev])
        return {"ms_per_step": wall * 1e3 / steps, "solves_per_s": steps / wall,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99))}

    def run(precision, passes=2, warm_steps=0, mmax=30):
        plan = P.PotentialPlan(numr, numz, numphi, precision=precision, refine_passes=passes, mmax=mmax)
        plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), 0, sp)
        torch.cuda.synchronize()
        pot_cold = d_pot.cpu().numpy()
        # the warm solve that follows a cold one (tstep 2, the reference's
        # 5 ADI steps from the kept potential), on a perturbed density
        d_rho.mul_(1.01).add_(d_rho2, alpha=1e-3)
        plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), 2, sp)
        torch.cuda.synchronize()
        pot_warm = d_pot.cpu().numpy()
        d_rho.copy_(d_rho1)
        for _ in range(args.warmup):
            plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), tstep, sp)
        res = timed(plan, tstep, args.steps)
        # stage and kernel times: a second timed region of the same solves in
        # which the library records HIP events on the solve's stream at the
        # stage boundaries and around the boundary contraction (kept out of
        # the region above, whose wall time is the headline)
        plan.set_timing(True)
        for _ in range(args.steps):
            plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), tstep, sp)
        torch.cuda.synchronize()
        res["stage_ms"] = plan.last_timings()
        plan.set_timing(False)
        if warm_steps:
            w = timed(plan, 2, warm_steps)
            res["warm"] = {k: w[k] for k in ("ms_per_step", "p50_ms", "p99_ms")}
            # a cold solve and the warm one that follows it (the hydro loop's
            # first two steps): a mixed cold solve leaves its warm-start
            # state as its corrections, which that warm solve writes out
            ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(10)]
            for e in ev:
                e[0].record(stream)
                plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), 0, sp)
                e[1].record(stream)
                plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), 2, sp)
                e[2].record(stream)
            torch.cuda.synchronize()
            res["warm"]["cold_then_warm_ms"] = [float(np.median([e[0].elapsed_time(e[1]) for e in ev])),
                                                float(np.median([e[1].elapsed_time(e[2]) for e in ev]))]
        res["error_vs_reference_cold"] = reference_errors(pot_cold, "1")
        res["error_vs_reference_warm"] = reference_errors(pot_warm, "2")
        plan.close()
        return res, pot_cold

    d_rho1 = d_rho.clone()
    if (numr, numz, numphi) == (258, 130, 512):
        d_rho2 = torch.from_numpy(density.config2_density(seed=99)).to(dev)
    else:
        d_rho2 = torch.from_numpy(density.scaled_binary(numr, numz, numphi, seed=99)).to(dev)
    if args.only:
        # one precision alone (for profiling runs): the headline fields of
        # that solve, nothing else
        prec = args.only
        res, _ = run(prec, passes=args.passes, mmax=args.mmax)
        print(json.dumps({"metric": METRIC, "value": res["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "f64" if prec in ("fp64", "spectral64") else "f32+f64", "data": "synthetic",
                          "config": {"workload": f"potential_solver {numr}x{numz}x{numphi} tstep={tstep}, "
                                                 f"{prec} only, mmax={args.mmax}" +
                                                 (f", {args.passes} refinement passes" if prec.startswith("mixed")
                                                  else "")},
                          "p50_ms": res["p50_ms"], "p99_ms": res["p99_ms"], "stage_ms": res["stage_ms"],
                          "error_vs_reference_cold": res["error_vs_reference_cold"],
                          "error_vs_reference_warm": res["error_vs_reference_warm"]}), flush=True)
        return
    fp64, pot64 = run("fp64", warm_steps=args.steps)
    mixed, potmx = run("mixed", passes=args.passes, warm_steps=args.steps)
    mixed["relL2_vs_fp64_gpu"] = rell2(potmx, pot64)
    ms_per_step = mixed["ms_per_step"]
    lat_p50, lat_p99 = mixed["p50_ms"], mixed["p99_ms"]
    # the error after each refinement pass (the solve with that many passes),
    # its time, and the opt-in FP32-table variant at one pass
    per_pass = []
    for npass in (1, 2, 3, 4):
        plan = P.PotentialPlan(numr, numz, numphi, precision="mixed", refine_passes=npass)
        plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), 0, sp)
        torch.cuda.synchronize()
        pot = d_pot.cpu().numpy()
        t = timed(plan, 0, max(10, args.steps // 5))
        plan.close()
        per_pass.append({"passes": npass, "ms_per_solve": t["ms_per_step"],
                         "error_vs_reference": reference_errors(pot, "1"), "relL2_vs_fp64_gpu": rell2(pot, pot64)})
    # the same spectral form of the iteration evaluated once in FP64: what
    # the FP32/bf16 transforms and the refinement passes buy
    s64, pots64 = run("spectral64", warm_steps=max(10, args.steps // 5))
    s64 = {"ms_per_step": s64["ms_per_step"], "warm_ms": s64["warm"]["ms_per_step"],
           "stage_ms": s64["stage_ms"], "error_vs_reference_cold": s64["error_vs_reference_cold"],
           "relL2_vs_fp64_gpu": rell2(pots64, pot64),
           "precision": "FP64 spectral solve: FP64 MFMA sine transforms and M_r products, FP64 Thomas, one pass"}
    # boundary modes m = 0..255 (configuration 3's table range; the
    # reference keeps 30): plan set-up with the device table generators, the
    # mixed solve against the FP64 solve with the same modes, and how far the
    # modes past 30 move the potential
    def run_modes(mm):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan = P.PotentialPlan(numr, numz, numphi, precision="mixed", refine_passes=args.passes, mmax=mm)
        torch.cuda.synchronize()
        setup_s = time.perf_counter() - t0
        plan.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), 0, sp)
        torch.cuda.synchronize()
        pot = d_pot.cpu().numpy()
        t = timed(plan, 0, max(10, args.steps // 5))
        plan.close()
        with P.PotentialPlan(numr, numz, numphi, precision="fp64", mmax=mm) as p64:
            p64.potential_solver_device(d_rho.data_ptr(), d_pot.data_ptr(), 0, sp)
            torch.cuda.synchronize()
            pot64m = d_pot.cpu().numpy()
        nb_ = (numr - 2 + 15) // 16
        return {"mmax": mm, "ms_per_step": t["ms_per_step"], "p50_ms": t["p50_ms"], "plan_setup_s": setup_s,
                "table_bytes": mm * nb_ * (nb_ + 1) // 2 * (numz - 2) * 256 * 8 + mm * (2 * numz - 5) * (numr - 2) * 8,
                "relL2_vs_fp64_gpu_same_modes": rell2(pot, pot64m),
                "fp64_mmax30_vs_fp64_this_relL2": rell2(pot64, pot64m)}
    modes = run_modes(min(256, numphi // 2)) if numphi // 2 > 30 else None
    t32, pot32 = run("mixed_t32", passes=1)
    t32 = {"ms_per_step": t32["ms_per_step"], "error_vs_reference_cold": t32["error_vs_reference_cold"],
           "precision": "FP32 top/bottom Green's table (opt-in), 1 refinement pass"}

    # Dominant kernel of the measured (mixed) solve: the top/bottom boundary
    # contraction, bound by reading the FP64 Green's table once.  The table
    # is stored in blocked symmetric form (16 x 16 blocks (I, J), J <= I, of
    # targets x sources per source plane and mode): mmax x NB(NB+1)/2 x
    # (numz-2) x 256 x 8 B, NB = ceil((numr-2)/16).  The timed
    # span (events around it) is the contraction alone (the projections are
    # scaled as they are read; the face kernel sums its partials).
    nri, nzi, mmax = numr - 2, numz - 2, 30
    nb = (nri + 15) // 16
    ct_bytes = mmax * nb * (nb + 1) // 2 * nzi * 256 * 8
    ct_ms = mixed["stage_ms"]["contract_topbot"]
    ct_traffic = CONTRACT_TRAFFIC if (numr, numz, numphi) == (258, 130, 512) else None
    # the inverse DFT of the mixed cold solve (the stage from mark 3 to 4 is
    # this one kernel at N = 512): x's faces, delta1 and the correction sum
    # in (FP32 interior), phi out (a cold solve keeps its warm-start state as
    # the two corrections)
    c2 = (numr, numz, numphi) == (258, 130, 512)
    inv_bytes = 2 * 4 * (numr - 2) * (numz - 2) * numphi + (numr * numz - (numr - 2) * (numz - 2)) * numphi * 8 + \
        8 * numr * numz * numphi
    inv_ms = mixed["stage_ms"]["inverse_dft"]
    inv_traffic = INV_TRAFFIC if c2 else None
    inv_flop = INV_MFMA_FLOP if c2 else None
    # ADI half-step kernel of the FP64 path: 3 streams (read phi, read
    # rho-hat, write phi) x 8 B x interior points x modes per launch
    nsteps = 5 if args.warm else 20
    interior = (numr - 2) * (numz - 2) * numphi
    sw_bytes = 3 * 8 * interior
    sw_ms = fp64["stage_ms"]["adi"] / (2 * nsteps)
    sw_traffic = SWEEP_TRAFFIC if (numr, numz, numphi) == (258, 130, 512) else None

    # the host entry point (bh_solve): rho copied in from host memory, the
    # potential copied back, into caller-owned pageable arrays touched once
    # beforehand (as a Fortran caller's COMMON blocks are): the
    # PCIe-inclusive cost per solve, never the headline
    with P.PotentialPlan(numr, numz, numphi, precision="mixed", refine_passes=args.passes) as hp:
        pot_host = np.zeros_like(rho)
        hp.potential_solver(rho, tstep=0, out=pot_host)
        nh = max(5, args.steps // 10)
        t0 = time.perf_counter()
        for _ in range(nh):
            hp.potential_solver(rho, tstep=0, out=pot_host)
        host_ms = (time.perf_counter() - t0) * 1e3 / nh
    host_entry = {"ms_per_solve": host_ms, "solves_per_s": 1e3 / host_ms, "solves": nh,
                  "device_solve_ms": ms_per_step,
                  "copy_bytes_per_solve": 2 * rho.nbytes,
                  "note": "bh_solve with pageable host arrays: rho in and phi out over PCIe around the same "
                          "mixed cold solve"}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(numr, numz, numphi, rho, threads)

    out = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"potential_solver {'warm (5 ADI steps)' if args.warm else 'cold (20 ADI steps)'}"
                               f" {numr}x{numz}x{numphi}, isym=1, mmax=30, binary n=1 polytropes, "
                               "inputs resident in HBM",
                   "grid": [numr, numz, numphi],
                   "precision": f"mixed: FP32 MFMA DFT of rho, FP32 spectral ADI (sine transforms, split-bf16 "
                                f"M_r filter, Thomas sweeps), FP64 Green's tables, FP64 residual and "
                                f"{args.passes} refinement passes, FP64 inverse DFT", "tstep": tstep},
        "p50_ms": lat_p50,
        "p99_ms": lat_p99,
        "relL2_vs_reference_slices": (mixed["error_vs_reference_cold"] or {}).get("slices"),
        "error_per_pass": per_pass,
        "mixed_over_fp64": fp64["ms_per_step"] / mixed["ms_per_step"],
        "warm": {"mixed_ms": mixed["warm"]["ms_per_step"], "fp64_ms": fp64["warm"]["ms_per_step"],
                 "mixed_cold_then_warm_ms": mixed["warm"]["cold_then_warm_ms"],
                 "fp64_cold_then_warm_ms": fp64["warm"]["cold_then_warm_ms"],
                 "mixed_over_fp64": fp64["warm"]["ms_per_step"] / mixed["warm"]["ms_per_step"]},
        "mixed_table32": t32,
        "fp64_spectral": s64,
        "boundary_modes": modes,
        "mixed_over_fp64_spectral": s64["ms_per_step"] / mixed["ms_per_step"],
        "roofline": {"bound": "hbm", "kernel": "boundary Green's-table contraction (contract_tb64, the projections scaled as read)",
                     "achieved": ct_bytes / (ct_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ct_bytes / (ct_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": ct_traffic,
                     "bytes_per_launch": ct_bytes, "launch_ms": ct_ms},
        "roofline_dft_inverse": {"bound": "hbm", "kernel": "dft_inv8 (mixed cold inverse real DFT, radix 8 on FP64 MFMA)",
                                 "achieved": inv_bytes / (inv_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": inv_bytes / (inv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": inv_traffic,
                                 "bytes_per_launch": inv_bytes, "launch_ms": inv_ms,
                                 "mfma_fp64_tflops": inv_flop / (inv_ms * 1e-3) / 1e12 if inv_flop else None,
                                 "mfma_fp64_frac": inv_flop / (inv_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS if inv_flop else None},
        "roofline_fp64_sweep": {"bound": "hbm", "kernel": "line_sweep (FP64 ADI half-step)",
                                "achieved": sw_bytes / (sw_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": sw_bytes / (sw_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "traffic": sw_traffic, "bytes_per_launch": sw_bytes, "launch_ms": sw_ms},
        "cpu_baseline": cpu,
        "host_entry": host_entry,
        "mixed": mixed,
        "fp64": fp64,
    }
    # the hydro's source step on the device, fed by the potential in HBM
    out["source_step"] = run_source_step(numr, numz, numphi, args)
    if not args.no_workloads:
        # the batched (64 uneven grids) and served (Poisson SCF stream)
        # workloads, short runs
        import copy
        ab = copy.copy(args)
        ab.steps, ab.warmup = 10, 2
        bt = run_batched(ab, emit=False)
        out["batched"] = {k: bt[k] for k in ("value", "unit", "ms_per_step", "p50_ms", "p99_ms",
                                             "grouped_launches_per_step", "one_plan_at_a_time_ms", "batch_speedup",
                                             "batch_equals_single_plans_bitwise", "autotune")}
        out["batched"]["workload"] = bt["config"]["workload"]
        asv = copy.copy(args)
        asv.requests, asv.rate = 32, 40.0
        sv = run_serve(asv, emit=False)
        out["served"] = {k: sv[k] for k in ("value", "unit", "p50_ms", "p99_ms", "mean_ms", "solves", "wall_s",
                                            "peak_active", "tail")}
        out["served"]["workload"] = sv["config"]["workload"]
    print(json.dumps(out), flush=True)


def run_source_step(numr, numz, numphi, args):
    """bh_source_device(source_time = 2) on seeded fields of the grid: its
    time per call (HIP events on the stream), and the HBM bytes it moves
    (make_source_temp: s, a, p, rho, pot in, temps, tempa out; source: t, s,
    a, p, rho, pot, temps, tempa in, t, s, a out: 18 doubles per point)
    against the 8 TB/s peak; beside it the PCIe download of the potential
    that keeping the step on the device avoids."""
    import torch
    from bipoly_hydro_amd import potential as P
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from golden_cases import source_inputs
    f, sc = source_inputs((numr, numz, numphi), seed=5)
    d = {k: torch.from_numpy(v).cuda() for k, v in f.items()}
    ptrs = {k: v.data_ptr() for k, v in d.items()}
    prm = dict(dt=sc["dt"], omega=sc["omega"], tstep=20, bc=(1, 1, 1))
    stream = torch.cuda.current_stream()
    plan = P.PotentialPlan(numr, numz, numphi, precision="mixed")
    for _ in range(3):
        plan.source_device(2, ptrs, prm, stream.cuda_stream)
    steps = max(5, args.steps // 5)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        plan.source_device(2, ptrs, prm, stream.cuda_stream)
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    host = torch.empty(numphi, numz, numr, dtype=torch.float64, pin_memory=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    host.copy_(d["pot"])
    torch.cuda.synchronize()
    d2h_ms = (time.perf_counter() - t0) * 1e3
    plan.close()
    nbytes = 18 * 8 * numr * numz * numphi
    return {"ms_per_call": ms, "bytes_per_call": nbytes, "achieved_GBs": nbytes / (ms * 1e-3) / 1e9,
            "frac_of_hbm_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "potential_download_ms_pinned": d2h_ms,
            "workload": f"source(2) with make_source_temp and boundary conditions, {numr}x{numz}x{numphi}, "
                        "wall boundaries, device fields"}


def run_batched(args, emit=True):
    """64 independent grids of seeded uneven sizes, mixed precision, solved
    as one batch; a step is one solve of every grid (cold, 20 ADI steps)."""
    import ctypes
    import torch
    from bipoly_hydro_amd import density
    from bipoly_hydro_amd import potential as P

    dev = torch.device("cuda:0")
    grids = density.batched_grids(args.ngrids)
    print(f"batched: {len(grids)} grids, {sum(r * z * p for r, z, p, _ in grids) / 1e6:.1f} M points",
          file=sys.stderr, flush=True)
    d_rho = [torch.from_numpy(density.scaled_binary(r, z, p, seed=s)).to(dev) for r, z, p, s in grids]
    d_pot = [torch.empty_like(x) for x in d_rho]
    rp, pp = [x.data_ptr() for x in d_rho], [x.data_ptr() for x in d_pot]
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    t0 = time.perf_counter()
    b = P.PotentialBatch([(r, z, p) for r, z, p, _ in grids], precision="mixed", graph=True, autotune=True)
    setup_s = time.perf_counter() - t0
    b.solve_device(rp, pp, 0, sp)  # the first autotuning solve
    ntune = 1
    while b.stats()["tuning_left"] > 0:  # the others, one candidate each
        b.solve_device(rp, pp, 0, sp)
        ntune += 1
    for _ in range(args.warmup):  # then warm-up of the tuned graph
        b.solve_device(rp, pp, 0, sp)
    torch.cuda.synchronize()
    ref = [x.cpu() for x in d_pot]
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, c in ev:
        a.record(stream)
        b.solve_device(rp, pp, 0, sp)
        c.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(c) for a, c in ev])
    st = b.stats()
    same = all(torch.equal(x.cpu(), y) for x, y in zip(d_pot, ref))
    # the same grids one plan at a time, each launch on its own (no grouping)
    L = P.load_library()
    plans = []
    for i in range(len(grids)):
        h = ctypes.c_void_p()
        P._check(L.bh_batch_plan(b._h, i, ctypes.byref(h)))
        plans.append(h)
    for h, r, q in zip(plans, rp, pp):
        P._check(L.bh_solve_device(h, ctypes.c_void_p(r), ctypes.c_void_p(q), 0, ctypes.c_void_p(sp)))
    torch.cuda.synchronize()
    same_single = all(torch.equal(x.cpu(), y) for x, y in zip(d_pot, ref))
    nseq = max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(nseq):
        for h, r, q in zip(plans, rp, pp):
            L.bh_solve_device(h, ctypes.c_void_p(r), ctypes.c_void_p(q), 0, ctypes.c_void_p(sp))
    torch.cuda.synchronize()
    seq_ms = (time.perf_counter() - t0) * 1e3 / nseq
    ms = wall * 1e3 / args.steps
    # what the tuning buys: the same batch with every launch at its default
    # shape (graph on, no autotuning), timed the same way
    tune = b.tuning()
    b.close()
    bd = P.PotentialBatch([(r, z, p) for r, z, p, _ in grids], precision="mixed", graph=True, autotune=False)
    for _ in range(1 + args.warmup):
        bd.solve_device(rp, pp, 0, sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bd.solve_device(rp, pp, 0, sp)
    torch.cuda.synchronize()
    default_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    same_default = all(torch.equal(x.cpu(), y) for x, y in zip(d_pot, ref))
    bd.close()
    tuning = {}
    for ax, v in tune.items():
        ch = [c for c in v["choice"] if c >= 0]
        if ch:
            tuning[ax] = {"steps_tuned": len(ch), "choices": {str(c): ch.count(c) for c in sorted(set(ch))},
                          "gain_ms_in_tuning": v["gain_ms"]}
    out = {
        "metric": "batched uneven Poisson solves/s (grids solved per second), mixed precision",
        "value": len(grids) * 1e3 / ms, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(grids)} independent grids, sizes drawn by seed from numr in "
                               "{66,130,194,258}, numz in {34,66,130}, numphi in {64,128,256,512}; "
                               "n=1 polytrope binaries, q in [0.3,1], separation in [0.4,0.8] x grid "
                               "radius; cold solves", "grids": [g[:3] for g in grids]},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "grouped_launches_per_step": st["launches"], "graph": st["graph"],
        "autotune": {"tuning_solves": ntune, "axes": tuning, "default_shapes_ms": default_ms,
                     "tuned_ms": ms, "speedup_over_default": default_ms / ms},
        "setup_s": setup_s,
        "one_plan_at_a_time_ms": seq_ms, "batch_speedup": seq_ms / ms,
        "batch_equals_single_plans_bitwise": bool(same and same_single and same_default),
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def run_serve(args, emit=True):
    """Poisson arrivals at a fixed rate of SCF requests: grid sizes drawn as in
    the batched workload, 20-50 dependent solves each (cold, then warm) on a
    slowly changing binary of bipolytropes (core n = 3, envelope n = 1.5, core
    radius 0.3 R) generated on the device.  A warm-up pass of one short
    request per grid size builds the plans first (steady state)."""
    from bipoly_hydro_amd import density
    from bipoly_hydro_amd import potential as P

    grids = density.batched_grids(args.requests, seed=777)
    u = density.splitmix64_uniform(991, 6 * args.requests)
    reqs = []
    for i, (r, z, p, s) in enumerate(grids):
        q = 0.3 + 0.7 * u[6 * i]
        rgrid = min(r - 3, z - 2)
        sep = (0.4 + 0.4 * u[6 * i + 1]) * rgrid
        r1 = max(3.0, min(0.45 * sep, 0.45 * (z - 4), 0.45 * (r - 3)))
        r2 = max(2.0, r1 * q ** (1.0 / 3.0))
        nsolves = 20 + int(31 * u[6 * i + 2])
        model = dict(radius=(r1, r2), separation=sep, drift=-0.002 * (0.5 + u[6 * i + 3]), noise=1e-4, seed=s)
        reqs.append(((r, z, p), min(nsolves, 50), model))
    gaps = -np.log(1.0 - u[6 * np.arange(args.requests) + 4]) / args.rate  # exponential inter-arrival times
    srv = P.PotentialServer(precision="mixed")
    sizes = sorted(set(g for g, _, _ in reqs))
    t0 = time.perf_counter()
    for g in sizes:
        srv.wait(srv.submit(g, 1, dict(radius=(3.0, 2.0), separation=8.0)))
    setup_s = time.perf_counter() - t0
    print(f"serve: warm-up of {len(sizes)} grid sizes took {setup_s:.1f} s", file=sys.stderr, flush=True)
    base = srv.stats()
    ids = []
    t0 = time.perf_counter()
    t_next = t0
    for (g, n, m), gap in zip(reqs, gaps):
        t_next += gap
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        ids.append(srv.submit(g, n, m))
    lat, first, later = [], [], []
    by_lane = {0: [], 1: [], 2: []}
    for i, (g, _, _) in zip(ids, reqs):
        res, _, l = srv.wait(i)
        if res["status"] != 0:
            raise RuntimeError(f"request {i} failed: {res}")
        lat.extend(l.tolist())
        first.append(l[0])
        later.extend(l[1:].tolist())
        by_lane[serve_lane(g)].extend(l.tolist())
    wall = time.perf_counter() - t0
    st = srv.stats()
    log = srv.iterations()
    srv.close()
    lat = np.array(lat)
    tail = serve_tail(log, lat, len(sizes), (first, later))
    nsolves = int(st["solves"] - base["solves"])
    out = {
        "metric": "served SCF stream: Poisson solves/s with per-solve p50/p99 latency, mixed precision",
        "value": nsolves / wall, "unit": "solves/s", "n_gpus": 1, "steps": args.requests, "warmup": len(sizes),
        "ms_per_step": wall * 1e3 / args.requests, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{args.requests} requests, Poisson arrivals at {args.rate}/s; grid sizes drawn "
                               "by seed from numr in {66,130,194,258}, numz in {34,66,130}, numphi in "
                               "{64,128,256,512}; 20-50 dependent solves per request (cold, then warm); "
                               "bipolytrope binary (core n=3, envelope n=1.5, core radius 0.3 R) drifting "
                               "by 0.1-0.3% per solve, generated on the device"},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()), "solves": nsolves, "wall_s": wall,
        "iterations": int(st["iterations"] - base["iterations"]), "peak_active": st["peak_active"],
        "peak_device_bytes": st["peak_bytes"], "warmup_s": setup_s, "tail": tail,
        # per lane of the serving loop (grid points < 2^20, < 2^22, larger)
        "lanes_p50_p99_ms": {str(k): ([float(np.percentile(v, 50)), float(np.percentile(v, 99)), len(v)]
                                      if v else None) for k, v in by_lane.items()},
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def serve_lane(grid):
    """The serving loop's lane of a grid: its points below 2^20, below 2^22,
    or larger (each lane's iterations run on a stream of their own)."""
    pts = grid[0] * grid[1] * grid[2]
    return 0 if pts < 1 << 20 else 1 if pts < 1 << 22 else 2


def serve_tail(log, lat, skip, first=None):
    """Where the served latency tail comes from, from the loop's iteration
    log (rows: t_admit, t_enq0, t_enq1, t_done, solves, plans created,
    points, lane; the first `skip` iterations are the warm-up).  A solve's
    latency is its iteration's completion minus the previous one's in the
    same lane, so the interval between a lane's completions, weighted by the
    solves in the iteration, stands for the per-solve latency; it is split by
    whether the admission before the iteration created a plan, and related
    to the points solved."""
    later = None
    if first is not None:
        first, later = np.asarray(first[0]), np.asarray(first[1])
    log = log[skip:]
    if len(log) < 3:
        return None
    # from the later of the lane's previous completion and this iteration's
    # enqueue (an iteration after an idle gap starts when it is enqueued)
    ivs, rows = [], []
    for lane in np.unique(log[:, 7]):
        lg = log[log[:, 7] == lane]
        lg = lg[np.argsort(lg[:, 3])]
        if len(lg) < 2:
            continue
        ivs.append((lg[1:, 3] - np.maximum(lg[:-1, 3], lg[1:, 1])) * 1e3)
        rows.append(lg[1:])
    if not rows:
        return None
    iv = np.concatenate(ivs)
    rows = np.concatenate(rows)
    w = rows[:, 4]
    p99 = float(np.percentile(lat, 99))
    tail_it = iv >= p99
    created = rows[:, 5] > 0
    admit_ms = (rows[:, 1] - rows[:, 0]) * 1e3
    return {
        "iterations": int(len(rows)),
        "interval_ms_p50_p99": [float(np.percentile(iv, 50)), float(np.percentile(iv, 99))],
        "solves_per_iteration_p50_max": [float(np.median(w)), float(w.max())],
        "mpoints_per_iteration_p50_p99": [float(np.percentile(rows[:, 6], 50)) / 1e6,
                                          float(np.percentile(rows[:, 6], 99)) / 1e6],
        "tail_iterations": int(tail_it.sum()),
        "tail_iterations_with_plan_creation": int((tail_it & created).sum()),
        "iterations_with_plan_creation": int(created.sum()),
        "admission_ms_p50_max": [float(np.median(admit_ms)), float(admit_ms.max())],
        "mean_mpoints_tail_vs_all": [float(rows[tail_it, 6].mean()) / 1e6 if tail_it.any() else None,
                                     float(rows[:, 6].mean()) / 1e6],
        # first solves of a request count from the submission (admission at
        # the next iteration boundary, behind the iteration in flight)
        "first_solve_ms_p50_p99": ([float(np.percentile(first, 50)), float(np.percentile(first, 99))]
                                   if first is not None and len(first) else None),
        "later_solves_ms_p50_p99": ([float(np.percentile(later, 50)), float(np.percentile(later, 99))]
                                    if later is not None and len(later) else None),
    }


if __name__ == "__main__":
    main()
