"""Writes the golden vectors of tests/golden/ from the reference's own
Fortran/MPI potential_solver, built by oracle/ref/Makefile from the sources
under /root/reference (build container only).

    python tests/golden/make_golden.py [--c2 | --c2-only] [--cont] [--source] [--sym | --sym-only]

Each solver case: a seeded density (bipoly_hydro_amd.density), a cold
solve (tstep = 0) and a warm solve (tstep = 2) of a perturbed density
starting from the cold result.  Output: <case>.npz with rho1, pot1, rho2,
pot2 (small grids) or slices and norms (258 x 130 x 512).  --cont: the
continuation files the reference's writer makes of known fields
(cont_<grid>_<ranks>/).  --source: the reference's momentum source step
on seeded fields (source_<grid>.npz).  --sym: the solver cases of
golden_cases.SYM_CASES, with the reference's run-time flag isym = 2 and 3
(s<isym>_<grid>.npz).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from bipoly_hydro_amd import density  # noqa: E402
from golden_cases import (SOURCE_CASES, SOURCE_GRID, SOURCE_PROCS, SYM_CASES, c2_digests,  # noqa: E402
                          source_inputs)

# name, numr, numz, numphi, numr_procs, numz_procs, density builder
CASES = [
    ("c1_66x66x64", 66, 66, 64, 2, 2, lambda s: density.config1_density(seed=s)),
    ("g34x34x32", 34, 34, 32, 2, 2,
     lambda s: density.lane_emden_n1_binary(34, 34, 32, radius_cells=(6.0, 4.0), centre_cells=(8.0, -8.0),
                                            noise=1e-6, seed=s)),
    ("g42x26x16", 42, 26, 16, 2, 2,
     lambda s: density.lane_emden_n1_binary(42, 26, 16, radius_cells=(7.0, 5.0), centre_cells=(12.0, -12.0),
                                            noise=1e-6, seed=s)),
]


def run_case(name, numr, numz, numphi, nrp, nzp, rho1, rho2, workdir, isym=1, build=None):
    build = build or name
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle/ref/Makefile"), f"NAME={build}",
                    f"NUMR={numr}", f"NUMZ={numz}", f"NUMPHI={numphi}", f"NRP={nrp}", f"NZP={nzp}", "-j8"],
                   check=True)
    exe = os.path.join(ROOT, "oracle/_ref", build, "bh_ref")
    f = {k: os.path.join(workdir, f"{name}_{k}.bin") for k in ("rho1", "pot1", "rho2", "pot2")}
    rho1.tofile(f["rho1"])
    rho2.tofile(f["rho2"])
    cmd = (f"ulimit -s unlimited; BH_REF_ISYM={isym} /opt/conda/bin/mpiexec -n {nrp * nzp} {exe} "
           f"{f['rho1']} {f['pot1']} {f['rho2']} {f['pot2']}")
    out = subprocess.run(cmd, shell=True, check=True, capture_output=True, text=True).stdout
    shape = rho1.shape
    return (np.fromfile(f["pot1"]).reshape(shape), np.fromfile(f["pot2"]).reshape(shape), out)


def c2_case(wd):
    """258 x 130 x 512 on 4 x 2 ranks: slices, norms and digests (the full
    field is 137 MB).  Cold solve: the mid-plane row K = 65 (all J, all L),
    the meridional plane L = 1, and the top, bottom and side faces at every
    8th L.  Warm solve: row K = 65 at every 4th L and the plane L = 1.  Both
    solves: c2_digests of the whole field."""
    name, numr, numz, numphi, nrp, nzp = "c2_258x130x512", 258, 130, 512, 4, 2
    rho1 = density.config2_density(seed=1234)
    rho2 = rho1 * 1.01 + 1e-3 * density.config2_density(seed=99)
    pot1, pot2, log = run_case(name, numr, numz, numphi, nrp, nzp, rho1, rho2, wd)
    weights = np.random.default_rng(2024).standard_normal(numr)
    dig = {}
    for tag, pot in (("1", pot1), ("2", pot2)):
        for k, v in c2_digests(pot, weights).items():
            dig[f"{k}{tag}"] = v
    np.savez(os.path.join(HERE, f"{name}.npz"), digest_weights=weights, **dig,
             rho_sums=np.array([rho1.sum(), rho2.sum(), np.abs(rho1).max()]),
             seeds=np.array([1234, 99]), procs=np.array([nrp, nzp]),
             mid1=pot1[:, 64, :], mer1=pot1[0], top1=pot1[::8, -1, :], bot1=pot1[::8, 0, :],
             side1=pot1[::8, :, -1], norm1=np.array([np.linalg.norm(pot1), pot1.sum(), np.abs(pot1).max()]),
             mid2=pot2[::4, 64, :], mer2=pot2[0],
             norm2=np.array([np.linalg.norm(pot2), pot2.sum(), np.abs(pot2).max()]))
    print(name, log.strip().replace("\n", " | "))


CONT_CASES = [("cont_42x26x16_2x2", 42, 26, 16, 2, 2), ("cont_34x18x8_4x2", 34, 18, 8, 4, 2)]
CONT_NAMES = ["s", "t", "a", "rho", "tau", "pot", "tempa", "frac1", "frac2"]
CONT_TSTEP = 12345


def cont_fields(numr, numz, numphi):
    """The nine fields written to the continuation-file fixtures: field f at
    (J, K, L) (1-based) is (f+1) 1e9 + J 1e6 + K 1e3 + L, exact in FP64, so
    every value names its own field and position."""
    J = np.arange(1, numr + 1)[None, None, :]
    K = np.arange(1, numz + 1)[None, :, None]
    L = np.arange(1, numphi + 1)[:, None, None]
    return [((f + 1) * 1e9 + J * 1e6 + K * 1e3 + L).astype(np.float64) for f in range(len(CONT_NAMES))]


def cont_case(wd, name, numr, numz, numphi, nrp, nzp):
    """The reference's own continuation writer (ritecont_frac_new) on a
    process grid of nrp x nzp ranks; its files become the fixture."""
    import shutil
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle/ref/Makefile"), f"NAME=h_{name}",
                    f"NUMR={numr}", f"NUMZ={numz}", f"NUMPHI={numphi}", f"NRP={nrp}", f"NZP={nzp}", "hydro", "-j8"],
                   check=True)
    exe = os.path.join(ROOT, "oracle/_ref", f"h_{name}", "bh_hydro")
    run = os.path.join(wd, name)
    os.makedirs(os.path.join(run, "output", "conts"), exist_ok=True)
    for n, v in zip(CONT_NAMES, cont_fields(numr, numz, numphi)):
        v.tofile(os.path.join(run, f"in.{n}.bin"))
    cmd = f"ulimit -s unlimited; cd {run} && /opt/conda/bin/mpiexec -n {nrp * nzp} {exe} cont in {CONT_TSTEP}"
    subprocess.run(cmd, shell=True, check=True, capture_output=True, text=True)
    dst = os.path.join(HERE, name)
    os.makedirs(dst, exist_ok=True)
    for f in sorted(os.listdir(os.path.join(run, "output", "conts"))):
        shutil.copy(os.path.join(run, "output", "conts", f), os.path.join(dst, f))
    print(name, sorted(os.listdir(dst)))


def source_case(wd):
    """The reference's own momentum source step (source.F90 with
    make_source_temp, source_bc, make_source_temp_bc and comm) on a 2 x 2
    process grid, for each of SOURCE_CASES, on the seeded inputs of
    golden_cases.source_inputs; source_<grid>.npz holds the outputs."""
    numr, numz, numphi = SOURCE_GRID
    nrp, nzp = SOURCE_PROCS
    name = f"h_src_{numr}x{numz}x{numphi}"
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle/ref/Makefile"), f"NAME={name}",
                    f"NUMR={numr}", f"NUMZ={numz}", f"NUMPHI={numphi}", f"NRP={nrp}", f"NZP={nzp}", "hydro", "-j8"],
                   check=True)
    exe = os.path.join(ROOT, "oracle/_ref", name, "bh_hydro")
    f, sc = source_inputs()
    run = os.path.join(wd, "source")
    os.makedirs(run, exist_ok=True)
    for k, v in f.items():
        v.tofile(os.path.join(run, f"in.{k}.bin"))
    com = np.concatenate([[sc["dt"], sc["omega"], sc["R_com"]], sc["cylin_a_com"].T.reshape(-1),
                          sc["cylin_v_com"].T.reshape(-1)])
    com.tofile(os.path.join(run, "in.com.bin"))
    out = {}
    for tag, st, tstep, bc in SOURCE_CASES:
        cmd = (f"ulimit -s unlimited; cd {run} && /opt/conda/bin/mpiexec -n {nrp * nzp} {exe} source in out_{tag} "
               f"{st} {tstep} {bc[0]} {bc[1]} {bc[2]}")
        subprocess.run(cmd, shell=True, check=True, capture_output=True, text=True)
        for k in ("s", "t", "a", "temps", "tempa"):
            out[f"{tag}_{k}"] = np.fromfile(os.path.join(run, f"out_{tag}.{k}.bin")).reshape(numphi, numz, numr)
    np.savez(os.path.join(HERE, f"source_{numr}x{numz}x{numphi}.npz"), **out)
    print("source", sorted(out))


def main():
    os.makedirs(os.path.join(ROOT, "oracle", "_ref"), exist_ok=True)
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "oracle", "_ref")) as wd:
        only = "--c2-only" in sys.argv or "--sym-only" in sys.argv
        for name, numr, numz, numphi, nrp, nzp, build in ([] if only else CASES):
            rho1 = build(1234)
            rho2 = rho1 * 1.01 + 1e-3 * build(99)
            pot1, pot2, log = run_case(name, numr, numz, numphi, nrp, nzp, rho1, rho2, wd)
            np.savez(os.path.join(HERE, f"{name}.npz"), pot1=pot1, pot2=pot2,
                     rho_sums=np.array([rho1.sum(), rho2.sum(), np.abs(rho1).max()]),
                     seeds=np.array([1234, 99]), procs=np.array([nrp, nzp]))
            print(name, log.strip().replace("\n", " | "))
        if "--c2" in sys.argv or "--c2-only" in sys.argv:
            c2_case(wd)
        if "--source" in sys.argv:
            source_case(wd)
        if "--sym" in sys.argv or "--sym-only" in sys.argv:
            for name, (numr, numz, numphi, isym, build) in SYM_CASES.items():
                rho1 = build(1234)
                rho2 = rho1 * 1.01 + 1e-3 * build(99)
                pot1, pot2, log = run_case(name, numr, numz, numphi, 2, 2, rho1, rho2, wd, isym=isym,
                                           build=f"sym_{numr}x{numz}x{numphi}")
                np.savez(os.path.join(HERE, f"{name}.npz"), pot1=pot1, pot2=pot2, isym=np.array(isym),
                         rho_sums=np.array([rho1.sum(), rho2.sum(), np.abs(rho1).max()]),
                         seeds=np.array([1234, 99]), procs=np.array([2, 2]))
                print(name, log.strip().replace("\n", " | "))
        if "--cont" in sys.argv:
            for c in CONT_CASES:
                cont_case(wd, *c)


if __name__ == "__main__":
    main()
