"""Seeded inputs of the golden cases in tests/golden/ (see make_golden.py)."""
import os

import numpy as np

from bipoly_hydro_amd import density

SMALL_CASES = {
    "c1_66x66x64": (66, 66, 64, lambda s: density.config1_density(seed=s)),
    "g34x34x32": (34, 34, 32, lambda s: density.lane_emden_n1_binary(
        34, 34, 32, radius_cells=(6.0, 4.0), centre_cells=(8.0, -8.0), noise=1e-6, seed=s)),
    "g42x26x16": (42, 26, 16, lambda s: density.lane_emden_n1_binary(
        42, 26, 16, radius_cells=(7.0, 5.0), centre_cells=(12.0, -12.0), noise=1e-6, seed=s)),
}


def rell2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def load_case(golden_dir, name):
    numr, numz, numphi, build = SMALL_CASES[name]
    g = np.load(os.path.join(golden_dir, name + ".npz"))
    s1, s2 = (int(x) for x in g["seeds"])
    rho1 = build(s1)
    rho2 = rho1 * 1.01 + 1e-3 * build(s2)
    # the density generator must still produce the inputs the reference saw
    np.testing.assert_allclose([rho1.sum(), rho2.sum(), np.abs(rho1).max()], g["rho_sums"], rtol=1e-14)
    return (numr, numz, numphi), rho1, rho2, g["pot1"], g["pot2"]


def c2_slices(pot, tag):
    """The parts of a 258 x 130 x 512 potential kept in c2_258x130x512.npz
    for solve `tag` ("1" cold, "2" warm); see make_golden.c2_case."""
    if tag == "1":
        return {"mid1": pot[:, 64, :], "mer1": pot[0], "top1": pot[::8, -1, :], "bot1": pot[::8, 0, :],
                "side1": pot[::8, :, -1]}
    return {"mid2": pot[::4, 64, :], "mer2": pot[0]}


def c2_digests(pot, weights):
    """Digests of a whole 258 x 130 x 512 potential (numpy (L, K, J)): the L2
    norm of every plane L, of every K slab and of every J column, and the
    weighted sum over J of every (L, K) row, sum_J w_J pot[L, K, J] (first
    order in any error, so a fault confined to one row moves its entry)."""
    return {"plane_norm": np.sqrt(np.einsum("lkj,lkj->l", pot, pot)),
            "k_norm": np.sqrt(np.einsum("lkj,lkj->k", pot, pot)),
            "j_norm": np.sqrt(np.einsum("lkj,lkj->j", pot, pot)),
            "row_dot": pot @ weights}
