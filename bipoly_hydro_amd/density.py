"""Seeded synthetic densities on the solver's cylindrical grid.

The grid is the reference's (setup_frac.F90:428-463):
    r_J   = (J - 1.5) dr          J = 1..numr   (J = 1 is the axis ghost)
    z_K   = (K - 0.5 - numz/2) dz K = 1..numz   (isym = 1)
    z_K   = (K - 1.5) dz          K = 1..numz   (isym = 2, 3: K = 1 mirrors K = 2)
    phi_L = (L - 1) dphi          L = 1..numphi, dphi = 2 pi / numphi
with dr = dz (default 1/127, setup_frac.F90:374-376).  Arrays are returned
in the reference's Fortran order (numr, numz, numphi): a numpy array of
shape (numphi, numz, numr), C-contiguous, so that r is fastest in memory.

Random numbers come from splitmix64 written here, so that a seed names the
same density on every machine and numpy version.
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n doubles uniform in [0, 1) from a splitmix64 stream (Steele et al.,
    OOPSLA 2014), vectorised: state_i = seed + (i+1) * golden_gamma."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + idx * gamma
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def grid_coords(numr: int, numz: int, numphi: int, dr: float = 1.0 / 127.0, isym: int = 1):
    """Cell-centre coordinates (r, z, phi) as 1-D arrays."""
    r = (np.arange(1, numr + 1, dtype=np.float64) - 1.5) * dr
    if isym == 1:
        z = (np.arange(1, numz + 1, dtype=np.float64) - 0.5 - (numz // 2)) * dr
    else:
        z = (np.arange(1, numz + 1, dtype=np.float64) - 1.5) * dr
    phi = np.arange(numphi, dtype=np.float64) * (2.0 * math.pi / numphi)
    return r, z, phi


def _cartesian(numr, numz, numphi, dr, isym=1):
    r, z, phi = grid_coords(numr, numz, numphi, dr, isym)
    x = r[None, None, :] * np.cos(phi)[:, None, None]
    y = r[None, None, :] * np.sin(phi)[:, None, None]
    zz = np.broadcast_to(z[None, :, None], (numphi, numz, numr))
    return x, y, zz


def lane_emden_n1_binary(numr: int, numz: int, numphi: int, *, rho_c=(1.0, 0.7),
                         radius_cells=(10.0, 8.0), centre_cells=(15.0, -15.0),
                         noise=1e-6, seed: int = 1234, dr: float = 1.0 / 127.0, isym: int = 1) -> np.ndarray:
    """Two n = 1 polytropes, rho = rho_c sin(pi d/R)/(pi d/R) for d < R, centred
    on the x axis in the equatorial plane, plus seeded noise U(-1,1)*noise*rho_c
    inside each star.  Returns float64 of shape (numphi, numz, numr).  With
    isym = 2, 3 the grid holds the upper half (z >= 0 and the mirror cell)."""
    x, y, z = _cartesian(numr, numz, numphi, dr, isym)
    rho = np.zeros((numphi, numz, numr), dtype=np.float64)
    for s, (rc, rad, cx) in enumerate(zip(rho_c, radius_cells, centre_cells)):
        R = rad * dr
        d = np.sqrt((x - cx * dr) ** 2 + y ** 2 + z ** 2)
        inside = d < R
        u = math.pi * d[inside] / R
        val = np.where(u > 0, np.sin(u) / np.where(u > 0, u, 1.0), 1.0)
        rho[inside] += rc * val
        if noise:
            rnd = splitmix64_uniform(seed * 1000003 + s, int(inside.sum()))
            rho[inside] += (2.0 * rnd - 1.0) * noise * rc
    return rho


def config1_density(seed: int = 1234) -> np.ndarray:
    """66 x 66 x 64: stars of radius 10 and 8 cells at x = +-15 cells."""
    return lane_emden_n1_binary(66, 66, 64, seed=seed)


def config2_density(seed: int = 1234, numr=258, numz=130, numphi=512) -> np.ndarray:
    """258 x 130 x 512: stars of radius 40 and 32 cells at x = +-60 cells,
    noise 1e-4 rho_c."""
    return lane_emden_n1_binary(numr, numz, numphi, radius_cells=(40.0, 32.0),
                                centre_cells=(60.0, -60.0), noise=1e-4, seed=seed)


def scaled_binary(numr: int, numz: int, numphi: int, seed: int, *, mass_ratio=None,
                  separation=None, dr: float = 1.0 / 127.0) -> np.ndarray:
    """Binary of n = 1 spheres sized to the grid: mass ratio q in [0.3, 1] and
    separation in [0.4, 0.8] x grid radius drawn by seed unless given,
    rho_c = 1 for both stars (mass ratio set through the radii, M ~ R^3)."""
    u = splitmix64_uniform(seed * 7919 + 17, 2)
    q = mass_ratio if mass_ratio is not None else 0.3 + 0.7 * u[0]
    sep = separation if separation is not None else 0.4 + 0.4 * u[1]
    rgrid = min(numr - 3, (numz - 2) // 2 * 2) * 1.0
    a = sep * rgrid
    r1 = min(0.45 * a, 0.45 * (numz - 4) / 1.0, 0.45 * (numr - 3))
    r1 = max(r1, 3.0)
    r2 = max(r1 * q ** (1.0 / 3.0), 2.0)
    # centres about the centre of mass, inside the grid radius
    x1 = a * q / (1.0 + q)
    x2 = -a / (1.0 + q)
    lim = numr - 3 - 1.0
    x1 = min(x1, lim - r1)
    x2 = max(x2, -(lim - r2))
    return lane_emden_n1_binary(numr, numz, numphi, rho_c=(1.0, 1.0), radius_cells=(r1, r2),
                                centre_cells=(x1, x2), noise=1e-6, seed=seed, dr=dr)


def batched_grids(n: int = 64, seed: int = 4242):
    """Grid sizes of the batched workload: numr in {66, 130, 194, 258}, numz in
    {34, 66, 130}, numphi in {64, 128, 256, 512}, drawn by seed; returns
    [(numr, numz, numphi, density_seed)]."""
    u = splitmix64_uniform(seed, 3 * n)
    nr, nz, npf = (66, 130, 194, 258), (34, 66, 130), (64, 128, 256, 512)
    out = []
    for i in range(n):
        out.append((nr[min(3, int(u[3 * i] * 4))], nz[min(2, int(u[3 * i + 1] * 3))],
                    npf[min(3, int(u[3 * i + 2] * 4))], seed * 1000 + i))
    return out


def _splitmix_hash_uniform(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser of x + golden gamma, as a double in [0, 1)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def bipolytrope_binary(numr: int, numz: int, numphi: int, model: dict, solve_index: int, profile: np.ndarray,
                       dr: float = 1.0 / 127.0, isym: int = 1) -> np.ndarray:
    """Host statement of the served workload's density (the device generator
    of bh_density.hip): two stars with rho_c f(d/R), f the tabulated
    bipolytrope profile interpolated linearly, centred about the centre of
    mass at separation * (1 + drift * solve_index), plus hashed noise."""
    rho_c = model.get("rho_c", (1.0, 1.0))
    rad = model["radius"]
    sep = model["separation"] * (1.0 + model.get("drift", 0.0) * solve_index)
    m1, m2 = rho_c[0] * rad[0] ** 3, rho_c[1] * rad[1] ** 3
    cx = (sep * m2 / (m1 + m2) * dr, -sep * m1 / (m1 + m2) * dr)
    noise, seed = model.get("noise", 0.0), int(model.get("seed", 0))
    nprof = len(profile)
    J = np.arange(1, numr + 1, dtype=np.float64)
    K = np.arange(1, numz + 1, dtype=np.float64)
    r = (J - 1.5) * dr
    z = (K - 0.5 - float(numz // 2)) * dr if isym == 1 else (K - 1.5) * dr
    phi = np.arange(numphi, dtype=np.float64) * (2.0 * math.pi * (1.0 / numphi))
    x = r[None, None, :] * np.cos(phi)[:, None, None]
    y = r[None, None, :] * np.sin(phi)[:, None, None]
    zz = z[None, :, None]
    idx = np.arange(numphi * numz * numr, dtype=np.uint64).reshape(numphi, numz, numr)
    rho = np.zeros((numphi, numz, numr))
    for s in range(2):
        R = rad[s] * dr
        d = np.sqrt((x - cx[s]) ** 2 + y ** 2 + zz ** 2)
        inside = d < R
        u = d[inside] / R * (nprof - 1)
        i = np.minimum(u.astype(np.int64), nprof - 2)
        w = u - i
        rho[inside] += rho_c[s] * (profile[i] + w * (profile[i + 1] - profile[i]))
        if noise:
            h = _splitmix_hash_uniform(np.uint64(seed * 1000003 + s * (1 << 32)) + idx[inside])
            rho[inside] += (2.0 * h - 1.0) * noise * rho_c[s]
    return rho
