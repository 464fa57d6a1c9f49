"""ctypes bindings for the C restatement in bh_oracle.c.

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.

Arrays use the reference layout (Fortran (numr, numz, numphi)), held in
numpy as C-contiguous (numphi, numz, numr) float64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbh_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.bho_plan_create.restype = ctypes.c_void_p
        L.bho_plan_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.bho_plan_destroy.argtypes = [ctypes.c_void_p]
        L.bho_potential_solver.restype = ctypes.c_int
        L.bho_potential_solver.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.bho_bessel.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.bho_helmadi_n.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.bho_gammln.restype = ctypes.c_double
        L.bho_gammln.argtypes = [ctypes.c_double]
        for f in ("bho_rf", "bho_rd"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_double] * 3
        for f in ("bho_ellk", "bho_elle"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_double]
        for f in ("bho_tmr", "bho_bmr", "bho_smz", "bho_coefh", "bho_dcoefh", "bho_zhf_g", "bho_rhf_g"):
            getattr(L, f).restype = _dp
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.bho_qhalf_at.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.bho_dft_forward.argtypes = [ctypes.c_int, _dp, _dp]
        L.bho_dft_inverse.argtypes = [ctypes.c_int, _dp, _dp]
        L.bho_dt_schedule.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.bho_source.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int, ctypes.c_int, ip] +
                                 [ctypes.c_double] * 3 + [_dp] * 10)
        L.bho_make_source_temp.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int, ip] +
                                           [ctypes.c_double] * 3 + [_dp] * 9)
        L.bho_max_threads.restype = ctypes.c_int
        L.bho_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class OraclePlan:
    """Grid + coefficients (+ Green's tables) for one geometry and symmetry
    flag isym (1 none, 2 equatorial, 3 equatorial + pi)."""

    def __init__(self, numr, numz, numphi, dr=1.0 / 127.0, mmax=30, with_tables=True, isym=1):
        """with_tables: True computes the Green's tables, False skips them,
        "filled" allocates them filled with a nonzero constant, every page
        written (timing of the solve's arithmetic and table stream; the
        potential is then meaningless)."""
        self.shape = (numphi, numz, numr)
        self.numr, self.numz, self.numphi, self.mmax, self.isym = numr, numz, numphi, mmax, isym
        mode = 2 if with_tables == "filled" else (1 if with_tables else 0)
        self._h = lib().bho_plan_create(numr, numz, numphi, int(isym), dr, mmax, mode)
        if not self._h:
            raise ValueError("oracle: invalid geometry")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().bho_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def solve(self, rho: np.ndarray, potp: np.ndarray | None = None, tstep: int = 0) -> np.ndarray:
        """potential_solver: cold (tstep <= 1) or warm from potp."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        out = np.zeros(self.shape) if potp is None else np.array(potp, dtype=np.float64, order="C")
        lib().bho_potential_solver(self._h, _ptr(rho), _ptr(out), int(tstep))
        return out

    def bessel(self, rho: np.ndarray, potp: np.ndarray | None = None) -> np.ndarray:
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        out = np.zeros(self.shape) if potp is None else np.array(potp, dtype=np.float64, order="C")
        lib().bho_bessel(self._h, _ptr(rho), _ptr(out))
        return out

    def helmadi(self, nsteps: int, rhop: np.ndarray, potp: np.ndarray) -> np.ndarray:
        r = np.array(rhop, dtype=np.float64, order="C")
        p = np.array(potp, dtype=np.float64, order="C")
        lib().bho_helmadi_n(self._h, int(nsteps), _ptr(r), _ptr(p))
        return p

    def dt_schedule(self, nsteps: int) -> np.ndarray:
        dt = np.zeros(nsteps)
        lib().bho_dt_schedule(self._h, nsteps, _ptr(dt))
        return dt

    def tables(self):
        """(tmr, bmr, smz) as arrays [m][target][kz][jr] over interior sources
        (bmr is None for isym = 2, 3, which have no bottom face)."""
        nri, nzi = self.numr - 2, self.numz - 2
        L = lib()
        t = np.ctypeslib.as_array(L.bho_tmr(self._h), shape=(self.mmax, self.numr, nzi, nri)).copy()
        b = (np.ctypeslib.as_array(L.bho_bmr(self._h), shape=(self.mmax, self.numr, nzi, nri)).copy()
             if self.isym == 1 else None)
        s = np.ctypeslib.as_array(L.bho_smz(self._h), shape=(self.mmax, self.numz, nzi, nri)).copy()
        return t, b, s

    def series_coefficients(self):
        L = lib()
        coefh = np.ctypeslib.as_array(L.bho_coefh(self._h), shape=(self.mmax,)).copy()
        dcoefh = np.ctypeslib.as_array(L.bho_dcoefh(self._h), shape=(self.mmax, 111)).copy()
        return coefh, dcoefh

    def qhalf(self, x: float) -> np.ndarray:
        q = np.zeros(self.mmax)
        lib().bho_qhalf_at(self._h, float(x), _ptr(q))
        return q


def gammln(x: float) -> float:
    return lib().bho_gammln(float(x))


def rf(x, y, z) -> float:
    return lib().bho_rf(float(x), float(y), float(z))


def rd(x, y, z) -> float:
    return lib().bho_rd(float(x), float(y), float(z))


def ellk(k) -> float:
    return lib().bho_ellk(float(k))


def elle(k) -> float:
    return lib().bho_elle(float(k))


def dft_forward(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().bho_dft_forward(len(x), _ptr(x), _ptr(out))
    return out


def dft_inverse(c: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(c, dtype=np.float64)
    out = np.zeros_like(c)
    lib().bho_dft_inverse(len(c), _ptr(c), _ptr(out))
    return out


def set_threads(n: int) -> None:
    lib().bho_set_threads(int(n))


def max_threads() -> int:
    return lib().bho_max_threads()


def source(fields: dict, *, source_time: int, tstep: int, bc, dt: float, omega: float, R_com: float = 0.0,
           cylin_a_com=None, cylin_v_com=None, dr: float = 1.0 / 127.0) -> dict:
    """The momentum source step (source.F90 with make_source_temp,
    source_bc, make_source_temp_bc) on one global block.  fields: rho, p,
    pot, s, t, a, temps, tempa as (numphi, numz, numr) arrays; returns the
    updated s, t, a, temps, tempa.  cylin_*_com: (3, numphi)."""
    numphi, numz, numr = fields["rho"].shape
    f = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in fields.items()}
    cac = np.zeros((numphi, 3)) if cylin_a_com is None else np.ascontiguousarray(np.asarray(cylin_a_com).T)
    cvc = np.zeros((numphi, 3)) if cylin_v_com is None else np.ascontiguousarray(np.asarray(cylin_v_com).T)
    bcv = (ctypes.c_int * 3)(*bc)
    lib().bho_source(numr, numz, numphi, dr, source_time, tstep, bcv, dt, omega, R_com, _ptr(cac), _ptr(cvc),
                     _ptr(f["s"]), _ptr(f["t"]), _ptr(f["a"]), _ptr(f["temps"]), _ptr(f["tempa"]),
                     _ptr(f["rho"]), _ptr(f["p"]), _ptr(f["pot"]))
    return {k: f[k] for k in ("s", "t", "a", "temps", "tempa")}


def make_source_temp(fields: dict, *, tstep: int, bc, dt: float, omega: float, R_com: float = 0.0,
                     cylin_a_com=None, cylin_v_com=None, dr: float = 1.0 / 127.0) -> dict:
    """make_source_temp alone (the start-up call, main.F90:315): temps, tempa."""
    numphi, numz, numr = fields["rho"].shape
    f = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in fields.items()}
    cac = np.zeros((numphi, 3)) if cylin_a_com is None else np.ascontiguousarray(np.asarray(cylin_a_com).T)
    cvc = np.zeros((numphi, 3)) if cylin_v_com is None else np.ascontiguousarray(np.asarray(cylin_v_com).T)
    bcv = (ctypes.c_int * 3)(*bc)
    lib().bho_make_source_temp(numr, numz, numphi, dr, tstep, bcv, dt, omega, R_com, _ptr(cac), _ptr(cvc),
                               _ptr(f["s"]), _ptr(f["a"]), _ptr(f["temps"]), _ptr(f["tempa"]),
                               _ptr(f["rho"]), _ptr(f["p"]), _ptr(f["pot"]))
    return {k: f[k] for k in ("temps", "tempa")}
