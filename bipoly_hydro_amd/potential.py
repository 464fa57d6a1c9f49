"""Host-side mirror of bipoly_hydro's self-gravity interface over the C ABI.

The reference drives the solve through two argument-less Fortran routines on
COMMON blocks: ``potsetup`` once after grid setup (setup_frac.F90:689) and
``potential_solver`` every step (main.F90:395-397).  Here they are
``PotentialPlan(...)`` (potsetup) and ``PotentialPlan.potential_solver(rho,
tstep)``, with the same arguments made explicit: the density ``rho`` and the
step counter ``tstep`` that selects a cold (tstep <= 1, 20 ADI steps from
zero) or warm (tstep > 1, 5 steps from the previous potential) solve
(potential_solver.F90:80-94).

Arrays are the reference's (numr, numz, numphi) Fortran-order arrays, held
in numpy as C-contiguous (numphi, numz, numr) float64.  Everything runs in
libbh_potential.so on the GPU; there is no CPU path, and a missing library
raises at plan creation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbh_potential.so")

BH_OK = 0
BH_PREC_FP64 = 0
BH_PREC_MIXED = 1
_PREC = {"fp64": BH_PREC_FP64, "mixed": BH_PREC_MIXED}

EXPORTS = ("bh_plan_create", "bh_plan_destroy", "bh_solve", "bh_solve_device",
           "bh_set_potential", "bh_plan_device_bytes", "bh_plan_set_timing",
           "bh_last_timings", "bh_sync_timing", "bh_last_error")


class BHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bh error {code}: {msg}")
        self.code = code


class GridDesc(ctypes.Structure):
    _fields_ = [("numr", ctypes.c_int), ("numz", ctypes.c_int), ("numphi", ctypes.c_int),
                ("isym", ctypes.c_int), ("dr", ctypes.c_double), ("mmax", ctypes.c_int),
                ("nsteps_cold", ctypes.c_int), ("nsteps_warm", ctypes.c_int)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libbh_potential.so (raises OSError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is not built; run python -m bipoly_hydro_amd.build")
    # torch's wheel ships its own libamdhip64.so.7 under the same soname as
    # ROCm's; whichever is loaded first serves the whole process.  Load
    # torch's first so that torch and this library share one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    L.bh_plan_create.argtypes = [ctypes.POINTER(GridDesc), ctypes.c_int, ctypes.c_size_t,
                                 ctypes.POINTER(vp)]
    L.bh_plan_create.restype = ctypes.c_int
    L.bh_plan_destroy.argtypes = [vp]
    L.bh_solve.argtypes = [vp, dp, dp, ctypes.c_int]
    L.bh_solve_device.argtypes = [vp, vp, vp, ctypes.c_int, vp]
    L.bh_set_potential.argtypes = [vp, vp, ctypes.c_int]
    L.bh_plan_device_bytes.argtypes = [vp]
    L.bh_plan_device_bytes.restype = ctypes.c_size_t
    L.bh_plan_set_timing.argtypes = [vp, ctypes.c_int]
    L.bh_last_timings.argtypes = [vp, dp, ctypes.c_int]
    L.bh_sync_timing.argtypes = [vp]
    L.bh_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc: int):
    if rc != BH_OK:
        raise BHError(rc, load_library().bh_last_error().decode())


class PotentialPlan:
    """One geometry's solver state on the current GPU (potsetup + potp)."""

    def __init__(self, numr: int, numz: int, numphi: int, *, precision: str = "fp64",
                 hbm_budget: int = 0, dr: float = 1.0 / 127.0, mmax: int = 30,
                 nsteps_cold: int = 20, nsteps_warm: int = 5, isym: int = 1):
        L = load_library()
        self.shape = (numphi, numz, numr)
        self.desc = GridDesc(numr, numz, numphi, isym, dr, mmax, nsteps_cold, nsteps_warm)
        h = ctypes.c_void_p()
        _check(L.bh_plan_create(ctypes.byref(self.desc), _PREC[precision], hbm_budget, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load_library().bh_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def device_bytes(self) -> int:
        return load_library().bh_plan_device_bytes(self._h)

    def potential_solver(self, rho: np.ndarray, tstep: int = 0) -> np.ndarray:
        """Host arrays in, host potential out (synchronous)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.shape != self.shape:
            raise ValueError(f"rho shape {rho.shape} != {self.shape}")
        pot = np.empty(self.shape, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        _check(load_library().bh_solve(self._h, rho.ctypes.data_as(dp), pot.ctypes.data_as(dp), int(tstep)))
        return pot

    def potential_solver_device(self, rho_ptr: int, pot_ptr: int, tstep: int = 0, stream: int = 0):
        """Device pointers (e.g. torch tensor .data_ptr()), enqueued on stream."""
        _check(load_library().bh_solve_device(self._h, ctypes.c_void_p(rho_ptr), ctypes.c_void_p(pot_ptr),
                                              int(tstep), ctypes.c_void_p(stream)))

    def set_potential(self, pot, on_device: bool = False):
        if on_device:
            _check(load_library().bh_set_potential(self._h, ctypes.c_void_p(int(pot)), 1))
        else:
            pot = np.ascontiguousarray(pot, dtype=np.float64)
            _check(load_library().bh_set_potential(self._h, ctypes.c_void_p(pot.ctypes.data), 0))

    def set_timing(self, on: bool = True):
        _check(load_library().bh_plan_set_timing(self._h, 1 if on else 0))

    def last_timings(self) -> dict:
        """Mean stage times (ms, HIP events on the solve stream) over the
        solves timed since timing was switched on or last read."""
        L = load_library()
        _check(L.bh_sync_timing(self._h))
        ms = (ctypes.c_double * 6)()
        _check(L.bh_last_timings(self._h, ms, 6))
        return dict(zip(("boundary", "forward_dft", "adi", "inverse_dft", "total", "contract_topbot"), list(ms)))
