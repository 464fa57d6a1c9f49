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
BH_PREC_MIXED_TABLE32 = 2
BH_PREC_SPECTRAL64 = 3
# "mixed": FP32 transforms and sweeps, every Green's table in FP64, FP64
# residual and refinement; "mixed_t32": the same with the top/bottom table
# stored in FP32 (an opt-in variant); "spectral64": the mixed solve's
# spectral form of the iteration evaluated once in FP64 (isym = 1)
_PREC = {"fp64": BH_PREC_FP64, "mixed": BH_PREC_MIXED, "mixed_t32": BH_PREC_MIXED_TABLE32,
         "spectral64": BH_PREC_SPECTRAL64}

EXPORTS = ("bh_plan_create", "bh_plan_destroy", "bh_solve", "bh_solve_device",
           "bh_set_potential", "bh_plan_device_bytes", "bh_plan_set_timing",
           "bh_last_timings", "bh_sync_timing", "bh_last_error",
           "bh_batch_create", "bh_batch_destroy", "bh_batch_solve_device", "bh_batch_solve",
           "bh_batch_device_bytes", "bh_batch_get_stats", "bh_batch_set_timing", "bh_batch_tuning",
           "bh_batch_plan", "bh_server_create", "bh_server_submit", "bh_server_wait", "bh_server_drain",
           "bh_server_get_stats", "bh_server_destroy", "bh_set_potential_noalloc",
           "bh_cont_write", "bh_cont_read", "bh_cont_write_scalars", "bh_cont_read_scalars",
           "bh_set_potential_from_cont", "bh_source_device", "bh_make_source_temp_device")

BH_BATCH_GRAPH = 1
BH_BATCH_AUTOTUNE = 2


class BHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bh error {code}: {msg}")
        self.code = code


class GridDesc(ctypes.Structure):
    _fields_ = [("numr", ctypes.c_int), ("numz", ctypes.c_int), ("numphi", ctypes.c_int),
                ("isym", ctypes.c_int), ("dr", ctypes.c_double), ("mmax", ctypes.c_int),
                ("nsteps_cold", ctypes.c_int), ("nsteps_warm", ctypes.c_int), ("refine_passes", ctypes.c_int)]


class BatchStats(ctypes.Structure):
    _fields_ = [("ngrids", ctypes.c_int), ("launches", ctypes.c_int), ("graph", ctypes.c_int),
                ("tuning_left", ctypes.c_int), ("last_ms", ctypes.c_double)]


class BinaryModel(ctypes.Structure):
    _fields_ = [("rho_c", ctypes.c_double * 2), ("radius", ctypes.c_double * 2), ("separation", ctypes.c_double),
                ("drift", ctypes.c_double), ("n_core", ctypes.c_double), ("n_env", ctypes.c_double),
                ("core_frac", ctypes.c_double), ("noise", ctypes.c_double), ("seed", ctypes.c_ulonglong)]


class Request(ctypes.Structure):
    _fields_ = [("grid", GridDesc), ("nsolves", ctypes.c_int), ("model", BinaryModel),
                ("pot_out", ctypes.c_void_p), ("solve_ms", ctypes.c_void_p), ("start_pot", ctypes.c_void_p)]


class RequestResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("nsolves", ctypes.c_int),
                ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_active", ctypes.c_int), ("precision", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("solves", ctypes.c_longlong),
                ("iterations", ctypes.c_longlong), ("active", ctypes.c_int), ("waiting", ctypes.c_int),
                ("peak_active", ctypes.c_int), ("bytes_in_use", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t),
                ("elapsed_s", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double)]


class ContLayout(ctypes.Structure):
    _fields_ = [("numr", ctypes.c_int), ("numz", ctypes.c_int), ("numphi", ctypes.c_int),
                ("numr_procs", ctypes.c_int), ("numz_procs", ctypes.c_int)]


class ContScalars(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("frnum", ctypes.c_int), ("phi_com", ctypes.c_double),
                ("R_com", ctypes.c_double), ("com", ctypes.c_double * 9), ("v_com", ctypes.c_double * 3),
                ("a_com", ctypes.c_double * 3), ("cylin_a_com", ctypes.c_void_p), ("cylin_v_com", ctypes.c_void_p),
                ("delt", ctypes.c_double * 3)]


CONT_FIELDS = ("s", "t", "a", "rho", "tau", "pot", "tempa", "frac1", "frac2")


class HydroFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("s", "t", "a", "temps", "tempa", "rho", "p", "pot")]


class SourceParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("omega_frame", ctypes.c_double), ("tstep", ctypes.c_int),
                ("boundary_condition", ctypes.c_int * 3), ("R_com", ctypes.c_double),
                ("cylin_a_com", ctypes.c_void_p), ("cylin_v_com", ctypes.c_void_p)]

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
    L.bh_set_potential_noalloc.argtypes = [vp, vp]
    cp = ctypes.c_char_p
    L.bh_cont_write.argtypes = [cp, ctypes.c_int, ctypes.POINTER(ContLayout), ctypes.POINTER(vp)]
    L.bh_cont_read.argtypes = [cp, ctypes.POINTER(ContLayout), ctypes.POINTER(vp)]
    L.bh_cont_write_scalars.argtypes = [cp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ContScalars)]
    L.bh_cont_read_scalars.argtypes = [cp, ctypes.c_int, ctypes.POINTER(ContScalars)]
    L.bh_set_potential_from_cont.argtypes = [vp, cp, ctypes.c_int, ctypes.c_int]
    L.bh_source_device.argtypes = [vp, ctypes.c_int, ctypes.POINTER(HydroFields), ctypes.POINTER(SourceParams), vp]
    L.bh_make_source_temp_device.argtypes = [vp, ctypes.POINTER(HydroFields), ctypes.POINTER(SourceParams), vp]
    L.bh_plan_device_bytes.argtypes = [vp]
    L.bh_plan_device_bytes.restype = ctypes.c_size_t
    L.bh_plan_set_timing.argtypes = [vp, ctypes.c_int]
    L.bh_last_timings.argtypes = [vp, dp, ctypes.c_int]
    L.bh_sync_timing.argtypes = [vp]
    L.bh_last_error.restype = ctypes.c_char_p
    pp = ctypes.POINTER(vp)
    ip = ctypes.POINTER(ctypes.c_int)
    L.bh_batch_create.argtypes = [ctypes.POINTER(GridDesc), ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                  ctypes.c_int, pp]
    L.bh_batch_destroy.argtypes = [vp]
    L.bh_batch_solve_device.argtypes = [vp, pp, pp, ip, vp]
    L.bh_batch_solve.argtypes = [vp, pp, pp, ip]
    L.bh_batch_device_bytes.argtypes = [vp]
    L.bh_batch_device_bytes.restype = ctypes.c_size_t
    L.bh_batch_get_stats.argtypes = [vp, ctypes.POINTER(BatchStats)]
    L.bh_batch_set_timing.argtypes = [vp, ctypes.c_int]
    L.bh_batch_tuning.argtypes = [vp, ctypes.c_int, ip, ctypes.c_int, dp]
    L.bh_batch_plan.argtypes = [vp, ctypes.c_int, pp]
    L.bh_server_create.argtypes = [ctypes.POINTER(ServerConfig), pp]
    L.bh_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)]
    L.bh_server_wait.argtypes = [vp, ctypes.c_longlong, ctypes.POINTER(RequestResult)]
    L.bh_server_drain.argtypes = [vp]
    L.bh_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.bh_server_destroy.argtypes = [vp]
    _lib = L
    return L


def _check(rc: int):
    if rc != BH_OK:
        raise BHError(rc, load_library().bh_last_error().decode())


class PotentialPlan:
    """One geometry's solver state on the current GPU (potsetup + potp)."""

    def __init__(self, numr: int, numz: int, numphi: int, *, precision: str = "fp64",
                 hbm_budget: int = 0, dr: float = 1.0 / 127.0, mmax: int = 30,
                 nsteps_cold: int = 20, nsteps_warm: int = 5, isym: int = 1, refine_passes: int = 2):
        """precision "fp64", "mixed", "mixed_t32" or "spectral64";
        refine_passes (1..4) is the number of FP64 residual + refinement
        passes of a mixed solve."""
        L = load_library()
        self.shape = (numphi, numz, numr)
        self.desc = GridDesc(numr, numz, numphi, isym, dr, mmax, nsteps_cold, nsteps_warm, refine_passes)
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

    def potential_solver(self, rho: np.ndarray, tstep: int = 0, out: np.ndarray = None) -> np.ndarray:
        """Host arrays in, host potential out (synchronous).  `out`, when
        given, is a C-contiguous float64 array of the grid's shape that
        receives the potential (a caller-owned array, as the reference's
        COMMON block is); otherwise a new array is returned."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.shape != self.shape:
            raise ValueError(f"rho shape {rho.shape} != {self.shape}")
        if out is None:
            pot = np.empty(self.shape, dtype=np.float64)
        else:
            if out.shape != self.shape or out.dtype != np.float64 or not out.flags["C_CONTIGUOUS"]:
                raise ValueError("out must be a C-contiguous float64 array of the grid's shape")
            pot = out
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

    def set_potential_noalloc(self, pot):
        """As set_potential from a host array, staged through the library's
        fixed device array (no device allocation)."""
        pot = np.ascontiguousarray(pot, dtype=np.float64)
        if pot.shape != self.shape:
            raise ValueError(f"pot shape {pot.shape} != {self.shape}")
        _check(load_library().bh_set_potential_noalloc(self._h, ctypes.c_void_p(pot.ctypes.data)))

    def set_potential_from_cont(self, prefix: str, numr_procs: int, numz_procs: int):
        """Warm-start potential from a run's fort.12 continuation files
        (prefix: the path up to the rank number)."""
        _check(load_library().bh_set_potential_from_cont(self._h, prefix.encode(), numr_procs, numz_procs))

    @staticmethod
    def _source_args(fields: dict, params: dict):
        f = HydroFields(**{k: int(v) for k, v in fields.items()})
        keep = []
        prm = SourceParams()
        prm.dt, prm.omega_frame, prm.tstep = params["dt"], params["omega"], int(params["tstep"])
        prm.boundary_condition[:] = list(params["bc"])
        prm.R_com = params.get("R_com", 0.0)
        for k in ("cylin_a_com", "cylin_v_com"):
            if params.get(k) is not None:
                arr = np.ascontiguousarray(np.asarray(params[k], dtype=np.float64).T)  # (numphi, 3): Fortran (3, numphi)
                keep.append(arr)
                setattr(prm, k, arr.ctypes.data)
        return f, prm, keep

    def source_device(self, source_time: int, fields: dict, params: dict, stream: int = 0):
        """The hydro's momentum source step on device fields (device
        pointers: s, t, a, temps, tempa, rho, p, pot); params: dt, omega,
        tstep, bc (3), and R_com, cylin_a_com, cylin_v_com (3, numphi) when
        tstep > 1000."""
        f, prm, keep = self._source_args(fields, params)
        _check(load_library().bh_source_device(self._h, int(source_time), ctypes.byref(f), ctypes.byref(prm),
                                               ctypes.c_void_p(stream)))
        del keep

    def make_source_temp_device(self, fields: dict, params: dict, stream: int = 0):
        f, prm, keep = self._source_args(fields, params)
        _check(load_library().bh_make_source_temp_device(self._h, ctypes.byref(f), ctypes.byref(prm),
                                                         ctypes.c_void_p(stream)))
        del keep

    def set_timing(self, on: bool = True):
        _check(load_library().bh_plan_set_timing(self._h, 1 if on else 0))

    def last_timings(self) -> dict:
        """Mean stage times (ms, HIP events on the solve stream) over the
        solves timed since timing was switched on or last read.  The
        forward DFT of rho runs on a side stream beside the projections:
        "projections" is the main stream's projection kernel, and
        "boundary_and_forward_dft" the boundary contraction up to the join
        with the side stream (its DFT and side contraction overlapped)."""
        L = load_library()
        _check(L.bh_sync_timing(self._h))
        ms = (ctypes.c_double * 6)()
        _check(L.bh_last_timings(self._h, ms, 6))
        return dict(zip(("boundary_and_forward_dft", "projections", "adi", "inverse_dft", "total",
                         "contract_topbot"), list(ms)))


class PotentialBatch:
    """Independent grids of different sizes solved together (one plan per
    grid; step k of every grid's solve runs as one grouped launch per kernel
    kind).  A grid is (numr, numz, numphi) or (numr, numz, numphi, isym).  ``graph`` replays the launch sequence from a captured hipGraph;
    ``autotune`` times tile and pencil candidates over the first five solves
    and keeps the fastest per step."""

    def __init__(self, grids, *, precision: str = "mixed", hbm_budget: int = 0, graph: bool = False,
                 autotune: bool = False, dr: float = 1.0 / 127.0, mmax: int = 30,
                 nsteps_cold: int = 20, nsteps_warm: int = 5, refine_passes: int = 2):
        L = load_library()
        # (numr, numz, numphi) or (numr, numz, numphi, isym)
        grids = [tuple(g) + (1,) * (4 - len(g)) for g in grids]
        self.shapes = [(p, z, r) for (r, z, p, _) in grids]
        n = len(grids)
        descs = (GridDesc * n)(*[GridDesc(r, z, p, sym, dr, mmax, nsteps_cold, nsteps_warm, refine_passes)
                                 for (r, z, p, sym) in grids])
        flags = (BH_BATCH_GRAPH if graph else 0) | (BH_BATCH_AUTOTUNE if autotune else 0)
        h = ctypes.c_void_p()
        _check(L.bh_batch_create(descs, n, _PREC[precision], hbm_budget, flags, ctypes.byref(h)))
        self._h = h
        self.n = n

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load_library().bh_batch_destroy(self._h)
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

    def device_bytes(self) -> int:
        return load_library().bh_batch_device_bytes(self._h)

    @staticmethod
    def _ptrs(ptrs):
        return (ctypes.c_void_p * len(ptrs))(*[ctypes.c_void_p(int(x)) for x in ptrs])

    def _tsteps(self, tstep):
        if isinstance(tstep, int):
            tstep = [tstep] * self.n
        return (ctypes.c_int * self.n)(*[int(t) for t in tstep])

    def solve(self, rhos, tstep=0):
        """Host arrays in, host potentials out (synchronous)."""
        rhos = [np.ascontiguousarray(r, dtype=np.float64) for r in rhos]
        for r, shp in zip(rhos, self.shapes):
            if r.shape != shp:
                raise ValueError(f"rho shape {r.shape} != {shp}")
        pots = [np.empty(s, dtype=np.float64) for s in self.shapes]
        _check(load_library().bh_batch_solve(self._h, self._ptrs([r.ctypes.data for r in rhos]),
                                             self._ptrs([p.ctypes.data for p in pots]), self._tsteps(tstep)))
        return pots

    def solve_device(self, rho_ptrs, pot_ptrs, tstep=0, stream: int = 0):
        _check(load_library().bh_batch_solve_device(self._h, self._ptrs(rho_ptrs), self._ptrs(pot_ptrs),
                                                    self._tsteps(tstep), ctypes.c_void_p(stream)))

    def set_timing(self, on: bool = True):
        _check(load_library().bh_batch_set_timing(self._h, 1 if on else 0))

    def stats(self) -> dict:
        st = BatchStats()
        _check(load_library().bh_batch_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in BatchStats._fields_}

    TUNE_AXES = ("tile", "nl", "rows", "grid", "order")

    def tuning(self) -> dict:
        """Per tuning axis: the choice at each step of the solve (-1: not
        applicable) and the ms the axis's choices saved over its default."""
        L = load_library()
        out = {}
        for ax, name in enumerate(self.TUNE_AXES):
            m = L.bh_batch_tuning(self._h, ax, None, 0, None)
            if m < 0:
                _check(m)
            ch = (ctypes.c_int * max(m, 1))()
            gain = ctypes.c_double(0.0)
            L.bh_batch_tuning(self._h, ax, ch, m, ctypes.byref(gain))
            out[name] = {"choice": list(ch)[:m], "gain_ms": gain.value}
        return out


class PotentialServer:
    """Continuous-batching server of SCF solve sequences (bh_server_*)."""

    def __init__(self, *, hbm_budget: int = 0, max_active: int = 64, precision: str = "mixed"):
        L = load_library()
        cfg = ServerConfig(hbm_budget, max_active, _PREC[precision])
        h = ctypes.c_void_p()
        _check(L.bh_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load_library().bh_server_destroy(self._h)
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

    def submit(self, grid, nsolves: int, model: dict, *, want_potential: bool = False,
               want_latencies: bool = True, dr: float = 1.0 / 127.0, refine_passes: int = 2,
               start_pot=None) -> int:
        """grid = (numr, numz, numphi) or (numr, numz, numphi, isym); model keys as bh_binary_model;
        start_pot: a (numphi, numz, numr) potential to resume from (e.g. from
        cont_read), the first solve then warm."""
        numr, numz, numphi, isym = tuple(grid) + (1,) * (4 - len(grid))
        m = BinaryModel()
        m.rho_c[:] = list(model.get("rho_c", (0.0, 0.0)))
        m.radius[:] = list(model["radius"])
        m.separation = model["separation"]
        m.drift = model.get("drift", 0.0)
        m.n_core, m.n_env, m.core_frac = model.get("n_core", 0.0), model.get("n_env", 0.0), model.get("core_frac", 0.0)
        m.noise = model.get("noise", 0.0)
        m.seed = int(model.get("seed", 0))
        pot = np.empty((numphi, numz, numr)) if want_potential else None
        lat = np.zeros(nsolves) if want_latencies else None
        sp = None
        if start_pot is not None:
            sp = np.ascontiguousarray(start_pot, dtype=np.float64)
            if sp.shape != (numphi, numz, numr):
                raise ValueError(f"start_pot shape {sp.shape} != {(numphi, numz, numr)}")
        r = Request(GridDesc(numr, numz, numphi, isym, dr, 30, 20, 5, refine_passes), nsolves, m,
                    pot.ctypes.data if pot is not None else None, lat.ctypes.data if lat is not None else None,
                    sp.ctypes.data if sp is not None else None)
        rid = ctypes.c_longlong()
        _check(load_library().bh_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        self._keep[rid.value] = (pot, lat)
        return rid.value

    def wait(self, rid: int):
        """Block until request rid finishes: (result dict, last potential or
        None, per-solve latencies in ms or None)."""
        res = RequestResult()
        rc = load_library().bh_server_wait(self._h, rid, ctypes.byref(res))
        pot, lat = self._keep.pop(rid, (None, None))
        _check(rc)
        return {k: getattr(res, k) for k, _ in RequestResult._fields_}, pot, lat

    def drain(self):
        _check(load_library().bh_server_drain(self._h))

    def iterations(self) -> np.ndarray:
        """Host-side log of the serving loop's iterations, one row each:
        t_admit, t_enq0, t_enq1, t_done (seconds since the server started;
        the admission pass before the iteration, its enqueue, its completion),
        solves in it, plans created for it, grid points solved, and its lane
        (grid points below 2^20, below 2^22, larger: lanes run side by side)."""
        L = load_library()
        f = L.bhx_server_iterations
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        n = f(self._h, None, 0)
        out = np.zeros((max(n, 1), 8))
        n = f(self._h, out.ctypes.data, n)
        return out[:n]

    def stats(self) -> dict:
        st = ServerStats()
        _check(load_library().bh_server_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}


def bipolytrope_profile(n_core: float = 3.0, n_env: float = 1.5, core_frac: float = 0.3, ntab: int = 2049):
    """The served workload's normalised bipolytrope profile (host code of the
    library): (table on u = 0..1, surface radius, interface radius) in the
    core's Lane-Emden units."""
    L = load_library()
    f = L.bhx_bipolytrope_profile
    f.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    t = np.zeros(ntab)
    info = np.zeros(2)
    _check(f(n_core, n_env, core_frac, t.ctypes.data, ntab, info.ctypes.data))
    return t, float(info[0]), float(info[1])


# ---- continuation files (fort.12.<tstep>.<rank>, fort.13.<tstep>) ---------

def cont_write(directory: str, tstep: int, grid, procs, fields: dict):
    """Write fort.12.<tstep>.<rank> files in the reference's format; fields
    maps names of CONT_FIELDS to (numphi, numz, numr) arrays (missing: zeros)."""
    numr, numz, numphi = grid
    lay = ContLayout(numr, numz, numphi, procs[0], procs[1])
    keep = [np.ascontiguousarray(fields[n], dtype=np.float64) if n in fields else None for n in CONT_FIELDS]
    ptrs = (ctypes.c_void_p * len(CONT_FIELDS))(*[a.ctypes.data if a is not None else None for a in keep])
    _check(load_library().bh_cont_write(directory.encode(), int(tstep), ctypes.byref(lay), ptrs))


def cont_read(prefix: str, grid, procs, names=CONT_FIELDS) -> dict:
    """Read the named fields from fort.12 files (prefix up to the rank)."""
    numr, numz, numphi = grid
    lay = ContLayout(numr, numz, numphi, procs[0], procs[1])
    out = {n: np.zeros((numphi, numz, numr)) for n in names}
    ptrs = (ctypes.c_void_p * len(CONT_FIELDS))(*[out[n].ctypes.data if n in out else None for n in CONT_FIELDS])
    _check(load_library().bh_cont_read(prefix.encode(), ctypes.byref(lay), ptrs))
    return out


def cont_read_scalars(path: str, numphi: int) -> dict:
    ca, cv = np.zeros((numphi, 3)), np.zeros((numphi, 3))
    sc = ContScalars()
    sc.cylin_a_com, sc.cylin_v_com = ca.ctypes.data, cv.ctypes.data
    _check(load_library().bh_cont_read_scalars(path.encode(), numphi, ctypes.byref(sc)))
    return {"time": sc.time, "frnum": sc.frnum, "phi_com": sc.phi_com, "R_com": sc.R_com,
            "com": np.array(sc.com[:]).reshape(3, 3, order="F"), "v_com": np.array(sc.v_com[:]),
            "a_com": np.array(sc.a_com[:]), "cylin_a_com": ca.T.copy(), "cylin_v_com": cv.T.copy(),
            "delt": np.array(sc.delt[:])}


def cont_write_scalars(directory: str, tstep: int, numphi: int, sc: dict):
    ca = np.ascontiguousarray(np.asarray(sc["cylin_a_com"], dtype=np.float64).T)
    cv = np.ascontiguousarray(np.asarray(sc["cylin_v_com"], dtype=np.float64).T)
    c = ContScalars()
    c.time, c.frnum, c.phi_com, c.R_com = sc["time"], int(sc["frnum"]), sc["phi_com"], sc["R_com"]
    c.com[:] = list(np.asarray(sc["com"], dtype=np.float64).reshape(-1, order="F"))
    c.v_com[:], c.a_com[:], c.delt[:] = list(sc["v_com"]), list(sc["a_com"]), list(sc["delt"])
    c.cylin_a_com, c.cylin_v_com = ca.ctypes.data, cv.ctypes.data
    _check(load_library().bh_cont_write_scalars(directory.encode(), int(tstep), numphi, ctypes.byref(c)))
