"""ctypes mirror of the C ABI in include/dwf_hip.h.

Host arrays follow the reference layouts (interleaved complex128):
DWF spinors ``[nrhs][Ls][Nt][Nx][2]``, Wilson spinors ``[nrhs][Nt][Nx][2]`` and gauge
angles ``[2][Nt][Nx]`` (domain_wall/include/indexing.hpp:5-13,
wilson/include/indexing.hpp:5-12).  Device buffers are passed as raw pointers
(``int``), e.g. ``tensor.data_ptr()`` of a torch tensor on ``cuda:0``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libdwfhip.so"

DWF_OK = 0
DWF_ERR_ARG = 1
DWF_ERR_HIP = 2
DWF_ERR_NOT_CONVERGED = 3
DWF_ERR_ZERO_DENOM = 4
DWF_ERR_CERTIFY = 5

DWF_OP_D = 0
DWF_OP_DDAG = 1
DWF_OP_DDDAG = 2
DWF_OP_DDAGD = 3

DWF_PREC_FP64 = 0
DWF_PREC_FP32 = 1

DWF_SOLVE_CG64 = 0
DWF_SOLVE_MIXED = 1
DWF_SOLVE_EO_MIXED = 2
DWF_SOLVE_LEARNED = 3

DWF_TUNE_SCHUNKS = 0
DWF_TUNE_XCD_REMAP = 1
DWF_TUNE_RELIABLE_DELTA = 2
DWF_TUNE_PREFETCH = 3
DWF_TUNE_TILE_X = 4
DWF_TUNE_NT_STORE = 5
DWF_TUNE_GRAPHS = 6
DWF_TUNE_LDS = 7
DWF_TUNE_NET_FUSED = 8
DWF_TUNE_FCG_BETA = 9
DWF_TUNE_EO_FUSED = 10
DWF_TUNE_NET_TILE = 11
DWF_TUNE_STREAMS = 12
DWF_TUNE_NET_SPLIT = 13


class DwfError(RuntimeError):
    """A non-zero status from libdwfhip; ``code`` is the C ABI status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("reliable_updates", ctypes.c_int),
        ("fp64_iterations", ctypes.c_int),
        ("passes", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("relres", ctypes.c_double),
        ("true_relres", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


DWF_HMC_DOMAIN_WALL = 0
DWF_HMC_WILSON = 1


class HmcParams(ctypes.Structure):
    _fields_ = [
        ("beta", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("steps", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("method", ctypes.c_int),
        ("maxiter", ctypes.c_int),
        ("rtol", ctypes.c_double),
    ]


class HmcResult(ctypes.Structure):
    _fields_ = [
        ("accepted", ctypes.c_int),
        ("solves", ctypes.c_int),
        ("cg_iterations", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("acceptance_probability", ctypes.c_double),
        ("delta_h", ctypes.c_double),
        ("h_initial", ctypes.c_double),
        ("h_final", ctypes.c_double),
        ("plaquette", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("tile_x", ctypes.c_int),
        ("prefetch", ctypes.c_int),
        ("chunks", ctypes.c_int),
        ("nt_store", ctypes.c_int),
        ("from_cache", ctypes.c_int),
        ("microseconds", ctypes.c_double),
        ("lds", ctypes.c_int),
        ("pad", ctypes.c_int),
    ]


class BatchTuneResult(ctypes.Structure):
    _fields_ = [("nrhs", ctypes.c_int), ("from_cache", ctypes.c_int), ("solves_per_s", ctypes.c_double)]


class LatticeGroup(ctypes.Structure):
    _fields_ = [("Nt", ctypes.c_int), ("Nx", ctypes.c_int), ("ncfg", ctypes.c_int), ("pad", ctypes.c_int)]


class MeasureSpec(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int),
        ("maxiter", ctypes.c_int),
        ("rtol", ctypes.c_double),
        ("wall", ctypes.c_int),
        ("point", ctypes.c_int),
        ("max_momentum", ctypes.c_int),
        ("noise_vectors", ctypes.c_int),
        ("noise_seed", ctypes.c_ulonglong),
        ("point_site", ctypes.c_void_p),
    ]


class MeasureOut(ctypes.Structure):
    _fields_ = [
        ("CPP", ctypes.c_void_p),
        ("CJ5", ctypes.c_void_p),
        ("momentum_CPP", ctypes.c_void_p),
        ("loop", ctypes.c_void_p),
        ("iterations", ctypes.c_void_p),
        ("dirac_residual", ctypes.c_void_p),
        ("wall_nonconverged", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("point_CJ5", ctypes.c_void_p),
    ]


def library_path() -> str:
    return os.path.join(_HERE, _LIB_NAME)


def build_library(jobs: int = 8) -> str:
    """Compile libdwfhip.so for gfx950 in-tree (hipcc via csrc/Makefile)."""
    subprocess.run(["make", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"], check=True)
    return library_path()


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); kept in the order of include/dwf_hip.h
PROTOTYPES = {
    "dwf_ctx_create_dwf": (_I, [_PP, _I, _I, _I, _D, _D]),
    "dwf_ctx_create_wilson": (_I, [_PP, _I, _I, _D, _D, _D]),
    "dwf_ctx_destroy": (_I, [_P]),
    "dwf_ctx_stream": (_P, [_P]),
    "dwf_ctx_tune": (_I, [_P, _I, _D]),
    "dwf_autotune": (_I, [_P, _I, _I, ctypes.c_char_p, _P]),
    "dwf_autotune_batch": (_I, [_P, _I, _D, _I, ctypes.c_char_p, _P]),
    "dwf_set_gauge": (_I, [_P, _P]),
    "dwf_set_gauge_device": (_I, [_P, _P]),
    "dwf_set_gauge_batch": (_I, [_P, _I, _P]),
    "dwf_set_gauge_groups": (_I, [_P, _I, _P, _P]),
    "dwf_apply": (_I, [_P, _I, _P, _P, _I]),
    "dwf_apply_device": (_I, [_P, _I, _I, _P, _P, _I]),
    "dwf_apply_device_noalloc": (_I, [_P, _I, _I, _P, _P, _I]),
    "dwf_solve": (_I, [_P, _I, _I, _P, _P, _D, _I, _I, _P, _P]),
    "dwf_solve_device": (_I, [_P, _I, _I, _P, _P, _D, _I, _I, _P, _P]),
    "dwf_solve_dirac": (_I, [_P, _I, _P, _P, _D, _I, _I, _P, _P]),
    "dwf_solve_dirac_device": (_I, [_P, _I, _P, _P, _D, _I, _I, _P, _P]),
    "dwf_pf_action_force": (_I, [_P, _I, _P, _P, _D, _I, _P, _P, _P, _P]),
    "dwf_force_from_eta": (_I, [_P, _P, _P]),
    "dwf_gauge_action_force": (_I, [_P, _D, _P, _P]),
    "dwf_average_plaquette": (_I, [_P, _P]),
    "dwf_get_gauge": (_I, [_P, _P]),
    "dwf_hmc_trajectory": (_I, [_P, _P, _P, _P, _D, _P]),
    "dwf_measure": (_I, [_P, _P, _P]),
    "dwf_measure_batch": (_I, [_P, _P, _I, _P, _P]),
    "dwf_net_num_weights": (ctypes.c_longlong, [_I]),
    "dwf_net_create": (_I, [_P, _I, ctypes.c_ulonglong]),
    "dwf_net_destroy": (_I, [_P]),
    "dwf_net_set_weights": (_I, [_P, _P]),
    "dwf_net_get_weights": (_I, [_P, _P]),
    "dwf_net_apply_device": (_I, [_P, _P, _I, _P, _P, _I]),
    "dwf_net_loss_grad_device": (_I, [_P, _P, _I, _P, _P, _P, _I, _P, _P]),
    "dwf_net_train_device": (_I, [_P, _P, _P, _P, _P, _I, _P, _P]),
    "dwf_ctx_set_preconditioner": (_I, [_P, _P]),
    "dwf_device_alloc": (_P, [ctypes.c_ulonglong]),
    "dwf_device_free": (_I, [_P]),
    "dwf_memcpy_htod": (_I, [_P, _P, ctypes.c_ulonglong]),
    "dwf_memcpy_dtoh": (_I, [_P, _P, ctypes.c_ulonglong]),
    "dwf_ctx_synchronize": (_I, [_P]),
    "dwf_timer_create": (_P, []),
    "dwf_timer_destroy": (None, [_P]),
    "dwf_timer_start": (_I, [_P, _P]),
    "dwf_timer_stop": (_I, [_P, _P]),
    "dwf_timer_elapsed_ms": (_D, [_P]),
    "dwf_copy_device": (_I, [_P, _P, _P, ctypes.c_ulonglong]),
    "dwf_convert_f64_f32_device": (_I, [_P, _P, _P, ctypes.c_ulonglong]),
    "dwf_rng_create": (_P, [ctypes.c_ulonglong]),
    "dwf_rng_destroy": (None, [_P]),
    "dwf_rng_dwf_heatbath": (None, [_P, _I, _I, _P, _P]),
    "dwf_rng_wilson_heatbath": (None, [_P, _I, _I, _P, _P]),
    "dwf_rng_uniform": (_D, [_P]),
    "dwf_z4_noise": (None, [ctypes.c_ulonglong, _I, _I, _P]),
    "dwf_last_error": (ctypes.c_char_p, []),
}


def load_library(path: str | None = None):
    """Load libdwfhip.so.  Raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path()
    if not os.path.exists(p):
        raise DwfError(DWF_ERR_HIP, f"{p} is missing: build it first (__graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(lib, rc: int, what: str):
    if rc != DWF_OK:
        msg = lib.dwf_last_error().decode()
        raise DwfError(rc, f"{what}: {msg} (status {rc})")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class Geometry:
    Nt: int
    Nx: int
    Ls: int  # 1 for Wilson

    @property
    def V4(self) -> int:
        return self.Nt * self.Nx

    @property
    def ndof(self) -> int:
        """Complex components of one spinor."""
        return self.Ls * self.Nt * self.Nx * 2


class DwfContext:
    """One device context (HIP stream, gauge field, workspace) of libdwfhip."""

    def __init__(self, handle, geom: Geometry, kind: str, lib):
        self._h = handle
        self.geom = geom
        self._geom0 = geom
        self._groups = None  # [(Nt, Nx, ncfg)] while several lattice sizes are held
        self.kind = kind
        self._lib = lib

    # construction -----------------------------------------------------------------
    @classmethod
    def dwf(cls, Nt: int, Nx: int, Ls: int, M5: float, mf: float) -> "DwfContext":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib, lib.dwf_ctx_create_dwf(ctypes.byref(h), Nt, Nx, Ls, M5, mf), "dwf_ctx_create_dwf")
        return cls(h, Geometry(Nt, Nx, Ls), "dwf", lib)

    @classmethod
    def wilson(cls, Nt: int, Nx: int, m0: float, bc_t: float, bc_x: float) -> "DwfContext":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib, lib.dwf_ctx_create_wilson(ctypes.byref(h), Nt, Nx, m0, bc_t, bc_x),
               "dwf_ctx_create_wilson")
        return cls(h, Geometry(Nt, Nx, 1), "wilson", lib)

    def close(self):
        if self._h:
            self._lib.dwf_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return self._lib.dwf_ctx_stream(self._h) or 0

    def tune(self, param: int, value: float):
        _check(self._lib, self._lib.dwf_ctx_tune(self._h, param, value), "dwf_ctx_tune")

    def autotune(self, nrhs: int, precision: int = DWF_PREC_FP32, cache_path: str | None = None) -> dict:
        r = TuneResult()
        _check(self._lib, self._lib.dwf_autotune(self._h, nrhs, precision,
                                                 cache_path.encode() if cache_path else None, ctypes.byref(r)),
               "dwf_autotune")
        return {name: getattr(r, name) for name, _ in r._fields_}

    def autotune_batch(self, method: int, rtol: float, max_nrhs: int, cache_path: str | None = None) -> dict:
        """dwf_autotune_batch: the solve batch size with the most solves/s."""
        r = BatchTuneResult()
        _check(self._lib, self._lib.dwf_autotune_batch(self._h, method, rtol, max_nrhs,
                                                       cache_path.encode() if cache_path else None, ctypes.byref(r)),
               "dwf_autotune_batch")
        return {name: getattr(r, name) for name, _ in r._fields_}

    def synchronize(self):
        _check(self._lib, self._lib.dwf_ctx_synchronize(self._h), "dwf_ctx_synchronize")

    def copy_device(self, dst_ptr: int, src_ptr: int, nbytes: int):
        _check(self._lib, self._lib.dwf_copy_device(self._h, dst_ptr, src_ptr, nbytes), "dwf_copy_device")

    def convert_f64_f32(self, src_ptr: int, dst_ptr: int, n: int):
        _check(self._lib, self._lib.dwf_convert_f64_f32_device(self._h, src_ptr, dst_ptr, n),
               "dwf_convert_f64_f32_device")

    # gauge --------------------------------------------------------------------------
    def _single_lattice(self):
        self.geom = self._geom0
        self._groups = None

    def set_gauge_groups(self, groups):
        """Hold configurations of several lattice sizes (dwf_set_gauge_groups):
        groups = [(Nt, Nx, thetas[ncfg][2*Nt*Nx]), ...].  Vectors of every later call are
        the concatenation, configuration-major in group order, of each right-hand side's
        vector on its own lattice."""
        defs = (LatticeGroup * len(groups))()
        parts, lat = [], []
        for i, (Nt, Nx, th) in enumerate(groups):
            th = np.ascontiguousarray(th, dtype=np.float64).reshape(-1, 2 * Nt * Nx)
            defs[i] = LatticeGroup(Nt, Nx, th.shape[0], 0)
            parts.append(th.reshape(-1))
            lat.append((Nt, Nx, th.shape[0]))
        allth = np.ascontiguousarray(np.concatenate(parts))
        _check(self._lib, self._lib.dwf_set_gauge_groups(self._h, len(groups), ctypes.cast(defs, ctypes.c_void_p),
                                                         _ptr(allth)), "dwf_set_gauge_groups")
        big = max(lat, key=lambda g: g[0] * g[1])
        self.geom = Geometry(big[0], big[1], self._geom0.Ls)
        self._groups = lat if len(lat) > 1 else None

    def config_lattices(self) -> list:
        """(Nt, Nx) of every configuration held, in order."""
        if self._groups is None:
            return None
        return [(Nt, Nx) for Nt, Nx, n in self._groups for _ in range(n)]

    def set_gauge(self, theta: np.ndarray):
        self._single_lattice()
        th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        if th.size != 2 * self.geom.V4:
            raise ValueError("theta must hold 2*Nt*Nx angles")
        _check(self._lib, self._lib.dwf_set_gauge(self._h, _ptr(th)), "dwf_set_gauge")

    def set_preconditioner(self, net):
        """Attach the network DWF_SOLVE_LEARNED uses (None detaches); keeps a reference."""
        self._net = net
        _check(self._lib, self._lib.dwf_ctx_set_preconditioner(self._h, net.handle if net else None),
               "dwf_ctx_set_preconditioner")

    def set_gauge_batch(self, thetas: np.ndarray):
        self._single_lattice()
        th = np.ascontiguousarray(thetas, dtype=np.float64).reshape(-1)
        if th.size % (2 * self.geom.V4):
            raise ValueError("thetas must hold ncfg x 2*Nt*Nx angles")
        n = th.size // (2 * self.geom.V4)
        _check(self._lib, self._lib.dwf_set_gauge_batch(self._h, n, _ptr(th)), "dwf_set_gauge_batch")

    def set_gauge_device(self, theta_ptr: int):
        self._single_lattice()
        _check(self._lib, self._lib.dwf_set_gauge_device(self._h, theta_ptr), "dwf_set_gauge_device")

    # operators ----------------------------------------------------------------------
    def _as_batch(self, v: np.ndarray) -> tuple[np.ndarray, int]:
        a = np.ascontiguousarray(v, dtype=np.complex128)
        if self._groups is not None:
            # k right-hand sides per configuration over every configuration held
            unit = sum(n * self.geom.Ls * Nt * Nx * 2 for Nt, Nx, n in self._groups)
            ncfg = sum(n for _, _, n in self._groups)
            if a.size % unit:
                raise ValueError(f"vector size {a.size} is not a multiple of {unit}")
            return a.reshape(-1), a.size // unit * ncfg
        n = self.geom.ndof
        if a.size % n:
            raise ValueError(f"vector size {a.size} is not a multiple of {n}")
        return a.reshape(-1), a.size // n

    def apply(self, op: int, psi: np.ndarray) -> np.ndarray:
        a, nrhs = self._as_batch(psi)
        out = np.empty_like(a)
        _check(self._lib, self._lib.dwf_apply(self._h, op, _ptr(a), _ptr(out), nrhs), "dwf_apply")
        return out.reshape(np.shape(psi))

    def apply_device(self, op: int, precision: int, in_ptr: int, out_ptr: int, nrhs: int):
        _check(self._lib, self._lib.dwf_apply_device(self._h, op, precision, in_ptr, out_ptr, nrhs),
               "dwf_apply_device")

    # solvers ------------------------------------------------------------------------
    def solve(self, normal: int, b: np.ndarray, x0: np.ndarray | None = None, rtol: float = 1e-10,
              maxiter: int = 10000, method: int = DWF_SOLVE_CG64, raise_on_fail: bool = True):
        """Solve A x = b, A = D D^dag or D^dag D; returns (x, [stats dict per RHS])."""
        bb, nrhs = self._as_batch(b)
        x = np.empty_like(bb)
        x0p = None
        if x0 is not None:
            x0a, n0 = self._as_batch(x0)
            if n0 != nrhs:
                raise ValueError("x0 and b differ in batch size")
            x0p = _ptr(x0a)
        stats = (SolveStats * nrhs)()
        rc = self._lib.dwf_solve(self._h, normal, method, _ptr(bb), x0p, rtol, maxiter, nrhs,
                                 _ptr(x), ctypes.cast(stats, ctypes.c_void_p))
        if raise_on_fail:
            _check(self._lib, rc, "dwf_solve")
        return x.reshape(np.shape(b)), [s.as_dict() for s in stats], rc

    def solve_dirac(self, src: np.ndarray, rtol: float = 1e-10, maxiter: int = 10000,
                    method: int = DWF_SOLVE_EO_MIXED, x0: np.ndarray | None = None, raise_on_fail: bool = True):
        """x = D^-1 src (certified on the D^dag D system); returns (x, [stats], rc)."""
        bb, nrhs = self._as_batch(src)
        x = np.empty_like(bb)
        x0p = None
        if x0 is not None:
            x0a, _ = self._as_batch(x0)
            x0p = _ptr(x0a)
        stats = (SolveStats * nrhs)()
        rc = self._lib.dwf_solve_dirac(self._h, method, _ptr(bb), x0p, rtol, maxiter, nrhs, _ptr(x),
                                       ctypes.cast(stats, ctypes.c_void_p))
        if raise_on_fail:
            _check(self._lib, rc, "dwf_solve_dirac")
        return x.reshape(np.shape(src)), [s.as_dict() for s in stats], rc

    def solve_dirac_device(self, method: int, src_ptr: int, x0_ptr: int | None, rtol: float, maxiter: int,
                           nrhs: int, x_ptr: int, raise_on_fail: bool = True):
        stats = (SolveStats * nrhs)()
        rc = self._lib.dwf_solve_dirac_device(self._h, method, src_ptr, x0_ptr, rtol, maxiter, nrhs, x_ptr,
                                              ctypes.cast(stats, ctypes.c_void_p))
        if raise_on_fail:
            _check(self._lib, rc, "dwf_solve_dirac_device")
        return [s.as_dict() for s in stats], rc

    def solve_device(self, normal: int, method: int, b_ptr: int, x0_ptr: int | None, rtol: float,
                     maxiter: int, nrhs: int, x_ptr: int, raise_on_fail: bool = True):
        stats = (SolveStats * nrhs)()
        rc = self._lib.dwf_solve_device(self._h, normal, method, b_ptr, x0_ptr, rtol, maxiter, nrhs,
                                        x_ptr, ctypes.cast(stats, ctypes.c_void_p))
        if raise_on_fail:
            _check(self._lib, rc, "dwf_solve_device")
        return [s.as_dict() for s in stats], rc

    # force, gauge, HMC -----------------------------------------------------------------
    def pf_action_force(self, phi: np.ndarray, rtol: float, maxiter: int, eta_guess=None,
                        method: int = DWF_SOLVE_CG64, raise_on_fail: bool = True):
        """(action, force[2*V4], eta, stats, rc) as pseudofermion_action_force."""
        ph, n = self._as_batch(phi)
        if n != 1:
            raise ValueError("one pseudofermion field")
        g = None
        if eta_guess is not None:
            g, _ = self._as_batch(eta_guess)
        act = ctypes.c_double()
        force = np.zeros(2 * self.geom.V4)
        eta = np.empty_like(ph)
        st = SolveStats()
        rc = self._lib.dwf_pf_action_force(self._h, method, _ptr(ph), None if g is None else _ptr(g), rtol,
                                           maxiter, ctypes.byref(act), _ptr(force), _ptr(eta), ctypes.byref(st))
        if raise_on_fail:
            _check(self._lib, rc, "dwf_pf_action_force")
        return act.value, force, eta, st.as_dict(), rc

    def force_from_eta(self, eta: np.ndarray) -> np.ndarray:
        e, _ = self._as_batch(eta)
        force = np.zeros(2 * self.geom.V4)
        _check(self._lib, self._lib.dwf_force_from_eta(self._h, _ptr(e), _ptr(force)), "dwf_force_from_eta")
        return force

    def gauge_action_force(self, beta: float):
        act = ctypes.c_double()
        force = np.zeros(2 * self.geom.V4)
        _check(self._lib, self._lib.dwf_gauge_action_force(self._h, beta, ctypes.byref(act), _ptr(force)),
               "dwf_gauge_action_force")
        return act.value, force

    def average_plaquette(self) -> float:
        v = ctypes.c_double()
        _check(self._lib, self._lib.dwf_average_plaquette(self._h, ctypes.byref(v)), "dwf_average_plaquette")
        return v.value

    def get_gauge(self) -> np.ndarray:
        th = np.zeros(2 * self.geom.V4)
        _check(self._lib, self._lib.dwf_get_gauge(self._h, _ptr(th)), "dwf_get_gauge")
        return th

    def hmc_trajectory(self, chi: np.ndarray, momentum: np.ndarray, u: float, beta: float, epsilon: float,
                       steps: int, rtol: float, maxiter: int, integrator: int = DWF_HMC_DOMAIN_WALL,
                       method: int = DWF_SOLVE_CG64) -> dict:
        c, _ = self._as_batch(chi)
        p = np.ascontiguousarray(momentum, dtype=np.float64)
        P = HmcParams(beta, epsilon, steps, integrator, method, maxiter, rtol)
        R = HmcResult()
        _check(self._lib, self._lib.dwf_hmc_trajectory(self._h, ctypes.byref(P), _ptr(c), _ptr(p), u,
                                                       ctypes.byref(R)), "dwf_hmc_trajectory")
        return R.as_dict()

    # analyzer ----------------------------------------------------------------------------
    def measure(self, rtol: float, maxiter: int, wall: bool = True, point: bool = True, max_momentum: int = 0,
                noise_vectors: int = 0, noise_seed: int = 0, method: int = DWF_SOLVE_CG64) -> dict:
        Nt, Nx = self.geom.Nt, self.geom.Nx
        nrhs = (4 if wall else 0) + (4 if point else 0) + noise_vectors
        cpp, cj5 = np.zeros(Nt), np.zeros(Nt)
        mom = np.zeros((max_momentum + 1) * Nt)
        loop = np.zeros(Nt * Nx, np.complex128)
        iters = np.zeros(max(nrhs, 1), np.int32)
        dres, pj5 = np.zeros(4), np.zeros(Nt)
        spec = MeasureSpec(method, maxiter, rtol, int(wall), int(point), max_momentum, noise_vectors, noise_seed,
                           None)
        out = MeasureOut(_ptr(cpp), _ptr(cj5), _ptr(mom), _ptr(loop), _ptr(iters), _ptr(dres), 0, 0, _ptr(pj5))
        _check(self._lib, self._lib.dwf_measure(self._h, ctypes.byref(spec), ctypes.byref(out)), "dwf_measure")
        return dict(CPP=cpp, CJ5=cj5, momentum_CPP=mom.reshape(max_momentum + 1, Nt), loop=loop.reshape(Nt, Nx),
                    iterations=iters[:nrhs], dirac_residual=dres, wall_nonconverged=out.wall_nonconverged,
                    point_CJ5=pj5)

    def measure_batch(self, ncfg: int, rtol: float, maxiter: int, wall: bool = True, point: bool = True,
                      max_momentum: int = 0, noise_vectors: int = 0, seeds=None,
                      method: int = DWF_SOLVE_CG64, point_sites=None) -> list:
        """dwf_measure for the ncfg configurations held (set_gauge_batch or
        set_gauge_groups); one dict each.  point_sites: None (the reference's 4 point
        sources) or [ncfg][2] = (t0, x0)."""
        lats = self.config_lattices() or [(self.geom.Nt, self.geom.Nx)] * ncfg
        if len(lats) != ncfg:
            raise ValueError("ncfg differs from the configurations held")
        npoint = 0 if not point else (4 if point_sites is None else 2)
        K = (4 if wall else 0) + npoint + noise_vectors
        keep, outs = [], (MeasureOut * ncfg)()
        for k in range(ncfg):
            Nt, Nx = lats[k]
            d = dict(CPP=np.zeros(Nt), CJ5=np.zeros(Nt), momentum_CPP=np.zeros((max_momentum + 1) * Nt),
                     loop=np.zeros(Nt * Nx, np.complex128), iterations=np.zeros(max(K, 1), np.int32),
                     dirac_residual=np.zeros(4), point_CJ5=np.zeros(Nt))
            keep.append(d)
            outs[k] = MeasureOut(*(_ptr(d[n]) for n in ("CPP", "CJ5", "momentum_CPP", "loop", "iterations",
                                                        "dirac_residual")), 0, 0, _ptr(d["point_CJ5"]))
        sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint64)
        ps = None if point_sites is None else np.ascontiguousarray(point_sites, dtype=np.int32).reshape(-1)
        spec = MeasureSpec(method, maxiter, rtol, int(wall), int(point), max_momentum, noise_vectors, 0,
                           None if ps is None else _ptr(ps))
        _check(self._lib, self._lib.dwf_measure_batch(self._h, ctypes.byref(spec), ncfg,
                                                      None if sd is None else _ptr(sd),
                                                      ctypes.cast(outs, ctypes.c_void_p)), "dwf_measure_batch")
        for k, d in enumerate(keep):
            Nt, Nx = lats[k]
            d["momentum_CPP"] = d["momentum_CPP"].reshape(max_momentum + 1, Nt)
            d["loop"] = d["loop"].reshape(Nt, Nx)
            d["iterations"] = d["iterations"][:K]
            d["wall_nonconverged"] = outs[k].wall_nonconverged
        return keep


class HostRng:
    """The callers' host random stream (mt19937_64 + libstdc++ distributions)."""

    def __init__(self, seed: int):
        self._lib = load_library()
        self._h = self._lib.dwf_rng_create(seed)

    def __del__(self):
        try:
            if self._h:
                self._lib.dwf_rng_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def dwf_heatbath(self, ndof: int, ngauge: int):
        chi = np.zeros(ndof, np.complex128)
        mom = np.zeros(ngauge)
        self._lib.dwf_rng_dwf_heatbath(self._h, ndof, ngauge, _ptr(chi), _ptr(mom))
        return chi, mom

    def wilson_heatbath(self, ngauge: int, ntot: int):
        chi = np.zeros(ntot, np.complex128)
        mom = np.zeros(ngauge)
        self._lib.dwf_rng_wilson_heatbath(self._h, ngauge, ntot, _ptr(mom), _ptr(chi))
        return mom, chi

    def uniform(self) -> float:
        return self._lib.dwf_rng_uniform(self._h)


def z4_noise(seed: int, nnoise: int, Nt: int, Nx: int) -> np.ndarray:
    eta = np.zeros(nnoise * Nt * Nx * 2, np.complex128)
    load_library().dwf_z4_noise(seed, nnoise, Nt * Nx * 2, _ptr(eta))
    return eta.reshape(nnoise, Nt, Nx, 2)


class DeviceArray:
    """A device buffer owned through libdwfhip (no other GPU runtime needed)."""

    def __init__(self, nbytes: int):
        self._lib = load_library()
        self.nbytes = int(nbytes)
        self.ptr = self._lib.dwf_device_alloc(self.nbytes)
        if not self.ptr:
            raise DwfError(DWF_ERR_HIP, self._lib.dwf_last_error().decode())

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("host array larger than the device buffer")
        _check(self._lib, self._lib.dwf_memcpy_htod(self.ptr, _ptr(a), a.nbytes), "dwf_memcpy_htod")

    def download(self, dtype, count: int | None = None) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        n = self.nbytes // itemsize if count is None else count
        out = np.empty(n, dtype)
        _check(self._lib, self._lib.dwf_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes), "dwf_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            self._lib.dwf_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TrainParams(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("loss", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("lr", ctypes.c_double),
        ("beta1", ctypes.c_double),
        ("beta2", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
    ]


class Network:
    """The learned preconditioner (dwf_net_*): weights on the device."""

    def __init__(self, Ls: int, seed: int = 0):
        self._lib = load_library()
        self.Ls = Ls
        h = ctypes.c_void_p()
        _check(self._lib, self._lib.dwf_net_create(ctypes.byref(h), Ls, seed), "dwf_net_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def num_weights(self) -> int:
        return int(self._lib.dwf_net_num_weights(self.Ls))

    def get_weights(self) -> np.ndarray:
        w = np.empty(self.num_weights(), np.float32)
        _check(self._lib, self._lib.dwf_net_get_weights(self._h, _ptr(w)), "dwf_net_get_weights")
        return w

    def set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != self.num_weights():
            raise ValueError("wrong number of weights")
        _check(self._lib, self._lib.dwf_net_set_weights(self._h, _ptr(w)), "dwf_net_set_weights")

    def apply(self, ctx: "DwfContext", v: np.ndarray, adjoint: bool = False) -> np.ndarray:
        """M v (or M^dag v) for host complex spinors [nrhs][ndof] (or one spinor)."""
        vb, n = ctx._as_batch(v)
        d_in = DeviceArray.from_host(vb.view(np.float64))
        d_out = DeviceArray(d_in.nbytes)
        _check(self._lib, self._lib.dwf_net_apply_device(ctx.handle, self._h, int(adjoint), d_in.ptr, d_out.ptr, n),
               "dwf_net_apply_device")
        out = d_out.download(np.float64).view(np.complex128)
        return out.reshape(np.shape(v))

    @staticmethod
    def _pairs(b, x):
        bb = np.ascontiguousarray(b, np.complex128).reshape(np.shape(b)[0], -1)
        db = DeviceArray.from_host(bb.view(np.float64).astype(np.float32))
        dx = None
        if x is not None:
            xx = np.ascontiguousarray(x, np.complex128).reshape(bb.shape)
            dx = DeviceArray.from_host(xx.view(np.float64).astype(np.float32))
        return db, dx, bb.shape[0]

    def loss_grad(self, ctx: "DwfContext", b, x, pair_cfg=None, loss: int = 0):
        """(loss, grad) over host pairs b[i], x[i] (complex spinors)."""
        db, dx, n = self._pairs(b, x)
        pc = np.ascontiguousarray(np.zeros(n) if pair_cfg is None else pair_cfg, np.int32)
        g = np.empty(self.num_weights(), np.float32)
        L = ctypes.c_double()
        _check(self._lib, self._lib.dwf_net_loss_grad_device(ctx.handle, self._h, loss, db.ptr,
                                                             dx.ptr if dx else None, _ptr(pc), n,
                                                             ctypes.byref(L), _ptr(g)), "dwf_net_loss_grad_device")
        return L.value, g

    def train_device(self, ctx: "DwfContext", b_ptr: int, x_ptr: int | None, pair_cfg, npairs: int, steps: int,
                     batch: int, loss: int = 1, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                     eps: float = 1e-8, seed: int = 0) -> np.ndarray:
        """Adam steps on FP32 device pairs; returns the per-step batch losses."""
        pc = np.ascontiguousarray(pair_cfg, np.int32)
        hist = np.zeros(steps)
        p = TrainParams(steps, batch, loss, 0, lr, beta1, beta2, eps, seed)
        _check(self._lib, self._lib.dwf_net_train_device(ctx.handle, self._h, b_ptr, x_ptr, _ptr(pc), npairs,
                                                         ctypes.byref(p), _ptr(hist)), "dwf_net_train_device")
        return hist

    def train(self, ctx: "DwfContext", b, x=None, pair_cfg=None, **kw) -> np.ndarray:
        db, dx, n = self._pairs(b, x)
        pc = np.zeros(n, np.int32) if pair_cfg is None else pair_cfg
        return self.train_device(ctx, db.ptr, dx.ptr if dx else None, pc, n, **kw)

    def close(self):
        if self._h:
            self._lib.dwf_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Timer:
    """HIP events on a context's stream."""

    def __init__(self):
        self._lib = load_library()
        self._h = self._lib.dwf_timer_create()

    def start(self, ctx: DwfContext):
        _check(self._lib, self._lib.dwf_timer_start(self._h, ctx.handle), "dwf_timer_start")

    def stop(self, ctx: DwfContext):
        _check(self._lib, self._lib.dwf_timer_stop(self._h, ctx.handle), "dwf_timer_stop")

    def elapsed_ms(self) -> float:
        v = self._lib.dwf_timer_elapsed_ms(self._h)
        if v < 0:
            raise DwfError(DWF_ERR_HIP, "dwf_timer_elapsed_ms failed")
        return v

    def __del__(self):
        try:
            self._lib.dwf_timer_destroy(self._h)
        except Exception:
            pass
