"""The oracle: ctypes wrappers of oracle/liboracle.so (oracle/oracle.cpp).

TEST INFRASTRUCTURE.  Only tests/, the smoke check and bench.py's cpu_baseline leg
use this module, as the checker of libdwfhip.so or as the timed CPU baseline; the
product never imports it.  Arrays use the reference layouts (interleaved complex128
spinors, [2][Nt][Nx] angles).  Build with ``make -C oracle``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

ORC_OK, ORC_ARG, ORC_NOT_CONVERGED, ORC_ZERO_DENOM = 0, 1, 3, 4

_P, _I, _D, _U = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong

_PROTOS = {
    "orc_dwf_apply": (_I, [_I, _I, _I, _I, _D, _D, _P, _P, _P]),
    "orc_dwf_cg": (_I, [_I, _I, _I, _I, _D, _D, _P, _P, _P, _D, _I, _P, _P, _P, _P]),
    "orc_dwf_leapfrog": (_I, [_I, _I, _I, _D, _D, _D, _D, _I, _P, _P, _P, _D, _I, _P, _P]),
    "orc_dwf_cg_many": (_I, [_I, _P, _P, _I, _D, _D, _P, _P, _D, _I, _P, _P, _P]),
    "orc_dwf_pf_action_force": (_I, [_I, _I, _I, _D, _D, _P, _P, _P, _D, _I, _P, _P, _P, _P, _P]),
    "orc_dwf_force_from_eta": (_I, [_I, _I, _I, _D, _D, _P, _P, _P]),
    "orc_gauge_action_force": (_D, [_I, _I, _D, _P, _P]),
    "orc_average_plaquette": (_D, [_I, _I, _P]),
    "orc_dwf_hmc": (_I, [_I, _I, _I, _D, _D, _D, _D, _I, _I, _U, _D, _I, _P, _P, _P, _P, _P, _P]),
    "orc_dwf_cpp_cj5": (_I, [_I, _I, _I, _D, _D, _P, _D, _I, _P, _P, _P, _P, _P]),
    "orc_z4_noise": (None, [_I, _I, _I, _U, _P]),
    "orc_dwf_loop_density": (_I, [_I, _I, _I, _D, _D, _P, _U, _I, _D, _I, _P]),
    "orc_dwf_momentum_cpp": (_I, [_I, _I, _I, _D, _D, _P, _I, _D, _I, _P]),
    "orc_wilson_apply": (_I, [_I, _I, _I, _D, _D, _D, _P, _P, _P]),
    "orc_wilson_solve_cgnr": (_I, [_I, _I, _D, _D, _D, _P, _P, _D, _I, _P, _P, _P, _P]),
    "orc_wilson_solve_dddag": (_I, [_I, _I, _D, _D, _D, _P, _P, _D, _I, _P, _P, _P]),
    "orc_wilson_pf_action_force": (_I, [_I, _I, _D, _D, _D, _P, _P, _D, _I, _P, _P, _P]),
    "orc_wilson_force_from_eta": (_I, [_I, _I, _D, _D, _D, _P, _P, _P]),
    "orc_wrap_angle": (_D, [_D]),
    "orc_wilson_hmc": (_I, [_I, _I, _D, _D, _D, _D, _D, _I, _I, _U, _D, _I, _P, _P, _P, _P, _P]),
    "orc_wilson_pseudoscalar_correlator": (_I, [_I, _I, _D, _D, _D, _P, _I, _I, _D, _I, _P]),
    "orc_mt19937_64_nth": (_U, [_U, _I]),
    "orc_dwf_heatbath_draws": (None, [_I, _I, _U, _P, _P, _P]),
    "orc_net_num_weights": (ctypes.c_longlong, [_I]),
    "orc_net_apply": (_I, [_I, _I, _I, _P, _P, _P, _P, _I, _I, _I]),
    "orc_net_loss_grad": (_I, [_I, _I, _I, _D, _D, _P, _P, _P, _P, _I, _I, _I, _P, _P]),
}


def build(jobs: int = 1):
    subprocess.run(["make", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _c(a, dtype=np.complex128):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return None if a is None else a.ctypes.data


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


# ------------------------------------------------------------------------------ DWF
def dwf_apply(op, theta, psi, Nt, Nx, Ls, M5, mf):
    psi = _c(psi)
    out = np.empty_like(psi)
    th = _c(theta, np.float64)
    rc = lib().orc_dwf_apply(op, Nt, Nx, Ls, M5, mf, _p(th), _p(psi), _p(out))
    if rc:
        raise OracleError(rc, "orc_dwf_apply")
    return out


def dwf_cg(normal, theta, b, Nt, Nx, Ls, M5, mf, rtol, maxiter, guess=None):
    b = _c(b)
    x = np.empty_like(b)
    g = None if guess is None else _c(guess)
    it, rr, cv = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    th = _c(theta, np.float64)
    rc = lib().orc_dwf_cg(normal, Nt, Nx, Ls, M5, mf, _p(th), _p(b), _p(g), rtol, maxiter, _p(x),
                          ctypes.byref(it), ctypes.byref(rr), ctypes.byref(cv))
    return x, it.value, rr.value, bool(cv.value), rc


def dwf_cg_many(thetas, bs, lattices, Ls, M5, mf, rtol, maxiter):
    """Independent D^dag D solves (one per right-hand side b[i] on lattices[i] = (Nt, Nx)
    with links thetas[i]) spread over the OpenMP threads, as the reference's analyzers
    spread configurations.  Returns (solutions, iterations, converged, threads)."""
    n = len(bs)
    th = [_c(t, np.float64) for t in thetas]
    bb = [_c(b) for b in bs]
    xs = [np.empty_like(b) for b in bb]
    Nt = np.array([l[0] for l in lattices], np.int32)
    Nx = np.array([l[1] for l in lattices], np.int32)
    def ptrs(arrs):
        return ctypes.cast((ctypes.c_void_p * n)(*[_p(a) for a in arrs]), ctypes.c_void_p)

    it = np.zeros(n, np.int32)
    cv = np.zeros(n, np.int32)
    pt, pb, px = ptrs(th), ptrs(bb), ptrs(xs)
    threads = lib().orc_dwf_cg_many(n, _p(Nt), _p(Nx), Ls, M5, mf, pt, pb, rtol, maxiter, px, _p(it), _p(cv))
    return xs, it, cv.astype(bool), threads


def dwf_pf_action_force(theta, phi, Nt, Nx, Ls, M5, mf, rtol, maxiter, guess=None):
    phi = _c(phi)
    th = _c(theta, np.float64)
    g = None if guess is None else _c(guess)
    act, it, rr = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
    force = np.zeros(2 * Nt * Nx)
    eta = np.empty_like(phi)
    rc = lib().orc_dwf_pf_action_force(Nt, Nx, Ls, M5, mf, _p(th), _p(phi), _p(g), rtol, maxiter,
                                       ctypes.byref(act), _p(force), _p(eta), ctypes.byref(it),
                                       ctypes.byref(rr))
    return act.value, force, eta, it.value, rr.value, rc


def dwf_force_from_eta(theta, eta, Nt, Nx, Ls, M5, mf):
    th, eta = _c(theta, np.float64), _c(eta)
    force = np.zeros(2 * Nt * Nx)
    lib().orc_dwf_force_from_eta(Nt, Nx, Ls, M5, mf, _p(th), _p(eta), _p(force))
    return force


def gauge_action_force(theta, Nt, Nx, beta):
    th = _c(theta, np.float64)
    force = np.zeros(2 * Nt * Nx)
    s = lib().orc_gauge_action_force(Nt, Nx, beta, _p(th), _p(force))
    return s, force


def average_plaquette(theta, Nt, Nx):
    th = _c(theta, np.float64)
    return lib().orc_average_plaquette(Nt, Nx, _p(th))


def dwf_hmc(theta, Nt, Nx, Ls, M5, mf, beta, epsilon, steps, ntraj, seed, pf_rtol, pf_maxiter):
    th = _c(theta, np.float64).copy()
    acc = np.zeros(ntraj, np.int32)
    pacc, dh, plaq = np.zeros(ntraj), np.zeros(ntraj), np.zeros(ntraj)
    iters = ctypes.c_int()
    rc = lib().orc_dwf_hmc(Nt, Nx, Ls, M5, mf, beta, epsilon, steps, ntraj, seed, pf_rtol, pf_maxiter,
                           _p(th), _p(acc), _p(pacc), _p(dh), _p(plaq), ctypes.byref(iters))
    if rc:
        raise OracleError(rc, "orc_dwf_hmc")
    return dict(theta=th, accepted=acc, acc_prob=pacc, delta_h=dh, plaquette=plaq, cg_iterations=iters.value)


def dwf_leapfrog(theta, momentum, phi, Nt, Nx, Ls, M5, mf, beta, epsilon, steps, pf_rtol, pf_maxiter):
    """One leapfrog trajectory of the DWF HMC; returns (theta, momentum, H_initial, H_final)."""
    th = _c(theta, np.float64).copy()
    p = _c(momentum, np.float64).copy()
    ph = _c(phi)
    h0, h1 = ctypes.c_double(), ctypes.c_double()
    rc = lib().orc_dwf_leapfrog(Nt, Nx, Ls, M5, mf, beta, epsilon, steps, _p(th), _p(p), _p(ph), pf_rtol, pf_maxiter,
                                ctypes.byref(h0), ctypes.byref(h1))
    if rc:
        raise OracleError(rc, "orc_dwf_leapfrog")
    return th, p, h0.value, h1.value


def dwf_cpp_cj5(theta, Nt, Nx, Ls, M5, mf, rtol, maxiter):
    th = _c(theta, np.float64)
    cpp, cj5 = np.zeros(Nt), np.zeros(Nt)
    nc = ctypes.c_int()
    iters = np.zeros(4, np.int32)
    dres = np.zeros(4)
    lib().orc_dwf_cpp_cj5(Nt, Nx, Ls, M5, mf, _p(th), rtol, maxiter, _p(cpp), _p(cj5), ctypes.byref(nc),
                          _p(iters), _p(dres))
    return cpp, cj5, nc.value, iters, dres


def z4_noise(Nt, Nx, nnoise, seed):
    eta = np.zeros(nnoise * Nt * Nx * 2, np.complex128)
    lib().orc_z4_noise(Nt, Nx, nnoise, seed, _p(eta))
    return eta.reshape(nnoise, Nt, Nx, 2)


def dwf_loop_density(theta, Nt, Nx, Ls, M5, mf, seed, nnoise, rtol, maxiter):
    th = _c(theta, np.float64)
    loop = np.zeros(Nt * Nx, np.complex128)
    rc = lib().orc_dwf_loop_density(Nt, Nx, Ls, M5, mf, _p(th), seed, nnoise, rtol, maxiter, _p(loop))
    if rc:
        raise OracleError(rc, "orc_dwf_loop_density")
    return loop


def dwf_momentum_cpp(theta, Nt, Nx, Ls, M5, mf, max_momentum, rtol, maxiter):
    th = _c(theta, np.float64)
    res = np.zeros((max_momentum + 1) * Nt)
    rc = lib().orc_dwf_momentum_cpp(Nt, Nx, Ls, M5, mf, _p(th), max_momentum, rtol, maxiter, _p(res))
    if rc:
        raise OracleError(rc, "orc_dwf_momentum_cpp")
    return res.reshape(max_momentum + 1, Nt)


# --------------------------------------------------------------------------- Wilson
def wilson_apply(op, theta, psi, Nt, Nx, m0, bct, bcx):
    psi = _c(psi)
    out = np.empty_like(psi)
    th = _c(theta, np.float64)
    lib().orc_wilson_apply(op, Nt, Nx, m0, bct, bcx, _p(th), _p(psi), _p(out))
    return out


def wilson_solve_cgnr(theta, src, Nt, Nx, m0, bct, bcx, tol, maxit):
    th, src = _c(theta, np.float64), _c(src)
    x = np.empty_like(src)
    it, nr, dr = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    rc = lib().orc_wilson_solve_cgnr(Nt, Nx, m0, bct, bcx, _p(th), _p(src), tol, maxit, _p(x),
                                     ctypes.byref(it), ctypes.byref(nr), ctypes.byref(dr))
    return x, it.value, nr.value, dr.value, rc


def wilson_solve_dddag(theta, phi, Nt, Nx, m0, bct, bcx, tol, maxit):
    th, phi = _c(theta, np.float64), _c(phi)
    eta = np.empty_like(phi)
    it, res = ctypes.c_int(), ctypes.c_double()
    rc = lib().orc_wilson_solve_dddag(Nt, Nx, m0, bct, bcx, _p(th), _p(phi), tol, maxit, _p(eta),
                                      ctypes.byref(it), ctypes.byref(res))
    return eta, it.value, res.value, rc


def wilson_pf_action_force(theta, phi, Nt, Nx, m0, bct, bcx, tol, maxit):
    th, phi = _c(theta, np.float64), _c(phi)
    act = ctypes.c_double()
    force = np.zeros(2 * Nt * Nx)
    eta = np.empty_like(phi)
    rc = lib().orc_wilson_pf_action_force(Nt, Nx, m0, bct, bcx, _p(th), _p(phi), tol, maxit,
                                          ctypes.byref(act), _p(force), _p(eta))
    return act.value, force, eta, rc


def wilson_force_from_eta(theta, eta, Nt, Nx, m0, bct, bcx):
    th, eta = _c(theta, np.float64), _c(eta)
    force = np.zeros(2 * Nt * Nx)
    lib().orc_wilson_force_from_eta(Nt, Nx, m0, bct, bcx, _p(th), _p(eta), _p(force))
    return force


def wrap_angle(a):
    return lib().orc_wrap_angle(a)


def wilson_hmc(theta, Nt, Nx, m0, bct, bcx, beta, tau, steps, ntraj, seed, tol, maxit):
    th = _c(theta, np.float64).copy()
    acc = np.zeros(ntraj, np.int32)
    pacc, dh, plaq = np.zeros(ntraj), np.zeros(ntraj), np.zeros(ntraj)
    rc = lib().orc_wilson_hmc(Nt, Nx, m0, bct, bcx, beta, tau, steps, ntraj, seed, tol, maxit, _p(th),
                              _p(acc), _p(pacc), _p(dh), _p(plaq))
    if rc:
        raise OracleError(rc, "orc_wilson_hmc")
    return dict(theta=th, accepted=acc, acc_prob=pacc, delta_h=dh, plaquette=plaq)


def wilson_pseudoscalar_correlator(theta, Nt, Nx, m0, bct, bcx, source_t, source_x, tol, maxit):
    th = _c(theta, np.float64)
    C = np.zeros(Nt)
    rc = lib().orc_wilson_pseudoscalar_correlator(Nt, Nx, m0, bct, bcx, _p(th), source_t, source_x,
                                                  tol, maxit, _p(C))
    if rc:
        raise OracleError(rc, "orc_wilson_pseudoscalar_correlator")
    return C


def mt19937_64_nth(seed, n):
    return lib().orc_mt19937_64_nth(seed, n)


def dwf_heatbath_draws(ndof, ngauge, seed):
    chi = np.zeros(ndof, np.complex128)
    mom = np.zeros(ngauge)
    u = ctypes.c_double()
    lib().orc_dwf_heatbath_draws(ndof, ngauge, seed, _p(chi), _p(mom), ctypes.byref(u))
    return chi, mom, u.value


# ------------------------------------------------------------ learned preconditioner

def net_num_weights(Ls):
    """Real parameters of the 2 -> 8 -> 8 -> 2 network (oracle_net.cpp)."""
    return int(lib().orc_net_num_weights(Ls))


def net_apply(theta, w, v, Nt, Nx, Ls, adj=False, bf16=False):
    """z = M v (or M^dag v); v [nrhs][ndof] complex (or one spinor)."""
    v = _c(v)
    one = v.ndim == 1
    vb = v.reshape(-1, Ls * Nt * Nx * 2)
    z = np.empty_like(vb)
    w = _c(w, np.float64)
    rc = lib().orc_net_apply(Nt, Nx, Ls, _p(_c(theta, np.float64)), _p(w), _p(vb), _p(z), vb.shape[0],
                             int(adj), int(bf16))
    if rc:
        raise OracleError(rc, "orc_net_apply: bad argument")
    return z[0] if one else z


def net_loss_grad(theta, w, b, x, Nt, Nx, Ls, bf16=False, loss=0, M5=1.0, mf=0.05):
    """(loss, grad): loss 0 = mean_i ||M b_i - x_i||^2 / ||x_i||^2, loss 1 =
    mean_i ||D^dag D M b_i - b_i||^2 / ||b_i||^2 (x unused)."""
    b = _c(b).reshape(-1, Ls * Nt * Nx * 2)
    xx = None if x is None else _c(x).reshape(b.shape)
    w = _c(w, np.float64)
    g = np.zeros_like(w)
    L = ctypes.c_double()
    rc = lib().orc_net_loss_grad(Nt, Nx, Ls, M5, mf, _p(_c(theta, np.float64)), _p(w), _p(b), _p(xx), b.shape[0],
                                 loss, int(bf16), ctypes.byref(L), _p(g))
    if rc:
        raise OracleError(rc, "orc_net_loss_grad: bad argument")
    return L.value, g
