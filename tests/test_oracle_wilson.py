"""Pins the oracle's Wilson restatement (wilson/include/*.hpp) against mathematics:
an independent Kronecker-product matrix with runtime boundary signs, dense solves,
finite-difference forces, the dense-inverse point-source correlator, wrap_angle and
HMC dH scaling.  Eigen's CG iteration counting is library behaviour (Eigen 3 is absent
from the reference tree and from this image): parity unpinned, the solution is pinned.
"""

import numpy as np
import pytest

import oracle
from dense_ref import rand_spinor, wilson_4d

NT, NX, M0 = 4, 6, 0.1


@pytest.mark.parametrize("bct,bcx", [(-1.0, 1.0), (1.0, -1.0), (-1.0, -1.0)])
def test_wilson_matrix_matches_kronecker(rng, bct, bcx):
    th = rng.uniform(-np.pi, np.pi, 2 * NT * NX)
    n = NT * NX * 2
    W = wilson_4d(th, NT, NX, 2.0 + M0, bct, bcx)
    E = np.eye(n, dtype=complex)
    D = np.stack([oracle.wilson_apply(0, th, E[:, j], NT, NX, M0, bct, bcx) for j in range(n)], axis=1)
    Dd = np.stack([oracle.wilson_apply(1, th, E[:, j], NT, NX, M0, bct, bcx) for j in range(n)], axis=1)
    assert np.max(np.abs(D - W)) < 1e-14
    assert np.max(np.abs(Dd - W.conj().T)) < 1e-14


def test_dwf_4d_part_is_wilson_with_shifted_mass(rng):
    """The DWF operator at Ls=1 without wall hops is the Wilson operator with 2+m0 = 3-M5."""
    th = rng.normal(0, 0.5, 2 * NT * NX)
    psi = rand_spinor(rng, NT * NX * 2)
    # Ls = 1: the fifth-dimension hops couple s=0 to itself with +mf on both spins
    got = oracle.dwf_apply(0, th, psi, NT, NX, 1, 1.0 - M0, 0.0)
    want = oracle.wilson_apply(0, th, psi, NT, NX, M0, -1.0, 1.0)
    assert np.allclose(got, want, atol=1e-14)


def test_cgnr_and_dddag_match_dense(rng):
    th = rng.uniform(-np.pi, np.pi, 2 * NT * NX)
    W = wilson_4d(th, NT, NX, 2.0 + M0, -1.0, 1.0)
    src = rand_spinor(rng, NT * NX * 2)
    x, it, nres, dres, rc = oracle.wilson_solve_cgnr(th, src, NT, NX, M0, -1.0, 1.0, 1e-12, 2000)
    assert rc == 0 and nres < 20e-12
    assert np.linalg.norm(x - np.linalg.solve(W, src)) / np.linalg.norm(x) < 1e-9
    assert abs(dres - np.linalg.norm(W @ x - src) / np.linalg.norm(src)) < 1e-12
    eta, it, res, rc = oracle.wilson_solve_dddag(th, src, NT, NX, M0, -1.0, 1.0, 1e-12, 2000)
    assert rc == 0
    assert np.linalg.norm(eta - np.linalg.solve(W @ W.conj().T, src)) / np.linalg.norm(eta) < 1e-9
    # failure path: the reference throws when the residual exceeds 20 x tol
    *_, rc = oracle.wilson_solve_cgnr(th, src, NT, NX, M0, -1.0, 1.0, 1e-14, 2)
    assert rc == oracle.ORC_NOT_CONVERGED


def test_wilson_force_finite_difference(rng):
    th = rng.uniform(-np.pi, np.pi, 2 * NT * NX)
    phi = rand_spinor(rng, NT * NX * 2)

    def S(t):
        W = wilson_4d(t, NT, NX, 2.0 + M0, -1.0, 1.0)
        return np.real(np.vdot(phi, np.linalg.solve(W @ W.conj().T, phi)))

    act, F, eta, rc = oracle.wilson_pf_action_force(th, phi, NT, NX, M0, -1.0, 1.0, 1e-13, 5000)
    assert rc == 0 and abs(act - S(th)) < 1e-9 * abs(act)
    h = 1e-5
    for link in range(0, 2 * NT * NX, 5):
        tp, tm = th.copy(), th.copy()
        tp[link] += h
        tm[link] -= h
        fd = (S(tp) - S(tm)) / (2 * h)
        assert abs(F[link] - fd) < 1e-6 * max(1, abs(fd)), (link, F[link], fd)


def test_point_correlator_against_dense_inverse(rng):
    th = rng.uniform(-np.pi, np.pi, 2 * NT * NX)
    W = wilson_4d(th, NT, NX, 2.0 + M0, -1.0, 1.0)
    G = np.linalg.inv(W)
    t0, x0 = 1, 2
    C = oracle.wilson_pseudoscalar_correlator(th, NT, NX, M0, -1.0, 1.0, t0, x0, 1e-13, 5000)
    cols = G[:, [(t0 * NX + x0) * 2, (t0 * NX + x0) * 2 + 1]].reshape(NT, NX, 2, 2)
    want = np.sum(np.abs(cols) ** 2, axis=(1, 2, 3))
    assert np.allclose(C, want, rtol=1e-9)


def test_wrap_angle():
    for a in (-7.0, -np.pi, -1.0, 0.0, 2.0, np.pi - 1e-12, 9.5):
        w = oracle.wrap_angle(a)
        assert -np.pi <= w < np.pi
        assert abs(np.exp(1j * w) - np.exp(1j * a)) < 1e-12


def test_wilson_hmc_delta_h_scaling_and_determinism():
    th0 = np.zeros(2 * NT * NX)
    dh = []
    for steps in (4, 8):
        r = oracle.wilson_hmc(th0, NT, NX, M0, -1.0, 1.0, 2.0, 0.5, steps, 1, 1234, 1e-12, 5000)
        dh.append(abs(r["delta_h"][0]))
    assert 2.5 < dh[0] / dh[1] < 6.0, dh
    a = oracle.wilson_hmc(th0, NT, NX, M0, -1.0, 1.0, 2.0, 1.0, 10, 4, 1234, 1e-10, 5000)
    b = oracle.wilson_hmc(th0, NT, NX, M0, -1.0, 1.0, 2.0, 1.0, 10, 4, 1234, 1e-10, 5000)
    assert np.array_equal(a["delta_h"], b["delta_h"]) and np.array_equal(a["theta"], b["theta"])
    assert np.all(np.abs(a["theta"]) <= np.pi)
