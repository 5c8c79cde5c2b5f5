"""Independent dense constructions used to pin the oracle (tests only).

The Dirac operators are assembled from Kronecker products of shift, link and Pauli
matrices -- a construction that shares no code with the oracle's site loops -- in the
reference's index order ((s*Nt+t)*Nx+x)*2+spin (domain_wall/include/indexing.hpp:5-8).
"""

from __future__ import annotations

import numpy as np

S1 = np.array([[0, 1], [1, 0]], dtype=complex)
S2 = np.array([[0, -1j], [1j, 0]], dtype=complex)
S3 = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2, dtype=complex)
PR = (I2 + S3) / 2
PL = (I2 - S3) / 2


def shift_matrices(Nt, Nx, bc_t, bc_x):
    """T_t, T_x on V4 sites: (T_mu v)(x) = bc * v(x + mu)."""
    V4 = Nt * Nx
    Tt = np.zeros((V4, V4), dtype=complex)
    Tx = np.zeros((V4, V4), dtype=complex)
    for t in range(Nt):
        for x in range(Nx):
            i = t * Nx + x
            Tt[i, ((t + 1) % Nt) * Nx + x] += bc_t if t == Nt - 1 else 1.0
            Tx[i, t * Nx + (x + 1) % Nx] += bc_x if x == Nx - 1 else 1.0
    return Tt, Tx


def wilson_4d(theta, Nt, Nx, diag, bc_t, bc_x):
    """diag - 1/2 sum_mu [U_mu T_mu (1-g_mu) + T_mu^+ U_mu^+ (1+g_mu)], gamma_t=s1, gamma_x=s2."""
    th = np.asarray(theta, dtype=float).reshape(2, Nt * Nx)
    Ut = np.diag(np.exp(1j * th[0]))
    Ux = np.diag(np.exp(1j * th[1]))
    Tt, Tx = shift_matrices(Nt, Nx, bc_t, bc_x)
    V4 = Nt * Nx
    D = diag * np.eye(2 * V4, dtype=complex)
    D -= 0.5 * (np.kron(Ut @ Tt, I2 - S1) + np.kron(Tt.conj().T @ Ut.conj().T, I2 + S1))
    D -= 0.5 * (np.kron(Ux @ Tx, I2 - S2) + np.kron(Tx.conj().T @ Ux.conj().T, I2 + S2))
    return D


def dwf_dense(theta, Nt, Nx, Ls, M5, mf):
    """Shamir D = 1_Ls (x) D_W(-M5 + 1) + S_up (x) 1 (x) P_L + S_dn (x) 1 (x) P_R."""
    V4 = Nt * Nx
    DW = wilson_4d(theta, Nt, Nx, 3.0 - M5, -1.0, 1.0)
    Sup = np.zeros((Ls, Ls))
    Sdn = np.zeros((Ls, Ls))
    for s in range(Ls - 1):
        Sup[s, s + 1] = -1.0
        Sdn[s + 1, s] = -1.0
    Sup[Ls - 1, 0] = mf
    Sdn[0, Ls - 1] = mf
    I4 = np.eye(V4)
    return (np.kron(np.eye(Ls), DW) + np.kron(np.kron(Sup, I4), PL) + np.kron(np.kron(Sdn, I4), PR))


def reflect_gamma5(Nt, Nx, Ls):
    """R gamma5 with R: s -> Ls-1-s."""
    R = np.fliplr(np.eye(Ls))
    return np.kron(np.kron(R, np.eye(Nt * Nx)), S3)


def rand_theta(rng, Nt, Nx, sigma=0.5):
    return rng.normal(0.0, sigma, size=2 * Nt * Nx)


def rand_spinor(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)
