"""Pins the oracle of the learned preconditioner (oracle/oracle_net.cpp) by mathematics:

* the whole network equals a product of dense layer matrices assembled here from the
  independent shift/link matrices of dense_ref (the same transporters as the Dirac
  operator's hops), forward and adjoint;
* gauge covariance M[theta^g](g v) = g M[theta](v);
* the analytic loss gradient matches central finite differences;
* the bf16-quantized mode stays within bf16 accuracy of the FP64 mode.
"""

import numpy as np
import pytest

import oracle
from dense_ref import rand_spinor, rand_theta, shift_matrices

CIN, COUT, NP = (2, 8, 8), (8, 8, 2), 7


def rand_weights(rng, Ls, scale=0.3):
    n = oracle.net_num_weights(Ls)
    return rng.normal(size=n) * scale


def split_weights(w, Ls):
    """[(Ls, 7, Cout, Cin) complex] per layer from the flat real array."""
    c = w[0::2] + 1j * w[1::2]
    out, o = [], 0
    for l in range(3):
        n = Ls * NP * COUT[l] * CIN[l]
        out.append(c[o:o + n].reshape(Ls, NP, COUT[l], CIN[l]))
        o += n
    assert o == c.size
    return out


def transport_matrices(theta, Nt, Nx):
    """V4 x V4 transporters of the 7 stencil points, from dense_ref's shifts and links:
    +mu: U_mu T_mu, -mu: T_mu^dag U_mu^dag (antiperiodic t, periodic x)."""
    th = np.asarray(theta, float).reshape(2, Nt * Nx)
    Ut, Ux = np.diag(np.exp(1j * th[0])), np.diag(np.exp(1j * th[1]))
    Tt, Tx = shift_matrices(Nt, Nx, -1.0, 1.0)
    I = np.eye(Nt * Nx, dtype=complex)
    return [I, Ut @ Tt, Tt.conj().T @ Ut.conj().T, Ux @ Tx, Tx.conj().T @ Ux.conj().T, I, I]


def s_shifts(Ls):
    I = np.eye(Ls)
    up = np.roll(I, 1, axis=1)  # (S v)(s) = v(s+1)
    return [I, I, I, I, I, up, up.T]


def dense_layer(W, P, S, Ls):
    """sum_p sum_s (e_s e_s^T S_p) (x) P_p (x) W_p(s) in the [s][site][channel] order."""
    V4 = P[0].shape[0]
    co, ci = W.shape[2], W.shape[3]
    L = np.zeros((Ls * V4 * co, Ls * V4 * ci), complex)
    for p in range(NP):
        for s in range(Ls):
            E = np.zeros((Ls, Ls))
            E[s, s] = 1.0
            L += np.kron(np.kron(E @ S[p], P[p]), W[s, p])
    return L


def dense_net(theta, w, Nt, Nx, Ls):
    P, S = transport_matrices(theta, Nt, Nx), s_shifts(Ls)
    Ws = split_weights(w, Ls)
    M = np.eye(Ls * Nt * Nx * 2, dtype=complex)
    for l in range(3):
        M = dense_layer(Ws[l], P, S, Ls) @ M
    return M


@pytest.mark.parametrize("Nt,Nx,Ls", [(4, 3, 3), (2, 4, 2)])
def test_net_matches_dense_product(Nt, Nx, Ls):
    rng = np.random.default_rng(Nt * 10 + Nx)
    th = rand_theta(rng, Nt, Nx, sigma=1.0)
    w = rand_weights(rng, Ls)
    M = dense_net(th, w, Nt, Nx, Ls)
    v = rand_spinor(rng, 2 * Ls * Nt * Nx * 2).reshape(2, -1)
    z = oracle.net_apply(th, w, v, Nt, Nx, Ls)
    for r in range(2):
        assert np.linalg.norm(z[r] - M @ v[r]) <= 1e-12 * np.linalg.norm(M @ v[r])
    za = oracle.net_apply(th, w, v, Nt, Nx, Ls, adj=True)
    for r in range(2):
        want = M.conj().T @ v[r]
        assert np.linalg.norm(za[r] - want) <= 1e-12 * np.linalg.norm(want)


def test_plane_wave_closed_form():
    """Only the +t point in the middle layer (identity embeddings around it), constant
    theta_t = phi: M v = e^{i phi} e^{i k_t} v for an antiperiodic plane wave."""
    Nt, Nx, Ls = 6, 4, 2
    phi = 0.37
    th = np.zeros(2 * Nt * Nx)
    th[: Nt * Nx] = phi
    Ws = [np.zeros((Ls, NP, COUT[l], CIN[l]), complex) for l in range(3)]
    for s in range(Ls):
        Ws[0][s, 0, :2, :2] = np.eye(2)
        Ws[1][s, 1] = np.eye(8)
        Ws[2][s, 0, :2, :2] = np.eye(2)
    w = np.concatenate([x.reshape(-1) for x in Ws])
    w = np.stack([w.real, w.imag], axis=1).reshape(-1)
    kt, kx = 3 * np.pi / Nt, 2 * np.pi / Nx
    t, x = np.meshgrid(np.arange(Nt), np.arange(Nx), indexing="ij")
    wave = np.exp(1j * (kt * t + kx * x))
    v = np.zeros((Ls, Nt, Nx, 2), complex)
    v[..., 0] = wave
    v[..., 1] = 0.5 * wave
    z = oracle.net_apply(th, w, v.reshape(-1), Nt, Nx, Ls)
    assert np.allclose(z, np.exp(1j * (phi + kt)) * v.reshape(-1), atol=1e-13)


def test_gauge_covariance():
    Nt, Nx, Ls = 6, 4, 3
    rng = np.random.default_rng(5)
    th = rand_theta(rng, Nt, Nx, sigma=1.0)
    w = rand_weights(rng, Ls)
    alpha = rng.uniform(-np.pi, np.pi, size=(Nt, Nx))
    t2 = th.reshape(2, Nt, Nx).copy()
    t2[0] += alpha - np.roll(alpha, -1, axis=0)
    t2[1] += alpha - np.roll(alpha, -1, axis=1)
    g = np.exp(1j * alpha)[None, :, :, None]
    v = rand_spinor(rng, Ls * Nt * Nx * 2).reshape(Ls, Nt, Nx, 2)
    z = oracle.net_apply(th, w, v.reshape(-1), Nt, Nx, Ls).reshape(v.shape)
    zg = oracle.net_apply(t2.reshape(-1), w, (g * v).reshape(-1), Nt, Nx, Ls).reshape(v.shape)
    assert np.allclose(zg, g * z, atol=1e-12 * np.abs(z).max())


@pytest.mark.parametrize("loss", [0, 1])
def test_loss_gradient_finite_differences(loss):
    Nt, Nx, Ls = 4, 4, 2
    rng = np.random.default_rng(8 + loss)
    th = rand_theta(rng, Nt, Nx)
    w = rand_weights(rng, Ls)
    n = Ls * Nt * Nx * 2
    b = rand_spinor(rng, 3 * n).reshape(3, n)
    x = rand_spinor(rng, 3 * n).reshape(3, n)
    L, g = oracle.net_loss_grad(th, w, b, x, Nt, Nx, Ls, loss=loss)
    # the loss itself from net_apply (and the oracle's D^dag D for loss 1)
    z = oracle.net_apply(th, w, b, Nt, Nx, Ls)
    if loss == 0:
        want = np.mean([np.linalg.norm(z[i] - x[i]) ** 2 / np.linalg.norm(x[i]) ** 2 for i in range(3)])
    else:
        A = lambda v: oracle.dwf_apply(3, th, v, Nt, Nx, Ls, 1.0, 0.05)
        want = np.mean([np.linalg.norm(A(z[i]) - b[i]) ** 2 / np.linalg.norm(b[i]) ** 2 for i in range(3)])
    assert abs(L - want) <= 1e-12 * want
    n1 = 2 * Ls * NP * 8 * 2
    n2 = n1 + 2 * Ls * NP * 64
    idx = np.concatenate([rng.choice(n1, 10, replace=False), n1 + rng.choice(n2 - n1, 10, replace=False),
                          n2 + rng.choice(w.size - n2, 10, replace=False)])
    h = 1e-5
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fp = oracle.net_loss_grad(th, wp, b, x, Nt, Nx, Ls, loss=loss)[0]
        fm = oracle.net_loss_grad(th, wm, b, x, Nt, Nx, Ls, loss=loss)[0]
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g).max()), (k, fd, g[k])


def test_quantized_mode_is_bf16_close():
    Nt, Nx, Ls = 8, 6, 4
    rng = np.random.default_rng(3)
    th = rand_theta(rng, Nt, Nx)
    w = rand_weights(rng, Ls)
    n = Ls * Nt * Nx * 2
    v = rand_spinor(rng, 2 * n).reshape(2, n)
    z = oracle.net_apply(th, w, v, Nt, Nx, Ls)
    zq = oracle.net_apply(th, w, v, Nt, Nx, Ls, bf16=True)
    assert 0 < np.linalg.norm(zq - z) / np.linalg.norm(z) < 2e-2
    x = rand_spinor(rng, 2 * n).reshape(2, n)
    l, g = oracle.net_loss_grad(th, w, v, x, Nt, Nx, Ls)
    lq, gq = oracle.net_loss_grad(th, w, v, x, Nt, Nx, Ls, bf16=True)
    assert abs(lq - l) <= 1e-2 * l
    assert np.linalg.norm(gq - g) <= 3e-2 * np.linalg.norm(g)
