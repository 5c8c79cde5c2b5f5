"""Pins the oracle's DWF restatement (oracle/oracle.cpp) against mathematics.

Each check is chosen so that a plausible slip -- a sign, a swapped projector, a wrong
neighbour, a transposed Nt/Nx, a dropped wall term -- fails at least one of them:
an independent Kronecker-product matrix, the plane-wave closed form, adjointness,
gamma5-R hermiticity, gauge covariance, dense solves and central finite differences.
The lattices are non-square (Nt != Nx) so t/x transpositions show.
"""

import numpy as np
import pytest

import oracle
from dense_ref import PL, PR, S1, S2, dwf_dense, rand_spinor, rand_theta, reflect_gamma5

NT, NX, LS, M5, MF = 4, 6, 4, 1.0, 0.05


def _apply(op, th, v, Nt=NT, Nx=NX, Ls=LS, M5_=M5, mf=MF):
    return oracle.dwf_apply(op, th, v, Nt, Nx, Ls, M5_, mf)


def test_dense_matrix_matches_kronecker_construction(rng):
    th = rand_theta(rng, NT, NX, 1.0)
    n = LS * NT * NX * 2
    Dk = dwf_dense(th, NT, NX, LS, M5, MF)
    E = np.eye(n, dtype=complex)
    Dcols = np.stack([_apply(0, th, E[:, j]) for j in range(n)], axis=1)
    Ddcols = np.stack([_apply(1, th, E[:, j]) for j in range(n)], axis=1)
    assert np.max(np.abs(Dcols - Dk)) < 1e-14
    assert np.max(np.abs(Ddcols - Dk.conj().T)) < 1e-14


def test_normal_operators_compose(rng):
    th = rand_theta(rng, NT, NX)
    v = rand_spinor(rng, LS * NT * NX * 2)
    D = dwf_dense(th, NT, NX, LS, M5, MF)
    assert np.allclose(_apply(2, th, v), D @ (D.conj().T @ v), atol=1e-12)
    assert np.allclose(_apply(3, th, v), D.conj().T @ (D @ v), atol=1e-12)


@pytest.mark.parametrize("n_t,n_x", [(0, 0), (1, 2), (3, 5)])
def test_plane_wave_closed_form(rng, n_t, n_x):
    """Constant theta_mu = c_mu: D acts on e^{i(p_t t + p_x x)} chi(s) as a 2Ls x 2Ls
    matrix with D_W(k) = (3-M5) - cos k_t - cos k_x + i(s1 sin k_t + s2 sin k_x), k = p + c,
    p_t = 2 pi (n_t + 1/2) / Nt (antiperiodic), p_x = 2 pi n_x / Nx."""
    ct, cx = 0.37, -1.1
    th = np.concatenate([np.full(NT * NX, ct), np.full(NT * NX, cx)])
    pt = 2 * np.pi * (n_t + 0.5) / NT
    px = 2 * np.pi * n_x / NX
    kt, kx = pt + ct, px + cx
    DW = (3 - M5 - np.cos(kt) - np.cos(kx)) * np.eye(2) + 1j * (S1 * np.sin(kt) + S2 * np.sin(kx))
    chi = rand_spinor(rng, 2 * LS).reshape(LS, 2)
    # fifth dimension
    out_chi = chi @ DW.T
    for s in range(LS):
        up = -chi[s + 1, 1] if s < LS - 1 else MF * chi[0, 1]
        dn = -chi[s - 1, 0] if s > 0 else MF * chi[LS - 1, 0]
        out_chi[s, 1] += up
        out_chi[s, 0] += dn
    t = np.arange(NT)[:, None]
    x = np.arange(NX)[None, :]
    phase = np.exp(1j * (pt * t + px * x))  # [Nt, Nx]
    psi = phase[None, :, :, None] * chi[:, None, None, :]
    want = phase[None, :, :, None] * out_chi[:, None, None, :]
    got = _apply(0, th, psi.reshape(-1)).reshape(LS, NT, NX, 2)
    assert np.max(np.abs(got - want)) < 1e-12


def test_adjointness(rng):
    th = rand_theta(rng, NT, NX, 1.3)
    n = LS * NT * NX * 2
    chi, psi = rand_spinor(rng, n), rand_spinor(rng, n)
    lhs = np.vdot(chi, _apply(0, th, psi))
    rhs = np.vdot(_apply(1, th, chi), psi)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_gamma5_R_hermiticity(rng):
    th = rand_theta(rng, NT, NX, 0.8)
    n = LS * NT * NX * 2
    psi = rand_spinor(rng, n)
    G = reflect_gamma5(NT, NX, LS)
    assert np.allclose(_apply(1, th, psi), G @ _apply(0, th, G @ psi), atol=1e-13)


def test_normal_operators_related_by_gamma5_R(rng):
    """D D^dag = (R gamma5) D^dag D (R gamma5): the learned solver runs D D^dag solves
    (pseudofermion force, HMC) as D^dag D solves of the rotated source."""
    th = rand_theta(rng, NT, NX, 0.8)
    psi = rand_spinor(rng, LS * NT * NX * 2)
    G = reflect_gamma5(NT, NX, LS)
    lhs = _apply(2, th, psi)
    assert np.allclose(lhs, G @ _apply(3, th, G @ psi), atol=1e-13 * np.abs(lhs).max())
    assert not np.allclose(lhs, _apply(3, th, psi), atol=1e-3)  # the two operators differ


def test_gauge_covariance(rng):
    th = rand_theta(rng, NT, NX, 0.9).reshape(2, NT, NX)
    alpha = rng.uniform(-np.pi, np.pi, size=(NT, NX))
    thp = th.copy()
    thp[0] += alpha - np.roll(alpha, -1, axis=0)
    thp[1] += alpha - np.roll(alpha, -1, axis=1)
    n = LS * NT * NX * 2
    psi = rand_spinor(rng, n).reshape(LS, NT, NX, 2)
    ph = np.exp(1j * alpha)[None, :, :, None]
    for op in (0, 1):
        lhs = _apply(op, thp.reshape(-1), (ph * psi).reshape(-1)).reshape(psi.shape)
        rhs = ph * _apply(op, th.reshape(-1), psi.reshape(-1)).reshape(psi.shape)
        assert np.max(np.abs(lhs - rhs)) < 1e-12


@pytest.mark.parametrize("normal", [2, 3])
def test_cg_matches_dense_solve(rng, normal):
    th = rand_theta(rng, NT, NX)
    n = LS * NT * NX * 2
    b = rand_spinor(rng, n)
    D = dwf_dense(th, NT, NX, LS, M5, MF)
    A = D @ D.conj().T if normal == 2 else D.conj().T @ D
    x, it, rr, conv, rc = oracle.dwf_cg(normal, th, b, NT, NX, LS, M5, MF, 1e-12, 5000)
    assert rc == 0 and conv and rr <= 1e-12 and it > 0
    xs = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-9
    # the true residual agrees with the recursive one at this tolerance
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-11


def test_cg_special_cases(rng):
    th = rand_theta(rng, NT, NX)
    n = LS * NT * NX * 2
    zero = np.zeros(n, complex)
    x, it, rr, conv, rc = oracle.dwf_cg(3, th, zero, NT, NX, LS, M5, MF, 1e-8, 100)
    assert (it, rr, conv, rc) == (0, 0.0, True, 0) and not np.any(x)
    b = rand_spinor(rng, n)
    D = dwf_dense(th, NT, NX, LS, M5, MF)
    xs = np.linalg.solve(D.conj().T @ D, b)
    x, it, rr, conv, rc = oracle.dwf_cg(3, th, b, NT, NX, LS, M5, MF, 1e-8, 100, guess=xs)
    assert it == 0 and conv
    x, it, rr, conv, rc = oracle.dwf_cg(3, th, b, NT, NX, LS, M5, MF, 1e-14, 3)
    assert it == 3 and not conv and rc == 0


def _pf_action_dense(th, phi):
    D = dwf_dense(th, NT, NX, LS, M5, MF)
    return np.real(np.vdot(phi, np.linalg.solve(D @ D.conj().T, phi)))


def test_pseudofermion_force_finite_difference(rng):
    th = rand_theta(rng, NT, NX, 0.7)
    n = LS * NT * NX * 2
    phi = rand_spinor(rng, n)
    S, F, eta, it, rr, rc = oracle.dwf_pf_action_force(th, phi, NT, NX, LS, M5, MF, 1e-13, 10000)
    assert rc == 0
    assert abs(S - _pf_action_dense(th, phi)) <= 1e-10 * abs(S)
    h = 1e-5
    for link in [0, 5, NT * NX - 1, NT * NX + 3, 2 * NT * NX - 1]:
        tp, tm = th.copy(), th.copy()
        tp[link] += h
        tm[link] -= h
        fd = (_pf_action_dense(tp, phi) - _pf_action_dense(tm, phi)) / (2 * h)
        assert abs(F[link] - fd) <= 1e-6 * max(1.0, abs(fd)), (link, F[link], fd)


def test_pseudofermion_force_not_converged(rng):
    th = rand_theta(rng, NT, NX)
    phi = rand_spinor(rng, LS * NT * NX * 2)
    *_, rc = oracle.dwf_pf_action_force(th, phi, NT, NX, LS, M5, MF, 1e-14, 2)
    assert rc == oracle.ORC_NOT_CONVERGED


def _gauge_action_direct(th, Nt, Nx, beta):
    t = th.reshape(2, Nt, Nx)
    P = t[0] + np.roll(t[1], -1, axis=0) - np.roll(t[0], -1, axis=1) - t[1]
    return beta * np.sum(1 - np.cos(P)), np.mean(np.cos(P))


def test_gauge_action_force_and_plaquette(rng):
    beta = 2.5
    th = rand_theta(rng, NT, NX, 1.0)
    S, F = oracle.gauge_action_force(th, NT, NX, beta)
    Sd, plaq = _gauge_action_direct(th, NT, NX, beta)
    assert abs(S - Sd) < 1e-12 * max(1, abs(S))
    assert abs(oracle.average_plaquette(th, NT, NX) - plaq) < 1e-14
    h = 1e-6
    for link in range(0, 2 * NT * NX, 7):
        tp, tm = th.copy(), th.copy()
        tp[link] += h
        tm[link] -= h
        fd = (_gauge_action_direct(tp, NT, NX, beta)[0] - _gauge_action_direct(tm, NT, NX, beta)[0]) / (2 * h)
        assert abs(F[link] - fd) < 1e-7
    # cold start: theta = const gives plaquette 1 and zero action and force
    S0, F0 = oracle.gauge_action_force(np.ones(2 * NT * NX), NT, NX, beta)
    assert S0 == 0.0 and not np.any(F0)
    assert oracle.average_plaquette(np.ones(2 * NT * NX), NT, NX) == 1.0


def test_mt19937_64_standard_value():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937_64
    assert oracle.mt19937_64_nth(5489, 10000) == 9981545732273789042


def test_hmc_delta_h_scales_as_epsilon_squared():
    Nt, Nx, Ls = 4, 4, 4
    th0 = np.ones(2 * Nt * Nx)
    dh = []
    for eps in (0.2, 0.1):
        steps = int(round(0.4 / eps))
        r = oracle.dwf_hmc(th0, Nt, Nx, Ls, 1.0, 0.2, 2.0, eps, steps, 1, 1234, 1e-12, 5000)
        dh.append(abs(r["delta_h"][0]))
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_hmc_runs_and_records(rng):
    Nt, Nx, Ls = 4, 4, 4
    r = oracle.dwf_hmc(np.ones(2 * Nt * Nx), Nt, Nx, Ls, 1.0, 0.2, 3.0, 0.05, 5, 3, 1234, 1e-10, 5000)
    assert r["accepted"].shape == (3,)
    assert np.all((r["acc_prob"] > 0) & (r["acc_prob"] <= 1))
    assert np.all(np.abs(r["delta_h"]) < 1.0)
    # deterministic for a seed
    r2 = oracle.dwf_hmc(np.ones(2 * Nt * Nx), Nt, Nx, Ls, 1.0, 0.2, 3.0, 0.05, 5, 3, 1234, 1e-10, 5000)
    assert np.array_equal(r["delta_h"], r2["delta_h"])


def test_hmc_leapfrog_reversible(rng):
    """The integrator is reversible: integrate, flip the momenta, integrate again -- the
    angles return and the momenta come back flipped, with H back to its start (tight
    solves, so the chronological guesses cost nothing measurable)."""
    Nt, Nx, Ls, M5, mf, beta, eps, steps, rtol = 8, 6, 4, 1.0, 0.1, 3.0, 0.05, 10, 1e-14
    th0 = rng.normal(0.0, 0.5, size=2 * Nt * Nx)
    p0 = rng.normal(size=2 * Nt * Nx)
    chi = (rng.normal(size=Ls * Nt * Nx * 2) + 1j * rng.normal(size=Ls * Nt * Nx * 2)) / np.sqrt(2)
    phi = oracle.dwf_apply(0, th0, chi, Nt, Nx, Ls, M5, mf)
    th1, p1, h0, h1 = oracle.dwf_leapfrog(th0, p0, phi, Nt, Nx, Ls, M5, mf, beta, eps, steps, rtol, 20000)
    assert np.max(np.abs(th1 - th0)) > 0.05  # it moved
    th2, p2, h1b, h2 = oracle.dwf_leapfrog(th1, -p1, phi, Nt, Nx, Ls, M5, mf, beta, eps, steps, rtol, 20000)
    assert abs(h1b - h1) <= 1e-10 * abs(h1)
    assert np.max(np.abs(th2 - th0)) <= 1e-11
    assert np.max(np.abs(p2 + p0)) <= 1e-11
    assert abs(h2 - h0) <= 1e-10 * abs(h0)


def test_hmc_trajectory_is_the_leapfrog():
    """orc_dwf_hmc's Delta H is the leapfrog's H_final - H_initial from the same draws
    (chi, then the momenta: generate_configs.cpp:56-69) and the cold start theta = 1."""
    Nt, Nx, Ls, M5, mf, beta, eps, steps, seed, rtol = 6, 6, 4, 1.0, 0.2, 3.0, 0.05, 5, 77, 1e-12
    th0 = np.ones(2 * Nt * Nx)
    r = oracle.dwf_hmc(th0, Nt, Nx, Ls, M5, mf, beta, eps, steps, 1, seed, rtol, 5000)
    chi, mom, u = oracle.dwf_heatbath_draws(Ls * Nt * Nx * 2, 2 * Nt * Nx, seed)
    phi = oracle.dwf_apply(0, th0, chi, Nt, Nx, Ls, M5, mf)
    th1, _, h0, h1 = oracle.dwf_leapfrog(th0, mom, phi, Nt, Nx, Ls, M5, mf, beta, eps, steps, rtol, 5000)
    assert abs((h1 - h0) - r["delta_h"][0]) <= 1e-12 * abs(h0)
    assert bool(r["accepted"][0]) == (u < min(1.0, np.exp(-(h1 - h0))))
    if r["accepted"][0]:
        assert np.array_equal(r["theta"], th1)


def test_cg_many_equals_single_solves(rng):
    """The OpenMP batch of independent solves (the CPU baseline of the batched analyzer)
    gives each solve exactly what a single call gives, on mixed lattice sizes."""
    Ls, M5, mf = 4, 1.0, 0.05
    lats = [(6, 6), (8, 6), (6, 6), (10, 8)]
    ths = [rng.normal(0.0, 0.5, size=2 * Nt * Nx) for Nt, Nx in lats]
    bs = [rng.normal(size=Ls * Nt * Nx * 2) + 1j * rng.normal(size=Ls * Nt * Nx * 2) for Nt, Nx in lats]
    xs, its, cv, threads = oracle.dwf_cg_many(ths, bs, lats, Ls, M5, mf, 1e-10, 5000)
    assert threads >= 1 and cv.all()
    for i, (Nt, Nx) in enumerate(lats):
        x1, it1, _, _, _ = oracle.dwf_cg(3, ths[i], bs[i], Nt, Nx, Ls, M5, mf, 1e-10, 5000)
        assert np.array_equal(xs[i], x1) and its[i] == it1


# ------------------------------------------------------------------- contractions
def _wall_or_point_source(Nt, Nx, Ls, t_src, ss, xs):
    src = np.zeros((Ls, Nt, Nx, 2), complex)
    for x in xs:
        src[0, t_src, x] += PL[:, ss]
        src[Ls - 1, t_src, x] += PR[:, ss]
    return src.reshape(-1)


def test_cpp_cj5_against_dense_inverse(rng):
    th = rand_theta(rng, NT, NX)
    Dinv = np.linalg.inv(dwf_dense(th, NT, NX, LS, M5, MF))
    cpp, cj5, nc, iters, dres = oracle.dwf_cpp_cj5(th, NT, NX, LS, M5, MF, 1e-13, 10000)
    assert nc == 0 and np.all(dres < 1e-10)
    want_pp, want_j5 = np.zeros(NT), np.zeros(NT)
    for t_src in (0, NT // 2):
        for ss in range(2):
            G = (Dinv @ _wall_or_point_source(NT, NX, LS, t_src, ss, range(NX))).reshape(LS, NT, NX, 2)
            q = G[0] @ PL.T + G[LS - 1] @ PR.T  # [Nt, Nx, 2]
            mid = G[LS // 2]
            for t in range(NT):
                dt = (t - t_src) % NT
                want_pp[dt] += np.sum(np.abs(q[t]) ** 2)
                want_j5[dt] += np.real(np.sum(np.conj(mid[t]) * q[t]))
    assert np.allclose(cpp, want_pp / 2, rtol=1e-9, atol=1e-13)
    assert np.allclose(cj5, want_j5 / 2, rtol=1e-8, atol=1e-12)


def test_momentum_cpp_against_dense_inverse(rng):
    th = rand_theta(rng, NT, NX)
    Dinv = np.linalg.inv(dwf_dense(th, NT, NX, LS, M5, MF))
    mm = NX // 2
    res = oracle.dwf_momentum_cpp(th, NT, NX, LS, M5, MF, mm, 1e-13, 10000)
    want = np.zeros((mm + 1, NT))
    for t_src in (0, NT // 2):
        local = np.zeros((NT, NX))
        for ss in range(2):
            G = (Dinv @ _wall_or_point_source(NT, NX, LS, t_src, ss, [0])).reshape(LS, NT, NX, 2)
            q = G[0] @ PL.T + G[LS - 1] @ PR.T
            local += np.sum(np.abs(q) ** 2, axis=-1)
        for n in range(mm + 1):
            proj = local @ np.cos(2 * np.pi * n * np.arange(NX) / NX)
            for t in range(NT):
                want[n, (t - t_src) % NT] += proj[t]
    assert np.allclose(res, want / 2, rtol=1e-9, atol=1e-13)
    with pytest.raises(oracle.OracleError):
        oracle.dwf_momentum_cpp(th, NT, NX, LS, M5, MF, NX // 2 + 1, 1e-8, 100)


def test_loop_density_against_dense_inverse(rng):
    th = rand_theta(rng, NT, NX)
    Dinv = np.linalg.inv(dwf_dense(th, NT, NX, LS, M5, MF))
    seed, nn = 918273, 3
    loop = oracle.dwf_loop_density(th, NT, NX, LS, M5, MF, seed, nn, 1e-13, 10000)
    noise = oracle.z4_noise(NT, NX, nn, seed)
    assert set(np.round(noise.reshape(-1), 12)) <= {1, 1j, -1, -1j}
    want = np.zeros((NT, NX), complex)
    for k in range(nn):
        eta = noise[k]
        src = np.zeros((LS, NT, NX, 2), complex)
        src[0] = eta @ PL.T
        src[LS - 1] += eta @ PR.T
        X = (Dinv @ src.reshape(-1)).reshape(LS, NT, NX, 2)
        q = X[0] @ PL.T + X[LS - 1] @ PR.T
        want += np.sum(np.conj(eta) * (q * np.array([1, -1])), axis=-1)
    assert np.allclose(loop.reshape(NT, NX), want / nn, rtol=1e-9, atol=1e-12)
    with pytest.raises(oracle.OracleError):
        oracle.dwf_loop_density(th, NT, NX, LS, M5, MF, seed, 0, 1e-8, 100)


def test_z4_noise_statistics():
    noise = oracle.z4_noise(8, 8, 64, 918273).reshape(-1)
    counts = [np.sum(np.isclose(noise, v)) for v in (1, 1j, -1, -1j)]
    assert sum(counts) == noise.size
    assert all(abs(c / noise.size - 0.25) < 0.02 for c in counts)
