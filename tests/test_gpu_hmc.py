"""GPU parity of the pseudofermion force, gauge action/force, plaquette and full HMC
trajectories (through the C ABI) against the oracle.  Bars (north star): force <= 1e-10
relative for a given eta, Delta H <= 1e-8 absolute on same-seed trajectories."""

import os

import numpy as np
import pytest

import oracle
from conftest import print_case
from dense_ref import rand_spinor, rand_theta

import three_part_dwf_study_amd as dwf
from three_part_dwf_study_amd import capi

pytestmark = pytest.mark.gpu

M5, MF = 1.0, 0.05


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("Nt,Nx,Ls", [(16, 16, 8), (6, 10, 4), (96, 96, 16)])
def test_dwf_force_from_eta(Nt, Nx, Ls):
    print_case(test="dwf_force_from_eta", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(Nt + Ls)
    th = rand_theta(rng, Nt, Nx, 0.7)
    eta = rand_spinor(rng, Ls * Nt * Nx * 2)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        got = ctx.force_from_eta(eta)
    want = oracle.dwf_force_from_eta(th, eta, Nt, Nx, Ls, M5, MF)
    assert rel(got, want) <= 1e-10, rel(got, want)


@pytest.mark.parametrize("bct,bcx", [(-1.0, 1.0), (1.0, -1.0)])
def test_wilson_force_from_eta(bct, bcx):
    Nt, Nx, m0 = 8, 12, 0.15
    print_case(test="wilson_force_from_eta", bct=bct, bcx=bcx)
    rng = np.random.default_rng(3)
    th = rng.uniform(-np.pi, np.pi, 2 * Nt * Nx)
    eta = rand_spinor(rng, Nt * Nx * 2)
    with dwf.DwfContext.wilson(Nt, Nx, m0, bct, bcx) as ctx:
        ctx.set_gauge(th)
        got = ctx.force_from_eta(eta)
    want = oracle.wilson_force_from_eta(th, eta, Nt, Nx, m0, bct, bcx)
    assert rel(got, want) <= 1e-10


def test_pf_action_force_and_gauge():
    Nt, Nx, Ls, beta = 16, 12, 8, 3.0
    print_case(test="pf_action_force")
    rng = np.random.default_rng(21)
    th = rand_theta(rng, Nt, Nx)
    phi = rand_spinor(rng, Ls * Nt * Nx * 2)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        S, F, eta, st, rc = ctx.pf_action_force(phi, 1e-12, 5000)
        So, Fo, etao, *_ = oracle.dwf_pf_action_force(th, phi, Nt, Nx, Ls, M5, MF, 1e-12, 5000)
        assert abs(S - So) <= 1e-10 * abs(So)
        assert rel(F, Fo) <= 1e-9
        Sm, Fm, *_ = ctx.pf_action_force(phi, 1e-12, 5000, method=dwf.DWF_SOLVE_MIXED)
        assert abs(Sm - So) <= 1e-10 * abs(So) and rel(Fm, Fo) <= 1e-9
        Sg, Fg = ctx.gauge_action_force(beta)
        Sgo, Fgo = oracle.gauge_action_force(th, Nt, Nx, beta)
        assert abs(Sg - Sgo) <= 1e-12 * abs(Sgo) and rel(Fg, Fgo) <= 1e-12
        assert abs(ctx.average_plaquette() - oracle.average_plaquette(th, Nt, Nx)) < 1e-14
        # non-convergence surfaces the reference's message
        _, _, _, st, rc = ctx.pf_action_force(phi, 1e-14, 2, raise_on_fail=False)
        assert rc == capi.DWF_ERR_NOT_CONVERGED
        assert b"Pseudofermion CG failed to converge" in ctx._lib.dwf_last_error()


def _gpu_dwf_hmc(Nt, Nx, Ls, beta, eps, steps, ntraj, seed, rtol, method=dwf.DWF_SOLVE_CG64):
    r = capi.HostRng(seed)
    out = {"delta_h": [], "plaquette": [], "accepted": []}
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(np.ones(2 * Nt * Nx))  # cold start theta = 1 (generate_configs.cpp:27-28)
        for _ in range(ntraj):
            chi, mom = r.dwf_heatbath(Ls * Nt * Nx * 2, 2 * Nt * Nx)
            res = ctx.hmc_trajectory(chi, mom, r.uniform(), beta, eps, steps, rtol, 10000, method=method)
            for k in out:
                out[k].append(res[k])
        out["theta"] = ctx.get_gauge()
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.parametrize("Nt,Nx,Ls", [(8, 8, 4), (16, 12, 8)])
def test_dwf_hmc_same_seed_history(Nt, Nx, Ls):
    beta, eps, steps, ntraj, seed, rtol = 3.0, 0.05, 8, 4, 1234, 1e-12
    print_case(test="dwf_hmc_history", Nt=Nt, Nx=Nx, Ls=Ls)
    g = _gpu_dwf_hmc(Nt, Nx, Ls, beta, eps, steps, ntraj, seed, rtol)
    o = oracle.dwf_hmc(np.ones(2 * Nt * Nx), Nt, Nx, Ls, M5, MF, beta, eps, steps, ntraj, seed, rtol, 10000)
    assert np.max(np.abs(g["delta_h"] - o["delta_h"])) <= 1e-8
    assert np.array_equal(g["accepted"], o["accepted"])
    assert np.max(np.abs(g["plaquette"] - o["plaquette"])) <= 1e-10
    assert np.max(np.abs(g["theta"] - o["theta"])) <= 1e-9


HMC96 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dwf_hmc_96x96x16.npz")


@pytest.mark.parametrize("method", ["cg64", "mixed", "eo_mixed", "learned"])
def test_dwf_hmc_96_matches_oracle_fixture(method):
    """cfg5 at its real size against the oracle: two trajectories at 96x96, Ls=16, beta=4,
    eps=0.05, 20 leapfrog steps from the reference's cold start theta = 1, seed 1234,
    force solves to 1e-12 (tests/golden/make_dwf_hmc96.py wrote the oracle's history).
    The GPU path draws the same momenta and pseudofermions; CG64, MIXED, EO_MIXED (two
    even-odd Dirac solves per force solve) and learned force solves (a network trained
    briefly here; every MIXED, EO_MIXED and learned solve is certified on its FP64 true
    residual) must reproduce the oracle's Delta H, accept/reject decisions, plaquettes
    and final angles.  Delta H: within 1e-8 for CG64 (the oracle's algorithm, differing in
    summation order only); the others stop on their FP64 true residual at the same 1e-12
    but at a different iterate, so their pseudofermion actions (about 3e5 here) may differ
    from the oracle's by ~rtol |S_f| kappa: within 1e-6 (measured: CG64 3e-9, MIXED 3e-8,
    learned 3e-7)."""
    f = np.load(HMC96)
    Nt, Nx, Ls, m5, mf, beta, eps, steps, ntraj, seed, rtol = f["params"]
    Nt, Nx, Ls, steps, ntraj, seed = int(Nt), int(Nx), int(Ls), int(steps), int(ntraj), int(seed)
    assert (m5, mf) == (M5, MF)
    print_case(test="dwf_hmc96", method=method)
    kind = {"cg64": dwf.DWF_SOLVE_CG64, "mixed": dwf.DWF_SOLVE_MIXED, "eo_mixed": dwf.DWF_SOLVE_EO_MIXED,
            "learned": dwf.DWF_SOLVE_LEARNED}[method]
    net = None
    if method == "learned":
        net = dwf.Network(Ls, seed=7)
        rng = np.random.default_rng(4242)
        ths = np.stack([1.0 + rng.normal(0.0, 0.3, size=2 * Nt * Nx) for _ in range(4)])
        with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
            ctx.set_gauge_batch(ths)
            src = rng.normal(size=(16, Ls * Nt * Nx * 2)) + 1j * rng.normal(size=(16, Ls * Nt * Nx * 2))
            b = ctx.apply(dwf.DWF_OP_DDAG, src)
            net.train(ctx, b, None, pair_cfg=np.repeat(np.arange(4), 4), steps=200, batch=8, loss=1, lr=3e-3, seed=5)
    r = capi.HostRng(seed)
    dh, acc, plaq = [], [], []
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(np.ones(2 * Nt * Nx))
        if net is not None:
            ctx.set_preconditioner(net)
        for _ in range(ntraj):
            chi, mom = r.dwf_heatbath(Ls * Nt * Nx * 2, 2 * Nt * Nx)
            res = ctx.hmc_trajectory(chi, mom, r.uniform(), float(beta), float(eps), steps, float(rtol), 20000,
                                     method=kind)
            dh.append(res["delta_h"])
            acc.append(res["accepted"])
            plaq.append(res["plaquette"])
        th = ctx.get_gauge()
    dh = np.array(dh)
    print(f"96x96x16 {method}: dH {dh.tolist()} vs oracle {f['delta_h'].tolist()}, "
          f"|diff| {np.abs(dh - f['delta_h']).tolist()}", flush=True)
    assert np.max(np.abs(dh - f["delta_h"])) <= (1e-8 if method == "cg64" else 1e-6)
    assert np.array_equal(np.array(acc), f["accepted"])
    assert np.max(np.abs(np.array(plaq) - f["plaquette"])) <= 1e-10
    assert np.max(np.abs(th - f["theta"])) <= 1e-9


def test_dwf_hmc_mixed_solver_matches_fp64():
    Nt, Nx, Ls = 16, 12, 8
    print_case(test="dwf_hmc_mixed")
    a = _gpu_dwf_hmc(Nt, Nx, Ls, 3.0, 0.05, 8, 3, 99, 1e-11)
    b = _gpu_dwf_hmc(Nt, Nx, Ls, 3.0, 0.05, 8, 3, 99, 1e-11, method=dwf.DWF_SOLVE_MIXED)
    assert np.max(np.abs(a["delta_h"] - b["delta_h"])) <= 1e-8
    assert np.array_equal(a["accepted"], b["accepted"])


@pytest.mark.parametrize("start", ["cold", "hot"])
def test_wilson_hmc_same_seed_history(start):
    """cfg1: 16x16 Wilson HMC, beta = 2, m0 = 0.1, tau = 1 in 10 steps.  The reference starts
    cold (theta = 0, wilson/generate_configs.cpp:20); the benchmark's hot start draws theta
    uniform on [-pi, pi) with seed 1234."""
    Nt, Nx, m0, beta, tau, steps, ntraj, seed, tol = 16, 16, 0.1, 2.0, 1.0, 10, 5, 1234, 1e-12
    print_case(test="wilson_hmc_history", start=start)
    th0 = (np.zeros(2 * Nt * Nx) if start == "cold"
           else np.random.default_rng(1234).uniform(-np.pi, np.pi, size=2 * Nt * Nx))
    o = oracle.wilson_hmc(th0, Nt, Nx, m0, -1.0, 1.0, beta, tau, steps, ntraj, seed, tol, 5000)
    r = capi.HostRng(seed)
    dh, acc, plaq = [], [], []
    with dwf.DwfContext.wilson(Nt, Nx, m0, -1.0, 1.0) as ctx:
        ctx.set_gauge(th0)
        for _ in range(ntraj):
            mom, chi = r.wilson_heatbath(2 * Nt * Nx, Nt * Nx * 2)
            res = ctx.hmc_trajectory(chi, mom, r.uniform(), beta, tau / steps, steps, tol, 5000,
                                     integrator=capi.DWF_HMC_WILSON)
            dh.append(res["delta_h"])
            acc.append(res["accepted"])
            plaq.append(res["plaquette"])
        th = ctx.get_gauge()
    assert np.max(np.abs(np.array(dh) - o["delta_h"])) <= 1e-8
    assert np.array_equal(np.array(acc), o["accepted"])
    assert np.max(np.abs(np.array(plaq) - o["plaquette"])) <= 1e-10
    assert np.max(np.abs(th - o["theta"])) <= 1e-9
