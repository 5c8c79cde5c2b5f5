"""End-to-end runs of the generate/analyze programs on the GPU against the oracle:
same-seed HMC histories and configurations, and the measured correlators, with the
reference's output files and CSV layouts."""

import os

import numpy as np
import pytest

import oracle
from apps_util import BIN, DEFAULTS, read_csv, run, write_parameters
from conftest import print_case

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "generate_domain_wall")), reason="apps not built")]


def close(a, b):  # CSV values carry 6 significant digits
    return np.allclose(a, b, rtol=2e-5, atol=1e-9)


def test_domain_wall_generate_and_analyze(tmp_path):
    print_case(test="apps_dwf")
    write_parameters(tmp_path)
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 0, err
    d = tmp_path / "output" / "domain_wall"
    P = DEFAULTS
    Nt, Nx, Ls, T = P["dwf.Nt"], P["dwf.Nx"], P["dwf.N5"], P["dwf.trajectories"]
    configs = np.load(d / "configs.npy")
    ana = np.load(d / "analysis_configs.npy")
    assert configs.shape == (T, 2, Nt, Nx) and ana.shape == (T - 1, 2, Nt, Nx)
    o = oracle.dwf_hmc(np.ones(2 * Nt * Nx), Nt, Nx, Ls, P["dwf.M5"], P["dwf.mf"], P["dwf.beta"], P["dwf.epsilon"],
                       P["dwf.leapfrog_steps"], T, P["dwf.random_seed"], P["dwf.pseudofermion_rtol"],
                       P["dwf.pseudofermion_maxiter"])
    assert np.max(np.abs(configs[-1].reshape(-1) - o["theta"])) <= 1e-9
    hdr, h = read_csv(d / "run_info_history.csv")
    assert hdr == ["trajectory", "accepted", "acceptance_probability", "delta_h", "plaquette"]
    assert np.array_equal(h[:, 1], o["accepted"]) and close(h[:, 3], o["delta_h"]) and close(h[:, 4], o["plaquette"])

    rc, out, err = run("analyze_domain_wall", tmp_path)
    assert rc == 0, err
    hdr, c = read_csv(d / "correlator.csv")
    assert hdr == ["t", "CPP_average", "CJ5_average", "m_res", "disconnected", "eta_singlet"] and c.shape == (Nt, 6)
    hdr, bc = read_csv(d / "channels_by_config.csv")
    assert hdr == (["config"] + [f"pion_t{t}" for t in range(Nt)] + [f"disconnected_t{t}" for t in range(Nt)]
                   + [f"eta_t{t}" for t in range(Nt)])
    assert bc.shape == (T - 1, 1 + 3 * Nt)
    hdr, mc = read_csv(d / "momentum_correlators.csv")
    assert hdr == ["momentum_index", "p", "p_squared", "t", "pion", "disconnected", "eta"]
    assert mc.shape == ((P["analysis.max_momentum"] + 1) * Nt, 7)
    # per-configuration measurements from the oracle, then the ensemble contractions of
    # domain_wall/analyze_configs.cpp:78-271 restated here in numpy
    M5, mf, mm, nn = P["dwf.M5"], P["dwf.mf"], P["analysis.max_momentum"], P["eta.noise_vectors"]
    cpp_all, cj5_all, loops, moms = [], [], [], []
    for k in range(T - 1):
        th = ana[k].reshape(-1)
        cpp, cj5, *_ = oracle.dwf_cpp_cj5(th, Nt, Nx, Ls, M5, mf, 1e-12, 5000)
        assert close(bc[k, 1:1 + Nt], cpp)
        cpp_all.append(cpp)
        cj5_all.append(cj5)
        loops.append(oracle.dwf_loop_density(th, Nt, Nx, Ls, M5, mf, P["eta.random_seed"] + 104729 * k, nn, 1e-12,
                                             5000).reshape(Nt, Nx))
        moms.append(oracle.dwf_momentum_cpp(th, Nt, Nx, Ls, M5, mf, mm, 1e-12, 5000))
    ncfg = T - 1
    cpp_avg, cj5_avg = np.mean(cpp_all, axis=0), np.mean(cj5_all, axis=0)
    assert close(c[:, 0], np.arange(Nt))
    assert close(c[:, 1], cpp_avg) and close(c[:, 2], cj5_avg)
    assert close(c[:, 3], cj5_avg / cpp_avg)  # m_res(t)
    loops = np.array(loops)
    dev = (loops - loops.mean(axis=0)).sum(axis=2)  # [config][t]: sum_x (L - <L>)
    disc = np.array([[np.mean([np.real(dev[k, (t0 + dt) % Nt] * np.conj(dev[k, t0])) for t0 in range(Nt)])
                      for dt in range(Nt)] for k in range(ncfg)])
    eta = np.array([moms[k][0] - P["eta.flavors"] * disc[k] for k in range(ncfg)])
    scale = np.max(np.abs(disc))
    assert np.allclose(c[:, 4], disc.mean(axis=0), rtol=2e-5, atol=2e-5 * scale)  # disconnected
    assert np.allclose(c[:, 5], eta.mean(axis=0), rtol=2e-5, atol=2e-5 * scale)   # eta_singlet
    assert np.allclose(bc[:, 1 + Nt:1 + 2 * Nt], disc, rtol=2e-5, atol=2e-5 * scale)
    assert np.allclose(bc[:, 1 + 2 * Nt:], eta, rtol=2e-5, atol=2e-5 * scale)
    # momentum-projected channels (analyze_configs.cpp:172-227)
    for n in range(mm + 1):
        p = 2 * np.pi * n / Nx
        lp = (loops * np.exp(-1j * p * np.arange(Nx))[None, None, :]).sum(axis=2)  # [config][t]
        d = lp - lp.mean(axis=0)
        dis = np.array([np.mean([np.mean([np.real(d[k, (t0 + dt) % Nt] * np.conj(d[k, t0])) for t0 in range(Nt)])
                                 for k in range(ncfg)]) for dt in range(Nt)])
        pion = np.mean([moms[k][n] for k in range(ncfg)], axis=0)
        rows = mc[n * Nt:(n + 1) * Nt]
        assert np.array_equal(rows[:, 0], np.full(Nt, n)) and close(rows[:, 1], np.full(Nt, p))
        assert close(rows[:, 2], np.full(Nt, p * p)) and np.array_equal(rows[:, 3], np.arange(Nt))
        sc = max(np.max(np.abs(dis)), 1e-300)
        assert close(rows[:, 4], pion), n
        assert np.allclose(rows[:, 5], dis, rtol=2e-5, atol=2e-5 * sc), n
        assert np.allclose(rows[:, 6], pion - P["eta.flavors"] * dis, rtol=2e-5, atol=2e-5 * max(sc, 1e-12)), n


def test_wilson_generate_and_analyze(tmp_path):
    print_case(test="apps_wilson")
    # Nt = 16 leaves room for a plateau window (t = 2 .. 8); any drift is accepted
    over = {"wilson.Nt": 16, "analysis.plateau_flatness_tolerance": 10.0}
    write_parameters(tmp_path, **{k.replace(".", "__"): v for k, v in over.items()})
    rc, out, err = run("generate_wilson", tmp_path)
    assert rc == 0, err
    P = dict(DEFAULTS, **over)
    Nt, Nx, T = P["wilson.Nt"], P["wilson.Nx"], P["wilson.trajectories"]
    o = oracle.wilson_hmc(np.zeros(2 * Nt * Nx), Nt, Nx, P["wilson.m0"], -1.0, 1.0, P["wilson.beta"],
                          P["wilson.trajectory_length"], P["wilson.hmc_steps"], T, P["wilson.random_seed"], 1e-12, 5000)
    hdr, h = read_csv(tmp_path / "output" / "wilson" / "hmc_history.csv")
    assert hdr == ["trajectory", "accepted", "dH", "acceptance_probability", "plaquette", "saved"]
    assert np.array_equal(h[:, 1], o["accepted"]) and close(h[:, 2], o["delta_h"])
    raw = open(tmp_path / "output" / "wilson" / "wilson_ensemble.bin", "rb").read()
    assert raw[:8] == b"WILCFG1\0" and len(raw) == 40 + (T - 1) * 2 * Nt * Nx * 8
    last = np.frombuffer(raw[-2 * Nt * Nx * 8:], dtype=np.float64)
    assert np.max(np.abs(last - o["theta"])) <= 1e-9
    rc, out, err = run("analyze_wilson", tmp_path)
    assert rc == 0, err
    hdr, cc = read_csv(tmp_path / "output" / "wilson" / "correlators_by_config.csv")
    assert hdr == ["config"] + [f"t{t}" for t in range(Nt)] and cc.shape == (T - 1, Nt + 1)
    C = oracle.wilson_pseudoscalar_correlator(last, Nt, Nx, P["wilson.m0"], -1.0, 1.0, 0, 0, 1e-12, 5000)
    assert close(cc[-1, 1:], C)
    # jackknife effective mass, plateau and mass (wilson/analyze_configs.cpp:36-79)
    line = [ln for ln in out.splitlines() if ln.startswith("Wilson pseudoscalar mass = ")]
    plat = [ln for ln in out.splitlines() if ln.startswith("Selected plateau: t=")]
    assert len(line) == 1 and len(plat) == 1, out
    mass, err_m = (float(v) for v in line[0].split("=")[1].split("+/-"))
    a, b = (int(v) for v in plat[0].split("t=")[1].split(" through "))
    hdr, em = read_csv(tmp_path / "output" / "wilson" / "effective_mass.csv")
    assert hdr == ["t", "C_average", "C_error", "m_eff", "m_eff_error"] and em.shape == (Nt, 5)
    assert close(em[:, 1], cc[:, 1:].mean(axis=0))
    # the jackknife mass is the window mean of the jackknife effective masses (linear)
    assert abs(mass - em[a:b + 1, 3].mean()) <= 1e-6 * mass and err_m > 0
    # and the jackknife mean of m_eff is close to m_eff of the mean correlator
    cbar = em[:, 1]
    direct = np.arccosh((cbar[a - 1:b] + cbar[a + 1:b + 2]) / (2 * cbar[a:b + 1]))
    assert np.all(np.abs(em[a:b + 1, 3] - direct) <= 3 * em[a:b + 1, 4] + 1e-6)


def test_domain_wall_learned_solvers(tmp_path):
    """The drop-in programs with the learned preconditioner: train_domain_wall_network on
    the generator's own ensemble, then the HMC with gpu.force_solver = learned and the
    analyzer with gpu.propagator_solver = learned.  Every solve is certified at the FP64
    true residual, so the accept/reject history, Delta H and the correlators equal those of
    plain FP64 CG (to the CSV files' 6 digits)."""
    print_case(test="apps_learned")
    d = tmp_path / "output" / "domain_wall"
    write_parameters(tmp_path)
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 0, err
    _, h64 = read_csv(d / "run_info_history.csv")
    rc, out, err = run("analyze_domain_wall", tmp_path)
    assert rc == 0, err
    _, c64 = read_csv(d / "correlator.csv")
    _, m64 = read_csv(d / "momentum_correlators.csv")
    net = {"gpu__network_weights": "net.npy", "gpu__train_steps": 200, "gpu__train_sources": 4, "gpu__train_batch": 4}
    write_parameters(tmp_path, **net)
    rc, out, err = run("train_domain_wall_network", tmp_path)
    assert rc == 0, err
    w = np.load(tmp_path / "net.npy")
    assert w.dtype == np.float64 and w.ndim == 1 and np.all(np.isfinite(w))
    write_parameters(tmp_path, gpu__force_solver="learned", **net)
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 0, err
    _, hl = read_csv(d / "run_info_history.csv")
    assert np.array_equal(hl[:, 1], h64[:, 1]) and close(hl[:, 3], h64[:, 3]) and close(hl[:, 4], h64[:, 4])
    write_parameters(tmp_path, gpu__propagator_solver="learned", **net)
    rc, out, err = run("analyze_domain_wall", tmp_path)
    assert rc == 0, err
    assert "right-hand sides per batched solve" in out
    _, cl = read_csv(d / "correlator.csv")
    _, ml = read_csv(d / "momentum_correlators.csv")
    assert close(cl[:, :4], c64[:, :4])
    scale = np.max(np.abs(c64[:, 4:]))
    assert np.allclose(cl[:, 4:], c64[:, 4:], rtol=2e-5, atol=2e-5 * scale)
    assert close(ml[:, :5], m64[:, :5])
