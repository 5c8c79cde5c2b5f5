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
    cpp_all = []
    for k in range(T - 1):
        th = ana[k].reshape(-1)
        cpp, cj5, *_ = oracle.dwf_cpp_cj5(th, Nt, Nx, Ls, P["dwf.M5"], P["dwf.mf"], 1e-12, 5000)
        assert close(bc[k, 1:1 + Nt], cpp)
        cpp_all.append((cpp, cj5))
    assert close(c[:, 1], np.mean([a for a, _ in cpp_all], axis=0))
    assert close(c[:, 2], np.mean([b for _, b in cpp_all], axis=0))
    mom0 = oracle.dwf_momentum_cpp(ana[0].reshape(-1), Nt, Nx, Ls, P["dwf.M5"], P["dwf.mf"], P["analysis.max_momentum"],
                                   1e-12, 5000)
    assert np.all(np.isfinite(mc)) and mom0.shape == (P["analysis.max_momentum"] + 1, Nt)


def test_wilson_generate_and_analyze(tmp_path):
    print_case(test="apps_wilson")
    write_parameters(tmp_path)
    rc, out, err = run("generate_wilson", tmp_path)
    assert rc == 0, err
    P = DEFAULTS
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
