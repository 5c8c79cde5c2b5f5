"""The generate/analyze programs on the CPU: parameter and file errors carry the
reference's messages (common/config.hpp:33-46, 93-102; analyze_configs.cpp:24-31;
npy_io.hpp:155; config_io.hpp:41-49) and, past them, the missing GPU fails loudly."""

import os

import numpy as np
import pytest

from apps_util import BIN, run, write_parameters

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "generate_domain_wall")),
                                reason="apps not built")


def test_missing_parameter_file(tmp_path):
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 1 and "error: Cannot open parameter file: parameters.txt" in err


def test_missing_and_invalid_keys(tmp_path):
    write_parameters(tmp_path, drop=("dwf.N5",))
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 1 and "Missing required parameter: dwf.N5" in err
    write_parameters(tmp_path, dwf__Nt="8x")
    rc, out, err = run("analyze_domain_wall", tmp_path)
    assert rc == 1 and "Invalid numeric value for dwf.Nt: 8x" in err
    write_parameters(tmp_path, drop=("wilson.random_seed",))
    rc, out, err = run("generate_wilson", tmp_path)
    assert rc == 1 and "Missing required parameter: wilson.random_seed" in err
    write_parameters(tmp_path, gpu__force_solver="fast")
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 1 and "Invalid value for gpu.force_solver" in err


def test_file_format_errors(tmp_path):
    write_parameters(tmp_path)
    os.makedirs(tmp_path / "output" / "domain_wall")
    np.save(tmp_path / "output" / "domain_wall" / "analysis_configs.npy", np.zeros((3, 2, 8, 6)))
    rc, out, err = run("analyze_domain_wall", tmp_path)
    assert rc == 1 and "analysis_configs.npy has the wrong shape" in err
    np.save(tmp_path / "output" / "domain_wall" / "analysis_configs.npy", np.zeros((3, 2, 8, 8)))
    with open(tmp_path / "output" / "domain_wall" / "analysis_configs.npy", "r+b") as f:
        f.truncate(200)
    rc, out, err = run("analyze_domain_wall", tmp_path)
    assert rc == 1 and "Truncated NPY data" in err
    os.makedirs(tmp_path / "output" / "wilson")
    with open(tmp_path / "output" / "wilson" / "wilson_ensemble.bin", "wb") as f:
        f.write(b"NOTWILS\0" + bytes(32))
    rc, out, err = run("analyze_wilson", tmp_path)
    assert rc == 1 and "Wrong ensemble file format" in err


def test_no_gpu_fails_loudly(tmp_path):
    import ctypes

    import three_part_dwf_study_amd as dwf

    lib = dwf.load_library()
    h = ctypes.c_void_p()
    if lib.dwf_ctx_create_dwf(ctypes.byref(h), 4, 4, 4, 1.0, 0.1) == 0:
        lib.dwf_ctx_destroy(h)
        pytest.skip("a GPU is present")
    write_parameters(tmp_path)
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 1 and "no HIP device available" in err


# ------------------------------------------------ Wilson analyzer statistics (host only)
def _write_correlators(d, corrs):
    os.makedirs(d / "output" / "wilson", exist_ok=True)
    Nt = corrs.shape[1]
    with open(d / "output" / "wilson" / "correlators_by_config.csv", "w") as f:
        f.write("config," + ",".join(f"t{t}" for t in range(Nt)) + "\n")
        for k, c in enumerate(corrs):
            f.write(f"{k}," + ",".join(repr(float(v)) for v in c) + "\n")


def _mass_line(out):
    line = [ln for ln in out.splitlines() if ln.startswith("Wilson pseudoscalar mass = ")][0]
    m, e = line.split("=")[1].split("+/-")
    plat = [ln for ln in out.splitlines() if ln.startswith("Selected plateau: t=")][0]
    a, b = plat.split("t=")[1].split(" through ")
    return float(m), float(e), int(a), int(b)


def test_wilson_statistics_pure_cosh(tmp_path):
    """Closed form: correlators A_k cosh(m (t - Nt/2)) give m_eff(t) = m at every interior t
    for every leave-one-out mean (they differ only by amplitude), so the jackknife mass is
    m with zero error, and the first window (t = earliest .. earliest + window - 1) is flat."""
    Nt, m = 16, 0.37
    t = np.arange(Nt)
    corrs = np.array([A * np.cosh(m * (t - Nt / 2)) for A in (1.0, 1.3, 0.8, 2.1)])
    write_parameters(tmp_path)
    _write_correlators(tmp_path, corrs)
    rc, out, err = run("wilson_statistics", tmp_path)
    assert rc == 0, err
    mass, error, a, b = _mass_line(out)
    assert abs(mass - m) <= 1e-7 and error <= 1e-7 and (a, b) == (2, 5)
    with open(tmp_path / "output" / "wilson" / "effective_mass.csv") as f:
        assert f.readline().strip() == "t,C_average,C_error,m_eff,m_eff_error"
        rows = [ln.strip().split(",") for ln in f]
    meff = np.array([float(r[3]) for r in rows])
    assert np.isnan(meff[0]) and np.isnan(meff[-1])
    assert np.allclose(meff[1:-1], m, rtol=0, atol=1e-9)
    assert np.allclose([float(r[1]) for r in rows], corrs.mean(axis=0), rtol=1e-10)


def test_wilson_statistics_plateau_selection(tmp_path):
    """An excited state (mass M >> m) lifts the early effective mass: the selected window is
    the first one, from earliest_plateau_time, whose (max - min) / mean <= the tolerance,
    computed here from m_eff(t) = acosh((C(t-1) + C(t+1)) / 2 C(t)) of the exact
    correlator (identical configurations up to amplitude, so jackknife = exact)."""
    Nt, m, M, tol = 24, 0.3, 1.4, 0.02
    t = np.arange(Nt)
    C = np.cosh(m * (t - Nt / 2)) + 40.0 * np.cosh(M * (t - Nt / 2)) / np.cosh(M * Nt / 2)
    meff = np.full(Nt, np.nan)
    meff[1:-1] = np.arccosh((C[:-2] + C[2:]) / (2 * C[1:-1]))
    want = None
    for a in range(2, Nt // 2 - 3 + 1):
        w = meff[a:a + 4]
        if (w.max() - w.min()) / w.mean() <= tol:
            want = (a, a + 3, w.mean())
            break
    assert want is not None and want[0] > 2  # the early windows are rejected
    write_parameters(tmp_path, analysis__plateau_flatness_tolerance=tol)
    _write_correlators(tmp_path, np.array([C, 2.0 * C, 0.5 * C]))
    rc, out, err = run("wilson_statistics", tmp_path)
    assert rc == 0, err
    mass, error, a, b = _mass_line(out)
    assert (a, b) == want[:2] and abs(mass - want[2]) <= 1e-7 * want[2] and error <= 1e-7
    assert "rejected" in out and "ACCEPTED" in out


def test_wilson_statistics_errors(tmp_path):
    Nt = 16
    t = np.arange(Nt)
    write_parameters(tmp_path)
    _write_correlators(tmp_path, np.array([np.cosh(0.3 * (t - 8))]))
    rc, out, err = run("wilson_statistics", tmp_path)
    assert rc == 1 and "error: Need at least two configurations" in err
    # no flat window: a correlator whose effective mass alternates
    bad = np.cosh(0.3 * (t - 8)) * (1.0 + 0.5 * (t % 2))
    _write_correlators(tmp_path, np.array([bad, bad]))
    rc, out, err = run("wilson_statistics", tmp_path)
    assert rc == 1 and "No contiguous effective-mass window passed the flatness test" in err


@pytest.mark.parametrize("prog", ["generate_domain_wall", "generate_wilson"])
def test_force_solver_takes_eo_mixed(tmp_path, prog):
    """gpu.force_solver = eo_mixed passes the parameter checks (the run then stops at the
    first device call on a machine without a GPU, and not on the key)."""
    import ctypes

    import three_part_dwf_study_amd as dwf

    lib = dwf.load_library()
    h = ctypes.c_void_p()
    if lib.dwf_ctx_create_dwf(ctypes.byref(h), 4, 4, 4, 1.0, 0.1) == 0:
        lib.dwf_ctx_destroy(h)
        pytest.skip("a GPU is present")
    write_parameters(tmp_path, gpu__force_solver="eo_mixed")
    rc, out, err = run(prog, tmp_path)
    assert rc == 1 and "no HIP device available" in err, err


def test_learned_solver_needs_weights(tmp_path):
    write_parameters(tmp_path, gpu__force_solver="learned")
    rc, out, err = run("generate_domain_wall", tmp_path)
    assert rc == 1 and "gpu.force_solver = learned needs gpu.network_weights" in err
    write_parameters(tmp_path, gpu__propagator_solver="learned")
    rc, out, err = run("analyze_domain_wall", tmp_path)
    assert rc == 1 and "gpu.propagator_solver = learned needs gpu.network_weights" in err
    write_parameters(tmp_path)
    rc, out, err = run("train_domain_wall_network", tmp_path)
    assert rc == 1 and "Missing required parameter: gpu.network_weights" in err
