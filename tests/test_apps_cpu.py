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
