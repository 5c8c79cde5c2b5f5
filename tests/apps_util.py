"""Helpers for running the generate/analyze programs in a scratch directory."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "three_part_dwf_study_amd", "bin")

DEFAULTS = {
    "wilson.Nt": 8, "wilson.Nx": 8, "wilson.beta": 2.0, "wilson.m0": 0.1, "wilson.trajectories": 4,
    "wilson.thermalization_cut": 1, "wilson.thin": 1, "wilson.hmc_steps": 10, "wilson.trajectory_length": 1.0,
    "wilson.source_t": 0, "wilson.source_x": 0, "wilson.time_boundary_sign": -1.0, "wilson.space_boundary_sign": 1.0,
    "wilson.force_cg_tolerance": 1e-12, "wilson.force_cg_max_iterations": 5000,
    "wilson.measurement_cg_tolerance": 1e-12, "wilson.measurement_cg_max_iterations": 5000, "wilson.random_seed": 1234,
    "dwf.Nt": 8, "dwf.Nx": 8, "dwf.N5": 4, "dwf.beta": 3.0, "dwf.M5": 1.0, "dwf.mf": 0.2, "dwf.epsilon": 0.05,
    "dwf.leapfrog_steps": 6, "dwf.trajectories": 4, "dwf.thermalization_cut": 1, "dwf.thin": 1,
    "dwf.pseudofermion_rtol": 1e-12, "dwf.pseudofermion_maxiter": 5000, "dwf.propagator_rtol": 1e-12,
    "dwf.propagator_maxiter": 5000, "dwf.random_seed": 1234, "eta.noise_vectors": 2, "eta.random_seed": 918273,
    "eta.flavors": 2, "analysis.threads": 0, "analysis.plateau_window_size": 4,
    "analysis.plateau_flatness_tolerance": 0.08, "analysis.earliest_plateau_time": 2, "analysis.fit_t_min": 3,
    "analysis.fit_t_max": 11, "analysis.gevp_t0": 2, "analysis.max_momentum": 2,
    "analysis.bootstrap_resamples": 2000, "analysis.bootstrap_block_size": 2, "analysis.bootstrap_seed": 24681357,
}


def write_parameters(d, drop=(), **over):
    p = dict(DEFAULTS)
    p.update({k.replace("__", "."): v for k, v in over.items()})
    with open(os.path.join(d, "parameters.txt"), "w") as f:
        f.write("# test parameters\n")
        for k, v in p.items():
            if k not in drop:
                f.write(f"{k} = {v}\n")


def run(prog, cwd, timeout=300):
    r = subprocess.run([os.path.join(BIN, prog)], cwd=cwd, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def read_csv(path):
    import numpy as np

    with open(path) as f:
        header = f.readline().strip().split(",")
        rows = [list(map(float, line.strip().split(","))) for line in f if line.strip()]
    return header, np.array(rows)
