// dwf_params.hpp -- the domain-wall and Wilson keys of parameters.txt, read in the
// reference's declaration order (domain_wall/include/parameters.hpp:7-34,
// wilson/include/parameters.hpp:6-38), so a missing or malformed key is reported for
// the same key first.
#pragma once

#include "app_common.hpp"

namespace app {

struct DwfParams {
    int Nt, Nx, N5;
    double beta, M5, mf, epsilon;
    int leapfrog_steps, trajectories, thermalization_cut, thin;
    double pseudofermion_rtol;
    int pseudofermion_maxiter;
    double propagator_rtol;
    int propagator_maxiter, analysis_threads, eta_noise_vectors, eta_flavors;
    unsigned long long eta_random_seed;
    int max_momentum;
    unsigned long long random_seed;
    int force_method, propagator_method;  // gpu.* (additive keys)
    int analysis_batch;                   // configurations per batched solve (0 = autotuned)
    std::string network_weights;          // learned preconditioner weights (NPY), for "learned" solvers
    std::string tune_cache;               // autotuner result file (empty: none)

    explicit DwfParams(const Parameters& p)
    {
        Nt = p.integer("dwf.Nt");
        Nx = p.integer("dwf.Nx");
        N5 = p.integer("dwf.N5");
        beta = p.real("dwf.beta");
        M5 = p.real("dwf.M5");
        mf = p.real("dwf.mf");
        epsilon = p.real("dwf.epsilon");
        leapfrog_steps = p.integer("dwf.leapfrog_steps");
        trajectories = p.integer("dwf.trajectories");
        thermalization_cut = p.integer("dwf.thermalization_cut");
        thin = p.integer("dwf.thin");
        pseudofermion_rtol = p.real("dwf.pseudofermion_rtol");
        pseudofermion_maxiter = p.integer("dwf.pseudofermion_maxiter");
        propagator_rtol = p.real("dwf.propagator_rtol");
        propagator_maxiter = p.integer("dwf.propagator_maxiter");
        analysis_threads = p.integer("analysis.threads");
        eta_noise_vectors = p.integer("eta.noise_vectors");
        eta_flavors = p.integer("eta.flavors");
        eta_random_seed = p.unsigned_integer("eta.random_seed");
        max_momentum = p.integer("analysis.max_momentum");
        random_seed = p.unsigned_integer("dwf.random_seed");
        force_method = force_solver_method(p, "gpu.force_solver");
        propagator_method = solver_method(p, "gpu.propagator_solver", 0);
        analysis_batch = std::stoi(p.text_or("gpu.analysis_batch", "0"));
        if (analysis_batch < 0) throw std::runtime_error("gpu.analysis_batch must be >= 0");
        network_weights = p.text_or("gpu.network_weights", "");
        tune_cache = p.text_or("gpu.tune_cache", "");
    }
    // The weights file a "learned" solver needs.
    const std::string& weights_for(int method, const char* key) const
    {
        if (method == 3 && network_weights.empty())
            throw std::runtime_error(std::string(key) + " = learned needs gpu.network_weights");
        return network_weights;
    }
};

struct WilsonParams {
    int Nt, Nx;
    double beta, m0;
    int trajectories, thermalization_cut, thin, hmc_steps;
    double trajectory_length, hmc_epsilon;
    int source_t, source_x;
    double time_boundary_sign, space_boundary_sign, force_cg_tolerance;
    int force_cg_max_iterations;
    double measurement_cg_tolerance;
    int measurement_cg_max_iterations, plateau_window_size;
    double plateau_flatness_tolerance;
    int earliest_plateau_time, analysis_threads;
    unsigned long long random_seed;
    int force_method, propagator_method;

    explicit WilsonParams(const Parameters& p)
    {
        Nt = p.integer("wilson.Nt");
        Nx = p.integer("wilson.Nx");
        beta = p.real("wilson.beta");
        m0 = p.real("wilson.m0");
        trajectories = p.integer("wilson.trajectories");
        thermalization_cut = p.integer("wilson.thermalization_cut");
        thin = p.integer("wilson.thin");
        hmc_steps = p.integer("wilson.hmc_steps");
        trajectory_length = p.real("wilson.trajectory_length");
        hmc_epsilon = trajectory_length / hmc_steps;
        source_t = p.integer("wilson.source_t");
        source_x = p.integer("wilson.source_x");
        time_boundary_sign = p.real("wilson.time_boundary_sign");
        space_boundary_sign = p.real("wilson.space_boundary_sign");
        force_cg_tolerance = p.real("wilson.force_cg_tolerance");
        force_cg_max_iterations = p.integer("wilson.force_cg_max_iterations");
        measurement_cg_tolerance = p.real("wilson.measurement_cg_tolerance");
        measurement_cg_max_iterations = p.integer("wilson.measurement_cg_max_iterations");
        plateau_window_size = p.integer("analysis.plateau_window_size");
        plateau_flatness_tolerance = p.real("analysis.plateau_flatness_tolerance");
        earliest_plateau_time = p.integer("analysis.earliest_plateau_time");
        analysis_threads = p.integer("analysis.threads");
        random_seed = p.unsigned_integer("wilson.random_seed");
        force_method = force_solver_method(p, "gpu.force_solver");
        propagator_method = solver_method(p, "gpu.propagator_solver", 0);
    }
};

}  // namespace app
