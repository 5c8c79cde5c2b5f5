// generate_wilson -- drop-in for the reference's wilson/generate_configs.cpp: Nf=2
// Wilson pseudofermion HMC on the GPU (hmc_step of wilson/include/gauge.hpp:132-180:
// momenta drawn before chi, wrap_angle after each drift).  Same parameters.txt keys,
// the same random stream, the same outputs: output/wilson/wilson_ensemble.bin (WILCFG1)
// and output/wilson/hmc_history.csv.

#include "dwf_hip.hpp"
#include "dwf_params.hpp"

#include <iomanip>
#include <iostream>

int main()
{
    try {
        const app::Parameters prm;
        const app::WilsonParams P(prm);
        std::filesystem::create_directories("output/wilson");
        std::ofstream history("output/wilson/hmc_history.csv");
        history << "trajectory,accepted,dH,acceptance_probability,plaquette,saved\n";
        dwfhip::HostRng rng(P.random_seed);
        dwfhip::Context& ctx = dwfhip::Context::wilson(P.Nt, P.Nx, P.m0, P.time_boundary_sign, P.space_boundary_sign);
        std::vector<double> theta((size_t)2 * P.Nt * P.Nx, 0.0);  // cold start
        std::vector<std::vector<double>> saved;
        int accepted_count = 0;
        dwf_hmc_params hp{P.beta, P.hmc_epsilon, P.hmc_steps, DWF_HMC_WILSON, P.force_method,
                          P.force_cg_max_iterations, P.force_cg_tolerance};
        for (int traj = 0; traj < P.trajectories; ++traj) {
            const dwfhip::HmcResult r = dwfhip::hmc_trajectory(ctx, theta, rng, hp);
            if (r.accepted) ++accepted_count;
            if (traj >= P.thermalization_cut && (traj - P.thermalization_cut) % P.thin == 0) saved.push_back(theta);
            history << traj << "," << r.accepted << "," << r.dH << "," << r.acceptance_probability << "," << r.plaquette
                    << "," << saved.size() << "\n";
            std::cout << "traj " << std::setw(4) << traj + 1 << "/" << P.trajectories << " accepted=" << r.accepted
                      << " dH=" << std::scientific << r.dH << " Pacc=" << r.acceptance_probability << " plaq=" << std::fixed
                      << r.plaquette << " saved=" << saved.size() << "\n";
        }
        app::save_wilcfg("output/wilson/wilson_ensemble.bin", P.Nt, P.Nx, P.beta, P.m0, saved);
        std::cout << "\nSaved " << saved.size() << " configurations to output/wilson/wilson_ensemble.bin\n"
                  << "Acceptance rate = " << (double)accepted_count / P.trajectories << "\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
