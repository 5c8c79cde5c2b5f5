// generate_domain_wall -- drop-in for the reference's domain_wall/generate_configs.cpp:
// Nf=2 Shamir domain-wall pseudofermion HMC on the GPU.  Same parameters.txt keys, the
// same random stream (host mt19937_64 in the reference's draw order), the same outputs:
// output/domain_wall/configs.npy (every trajectory's theta0), analysis_configs.npy
// (after the thermalization cut, thinned) and run_info_history.csv.  Each trajectory
// runs on the device in FP64 (dwf_hmc_trajectory); the force solves use
// gpu.force_solver (cg64 by default, the reference's algorithm; "mixed"; "eo_mixed", two
// even-odd Dirac solves per solve; "learned" with the network weights of
// gpu.network_weights, e.g. from train_domain_wall_network).

#include "dwf_hip.hpp"
#include "dwf_params.hpp"

#include <cmath>
#include <filesystem>
#include <iomanip>
#include <iostream>

int main()
{
    try {
        const app::Parameters prm;
        const app::DwfParams P(prm);
        const std::string weights = P.weights_for(P.force_method, "gpu.force_solver");  // checked before any device work
        std::filesystem::create_directories("output/domain_wall");
        dwfhip::HostRng rng(P.random_seed);
        dwfhip::Context& ctx = dwfhip::Context::dwf(P.Nt, P.Nx, P.N5, P.M5, P.mf);
        std::unique_ptr<dwfhip::Network> net;
        if (P.force_method == DWF_SOLVE_LEARNED) {
            net.reset(new dwfhip::Network(P.N5));
            net->set_weights(app::load_network_weights(weights,
                                                       dwfhip::Network::size(P.N5)));
            ctx.use_network(net.get());
        }
        const size_t ngauge = (size_t)2 * P.Nt * P.Nx;
        std::vector<double> theta0(ngauge, 1.0);  // cold start, generate_configs.cpp:27-28

        std::vector<double> configs, analysis;
        configs.reserve((size_t)P.trajectories * ngauge);
        std::vector<double> plaq(P.trajectories, 0.0), dh(P.trajectories, 0.0), pacc(P.trajectories, 0.0);
        std::vector<int> acc(P.trajectories, 0);
        int accepted_count = 0;
        dwf_hmc_params hp{P.beta, P.epsilon, P.leapfrog_steps, DWF_HMC_DOMAIN_WALL, P.force_method,
                          P.pseudofermion_maxiter, P.pseudofermion_rtol};
        for (int traj = 0; traj < P.trajectories; ++traj) {
            const dwfhip::HmcResult r = dwfhip::hmc_trajectory(ctx, theta0, rng, hp);
            if (r.accepted) ++accepted_count;
            std::cout << "traj " << std::setw(4) << traj << "  accepted=" << r.accepted << "  acc_prob=" << std::fixed
                      << std::setprecision(3) << r.acceptance_probability << "  dH=" << std::scientific
                      << std::setprecision(3) << r.dH << "  plaq=" << std::fixed << std::setprecision(5) << r.plaquette
                      << '\n';
            configs.insert(configs.end(), theta0.begin(), theta0.end());
            if (traj >= P.thermalization_cut && (traj - P.thermalization_cut) % P.thin == 0)
                analysis.insert(analysis.end(), theta0.begin(), theta0.end());
            plaq[traj] = r.plaquette;
            acc[traj] = r.accepted ? 1 : 0;
            dh[traj] = r.dH;
            pacc[traj] = r.acceptance_probability;
        }
        std::cout.unsetf(std::ios::floatfield);
        std::cout << std::setprecision(6);
        const size_t nan = analysis.size() / ngauge;
        app::save_npy("output/domain_wall/configs.npy", {(size_t)P.trajectories, 2, (size_t)P.Nt, (size_t)P.Nx}, configs);
        app::save_npy("output/domain_wall/analysis_configs.npy", {nan, 2, (size_t)P.Nt, (size_t)P.Nx}, analysis);
        std::ofstream hist("output/domain_wall/run_info_history.csv");
        hist << "trajectory,accepted,acceptance_probability,delta_h,plaquette\n";
        for (int i = 0; i < P.trajectories; ++i)
            hist << i << "," << acc[i] << "," << pacc[i] << "," << dh[i] << "," << plaq[i] << "\n";
        std::cout << "\nacceptance rate = " << (double)accepted_count / P.trajectories << "\nall configs = ("
                  << P.trajectories << ", 2, " << P.Nt << ", " << P.Nx << ")"
                  << "\nanalysis configs = (" << nan << ", 2, " << P.Nt << ", " << P.Nx << ")\n"
                  << "saved configs.npy\nsaved analysis_configs.npy\nsaved run_info_history.csv\n";
        ctx.use_network(nullptr);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << '\n';
        return 1;
    }
}
