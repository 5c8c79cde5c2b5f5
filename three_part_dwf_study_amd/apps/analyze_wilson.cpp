// analyze_wilson -- drop-in for the reference's wilson/analyze_configs.cpp: for every
// configuration of the WILCFG1 ensemble, the point-source pseudoscalar correlator
// C(t) = sum_x,a,b |G(t,x;a,b)|^2 from two CGNR solves (wilson/include/correlator.hpp:10-40)
// on the GPU, written to output/wilson/correlators_by_config.csv in the reference's layout;
// then on the host the jackknife cosh effective mass, the flat-plateau search and the
// printed "Wilson pseudoscalar mass = m +/- dm" (analyze_configs.cpp:36-79,
// wilson/include/statistics.hpp), with the plotted series written to
// output/wilson/effective_mass.csv in place of the reference's SVG plots.

#include "dwf_hip.hpp"
#include "dwf_params.hpp"
#include "wilson_stats.hpp"

#include <complex>
#include <iomanip>
#include <iostream>

int main()
{
    try {
        const app::Parameters prm;
        const app::WilsonParams P(prm);
        const int Nt = P.Nt, Nx = P.Nx;
        const auto configs = app::load_wilcfg("output/wilson/wilson_ensemble.bin", Nt, Nx);
        if (configs.size() < 2) throw std::runtime_error("Need at least two configurations");
        dwfhip::Context& ctx = dwfhip::Context::wilson(Nt, Nx, P.m0, P.time_boundary_sign, P.space_boundary_sign);
        const size_t ntot = (size_t)Nt * Nx * 2;
        std::vector<std::vector<double>> all(configs.size(), std::vector<double>(Nt, 0.0));
        for (size_t i = 0; i < configs.size(); ++i) {
            std::cout << "Measuring configuration " << i + 1 << "/" << configs.size() << "\n";
            std::vector<std::vector<std::complex<double>>> prop(2);
            for (int ss = 0; ss < 2; ++ss) {
                std::vector<std::complex<double>> src(ntot, 0.0);
                src[(size_t)((P.source_t * Nx + P.source_x) * 2 + ss)] = 1.0;
                const auto r = dwfhip::solve_CGNR(ctx, configs[i], src, P.measurement_cg_tolerance,
                                                  P.measurement_cg_max_iterations, P.propagator_method);
                std::cout << "    spin=" << ss << " CG iterations=" << r.iterations
                          << " normal residual=" << r.normal_residual << " D residual=" << r.dirac_residual << "\n";
                prop[ss] = r.x;
            }
            for (int t = 0; t < Nt; ++t)
                for (int x = 0; x < Nx; ++x)
                    for (int sink = 0; sink < 2; ++sink)
                        for (int ss = 0; ss < 2; ++ss) all[i][t] += std::norm(prop[ss][(size_t)((t * Nx + x) * 2 + sink)]);
        }
        const app::WilsonMass m = app::wilson_mass(
            all, app::PlateauSettings{P.plateau_window_size, P.plateau_flatness_tolerance, P.earliest_plateau_time},
            &std::cout);
        app::write_effective_mass_csv("output/wilson/effective_mass.csv", m);
        std::ofstream corr("output/wilson/correlators_by_config.csv");
        corr << "config";
        for (int t = 0; t < Nt; ++t) corr << ",t" << t;
        corr << "\n";
        for (size_t i = 0; i < all.size(); ++i) {
            corr << i;
            for (double v : all[i]) corr << "," << v;
            corr << "\n";
        }
        std::cout << "Wrote:\n  output/wilson/effective_mass.csv\n  output/wilson/correlators_by_config.csv\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
