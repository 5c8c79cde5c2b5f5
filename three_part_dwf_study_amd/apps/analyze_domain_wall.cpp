// analyze_domain_wall -- drop-in for the reference's domain_wall/analyze_configs.cpp.
// Configurations are measured in groups (gpu.analysis_batch configurations, or as many as
// the batch size dwf_autotune_batch picks for this lattice and solver): one
// device call solves all 8 + eta.noise_vectors propagators of every configuration in the
// group as one batch and contracts C_PP, C_J5, the momentum-projected point correlator
// and the Z4 loop density in FP64 (dwf_measure_batch).  The ensemble contractions -- averages, m_res,
// the vacuum-subtracted disconnected loop, eta, momentum projection -- stay on the host
// in FP64 (analyze_configs.cpp:78-271) and the three CSV files keep the reference's
// layouts: correlator.csv, channels_by_config.csv, momentum_correlators.csv.

#include "dwf_hip.hpp"
#include "dwf_params.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <iomanip>
#include <iostream>

using Complex = std::complex<double>;
using Corr = std::vector<double>;

int main()
{
    try {
        const app::Parameters prm;
        const app::DwfParams P(prm);
        const std::string weights = P.weights_for(P.propagator_method, "gpu.propagator_solver");  // checked before any device work
        const int Nt = P.Nt, Nx = P.Nx;
        const double pi = 3.141592653589793238462643383279502884;
        const app::NpyArray arr = app::load_npy("output/domain_wall/analysis_configs.npy");
        if (arr.shape.size() != 4 || arr.shape[1] != 2 || arr.shape[2] != (size_t)Nt || arr.shape[3] != (size_t)Nx)
            throw std::runtime_error("analysis_configs.npy has the wrong shape");
        const int ncfg = (int)arr.shape[0];
        const size_t ngauge = (size_t)2 * Nt * Nx;
        if (P.eta_noise_vectors < 1) throw std::runtime_error("eta.noise_vectors must be positive");
        if (P.max_momentum < 0 || P.max_momentum > Nx / 2)
            throw std::runtime_error("analysis.max_momentum is outside 0..Nx/2");

        dwfhip::Context& ctx = dwfhip::Context::dwf(Nt, Nx, P.N5, P.M5, P.mf);
        std::unique_ptr<dwfhip::Network> net;
        if (P.propagator_method == DWF_SOLVE_LEARNED) {
            net.reset(new dwfhip::Network(P.N5));
            net->set_weights(app::load_network_weights(weights,
                                                       dwfhip::Network::size(P.N5)));
            ctx.use_network(net.get());
        }
        std::vector<Corr> CPP_all(ncfg), CJ5_all(ncfg);
        std::vector<std::vector<Complex>> loops_all(ncfg);
        std::vector<std::vector<Corr>> mom_all(ncfg);
        // configurations per batched solve: gpu.analysis_batch, else the right-hand-side
        // batch with the most solves/s on this lattice (autotuned on the first configuration)
        const int per_cfg = 8 + P.eta_noise_vectors;
        int group = P.analysis_batch;
        if (group <= 0 && ncfg > 0) {
            const dwfhip::Gauge first(arr.data.begin(), arr.data.begin() + ngauge);
            const int rhs = dwfhip::tuned_batch(ctx, first, P.propagator_method, P.propagator_rtol, per_cfg * ncfg,
                                                P.tune_cache.empty() ? nullptr : P.tune_cache.c_str());
            group = std::max(1, rhs / per_cfg);
            std::cout << "right-hand sides per batched solve: " << rhs << " (" << group << " configurations)\n";
        }
        for (int c0 = 0; c0 < ncfg; c0 += group) {
            const int n = std::min(group, ncfg - c0);
            std::vector<unsigned long long> seeds(n);
            for (int k = 0; k < n; ++k) {
                std::cout << "measuring config " << c0 + k + 1 << "/" << ncfg << '\n';
                seeds[k] = P.eta_random_seed + (unsigned long long)(c0 + k) * 104729ULL;  // analyze_configs.cpp:67-68
            }
            const std::vector<dwfhip::Measurements> ms =
                dwfhip::measure_batch(ctx, arr.data.data() + (size_t)c0 * ngauge, n, seeds, P.propagator_rtol,
                                      P.propagator_maxiter, P.max_momentum, P.eta_noise_vectors, P.propagator_method);
            for (int k = 0; k < n; ++k) {
                const dwfhip::Measurements& m = ms[k];
                for (int j = 0; j < m.wall_nonconverged; ++j) std::cout << "    CGNR info = nonzero\n";
                CPP_all[c0 + k] = m.pair.CPP;
                CJ5_all[c0 + k] = m.pair.CJ5;
                loops_all[c0 + k] = m.loop;
                mom_all[c0 + k] = m.momentum_CPP;
            }
        }

        // ensemble averages and m_res (analyze_configs.cpp:78-113)
        Corr CPP_avg(Nt, 0.0), CJ5_avg(Nt, 0.0);
        for (int c = 0; c < ncfg; ++c)
            for (int t = 0; t < Nt; ++t) {
                CPP_avg[t] += CPP_all[c][t];
                CJ5_avg[t] += CJ5_all[c][t];
            }
        for (int t = 0; t < Nt; ++t) {
            CPP_avg[t] /= ncfg;
            CJ5_avg[t] /= ncfg;
        }
        Corr mres(Nt, 0.0);
        for (int t = 0; t < Nt; ++t) mres[t] = CJ5_avg[t] / CPP_avg[t];
        double mres_avg = 0.0;
        int mres_n = 0;
        for (int t = 3; t < Nt - 3; ++t) {
            mres_avg += mres[t];
            ++mres_n;
        }
        mres_avg /= mres_n;

        // vacuum-subtracted disconnected loop over all time origins (:115-160)
        std::vector<Complex> loop_mean((size_t)Nt * Nx, Complex(0.0, 0.0));
        for (const auto& L : loops_all)
            for (size_t i = 0; i < loop_mean.size(); ++i) loop_mean[i] += L[i] / (double)ncfg;
        std::vector<Corr> disc_all(ncfg, Corr(Nt, 0.0)), eta_all(ncfg, Corr(Nt, 0.0));
        Corr disc_avg(Nt, 0.0), eta_avg(Nt, 0.0);
        for (int c = 0; c < ncfg; ++c)
            for (int dt = 0; dt < Nt; ++dt) {
                double v = 0.0;
                for (int t0 = 0; t0 < Nt; ++t0) {
                    const int t1 = (t0 + dt) % Nt;
                    Complex l1 = 0.0, l0 = 0.0;
                    for (int x = 0; x < Nx; ++x) {
                        l1 += loops_all[c][t1 * Nx + x] - loop_mean[t1 * Nx + x];
                        l0 += loops_all[c][t0 * Nx + x] - loop_mean[t0 * Nx + x];
                    }
                    v += std::real(l1 * std::conj(l0));
                }
                v /= (double)Nt;
                disc_all[c][dt] = v;
                eta_all[c][dt] = mom_all[c][0][dt] - P.eta_flavors * v;
                disc_avg[dt] += v / (double)ncfg;
                eta_avg[dt] += eta_all[c][dt] / (double)ncfg;
            }

        // momentum-projected channels (:172-227)
        std::filesystem::create_directories("output/domain_wall");
        std::ofstream mcsv("output/domain_wall/momentum_correlators.csv");
        mcsv << "momentum_index,p,p_squared,t,pion,disconnected,eta\n";
        for (int n = 0; n <= P.max_momentum; ++n) {
            const double p = 2.0 * pi * n / Nx;
            std::vector<std::vector<Complex>> lp(ncfg, std::vector<Complex>(Nt, Complex(0.0, 0.0)));
            std::vector<Complex> lp_mean(Nt, Complex(0.0, 0.0));
            for (int c = 0; c < ncfg; ++c)
                for (int t = 0; t < Nt; ++t) {
                    for (int x = 0; x < Nx; ++x) lp[c][t] += std::exp(Complex(0.0, -p * x)) * loops_all[c][t * Nx + x];
                    lp_mean[t] += lp[c][t] / (double)ncfg;
                }
            for (int dt = 0; dt < Nt; ++dt) {
                double pion = 0.0, disc = 0.0;
                for (int c = 0; c < ncfg; ++c) {
                    pion += mom_all[c][n][dt] / (double)ncfg;
                    double one = 0.0;
                    for (int t0 = 0; t0 < Nt; ++t0) {
                        const int t1 = (t0 + dt) % Nt;
                        one += std::real((lp[c][t1] - lp_mean[t1]) * std::conj(lp[c][t0] - lp_mean[t0]));
                    }
                    disc += one / (double)(Nt * ncfg);
                }
                mcsv << n << "," << p << "," << p * p << "," << dt << "," << pion << "," << disc << ","
                     << pion - P.eta_flavors * disc << "\n";
            }
        }

        std::ofstream csv("output/domain_wall/correlator.csv");
        csv << "t,CPP_average,CJ5_average,m_res,disconnected,eta_singlet\n";
        for (int t = 0; t < Nt; ++t)
            csv << t << "," << CPP_avg[t] << "," << CJ5_avg[t] << "," << mres[t] << "," << disc_avg[t] << ","
                << eta_avg[t] << "\n";
        std::ofstream bc("output/domain_wall/channels_by_config.csv");
        bc << "config";
        for (int t = 0; t < Nt; ++t) bc << ",pion_t" << t;
        for (int t = 0; t < Nt; ++t) bc << ",disconnected_t" << t;
        for (int t = 0; t < Nt; ++t) bc << ",eta_t" << t;
        bc << "\n";
        for (int c = 0; c < ncfg; ++c) {
            bc << c;
            for (double v : CPP_all[c]) bc << "," << v;
            for (double v : disc_all[c]) bc << "," << v;
            for (double v : eta_all[c]) bc << "," << v;
            bc << "\n";
        }
        std::cout << std::setprecision(10) << "\nall configs = (" << P.trajectories << ", 2, " << Nt << ", " << Nx << ")"
                  << "\nanalysis configs = (" << ncfg << ", 2, " << Nt << ", " << Nx << ")"
                  << "\nm_res = " << mres_avg << "\n\nWrote:\n"
                  << "  output/domain_wall/correlator.csv\n"
                  << "  output/domain_wall/channels_by_config.csv\n"
                  << "  output/domain_wall/momentum_correlators.csv\n";
        ctx.use_network(nullptr);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << '\n';
        return 1;
    }
}
