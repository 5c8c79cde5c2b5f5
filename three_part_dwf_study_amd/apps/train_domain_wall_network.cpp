// train_domain_wall_network -- trains the learned preconditioner of DWF_SOLVE_LEARNED for
// the lattice of parameters.txt and writes its weights to gpu.network_weights (NPY <f8),
// where generate_domain_wall (gpu.force_solver = learned) and analyze_domain_wall
// (gpu.propagator_solver = learned) read them.  No counterpart in the reference.
//
// Training data: the configurations of gpu.train_configs (default
// output/domain_wall/configs.npy, shape (n, 2, Nt, Nx); at most gpu.train_max_configs of
// them, evenly spaced), gpu.train_sources full-volume Gaussian sources eta each (host
// mt19937_64(gpu.train_seed)), right-hand sides b = D^dag eta.  The network minimises
// mean ||D^dag D M b - b||^2 / ||b||^2 by Adam on the device (gpu.train_steps steps of
// gpu.train_batch pairs, learning rate gpu.train_lr).

#include "dwf_hip.hpp"
#include "dwf_params.hpp"

#include <iomanip>
#include <iostream>
#include <random>

namespace {
struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes)
    {
        p = dwf_device_alloc(bytes);
        if (!p) throw std::runtime_error(dwf_last_error());
    }
    ~DeviceBuffer() { dwf_device_free(p); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};
}  // namespace

int main()
{
    try {
        const app::Parameters prm;
        const app::DwfParams P(prm);
        if (P.network_weights.empty()) throw std::runtime_error("Missing required parameter: gpu.network_weights");
        const std::string cfg_path = prm.text_or("gpu.train_configs", "output/domain_wall/configs.npy");
        const int max_cfg = std::stoi(prm.text_or("gpu.train_max_configs", "64"));
        const int per = std::stoi(prm.text_or("gpu.train_sources", "16"));
        const int steps = std::stoi(prm.text_or("gpu.train_steps", "1000"));
        const int batch = std::stoi(prm.text_or("gpu.train_batch", "16"));
        const double lr = std::stod(prm.text_or("gpu.train_lr", "3e-3"));
        const unsigned long long seed = std::stoull(prm.text_or("gpu.train_seed", "918273"));
        if (max_cfg < 1 || per < 1 || steps < 1 || batch < 1 || !(lr > 0.0))
            throw std::runtime_error("gpu.train_* values must be positive");
        const int Nt = P.Nt, Nx = P.Nx, Ls = P.N5;
        const app::NpyArray arr = app::load_npy(cfg_path);
        if (arr.shape.size() != 4 || arr.shape[1] != 2 || arr.shape[2] != (size_t)Nt || arr.shape[3] != (size_t)Nx)
            throw std::runtime_error(cfg_path + " has the wrong shape");
        const size_t ngauge = (size_t)2 * Nt * Nx, have = arr.shape[0];
        if (have < 1) throw std::runtime_error(cfg_path + " holds no configurations");
        const int ncfg = (int)std::min<size_t>(have, (size_t)max_cfg);
        std::vector<double> thetas;
        for (int k = 0; k < ncfg; ++k) {  // evenly spaced over the file
            const size_t i = ncfg == 1 ? have - 1 : (size_t)k * (have - 1) / (ncfg - 1);
            thetas.insert(thetas.end(), arr.data.begin() + i * ngauge, arr.data.begin() + (i + 1) * ngauge);
        }
        dwfhip::Context& ctx = dwfhip::Context::dwf(Nt, Nx, Ls, P.M5, P.mf);
        dwfhip::check(dwf_set_gauge_batch(ctx.get(), ncfg, thetas.data()));
        ctx.invalidate_gauge();
        const size_t n4 = (size_t)Ls * Nt * Nx * 4;  // reals per spinor
        const size_t npairs = (size_t)ncfg * per;
        DeviceBuffer b32(npairs * n4 * sizeof(float)), eta(ncfg * n4 * sizeof(double)), b(ncfg * n4 * sizeof(double));
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> normal(0.0, 1.0);
        std::vector<double> h(ncfg * n4);
        // pairs: block j holds source j of every configuration (right-hand side c uses config c)
        for (int j = 0; j < per; ++j) {
            for (double& v : h) v = normal(rng);
            dwfhip::check(dwf_memcpy_htod(eta.p, h.data(), h.size() * sizeof(double)));
            dwfhip::check(dwf_apply_device(ctx.get(), DWF_OP_DDAG, DWF_PREC_FP64, eta.p, b.p, ncfg));
            dwfhip::check(dwf_convert_f64_f32_device(ctx.get(), (const double*)b.p,
                                                     (float*)b32.p + (size_t)j * ncfg * n4, ncfg * n4));
        }
        std::vector<int> pair_cfg(npairs);
        for (size_t i = 0; i < npairs; ++i) pair_cfg[i] = (int)(i % ncfg);
        dwfhip::Network net(Ls, 11);
        dwf_train_params tp{steps, batch, 1, 0, lr, 0.9, 0.999, 1e-8, seed + 1};
        std::vector<double> hist(steps);
        dwfhip::check(dwf_net_train_device(ctx.get(), net.get(), (const float*)b32.p, nullptr, pair_cfg.data(),
                                           (int)npairs, &tp, hist.data()));
        double last = 0.0;
        const int tail = std::min(20, steps);
        for (int i = steps - tail; i < steps; ++i) last += hist[i] / tail;
        app::save_network_weights(P.network_weights, net.weights(Ls));
        std::cout << std::setprecision(6) << "trained on " << npairs << " sources from " << ncfg << " configurations, "
                  << steps << " steps: loss " << hist[0] << " -> " << last << "\nsaved " << P.network_weights << "\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << '\n';
        return 1;
    }
}
