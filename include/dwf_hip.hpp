// dwf_hip.hpp -- C++17 drop-in layer over libdwfhip's C ABI (dwf_hip.h).
//
// The reference's hot path is a set of inline functions over Eigen types
// (domain_wall/include/{dwf_operator,cg_solver,pseudofermion,gauge,correlator}.hpp,
// wilson/include/{wilson_operator,cg_solver,gauge,correlator}.hpp).  This header gives
// the same names, result structs and exception messages, backed by the GPU.  Spinor
// containers are templates: anything holding contiguous std::complex<double> with
// data(), size() and a size constructor -- Eigen::VectorXcd (the reference's VectorC)
// or std::vector<std::complex<double>>.  Gauge fields are std::vector<double> in the
// reference's [mu][t][x] order.
//
// The operator parameters are explicit (the reference reads them from process-global
// constants, domain_wall/include/parameters.hpp:7-34); Context::dwf / Context::wilson
// return a context per host thread and parameter set, so OpenMP analyzers that call
// these functions from several threads each get their own device context and stream
// (domain_wall/analyze_configs.cpp:45).  The gauge field is uploaded only when it
// changes.
#pragma once

#include "dwf_hip.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace dwfhip {

using Complex = std::complex<double>;
using Gauge = std::vector<double>;
using Corr = std::vector<double>;
using ComplexCorr = std::vector<Complex>;

// Throws std::runtime_error carrying the library's message (which is the reference's
// own text where the reference throws).
inline void check(int rc)
{
    if (rc != DWF_OK) throw std::runtime_error(dwf_last_error());
}

class Context {
public:
    // One context per (thread, parameter set), created on first use.
    static Context& dwf(int Nt, int Nx, int Ls, double M5, double mf)
    {
        return lookup(Key{0, Nt, Nx, Ls, M5, mf, 0.0, 0.0});
    }
    static Context& wilson(int Nt, int Nx, double m0, double bc_t, double bc_x)
    {
        return lookup(Key{1, Nt, Nx, 1, m0, 0.0, bc_t, bc_x});
    }

    dwf_ctx* get() const { return ctx_.get(); }
    int Nt() const { return std::get<1>(key_); }
    int Nx() const { return std::get<2>(key_); }
    int Ls() const { return std::get<3>(key_); }
    size_t ndof() const { return (size_t)Ls() * Nt() * Nx() * 2; }  // complex components
    size_t ngauge() const { return (size_t)2 * Nt() * Nx(); }

    // Upload theta if it differs from the field already on the device.
    void use_gauge(const Gauge& theta)
    {
        if (theta.size() != ngauge()) throw std::runtime_error("gauge field has the wrong size");
        if (have_ && std::memcmp(theta.data(), last_.data(), theta.size() * sizeof(double)) == 0) return;
        check(dwf_set_gauge(get(), theta.data()));
        last_ = theta;
        have_ = true;
    }
    // The device field changed behind our back (an HMC trajectory): forget the cache.
    void invalidate_gauge() { have_ = false; }
    // The device already holds theta (read back after an HMC trajectory).
    void adopt_gauge(const Gauge& theta)
    {
        last_ = theta;
        have_ = true;
    }
    // The learned preconditioner DWF_SOLVE_LEARNED uses (nullptr detaches); the network
    // must outlive its use here.
    void use_network(const class Network* net);

private:
    using Key = std::tuple<int, int, int, int, double, double, double, double>;
    struct Del {
        void operator()(dwf_ctx* c) const { dwf_ctx_destroy(c); }
    };
    std::unique_ptr<dwf_ctx, Del> ctx_;
    Key key_;
    Gauge last_;
    bool have_ = false;

    explicit Context(const Key& k) : key_(k)
    {
        dwf_ctx* c = nullptr;
        if (std::get<0>(k) == 0)
            check(dwf_ctx_create_dwf(&c, std::get<1>(k), std::get<2>(k), std::get<3>(k), std::get<4>(k), std::get<5>(k)));
        else
            check(dwf_ctx_create_wilson(&c, std::get<1>(k), std::get<2>(k), std::get<4>(k), std::get<6>(k),
                                        std::get<7>(k)));
        ctx_.reset(c);
    }
    static Context& lookup(const Key& k)
    {
        thread_local std::map<Key, std::unique_ptr<Context>> cache;
        auto it = cache.find(k);
        if (it == cache.end()) it = cache.emplace(k, std::unique_ptr<Context>(new Context(k))).first;
        return *it->second;
    }
};

// The learned preconditioner's network (no counterpart in the reference): weights are
// dwf_net_num_weights(Ls) floats, e.g. written by train_domain_wall_network.
class Network {
public:
    explicit Network(int Ls, unsigned long long seed = 0)
    {
        dwf_net* n = nullptr;
        check(dwf_net_create(&n, Ls, seed));
        net_.reset(n);
    }
    static size_t size(int Ls) { return (size_t)dwf_net_num_weights(Ls); }
    void set_weights(const std::vector<float>& w) { check(dwf_net_set_weights(net_.get(), w.data())); }
    std::vector<float> weights(int Ls) const
    {
        std::vector<float> w(size(Ls));
        check(dwf_net_get_weights(net_.get(), w.data()));
        return w;
    }
    dwf_net* get() const { return net_.get(); }

private:
    struct Del {
        void operator()(dwf_net* n) const { dwf_net_destroy(n); }
    };
    std::unique_ptr<dwf_net, Del> net_;
};

inline void Context::use_network(const Network* net) { check(dwf_ctx_set_preconditioner(get(), net ? net->get() : nullptr)); }

// Right-hand sides per batched solve with the most solves/s for this lattice and method
// (dwf_autotune_batch, on the given configuration).
inline int tuned_batch(Context& c, const Gauge& theta, int method, double rtol, int max_nrhs,
                       const char* cache_path = nullptr)
{
    c.use_gauge(theta);
    dwf_batch_tune_result r{};
    check(dwf_autotune_batch(c.get(), method, rtol, max_nrhs, cache_path, &r));
    return r.nrhs;
}

namespace detail {
template <class V> V make(size_t n) { return V(static_cast<decltype(std::declval<V>().size())>(n)); }
template <class V> const double* dptr(const V& v) { return reinterpret_cast<const double*>(v.data()); }
template <class V> double* dptr(V& v) { return reinterpret_cast<double*>(v.data()); }
template <class V> V apply(Context& c, int op, const Gauge& theta, const V& psi)
{
    c.use_gauge(theta);
    V out = make<V>((size_t)psi.size());
    check(dwf_apply(c.get(), op, dptr(psi), dptr(out), 1));
    return out;
}
}  // namespace detail

// ------------------------------------------------------ dwf_operator.hpp:15-181
template <class V> V apply_D(Context& c, const Gauge& theta, const V& psi) { return detail::apply(c, DWF_OP_D, theta, psi); }
template <class V> V apply_D_dagger(Context& c, const Gauge& theta, const V& psi)
{
    return detail::apply(c, DWF_OP_DDAG, theta, psi);
}
template <class V> V apply_DD_dagger(Context& c, const Gauge& theta, const V& psi)
{
    return detail::apply(c, DWF_OP_DDDAG, theta, psi);
}
template <class V> V apply_D_dagger_D(Context& c, const Gauge& theta, const V& psi)
{
    return detail::apply(c, DWF_OP_DDAGD, theta, psi);
}

// --------------------------------------------------------- cg_solver.hpp:9-64
template <class V> struct CgResult {
    V x;
    int iterations;
    double relative_residual;
    bool converged;
};

// conjugate_gradient(apply_DD_dagger | apply_D_dagger_D, b, guess, rtol, maxiter): the
// operator is named by dwf_op instead of passed as a std::function.  Throws "CG
// encountered a zero denominator" as the reference; non-convergence is reported, not
// thrown.
template <class V>
CgResult<V> conjugate_gradient(Context& c, const Gauge& theta, int normal_op, const V& b, const V* guess, double rtol,
                               int maxiter, int method = DWF_SOLVE_CG64)
{
    c.use_gauge(theta);
    CgResult<V> r{detail::make<V>((size_t)b.size()), 0, 0.0, false};
    dwf_solve_stats st{};
    const int rc = dwf_solve(c.get(), normal_op, method, detail::dptr(b), guess ? detail::dptr(*guess) : nullptr, rtol,
                             maxiter, 1, detail::dptr(r.x), &st);
    if (rc == DWF_ERR_ZERO_DENOM || rc == DWF_ERR_HIP || rc == DWF_ERR_ARG) check(rc);
    r.iterations = st.iterations;
    r.relative_residual = st.relres;
    r.converged = st.converged != 0;
    return r;
}

// --------------------------------------------------- pseudofermion.hpp:13-166
template <class V> struct PseudofermionActionForce {
    double action;
    Gauge force;
    V eta;
    int iterations;
    double relative_residual;
};

template <class V>
PseudofermionActionForce<V> pseudofermion_action_force(Context& c, const Gauge& theta, const V& phi,
                                                       const V* eta_guess, double rtol, int maxiter,
                                                       int method = DWF_SOLVE_CG64)
{
    c.use_gauge(theta);
    PseudofermionActionForce<V> r{0.0, Gauge(c.ngauge(), 0.0), detail::make<V>((size_t)phi.size()), 0, 0.0};
    dwf_solve_stats st{};
    check(dwf_pf_action_force(c.get(), method, detail::dptr(phi), eta_guess ? detail::dptr(*eta_guess) : nullptr, rtol,
                              maxiter, &r.action, r.force.data(), detail::dptr(r.eta), &st));
    r.iterations = st.iterations;
    r.relative_residual = st.relres;
    return r;
}

// --------------------------------------------------------------- gauge.hpp:10-73
struct GaugeActionForce {
    double action;
    Gauge force;
};

inline GaugeActionForce gauge_action_force(Context& c, const Gauge& theta, double beta)
{
    c.use_gauge(theta);
    GaugeActionForce r{0.0, Gauge(c.ngauge(), 0.0)};
    check(dwf_gauge_action_force(c.get(), beta, &r.action, r.force.data()));
    return r;
}

inline double average_plaquette(Context& c, const Gauge& theta)
{
    c.use_gauge(theta);
    double p = 0.0;
    check(dwf_average_plaquette(c.get(), &p));
    return p;
}

// ----------------------------------------------------------- correlator.hpp:17-310
struct CorrelatorPair {
    Corr CPP;
    Corr CJ5;
};

struct Measurements {
    CorrelatorPair pair;
    ComplexCorr loop;
    std::vector<Corr> momentum_CPP;
    int wall_nonconverged = 0;
};

// One configuration's analyzer measurements in one batched solve (the loop body of
// domain_wall/analyze_configs.cpp:57-75).  noise_vectors = 0 skips the loop density.
inline Measurements measure(Context& c, const Gauge& theta, double rtol, int maxiter, int max_momentum,
                            int noise_vectors, unsigned long long noise_seed, int method = DWF_SOLVE_CG64)
{
    c.use_gauge(theta);
    Measurements m;
    const int Nt = c.Nt(), Nx = c.Nx();
    m.pair.CPP.assign(Nt, 0.0);
    m.pair.CJ5.assign(Nt, 0.0);
    std::vector<double> mom((size_t)(max_momentum + 1) * Nt, 0.0);
    std::vector<double> loop((size_t)Nt * Nx * 2, 0.0);
    dwf_measure_spec spec{method, maxiter, rtol, 1, 1, max_momentum, noise_vectors, noise_seed};
    dwf_measure_out out{m.pair.CPP.data(), m.pair.CJ5.data(), mom.data(), loop.data(), nullptr, nullptr, 0, 0};
    check(dwf_measure(c.get(), &spec, &out));
    m.wall_nonconverged = out.wall_nonconverged;
    m.momentum_CPP.assign(max_momentum + 1, Corr(Nt, 0.0));
    for (int n = 0; n <= max_momentum; ++n)
        for (int t = 0; t < Nt; ++t) m.momentum_CPP[n][t] = mom[(size_t)n * Nt + t];
    m.loop.resize((size_t)Nt * Nx);
    for (size_t i = 0; i < m.loop.size(); ++i) m.loop[i] = Complex(loop[2 * i], loop[2 * i + 1]);
    return m;
}

// The same for ncfg configurations at once (thetas: ncfg consecutive fields, seeds: one
// Z4 seed each): every configuration's sources go through one batched solve.
inline std::vector<Measurements> measure_batch(Context& c, const double* thetas, int ncfg,
                                               const std::vector<unsigned long long>& seeds, double rtol,
                                               int maxiter, int max_momentum, int noise_vectors,
                                               int method = DWF_SOLVE_CG64)
{
    if (ncfg < 1 || (int)seeds.size() != ncfg) throw std::runtime_error("measure_batch: one seed per configuration");
    const int Nt = c.Nt(), Nx = c.Nx();
    check(dwf_set_gauge_batch(c.get(), ncfg, thetas));
    c.invalidate_gauge();
    std::vector<Measurements> ms(ncfg);
    std::vector<std::vector<double>> mom(ncfg, std::vector<double>((size_t)(max_momentum + 1) * Nt, 0.0));
    std::vector<std::vector<double>> loop(ncfg, std::vector<double>((size_t)Nt * Nx * 2, 0.0));
    std::vector<dwf_measure_out> outs(ncfg);
    for (int k = 0; k < ncfg; ++k) {
        ms[k].pair.CPP.assign(Nt, 0.0);
        ms[k].pair.CJ5.assign(Nt, 0.0);
        outs[k] = dwf_measure_out{ms[k].pair.CPP.data(), ms[k].pair.CJ5.data(), mom[k].data(), loop[k].data(),
                                  nullptr, nullptr, 0, 0};
    }
    dwf_measure_spec spec{method, maxiter, rtol, 1, 1, max_momentum, noise_vectors, 0};
    const int rc = dwf_measure_batch(c.get(), &spec, ncfg, seeds.data(), outs.data());
    check(dwf_set_gauge_batch(c.get(), 1, thetas));  // leave the context with one configuration
    check(rc);
    for (int k = 0; k < ncfg; ++k) {
        Measurements& m = ms[k];
        m.wall_nonconverged = outs[k].wall_nonconverged;
        m.momentum_CPP.assign(max_momentum + 1, Corr(Nt, 0.0));
        for (int n = 0; n <= max_momentum; ++n)
            for (int t = 0; t < Nt; ++t) m.momentum_CPP[n][t] = mom[k][(size_t)n * Nt + t];
        m.loop.resize((size_t)Nt * Nx);
        for (size_t i = 0; i < m.loop.size(); ++i) m.loop[i] = Complex(loop[k][2 * i], loop[k][2 * i + 1]);
    }
    return ms;
}

inline CorrelatorPair compute_CPP_CJ5(Context& c, const Gauge& theta, double rtol, int maxiter,
                                      int method = DWF_SOLVE_CG64)
{
    c.use_gauge(theta);
    CorrelatorPair p{Corr(c.Nt(), 0.0), Corr(c.Nt(), 0.0)};
    dwf_measure_spec spec{method, maxiter, rtol, 1, 0, 0, 0, 0};
    dwf_measure_out out{p.CPP.data(), p.CJ5.data(), nullptr, nullptr, nullptr, nullptr, 0, 0};
    check(dwf_measure(c.get(), &spec, &out));
    return p;
}

inline ComplexCorr estimate_pseudoscalar_loop_density(Context& c, const Gauge& theta, unsigned long long seed,
                                                      int noise_vectors, double rtol, int maxiter,
                                                      int method = DWF_SOLVE_CG64)
{
    if (noise_vectors < 1) throw std::runtime_error("eta.noise_vectors must be positive");  // correlator.hpp:31-32
    c.use_gauge(theta);
    std::vector<double> loop((size_t)c.Nt() * c.Nx() * 2, 0.0);
    dwf_measure_spec spec{method, maxiter, rtol, 0, 0, 0, noise_vectors, seed};
    dwf_measure_out out{nullptr, nullptr, nullptr, loop.data(), nullptr, nullptr, 0, 0};
    check(dwf_measure(c.get(), &spec, &out));
    ComplexCorr r((size_t)c.Nt() * c.Nx());
    for (size_t i = 0; i < r.size(); ++i) r[i] = Complex(loop[2 * i], loop[2 * i + 1]);
    return r;
}

inline std::vector<Corr> compute_momentum_CPP(Context& c, const Gauge& theta, int max_momentum, double rtol,
                                              int maxiter, int method = DWF_SOLVE_CG64)
{
    c.use_gauge(theta);
    std::vector<double> mom((size_t)(max_momentum + 1) * c.Nt(), 0.0);
    dwf_measure_spec spec{method, maxiter, rtol, 0, 1, max_momentum, 0, 0};
    dwf_measure_out out{nullptr, nullptr, mom.data(), nullptr, nullptr, nullptr, 0, 0};
    check(dwf_measure(c.get(), &spec, &out));
    std::vector<Corr> r(max_momentum + 1, Corr(c.Nt(), 0.0));
    for (int n = 0; n <= max_momentum; ++n)
        for (int t = 0; t < c.Nt(); ++t) r[n][t] = mom[(size_t)n * c.Nt() + t];
    return r;
}

// ------------------------------------------- wilson/include/cg_solver.hpp:11-78
template <class V> struct SolveResult {
    V x;
    int iterations;
    double normal_residual;
    double dirac_residual;
};

// solve_CGNR(D, source, tolerance, max_iterations): CG on D^dag D with rhs D^dag source;
// throws "CGNR failed: iterations=..., normal residual=..., D residual=..." when the
// normal residual exceeds 20 x tolerance, as the reference.
template <class V>
SolveResult<V> solve_CGNR(Context& c, const Gauge& theta, const V& source, double tolerance, int max_iterations,
                          int method = DWF_SOLVE_CG64)
{
    c.use_gauge(theta);
    const size_t n = (size_t)source.size();
    SolveResult<V> r{detail::make<V>(n), 0, 0.0, 0.0};
    dwf_solve_stats st{};
    const int rc = dwf_solve_dirac(c.get(), method, detail::dptr(source), nullptr, tolerance, max_iterations, 1,
                                   detail::dptr(r.x), &st);
    if (rc == DWF_ERR_ZERO_DENOM || rc == DWF_ERR_HIP || rc == DWF_ERR_ARG) check(rc);
    V Dx = apply_D(c, theta, r.x);
    double dn = 0.0, sn = 0.0;
    for (size_t i = 0; i < n; ++i) {
        dn += std::norm(Dx[i] - source[i]);
        sn += std::norm(source[i]);
    }
    r.iterations = st.iterations;
    r.normal_residual = st.true_relres;
    r.dirac_residual = std::sqrt(dn) / std::max(std::sqrt(sn), 1.0e-300);
    if (!std::isfinite(r.normal_residual) || r.normal_residual > 20.0 * tolerance) {
        std::ostringstream msg;
        msg.setf(std::ios::scientific);
        msg << "CGNR failed: iterations=" << r.iterations << ", normal residual=" << r.normal_residual
            << ", D residual=" << r.dirac_residual;
        throw std::runtime_error(msg.str());
    }
    return r;
}

// ------------------------------------------------------------- HMC trajectories
struct HmcResult {
    bool accepted;
    double dH;
    double acceptance_probability;
    double plaquette;
    int solves;
    int cg_iterations;
};

// The callers' random stream: std::mt19937_64 + libstdc++ normal/uniform, drawing in
// the reference's order (compiled by g++ inside libdwfhip).
class HostRng {
public:
    explicit HostRng(unsigned long long seed) : r_(dwf_rng_create(seed)) {}
    ~HostRng() { dwf_rng_destroy(r_); }
    HostRng(const HostRng&) = delete;
    HostRng& operator=(const HostRng&) = delete;
    dwf_rng* get() const { return r_; }

private:
    dwf_rng* r_;
};

// One trajectory from `theta` (domain_wall/generate_configs.cpp:52-169 for the DWF
// integrator, wilson/include/gauge.hpp:132-180 for the Wilson one); theta becomes the
// configuration after the accept step.
inline HmcResult hmc_trajectory(Context& c, Gauge& theta, HostRng& rng, const dwf_hmc_params& p)
{
    c.use_gauge(theta);
    const size_t ndof = c.ndof();  // complex components of a spinor
    std::vector<double> chi(2 * ndof), mom(c.ngauge());
    if (p.integrator == DWF_HMC_DOMAIN_WALL)
        dwf_rng_dwf_heatbath(rng.get(), (int)ndof, (int)c.ngauge(), chi.data(), mom.data());
    else
        dwf_rng_wilson_heatbath(rng.get(), (int)c.ngauge(), (int)ndof, mom.data(), chi.data());
    const double u = dwf_rng_uniform(rng.get());
    dwf_hmc_result res{};
    check(dwf_hmc_trajectory(c.get(), &p, chi.data(), mom.data(), u, &res));
    check(dwf_get_gauge(c.get(), theta.data()));
    c.adopt_gauge(theta);  // the device already holds theta
    return HmcResult{res.accepted != 0, res.delta_h, res.acceptance_probability, res.plaquette, res.solves,
                     res.cg_iterations};
}

}  // namespace dwfhip
