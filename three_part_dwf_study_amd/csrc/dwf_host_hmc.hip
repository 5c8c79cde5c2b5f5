// dwf_host_hmc.hip -- host driver of the pseudofermion force, gauge action and HMC
// trajectory (C ABI in include/dwf_hip.h).  Angles, momenta, forces and actions stay
// on the device in FP64; the host reads back a handful of FP64 scalars per solve.

#include "../../include/dwf_hip.h"
#include "dwf_context.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <string>

using namespace dwfk;

namespace {

int guard(const std::function<int()>& f)
{
    try {
        return f();
    } catch (const DwfError& e) {
        dwf_set_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        dwf_set_error(e.what());
        return DWF_ERR_ARG;
    }
}

struct Dev {
    dwf_ctx* c;
    size_t ncx;   // complex components of one spinor
    int ng;       // gauge links
    double* scal;
    double* rpart;
    explicit Dev(dwf_ctx* ctx) : c(ctx)
    {
        ncx = (size_t)c->L.V5() * 2;
        ng = 2 * c->L.V4;
        scal = (double*)c->buf(B_SCAL, S_COUNT * sizeof(double));
        // two sets of up to 1024 block partials (reduce_blocks caps the grids)
        rpart = (double*)c->buf(B_RPART, 2048 * sizeof(double));
    }
    double* vec(int id) { return (double*)c->buf(id, ncx * 2 * sizeof(double)); }
    double* gauge(int id) { return (double*)c->buf(id, ng * sizeof(double)); }
    void read_scalars(double* out)
    {
        c->check(hipMemcpyAsync(out, scal, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                 "scalar readback");
        c->check(hipStreamSynchronize(c->stream), "scalar sync");
    }
};

// Gauge action, force (nullable) and plaquette of the context's current field.
void gauge_dev(Dev& d, double beta, double* force)
{
    launch_gauge(d.c->L, d.c->theta, beta, force, d.rpart, d.rpart + 1024, d.scal, S_GAUGE, S_PLAQ, d.c->stream);
    d.c->check_launch("gauge action");
}

struct PfOut {
    int iterations = 0;
    dwf_solve_stats stats{};
    int rc = DWF_OK;
};

PfOut check_cg(dwf_ctx* c, int method, double rtol);

// eta = (D D^dag)^-1 phi (guess nullable, may alias eta), S -> scal[S_PF], force.
PfOut pf_dev(Dev& d, int method, const double* phi, const double* guess, double rtol, int maxiter, double* eta,
             double* force)
{
    dwf_ctx* c = d.c;
    PfOut o;
    if (method == DWF_SOLVE_LEARNED || method == DWF_SOLVE_EO_MIXED) {
        // learned preconditioner, or two even-odd Dirac solves per correction pass, each
        // with its certified CG fallback (certified on D D^dag or failed)
        const int rc = dwf_solve_normal_impl(c, DWF_OP_DDDAG, method, phi, guess, rtol, maxiter, 1, eta, &o.stats);
        o.iterations = o.stats.iterations + o.stats.fp64_iterations;
        if (rc != DWF_OK || !o.stats.converged) {
            o.rc = DWF_ERR_NOT_CONVERGED;  // pseudofermion.hpp:39-41
            dwf_set_error("Pseudofermion CG failed to converge");
            return o;
        }
    } else {
        dwf_solve_device_impl(c, DWF_OP_DDDAG, method, phi, guess, rtol, maxiter, 1, eta);
        o = check_cg(c, method, rtol);
        if (o.rc != DWF_OK) return o;
    }
    launch_redot(phi, eta, (long long)d.ncx, d.rpart, d.scal, S_PF, c->stream);  // S = Re phi^dag eta
    double* xi = d.vec(B_XI);
    dwf_apply_device_impl(c, DWF_OP_DDAG, DWF_PREC_FP64, eta, xi, 1);
    launch_pf_force(c->L, c->l64, eta, xi, force, c->stream);
    c->check_launch("pseudofermion force");
    return o;
}

// State of the CG64 / MIXED solve just run: stats, and the reference's failure rules.
PfOut check_cg(dwf_ctx* c, int method, double rtol)
{
    PfOut o;
    const CgState& S = c->st_host[0];
    const bool done = (S.flags[c->last_slot] & F_DONE) && S.status == 0;
    o.stats.iterations = S.iters;
    o.stats.converged = done ? 1 : 0;
    o.stats.reliable_updates = S.reliable;
    o.stats.relres = S.relres;
    o.stats.true_relres = S.true_relres;
    o.stats.fp64_iterations = method == DWF_SOLVE_CG64 ? S.iters : 0;
    o.iterations = S.iters;
    if (S.status == 4) throw DwfError(DWF_ERR_ZERO_DENOM, "CG encountered a zero denominator");
    if (c->kind == 0) {
        if (!done) {
            o.rc = DWF_ERR_NOT_CONVERGED;  // pseudofermion.hpp:39-41
            dwf_set_error("Pseudofermion CG failed to converge");
            return o;
        }
    } else if (!std::isfinite(S.true_relres) || S.true_relres > 20.0 * rtol) {
        std::ostringstream msg;  // wilson/include/cg_solver.hpp:69-76
        msg.setf(std::ios::scientific);
        msg << "pseudofermion CG failed: iterations=" << S.iters << ", residual=" << S.true_relres;
        o.rc = DWF_ERR_NOT_CONVERGED;
        dwf_set_error(msg.str());
        return o;
    }
    return o;
}

void upload(dwf_ctx* c, void* dst, const void* src, size_t bytes)
{
    c->check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
}
void download(dwf_ctx* c, void* dst, const void* src, size_t bytes)
{
    c->check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "download");
    c->check(hipStreamSynchronize(c->stream), "download sync");
}

void require(dwf_ctx* c)
{
    if (!c) throw DwfError(DWF_ERR_ARG, "null context");
    if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "no gauge field set");
    c->require_single_config("the force, gauge and HMC entry points");
    c->bind_rhs(1);
}

}  // namespace

extern "C" {

int dwf_pf_action_force(dwf_ctx* c, int method, const double* phi, const double* eta_guess, double rtol,
                        int maxiter, double* action, double* force, double* eta, dwf_solve_stats* stats)
{
    return guard([&]() -> int {
        require(c);
        if (!phi || !action || !force || maxiter < 0 ||
            (method != DWF_SOLVE_CG64 && method != DWF_SOLVE_MIXED && method != DWF_SOLVE_EO_MIXED &&
             method != DWF_SOLVE_LEARNED))
            throw DwfError(DWF_ERR_ARG, "dwf_pf_action_force: bad argument");
        Dev d(c);
        double* dphi = d.vec(B_PHI);
        double* deta = d.vec(B_ETA);
        double* dF = d.gauge(B_FPF);
        upload(c, dphi, phi, d.ncx * 16);
        if (eta_guess) upload(c, deta, eta_guess, d.ncx * 16);
        PfOut o = pf_dev(d, method, dphi, eta_guess ? deta : nullptr, rtol, maxiter, deta, dF);
        if (stats) *stats = o.stats;
        if (eta) download(c, eta, deta, d.ncx * 16);
        if (o.rc != DWF_OK) return o.rc;
        double sc[S_COUNT];
        d.read_scalars(sc);
        *action = sc[S_PF];
        download(c, force, dF, d.ng * sizeof(double));
        return DWF_OK;
    });
}

int dwf_force_from_eta(dwf_ctx* c, const double* eta, double* force)
{
    return guard([&]() -> int {
        require(c);
        if (!eta || !force) throw DwfError(DWF_ERR_ARG, "dwf_force_from_eta: bad argument");
        Dev d(c);
        double* deta = d.vec(B_ETA);
        double* xi = d.vec(B_XI);
        double* dF = d.gauge(B_FPF);
        upload(c, deta, eta, d.ncx * 16);
        dwf_apply_device_impl(c, DWF_OP_DDAG, DWF_PREC_FP64, deta, xi, 1);
        launch_pf_force(c->L, c->l64, deta, xi, dF, c->stream);
        c->check_launch("pseudofermion force");
        download(c, force, dF, d.ng * sizeof(double));
        return DWF_OK;
    });
}

int dwf_gauge_action_force(dwf_ctx* c, double beta, double* action, double* force)
{
    return guard([&]() -> int {
        require(c);
        if (!action) throw DwfError(DWF_ERR_ARG, "dwf_gauge_action_force: bad argument");
        Dev d(c);
        double* dF = force ? d.gauge(B_FG) : nullptr;
        gauge_dev(d, beta, dF);
        double sc[S_COUNT];
        d.read_scalars(sc);
        *action = sc[S_GAUGE];
        if (force) download(c, force, dF, d.ng * sizeof(double));
        return DWF_OK;
    });
}

int dwf_average_plaquette(dwf_ctx* c, double* plaquette)
{
    return guard([&]() -> int {
        require(c);
        if (!plaquette) throw DwfError(DWF_ERR_ARG, "dwf_average_plaquette: bad argument");
        Dev d(c);
        gauge_dev(d, 1.0, nullptr);
        double sc[S_COUNT];
        d.read_scalars(sc);
        *plaquette = sc[S_PLAQ];
        return DWF_OK;
    });
}

int dwf_get_gauge(dwf_ctx* c, double* theta)
{
    return guard([&]() -> int {
        require(c);
        if (!theta) throw DwfError(DWF_ERR_ARG, "dwf_get_gauge: bad argument");
        download(c, theta, c->theta, (size_t)2 * c->L.V4 * sizeof(double));
        return DWF_OK;
    });
}

int dwf_hmc_trajectory(dwf_ctx* c, const dwf_hmc_params* P, const double* chi, const double* momentum, double u,
                       dwf_hmc_result* res)
{
    return guard([&]() -> int {
        require(c);
        if (!P || !chi || !momentum || !res || P->steps < 0 || P->maxiter < 0 ||
            (P->integrator != DWF_HMC_DOMAIN_WALL && P->integrator != DWF_HMC_WILSON) ||
            (P->method != DWF_SOLVE_CG64 && P->method != DWF_SOLVE_MIXED && P->method != DWF_SOLVE_EO_MIXED &&
             P->method != DWF_SOLVE_LEARNED))
            throw DwfError(DWF_ERR_ARG, "dwf_hmc_trajectory: bad argument");
        Dev d(c);
        const int ng = d.ng;
        const bool dwf_int = P->integrator == DWF_HMC_DOMAIN_WALL;
        const double eps = P->epsilon;
        double* phi = d.vec(B_PHI);
        double* eta = d.vec(B_ETA);
        double* mom = d.gauge(B_MOM);
        double* Fg = d.gauge(B_FG);
        double* Fpf = d.gauge(B_FPF);
        double* th0 = d.gauge(B_THETA0);
        double* chi_d = d.vec(B_XI);  // free until the first force pass
        *res = dwf_hmc_result{};

        c->check(hipMemcpyAsync(th0, c->theta, ng * sizeof(double), hipMemcpyDeviceToDevice, c->stream), "theta0");
        upload(c, chi_d, chi, d.ncx * 16);
        upload(c, mom, momentum, ng * sizeof(double));
        dwf_apply_device_impl(c, DWF_OP_D, DWF_PREC_FP64, chi_d, phi, 1);  // phi = D chi

        auto restore = [&] {
            c->check(hipMemcpyAsync(c->theta, th0, ng * sizeof(double), hipMemcpyDeviceToDevice, c->stream), "restore");
            launch_theta(c->theta, mom, ng, 0.0, 0, c->l64, c->l32, c->stream);  // repack links
        };
        auto forces = [&](const double* guess) -> bool {
            gauge_dev(d, P->beta, Fg);
            PfOut o = pf_dev(d, P->method, phi, guess, P->rtol, P->maxiter, eta, Fpf);
            res->solves += 1;
            res->cg_iterations += o.iterations;
            return o.rc == DWF_OK;
        };
        double sc[S_COUNT];
        if (!forces(nullptr)) {
            restore();
            return static_cast<int>(DWF_ERR_NOT_CONVERGED);
        }
        launch_mom(mom, Fg, Fpf, ng, 0.0, d.rpart, d.scal, S_KIN, c->stream);  // K0
        d.read_scalars(sc);
        const double H0 = sc[S_KIN] + sc[S_GAUGE] + sc[S_PF];
        double Sg = sc[S_GAUGE], Spf = sc[S_PF];

        launch_mom(mom, Fg, Fpf, ng, 0.5 * eps, nullptr, nullptr, 0, c->stream);
        for (int step = 0; step < P->steps; ++step) {
            launch_theta(c->theta, mom, ng, eps, dwf_int ? 0 : 1, c->l64, c->l32, c->stream);
            // DWF: chronological guess (generate_configs.cpp:119-126); Wilson: zero guess.
            if (!forces(dwf_int ? eta : nullptr)) {
                restore();
                return static_cast<int>(DWF_ERR_NOT_CONVERGED);
            }
            const bool last = step == P->steps - 1;
            if (dwf_int) {
                if (!last) launch_mom(mom, Fg, Fpf, ng, eps, nullptr, nullptr, 0, c->stream);
            } else {
                launch_mom(mom, Fg, Fpf, ng, (last ? 0.5 : 1.0) * eps, nullptr, nullptr, 0, c->stream);
            }
        }
        if (dwf_int) {
            launch_mom(mom, Fg, Fpf, ng, 0.5 * eps, d.rpart, d.scal, S_KIN, c->stream);
        } else {
            // The reference recomputes both actions at the final angles with a fresh
            // zero-guess solve (gauge.hpp:171-172): the last drift's solve is exactly that
            // solve, so its actions are reused.  With no steps, the start values stand.
            launch_mom(mom, Fg, Fpf, ng, 0.0, d.rpart, d.scal, S_KIN, c->stream);
        }
        d.read_scalars(sc);
        if (P->steps > 0) {
            Sg = sc[S_GAUGE];
            Spf = sc[S_PF];
        }
        const double H = sc[S_KIN] + Sg + Spf;
        const double dH = H - H0;
        const double pacc = std::min(std::exp(-dH), 1.0);
        const bool acc = u < pacc;
        if (!acc) restore();
        gauge_dev(d, 1.0, nullptr);
        d.read_scalars(sc);
        res->accepted = acc ? 1 : 0;
        res->acceptance_probability = pacc;
        res->delta_h = dH;
        res->h_initial = H0;
        res->h_final = H;
        res->plaquette = sc[S_PLAQ];
        return static_cast<int>(DWF_OK);
    });
}

}  // extern "C"
