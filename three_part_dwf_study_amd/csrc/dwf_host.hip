// dwf_host.hip -- host driver and C ABI of libdwfhip.so (declared in include/dwf_hip.h).
//
// A context owns one HIP stream, the gauge field (FP64 angles and cos/sin link arrays
// in FP64 and FP32) and a grow-only device workspace.  Solvers keep every vector and
// every per-right-hand-side scalar on the device; the host only launches kernels and
// reads the per-RHS flags once per chunk of iterations to decide when to stop.

#include "../../include/dwf_hip.h"
#include "dwf_context.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace dwfk;

namespace {
thread_local std::string g_last_error;
}

void dwf_set_error(const std::string& msg) { g_last_error = msg; }

int dwf_hip_fail(hipError_t e, const char* what)
{
    dwf_set_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
    return DWF_ERR_HIP;
}

dwf_ctx::~dwf_ctx()
{
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& b : bufs)
        if (b.p) (void)hipFree(b.p);
    if (theta) (void)hipFree(theta);
    if (l64) (void)hipFree(l64);
    if (l32) (void)hipFree(l32);
    if (st_host) (void)hipHostFree(st_host);
    if (stream) (void)hipStreamDestroy(stream);
}

void* dwf_ctx::buf(int id, size_t bytes)
{
    Buf& b = bufs[id];
    if (b.n < bytes) {
        if (b.p) {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(b.p);
            b.p = nullptr;
            b.n = 0;
        }
        const hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            b.p = nullptr;
            throw DwfError(DWF_ERR_HIP, std::string("hipMalloc of ") + std::to_string(bytes) +
                                            " bytes failed: " + hipGetErrorString(e));
        }
        b.n = bytes;
    }
    return b.p;
}

CgState* dwf_ctx::host_states(int nrhs)
{
    if (st_host_n < nrhs) {
        if (st_host) (void)hipHostFree(st_host);
        st_host = nullptr;
        check(hipHostMalloc((void**)&st_host, sizeof(CgState) * nrhs, hipHostMallocDefault),
              "hipHostMalloc");
        st_host_n = nrhs;
    }
    return st_host;
}

void dwf_ctx::check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw DwfError(DWF_ERR_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

void dwf_ctx::check_launch(const char* what) { check(hipGetLastError(), what); }

// ------------------------------------------------------------------- operators
namespace {

template <typename R> HopArgs<R> hop_base(const dwf_ctx* c, const HopCfg& cfg)
{
    HopArgs<R> a{};
    a.L = c->L;
    a.cfg = cfg;
    a.links = sizeof(R) == 8 ? (const void*)c->l64 : (const void*)c->l32;
    return a;
}

// out = op(in) on device buffers; tmp is needed for the normal operators.
template <typename R> void apply_op(dwf_ctx* c, int op, const R* in, R* out, R* tmp, int nrhs)
{
    const HopCfg cfg = hop_cfg(c->L, nrhs, c->chunks_override, c->xcd_remap);
    HopArgs<R> a = hop_base<R>(c, cfg);
    switch (op) {
        case DWF_OP_D:
        case DWF_OP_DDAG:
            a.in = in;
            a.out = out;
            launch_hop<R>(op == DWF_OP_DDAG, EPI_STORE, a, nrhs, c->stream);
            break;
        default: {
            const bool first_dag = (op == DWF_OP_DDDAG);
            a.in = in;
            a.out = tmp;
            launch_hop<R>(first_dag, EPI_STORE, a, nrhs, c->stream);
            a.in = tmp;
            a.out = out;
            launch_hop<R>(!first_dag, EPI_STORE, a, nrhs, c->stream);
        }
    }
    c->check_launch("stencil launch");
}

}  // namespace

void dwf_apply_device_impl(dwf_ctx* c, int op, int prec, const void* in, void* out, int nrhs)
{
    const size_t n = (size_t)c->L.V5() * 4 * nrhs;
    if (prec == DWF_PREC_FP64) {
        double* tmp = op >= DWF_OP_DDDAG ? (double*)c->buf(B_APPLY_TMP, n * sizeof(double)) : nullptr;
        apply_op<double>(c, op, (const double*)in, (double*)out, tmp, nrhs);
    } else {
        float* tmp = op >= DWF_OP_DDDAG ? (float*)c->buf(B_APPLY_TMP, n * sizeof(float)) : nullptr;
        apply_op<float>(c, op, (const float*)in, (float*)out, tmp, nrhs);
    }
}

// ----------------------------------------------------------------------- solvers
namespace {

struct SolveBufs {
    double *x, *b;
    double *r64, *t64, *p64;
    float *r32, *t32, *p32, *x32;
    double *partA, *partB, *partR;
    CgState* st;
};

template <typename R>
void cg_iteration(dwf_ctx* c, const HopCfg& cfg, bool first_dag, R* p, R* t, R* x, R* r,
                  const SolveBufs& B, int slot, int mode, int nrhs)
{
    HopArgs<R> h = hop_base<R>(c, cfg);
    h.in = p;
    h.out = t;
    h.part = B.partA;
    h.st = B.st;
    h.slot = slot;
    h.skip = F_DONE | F_FROZEN;
    launch_hop<R>(first_dag, EPI_NORM, h, nrhs, c->stream);
    HopArgs<R> g = hop_base<R>(c, cfg);
    g.in = t;
    g.x = x;
    g.r = r;
    g.p = p;
    g.partA = B.partA;
    g.part = B.partB;
    g.st = B.st;
    g.slot = slot;
    g.skip = F_DONE | F_FROZEN;
    launch_hop<R>(!first_dag, EPI_CGUPD, g, nrhs, c->stream);
    launch_cg_p<R>(c->L, cfg, r, p, B.partA, B.partB, B.st, slot, mode, nrhs, c->stream);
}

// r64 = b - A x in FP64 (+ optional copies), partials |r|^2, |b|^2 into partR.
void fp64_residual(dwf_ctx* c, const HopCfg& cfg, bool first_dag, const SolveBufs& B, double* p64,
                   float* r32, float* p32, const CgState* st, int slot, int skip, int nrhs)
{
    HopArgs<double> a = hop_base<double>(c, cfg);
    a.in = B.x;
    a.out = B.t64;
    a.st = st;
    a.slot = slot;
    a.skip = skip;
    launch_hop<double>(first_dag, EPI_STORE, a, nrhs, c->stream);
    HopArgs<double> e = hop_base<double>(c, cfg);
    e.in = B.t64;
    e.b = B.b;
    e.r64 = B.r64;
    e.p64 = p64;
    e.r32 = r32;
    e.p32 = p32;
    e.part = B.partR;
    e.st = st;
    e.slot = slot;
    e.skip = skip;
    launch_hop<double>(!first_dag, EPI_RESID, e, nrhs, c->stream);
}

void read_states(dwf_ctx* c, const CgState* st, int nrhs)
{
    CgState* h = c->host_states(nrhs);
    c->check(hipMemcpyAsync(h, st, sizeof(CgState) * nrhs, hipMemcpyDeviceToHost, c->stream), "state readback");
    c->check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

bool all_done(const CgState* h, int slot, int nrhs)
{
    for (int i = 0; i < nrhs; ++i)
        if (!(h[i].flags[slot] & F_DONE)) return false;
    return true;
}

}  // namespace

// Solve A x = b on device buffers.  Returns the per-RHS states in the context's host
// state mirror (c->st_host).
int dwf_solve_device_impl(dwf_ctx* c, int normal, int method, const double* b, const double* x0,
                          double rtol, int maxiter, int nrhs, double* x)
{
    const Lat& L = c->L;
    const long long n4 = L.V5() * 4;
    const size_t nv = (size_t)n4 * nrhs;
    const bool first_dag = (normal == DWF_OP_DDDAG);
    const bool mixed = (method == DWF_SOLVE_MIXED);
    const HopCfg cfg = hop_cfg(L, nrhs, c->chunks_override, c->xcd_remap);
    const size_t npart = (size_t)nrhs * cfg.nblk() * 2;

    SolveBufs B{};
    B.x = x;
    B.b = const_cast<double*>(b);
    B.r64 = (double*)c->buf(B_R64, nv * 8);
    B.t64 = (double*)c->buf(B_T64, nv * 8);
    B.partA = (double*)c->buf(B_PA, npart * 8);
    B.partB = (double*)c->buf(B_PB, npart * 8);
    B.partR = (double*)c->buf(B_PR, npart * 8);
    B.st = (CgState*)c->buf(B_STATE, sizeof(CgState) * nrhs);
    if (mixed) {
        B.r32 = (float*)c->buf(B_R32, nv * 4);
        B.t32 = (float*)c->buf(B_T32, nv * 4);
        B.p32 = (float*)c->buf(B_P32, nv * 4);
        B.x32 = (float*)c->buf(B_X32, nv * 4);
    } else {
        B.p64 = (double*)c->buf(B_P64, nv * 8);
    }

    if (x0) {
        if (x0 != x) c->check(hipMemcpyAsync(x, x0, nv * 8, hipMemcpyDeviceToDevice, c->stream), "copy x0");
    } else {
        launch_zero<double>(x, (long long)nv, c->stream);
    }
    // r = b - A x0 (the reference applies A even to a zero guess, cg_solver.hpp:24-27)
    fp64_residual(c, cfg, first_dag, B, mixed ? nullptr : B.p64, mixed ? B.r32 : nullptr,
                  mixed ? B.p32 : nullptr, nullptr, 0, 0, nrhs);
    const double delta = c->reliable_delta;
    launch_cg_start(cfg, B.partR, B.st, rtol, delta, nrhs, c->stream);
    if (mixed) launch_zero<float>(B.x32, (long long)nv, c->stream);
    c->check_launch("solver start");

    int it = 0;
    int chunk = mixed ? 16 : 32;
    read_states(c, B.st, nrhs);
    std::vector<double> last_true(nrhs);
    std::vector<int> last_iters(nrhs);
    for (int i = 0; i < nrhs; ++i) {
        last_true[i] = c->st_host[i].true_relres;
        last_iters[i] = 0;
    }
    while (!all_done(c->st_host, it & 1, nrhs)) {
        const int n = std::min(chunk, maxiter - it);
        if (n <= 0) break;
        for (int j = 0; j < n; ++j, ++it) {
            if (mixed)
                cg_iteration<float>(c, cfg, first_dag, B.p32, B.t32, B.x32, B.r32, B, it & 1, MODE_MIXED, nrhs);
            else
                cg_iteration<double>(c, cfg, first_dag, B.p64, B.t64, B.x, B.r64, B, it & 1, MODE_CG64, nrhs);
        }
        c->check_launch("CG iteration");
        if (mixed) {
            // Reliable update: x64 += x32, FP64 true residual, certify or restart the cycle.
            const int slot = it & 1;
            launch_accum_x(n4, B.x, B.x32, B.st, slot, nrhs, c->stream);
            fp64_residual(c, cfg, first_dag, B, nullptr, B.r32, nullptr, B.st, slot, F_DONE, nrhs);
            launch_ru_commit(cfg, B.partR, B.st, slot, nrhs, c->stream);
            c->check_launch("reliable update");
            read_states(c, B.st, nrhs);
            // Next cycle length: iterations the fastest active RHS needs to shrink its
            // residual by delta (or to reach the tolerance), from the last cycle's rate.
            double best = 1e30;
            for (int i = 0; i < nrhs; ++i) {
                const CgState& S = c->st_host[i];
                if (S.flags[slot] & F_DONE) continue;
                const int di = S.iters - last_iters[i];
                if (di > 0 && S.true_relres > 0.0 && last_true[i] > 0.0 && S.true_relres < last_true[i]) {
                    const double rate = std::log(S.true_relres / last_true[i]) / di;  // < 0
                    const double need_delta = std::log(delta) / rate;
                    const double need_tol = std::log(rtol / S.true_relres) / rate + 2.0;
                    best = std::min(best, std::min(need_delta, need_tol));
                }
                last_true[i] = S.true_relres;
                last_iters[i] = S.iters;
            }
            if (best < 1e29) chunk = std::max(8, std::min(1024, (int)std::ceil(best)));
        } else {
            read_states(c, B.st, nrhs);
            chunk = 32;
        }
    }

    // FP64 true residual of the final x for the statistics (and certification).
    fp64_residual(c, cfg, first_dag, B, nullptr, nullptr, nullptr, nullptr, 0, 0, nrhs);
    launch_true_resid(cfg, B.partR, B.st, nrhs, c->stream);
    c->check_launch("true residual");
    read_states(c, B.st, nrhs);
    c->last_slot = it & 1;
    return DWF_OK;
}

// ------------------------------------------------------------------------- C ABI
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

int create(dwf_ctx** out, const Lat& L, int kind)
{
    if (!out) {
        dwf_set_error("null context pointer");
        return DWF_ERR_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) {
        dwf_set_error("no HIP device available: libdwfhip needs an MI355X (gfx950) GPU");
        return DWF_ERR_HIP;
    }
    dwf_ctx* c = new dwf_ctx();
    c->L = L;
    c->kind = kind;
    e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    const size_t nl = (size_t)2 * L.V4;
    if (e == hipSuccess) e = hipMalloc(&c->theta, nl * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->l64, nl * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc(&c->l32, nl * sizeof(float2));
    if (e != hipSuccess) {
        delete c;
        return dwf_hip_fail(e, "context creation");
    }
    *out = c;
    return DWF_OK;
}

}  // namespace

extern "C" {

int dwf_ctx_create_dwf(dwf_ctx** out, int Nt, int Nx, int Ls, double M5, double mf)
{
    if (Nt < 1 || Nx < 1 || Ls < 1 || (long long)Nt * Nx * Ls * 2 > (1LL << 31) - 1) {
        dwf_set_error("dwf_ctx_create_dwf: lattice extents must be positive and Ndof < 2^31");
        return DWF_ERR_ARG;
    }
    Lat L{};
    L.Nt = Nt;
    L.Nx = Nx;
    L.Ls = Ls;
    L.V4 = Nt * Nx;
    L.d5 = 1;
    L.diag = 3.0 - M5;
    L.mf = mf;
    L.bct = -1.0;  // antiperiodic in t (dwf_operator.hpp:29-30)
    L.bcx = 1.0;   // periodic in x
    return create(out, L, 0);
}

int dwf_ctx_create_wilson(dwf_ctx** out, int Nt, int Nx, double m0, double bc_t, double bc_x)
{
    if (Nt < 1 || Nx < 1) {
        dwf_set_error("dwf_ctx_create_wilson: lattice extents must be positive");
        return DWF_ERR_ARG;
    }
    Lat L{};
    L.Nt = Nt;
    L.Nx = Nx;
    L.Ls = 1;
    L.V4 = Nt * Nx;
    L.d5 = 0;
    L.diag = 2.0 + m0;
    L.mf = 0.0;
    L.bct = bc_t;
    L.bcx = bc_x;
    return create(out, L, 1);
}

int dwf_ctx_destroy(dwf_ctx* ctx)
{
    delete ctx;
    return DWF_OK;
}

void* dwf_ctx_stream(dwf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int dwf_ctx_tune(dwf_ctx* ctx, int param, double value)
{
    if (!ctx) return DWF_ERR_ARG;
    switch (param) {
        case DWF_TUNE_SCHUNKS:
            if (value < 0 || value > ctx->L.Ls) break;
            ctx->chunks_override = (int)value;
            return DWF_OK;
        case DWF_TUNE_XCD_REMAP:
            ctx->xcd_remap = value != 0.0 ? 1 : 0;
            return DWF_OK;
        case DWF_TUNE_RELIABLE_DELTA:
            if (!(value > 0.0 && value < 1.0)) break;
            ctx->reliable_delta = value;
            return DWF_OK;
        default:
            break;
    }
    dwf_set_error("dwf_ctx_tune: bad parameter or value");
    return DWF_ERR_ARG;
}

int dwf_set_gauge_device(dwf_ctx* c, const double* theta_device)
{
    return guard([&] {
        if (!c || !theta_device) throw DwfError(DWF_ERR_ARG, "dwf_set_gauge: null argument");
        const size_t nl = (size_t)2 * c->L.V4;
        if (theta_device != c->theta)
            c->check(hipMemcpyAsync(c->theta, theta_device, nl * 8, hipMemcpyDeviceToDevice, c->stream), "gauge copy");
        launch_pack_links(c->L, c->theta, c->l64, c->l32, c->stream);
        c->check_launch("pack links");
        c->have_gauge = true;
        return DWF_OK;
    });
}

int dwf_set_gauge(dwf_ctx* c, const double* theta)
{
    return guard([&] {
        if (!c || !theta) throw DwfError(DWF_ERR_ARG, "dwf_set_gauge: null argument");
        const size_t nl = (size_t)2 * c->L.V4;
        c->check(hipMemcpyAsync(c->theta, theta, nl * 8, hipMemcpyHostToDevice, c->stream), "gauge upload");
        launch_pack_links(c->L, c->theta, c->l64, c->l32, c->stream);
        c->check_launch("pack links");
        c->check(hipStreamSynchronize(c->stream), "gauge upload sync");
        c->have_gauge = true;
        return DWF_OK;
    });
}

int dwf_apply_device(dwf_ctx* c, int op, int precision, const void* in, void* out, int nrhs)
{
    return guard([&] {
        if (!c || !in || !out || nrhs < 1 || op < DWF_OP_D || op > DWF_OP_DDAGD ||
            (precision != DWF_PREC_FP64 && precision != DWF_PREC_FP32) || in == out)
            throw DwfError(DWF_ERR_ARG, "dwf_apply_device: bad argument");
        if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_apply_device: no gauge field set");
        dwf_apply_device_impl(c, op, precision, in, out, nrhs);
        return DWF_OK;
    });
}

int dwf_apply(dwf_ctx* c, int op, const double* in, double* out, int nrhs)
{
    return guard([&] {
        if (!c || !in || !out || nrhs < 1 || op < DWF_OP_D || op > DWF_OP_DDAGD)
            throw DwfError(DWF_ERR_ARG, "dwf_apply: bad argument");
        if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_apply: no gauge field set");
        const size_t n = (size_t)c->L.V5() * 4 * nrhs;
        double* din = (double*)c->buf(B_HOST_IN, n * 8);
        double* dout = (double*)c->buf(B_HOST_OUT, n * 8);
        c->check(hipMemcpyAsync(din, in, n * 8, hipMemcpyHostToDevice, c->stream), "apply upload");
        dwf_apply_device_impl(c, op, DWF_PREC_FP64, din, dout, nrhs);
        c->check(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, c->stream), "apply download");
        c->check(hipStreamSynchronize(c->stream), "apply sync");
        return DWF_OK;
    });
}

}  // extern "C"

namespace {

int finish_stats(dwf_ctx* c, int method, int nrhs, dwf_solve_stats* stats, int fallback_iters_from)
{
    (void)fallback_iters_from;
    int rc = DWF_OK;
    for (int i = 0; i < nrhs; ++i) {
        const CgState& S = c->st_host[i];
        dwf_solve_stats s{};
        const bool done = (S.flags[c->last_slot] & F_DONE) != 0;
        s.iterations = S.iters;
        s.reliable_updates = S.reliable;
        s.relres = S.relres;
        s.true_relres = S.true_relres;
        s.fp64_iterations = method == DWF_SOLVE_CG64 ? S.iters : 0;
        s.converged = done && S.status == 0 ? 1 : 0;
        if (S.status == 4) {
            dwf_set_error("CG encountered a zero denominator");
            rc = DWF_ERR_ZERO_DENOM;
        } else if (!s.converged && rc == DWF_OK) {
            dwf_set_error("CG did not converge within maxiter");
            rc = DWF_ERR_NOT_CONVERGED;
        }
        if (stats) stats[i] = s;
    }
    return rc;
}

int solve_checked(dwf_ctx* c, int normal, int method, const double* b, const double* x0, double rtol,
                  int maxiter, int nrhs, double* x, dwf_solve_stats* stats)
{
    if (!c || !b || !x || nrhs < 1 || maxiter < 0 || !(rtol >= 0.0) ||
        (normal != DWF_OP_DDDAG && normal != DWF_OP_DDAGD) ||
        (method != DWF_SOLVE_CG64 && method != DWF_SOLVE_MIXED))
        throw DwfError(DWF_ERR_ARG, "dwf_solve: bad argument");
    if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_solve: no gauge field set");
    dwf_solve_device_impl(c, normal, method, b, x0, rtol, maxiter, nrhs, x);
    int rc = finish_stats(c, method, nrhs, stats, 0);
    if (method == DWF_SOLVE_MIXED && rc == DWF_ERR_NOT_CONVERGED) {
        // The FP32 cycle stalled: plain FP64 CG from the current iterate decides.
        std::vector<dwf_solve_stats> first(nrhs);
        finish_stats(c, method, nrhs, first.data(), 0);
        dwf_solve_device_impl(c, normal, DWF_SOLVE_CG64, b, x, rtol, maxiter, nrhs, x);
        rc = finish_stats(c, DWF_SOLVE_CG64, nrhs, stats, 0);
        if (stats)
            for (int i = 0; i < nrhs; ++i) {
                stats[i].fp64_iterations = stats[i].iterations;
                stats[i].iterations = first[i].iterations;
                stats[i].reliable_updates = first[i].reliable_updates;
            }
        if (rc == DWF_ERR_NOT_CONVERGED) rc = DWF_ERR_CERTIFY;
    }
    return rc;
}

}  // namespace

extern "C" {

int dwf_solve_device(dwf_ctx* c, int normal_op, int method, const double* b_device,
                     const double* x0_device, double rtol, int maxiter, int nrhs, double* x_device,
                     dwf_solve_stats* stats)
{
    return guard([&] {
        return solve_checked(c, normal_op, method, b_device, x0_device, rtol, maxiter, nrhs, x_device, stats);
    });
}

int dwf_solve(dwf_ctx* c, int normal_op, int method, const double* b, const double* x0, double rtol,
              int maxiter, int nrhs, double* x, dwf_solve_stats* stats)
{
    return guard([&] {
        if (!c || !b || !x || nrhs < 1) throw DwfError(DWF_ERR_ARG, "dwf_solve: bad argument");
        const size_t n = (size_t)c->L.V5() * 4 * nrhs;
        double* db = (double*)c->buf(B_HOST_IN, n * 8);
        double* dx = (double*)c->buf(B_HOST_OUT, n * 8);
        double* dx0 = nullptr;
        c->check(hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, c->stream), "solve upload b");
        if (x0) {
            dx0 = dx;
            c->check(hipMemcpyAsync(dx, x0, n * 8, hipMemcpyHostToDevice, c->stream), "solve upload x0");
        }
        const int rc = solve_checked(c, normal_op, method, db, dx0, rtol, maxiter, nrhs, dx, stats);
        if (rc == DWF_OK || rc == DWF_ERR_NOT_CONVERGED || rc == DWF_ERR_CERTIFY) {
            c->check(hipMemcpyAsync(x, dx, n * 8, hipMemcpyDeviceToHost, c->stream), "solve download");
            c->check(hipStreamSynchronize(c->stream), "solve sync");
        }
        return rc;
    });
}

const char* dwf_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
