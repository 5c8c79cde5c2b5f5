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
#include <mutex>
#include <string>
#include <vector>

using namespace dwfk;

namespace {
thread_local std::string g_last_error;
}

// ------------------------------------------------------------------ fixed scratch
// Device memory of dwf_apply_device_noalloc (the normal operators' intermediate vector):
// one copy per GPU, loaded with the code object and held until the process ends.  Host
// code never uses the array's name as an address: each device's address comes from
// hipGetSymbolAddress at context creation (with that device current) and is kept in
// g_scratch_ptr.  Every call on one GPU shares that GPU's array, so calls that overlap
// in time on one GPU (several host threads or streams) race on it.
static __device__ __align__(256) unsigned char scratch[DWF_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Looks up (once per device) the device address of `scratch` on device `dev`, which must
// be the current device.
hipError_t scratch_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[dev] = p;
    return e;
}

void* scratch_ptr(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[dev];
}
}  // namespace

void dwf_set_error(const std::string& msg) { g_last_error = msg; }

int dwf_hip_fail(hipError_t e, const char* what)
{
    dwf_set_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
    return DWF_ERR_HIP;
}

void dwf_ctx::drop_graphs()
{
    // An executable graph may still be running on the stream (launches are asynchronous):
    // finish it before destroying it.
    if (!graphs.empty() && stream) (void)hipStreamSynchronize(stream);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
    for (auto& k : kids) k->drop_graphs();  // their graphs hold this context's settings too
}

hipGraphExec_t dwf_ctx::cached_graph(const std::vector<unsigned long long>& key, const std::function<void()>& body)
{
    auto f = graphs.find(key);
    if (f != graphs.end()) return f->second;
    if (graphs.size() >= 64) drop_graphs();
    hipGraphExec_t ex = capture_graph(body);
    graphs.emplace(key, ex);
    return ex;
}

hipGraphExec_t dwf_ctx::capture_graph(const std::function<void()>& body)
{
    check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipGraph_t g = nullptr;
    try {
        body();
    } catch (...) {
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    check(e, "hipGraphInstantiate");
    return ex;
}

dwf_ctx::~dwf_ctx()
{
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    kids.clear();  // sub-contexts first: they borrow this context's gauge buffers
    for (auto& b : bufs)
        if (b.p) (void)hipFree(b.p);
    if (!borrowed) {
        if (theta) (void)hipFree(theta);
        if (l64) (void)hipFree(l64);
        if (l32) (void)hipFree(l32);
    }
    if (ev) (void)hipEventDestroy(ev);
    if (gr_dev) (void)hipFree(gr_dev);
    if (st_host) (void)hipHostFree(st_host);
    if (stream) (void)hipStreamDestroy(stream);
}

dwf_ctx* dwf_ctx::kid(int i, int cfg0, int kid_ncfg)
{
    if (multi()) throw DwfError(DWF_ERR_ARG, "sub-contexts hold one lattice");
    while ((int)kids.size() <= i) {
        std::unique_ptr<dwf_ctx> k(new dwf_ctx());
        k->borrowed = true;
        k->device = device;
        check(hipStreamCreateWithFlags(&k->stream, hipStreamNonBlocking), "sub-context stream");
        check(hipEventCreateWithFlags(&k->ev, hipEventDisableTiming), "sub-context event");
        kids.push_back(std::move(k));
    }
    if (!ev) check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "context event");
    dwf_ctx* k = kids[i].get();
    // the gauge field of configurations cfg0 ..: the parent's buffers at an offset
    const long long nl = 2LL * L.V4;
    double* th = theta + cfg0 * nl;
    double2* a64 = l64 + cfg0 * nl;
    float2* a32 = l32 + cfg0 * nl;
    if (k->l32 != a32 || k->l64 != a64 || k->theta != th) k->drop_graphs();  // graphs hold link pointers
    k->theta = th;
    k->l64 = a64;
    k->l32 = a32;
    k->L = L;
    k->L0 = L0;
    k->kind = kind;
    k->have_gauge = have_gauge;
    k->ncfg = kid_ncfg;
    k->cfg_cap = kid_ncfg;
    k->link_cap = (size_t)kid_ncfg * nl;
    k->gdefs.clear();
    k->chunks_override = chunks_override;
    k->xcd_remap = xcd_remap;
    k->prefetch = prefetch;
    k->tile_x = tile_x;
    k->nt_store = nt_store;
    k->lds = lds;
    k->net_fused = net_fused;
    k->fcg_beta = fcg_beta;
    k->eo_fused = eo_fused;
    k->eo_pipe = eo_pipe;
    k->net_tile = net_tile;
    k->net_split = net_split;
    k->reliable_delta = reliable_delta;
    k->delta_set = delta_set;
    k->use_graphs = use_graphs;
    k->net = net;
    k->streams = 1;
    return k;
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

void dwf_ctx::bind_rhs(int nrhs)
{
    if (nrhs < 1) throw DwfError(DWF_ERR_ARG, "the number of right-hand sides must be positive");
    if (ncfg <= 1) {
        rhs_per_cfg = 1 << 30;
    } else {
        if (nrhs % ncfg != 0)
            throw DwfError(DWF_ERR_ARG, "the number of right-hand sides must be a multiple of the configurations held (" +
                                            std::to_string(ncfg) + ")");
        rhs_per_cfg = nrhs / ncfg;
    }
    if (!multi()) {
        gr = single_group(L, nrhs, rhs_per_cfg);
        return;
    }
    // right-hand sides configuration-major in group order, rhs_per_cfg per configuration
    Groups G{};
    G.n = (int)gdefs.size();
    long long site = 0, link = 0;
    int r = 0;
    for (int g = 0; g < G.n; ++g) {
        const GroupDef& d = gdefs[g];
        G.rhs0[g] = r;
        G.Nt[g] = d.Nt;
        G.Nx[g] = d.Nx;
        G.rpc[g] = rhs_per_cfg;
        G.site0[g] = site;
        G.link0[g] = link;
        r += d.ncfg * rhs_per_cfg;
        site += (long long)d.ncfg * rhs_per_cfg * L.Ls * d.Nt * d.Nx;
        link += 2LL * d.ncfg * d.Nt * d.Nx;
    }
    G.rhs0[G.n] = r;
    gr = G;
    groups_device();
}

void dwf_ctx::require_single_lattice(const char* what) const
{
    if (multi()) throw DwfError(DWF_ERR_ARG, std::string(what) + " needs a context holding a single lattice size");
}

Groups dwf_ctx::groups_for(int nrhs) const
{
    if (!multi()) return single_group(L, nrhs, rhs_per_cfg);
    if (gr.n < 2 || gr.rhs0[gr.n] != nrhs) throw DwfError(DWF_ERR_ARG, "lattice groups: layout not bound to this call");
    return gr;
}

VecLayout dwf_ctx::layout(int nrhs, int shift) const
{
    VecLayout v{};
    v.gr = groups_for(nrhs);
    v.Ls = L.Ls;
    v.shift = shift;
    v.dev = multi() ? gr_dev : nullptr;
    return v;
}

const Groups* dwf_ctx::groups_device()
{
    if (!multi()) return nullptr;
    if (!gr_dev) check(hipMalloc(&gr_dev, sizeof(Groups)), "group table allocation");
    if (!gr_dev_valid || std::memcmp(&gr_dev_copy, &gr, sizeof(Groups)) != 0) {
        check(hipStreamSynchronize(stream), "group table update");  // no kernel still reads the old one
        check(hipMemcpy(gr_dev, &gr, sizeof(Groups), hipMemcpyHostToDevice), "group table update");
        gr_dev_copy = gr;
        gr_dev_valid = true;
    }
    return gr_dev;
}

void dwf_ctx::require_single_config(const char* what) const
{
    if (ncfg != 1) throw DwfError(DWF_ERR_ARG, std::string(what) + " needs a context holding one configuration");
}

void dwf_ctx::reserve_configs(int n)
{
    reserve_links((size_t)2 * L.V4 * n);
    cfg_cap = std::max(cfg_cap, n);
}

void dwf_ctx::reserve_links(size_t nl)
{
    if (nl <= link_cap) return;
    drop_graphs();  // captured kernels hold the old link buffers
    (void)hipStreamSynchronize(stream);
    (void)hipFree(theta);
    (void)hipFree(l64);
    (void)hipFree(l32);
    theta = nullptr;
    l64 = nullptr;
    l32 = nullptr;
    link_cap = 0;
    check(hipMalloc(&theta, nl * sizeof(double)), "gauge allocation");
    check(hipMalloc(&l64, nl * sizeof(double2)), "gauge allocation");
    check(hipMalloc(&l32, nl * sizeof(float2)), "gauge allocation");
    link_cap = nl;
}

// ------------------------------------------------------------------- operators
namespace {

// out = op(in) on device buffers; tmp is needed for the normal operators.
template <typename R> void apply_op(dwf_ctx* c, int op, const R* in, R* out, R* tmp, int nrhs)
{
    const HopCfg cfg = ctx_hop_cfg(c, nrhs);
    HopArgs<R> a = ctx_hop_args<R>(c, cfg, nrhs);
    switch (op) {
        case DWF_OP_D:
        case DWF_OP_DDAG:
            a.in = in;
            a.out = out;
            launch_hop<R>(op == DWF_OP_DDAG, EPI_STORE, a, nrhs, c->stream, c->groups_host());
            break;
        default: {
            const bool first_dag = (op == DWF_OP_DDDAG);
            a.in = in;
            a.out = tmp;
            launch_hop<R>(first_dag, EPI_STORE, a, nrhs, c->stream, c->groups_host());
            a.in = tmp;
            a.out = out;
            launch_hop<R>(!first_dag, EPI_STORE, a, nrhs, c->stream, c->groups_host());
        }
    }
    c->check_launch("stencil launch");
}

}  // namespace

HopCfg ctx_hop_cfg(const dwf_ctx* c, int nrhs)
{
    if (c->multi()) return hop_cfg(c->L, nrhs, c->chunks_override, c->xcd_remap, c->prefetch, 0, c->nt_store, 0);
    return hop_cfg(c->L, nrhs, c->chunks_override, c->xcd_remap, c->prefetch, c->tile_x, c->nt_store, c->lds);
}

template <typename R> HopArgs<R> ctx_hop_args(const dwf_ctx* c, const HopCfg& cfg, int nrhs)
{
    HopArgs<R> a{};
    a.L = c->L;
    a.cfg = cfg;
    a.links = sizeof(R) == 8 ? (const void*)c->l64 : (const void*)c->l32;
    a.link_stride = 2LL * c->L.V4;
    a.rhs_per_cfg = c->rhs_per_cfg;
    if (c->multi()) {
        (void)c->groups_for(nrhs);  // checks that the call's layout is bound
        a.groups = c->gr_dev;
    }
    return a;
}
template HopArgs<float> ctx_hop_args<float>(const dwf_ctx*, const HopCfg&, int);
template HopArgs<double> ctx_hop_args<double>(const dwf_ctx*, const HopCfg&, int);

void dwf_apply_device_impl(dwf_ctx* c, int op, int prec, const void* in, void* out, int nrhs)
{
    const size_t n = (size_t)c->vec_sites(nrhs) * 4;
    if (prec == DWF_PREC_FP64) {
        double* tmp = op >= DWF_OP_DDDAG ? (double*)c->buf(B_APPLY_TMP, n * sizeof(double)) : nullptr;
        apply_op<double>(c, op, (const double*)in, (double*)out, tmp, nrhs);
    } else {
        float* tmp = op >= DWF_OP_DDDAG ? (float*)c->buf(B_APPLY_TMP, n * sizeof(float)) : nullptr;
        apply_op<float>(c, op, (const float*)in, (float*)out, tmp, nrhs);
    }
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
    c->L0 = L;
    c->kind = kind;
    e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = scratch_setup(c->device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    const size_t nl = (size_t)2 * L.V4;
    if (e == hipSuccess) e = hipMalloc(&c->theta, nl * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->l64, nl * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc(&c->l32, nl * sizeof(float2));
    if (e != hipSuccess) {
        delete c;
        return dwf_hip_fail(e, "context creation");
    }
    c->link_cap = nl;
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
        case DWF_TUNE_PREFETCH:
            ctx->prefetch = value != 0.0 ? 1 : 0;
            return DWF_OK;
        case DWF_TUNE_TILE_X:
            if (value < 0 || value > 256) break;
            ctx->tile_x = (int)value;
            return DWF_OK;
        case DWF_TUNE_NT_STORE:
            ctx->nt_store = value != 0.0 ? 1 : 0;
            return DWF_OK;
        case DWF_TUNE_RELIABLE_DELTA:
            if (value == 0.0) {  // back to each operator's default
                ctx->reliable_delta = 0.01;
                ctx->delta_set = false;
                return DWF_OK;
            }
            if (!(value > 0.0 && value < 1.0)) break;
            ctx->reliable_delta = value;
            ctx->delta_set = true;
            return DWF_OK;
        case DWF_TUNE_GRAPHS:
            ctx->use_graphs = value != 0.0 ? 1 : 0;
            return DWF_OK;
        case DWF_TUNE_LDS:
            if (!(value >= 0.0 && value <= 6.0 && value == (double)(int)value)) break;
            ctx->lds = (int)value;
            return DWF_OK;
        case DWF_TUNE_NET_FUSED:
            if (!(value == 0.0 || value == 1.0 || value == 2.0)) break;
            ctx->net_fused = (int)value;
            return DWF_OK;
        case DWF_TUNE_FCG_BETA:
            if (!(value == 0.0 || value == 1.0)) break;
            ctx->fcg_beta = (int)value;
            return DWF_OK;
        case DWF_TUNE_EO_FUSED:
            if (!(value == 0.0 || value == 1.0 || value == 2.0 || value == 3.0)) break;
            ctx->eo_fused = value != 0.0 ? 1 : 0;
            ctx->eo_pipe = value >= 2.0 ? (int)value - 1 : 0;
            return DWF_OK;
        case DWF_TUNE_STREAMS:
            if (!(value >= 1.0 && value <= 8.0 && value == (double)(int)value)) break;
            ctx->streams = (int)value;
            return DWF_OK;
        case DWF_TUNE_NET_TILE: {
            if (!(value >= 0.0 && value == (double)(int)value)) break;
            const int v = (int)value, ty = v >> 8, tx = v & 255;
            if (v != 0 && (ty < 1 || ty > 8 || tx < 1 || tx > 16 || ctx->L.Nt % ty || ctx->L.Nx % tx)) break;
            if (v != ctx->net_tile) ctx->drop_graphs();
            ctx->net_tile = v;
            return DWF_OK;
        }
        case DWF_TUNE_NET_SPLIT: {
            if (!(value >= 0.0 && value == (double)(int)value)) break;
            const int v = (int)value;
            if (v != 0 && (v > ctx->L.Ls || ctx->L.Ls % v)) break;
            if (v != ctx->net_split) ctx->drop_graphs();
            ctx->net_split = v;
            return DWF_OK;
        }
        default:
            break;
    }
    dwf_set_error("dwf_ctx_tune: bad parameter or value");
    return DWF_ERR_ARG;
}

namespace {
// Back to the context's own lattice after lattice groups.
void single_lattice(dwf_ctx* c)
{
    if (c->gdefs.empty() && c->L.Nt == c->L0.Nt && c->L.Nx == c->L0.Nx) return;
    c->drop_graphs();
    c->gdefs.clear();
    c->L = c->L0;
}
}  // namespace

int dwf_set_gauge_device(dwf_ctx* c, const double* theta_device)
{
    return guard([&] {
        if (!c || !theta_device) throw DwfError(DWF_ERR_ARG, "dwf_set_gauge: null argument");
        single_lattice(c);
        c->reserve_configs(1);
        const size_t nl = (size_t)2 * c->L.V4;
        if (theta_device != c->theta)
            c->check(hipMemcpyAsync(c->theta, theta_device, nl * 8, hipMemcpyDeviceToDevice, c->stream), "gauge copy");
        launch_pack_links((long long)nl, c->theta, c->l64, c->l32, c->stream);
        c->check_launch("pack links");
        c->ncfg = 1;
        c->have_gauge = true;
        return DWF_OK;
    });
}

int dwf_set_gauge(dwf_ctx* c, const double* theta) { return dwf_set_gauge_batch(c, 1, theta); }

int dwf_set_gauge_batch(dwf_ctx* c, int ncfg, const double* theta)
{
    return guard([&] {
        if (!c || !theta || ncfg < 1) throw DwfError(DWF_ERR_ARG, "dwf_set_gauge: bad argument");
        single_lattice(c);
        const size_t nl = (size_t)2 * c->L.V4;
        c->reserve_configs(ncfg);
        c->check(hipMemcpyAsync(c->theta, theta, nl * ncfg * 8, hipMemcpyHostToDevice, c->stream), "gauge upload");
        launch_pack_links((long long)(nl * ncfg), c->theta, c->l64, c->l32, c->stream);
        c->check_launch("pack links");
        c->check(hipStreamSynchronize(c->stream), "gauge upload sync");
        c->ncfg = ncfg;
        c->have_gauge = true;
        return DWF_OK;
    });
}

int dwf_set_gauge_groups(dwf_ctx* c, int ngroups, const dwf_lattice_group* groups, const double* theta)
{
    return guard([&] {
        if (!c || !groups || !theta || ngroups < 1 || ngroups > kMaxGroups)
            throw DwfError(DWF_ERR_ARG, "dwf_set_gauge_groups: bad argument (1.." + std::to_string(kMaxGroups) +
                                            " lattice groups)");
        if (c->kind != 0) throw DwfError(DWF_ERR_ARG, "dwf_set_gauge_groups: needs a domain-wall context");
        size_t nl = 0;
        int ncfg = 0, big = 0;
        for (int g = 0; g < ngroups; ++g) {
            const dwf_lattice_group& d = groups[g];
            if (d.Nt < 1 || d.Nx < 1 || d.ncfg < 1 || (long long)d.Nt * d.Nx * c->L0.Ls * 2 > (1LL << 31) - 1)
                throw DwfError(DWF_ERR_ARG, "dwf_set_gauge_groups: lattice group " + std::to_string(g) +
                                                " needs positive extents and configurations");
            nl += (size_t)2 * d.Nt * d.Nx * d.ncfg;
            ncfg += d.ncfg;
            if ((long long)d.Nt * d.Nx > (long long)groups[big].Nt * groups[big].Nx) big = g;
        }
        c->drop_graphs();
        c->L = c->L0;
        c->L.Nt = groups[big].Nt;  // the largest lattice (shared Ls, masses, signs)
        c->L.Nx = groups[big].Nx;
        c->L.V4 = c->L.Nt * c->L.Nx;
        c->gdefs.clear();
        if (ngroups > 1)
            for (int g = 0; g < ngroups; ++g) c->gdefs.push_back({groups[g].Nt, groups[g].Nx, groups[g].ncfg});
        c->reserve_links(nl);
        c->cfg_cap = std::max(c->cfg_cap, ncfg);
        c->check(hipMemcpyAsync(c->theta, theta, nl * 8, hipMemcpyHostToDevice, c->stream), "gauge upload");
        launch_pack_links((long long)nl, c->theta, c->l64, c->l32, c->stream);
        c->check_launch("pack links");
        c->check(hipStreamSynchronize(c->stream), "gauge upload sync");
        c->ncfg = ncfg;
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
        c->bind_rhs(nrhs);
        dwf_apply_device_impl(c, op, precision, in, out, nrhs);
        return DWF_OK;
    });
}

int dwf_apply_device_noalloc(dwf_ctx* c, int op, int precision, const void* in, void* out, int nrhs)
{
    return guard([&] {
        if (!c || !in || !out || nrhs < 1 || op < DWF_OP_D || op > DWF_OP_DDAGD ||
            (precision != DWF_PREC_FP64 && precision != DWF_PREC_FP32) || in == out)
            throw DwfError(DWF_ERR_ARG, "dwf_apply_device_noalloc: bad argument");
        if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_apply_device_noalloc: no gauge field set");
        c->bind_rhs(nrhs);
        const unsigned long long real = precision == DWF_PREC_FP64 ? 8 : 4;
        const unsigned long long need =
            op >= DWF_OP_DDDAG ? (unsigned long long)c->vec_sites(nrhs) * 4ULL * real : 0ULL;
        if (need > DWF_NOALLOC_SCRATCH_BYTES)
            throw DwfError(DWF_ERR_ARG, "dwf_apply_device_noalloc: the intermediate vector needs " +
                                            std::to_string(need) + " bytes, more than the " +
                                            std::to_string(DWF_NOALLOC_SCRATCH_BYTES) +
                                            "-byte scratch; pass fewer right-hand sides or use dwf_apply_device");
        void* tmp = scratch_ptr(c->device);
        if (!tmp) throw DwfError(DWF_ERR_HIP, "dwf_apply_device_noalloc: no scratch address for this device");
        if (precision == DWF_PREC_FP64)
            apply_op<double>(c, op, (const double*)in, (double*)out, op >= DWF_OP_DDDAG ? (double*)tmp : nullptr, nrhs);
        else
            apply_op<float>(c, op, (const float*)in, (float*)out, op >= DWF_OP_DDDAG ? (float*)tmp : nullptr, nrhs);
        return DWF_OK;
    });
}

int dwf_apply(dwf_ctx* c, int op, const double* in, double* out, int nrhs)
{
    return guard([&] {
        if (!c || !in || !out || nrhs < 1 || op < DWF_OP_D || op > DWF_OP_DDAGD)
            throw DwfError(DWF_ERR_ARG, "dwf_apply: bad argument");
        if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_apply: no gauge field set");
        c->bind_rhs(nrhs);
        const size_t n = (size_t)c->vec_sites(nrhs) * 4;
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

extern "C" {

const char* dwf_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
