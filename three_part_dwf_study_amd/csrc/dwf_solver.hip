// dwf_solver.hip -- conjugate-gradient drivers of libdwfhip.
//
// One driver (cg_run) runs CG on any operator A = Op2 Op1 whose first half yields
// p^dag A p = |Op1 p|^2 (the normal operators D^dag D, D D^dag and the even-odd Schur
// normal operator Dhat^dag Dhat).  Every vector and per-RHS scalar stays on the device;
// per iteration three launches (Op1 + norm partials, Op2 fused with the x/r update and
// |r|^2 partials, direction update) for the full operator, five for the Schur one.
//   CG64:  FP64 throughout, the reference's recursive-residual stop
//          (domain_wall/include/cg_solver.hpp:17-64);
//   MIXED: FP32 iterations, FP64 reliable updates certifying on the FP64 true residual.
// dwf_solve_dirac adds the propagator pattern x = D^-1 src of correlator.hpp (rhs = D^dag
// src, CG on D^dag D), optionally through the even-odd Schur complement with defect
// correction, certified on the reference's own system ||D^dag src - D^dag D x||.

#include "../../include/dwf_hip.h"
#include "dwf_context.hpp"
#include "dwf_eo.hpp"

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <functional>
#include <string>
#include <thread>
#include <vector>

using namespace dwfk;

namespace {

// ----------------------------------------------------------------- operators
struct CgOperator {
    VecLayout lay{};      // extent and offset of each right-hand side's x, b, r, p
    int nblk = 0;         // block partials per right-hand side (the partial arrays' stride)
    virtual ~CgOperator() = default;
    // t = Op1 p, partials |t|^2
    virtual void first(bool f64, const void* p, void* t, double* part, const CgState* st, int slot, int skip) = 0;
    // x += alpha p, r -= alpha Op2 t, partials |r|^2 (alpha = rr / sum partA)
    virtual void second(bool f64, const void* t, void* x, void* r, const void* p, const double* partA, double* part,
                        const CgState* st, int slot, int skip) = 0;
    // r64 = b - A x in FP64 (copies into p64/r32/p32 if given), partials |r|^2, |b|^2
    virtual void residual(const double* x, const double* b, double* r64, double* p64, float* r32, float* p32,
                          double* part, const CgState* st, int slot, int skip) = 0;
    // Every value first()/second() pass to their kernels beyond their own arguments.
    virtual void signature(std::vector<unsigned long long>& k) const = 0;
    // Pipelined FP32 iteration (optional): first_pipe forms p = r + beta p_old from the
    // states and the two direction buffers (CgState::pbuf) inside its operator launch and
    // computes t = Op1 p; second_pipe is second() with that direction; the iteration ends
    // with the state update alone (launch_cg_flags), or leaves it to the next first_pipe
    // (fold: that launch does it from the previous iteration's partials partAp, partB).
    virtual bool pipelined() const { return false; }
    // The mixed solver's reliable-update drop for this operator when none is set by hand
    // (0: the context's default).
    virtual double default_delta() const { return 0.0; }
    virtual void first_pipe(float* p0, float* p1, const float* r, float* t, double* part, CgState* st, int slot,
                            int skip, int fold, const double* partAp, const double* partB)
    {
        (void)p0, (void)p1, (void)r, (void)t, (void)part, (void)st, (void)slot, (void)skip;
        (void)fold, (void)partAp, (void)partB;
    }
    virtual void second_pipe(const float* t, float* x, float* r, float* p0, float* p1, const double* partA,
                             double* part, CgState* st, int slot, int skip)
    {
        (void)t, (void)x, (void)r, (void)p0, (void)p1, (void)partA, (void)part, (void)st, (void)slot, (void)skip;
    }
};

using Sig = std::vector<unsigned long long>;
inline unsigned long long sig_ptr(const void* p) { return (unsigned long long)(uintptr_t)p; }

void sig_ctx(const dwf_ctx* c, Sig& k)
{
    k.push_back(sig_ptr(c->l64));
    k.push_back(sig_ptr(c->l32));
    k.push_back((unsigned long long)c->rhs_per_cfg);
    k.push_back((unsigned long long)c->xcd_remap);
    if (c->multi()) {  // the lattice-group layout the kernels receive
        const Groups& G = c->gr;
        k.push_back(0x6705ULL + (unsigned long long)G.n);
        k.push_back(sig_ptr(c->gr_dev));
        for (int g = 0; g < G.n; ++g)
            k.insert(k.end(), {(unsigned long long)G.Nt[g], (unsigned long long)G.Nx[g], (unsigned long long)G.rhs0[g],
                               (unsigned long long)G.rpc[g], (unsigned long long)G.site0[g],
                               (unsigned long long)G.link0[g]});
        k.push_back((unsigned long long)G.rhs0[G.n]);
    }
}

// A = Op2 Op1 = D^dag D (first_dag = false) or D D^dag (first_dag = true).
struct FullOp : CgOperator {
    dwf_ctx* c;
    HopCfg cfg;
    bool first_dag;
    int nrhs;
    double* t64;
    FullOp(dwf_ctx* ctx, bool fd, int n, double* tmp64) : c(ctx), first_dag(fd), nrhs(n), t64(tmp64)
    {
        cfg = ctx_hop_cfg(c, nrhs);
        lay = c->layout(nrhs, 0);
        nblk = hop_pstride(lay.gr, cfg);
    }
    template <typename R> HopArgs<R> args() const { return ctx_hop_args<R>(c, cfg, nrhs); }
    template <typename R> void first_t(const R* p, R* t, double* part, const CgState* st, int slot, int skip)
    {
        HopArgs<R> a = args<R>();
        a.in = p;
        a.out = t;
        a.part = part;
        a.st = st;
        a.slot = slot;
        a.skip = skip;
        launch_hop<R>(first_dag, EPI_NORM, a, nrhs, c->stream, c->groups_host());
    }
    template <typename R>
    void second_t(const R* t, R* x, R* r, const R* p, const double* partA, double* part, const CgState* st, int slot,
                  int skip)
    {
        HopArgs<R> a = args<R>();
        a.in = t;
        a.x = x;
        a.r = r;
        a.p = p;
        a.partA = partA;
        a.part = part;
        a.st = st;
        a.slot = slot;
        a.skip = skip;
        launch_hop<R>(!first_dag, EPI_CGUPD, a, nrhs, c->stream, c->groups_host());
    }
    void first(bool f64, const void* p, void* t, double* part, const CgState* st, int slot, int skip) override
    {
        if (f64)
            first_t<double>((const double*)p, (double*)t, part, st, slot, skip);
        else
            first_t<float>((const float*)p, (float*)t, part, st, slot, skip);
    }
    void second(bool f64, const void* t, void* x, void* r, const void* p, const double* partA, double* part,
                const CgState* st, int slot, int skip) override
    {
        if (f64)
            second_t<double>((const double*)t, (double*)x, (double*)r, (const double*)p, partA, part, st, slot, skip);
        else
            second_t<float>((const float*)t, (float*)x, (float*)r, (const float*)p, partA, part, st, slot, skip);
    }
    void residual(const double* x, const double* b, double* r64, double* p64, float* r32, float* p32, double* part,
                  const CgState* st, int slot, int skip) override
    {
        HopArgs<double> a = args<double>();
        a.in = x;
        a.out = t64;
        a.st = st;
        a.slot = slot;
        a.skip = skip;
        launch_hop<double>(first_dag, EPI_STORE, a, nrhs, c->stream, c->groups_host());
        HopArgs<double> e = args<double>();
        e.in = t64;
        e.b = b;
        e.r64 = r64;
        e.p64 = p64;
        e.r32 = r32;
        e.p32 = p32;
        e.part = part;
        e.st = st;
        e.slot = slot;
        e.skip = skip;
        launch_hop<double>(!first_dag, EPI_RESID, e, nrhs, c->stream, c->groups_host());
    }
    void signature(Sig& k) const override
    {
        sig_ctx(c, k);
        k.insert(k.end(), {1ULL, (unsigned long long)first_dag, (unsigned long long)nrhs});
        const int v[] = {cfg.bx,       cfg.chunks,  cfg.spt,      cfg.remap, cfg.prefetch,
                         cfg.tx,       cfg.tiles_x, cfg.nt_store, cfg.lds};
        for (int x : v) k.push_back((unsigned long long)(unsigned)x);
    }
};

// A = Dhat^dag Dhat on the odd half-lattice.
struct SchurOp : CgOperator {
    dwf_ctx* c;
    int nrhs;
    int bx, chunks, spt;
    float* te32;
    double* te64;
    double* q64;
    bool fused = false;  // one k_eo_schur launch per application (A^-1 H psi kept on the chip)
    EoTile tile{};
    SchurOp(dwf_ctx* ctx, int n, float* t32, double* t64, double* q) : c(ctx), nrhs(n), te32(t32), te64(t64), q64(q)
    {
        const Lat& L = c->L;
        lay = c->layout(nrhs, 1);
        long long blocks = 0;  // blocks of one s-chunk over every right-hand side
        for (int g = 0; g < lay.gr.n; ++g)
            blocks += (long long)(lay.gr.rhs0[g + 1] - lay.gr.rhs0[g]) *
                      ((lay.gr.Nt[g] * lay.gr.Nx[g] / 2 + kBlock - 1) / kBlock);
        bx = (int)std::min<long long>(blocks / std::max(1, nrhs), 1 << 30);
        int ch = 1;
        if (c->chunks_override > 0)
            ch = c->chunks_override;
        else
            while (ch < L.Ls && blocks * ch < 2048) ch *= 2;
        ch = std::min(ch, L.Ls);
        spt = (L.Ls + ch - 1) / ch;
        chunks = (L.Ls + spt - 1) / spt;
        nblk = eo_pstride(lay.gr, chunks);
        if (c->eo_fused && eo_fused_tile(lay.gr, L.Ls, tile)) {
            fused = true;
            nblk = eo_fused_pstride(lay.gr, tile);
        }
    }
    // the fused kernel's arguments for out = [Dhat | Dhat^dag] in (epilogue fields set by the caller)
    template <typename R> EoArgs<R> fused_args(const R* in, const CgState* st, int slot, int skip) const
    {
        EoArgs<R> b = base<R>(1);
        b.in = in;
        b.ca = (R)1;
        b.ch = (R)-0.25;
        b.st = st;
        b.slot = slot;
        b.skip = skip;
        return b;
    }
    template <typename R> EoArgs<R> base(int tpar) const
    {
        EoArgs<R> a{};
        a.L = c->L;
        a.Nxh = c->L.Nx / 2;
        a.V4h = c->L.V4 / 2;
        a.bx = (a.V4h + kBlock - 1) / kBlock;
        a.bxa = (a.V4h + kBlockA - 1) / kBlockA;
        a.groups = c->multi() ? c->gr_dev : nullptr;  // then each group's sizes replace these
        a.chunks = chunks;
        a.spt = spt;
        a.remap = c->xcd_remap;
        a.tpar = tpar;
        a.links = sizeof(R) == 8 ? (const void*)c->l64 : (const void*)c->l32;
        a.link_stride = 2LL * c->L.V4;
        a.rhs_per_cfg = c->rhs_per_cfg;
        return a;
    }
    // out_o = [Dhat | Dhat^dag] in_o with epilogue `epi` (args in the returned struct)
    template <typename R> EoArgs<R> apply_args(bool dag, const R* in, R* te, const CgState* st, int slot, int skip)
    {
        EoArgs<R> a = base<R>(0);
        a.hin = in;
        a.ch = (R)1;
        a.out = te;
        a.st = st;
        a.slot = slot;
        a.skip = skip;
        launch_eo_a<R>(dag, true, a, nrhs, c->stream, c->groups_host());  // te = A^-1 H_eo in   (A^-dag H^dag when dag)
        EoArgs<R> b = base<R>(1);
        b.in = in;
        b.ca = (R)1;
        b.hin = te;
        b.ch = (R)-0.25;
        b.st = st;
        b.slot = slot;
        b.skip = skip;
        return b;
    }
    void first(bool f64, const void* p, void* t, double* part, const CgState* st, int slot, int skip) override
    {
        if (fused) {
            if (f64) {
                EoArgs<double> b = fused_args<double>((const double*)p, st, slot, skip);
                b.out = (double*)t;
                b.part = part;
                launch_eo_schur<double>(false, EPI_NORM, b, tile, nrhs, c->stream, c->groups_host());
            } else {
                EoArgs<float> b = fused_args<float>((const float*)p, st, slot, skip);
                b.out = (float*)t;
                b.part = part;
                launch_eo_schur<float>(false, EPI_NORM, b, tile, nrhs, c->stream, c->groups_host());
            }
            return;
        }
        if (f64) {
            EoArgs<double> b = apply_args<double>(false, (const double*)p, te64, st, slot, skip);
            b.out = (double*)t;
            b.part = part;
            launch_eo_b<double>(false, true, EPI_NORM, b, nrhs, c->stream, c->groups_host());
        } else {
            EoArgs<float> b = apply_args<float>(false, (const float*)p, te32, st, slot, skip);
            b.out = (float*)t;
            b.part = part;
            launch_eo_b<float>(false, true, EPI_NORM, b, nrhs, c->stream, c->groups_host());
        }
    }
    void second(bool f64, const void* t, void* x, void* r, const void* p, const double* partA, double* part,
                const CgState* st, int slot, int skip) override
    {
        if (fused) {
            if (f64) {
                EoArgs<double> b = fused_args<double>((const double*)t, st, slot, skip);
                b.x = (double*)x;
                b.r = (double*)r;
                b.p = (const double*)p;
                b.partA = partA;
                b.part = part;
                launch_eo_schur<double>(true, EPI_CGUPD, b, tile, nrhs, c->stream, c->groups_host());
            } else {
                EoArgs<float> b = fused_args<float>((const float*)t, st, slot, skip);
                b.x = (float*)x;
                b.r = (float*)r;
                b.p = (const float*)p;
                b.partA = partA;
                b.part = part;
                launch_eo_schur<float>(true, EPI_CGUPD, b, tile, nrhs, c->stream, c->groups_host());
            }
            return;
        }
        if (f64) {
            EoArgs<double> b = apply_args<double>(true, (const double*)t, te64, st, slot, skip);
            b.x = (double*)x;
            b.r = (double*)r;
            b.p = (const double*)p;
            b.partA = partA;
            b.part = part;
            launch_eo_b<double>(true, true, EPI_CGUPD, b, nrhs, c->stream, c->groups_host());
        } else {
            EoArgs<float> b = apply_args<float>(true, (const float*)t, te32, st, slot, skip);
            b.x = (float*)x;
            b.r = (float*)r;
            b.p = (const float*)p;
            b.partA = partA;
            b.part = part;
            launch_eo_b<float>(true, true, EPI_CGUPD, b, nrhs, c->stream, c->groups_host());
        }
    }
    bool pipelined() const override { return fused && c->eo_pipe; }
    // 3e-4: at 96^2 x 16, 32 right-hand sides, 5 reliable updates per solve instead of 8 at
    // 1e-2 with the same FP32 iterations (+2% solves/s)
    double default_delta() const override { return 3e-4; }
    void first_pipe(float* p0, float* p1, const float* r, float* t, double* part, CgState* st, int slot,
                    int skip, int fold, const double* partAp, const double* partB) override
    {
        EoArgs<float> b = fused_args<float>(nullptr, st, slot, skip);
        b.r = const_cast<float*>(r);
        b.pd0 = p0;
        b.pd1 = p1;
        b.out = t;
        b.part = part;
        b.fold = fold;
        b.partAp = partAp;
        b.partB = partB;
        b.stw = st;
        launch_eo_schur_pipe(b, tile, nrhs, c->stream, c->groups_host());
    }
    void second_pipe(const float* t, float* x, float* r, float* p0, float* p1, const double* partA, double* part,
                     CgState* st, int slot, int skip) override
    {
        EoArgs<float> b = fused_args<float>(t, st, slot, skip);
        b.x = x;
        b.r = r;
        b.pd0 = p0;
        b.pd1 = p1;
        b.partA = partA;
        b.part = part;
        launch_eo_schur<float>(true, EPI_CGUPD, b, tile, nrhs, c->stream, c->groups_host());
    }
    void residual(const double* x, const double* b64, double* r64, double* p64, float* r32, float* p32, double* part,
                  const CgState* st, int slot, int skip) override
    {
        if (fused) {
            EoArgs<double> b = fused_args<double>(x, st, slot, skip);
            b.out = q64;
            launch_eo_schur<double>(false, EPI_STORE, b, tile, nrhs, c->stream, c->groups_host());
            EoArgs<double> e = fused_args<double>(q64, st, slot, skip);
            e.b = b64;
            e.r64 = r64;
            e.p64 = p64;
            e.r32 = r32;
            e.p32 = p32;
            e.part = part;
            launch_eo_schur<double>(true, EPI_RESID, e, tile, nrhs, c->stream, c->groups_host());
            return;
        }
        EoArgs<double> b = apply_args<double>(false, x, te64, st, slot, skip);
        b.out = q64;
        launch_eo_b<double>(false, true, EPI_STORE, b, nrhs, c->stream, c->groups_host());
        EoArgs<double> e = apply_args<double>(true, q64, te64, st, slot, skip);
        e.b = b64;
        e.r64 = r64;
        e.p64 = p64;
        e.r32 = r32;
        e.p32 = p32;
        e.part = part;
        launch_eo_b<double>(true, true, EPI_RESID, e, nrhs, c->stream, c->groups_host());
    }
    void signature(Sig& k) const override
    {
        sig_ctx(c, k);
        k.insert(k.end(), {2ULL, (unsigned long long)nrhs, (unsigned long long)bx, (unsigned long long)chunks,
                           (unsigned long long)spt, sig_ptr(te32), sig_ptr(te64), sig_ptr(q64),
                           (unsigned long long)fused, (unsigned long long)tile.ty, (unsigned long long)tile.tx,
                           (unsigned long long)tile.nch, (unsigned long long)tile.spt,
                           (unsigned long long)pipelined()});
    }
};

// ------------------------------------------------------------------- driver
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

__global__ void k_mark_done(CgState* st, const CgState* pre, int nrhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrhs && (pre[i].flags[0] & F_DONE)) {
        st[i].flags[0] |= F_DONE;
        st[i].flags[1] |= F_DONE;
    }
}

__global__ void k_axpy_masked(double* x, const double* d, const VecLayout lay, const CgState* pre)
{
    const int rhs = blockIdx.y;
    if (pre && (pre[rhs].flags[0] & F_DONE)) return;
    long long n, off;
    lay.locate(rhs, n, off);
    const long long n_per = n * 4;
    double* xr = x + off * 4;
    const double* dr = d + off * 4;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_per; i += (long long)gridDim.x * blockDim.x)
        xr[i] += dr[i];
}

__global__ void k_sub_into(const double* a, const double* b, double* out, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = a[i] - b[i];
}

// CG on op.  Leaves the per-RHS states in c->st_host (slot c->last_slot).  `pre`
// (nullable) marks right-hand sides that are already finished.
void cg_run(dwf_ctx* c, CgOperator& op, bool mixed, const double* b, const double* x0, double rtol, int maxiter,
            int nrhs, double* x, const CgState* pre = nullptr)
{
    const size_t nv = (size_t)op.lay.total_sites() * 4;
    const size_t npart = (size_t)nrhs * op.nblk * 2;
    double* r64 = (double*)c->buf(B_R64, nv * 8);
    double* partA = (double*)c->buf(B_PA, npart * 8);
    double* partB = (double*)c->buf(B_PB, npart * 8);
    double* partR = (double*)c->buf(B_PR, npart * 8);
    CgState* st = (CgState*)c->buf(B_STATE, sizeof(CgState) * nrhs);
    double *p64 = nullptr, *t64 = nullptr;
    float *r32 = nullptr, *t32 = nullptr, *p32 = nullptr, *x32 = nullptr;
    const bool pipe = mixed && op.pipelined();
    float* p32b = nullptr;
    // pipelined: the operator's partials by iteration parity (a launch that folds the previous
    // iteration's state update reads that iteration's while it writes its own)
    double* partA2 = pipe ? (double*)c->buf(B_PA2, npart * 8) : nullptr;
    if (mixed) {
        r32 = (float*)c->buf(B_R32, nv * 4);
        t32 = (float*)c->buf(B_T32, nv * 4);
        p32 = (float*)c->buf(B_P32, nv * 4);
        x32 = (float*)c->buf(B_X32, nv * 4);
        if (pipe) p32b = (float*)c->buf(B_P32B, nv * 4);
    } else {
        p64 = (double*)c->buf(B_P64, nv * 8);
        t64 = (double*)c->buf(B_TQ64, nv * 8);
    }
    hipStream_t s = c->stream;
    if (op.lay.gr.n > 1) {
        // a group with fewer blocks than the stride leaves partials unwritten: they must read 0
        launch_zero<double>(partA, (long long)npart, s);
        if (partA2) launch_zero<double>(partA2, (long long)npart, s);
        launch_zero<double>(partB, (long long)npart, s);
        launch_zero<double>(partR, (long long)npart, s);
    }
    if (x0) {
        if (x0 != x) c->check(hipMemcpyAsync(x, x0, nv * 8, hipMemcpyDeviceToDevice, s), "copy x0");
    } else {
        launch_zero<double>(x, (long long)nv, s);
    }
    // r = b - A x0 (the reference applies A even to a zero guess, cg_solver.hpp:24-27)
    op.residual(x, b, r64, mixed ? nullptr : p64, mixed ? r32 : nullptr, mixed ? p32 : nullptr, partR, nullptr, 0, 0);
    const double delta = (!c->delta_set && op.default_delta() > 0.0) ? op.default_delta() : c->reliable_delta;
    launch_cg_start(op.nblk, partR, st, rtol, delta, nrhs, s);
    if (pre) k_mark_done<<<(nrhs + 63) / 64, 64, 0, s>>>(st, pre, nrhs);
    if (pipe) launch_set_fresh(st, nrhs, s);
    if (mixed) launch_zero<float>(x32, (long long)nv, s);
    c->check_launch("solver start");

    int it = 0;
    int chunk = mixed ? 16 : 32;
    read_states(c, st, nrhs);
    // per right-hand side: residual and iteration count at the start of its current FP32
    // cycle, and the length of its last cycle
    std::vector<double> c0_rel(nrhs);
    std::vector<int> c0_it(nrhs, 0), cycle_len(nrhs, 0);
    std::vector<char> upd(nrhs, 0);
    std::vector<int> hold_h(mixed ? nrhs : 0);
    int* hold_d = mixed ? (int*)c->buf(B_HOLD, sizeof(int) * nrhs) : nullptr;
    for (int i = 0; i < nrhs; ++i) c0_rel[i] = c->st_host[i].relres;
    std::vector<double> prev_rel(c0_rel);  // residual at the start of the previous cycle
    const int mode = mixed ? MODE_MIXED : MODE_CG64;
    // fold: the pipelined operator launch also does the previous iteration's state update;
    // flags: this iteration's state update is launched (else the next iteration folds it)
    auto iteration = [&](int i, bool fold, bool flags) {
        const int slot = i & 1, skip = F_DONE | F_FROZEN;
        if (pipe) {
            double* pa = slot ? partA2 : partA;
            op.first_pipe(p32, p32b, r32, t32, pa, st, slot, skip, fold ? 1 : 0, slot ? partA : partA2, partB);
            op.second_pipe(t32, x32, r32, p32, p32b, pa, partB, st, slot, skip);
            if (flags) launch_cg_flags(op.nblk, pa, partB, st, slot, mode, nrhs, s);
        } else if (mixed) {
            op.first(false, p32, t32, partA, st, slot, skip);
            op.second(false, t32, x32, r32, p32, partA, partB, st, slot, skip);
            launch_cg_p<float>(op.lay, op.nblk, r32, p32, partA, partB, st, slot, mode, nrhs, s);
        } else {
            op.first(true, p64, t64, partA, st, slot, skip);
            op.second(true, t64, x, r64, p64, partA, partB, st, slot, skip);
            launch_cg_p<double>(op.lay, op.nblk, r64, p64, partA, partB, st, slot, mode, nrhs, s);
        }
    };
    // Blocks of kGraphIters iterations starting on an even iteration are one graph launch.
    constexpr int kGraphIters = 8;
    // Blocks starting on an odd iteration have a graph of their own (the slots alternate
    // the other way), so every full block of a chunk is replayed whatever the parity the
    // previous chunk ended on.
    hipGraphExec_t block[2] = {nullptr, nullptr};
    if (c->use_graphs && maxiter >= kGraphIters) {
        Sig key{(unsigned long long)mixed, (unsigned long long)nrhs, (unsigned long long)nv,
                (unsigned long long)op.nblk, (unsigned long long)mode, sig_ptr(x),   sig_ptr(r64),  sig_ptr(p64),
                sig_ptr(t64),            sig_ptr(r32),            sig_ptr(t32),  sig_ptr(p32), sig_ptr(x32),
                sig_ptr(partA),          sig_ptr(partB),          sig_ptr(st),   sig_ptr(p32b),
                sig_ptr(partA2)};
        op.signature(key);
        // pipelined: iterations 1..7 of the block fold their predecessor's state update, the
        // block's last one launches its own (so a block always ends with complete states)
        const bool fold = pipe && c->eo_pipe == 2;
        key.push_back((unsigned long long)fold);
        // room for both in the cache, so capturing the second cannot evict the first
        if (c->graphs.size() + 2 > 64) c->drop_graphs();
        for (int par = 0; par < 2; ++par) {
            Sig kp = key;
            kp.push_back((unsigned long long)par);
            block[par] = c->cached_graph(kp, [&] {
                for (int i = 0; i < kGraphIters; ++i)
                    iteration(par + i, fold && i > 0, !fold || i == kGraphIters - 1);
            });
        }
    }
    while (!all_done(c->st_host, it & 1, nrhs)) {
        const int n = std::min(chunk, maxiter - it);
        if (n <= 0) break;
        int j = 0;
        if (block[it & 1])
            for (; j + kGraphIters <= n; j += kGraphIters, it += kGraphIters)
                c->check(hipGraphLaunch(block[it & 1], s), "hipGraphLaunch");
        for (; j < n; ++j, ++it) iteration(it, false, true);
        c->check_launch("CG iteration");
        if (mixed && op.lay.gr.n == 1) {
            // One lattice: the right-hand sides converge at similar rates, so every one not
            // done gets its reliable update (x64 += x32, FP64 true residual, certify or start
            // the next cycle from it) at each chunk end, which keeps their cycles in step;
            // the next chunk is the fewest iterations any of them needs to shrink its true
            // residual by delta (or reach the tolerance) at its last cycle's rate.
            const int slot = it & 1;
            launch_accum_x(op.lay, x, x32, st, slot, F_DONE, nrhs, s);
            op.residual(x, b, r64, nullptr, r32, nullptr, partR, st, slot, F_DONE);
            launch_ru_commit(op.nblk, partR, st, slot, nrhs, s);
            c->check_launch("reliable update");
            read_states(c, st, nrhs);
            double best = 1e30;
            for (int i = 0; i < nrhs; ++i) {
                const CgState& S = c->st_host[i];
                if (S.flags[slot] & F_DONE) continue;
                const int di = S.iters - c0_it[i];
                if (di > 0 && S.true_relres > 0.0 && c0_rel[i] > 0.0 && S.true_relres < c0_rel[i]) {
                    const double rate = std::log(S.true_relres / c0_rel[i]) / di;
                    const double need_delta = std::log(delta) / rate;
                    const double need_tol = std::log(rtol / S.true_relres) / rate + 2.0;
                    best = std::min(best, std::min(need_delta, need_tol));
                }
                c0_rel[i] = S.true_relres;
                c0_it[i] = S.iters;
            }
            if (best < 1e29) chunk = std::max(8, std::min(1024, (int)std::ceil(best)));
            // whole blocks of kGraphIters (replayed as graphs when they are on; the same
            // schedule without them, so both give the same iterates): a right-hand side that
            // freezes before the chunk ends skips the rest, where a remainder would run as
            // eager launches
            chunk = (chunk + kGraphIters - 1) / kGraphIters * kGraphIters;
        } else if (mixed) {
            // Several lattices converge at different rates: per right-hand side reliable
            // updates (x64 += x32, FP64 true residual, certify or start the next
            // cycle from it) for the right-hand sides whose FP32 cycle ended (frozen: the
            // residual fell by delta or met the tolerance), and for those past half of their
            // expected cycle: they would freeze soon and, updated together, stay in step, so
            // one chunk end serves them all.  The others keep iterating (F_HOLD).
            const int slot = it & 1;
            read_states(c, st, nrhs);
            int frozen = 0, updated = 0, running = 0;
            bool stalled = false;
            std::vector<double> rem(nrhs, -1.0);  // iterations until each running one freezes (-1: unknown)
            for (int i = 0; i < nrhs; ++i) {
                const CgState& S = c->st_host[i];
                const int f = S.flags[slot];
                upd[i] = 0;
                if (f & F_DONE) continue;
                const int di = S.iters - c0_it[i];
                if (f & F_FROZEN) {
                    ++frozen;
                    upd[i] = 1;
                    cycle_len[i] = di;
                    continue;
                }
                ++running;
                if (di > 0 && S.relres > 0.0 && S.relres < c0_rel[i]) {
                    const double rate = std::log(S.relres / c0_rel[i]) / di;
                    const double target = std::max(delta * std::sqrt(S.rr_ref / S.bnorm2), S.tol);
                    rem[i] = std::max(0.0, std::log(target / S.relres) / rate);
                } else if (di >= 64) {
                    stalled = true;  // no progress in a whole chunk: refresh it from the FP64 residual
                }
            }
            if (frozen > 0 || stalled) {
                for (int i = 0; i < nrhs; ++i) {
                    const CgState& S = c->st_host[i];
                    if ((S.flags[slot] & (F_DONE | F_FROZEN)) || upd[i]) continue;
                    const int di = S.iters - c0_it[i];
                    // past half of its cycle (remaining at most what it has done), or stalled
                    if (stalled || (rem[i] >= 0.0 && rem[i] <= di)) {
                        upd[i] = 1;
                        cycle_len[i] = (int)std::ceil(di + rem[i]);
                    }
                }
                for (int i = 0; i < nrhs; ++i) hold_h[i] = upd[i] ? 0 : 1;
                c->check(hipMemcpyAsync(hold_d, hold_h.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s),
                         "hold upload");
                launch_hold(st, hold_d, slot, nrhs, s);
                const int skip = F_DONE | F_HOLD;
                    launch_accum_x(op.lay, x, x32, st, slot, skip, nrhs, s);
                op.residual(x, b, r64, nullptr, r32, nullptr, partR, st, slot, skip);
                launch_ru_commit(op.nblk, partR, st, slot, nrhs, s);
                c->check_launch("reliable update");
                read_states(c, st, nrhs);
                for (int i = 0; i < nrhs; ++i) {
                    const CgState& S = c->st_host[i];
                    if (!upd[i]) continue;
                    ++updated;
                    c0_rel[i] = S.true_relres;
                    c0_it[i] = S.iters;
                }
            }
            // next chunk end: the first expected freeze -- a held one's remaining
            // iterations, or a restarted one's cycle (as long as its last one, or shorter
            // if the tolerance is near)
            double best = 1e30;
            for (int i = 0; i < nrhs; ++i) {
                const CgState& S = c->st_host[i];
                if (S.flags[slot] & F_DONE) continue;
                if (upd[i]) {
                    double need = cycle_len[i] > 0 ? cycle_len[i] : 16;
                    if (cycle_len[i] > 0 && prev_rel[i] > 0.0 && S.true_relres > 0.0 && S.true_relres < prev_rel[i]) {
                        const double rate = std::log(S.true_relres / prev_rel[i]) / cycle_len[i];
                        need = std::min(need, std::log(S.tol / S.true_relres) / rate + 1.0);
                    }
                    best = std::min(best, need);
                } else if (rem[i] >= 0.0) {
                    best = std::min(best, rem[i] + 1.0);
                }
            }
            for (int i = 0; i < nrhs; ++i)
                if (upd[i]) prev_rel[i] = c->st_host[i].true_relres;
            (void)updated;
            (void)running;
            chunk = best < 1e29 ? std::max(8, std::min(1024, (int)std::ceil(best))) : 16;
        } else {
            read_states(c, st, nrhs);
            chunk = 32;
        }
    }
    // FP64 true residual of the final x
    op.residual(x, b, r64, nullptr, nullptr, nullptr, partR, nullptr, 0, 0);
    launch_true_resid(op.nblk, partR, st, nrhs, s);
    c->check_launch("true residual");
    read_states(c, st, nrhs);
    c->last_slot = it & 1;
}

void to_stats(const CgState& S, int slot, int method, dwf_solve_stats& s)
{
    s = dwf_solve_stats{};
    s.iterations = S.iters;
    s.reliable_updates = S.reliable;
    s.relres = S.relres;
    s.true_relres = S.true_relres;
    s.fp64_iterations = method == DWF_SOLVE_CG64 ? S.iters : 0;
    s.converged = ((S.flags[slot] & F_DONE) && S.status == 0) ? 1 : 0;
}

int collect(dwf_ctx* c, int method, int nrhs, dwf_solve_stats* stats)
{
    int rc = DWF_OK;
    for (int i = 0; i < nrhs; ++i) {
        dwf_solve_stats s;
        to_stats(c->st_host[i], c->last_slot, method, s);
        if (c->st_host[i].status == 4) {
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

}  // namespace

// Normal-equation solve A x = b with A = D^dag D or D D^dag (device buffers).
int dwf_solve_device_impl(dwf_ctx* c, int normal, int method, const double* b, const double* x0, double rtol,
                          int maxiter, int nrhs, double* x)
{
    const size_t nv = (size_t)c->vec_sites(nrhs) * 4;
    FullOp op(c, normal == DWF_OP_DDDAG, nrhs, (double*)c->buf(B_T64, nv * 8));
    cg_run(c, op, method == DWF_SOLVE_MIXED, b, x0, rtol, maxiter, nrhs, x);
    return DWF_OK;
}

namespace {

// v <- Gamma v in place, Gamma = sigma3 R (R: s -> Ls-1-s; sigma3 = diag(1, -1) on the
// two spin components), for nrhs spinors.  Gamma^2 = 1 and D^dag = Gamma D Gamma
// (dwf_operator.hpp:94-171), so D D^dag = Gamma (D^dag D) Gamma.
__global__ void k_gamma_r(double* v, long long V4, int Ls, int nrhs)
{
    const long long V5 = V4 * Ls, half = V4 * ((Ls + 1) / 2);
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= half) return;
    const int s = (int)(i / V4);
    const long long site = i - (long long)s * V4;
    double4* p = reinterpret_cast<double4*>(v + (long long)rhs * V5 * 4);
    const long long a = (long long)s * V4 + site, b = (long long)(Ls - 1 - s) * V4 + site;
    const double4 u = p[a], w = p[b];
    p[a] = make_double4(w.x, w.y, -w.z, -w.w);
    if (b != a) p[b] = make_double4(u.x, u.y, -u.z, -u.w);
}

void gamma_r(dwf_ctx* c, double* v, int nrhs)
{
    const long long half = (long long)c->L.V4 * ((c->L.Ls + 1) / 2);
    const dim3 grid((unsigned)((half + kBlock - 1) / kBlock), (unsigned)nrhs);
    k_gamma_r<<<grid, kBlock, 0, c->stream>>>(v, c->L.V4, c->L.Ls, nrhs);
}

// D^dag D x = b by flexible CG with the learned preconditioner (FP32 cycles, FP64
// reliable updates).  If it breaks down (the network is not positive on some residual)
// or misses rtol, plain FP64 CG continues from its iterate and decides; the stats then
// count both (fp64_iterations = the fallback's iterations, 0 without fallback).
// D D^dag x = b is the same solve of the Gamma-rotated system: b' = Gamma b,
// x = Gamma x' (the residual norms are the same, Gamma being unitary).
int solve_learned(dwf_ctx* c, int normal, const double* b, const double* x0, double rtol, int maxiter, int nrhs,
                  double* x, dwf_solve_stats* stats)
{
    const size_t nv = (size_t)c->L.V5() * 4 * nrhs;
    if (x0 && x0 != x) c->check(hipMemcpyAsync(x, x0, nv * 8, hipMemcpyDeviceToDevice, c->stream), "copy x0");
    const bool rot = normal == DWF_OP_DDDAG;
    if (rot) {
        double* gb = (double*)c->buf(B_GAMMA_B, nv * 8);
        c->check(hipMemcpyAsync(gb, b, nv * 8, hipMemcpyDeviceToDevice, c->stream), "copy b");
        gamma_r(c, gb, nrhs);
        if (x0) gamma_r(c, x, nrhs);
        b = gb;
    }
    std::vector<dwf_solve_stats> first(nrhs);
    int rc = dwf_fcg_solve(c, b, x, x0 != nullptr, rtol, maxiter, nrhs, first.data());
    if (rc == DWF_ERR_ZERO_DENOM || rc == DWF_ERR_NOT_CONVERGED) {
        dwf_solve_device_impl(c, DWF_OP_DDAGD, DWF_SOLVE_CG64, b, x, rtol, maxiter, nrhs, x);
        std::vector<dwf_solve_stats> second(nrhs);
        rc = collect(c, DWF_SOLVE_CG64, nrhs, second.data());
        for (int i = 0; i < nrhs; ++i) {  // iterations: the FCG ones; fp64_iterations: the fallback's
            second[i].fp64_iterations = second[i].iterations;
            second[i].iterations = first[i].iterations;
            second[i].reliable_updates = first[i].reliable_updates;
            second[i].passes = first[i].passes;
        }
        first = second;
        if (rc == DWF_ERR_NOT_CONVERGED) rc = DWF_ERR_CERTIFY;
    }
    if (rot) gamma_r(c, x, nrhs);
    c->check_launch("learned solve");
    if (stats) std::copy(first.begin(), first.end(), stats);
    return rc;
}

int solve_normal_eo(dwf_ctx* c, int normal, const double* b, const double* x0, double rtol, int maxiter, int nrhs,
                    double* x, dwf_solve_stats* stats);

int solve_normal(dwf_ctx* c, int normal, int method, const double* b, const double* x0, double rtol, int maxiter,
                 int nrhs, double* x, dwf_solve_stats* stats)
{
    if (!c || !b || !x || nrhs < 1 || maxiter < 0 || !(rtol >= 0.0) ||
        (normal != DWF_OP_DDDAG && normal != DWF_OP_DDAGD) ||
        (method != DWF_SOLVE_CG64 && method != DWF_SOLVE_MIXED && method != DWF_SOLVE_EO_MIXED &&
         method != DWF_SOLVE_LEARNED))
        throw DwfError(DWF_ERR_ARG, "dwf_solve: bad argument");
    if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_solve: no gauge field set");
    c->bind_rhs(nrhs);
    if (method == DWF_SOLVE_LEARNED) {
        c->require_single_lattice("DWF_SOLVE_LEARNED");
        return solve_learned(c, normal, b, x0, rtol, maxiter, nrhs, x, stats);
    }
    if (method == DWF_SOLVE_EO_MIXED) {
        c->require_single_lattice("DWF_SOLVE_EO_MIXED on a normal operator");
        if (eo_supported(c->L)) return solve_normal_eo(c, normal, b, x0, rtol, maxiter, nrhs, x, stats);
        method = DWF_SOLVE_MIXED;  // Wilson, odd extents or an Ls without an even-odd kernel
    }
    dwf_solve_device_impl(c, normal, method, b, x0, rtol, maxiter, nrhs, x);
    int rc = collect(c, method, nrhs, stats);
    if (method == DWF_SOLVE_MIXED && rc == DWF_ERR_NOT_CONVERGED) {
        // The FP32 cycle stalled: plain FP64 CG from the current iterate decides.
        std::vector<dwf_solve_stats> first(nrhs);
        collect(c, method, nrhs, first.data());
        dwf_solve_device_impl(c, normal, DWF_SOLVE_CG64, b, x, rtol, maxiter, nrhs, x);
        rc = collect(c, DWF_SOLVE_CG64, nrhs, stats);
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

// x = D^-1 src through the even-odd Schur complement with defect correction; certified
// on ||D^dag src - D^dag D x|| / ||D^dag src|| <= rtol in FP64.
int solve_dirac_eo(dwf_ctx* c, const double* src, const double* x0, double rtol, int maxiter, int nrhs, double* x,
                   dwf_solve_stats* stats)
{
    hipStream_t s = c->stream;
    const VecLayout layf = c->layout(nrhs, 0);
    const size_t nvf = (size_t)layf.total_sites() * 4, nvh = nvf / 2;  // doubles of the full / half batch
    double* B = (double*)c->buf(B_EO_B, nvf * 8);        // D^dag src (certification rhs)
    double* rsrc = (double*)c->buf(B_EO_RSRC, nvf * 8);  // current Dirac residual src - D x
    double* be = (double*)c->buf(B_EO_BE, nvh * 8);
    double* bo = (double*)c->buf(B_EO_BO, nvh * 8);
    double* ue = (double*)c->buf(B_EO_UE, nvh * 8);
    double* bh = (double*)c->buf(B_EO_BH, nvh * 8);
    double* Bh = (double*)c->buf(B_EO_BHAT, nvh * 8);
    double* xo = (double*)c->buf(B_EO_XO, nvh * 8);
    double* xe = (double*)c->buf(B_EO_XE, nvh * 8);
    double* q64 = (double*)c->buf(B_EO_Q, nvh * 8);
    float* te32 = (float*)c->buf(B_EO_TE32, nvh * 4);
    double* delta = (double*)c->buf(B_EO_DELTA, nvf * 8);
    double* t64 = (double*)c->buf(B_T64, nvf * 8);
    CgState* cert = (CgState*)c->buf(B_STATE2, sizeof(CgState) * nrhs);

    FullOp full(c, false, nrhs, t64);
    double* partC = (double*)c->buf(B_PR2, (size_t)nrhs * full.nblk * 2 * 8);
    if (layf.gr.n > 1) launch_zero<double>(partC, (long long)nrhs * full.nblk * 2, s);  // unwritten partials read 0
    SchurOp schur(c, nrhs, te32, ue, q64);
    // certification rhs B = D^dag src
    {
        HopArgs<double> a = full.args<double>();
        a.in = src;
        a.out = B;
        launch_hop<double>(true, EPI_STORE, a, nrhs, s, c->groups_host());
    }
    if (x0) {
        if (x0 != x) c->check(hipMemcpyAsync(x, x0, nvf * 8, hipMemcpyDeviceToDevice, s), "copy x0");
    } else {
        launch_zero<double>(x, (long long)nvf, s);
    }
    std::vector<int> iters(nrhs, 0), reliable(nrhs, 0);
    std::vector<double> srelres(nrhs, 0.0);
    bool have_cert = false;
    double worst = 1.0;
    const int passes = 4;
    int used = 0;
    bool guess_done = false;
    if (x0) {
        // certify the guess first: the first pass then reduces only what the guess leaves
        // (its tolerance scales with the worst residual), and a right-hand side the guess
        // already solves takes no pass
        full.residual(x, B, delta, nullptr, nullptr, nullptr, partC, nullptr, 0, 0);
        launch_cg_start(full.nblk, partC, cert, rtol, c->reliable_delta, nrhs, s);
        c->check_launch("certification of the guess");
        have_cert = true;
        read_states(c, cert, nrhs);
        worst = 0.0;
        guess_done = true;
        for (int i = 0; i < nrhs; ++i) {
            if (!(c->st_host[i].flags[0] & F_DONE)) {
                guess_done = false;
                worst = std::max(worst, c->st_host[i].relres);
            }
        }
    }
    for (int pass = 0; pass < passes && !guess_done; ++pass) {
        // Dirac residual of the current iterate
        const double* r_in = src;
        if (pass > 0 || x0) {
            HopArgs<double> a = full.args<double>();
            a.in = x;
            a.out = t64;
            launch_hop<double>(false, EPI_STORE, a, nrhs, s, c->groups_host());
            k_sub_into<<<2048, 256, 0, s>>>(src, t64, rsrc, (long long)nvf);
            r_in = rsrc;
        }
        // Schur right-hand side: bh = b_o + 1/2 H_oe A^-1 b_e, Bh = Dhat^dag bh
        launch_eo_split(layf, r_in, be, bo, nrhs, s);
        {
            EoArgs<double> a = schur.base<double>(0);
            a.b = be;
            a.cb = 1.0;
            a.out = ue;
            launch_eo_a<double>(false, true, a, nrhs, s, c->groups_host());
            EoArgs<double> b = schur.base<double>(1);
            b.in = bo;
            b.ca = 1.0;
            b.hin = ue;
            b.ch = 0.5;
            b.out = bh;
            launch_eo_b<double>(false, false, EPI_STORE, b, nrhs, s, c->groups_host());
            EoArgs<double> d = schur.apply_args<double>(true, bh, ue, nullptr, 0, 0);
            d.out = Bh;
            launch_eo_b<double>(true, true, EPI_STORE, d, nrhs, s, c->groups_host());
        }
        c->check_launch("Schur right-hand side");
        const double rtol_pass = std::min(0.5, 0.25 * rtol / std::max(worst, 1e-300));
        cg_run(c, schur, true, Bh, nullptr, rtol_pass, maxiter, nrhs, xo, have_cert ? cert : nullptr);
        used = pass + 1;
        for (int i = 0; i < nrhs; ++i) {
            iters[i] += c->st_host[i].iters;
            reliable[i] += c->st_host[i].reliable;
            srelres[i] = c->st_host[i].true_relres;
        }
        // x_e = A^-1 (b_e + 1/2 H_eo x_o), assemble the correction and add it
        {
            EoArgs<double> a = schur.base<double>(0);
            a.b = be;
            a.cb = 1.0;
            a.hin = xo;
            a.ch = 0.5;
            a.out = xe;
            launch_eo_a<double>(false, true, a, nrhs, s, c->groups_host());
        }
        launch_eo_merge(layf, xe, xo, delta, nrhs, s);
        k_axpy_masked<<<dim3(1024, nrhs), 256, 0, s>>>(x, delta, layf, have_cert ? cert : nullptr);
        // certify on the reference's system (delta is free again)
        full.residual(x, B, delta, nullptr, nullptr, nullptr, partC, nullptr, 0, 0);
        launch_cg_start(full.nblk, partC, cert, rtol, c->reliable_delta, nrhs, s);
        c->check_launch("certification");
        have_cert = true;
        read_states(c, cert, nrhs);
        worst = 0.0;
        bool all = true;
        for (int i = 0; i < nrhs; ++i) {
            if (!(c->st_host[i].flags[0] & F_DONE)) {
                all = false;
                worst = std::max(worst, c->st_host[i].relres);
            }
        }
        if (all) break;
    }
    int rc = DWF_OK;
    for (int i = 0; i < nrhs; ++i) {
        const CgState& S = c->st_host[i];
        dwf_solve_stats st{};
        st.iterations = iters[i];
        st.passes = used;
        st.reliable_updates = reliable[i];
        st.relres = srelres[i];
        st.true_relres = S.relres;
        st.converged = (S.flags[0] & F_DONE) ? 1 : 0;
        if (!st.converged && rc == DWF_OK) {
            dwf_set_error("even-odd solve did not certify within maxiter");
            rc = DWF_ERR_NOT_CONVERGED;
        }
        if (stats) stats[i] = st;
    }
    return rc;
}

int solve_dirac_one(dwf_ctx* c, int method, const double* src, const double* x0, double rtol, int maxiter, int nrhs,
                    double* x, dwf_solve_stats* stats)
{
    if (!c || !src || !x || nrhs < 1 || maxiter < 0 || !(rtol >= 0.0) ||
        (method != DWF_SOLVE_CG64 && method != DWF_SOLVE_MIXED && method != DWF_SOLVE_EO_MIXED &&
         method != DWF_SOLVE_LEARNED))
        throw DwfError(DWF_ERR_ARG, "dwf_solve_dirac: bad argument");
    if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_solve_dirac: no gauge field set");
    c->bind_rhs(nrhs);
    if (method == DWF_SOLVE_LEARNED) c->require_single_lattice("DWF_SOLVE_LEARNED");
    bool eo_ok = eo_supported(c->L);
    for (int g = 0; eo_ok && g < c->gr.n; ++g) eo_ok = (c->gr.Nt[g] % 2 == 0) && (c->gr.Nx[g] % 2 == 0);
    if (method == DWF_SOLVE_EO_MIXED && eo_ok) {
        const int rc = solve_dirac_eo(c, src, x0, rtol, maxiter, nrhs, x, stats);
        if (rc != DWF_ERR_NOT_CONVERGED) return rc;
        method = DWF_SOLVE_MIXED;  // fall back to the full-lattice solver from this iterate
        x0 = x;
    } else if (method == DWF_SOLVE_EO_MIXED) {
        method = DWF_SOLVE_MIXED;  // odd extents or an Ls without an even-odd kernel
    }
    // the reference's pattern: rhs = D^dag src, CG on D^dag D
    const size_t nv = (size_t)c->vec_sites(nrhs) * 4;
    double* B = (double*)c->buf(B_EO_B, nv * 8);
    {
        HopArgs<double> a = ctx_hop_args<double>(c, ctx_hop_cfg(c, nrhs), nrhs);
        a.in = src;
        a.out = B;
        launch_hop<double>(true, EPI_STORE, a, nrhs, c->stream, c->groups_host());
    }
    if (method == DWF_SOLVE_LEARNED) return solve_learned(c, DWF_OP_DDAGD, B, x0, rtol, maxiter, nrhs, x, stats);
    return solve_normal(c, DWF_OP_DDAGD, method, B, x0, rtol, maxiter, nrhs, x, stats);
}

// A batched solve of one lattice as c->streams concurrent sub-batches (contiguous ranges of
// right-hand sides, whole configurations when several are held), each on a sub-context of
// its own from a host thread of its own, writing the caller's buffers in place.  Every
// right-hand side is solved and certified exactly as in one batch; only the batch-level
// schedule (chunk lengths, reliable-update points) is per sub-batch.
int solve_dirac(dwf_ctx* c, int method, const double* src, const double* x0, double rtol, int maxiter, int nrhs,
                double* x, dwf_solve_stats* stats)
{
    const int k = c ? c->streams : 1;
    if (k < 2 || !src || !x || c->borrowed || c->multi() || !c->have_gauge || nrhs < 2 * k)
        return solve_dirac_one(c, method, src, x0, rtol, maxiter, nrhs, x, stats);
    c->bind_rhs(nrhs);
    const int unit = c->ncfg > 1 ? c->rhs_per_cfg : 1;  // right-hand sides per split unit
    const int units = nrhs / unit;
    if (units < k) return solve_dirac_one(c, method, src, x0, rtol, maxiter, nrhs, x, stats);
    const size_t per = (size_t)c->L.V5() * 4;  // doubles per right-hand side
    if (!c->ev) c->check(hipEventCreateWithFlags(&c->ev, hipEventDisableTiming), "split solve: event");
    c->check(hipEventRecord(c->ev, c->stream), "split solve: input event");
    std::vector<dwf_ctx*> kc(k);
    std::vector<int> r0(k + 1);
    for (int i = 0; i <= k; ++i) r0[i] = (int)((long long)units * i / k) * unit;
    for (int i = 0; i < k; ++i) {
        const int cfg0 = c->ncfg > 1 ? r0[i] / unit : 0, nc = c->ncfg > 1 ? (r0[i + 1] - r0[i]) / unit : 1;
        kc[i] = c->kid(i, cfg0, nc);
        c->check(hipStreamWaitEvent(kc[i]->stream, c->ev, 0), "split solve: wait for the inputs");
    }
    std::vector<int> rc(k, DWF_OK), code(k, DWF_OK);
    std::vector<std::string> err(k);
    std::vector<std::thread> th;
    for (int i = 0; i < k; ++i)
        th.emplace_back([&, i] {
            try {
                dwf_ctx::check(hipSetDevice(c->device), "split solve: hipSetDevice");
                const size_t o = (size_t)r0[i] * per;
                rc[i] = solve_dirac_one(kc[i], method, src + o, x0 ? x0 + o : nullptr, rtol, maxiter,
                                        r0[i + 1] - r0[i], x + o, stats ? stats + r0[i] : nullptr);
            } catch (const DwfError& e) {
                code[i] = e.code;
                err[i] = e.what();
            } catch (const std::exception& e) {
                code[i] = DWF_ERR_ARG;
                err[i] = e.what();
            }
        });
    for (auto& t : th) t.join();
    // the caller's stream continues after every sub-batch's work
    for (int i = 0; i < k; ++i) {
        c->check(hipEventRecord(kc[i]->ev, kc[i]->stream), "split solve: output event");
        c->check(hipStreamWaitEvent(c->stream, kc[i]->ev, 0), "split solve: wait for the outputs");
    }
    for (int i = 0; i < k; ++i)
        if (code[i] != DWF_OK) throw DwfError(code[i], err[i]);
    for (int i = 0; i < k; ++i)
        if (rc[i] != DWF_OK) return rc[i];
    return DWF_OK;
}

// A x = b with A = D D^dag or D^dag D through even-odd Dirac solves.  D^dag = Gamma D Gamma
// (Gamma = sigma3 R) gives (D D^dag)^-1 = Gamma D^-1 Gamma D^-1, and D^dag D = Gamma D D^dag
// Gamma turns the second system into the first.  Each pass certifies the iterate on the FP64
// residual of A itself, the system of the reference's CG (domain_wall/include/cg_solver.hpp:
// 17-64), and while a right-hand side misses rtol adds Gamma D^-1 Gamma D^-1 r, the two
// Dirac solves run to a tolerance scaled to what that pass has left: a Dirac solve's
// certified residual bounds the normal one only up to the condition number of D, so one pass
// need not reach rtol.  After four passes the mixed CG (with its FP64 CG fallback) continues
// from the iterate and decides.  stats: iterations and reliable updates summed over every
// Dirac solve, passes = the correction passes, fp64_iterations = the fallback's.
int solve_normal_eo(dwf_ctx* c, int normal, const double* b, const double* x0, double rtol, int maxiter, int nrhs,
                    double* x, dwf_solve_stats* stats)
{
    hipStream_t s = c->stream;
    const size_t nv = (size_t)c->L.V5() * 4 * nrhs;
    if (x0) {
        if (x0 != x) c->check(hipMemcpyAsync(x, x0, nv * 8, hipMemcpyDeviceToDevice, s), "copy x0");
    } else {
        launch_zero<double>(x, (long long)nv, s);
    }
    if (normal == DWF_OP_DDAGD) {  // D^dag D x = b  <=>  D D^dag (Gamma x) = Gamma b
        double* gb = (double*)c->buf(B_GAMMA_B, nv * 8);
        c->check(hipMemcpyAsync(gb, b, nv * 8, hipMemcpyDeviceToDevice, s), "copy b");
        gamma_r(c, gb, nrhs);
        gamma_r(c, x, nrhs);
        b = gb;
    }
    std::vector<dwf_solve_stats> acc(nrhs), one(nrhs);
    auto add = [&]() {
        for (int i = 0; i < nrhs; ++i) {
            acc[i].iterations += one[i].iterations;
            acc[i].reliable_updates += one[i].reliable_updates;
            acc[i].fp64_iterations += one[i].fp64_iterations;
        }
    };
    const int passes = 4;
    int used = 0;
    bool all = false;
    for (int pass = 0;; ++pass) {
        // workspace fetched per pass: the Dirac solves below size buffers of their own
        double* r = (double*)c->buf(B_NEO_R, nv * 8);
        double* y = (double*)c->buf(B_NEO_Y, nv * 8);
        FullOp op(c, true, nrhs, (double*)c->buf(B_T64, nv * 8));
        double* part = (double*)c->buf(B_NEO_PART, (size_t)nrhs * op.nblk * 2 * 8);
        CgState* cert = (CgState*)c->buf(B_NEO_STATE, sizeof(CgState) * nrhs);
        op.residual(x, b, r, nullptr, nullptr, nullptr, part, nullptr, 0, 0);  // r = b - D D^dag x
        launch_cg_start(op.nblk, part, cert, rtol, c->reliable_delta, nrhs, s);
        c->check_launch("even-odd normal solve: certification");
        read_states(c, cert, nrhs);
        double worst = 0.0;
        all = true;
        for (int i = 0; i < nrhs; ++i) {
            acc[i].relres = acc[i].true_relres = c->st_host[i].relres;
            if (!(c->st_host[i].flags[0] & F_DONE)) {
                all = false;
                worst = std::max(worst, c->st_host[i].relres);
            }
        }
        if (all || pass == passes) break;
        // the Dirac solves' outcome is judged by the next certification, not by their codes
        const double tol_in = std::min(0.5, 0.25 * rtol / worst);
        solve_dirac(c, DWF_SOLVE_EO_MIXED, r, nullptr, tol_in, maxiter, nrhs, y, one.data());  // y = D^-1 r
        add();
        gamma_r(c, y, nrhs);
        solve_dirac(c, DWF_SOLVE_EO_MIXED, y, nullptr, tol_in, maxiter, nrhs, r, one.data());  // r = D^-1 Gamma y
        add();
        gamma_r(c, r, nrhs);
        k_axpy_masked<<<dim3(1024, nrhs), 256, 0, s>>>(x, r, c->layout(nrhs, 0), cert);
        c->check_launch("even-odd normal solve: correction");
        used = pass + 1;
    }
    int rc = DWF_OK;
    if (all) {
        for (int i = 0; i < nrhs; ++i) acc[i].converged = 1;
    } else {
        std::vector<dwf_solve_stats> fb(nrhs);
        rc = solve_normal(c, DWF_OP_DDDAG, DWF_SOLVE_MIXED, b, x, rtol, maxiter, nrhs, x, fb.data());
        for (int i = 0; i < nrhs; ++i) {
            acc[i].iterations += fb[i].iterations;
            acc[i].reliable_updates += fb[i].reliable_updates;
            acc[i].fp64_iterations += fb[i].fp64_iterations;
            acc[i].converged = fb[i].converged;
            acc[i].relres = fb[i].relres;
            acc[i].true_relres = fb[i].true_relres;
        }
        if (rc == DWF_ERR_NOT_CONVERGED) rc = DWF_ERR_CERTIFY;
    }
    if (normal == DWF_OP_DDAGD) gamma_r(c, x, nrhs);
    c->check_launch("even-odd normal solve");
    for (int i = 0; i < nrhs; ++i) acc[i].passes = used;
    if (stats) std::copy(acc.begin(), acc.end(), stats);
    return rc;
}

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

}  // namespace

int dwf_solve_learned_impl(dwf_ctx* c, int normal, const double* b, const double* x0, double rtol, int maxiter,
                           int nrhs, double* x, dwf_solve_stats* stats)
{
    return solve_learned(c, normal, b, x0, rtol, maxiter, nrhs, x, stats);
}

int dwf_solve_normal_impl(dwf_ctx* c, int normal, int method, const double* b, const double* x0, double rtol,
                          int maxiter, int nrhs, double* x, dwf_solve_stats* stats)
{
    return solve_normal(c, normal, method, b, x0, rtol, maxiter, nrhs, x, stats);
}

int dwf_solve_dirac_impl(dwf_ctx* c, int method, const double* src, const double* x0, double rtol, int maxiter,
                         int nrhs, double* x, dwf_solve_stats* stats)
{
    return solve_dirac(c, method, src, x0, rtol, maxiter, nrhs, x, stats);
}

namespace {
// The solvers write x while they still read the right-hand side (every residual pass)
// and read a guess before x is first written, so x may not overlap the right-hand side,
// and a guess is either x itself or disjoint from it.
void check_device_operands(dwf_ctx* c, int nrhs, const double* in, const double* x0, const double* x,
                           const char* who)
{
    if (!c || !in || !x || nrhs < 1) return;  // the solver reports these itself
    c->bind_rhs(nrhs);
    const size_t n = (size_t)c->vec_sites(nrhs) * 4;
    auto overlap = [n](const double* a, const double* b) { return a < b + n && b < a + n; };
    if (overlap(in, x))
        throw DwfError(DWF_ERR_ARG, std::string(who) + ": x must not overlap the right-hand side");
    if (x0 && x0 != x && overlap(x0, x))
        throw DwfError(DWF_ERR_ARG, std::string(who) + ": x0 must be x itself or not overlap it");
}
}  // namespace

extern "C" {

int dwf_solve_device(dwf_ctx* c, int normal_op, int method, const double* b, const double* x0, double rtol,
                     int maxiter, int nrhs, double* x, dwf_solve_stats* stats)
{
    return guard([&]() -> int {
        check_device_operands(c, nrhs, b, x0, x, "dwf_solve_device");
        return solve_normal(c, normal_op, method, b, x0, rtol, maxiter, nrhs, x, stats);
    });
}

int dwf_solve(dwf_ctx* c, int normal_op, int method, const double* b, const double* x0, double rtol, int maxiter,
              int nrhs, double* x, dwf_solve_stats* stats)
{
    return guard([&]() -> int {
        if (!c || !b || !x || nrhs < 1) throw DwfError(DWF_ERR_ARG, "dwf_solve: bad argument");
        c->bind_rhs(nrhs);
        const size_t n = (size_t)c->vec_sites(nrhs) * 4;
        double* db = (double*)c->buf(B_HOST_IN, n * 8);
        double* dx = (double*)c->buf(B_HOST_OUT, n * 8);
        c->check(hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, c->stream), "solve upload b");
        if (x0) c->check(hipMemcpyAsync(dx, x0, n * 8, hipMemcpyHostToDevice, c->stream), "solve upload x0");
        const int rc = solve_normal(c, normal_op, method, db, x0 ? dx : nullptr, rtol, maxiter, nrhs, dx, stats);
        if (rc == DWF_OK || rc == DWF_ERR_NOT_CONVERGED || rc == DWF_ERR_CERTIFY) {
            c->check(hipMemcpyAsync(x, dx, n * 8, hipMemcpyDeviceToHost, c->stream), "solve download");
            c->check(hipStreamSynchronize(c->stream), "solve sync");
        }
        return rc;
    });
}

int dwf_solve_dirac_device(dwf_ctx* c, int method, const double* src, const double* x0, double rtol, int maxiter,
                           int nrhs, double* x, dwf_solve_stats* stats)
{
    return guard([&]() -> int {
        check_device_operands(c, nrhs, src, x0, x, "dwf_solve_dirac_device");
        return solve_dirac(c, method, src, x0, rtol, maxiter, nrhs, x, stats);
    });
}

int dwf_solve_dirac(dwf_ctx* c, int method, const double* src, const double* x0, double rtol, int maxiter, int nrhs,
                    double* x, dwf_solve_stats* stats)
{
    return guard([&]() -> int {
        if (!c || !src || !x || nrhs < 1)
            throw DwfError(DWF_ERR_ARG, "dwf_solve_dirac: bad argument");
        c->bind_rhs(nrhs);
        const size_t n = (size_t)c->vec_sites(nrhs) * 4;
        double* db = (double*)c->buf(B_HOST_IN, n * 8);
        double* dx = (double*)c->buf(B_HOST_OUT, n * 8);
        c->check(hipMemcpyAsync(db, src, n * 8, hipMemcpyHostToDevice, c->stream), "solve upload src");
        if (x0) c->check(hipMemcpyAsync(dx, x0, n * 8, hipMemcpyHostToDevice, c->stream), "solve upload x0");
        const int rc = solve_dirac(c, method, db, x0 ? dx : nullptr, rtol, maxiter, nrhs, dx, stats);
        if (rc == DWF_OK || rc == DWF_ERR_NOT_CONVERGED || rc == DWF_ERR_CERTIFY) {
            c->check(hipMemcpyAsync(x, dx, n * 8, hipMemcpyDeviceToHost, c->stream), "solve download");
            c->check(hipStreamSynchronize(c->stream), "solve sync");
        }
        return rc;
    });
}

}  // extern "C"
