// dwf_device.hip -- gfx950 kernels of the Wilson / Shamir domain-wall hot path.
//
// Operator (reference domain_wall/include/dwf_operator.hpp:15-171 and
// wilson/include/wilson_operator.hpp:10-62):
//   (D psi)(s,x) = diag psi - 1/2 sum_mu [ bc U_mu(x) (1 - g.gamma_mu) psi(x+mu)
//                                        + bc U_mu(x-mu)^* (1 + g.gamma_mu) psi(x-mu) ] + D5
// with gamma_t = sigma1, gamma_x = sigma2, g = +1 for D and -1 for D^dagger, and the
// Shamir fifth-dimension hops D5 (dwf_operator.hpp:71-86, 150-165).
//
// Layout in HBM: spinors keep the reference order [rhs][s][t][x][spin] with the two
// complex spin components of a site packed in one float4 / double4, so a wavefront's
// 64 consecutive 4D sites load 1 KiB (FP32) contiguously.  Links are [mu][t][x] complex
// (cos theta, sin theta), evaluated once per gauge update, never per slice.
//
// One thread owns one 4D site and walks `spt` consecutive s-slices: the four links are
// loaded once and reused for every slice, and the fifth-dimension neighbours come from
// a register window (the previous and next centre spinors), so each slice costs the
// centre + four 4D neighbours (the neighbours hit L1/L2; the algorithmic traffic is one
// read of the input, one write of the output and one read of the links).
//
// Reductions are deterministic: fixed wavefront shuffle trees, one partial per block,
// partials summed in a fixed order by whichever kernel consumes them.

#include "dwf_context.hpp"
#include "dwf_internal.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace dwfk {
namespace {

template <int V> using IC = std::integral_constant<int, V>;

template <typename R> struct Vec4;
template <> struct Vec4<float> { using T = float4; };
template <> struct Vec4<double> { using T = double4; };
template <typename R> struct Cx;
template <> struct Cx<float> { using T = float2; };
template <> struct Cx<double> { using T = double2; };

// Two complex spin components of one site: (a_re, a_im) spin 0, (b_re, b_im) spin 1.
template <typename R> struct Sp {
    R ar, ai, br, bi;
};

template <typename R> __device__ __forceinline__ Sp<R> ldsp(const R* base, long long i)
{
    const typename Vec4<R>::T v = reinterpret_cast<const typename Vec4<R>::T*>(base)[i];
    return Sp<R>{v.x, v.y, v.z, v.w};
}

template <typename R> __device__ __forceinline__ void stsp(R* base, long long i, const Sp<R>& s)
{
    typename Vec4<R>::T v;
    v.x = s.ar;
    v.y = s.ai;
    v.z = s.br;
    v.w = s.bi;
    reinterpret_cast<typename Vec4<R>::T*>(base)[i] = v;
}

// Streaming store (no L2 allocation on the way to HBM) for outputs nobody re-reads soon.
template <typename R> __device__ __forceinline__ void stsp_nt(R* base, long long i, const Sp<R>& s)
{
    typedef R nv4 __attribute__((ext_vector_type(4)));
    nv4 v;
    v.x = s.ar;
    v.y = s.ai;
    v.z = s.br;
    v.w = s.bi;
    __builtin_nontemporal_store(v, reinterpret_cast<nv4*>(base) + i);
}

// c ? a : b field by field (a select of whole structs can put both in scratch memory)
template <typename R> __device__ __forceinline__ Sp<R> pick(bool c, const Sp<R>& a, const Sp<R>& b)
{
    return Sp<R>{c ? a.ar : b.ar, c ? a.ai : b.ai, c ? a.br : b.br, c ? a.bi : b.bi};
}

template <typename R> __device__ __forceinline__ double norm2(const Sp<R>& s)
{
    const double a = s.ar, b = s.ai, c = s.br, d = s.bi;
    return a * a + b * b + c * c + d * d;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum of two accumulators; the result is valid in thread 0.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* sh)
{
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[w] = a;
        sh[4 + w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = (sh[0] + sh[1]) + (sh[2] + sh[3]);
        b = (sh[4] + sh[5]) + (sh[6] + sh[7]);
    }
}

// Sum of element `which` of n (value, value) partial pairs, same result in every block
// (fixed order).  Every thread of the block must call it.
__device__ __forceinline__ double sum_partials(const double* p, int n, int which, double* sh)
{
    if (threadIdx.x < 64) {
        double v = 0.0;
        for (int i = threadIdx.x; i < n; i += 64) v += p[2 * i + which];
        v = wave_sum(v);
        if (threadIdx.x == 0) sh[0] = v;
    }
    __syncthreads();
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Both sums of the (pair) partials p[2 i], p[2 i + 1], i < n, in fixed order, in one pass.
__device__ __forceinline__ void sum_partials2(const double* p, int n, double* sh, double& s0, double& s1)
{
    if (threadIdx.x < 64) {
        double v0 = 0.0, v1 = 0.0;
        for (int i = threadIdx.x; i < n; i += 64) {
            v0 += p[2 * i];
            v1 += p[2 * i + 1];
        }
        v0 = wave_sum(v0);
        v1 = wave_sum(v1);
        if (threadIdx.x == 0) {
            sh[0] = v0;
            sh[1] = v1;
        }
    }
    __syncthreads();
    s0 = sh[0];
    s1 = sh[1];
    __syncthreads();
}

// Block sum of two accumulators over NT threads (NT/64 wavefronts, fixed order); the
// result is valid in thread 0.  sh holds 2 * NT / 64 doubles.
template <int NT> __device__ __forceinline__ void block_sum2_n(double& a, double& b, double* sh)
{
    constexpr int nw = NT / 64;
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[w] = a;
        sh[nw + w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int i = 0; i < nw; ++i) {
            x += sh[i];
            y += sh[nw + i];
        }
        a = x;
        b = y;
    }
}

// ------------------------------------------------------------------ stencil kernels
// Per-thread constants of one 4D site: its four links and boundary signs.
template <typename R> struct SiteK {
    typename Cx<R>::T ut, ux, utm, uxm;
    R stp, stm, sxp, sxm;
};

// Links of right-hand side rhs of a single-lattice launch.
template <typename R> __device__ __forceinline__ const typename Cx<R>::T* rhs_links(const HopArgs<R>& a, int rhs)
{
    using C = typename Cx<R>::T;
    const int cfg = a.cfg_map ? a.cfg_map[rhs] : rhs / a.rhs_per_cfg;
    return static_cast<const C*>(a.links) + (long long)cfg * a.link_stride;
}

template <typename R>
__device__ __forceinline__ SiteK<R> site_consts(const Lat& L, const typename Cx<R>::T* lk, int t, int x)
{
    const int Nx = L.Nx, Nt = L.Nt, V4 = L.V4;
    const int tm = (t == 0) ? Nt - 1 : t - 1;
    const int xm = (x == 0) ? Nx - 1 : x - 1;
    const int site = t * Nx + x;
    SiteK<R> k;
    k.ut = lk[site];
    k.ux = lk[V4 + site];
    k.utm = lk[tm * Nx + x];
    k.uxm = lk[V4 + t * Nx + xm];
    k.stp = (t == Nt - 1) ? (R)L.bct : (R)1;
    k.stm = (t == 0) ? (R)L.bct : (R)1;
    k.sxp = (x == Nx - 1) ? (R)L.bcx : (R)1;
    k.sxm = (x == 0) ? (R)L.bcx : (R)1;
    return k;
}

// The operator at one site of slice s from the centre, its four 4D neighbours and (D5)
// its fifth-dimension neighbours.
template <typename R, bool DAG, bool D5>
__device__ __forceinline__ Sp<R> hop_site(const Lat& L, const SiteK<R>& k, const Sp<R>& cur, const Sp<R>& vtp,
                                          const Sp<R>& vtm, const Sp<R>& vxp, const Sp<R>& vxm, const Sp<R>& dn,
                                          const Sp<R>& up, int s)
{
    const R diag = (R)L.diag, mf = (R)L.mf, h = (R)-0.5;
    constexpr R sg = DAG ? (R)-1 : (R)1;
    const R stp = k.stp, stm = k.stm, sxp = k.sxp, sxm = k.sxm;
    const auto ut = k.ut, ux = k.ux, utm = k.utm, uxm = k.uxm;
    (void)dn;
    (void)up;
    (void)mf;
    Sp<R> o{diag * cur.ar, diag * cur.ai, diag * cur.br, diag * cur.bi};
    {   // forward t: U_t(x) (1 - g sigma1) psi(x+t): upper a = v0 - g v1, lower -g a
        const R ar = stp * (vtp.ar - sg * vtp.br), ai = stp * (vtp.ai - sg * vtp.bi);
        const R hr = ut.x * ar - ut.y * ai, hi = ut.x * ai + ut.y * ar;
        o.ar += h * hr;
        o.ai += h * hi;
        o.br -= h * sg * hr;
        o.bi -= h * sg * hi;
    }
    {   // backward t: U_t(x-t)^* (1 + g sigma1) psi(x-t): upper b = v0 + g v1, lower g b
        const R br = stm * (vtm.ar + sg * vtm.br), bi = stm * (vtm.ai + sg * vtm.bi);
        const R hr = utm.x * br + utm.y * bi, hi = utm.x * bi - utm.y * br;
        o.ar += h * hr;
        o.ai += h * hi;
        o.br += h * sg * hr;
        o.bi += h * sg * hi;
    }
    {   // forward x: U_x(x) (1 - g sigma2) psi(x+x): upper c = v0 + g i v1, lower -g i c
        const R cr = sxp * (vxp.ar - sg * vxp.bi), ci = sxp * (vxp.ai + sg * vxp.br);
        const R hr = ux.x * cr - ux.y * ci, hi = ux.x * ci + ux.y * cr;
        o.ar += h * hr;
        o.ai += h * hi;
        o.br += h * sg * hi;   // -g i (hr + i hi) = g hi - i g hr
        o.bi -= h * sg * hr;
    }
    {   // backward x: U_x(x-x)^* (1 + g sigma2) psi(x-x): upper d = v0 - g i v1, lower g i d
        const R dr = sxm * (vxm.ar + sg * vxm.bi), di = sxm * (vxm.ai - sg * vxm.br);
        const R hr = uxm.x * dr + uxm.y * di, hi = uxm.x * di - uxm.y * dr;
        o.ar += h * hr;
        o.ai += h * hi;
        o.br -= h * sg * hi;   // g i (hr + i hi) = -g hi + i g hr
        o.bi += h * sg * hr;
    }
    if constexpr (D5) {
        // D:   spin1 <- s+1 (wall: +mf psi(0)),    spin0 <- s-1 (wall: +mf psi(Ls-1))
        // D^+: spin1 <- s-1 (wall: +mf psi(Ls-1)), spin0 <- s+1 (wall: +mf psi(0))
        const R cup = (s + 1 == L.Ls) ? mf : (R)-1;
        const R cdn = (s == 0) ? mf : (R)-1;
        if constexpr (!DAG) {
            o.br += cup * up.br;
            o.bi += cup * up.bi;
            o.ar += cdn * dn.ar;
            o.ai += cdn * dn.ai;
        } else {
            o.br += cdn * dn.br;
            o.bi += cdn * dn.bi;
            o.ar += cup * up.ar;
            o.ai += cup * up.ai;
        }
    }

    return o;
}

// Epilogue at spinor index idx of right-hand side offset rb.
template <typename R, int EPI>
__device__ __forceinline__ void hop_epi(const HopArgs<R>& a, long long rb, long long idx, const Sp<R>& o, R alpha,
                                        double& acc0, double& acc1)
{
    (void)alpha;
    (void)acc1;
    if constexpr (EPI == EPI_STORE) {
        if (a.cfg.nt_store)
            stsp_nt(a.out + rb, idx, o);
        else
            stsp(a.out + rb, idx, o);
    } else if constexpr (EPI == EPI_NORM) {
        if (a.cfg.nt_store)
            stsp_nt(a.out + rb, idx, o);
        else
            stsp(a.out + rb, idx, o);
        acc0 += norm2(o);
    } else if constexpr (EPI == EPI_CGUPD) {
        if (a.out) {
            if (a.cfg.nt_store)
                stsp_nt(a.out + rb, idx, o);
            else
                stsp(a.out + rb, idx, o);
        }
        const Sp<R> pv = ldsp(a.p + rb, idx);
        Sp<R> xv = ldsp(a.x + rb, idx);
        Sp<R> rv = ldsp(a.r + rb, idx);
        xv.ar += alpha * pv.ar;
        xv.ai += alpha * pv.ai;
        xv.br += alpha * pv.br;
        xv.bi += alpha * pv.bi;
        rv.ar -= alpha * o.ar;
        rv.ai -= alpha * o.ai;
        rv.br -= alpha * o.br;
        rv.bi -= alpha * o.bi;
        stsp(a.x + rb, idx, xv);
        stsp(a.r + rb, idx, rv);
        acc0 += norm2(rv);
    } else {  // EPI_RESID (FP64 only)
        if constexpr (sizeof(R) == 8) {
            const Sp<double> bv = ldsp(a.b + rb, idx);
            const Sp<double> rv{bv.ar - o.ar, bv.ai - o.ai, bv.br - o.br, bv.bi - o.bi};
            stsp(a.r64 + rb, idx, rv);
            if (a.p64) stsp(a.p64 + rb, idx, rv);
            if (a.r32 || a.p32) {
                const Sp<float> fv{(float)rv.ar, (float)rv.ai, (float)rv.br, (float)rv.bi};
                if (a.r32) stsp(a.r32 + rb, idx, fv);
                if (a.p32) stsp(a.p32 + rb, idx, fv);
            }
            acc0 += norm2(rv);
            acc1 += norm2(bv);
        }
    }

}

// MG: a launch over several lattices (a.groups set); the single-lattice instantiations
// carry none of the group code.
template <typename R, bool DAG, bool D5, int EPI, bool PF, bool MG = false>
__global__ __launch_bounds__(kBlock) void k_hop(const HopArgs<R> a)
{
    // 1D grid over (rhs, s-chunk, tile) -- over (lattice group, rhs, s-chunk, block) with
    // several lattices -- optionally in XCD-aware order
    const int total = gridDim.x;
    const int lb = a.cfg.remap ? xcd_logical_block(blockIdx.x, total) : (int)blockIdx.x;
    Lat L = a.L;
    int rhs, bid, bx, pstride;
    long long rb;                       // this right-hand side's spinors (offset in reals)
    const typename Cx<R>::T* lk;        // and links
    if constexpr (!MG) {
        const int nblk = a.cfg.nblk();
        rhs = lb / nblk;
        bid = lb - rhs * nblk;
        bx = a.cfg.bx;
        pstride = nblk;
        rb = (long long)rhs * L.V5() * 4;
        lk = rhs_links(a, rhs);
    } else {
        const int chunks = a.cfg.chunks;
        int blk0, nblk;
        const GroupPos p = block_group(a.groups, lb, [chunks](int nt, int nx) { return (nt * nx + kBlock - 1) / kBlock * chunks; },
                                       blk0, nblk);
        const int local = lb - blk0, rl = local / nblk;  // right-hand side within the group
        rhs = p.rhs0 + rl;
        bid = local - rl * nblk;
        bx = nblk / chunks;
        pstride = a.pstride;
        L.Nt = p.Nt;
        L.Nx = p.Nx;
        L.V4 = L.Nt * L.Nx;
        rb = (p.site0 + (long long)rl * L.V5()) * 4;
        lk = static_cast<const typename Cx<R>::T*>(a.links) + p.link0 + (long long)(rl / p.rpc) * 2 * L.V4;
    }
    const int chunk = bid / bx;
    const int tile = bid - chunk * bx;
    if (a.st != nullptr && (a.st[rhs].flags[a.slot] & a.skip)) return;  // block-uniform

    __shared__ double sh[8];

    R alpha = 0;
    if constexpr (EPI == EPI_CGUPD) {
        const double pAp = sum_partials(a.partA + (size_t)rhs * pstride * 2, pstride, 0, sh);
        if (pAp == 0.0 || (a.anum && !(pAp > 0.0))) return;  // block-uniform; a later kernel records the failure
        alpha = (R)((a.anum ? a.anum[(size_t)rhs * a.anum_stride] : a.st[rhs].rr[a.slot]) / pAp);
    }

    const int V4 = L.V4;
    int t, x;
    if (a.cfg.tx) {  // 2D tile: (kBlock/tx) rows x tx columns
        const int tt = tile / a.cfg.tiles_x, tc = tile - tt * a.cfg.tiles_x;
        const int lt = threadIdx.x / a.cfg.tx;
        t = tt * (kBlock / a.cfg.tx) + lt;
        x = tc * a.cfg.tx + (threadIdx.x - lt * a.cfg.tx);
    } else {
        const int lin = tile * kBlock + threadIdx.x;
        t = lin / L.Nx;
        x = lin - t * L.Nx;
    }
    const int site = t * L.Nx + x;
    double acc0 = 0.0, acc1 = 0.0;

    const int s0 = chunk * a.cfg.spt;
    const int s1 = min(L.Ls, s0 + a.cfg.spt);
    if (t < L.Nt && s0 < s1) {
        const int Nx = L.Nx, Nt = L.Nt;
        const int tp = (t + 1 == Nt) ? 0 : t + 1;
        const int tm = (t == 0) ? Nt - 1 : t - 1;
        const int xp = (x + 1 == Nx) ? 0 : x + 1;
        const int xm = (x == 0) ? Nx - 1 : x - 1;
        const SiteK<R> k = site_consts<R>(L, lk, t, x);
        const int ntp = tp * Nx + x, ntm = tm * Nx + x, nxp = t * Nx + xp, nxm = t * Nx + xm;

        const R* in = a.in + rb;
        // PF (software pipeline): the loads of slice s+1 (four neighbours) and of the
        // centre of slice s+2 are issued before slice s is computed, so every wave keeps
        // two slices of memory traffic in flight, at the price of more registers.
        const int Ls = L.Ls;
        auto wrap = [Ls](int s) { return s >= Ls ? s - Ls : (s < 0 ? s + Ls : s); };
        Sp<R> cur = ldsp(in, (long long)s0 * V4 + site);
        Sp<R> dn{0, 0, 0, 0}, up{0, 0, 0, 0};
        if constexpr (D5) dn = ldsp(in, (long long)wrap(s0 - 1) * V4 + site);
        if (D5 || s0 + 1 < s1) up = ldsp(in, (long long)wrap(s0 + 1) * V4 + site);
        long long sl0 = (long long)s0 * V4;
        Sp<R> vtp{0, 0, 0, 0}, vtm{0, 0, 0, 0}, vxp{0, 0, 0, 0}, vxm{0, 0, 0, 0};
        if constexpr (PF) {
            vtp = ldsp(in, sl0 + ntp);
            vtm = ldsp(in, sl0 + ntm);
            vxp = ldsp(in, sl0 + nxp);
            vxm = ldsp(in, sl0 + nxm);
        }

        for (int s = s0; s < s1; ++s) {
            const long long sl = (long long)s * V4;
            const bool more = s + 1 < s1;
            Sp<R> up2{0, 0, 0, 0}, ntp_{0, 0, 0, 0}, ntm_{0, 0, 0, 0}, nxp_{0, 0, 0, 0}, nxm_{0, 0, 0, 0};
            if constexpr (!PF) {
                vtp = ldsp(in, sl + ntp);
                vtm = ldsp(in, sl + ntm);
                vxp = ldsp(in, sl + nxp);
                vxm = ldsp(in, sl + nxm);
                if (D5 ? s + 2 <= s1 : s + 2 < s1) up2 = ldsp(in, (long long)wrap(s + 2) * V4 + site);
            } else if (more) {
                const long long sn = sl + V4;
                ntp_ = ldsp(in, sn + ntp);
                ntm_ = ldsp(in, sn + ntm);
                nxp_ = ldsp(in, sn + nxp);
                nxm_ = ldsp(in, sn + nxm);
                if (D5 ? s + 2 <= s1 : s + 2 < s1) up2 = ldsp(in, (long long)wrap(s + 2) * V4 + site);
            }

            const Sp<R> o = hop_site<R, DAG, D5>(L, k, cur, vtp, vtm, vxp, vxm, dn, up, s);
            hop_epi<R, EPI>(a, rb, sl + site, o, alpha, acc0, acc1);

            dn = cur;
            cur = up;
            up = up2;
            if constexpr (PF) {
                vtp = ntp_;
                vtm = ntm_;
                vxp = nxp_;
                vxm = nxm_;
            }
        }
    }

    if constexpr (EPI != EPI_STORE) {
        block_sum2(acc0, acc1, sh);
        if (threadIdx.x == 0) {
            double* pp = a.part + ((size_t)rhs * pstride + bid) * 2;
            pp[0] = acc0;
            pp[1] = acc1;
        }
    }
}


// Raw buffer access.  A lane whose byte offset lies past the end of the buffer reads
// zeros and writes nothing, so optional loads and stores need no branch (a branch around
// a load makes the compiler wait for every load in flight where the paths join).
constexpr int kBufOff = 0x7ffffff0;  // past the end of every buffer (all are < 2 GiB)
constexpr int kRsrcWord3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, long long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, p ? (int)bytes : 0, kRsrcWord3);
}
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <typename R> __device__ __forceinline__ Sp<R> bld(__amdgpu_buffer_rsrc_t r, int off)
{
    if constexpr (sizeof(R) == 4) {
        const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        return Sp<R>{__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w)};
    } else {
        const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, 0);
        auto d = [](unsigned lo, unsigned hi) {
            return __hiloint2double((int)hi, (int)lo);
        };
        return Sp<R>{(R)d(v.x, v.y), (R)d(v.z, v.w), (R)d(w.x, w.y), (R)d(w.z, w.w)};
    }
}
// AUX 2: streaming (non-temporal) store
template <typename R, int AUX = 0> __device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, int off, const Sp<R>& s)
{
    if constexpr (sizeof(R) == 4) {
        const u32x4 v = {__float_as_uint(s.ar), __float_as_uint(s.ai), __float_as_uint(s.br), __float_as_uint(s.bi)};
        __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, AUX);
    } else {
        auto lo = [](double x) { return (unsigned)__double2loint(x); };
        auto hi = [](double x) { return (unsigned)__double2hiint(x); };
        const u32x4 v = {lo(s.ar), hi(s.ar), lo(s.ai), hi(s.ai)};
        const u32x4 w = {lo(s.br), hi(s.br), lo(s.bi), hi(s.bi)};
        __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, AUX);
        __builtin_amdgcn_raw_buffer_store_b128(w, r, off + 16, 0, AUX);
    }
}

// LDS-tiled variant: a block owns a full 2D tile of ty x tx sites (ty = NT / tx) of one
// right-hand side and walks its s-chunk.  Each slice of the tile plus its one-site halo
// is staged in LDS once (double-buffered, one barrier per slice): a site costs one
// global load of its own spinor plus (2 tx + 2 ty) / NT halo loads instead of five.  The
// fifth-dimension window (previous, current, next slice) stays in registers.  Pipeline:
// at the step of slice s the block loads its own sites of slice s + 2 and the halo of
// slice s + 2, and stages slice s + 1 (loaded one step earlier) into LDS, so every load
// has a full step to arrive.  The walk is unrolled four times, so the window's four
// register sets and the two halo sets change roles at compile time (no copies of
// registers a load is still writing), and every load and store is unconditional: steps
// past the chunk's end and lanes without a halo site use out-of-range buffer offsets.
// The output is always stored streaming (non-temporal): a choice between two store
// kinds would be a branch again.
// Each chunk fetches its spt slices plus the two boundary slices of the fifth-dimension
// hop, and no more; odd chunks walk downwards so that neighbouring chunks fetch their
// shared boundary slices at the same time (the second fetch hits L2).
// Needs 8 <= tx <= 128, tx dividing NT and Nx, Nt % (NT/tx) == 0 (every thread owns a
// site) and 2 tx + 2 NT/tx <= NT (every halo site has a loader).  A tile as wide as the
// lattice (tx = Nx) has no x halo of its own: the wrapped columns are its own edges.
constexpr bool lds_tile_ok(int nt, int tx)
{
    return tx >= 8 && tx <= 128 && nt % tx == 0 && 2 * tx + 2 * (nt / tx) <= nt;
}
// Largest (ty+2)(tx+2) over the tile widths lds_tile_ok accepts.
constexpr int lds_max(int nt)
{
    int m = 0;
    for (int tx = 8; tx <= 128; ++tx)
        if (lds_tile_ok(nt, tx)) {
            const int v = (nt / tx + 2) * (tx + 2);
            m = v > m ? v : m;
        }
    return m;
}
// Threads per block of LDS tile mode 1..6.
constexpr int lds_threads(int mode)
{
    return mode == 2 ? 512 : mode == 3 ? 64 : mode == 4 ? 128 : mode == 5 ? 384 : mode == 6 ? 768 : kBlock;
}

// s_waitcnt lgkmcnt(0) (gfx9 encoding; vmcnt and expcnt unconstrained)
constexpr int kWaitLgkm0 = 0xC07F;
// Block barrier that first completes this wave's LDS writes.  __syncthreads() alone is
// not enough at the top of a loop: the compiler drops the LDS wait of its fence on the
// path that enters the loop header over the back-edge, so a wave could pass the barrier
// while another wave's writes of the previous step are still in flight.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    __syncthreads();
}

template <typename R, bool DAG, bool D5, int EPI, int NT>
__global__ __launch_bounds__(NT) void k_hop_lds(const HopArgs<R> a)
{
    const int nblk = a.cfg.nblk();
    const int total = gridDim.x;
    const int lb = a.cfg.remap ? xcd_logical_block(blockIdx.x, total) : (int)blockIdx.x;
    const int rhs = lb / nblk;
    const int bid = lb - rhs * nblk;
    const int chunk = bid / a.cfg.bx;
    const int tile = bid - chunk * a.cfg.bx;
    const Lat& L = a.L;
    if (a.st != nullptr && (a.st[rhs].flags[a.slot] & a.skip)) return;  // block-uniform

    __shared__ double sh[2 * NT / 64 > 16 ? 2 * NT / 64 : 16];
    using V = typename Vec4<R>::T;
    constexpr int kSlots = lds_max(NT) + 1;  // + one sink slot for lanes without a halo site
    __shared__ V lds[2][kSlots];

    R alpha = 0;
    constexpr bool kPn = EPI == EPI_PNORM;
    R beta = 0, alpha_prev = 0;
    bool use_p = false, use_x = false;

    const int V4 = L.V4, Nx = L.Nx, Nt = L.Nt, Ls = L.Ls;
    const long long V5 = L.V5();
    const long long rb = V5 * 4 * rhs;
    constexpr int VB = (int)sizeof(V);
    const long long nbytes = V5 * VB;
    const __amdgpu_buffer_rsrc_t rin = make_rsrc(a.in + rb, nbytes);
    const __amdgpu_buffer_rsrc_t rpo = make_rsrc(kPn ? a.pold + rb : nullptr, nbytes);
    const __amdgpu_buffer_rsrc_t rpn = make_rsrc(kPn ? a.pnew + rb : nullptr, nbytes);
    // EPI_PNORM: the stencil's input is p = z + beta p_old (z when beta is 0: p_old may hold
    // anything then).  z and p_old are loaded into separate registers and combined where the
    // value is first used, a step later, so no instruction waits for a load in flight.
    auto comb = [&](const Sp<R>& z, const Sp<R>& q) {
        return Sp<R>{z.ar + (use_p ? beta * q.ar : (R)0), z.ai + (use_p ? beta * q.ai : (R)0),
                     z.br + (use_p ? beta * q.br : (R)0), z.bi + (use_p ? beta * q.bi : (R)0)};
    };
    const int tx = a.cfg.tx, ty = NT / tx, W = tx + 2;
    const int tt = tile / a.cfg.tiles_x, tc = tile - tt * a.cfg.tiles_x;
    const int lt = threadIdx.x / tx, lx = threadIdx.x - lt * tx;
    const int t0 = tt * ty, x0 = tc * tx;
    const int t = t0 + lt, x = x0 + lx;
    const int site = t * Nx + x;
    const int own = (lt + 1) * W + lx + 1;
    // this thread's halo site (2 tx + 2 ty <= NT of them), or none (then its halo
    // loads are out of range and it stages them into the sink slot)
    int hsite = -1, hslot = kSlots - 1;
    {
        const int hh = threadIdx.x;
        const int tb = t0 == 0 ? Nt - 1 : t0 - 1, ta = t0 + ty == Nt ? 0 : t0 + ty;
        const int xl = x0 == 0 ? Nx - 1 : x0 - 1, xr = x0 + tx == Nx ? 0 : x0 + tx;
        if (hh < tx) {
            hsite = tb * Nx + x0 + hh;
            hslot = hh + 1;
        } else if (hh < 2 * tx) {
            hsite = ta * Nx + x0 + (hh - tx);
            hslot = (ty + 1) * W + (hh - tx) + 1;
        } else if (hh < 2 * tx + ty) {
            hsite = (t0 + hh - 2 * tx) * Nx + xl;
            hslot = (hh - 2 * tx + 1) * W;
        } else if (hh < 2 * tx + 2 * ty) {
            hsite = (t0 + hh - 2 * tx - ty) * Nx + xr;
            hslot = (hh - 2 * tx - ty + 1) * W + tx + 1;
        }
    }
    const SiteK<R> k = site_consts<R>(L, rhs_links(a, rhs), t, x);
    double acc0 = 0.0, acc1 = 0.0;
    const int s0 = chunk * a.cfg.spt;
    const int s1 = min(Ls, s0 + a.cfg.spt);
    auto wrap = [Ls](int s) { return s >= Ls ? s - Ls : (s < 0 ? s + Ls : s); };
    auto put = [&](int b, int slot, const Sp<R>& v) {
        V q;
        q.x = v.ar;
        q.y = v.ai;
        q.z = v.br;
        q.w = v.bi;
        lds[b][slot] = q;
    };
    auto get = [&](int b, int slot) {
        const V q = lds[b][slot];
        return Sp<R>{q.x, q.y, q.z, q.w};
    };
    auto off_own = [&](int s) { return (s * V4 + site) * VB; };
    auto off_halo = [&](int s) { return hsite >= 0 ? (s * V4 + hsite) * VB : kBufOff; };
    // Odd chunks walk their slices downwards.  Step k computes slice s_k = sa + k dir.
    const int n = s1 - s0;
    const int dir = (chunk & 1) ? -1 : 1;
    const int sa = dir > 0 ? s0 : s1 - 1;
    // One chunk spanning all of s (then dir = +1, sa = 0): the wall hop needs psi(Ls-1)
    // first and psi(0) last, the slices the walk starts from -- keep them instead of
    // fetching them twice.
    const bool whole = D5 && n == Ls;
    const int s_prev0 = wrap(sa - dir);

    // window: at step k (d = k mod 4) win[d] = psi(s_k), win[d+3] = psi(s_k - dir),
    // win[d+1] = psi(s_k + dir), win[d+2] receives psi(s_k + 2 dir); halo: hreg[d & 1]
    // holds the halo of slice s_k + dir, hreg[(d + 1) & 1] receives that of s_k + 2 dir
    Sp<R> win[4], hreg[2];
    Sp<R> winp[4], hregp[2];  // EPI_PNORM: the p_old halves of loads in flight
    Sp<R> h0;
    {
        // The prologue's loads go out first; the block's scalars (alpha, beta: reductions
        // over the previous kernels' partials) are folded while they are in flight.
        const int o0 = off_own(sa), o3 = D5 ? off_own(s_prev0) : kBufOff;
        const int o1 = (D5 || n > 1) ? off_own(wrap(sa + dir)) : kBufOff;
        const int q0 = off_halo(sa), q1 = n > 1 ? off_halo(sa + dir) : kBufOff;
        win[0] = bld<R>(rin, o0);
        win[3] = bld<R>(rin, o3);
        win[1] = bld<R>(rin, o1);
        h0 = bld<R>(rin, q0);
        hreg[0] = bld<R>(rin, q1);
        Sp<R> p0, p3, p1, ph0, ph1;
        if constexpr (kPn) {
            p0 = bld<R>(rpo, o0);
            p3 = bld<R>(rpo, o3);
            p1 = bld<R>(rpo, o1);
            ph0 = bld<R>(rpo, q0);
            ph1 = bld<R>(rpo, q1);
        }
        if constexpr (EPI == EPI_CGUPD || EPI == EPI_RUPD) {
            const double pAp = sum_partials(a.partA + (size_t)rhs * nblk * 2, nblk, 0, sh);
            const bool ok = !(pAp == 0.0 || (a.anum && !(pAp > 0.0)));
            const double al = ok ? (a.anum ? a.anum[(size_t)rhs * a.anum_stride] : a.st[rhs].rr[a.slot]) / pAp : 0.0;
            if (a.aout && bid == 0 && threadIdx.x == 0) a.aout[(size_t)rhs * a.anum_stride] = al;
            if (!ok) return;  // block-uniform; a later kernel records the failure
            alpha = (R)al;
        }
        if constexpr (kPn) {
            // beta of the direction update p = z + beta p_old, from the network's partials; every
            // block folds them in the same order, block 0 records (z, r) and beta
            double zr, zq;
            sum_partials2(a.zpart + (size_t)rhs * a.nz * 2, a.nz, sh, zr, zq);
            if (!(zr > 0.0)) {  // the preconditioner is not positive on this residual
                if (bid == 0 && threadIdx.x == 0) {
                    a.stw[rhs].status = 5;
                    a.stw[rhs].flags[0] |= F_DONE;
                }
                return;  // block-uniform
            }
            const FcgScal& f = a.fs[rhs];
            const double b = a.fr ? zr / f.rz_prev : -zq / f.pq;
            const double ap = f.alpha;  // the previous iteration's step (0: none pending)
            if (bid == 0 && threadIdx.x == 0) {
                a.fs[rhs].rz = zr;
                a.fs[rhs].beta = b;
                a.fs[rhs].pbuf = a.pbuf;
            }
            beta = (R)b;
            use_p = b != 0.0;
            alpha_prev = (R)ap;
            use_x = a.x != nullptr && ap != 0.0;
        }
        if constexpr (kPn) {
            win[0] = comb(win[0], p0);
            win[3] = comb(win[3], p3);
            win[1] = comb(win[1], p1);
            h0 = comb(h0, ph0);
            hreg[0] = comb(hreg[0], ph1);
        }
    }
    win[2] = Sp<R>{0, 0, 0, 0};
    hreg[1] = Sp<R>{0, 0, 0, 0};
    const Sp<R> cur0 = win[0], prev0 = win[3];
    put(0, own, win[0]);
    put(0, hslot, h0);

    // epilogue buffers of this right-hand side (null pointers give empty buffers)
    constexpr bool kOut = EPI != EPI_RESID, kRes = EPI == EPI_RESID;
    constexpr bool kUpd = EPI == EPI_CGUPD || EPI == EPI_RUPD;  // r -= alpha (Op in)
    constexpr bool kXp = EPI == EPI_CGUPD || kPn;                // x += alpha p
    const __amdgpu_buffer_rsrc_t rout = make_rsrc(kOut && a.out ? a.out + rb : nullptr, nbytes);
    // x += alpha p: EPI_CGUPD unless x is null; EPI_PNORM the previous iteration's, with p_old
    const bool xupd = EPI == EPI_CGUPD ? a.x != nullptr : (kPn && use_x);
    const __amdgpu_buffer_rsrc_t rx = make_rsrc(xupd ? a.x + rb : nullptr, nbytes);
    const __amdgpu_buffer_rsrc_t rr = make_rsrc(kUpd ? a.r + rb : nullptr, nbytes);
    const __amdgpu_buffer_rsrc_t rp = make_rsrc(xupd ? (kPn ? a.pold : a.p) + rb : nullptr, nbytes);
    const __amdgpu_buffer_rsrc_t rbv = make_rsrc(kRes ? a.b + rb : nullptr, V5 * 32);
    const __amdgpu_buffer_rsrc_t rr64 = make_rsrc(kRes ? a.r64 + rb : nullptr, V5 * 32);
    const __amdgpu_buffer_rsrc_t rp64 = make_rsrc(kRes && a.p64 ? a.p64 + rb : nullptr, V5 * 32);
    const __amdgpu_buffer_rsrc_t rr32 = make_rsrc(kRes && a.r32 ? a.r32 + rb : nullptr, V5 * 16);
    const __amdgpu_buffer_rsrc_t rp32 = make_rsrc(kRes && a.p32 ? a.p32 + rb : nullptr, V5 * 16);

    // FP32 CG update: p, x, r of a slice are loaded one step before their epilogue
    // (ep[d & 1] serves step d, ep[(d + 1) & 1] receives the next slice's); EPI_PNORM's
    // x update the same way with p_old, x
    constexpr bool kPre = (kUpd && sizeof(R) == 4) || kPn;
    Sp<R> ep[2][3];
    if constexpr (kPre) {
        if constexpr (kXp) {
            ep[0][0] = bld<R>(rp, off_own(sa));
            ep[0][1] = bld<R>(rx, off_own(sa));
        }
        if constexpr (kUpd) ep[0][2] = bld<R>(rr, off_own(sa));
    }

    auto step = [&](const int kk, auto D) {
        constexpr int d = decltype(D)::value;
        constexpr int b = d & 1;
        const bool active = kk < n;
        const int s = sa + kk * dir;
        lds_barrier();  // slice s staged in lds[b]; nobody still reads lds[b ^ 1]
        // own sites of slice s + 2 dir (unless past the end or a kept wall slice)
        int oo = kBufOff;
        if (kk + 1 < n && (D5 || kk + 2 < n)) {
            const int sn = wrap(s + 2 * dir);
            if (!(whole && (sn == s_prev0 || sn == sa))) oo = off_own(sn);
        }
        const int oh = kk + 2 < n ? off_halo(s + 2 * dir) : kBufOff;
        win[(d + 2) & 3] = bld<R>(rin, oo);
        hreg[(d + 1) & 1] = bld<R>(rin, oh);
        if constexpr (kPn) {
            winp[(d + 2) & 3] = bld<R>(rpo, oo);
            hregp[(d + 1) & 1] = bld<R>(rpo, oh);
        }
        if constexpr (kPre) {
            const int en = kk + 1 < n ? off_own(s + dir) : kBufOff;
            if constexpr (kXp) {
                ep[(d + 1) & 1][0] = bld<R>(rp, en);
                ep[(d + 1) & 1][1] = bld<R>(rx, en);
            }
            if constexpr (kUpd) ep[(d + 1) & 1][2] = bld<R>(rr, en);
        }
        // psi(s + dir), loaded one step ago (at step 0: in the prologue, combined there), or
        // the kept wall slice it stands for
        {
            const int sn = wrap(s + dir);
            Sp<R> nv = win[(d + 1) & 3];
            if constexpr (kPn)
                if (kk > 0) nv = comb(nv, winp[(d + 1) & 3]);
            win[(d + 1) & 3] = pick(whole && sn == s_prev0, prev0, pick(whole && sn == sa, cur0, nv));
        }
        const Sp<R>& cur = win[d];
        const Sp<R>& nxt = win[(d + 1) & 3];
        const Sp<R>& prv = win[(d + 3) & 3];
        const Sp<R> vtp = get(b, own + W), vtm = get(b, own - W);
        const Sp<R> vxp = get(b, own + 1), vxm = get(b, own - 1);
        const Sp<R> dn = pick(dir > 0, prv, nxt), up = pick(dir > 0, nxt, prv);
        const Sp<R> o = hop_site<R, DAG, D5>(L, k, cur, vtp, vtm, vxp, vxm, dn, up, s);
        const int eo = active ? off_own(s) : kBufOff;
        if constexpr (EPI == EPI_STORE || EPI == EPI_NORM || EPI == EPI_PNORM) {
            bst<R, 2>(rout, eo, o);
            if constexpr (EPI != EPI_STORE) acc0 += active ? norm2(o) : 0.0;
            if constexpr (kPn) {
                bst<R>(rpn, eo, cur);
                const Sp<R>& pv = ep[d & 1][0];
                Sp<R> xv = ep[d & 1][1];
                xv.ar += alpha_prev * pv.ar;
                xv.ai += alpha_prev * pv.ai;
                xv.br += alpha_prev * pv.br;
                xv.bi += alpha_prev * pv.bi;
                bst<R>(rx, eo, xv);
            }
        } else if constexpr (kUpd) {
            bst<R, 2>(rout, eo, o);
            Sp<R> rv;
            if constexpr (kPre)
                rv = ep[d & 1][2];
            else
                rv = bld<R>(rr, eo);
            if constexpr (kXp) {
                Sp<R> pv, xv;
                if constexpr (kPre) {
                    pv = ep[d & 1][0];
                    xv = ep[d & 1][1];
                } else {
                    pv = bld<R>(rp, eo);
                    xv = bld<R>(rx, eo);
                }
                xv.ar += alpha * pv.ar;
                xv.ai += alpha * pv.ai;
                xv.br += alpha * pv.br;
                xv.bi += alpha * pv.bi;
                bst<R>(rx, eo, xv);
            }
            rv.ar -= alpha * o.ar;
            rv.ai -= alpha * o.ai;
            rv.br -= alpha * o.br;
            rv.bi -= alpha * o.bi;
            bst<R>(rr, eo, rv);
            acc0 += active ? norm2(rv) : 0.0;
        } else {  // EPI_RESID (FP64 only)
            if constexpr (sizeof(R) == 8) {
                const int eo32 = active ? (s * V4 + site) * 16 : kBufOff;
                const Sp<double> bv = bld<double>(rbv, eo);
                const Sp<double> rv{bv.ar - o.ar, bv.ai - o.ai, bv.br - o.br, bv.bi - o.bi};
                bst<double>(rr64, eo, rv);
                bst<double>(rp64, eo, rv);
                const Sp<float> fv{(float)rv.ar, (float)rv.ai, (float)rv.br, (float)rv.bi};
                bst<float>(rr32, eo32, fv);
                bst<float>(rp32, eo32, fv);
                acc0 += active ? norm2(rv) : 0.0;
                acc1 += active ? norm2(bv) : 0.0;
            }
        }
        // stage slice s + dir for the next step (its halo loaded one step ago; at step 0 in the
        // prologue, combined there)
        put(b ^ 1, own, nxt);
        if constexpr (kPn)
            put(b ^ 1, hslot, kk > 0 ? comb(hreg[d & 1], hregp[d & 1]) : hreg[d & 1]);
        else
            put(b ^ 1, hslot, hreg[d & 1]);
    };
    for (int k0 = 0; k0 < n; k0 += 4) {
        step(k0, IC<0>{});
        step(k0 + 1, IC<1>{});
        step(k0 + 2, IC<2>{});
        step(k0 + 3, IC<3>{});
    }

    if constexpr (EPI != EPI_STORE) {
        block_sum2_n<NT>(acc0, acc1, sh);
        if (threadIdx.x == 0) {
            double* pp = a.part + ((size_t)rhs * nblk + bid) * 2;
            pp[0] = acc0;
            pp[1] = acc1;
        }
    }
}

// ------------------------------------------------------------- CG direction update
// Reduces r.r and p^dag A p from the stencil partials, decides convergence, and sets
// p = r + beta p.  Block 0 writes the next slot's scalars; the other blocks only read
// the current slot, so nothing in one launch races.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cg_p(const long long sites, const int nblk, R* r, R* p,
                                                 const double* partA, const double* partB, CgState* st, const int slot,
                                                 const int mode, const Groups* groups, const int Ls, const int shift)
{
    const int rhs = blockIdx.z;
    CgState& S = st[rhs];
    const int f = S.flags[slot];
    const int ns = slot ^ 1;
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (f & (F_DONE | F_FROZEN)) {
        if (lead) {
            S.flags[ns] = f;
            S.rr[ns] = S.rr[slot];
        }
        return;
    }
    __shared__ double sh[2];
    const double pAp = sum_partials(partA + (size_t)rhs * nblk * 2, nblk, 0, sh);
    const bool zero = (pAp == 0.0);
    const double rr_new = zero ? 0.0 : sum_partials(partB + (size_t)rhs * nblk * 2, nblk, 0, sh);
    const double rr_old = S.rr[slot];
    const double relres = std::sqrt(rr_new) / std::sqrt(S.bnorm2);
    const bool conv = relres <= S.tol;
    const bool freeze = (mode == MODE_MIXED) && (conv || rr_new <= S.delta * S.delta * S.rr_ref);
    if (lead) {
        int nf = f;
        if (zero) {
            S.status = 4;
            nf |= F_DONE;
            S.rr[ns] = rr_old;
        } else {
            S.rr[ns] = rr_new;
            S.relres = relres;
            S.iters += 1;
            if (mode == MODE_CG64 && conv) nf |= F_DONE;
            if (freeze) nf |= F_FROZEN;
        }
        S.flags[ns] = nf;
    }
    if (zero || (mode == MODE_CG64 && conv)) return;
    const R beta = (R)(rr_new / rr_old);
    long long V5 = sites, rb = (long long)rhs * sites * 4;
    if (groups) {  // several lattices: this right-hand side's extent and offset
        const GroupPos q = group_for_rhs(*groups, rhs);
        const long long v5 = (long long)Ls * q.Nt * q.Nx;
        V5 = v5 >> shift;
        rb = ((q.site0 + (long long)(rhs - q.rhs0) * v5) >> shift) * 4;
    }
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < V5; i += (long long)gridDim.x * kBlock) {
        const Sp<R> rv = ldsp(r + rb, i);
        Sp<R> pv = ldsp(p + rb, i);
        pv.ar = rv.ar + beta * pv.ar;
        pv.ai = rv.ai + beta * pv.ai;
        pv.br = rv.br + beta * pv.br;
        pv.bi = rv.bi + beta * pv.bi;
        stsp(p + rb, i, pv);
    }
}

// ------------------------------------------------------------------ scalar kernels
// One 64-thread block per right-hand side.
__global__ void k_cg_start(const int nblk, const double* partR, CgState* st, const double tol,
                           const double delta)
{
    const int rhs = blockIdx.x;
    const double* p = partR + (size_t)rhs * nblk * 2;
    double rr = 0.0, bb = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 64) {
        rr += p[2 * i];
        bb += p[2 * i + 1];
    }
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    if (threadIdx.x == 0) {
        CgState S{};
        S.bnorm2 = bb;
        S.tol = tol;
        S.delta = delta;
        S.rr[0] = S.rr[1] = rr;
        S.rr_ref = rr;
        int flags = 0;
        const double bnorm = std::sqrt(bb);
        if (bnorm == 0.0) {
            flags = F_DONE;
            S.relres = 0.0;
            S.true_relres = 0.0;
        } else {
            S.relres = std::sqrt(rr) / bnorm;
            S.true_relres = S.relres;
            if (S.relres <= tol) flags = F_DONE;
        }
        S.flags[0] = S.flags[1] = flags;
        st[rhs] = S;
    }
}

// The bookkeeping of k_cg_p without its vector pass, for the pipelined even-odd CG (whose
// next operator launch forms p = r + beta p itself): convergence, freezing and the states
// of slot ^ 1, beta for the next direction, and the direction buffer flip.  One 64-thread
// block per right-hand side.  The operator launch of the next iteration can do the same
// itself (k_eo_schur's fold: the same arithmetic, in every block).
__global__ void k_cg_flags(const int nblk, const double* partA, const double* partB, CgState* st, const int slot,
                           const int mode)
{
    const int rhs = blockIdx.x;
    CgState& S = st[rhs];
    const int f = S.flags[slot];
    const int ns = slot ^ 1;
    if (f & (F_DONE | F_FROZEN)) {
        if (threadIdx.x == 0) {
            S.flags[ns] = f;
            S.rr[ns] = S.rr[slot];
            S.pbuf[ns] = S.pbuf[slot];
        }
        return;
    }
    const double* pa = partA + (size_t)rhs * nblk * 2;
    const double* pb = partB + (size_t)rhs * nblk * 2;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 64) {
        a += pa[2 * i];
        b += pb[2 * i];
    }
    const double pAp = wave_sum(a);
    const double rr_new = wave_sum(b);
    if (threadIdx.x != 0) return;
    const double rr_old = S.rr[slot];
    int nf = f & ~F_FRESH;
    if (pAp == 0.0) {
        S.status = 4;
        nf |= F_DONE;
        S.rr[ns] = rr_old;
        S.pbuf[ns] = S.pbuf[slot];
    } else {
        const double relres = std::sqrt(rr_new) / std::sqrt(S.bnorm2);
        const bool conv = relres <= S.tol;
        S.rr[ns] = rr_new;
        S.relres = relres;
        S.iters += 1;
        S.beta = rr_new / rr_old;
        S.pbuf[ns] = S.pbuf[slot] ^ 1;  // this iteration's direction (written to the other buffer) is current
        if (mode == MODE_CG64 && conv) nf |= F_DONE;
        if (mode == MODE_MIXED && (conv || rr_new <= S.delta * S.delta * S.rr_ref)) nf |= F_FROZEN;
    }
    S.flags[ns] = nf;
}

// Start of a pipelined even-odd CG: every right-hand side's first direction is its residual.
__global__ void k_set_fresh(CgState* st, const int nrhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrhs) {
        st[i].flags[0] |= F_FRESH;
        st[i].beta = 0.0;
        st[i].pbuf[0] = st[i].pbuf[1] = 0;
    }
}

// After a reliable update: FP64 true residual -> certify, or restart the FP32 cycle.
__global__ void k_ru_commit(const int nblk, const double* partR, CgState* st, const int slot)
{
    const int rhs = blockIdx.x;
    CgState& S = st[rhs];
    const int f = S.flags[slot];
    if (f & F_HOLD) {  // its FP32 cycle continues; this update skipped it
        if (threadIdx.x == 0) S.flags[slot] = f & ~F_HOLD;
        return;
    }
    if (f & F_DONE) return;
    const double* p = partR + (size_t)rhs * nblk * 2;
    double rr = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 64) rr += p[2 * i];
    rr = wave_sum(rr);
    if (threadIdx.x == 0) {
        const double tr = std::sqrt(rr) / std::sqrt(S.bnorm2);
        S.true_relres = tr;
        S.reliable += 1;
        if (tr <= S.tol) {
            S.flags[slot] |= F_DONE;
        } else {
            S.rr[slot] = rr;
            S.rr_ref = rr;
            S.flags[slot] &= ~F_FROZEN;
        }
    }
}

__global__ void k_true_resid(const int nblk, const double* partR, CgState* st)
{
    const int rhs = blockIdx.x;
    const double* p = partR + (size_t)rhs * nblk * 2;
    double rr = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 64) rr += p[2 * i];
    rr = wave_sum(rr);
    if (threadIdx.x == 0) {
        const double bn = std::sqrt(st[rhs].bnorm2);
        st[rhs].true_relres = bn == 0.0 ? 0.0 : std::sqrt(rr) / bn;
    }
}

// F_HOLD on the right-hand sides the host marked (hold[i] != 0) that are not done.
__global__ void k_hold(CgState* st, const int* hold, const int slot, const int nrhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrhs && hold[i] && !(st[i].flags[slot] & F_DONE)) st[i].flags[slot] |= F_HOLD;
}

__global__ void k_set_frozen_all(CgState* st, const int slot, const int nrhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrhs && !(st[i].flags[slot] & F_DONE)) st[i].flags[slot] |= F_FROZEN;
}

// x64 += x32; x32 = 0 for the right-hand sides still running.
__global__ __launch_bounds__(kBlock) void k_accum_x(const long long sites, double* x64, float* x32,
                                                    const CgState* st, const int slot, const int skip,
                                                    const Groups* groups, const int Ls, const int shift)
{
    const int rhs = blockIdx.z;
    if (st[rhs].flags[slot] & skip) return;
    long long n = sites, rb = (long long)rhs * sites * 4;
    if (groups) {
        const GroupPos q = group_for_rhs(*groups, rhs);
        const long long v5 = (long long)Ls * q.Nt * q.Nx;
        n = v5 >> shift;
        rb = ((q.site0 + (long long)(rhs - q.rhs0) * v5) >> shift) * 4;
    }
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;  // site
    if (i >= n) return;
    const Sp<float> d = ldsp(x32 + rb, i);
    Sp<double> v = ldsp(x64 + rb, i);
    v.ar += d.ar;
    v.ai += d.ai;
    v.br += d.br;
    v.bi += d.bi;
    stsp(x64 + rb, i, v);
    stsp(x32 + rb, i, Sp<float>{0.f, 0.f, 0.f, 0.f});
}

template <typename R> __global__ void k_zero(R* p, const long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (R)0;
}

__global__ void k_pack_links(const int n, const double* theta, double2* l64, float2* l32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s, c;
    sincos(theta[i], &s, &c);
    l64[i] = make_double2(c, s);
    l32[i] = make_float2((float)c, (float)s);
}

template <typename R, bool DAG, bool D5, bool PF, bool MG = false>
void dispatch_epi_pf(int epi, const HopArgs<R>& a, dim3 grid, hipStream_t s)
{
    switch (epi) {
        case EPI_STORE: k_hop<R, DAG, D5, EPI_STORE, PF, MG><<<grid, kBlock, 0, s>>>(a); break;
        case EPI_NORM: k_hop<R, DAG, D5, EPI_NORM, PF, MG><<<grid, kBlock, 0, s>>>(a); break;
        case EPI_CGUPD: k_hop<R, DAG, D5, EPI_CGUPD, PF, MG><<<grid, kBlock, 0, s>>>(a); break;
        default:
            if constexpr (sizeof(R) == 8) k_hop<R, DAG, D5, EPI_RESID, PF, MG><<<grid, kBlock, 0, s>>>(a);
            break;
    }
}

template <typename R, bool DAG, bool D5, int NT>
void dispatch_lds_nt(int epi, const HopArgs<R>& a, dim3 grid, hipStream_t s)
{
    switch (epi) {
        case EPI_STORE: k_hop_lds<R, DAG, D5, EPI_STORE, NT><<<grid, NT, 0, s>>>(a); break;
        case EPI_NORM: k_hop_lds<R, DAG, D5, EPI_NORM, NT><<<grid, NT, 0, s>>>(a); break;
        case EPI_CGUPD: k_hop_lds<R, DAG, D5, EPI_CGUPD, NT><<<grid, NT, 0, s>>>(a); break;
        case EPI_PNORM:  // the learned solver's D apply (FP32, domain wall)
            if constexpr (sizeof(R) == 4 && !DAG && D5) k_hop_lds<R, DAG, D5, EPI_PNORM, NT><<<grid, NT, 0, s>>>(a);
            else throw DwfError(DWF_ERR_ARG, "EPI_PNORM: FP32 domain-wall D only");
            break;
        case EPI_RUPD:  // the learned solver's D^dag apply (FP32, domain wall)
            if constexpr (sizeof(R) == 4 && DAG && D5) k_hop_lds<R, DAG, D5, EPI_RUPD, NT><<<grid, NT, 0, s>>>(a);
            else throw DwfError(DWF_ERR_ARG, "EPI_RUPD: FP32 domain-wall D^dag only");
            break;
        default:
            if constexpr (sizeof(R) == 8) k_hop_lds<R, DAG, D5, EPI_RESID, NT><<<grid, NT, 0, s>>>(a);
            break;
    }
}

template <typename R, bool DAG, bool D5>
void dispatch_lds(int epi, const HopArgs<R>& a, dim3 grid, hipStream_t s)
{
    switch (a.cfg.lds) {
        case 2: dispatch_lds_nt<R, DAG, D5, 512>(epi, a, grid, s); break;
        case 3: dispatch_lds_nt<R, DAG, D5, 64>(epi, a, grid, s); break;
        case 4: dispatch_lds_nt<R, DAG, D5, 128>(epi, a, grid, s); break;
        case 5: dispatch_lds_nt<R, DAG, D5, 384>(epi, a, grid, s); break;
        case 6: dispatch_lds_nt<R, DAG, D5, 768>(epi, a, grid, s); break;
        default: dispatch_lds_nt<R, DAG, D5, 256>(epi, a, grid, s); break;
    }
}

template <typename R, bool DAG, bool D5>
void dispatch_epi(int epi, const HopArgs<R>& a, dim3 grid, hipStream_t s)
{
    if ((epi == EPI_PNORM || epi == EPI_RUPD) && !a.cfg.lds)
        throw DwfError(DWF_ERR_ARG, "EPI_PNORM / EPI_RUPD need the LDS-staged stencil");
    if (epi == EPI_CGUPD && (!a.x || a.aout) && !a.cfg.lds)
        throw DwfError(DWF_ERR_ARG, "EPI_CGUPD without x or with aout needs the LDS-staged stencil");
    if (a.groups)
        dispatch_epi_pf<R, DAG, D5, false, true>(epi, a, grid, s);  // lattice groups: linear blocks, no prefetch
    else if (a.cfg.lds)
        dispatch_lds<R, DAG, D5>(epi, a, grid, s);
    else if (a.cfg.prefetch)
        dispatch_epi_pf<R, DAG, D5, true>(epi, a, grid, s);
    else
        dispatch_epi_pf<R, DAG, D5, false>(epi, a, grid, s);
}

}  // namespace

Groups single_group(const Lat& L, int nrhs, int rhs_per_cfg)
{
    Groups G{};
    G.n = 1;
    G.rhs0[0] = 0;
    G.rhs0[1] = nrhs;
    G.Nt[0] = L.Nt;
    G.Nx[0] = L.Nx;
    G.rpc[0] = rhs_per_cfg > 0 ? rhs_per_cfg : 1 << 30;
    return G;
}

int hop_pstride(const Groups& G, const HopCfg& cfg)
{
    if (G.n <= 1) return cfg.nblk();
    int m = 0;
    for (int g = 0; g < G.n; ++g) m = std::max(m, (G.Nt[g] * G.Nx[g] + kBlock - 1) / kBlock * cfg.chunks);
    return m;
}

HopCfg hop_cfg(const Lat& L, int nrhs, int chunks_override, int remap, int prefetch, int tile_x, int nt_store,
              int lds)
{
    HopCfg c{};
    c.nt_store = nt_store;
    c.remap = remap;
    c.prefetch = prefetch;
    c.lds = 0;
    c.tx = 0;
    c.tiles_x = 0;
    c.bx = (L.V4 + kBlock - 1) / kBlock;
    // (the LDS kernel addresses one right-hand side's spinors with 32-bit byte offsets)
    if (lds >= 1 && lds <= 6 && L.V5() * 32 < kBufOff) {
        const int nt = lds_threads(lds);
        if (lds_tile_ok(nt, tile_x) && L.Nx % tile_x == 0 && L.Nt % (nt / tile_x) == 0) {
            c.lds = lds;
            c.tx = tile_x;
            c.tiles_x = L.Nx / tile_x;
            c.bx = c.tiles_x * (L.Nt / (nt / tile_x));
            c.nt_store = 1;  // the LDS kernel always streams its output
        }
    }
    if (!c.lds && tile_x > 0 && tile_x <= kBlock && kBlock % tile_x == 0 && L.Nx % tile_x == 0) {
        c.tx = tile_x;
        c.tiles_x = L.Nx / tile_x;
        const int rows = kBlock / tile_x;
        c.bx = c.tiles_x * ((L.Nt + rows - 1) / rows);
    }
    int chunks = 1;
    if (chunks_override > 0) {
        chunks = chunks_override;
    } else {
        // Enough blocks to cover 256 CUs several times over; fewer chunks keep more
        // slices per thread (more link reuse) once the batch alone fills the chip.
        while (chunks < L.Ls && (long long)c.bx * chunks * nrhs < 2048) chunks *= 2;
    }
    if (chunks > L.Ls) chunks = L.Ls;
    if (chunks < 1) chunks = 1;
    c.spt = (L.Ls + chunks - 1) / chunks;
    c.chunks = (L.Ls + c.spt - 1) / c.spt;
    return c;
}

template <typename R>
void launch_hop(bool dagger, int epi, const HopArgs<R>& a0, int nrhs, hipStream_t s, const Groups* hg)
{
    HopArgs<R> a = a0;
    long long nb = 0;
    if (!a.groups) {
        // one lattice: the shape's own tiles (direct or LDS-staged)
        nb = (long long)a.cfg.nblk() * nrhs;
    } else {
        // several lattices: linear blocks of kBlock consecutive 4D sites per group
        if (!hg || a.cfg.lds || a.cfg.tx || a.cfg_map || hg->n < 1 || hg->n > kMaxGroups || hg->rhs0[hg->n] != nrhs)
            throw DwfError(DWF_ERR_ARG, "stencil launch: bad lattice-group layout");
        for (int g = 0; g < hg->n; ++g)
            nb += (long long)(hg->rhs0[g + 1] - hg->rhs0[g]) * ((hg->Nt[g] * hg->Nx[g] + kBlock - 1) / kBlock) *
                  a.cfg.chunks;
        a.pstride = hop_pstride(*hg, a.cfg);
    }
    if (nb <= 0 || nb > 0x7fffffffLL) throw DwfError(DWF_ERR_ARG, "stencil launch: grid size out of range");
    const dim3 grid((unsigned)nb);
    if (a.L.d5) {
        if (dagger)
            dispatch_epi<R, true, true>(epi, a, grid, s);
        else
            dispatch_epi<R, false, true>(epi, a, grid, s);
    } else {
        if (dagger)
            dispatch_epi<R, true, false>(epi, a, grid, s);
        else
            dispatch_epi<R, false, false>(epi, a, grid, s);
    }
}
template void launch_hop<float>(bool, int, const HopArgs<float>&, int, hipStream_t, const Groups*);
template void launch_hop<double>(bool, int, const HopArgs<double>&, int, hipStream_t, const Groups*);

namespace {
long long max_sites(const VecLayout& lay)
{
    long long m = 0;
    for (int g = 0; g < lay.gr.n; ++g) m = std::max(m, lay.sites(g));
    return m;
}
}  // namespace

template <typename R>
void launch_cg_p(const VecLayout& lay, int nblk, R* r, R* p, const double* partA, const double* partB, CgState* st,
                 int slot, int mode, int nrhs, hipStream_t s)
{
    // blocks of 4 sites per thread over the largest vector (grid-stride loops cover the rest)
    const long long per = (max_sites(lay) + 4LL * kBlock - 1) / (4LL * kBlock);
    const dim3 grid((unsigned)per, 1, nrhs);
    const bool one = lay.gr.n == 1;
    if (!one && !lay.dev) throw DwfError(DWF_ERR_ARG, "CG update: lattice groups without their device table");
    k_cg_p<R><<<grid, kBlock, 0, s>>>(one ? lay.sites(0) : 0, nblk, r, p, partA, partB, st, slot, mode,
                                      one ? nullptr : lay.dev, lay.Ls, lay.shift);
}
template void launch_cg_p<float>(const VecLayout&, int, float*, float*, const double*, const double*, CgState*, int,
                                 int, int, hipStream_t);
template void launch_cg_p<double>(const VecLayout&, int, double*, double*, const double*, const double*, CgState*,
                                  int, int, int, hipStream_t);

void launch_cg_start(int nblk, const double* partR, CgState* st, double tol, double delta, int nrhs, hipStream_t s)
{
    k_cg_start<<<nrhs, 64, 0, s>>>(nblk, partR, st, tol, delta);
}

void launch_cg_flags(int nblk, const double* partA, const double* partB, CgState* st, int slot, int mode, int nrhs,
                     hipStream_t s)
{
    k_cg_flags<<<nrhs, 64, 0, s>>>(nblk, partA, partB, st, slot, mode);
}

void launch_set_fresh(CgState* st, int nrhs, hipStream_t s)
{
    k_set_fresh<<<(nrhs + 63) / 64, 64, 0, s>>>(st, nrhs);
}

void launch_ru_commit(int nblk, const double* partR, CgState* st, int slot, int nrhs, hipStream_t s)
{
    k_ru_commit<<<nrhs, 64, 0, s>>>(nblk, partR, st, slot);
}

void launch_true_resid(int nblk, const double* partR, CgState* st, int nrhs, hipStream_t s)
{
    k_true_resid<<<nrhs, 64, 0, s>>>(nblk, partR, st);
}

void launch_set_frozen_all(CgState* st, int slot, int nrhs, hipStream_t s)
{
    k_set_frozen_all<<<(nrhs + 63) / 64, 64, 0, s>>>(st, slot, nrhs);
}

void launch_accum_x(const VecLayout& lay, double* x64, float* x32, const CgState* st, int slot, int skip, int nrhs,
                    hipStream_t s)
{
    const dim3 grid((unsigned)((max_sites(lay) + kBlock - 1) / kBlock), 1, nrhs);
    const bool one = lay.gr.n == 1;
    if (!one && !lay.dev) throw DwfError(DWF_ERR_ARG, "x update: lattice groups without their device table");
    k_accum_x<<<grid, kBlock, 0, s>>>(one ? lay.sites(0) : 0, x64, x32, st, slot, skip, one ? nullptr : lay.dev,
                                      lay.Ls, lay.shift);
}

void launch_hold(CgState* st, const int* hold, int slot, int nrhs, hipStream_t s)
{
    k_hold<<<(nrhs + 63) / 64, 64, 0, s>>>(st, hold, slot, nrhs);
}

template <typename R> void launch_zero(R* p, long long n, hipStream_t s)
{
    if (n <= 0) return;
    k_zero<R><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(p, n);
}
template void launch_zero<float>(float*, long long, hipStream_t);
template void launch_zero<double>(double*, long long, hipStream_t);

void launch_pack_links(long long nlinks, const double* theta, double2* l64, float2* l32, hipStream_t s)
{
    const int n = (int)nlinks;
    k_pack_links<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, theta, l64, l32);
}

}  // namespace dwfk
