// dwf_eo.hip -- even-odd (4D checkerboard) kernels of the Schur-complement DWF solver.
//
// The Shamir operator splits as D = A - 1/2 H, with A = (3-M5) + D5 local in 4D (it
// couples only the s-slices of one 4D site) and H the 4D hopping term, which connects
// the two 4D parities (dwf_operator.hpp:15-92).  With Nt, Nx even:
//
//   D x = b  <=>  Dhat x_o = b_o + 1/2 H_oe A^-1 b_e,   x_e = A^-1 (b_e + 1/2 H_eo x_o),
//   Dhat = A - 1/4 H_oe A^-1 H_eo,   Dhat^dag = A^dag - 1/4 H^dag_oe A^-dag H^dag_eo,
//
// where H^dag is the hopping term of D^dag (dwf_operator.hpp:94-171).
//
// A^-1 has a closed form per spin: spin 0 of A is d + S_dn (lower bidiagonal with the
// wall term +mf in the corner), spin 1 is d + S_up (upper), with d = 3-M5; A^dag swaps
// them.  Each is solved by one O(Ls) recurrence plus the wall correction
//   lower: beta_0 = 0, beta_s = (v_s + beta_{s-1})/d, y_0 = (v_0 - mf beta_{L-1})/(d + mf d^{1-L}),
//          y_s = d^-s y_0 + beta_s;
//   upper: the mirror image from s = L-1 downwards,
// so a thread that owns one 4D site keeps the whole s-column in registers (Ls is a
// template parameter).
//
// Half-lattice layout: parity p = (t+x)&1 holds V4/2 sites, row-major (t, xh) with
// x = 2 xh + ((t+p)&1); a vector is [rhs][s][t][xh] spinors.  Links stay lexicographic.

#include "dwf_context.hpp"
#include "dwf_internal.hpp"
#include "dwf_eo.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace dwfk {
namespace {

template <typename R> struct V4T;
template <> struct V4T<float> { using T = float4; };
template <> struct V4T<double> { using T = double4; };
template <typename R> struct C2T;
template <> struct C2T<float> { using T = float2; };
template <> struct C2T<double> { using T = double2; };

template <typename R> struct Sp {
    R ar, ai, br, bi;
};

template <typename R> __device__ __forceinline__ Sp<R> ld(const R* base, long long i)
{
    const typename V4T<R>::T v = reinterpret_cast<const typename V4T<R>::T*>(base)[i];
    return Sp<R>{v.x, v.y, v.z, v.w};
}
template <typename R> __device__ __forceinline__ void st(R* base, long long i, const Sp<R>& s)
{
    typename V4T<R>::T v;
    v.x = s.ar;
    v.y = s.ai;
    v.z = s.br;
    v.w = s.bi;
    reinterpret_cast<typename V4T<R>::T*>(base)[i] = v;
}
template <typename R> __device__ __forceinline__ double n2(const Sp<R>& s)
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
__device__ __forceinline__ double sum_partials(const double* p, int n, double* sh)
{
    if (threadIdx.x < 64) {
        double v = 0.0;
        for (int i = threadIdx.x; i < n; i += 64) v += p[2 * i];
        v = wave_sum(v);
        if (threadIdx.x == 0) sh[0] = v;
    }
    __syncthreads();
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Geometry of one target-parity site: its lexicographic (t, x), the half-lattice
// indices of its four neighbours (other parity), links and boundary signs.
template <typename R> struct EoSite {
    int t, x;
    int ntp, ntm, nxp, nxm;
    typename C2T<R>::T ut, ux, utm, uxm;
    R stp, stm, sxp, sxm;
};

template <typename R>
__device__ __forceinline__ EoSite<R> eo_site(const Lat& L, int Nxh, int tpar, int hs, const void* links)
{
    EoSite<R> g;
    const int Nt = L.Nt, Nx = L.Nx, V4 = L.V4;
    const int t = hs / Nxh, xh = hs - t * Nxh;
    const int x = 2 * xh + ((t + tpar) & 1);
    const int tp = (t + 1 == Nt) ? 0 : t + 1, tm = (t == 0) ? Nt - 1 : t - 1;
    const int xp = (x + 1 == Nx) ? 0 : x + 1, xm = (x == 0) ? Nx - 1 : x - 1;
    g.t = t;
    g.x = x;
    g.ntp = tp * Nxh + (x >> 1);
    g.ntm = tm * Nxh + (x >> 1);
    g.nxp = t * Nxh + (xp >> 1);
    g.nxm = t * Nxh + (xm >> 1);
    using C = typename C2T<R>::T;
    const C* lk = static_cast<const C*>(links);
    g.ut = lk[t * Nx + x];
    g.ux = lk[V4 + t * Nx + x];
    g.utm = lk[tm * Nx + x];
    g.uxm = lk[V4 + t * Nx + xm];
    g.stp = (t == Nt - 1) ? (R)L.bct : (R)1;
    g.stm = (t == 0) ? (R)L.bct : (R)1;
    g.sxp = (x == Nx - 1) ? (R)L.bcx : (R)1;
    g.sxm = (x == 0) ? (R)L.bcx : (R)1;
    return g;
}

// Where one block of an even-odd launch works: its right-hand side, block within it, the
// lattice, half-lattice sizes, spinor offset (reals) and links.
template <typename R> struct EoPos {
    int rhs, bid;
    Lat L;
    int Nxh, V4h;
    long long rb;
    const void* links;
};

// per: target sites per block; chunks: s-chunks (1 for k_eo_a); LS: the s-extent.
template <typename R>
__device__ __forceinline__ EoPos<R> eo_pos(const EoArgs<R>& a, int lb, int per, int chunks, int nblk1)
{
    using C = typename C2T<R>::T;
    EoPos<R> q;
    q.L = a.L;
    if (!a.groups) {
        q.rhs = lb / nblk1;
        q.bid = lb - q.rhs * nblk1;
        q.Nxh = a.Nxh;
        q.V4h = a.V4h;
        q.rb = (long long)q.rhs * a.V4h * a.L.Ls * 4;
        q.links = static_cast<const C*>(a.links) + (long long)(q.rhs / a.rhs_per_cfg) * a.link_stride;
    } else {
        int blk0, nblk;
        const GroupPos p = block_group(a.groups, lb, [per, chunks](int nt, int nx) { return (nt * nx / 2 + per - 1) / per * chunks; },
                                       blk0, nblk);
        const int local = lb - blk0, rl = local / nblk;
        q.rhs = p.rhs0 + rl;
        q.bid = local - rl * nblk;
        q.L.Nt = p.Nt;
        q.L.Nx = p.Nx;
        q.L.V4 = p.Nt * p.Nx;
        q.Nxh = p.Nx / 2;
        q.V4h = q.L.V4 / 2;
        q.rb = (p.site0 / 2 + (long long)rl * q.V4h * a.L.Ls) * 4;
        q.links = static_cast<const C*>(a.links) + p.link0 + (long long)(rl / p.rpc) * 2 * q.L.V4;
    }
    return q;
}

// o += c * H psi at one slice (H: the bracket of D's hop, or D^dag's when DAG), from the
// values of psi at the site's +t, -t, +x, -x neighbours.
// SIGNS false: the boundary signs are already folded into the links (fold_signs).
template <typename R, bool DAG, bool SIGNS = true>
__device__ __forceinline__ void hop_sum(Sp<R>& o, R c, const EoSite<R>& g, const Sp<R>& vtp, const Sp<R>& vtm,
                                        const Sp<R>& vxp, const Sp<R>& vxm)
{
    constexpr R sg = DAG ? (R)-1 : (R)1;
    auto sgn = [](R s, R v) { return SIGNS ? s * v : v; };
    {
        const R ar = sgn(g.stp, vtp.ar - sg * vtp.br), ai = sgn(g.stp, vtp.ai - sg * vtp.bi);
        const R hr = g.ut.x * ar - g.ut.y * ai, hi = g.ut.x * ai + g.ut.y * ar;
        o.ar += c * hr;
        o.ai += c * hi;
        o.br -= c * sg * hr;
        o.bi -= c * sg * hi;
    }
    {
        const R br = sgn(g.stm, vtm.ar + sg * vtm.br), bi = sgn(g.stm, vtm.ai + sg * vtm.bi);
        const R hr = g.utm.x * br + g.utm.y * bi, hi = g.utm.x * bi - g.utm.y * br;
        o.ar += c * hr;
        o.ai += c * hi;
        o.br += c * sg * hr;
        o.bi += c * sg * hi;
    }
    {
        const R cr = sgn(g.sxp, vxp.ar - sg * vxp.bi), ci = sgn(g.sxp, vxp.ai + sg * vxp.br);
        const R hr = g.ux.x * cr - g.ux.y * ci, hi = g.ux.x * ci + g.ux.y * cr;
        o.ar += c * hr;
        o.ai += c * hi;
        o.br += c * sg * hi;
        o.bi -= c * sg * hr;
    }
    {
        const R dr = sgn(g.sxm, vxm.ar + sg * vxm.bi), di = sgn(g.sxm, vxm.ai - sg * vxm.br);
        const R hr = g.uxm.x * dr + g.uxm.y * di, hi = g.uxm.x * di - g.uxm.y * dr;
        o.ar += c * hr;
        o.ai += c * hi;
        o.br -= c * sg * hi;
        o.bi += c * sg * hr;
    }
}

// Multiplies the links by the boundary signs (+-1, so exactly), for hop_sum<..., false>.
template <typename R> __device__ __forceinline__ void fold_signs(EoSite<R>& g)
{
    g.ut.x *= g.stp;
    g.ut.y *= g.stp;
    g.utm.x *= g.stm;
    g.utm.y *= g.stm;
    g.ux.x *= g.sxp;
    g.ux.y *= g.sxp;
    g.uxm.x *= g.sxm;
    g.uxm.y *= g.sxm;
}

template <typename R, bool DAG>
__device__ __forceinline__ void add_hop(Sp<R>& o, R c, const EoSite<R>& g, const R* in, long long sl)
{
    hop_sum<R, DAG>(o, c, g, ld(in, sl + g.ntp), ld(in, sl + g.ntm), ld(in, sl + g.nxp), ld(in, sl + g.nxm));
}

// In-register solve along s.  LOWER: d y_s - y_{s-1} = v_s (s>0), d y_0 + mf y_{L-1} = v_0.
// UPPER: d y_s - y_{s+1} = v_s (s<L-1), d y_{L-1} + mf y_0 = v_{L-1}.  re/im of one spin
// component are passed as two arrays.
template <typename R, int LS, bool UPPER>
__device__ __forceinline__ void wall_solve(R (&vr)[LS], R (&vi)[LS], R d, R mf)
{
    const R inv = (R)1 / d;
    R dpow = (R)1;  // d^{-(L-1)}
#pragma unroll
    for (int k = 1; k < LS; ++k) dpow *= inv;
    const R den = (R)1 / (d + mf * dpow);
    if constexpr (!UPPER) {
        R br = 0, bi = 0;
#pragma unroll
        for (int s = 1; s < LS; ++s) {
            br = (vr[s] + br) * inv;
            bi = (vi[s] + bi) * inv;
            vr[s] = br;
            vi[s] = bi;
        }
        const R y0r = (vr[0] - mf * br) * den, y0i = (vi[0] - mf * bi) * den;
        vr[0] = y0r;
        vi[0] = y0i;
        R a = (R)1;
#pragma unroll
        for (int s = 1; s < LS; ++s) {
            a *= inv;
            vr[s] += a * y0r;
            vi[s] += a * y0i;
        }
    } else {
        R br = 0, bi = 0;
#pragma unroll
        for (int s = LS - 2; s >= 0; --s) {
            br = (vr[s] + br) * inv;
            bi = (vi[s] + bi) * inv;
            vr[s] = br;
            vi[s] = bi;
        }
        const R ylr = (vr[LS - 1] - mf * br) * den, yli = (vi[LS - 1] - mf * bi) * den;
        vr[LS - 1] = ylr;
        vi[LS - 1] = yli;
        R a = (R)1;
#pragma unroll
        for (int s = LS - 2; s >= 0; --s) {
            a *= inv;
            vr[s] += a * ylr;
            vi[s] += a * yli;
        }
    }
}

// wall_solve on one spin component kept in LDS, (re, im) of slice s at col[s * cs],
// col[s * cs + 1]: the same operations in the same order, streamed through LDS so only
// the recurrence's carries live in registers.
template <typename R, int LS, bool UPPER>
__device__ __forceinline__ void wall_solve_col(R* col, int cs, R d, R mf)
{
    const R inv = (R)1 / d;
    R dpow = (R)1;  // d^{-(L-1)}
#pragma unroll
    for (int k = 1; k < LS; ++k) dpow *= inv;
    const R den = (R)1 / (d + mf * dpow);
    // lower: slices 1 .. L-1 upwards, wall slice 0; upper: slices L-2 .. 0 downwards, wall L-1
    auto at = [&](int k) { return UPPER ? LS - 1 - k : k; };
    R br = 0, bi = 0;
#pragma unroll 4
    for (int k = 1; k < LS; ++k) {
        R* c = col + at(k) * cs;
        br = (c[0] + br) * inv;
        bi = (c[1] + bi) * inv;
        c[0] = br;
        c[1] = bi;
    }
    R* cw = col + at(0) * cs;
    const R ywr = (cw[0] - mf * br) * den, ywi = (cw[1] - mf * bi) * den;
    cw[0] = ywr;
    cw[1] = ywi;
    R a = (R)1;
#pragma unroll 4
    for (int k = 1; k < LS; ++k) {
        a *= inv;
        R* c = col + at(k) * cs;
        c[0] += a * ywr;
        c[1] += a * ywi;
    }
}

// out = [A^-1 | A^-dag | 1] ( cb b + ch H psi_other ) on the target parity, whole
// s-column per thread.  b and the hop are optional (cb == 0 / ch == 0 skip them).
template <typename R, bool DAG, int LS, bool AINV>
__global__ __launch_bounds__(kBlockA) void k_eo_a(const EoArgs<R> a)
{
    const int total = gridDim.x;
    const int lb = a.remap ? xcd_logical_block(blockIdx.x, total) : (int)blockIdx.x;
    const EoPos<R> q = eo_pos(a, lb, kBlockA, 1, a.bxa);
    const int rhs = q.rhs;
    const int hs = q.bid * kBlockA + threadIdx.x;
    if (a.st != nullptr && (a.st[rhs].flags[a.slot] & a.skip)) return;
    const Lat& L = q.L;
    const int V4h = q.V4h;
    if (hs >= V4h) return;
    const long long rb = q.rb;
    const EoSite<R> g = eo_site<R>(L, q.Nxh, a.tpar, hs, q.links);
    R v0r[LS], v0i[LS], v1r[LS], v1i[LS];
#pragma unroll
    for (int s = 0; s < LS; ++s) {
        const long long sl = (long long)s * V4h;
        Sp<R> o{0, 0, 0, 0};
        if (a.cb != (R)0) {
            const Sp<R> bv = ld(a.b + rb, sl + hs);
            o = Sp<R>{a.cb * bv.ar, a.cb * bv.ai, a.cb * bv.br, a.cb * bv.bi};
        }
        if (a.ch != (R)0) add_hop<R, DAG>(o, a.ch, g, a.hin + rb, sl);
        v0r[s] = o.ar;
        v0i[s] = o.ai;
        v1r[s] = o.br;
        v1i[s] = o.bi;
    }
    if constexpr (AINV) {
        const R d = (R)a.L.diag, mf = (R)a.L.mf;
        // A: spin 0 lower, spin 1 upper; A^dag: spin 0 upper, spin 1 lower
        wall_solve<R, LS, DAG>(v0r, v0i, d, mf);
        wall_solve<R, LS, !DAG>(v1r, v1i, d, mf);
    }
#pragma unroll
    for (int s = 0; s < LS; ++s) st(a.out + rb, (long long)s * V4h + hs, Sp<R>{v0r[s], v0i[s], v1r[s], v1i[s]});
}

// The per-site epilogue of the even-odd operator kernels: o is the operator's value at
// index idx (within right-hand side offset rb) of the target parity.
template <typename R, int EPI>
__device__ __forceinline__ void eo_epi(const EoArgs<R>& a, const R* pdir, long long rb, long long idx, const Sp<R>& o,
                                       R alpha, double& acc0, double& acc1)
{
    if constexpr (EPI == EPI_STORE) {
        st(a.out + rb, idx, o);
    } else if constexpr (EPI == EPI_NORM) {
        st(a.out + rb, idx, o);
        acc0 += n2(o);
    } else if constexpr (EPI == EPI_CGUPD) {
        const Sp<R> pv = ld(pdir + rb, idx);
        Sp<R> xv = ld(a.x + rb, idx);
        xv.ar += alpha * pv.ar;
        xv.ai += alpha * pv.ai;
        xv.br += alpha * pv.br;
        xv.bi += alpha * pv.bi;
        st(a.x + rb, idx, xv);
        Sp<R> rv = ld(a.r + rb, idx);
        rv.ar -= alpha * o.ar;
        rv.ai -= alpha * o.ai;
        rv.br -= alpha * o.br;
        rv.bi -= alpha * o.bi;
        st(a.r + rb, idx, rv);
        acc0 += n2(rv);
    } else {
        if constexpr (sizeof(R) == 8) {
            const Sp<double> bv = ld(a.b + rb, idx);
            const Sp<double> rv{bv.ar - o.ar, bv.ai - o.ai, bv.br - o.br, bv.bi - o.bi};
            st(a.r64 + rb, idx, rv);
            if (a.p64) st(a.p64 + rb, idx, rv);
            if (a.r32 || a.p32) {
                const Sp<float> fv{(float)rv.ar, (float)rv.ai, (float)rv.br, (float)rv.bi};
                if (a.r32) st(a.r32 + rb, idx, fv);
                if (a.p32) st(a.p32 + rb, idx, fv);
            }
            acc0 += n2(rv);
            acc1 += n2(bv);
        }
    }
}

// out = ca [A | A^dag | 1] psi + ch H hin on the target parity, with the CG epilogues
// of the full-lattice stencil (store / norm / CG update / FP64 residual).
template <typename R, bool DAG, bool AOP, int EPI>
__global__ __launch_bounds__(kBlock) void k_eo_b(const EoArgs<R> a)
{
    const int total = gridDim.x;
    const int lb = a.remap ? xcd_logical_block(blockIdx.x, total) : (int)blockIdx.x;
    const EoPos<R> q = eo_pos(a, lb, kBlock, a.chunks, a.bx * a.chunks);
    const int rhs = q.rhs, bid = q.bid;
    const int bx = a.groups ? (q.V4h + kBlock - 1) / kBlock : a.bx;
    const int pstride = a.groups ? a.pstride : a.bx * a.chunks;
    const int chunk = bid / bx;
    const int hs = (bid - chunk * bx) * kBlock + threadIdx.x;
    if (a.st != nullptr && (a.st[rhs].flags[a.slot] & a.skip)) return;
    __shared__ double sh[8];
    R alpha = 0;
    if constexpr (EPI == EPI_CGUPD) {
        const double pAp = sum_partials(a.partA + (size_t)rhs * pstride * 2, pstride, sh);
        if (pAp == 0.0) return;
        alpha = (R)(a.st[rhs].rr[a.slot] / pAp);
    }
    const Lat& L = q.L;
    const int Ls = L.Ls, V4h = q.V4h;
    const long long rb = q.rb;
    double acc0 = 0.0, acc1 = 0.0;
    const int s0 = chunk * a.spt, s1 = min(Ls, s0 + a.spt);
    if (hs < V4h && s0 < s1) {
        const EoSite<R> g = eo_site<R>(L, q.Nxh, a.tpar, hs, q.links);
        const R d = (R)L.diag, mf = (R)L.mf;
        const R* in = a.in + rb;
        Sp<R> cur = ld(in, (long long)s0 * V4h + hs);
        Sp<R> dn{0, 0, 0, 0};
        if constexpr (AOP) dn = ld(in, (long long)((s0 == 0) ? Ls - 1 : s0 - 1) * V4h + hs);
        for (int s = s0; s < s1; ++s) {
            const long long sl = (long long)s * V4h;
            Sp<R> up{0, 0, 0, 0};
            if constexpr (AOP) up = ld(in, (long long)((s + 1 == Ls) ? 0 : s + 1) * V4h + hs);
            Sp<R> o;
            if constexpr (AOP) {
                // A psi: d psi + D5 (D: spin1 <- s+1, spin0 <- s-1; D^dag swapped)
                const R cup = (s + 1 == Ls) ? mf : (R)-1;
                const R cdn = (s == 0) ? mf : (R)-1;
                o = Sp<R>{d * cur.ar, d * cur.ai, d * cur.br, d * cur.bi};
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
            } else {
                o = cur;
            }
            o.ar *= a.ca;
            o.ai *= a.ca;
            o.br *= a.ca;
            o.bi *= a.ca;
            add_hop<R, DAG>(o, a.ch, g, a.hin + rb, sl);
            eo_epi<R, EPI>(a, a.p, rb, sl + hs, o, alpha, acc0, acc1);
            if constexpr (AOP) {
                dn = cur;
                cur = up;
            } else if (s + 1 < s1) {
                cur = ld(in, sl + V4h + hs);
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

// The Schur operator in one launch: out_o = [A | A^dag] psi_o + ch H w with
// w = [A^-1 | A^-dag] H psi_o on the other parity, w never leaving the chip (the
// two-kernel form writes w to device memory and reads it back: three vector passes more).
//
// A block owns a tile of ty x tx 4D sites of one right-hand side (ty, tx even, so a
// site's parity is that of its coordinates in the tile; at most 64 target sites).
//   1a  psi on the region of target-parity sites around the tile ([-2, ty+1] x
//       [-2, tx+1]: every neighbour of the box below), all Ls slices, into LDS; two
//       threads per region site, each issuing half of its slices' loads at once.
//   1b  the hop H psi at every other-parity site of the box [-1, ty] x [-1, tx] (all of
//       them neighbour a target site but two corners), two threads per site (half of
//       the slices each; the hop is local to a slice), from LDS; each phase-2 thread
//       also takes its target's psi window from LDS.  Then w overwrites psi in LDS.
//   1c  A^-1 (or A^-dag) along s, one thread per box site and spin component (the two
//       O(Ls) recurrences are independent), in place in LDS.
//   2   wavefront c walks s-chunk c (Ls / 4 slices) of target site `lane`: the A term
//       from its window in registers, the hop of w from LDS, k_eo_b's epilogue.
// The boundary signs are folded into the links (exact).  The arithmetic otherwise
// follows k_eo_a and k_eo_b; box or region sites that wrap around a small lattice are
// computed twice, into two LDS entries.
//
// TYC, TXC > 0: the tile is TYC x TXC at compile time (one lattice only), so the region and
// box strides, the index divisions and the LDS offsets fold into constants; 0: from tl.
template <typename R, bool DAG, int LS, int EPI, bool PIPE = false, int TYC = 0, int TXC = 0>
__global__ __launch_bounds__(kBlockS, sizeof(R) == 4 ? 5 : 2) void k_eo_schur(const EoArgs<R> a, const EoTile tl)
{
    extern __shared__ __align__(16) unsigned char eo_smem[];
    Sp<R>* w = reinterpret_cast<Sp<R>*>(eo_smem);  // [LS][nbox]
    __shared__ double sh[8];
    using C = typename C2T<R>::T;
    const int total = gridDim.x;
    const int lb = a.remap ? xcd_logical_block(blockIdx.x, total) : (int)blockIdx.x;
    int TY = TYC > 0 ? TYC : tl.ty, TX = TXC > 0 ? TXC : tl.tx;  // one lattice; several: each group's own tile
    Lat L = a.L;
    int rhs, tile, pstride;
    long long rb;
    const C* links;
    if (!a.groups) {
        const int tiles = (L.Nt / TY) * (L.Nx / TX);
        rhs = lb / tiles;
        tile = lb - rhs * tiles;
        pstride = tiles;
        rb = (long long)rhs * (L.V4 / 2) * LS * 4;
        links = static_cast<const C*>(a.links) + (long long)(rhs / a.rhs_per_cfg) * a.link_stride;
    } else {
        int blk0, per;
        const GroupPos p = block_group(
            a.groups, lb, [](int nt, int nx) { return (nt / eo_tile_dim(nt, 8)) * (nx / eo_tile_dim(nx, 16)); }, blk0,
            per);
        if constexpr (TYC == 0) {  // (a compile-time tile is launched for one lattice only)
            TY = eo_tile_dim(p.Nt, 8);
            TX = eo_tile_dim(p.Nx, 16);
        }
        const int local = lb - blk0, rl = local / per;
        rhs = p.rhs0 + rl;
        tile = local - rl * per;
        pstride = a.pstride;
        L.Nt = p.Nt;
        L.Nx = p.Nx;
        L.V4 = p.Nt * p.Nx;
        rb = (p.site0 / 2 + (long long)rl * (L.V4 / 2) * LS) * 4;
        links = static_cast<const C*>(a.links) + p.link0 + (long long)(rl / p.rpc) * 2 * L.V4;
    }
    // PIPE: the flags, beta and direction buffer this iteration starts from -- from the
    // state, or (fold) computed here from the previous iteration's partials exactly as
    // k_cg_flags does, by every wave alike; block 0 of the right-hand side records them.
    // The fold runs after this block's loads of r and p_old are issued, in wave sums with
    // no barrier, so the loads stay in flight meanwhile.
    int pflags = 0, pb = 0;
    R pbeta = 0;
    bool fold_pending = false;
    const bool lead = tile == 0 && threadIdx.x == 0;
    // fold: the previous slot's state, read here (before psi's loads) for the same reason
    int f_prev = 0, it_prev = 0, pb_prev = 0;
    double rr_prev = 0.0, bn2 = 1.0, tol = 0.0, del2rr = 0.0;
    if constexpr (PIPE) {
        const CgState& S = a.st[rhs];
        if (!a.fold) {
            pflags = S.flags[a.slot];
            pbeta = (R)S.beta;
            pb = S.pbuf[a.slot];
            if (pflags & a.skip) return;  // block-uniform
        } else {
            const int ps = a.slot ^ 1, f = S.flags[ps];
            if (f & (F_DONE | F_FROZEN)) {  // block-uniform: carried to this slot
                if (lead) {
                    CgState& W = a.stw[rhs];
                    W.rr[a.slot] = S.rr[ps];
                    W.pbuf[a.slot] = S.pbuf[ps];
                    W.flags[a.slot] = f;
                }
                return;
            }
            f_prev = f;
            pb_prev = S.pbuf[ps];
            it_prev = S.iters;
            rr_prev = S.rr[ps];
            bn2 = S.bnorm2;
            tol = S.tol;
            del2rr = S.delta * S.delta * S.rr_ref;
            pb = pb_prev ^ 1;  // the direction the previous iteration formed is current
            fold_pending = true;
        }
    } else {
        if (a.st != nullptr && (a.st[rhs].flags[a.slot] & a.skip)) return;  // block-uniform
    }
    R alpha = 0;
    // CGUPD: the partials of p^dag A p are loaded first and folded (in every wave, no
    // barrier) after psi's loads are issued, so their latency overlaps; entries lane and
    // lane + 64, clamped (a remainder loop covers longer partial arrays)
    double cpa0 = 0.0, cpa1 = 0.0, crr = 0.0;
    if constexpr (EPI == EPI_CGUPD) {
        const int lane = threadIdx.x & 63;
        const double* pa = a.partA + (size_t)rhs * pstride * 2;
        cpa0 = pa[2 * min(lane, pstride - 1)];
        cpa1 = pa[2 * min(lane + 64, pstride - 1)];
        crr = a.st[rhs].rr[a.slot];
    }
    const int Nxh = L.Nx / 2, V4h = L.V4 / 2;
    const int tiles_x = L.Nx / TX;
    const int tyi = tile / tiles_x;
    const int t0 = tyi * TY, x0 = (tile - tyi * tiles_x) * TX;
    const int WB = (TX + 2) / 2, nbox = (TY + 2) * WB;
    const int opar = 1 - a.tpar;
    const R* in = PIPE ? nullptr : a.in + rb;
    const R d = (R)L.diag, mf = (R)L.mf;

    // ---- phase 1a: psi on the region [-2, ty + 1] x [-2, tx + 1] of target-parity sites
    // (every neighbour of the box), all slices, into LDS: two threads per region site,
    // each loading half of its slices (independent loads, all in flight at once)
    constexpr int LH = LS / 2;
    const int WR = (TX + 4) / 2, nreg = (TY + 4) * WR;
    Sp<R>* P = w;  // [LS][nreg], later overwritten by w [LS][nbox]
    {
        const int r = threadIdx.x >> 1, h = threadIdx.x & 1;
        const bool act = r < nreg;
        long long i0 = 0;
        bool own = false;
        Sp<R> v[LH], po[LH];
        // PIPE: p = r + beta p_old (k_cg_p's update), on the fly; the tile's own sites are
        // stored to the new direction buffer (a wrapped duplicate outside the tile is not).
        // A folding launch never starts a solve, so it always has a p_old.
        const bool with_old = PIPE && (fold_pending || !(pflags & F_FRESH));
        // Loads are issued by every thread (a thread past the region's end loads its last site
        // again and stores nothing) and p_old always (in bounds; unused on a fresh start): no
        // branch around them, so the waits below count them exactly (at a join the compiler
        // would wait for every load in flight)
        // the fold's partials, loaded before psi so that waiting for them (vmcnt counts loads
        // in issue order) leaves psi's loads in flight: entries lane and lane + 64, clamped
        // (a remainder loop covers longer partial arrays)
        double fa0 = 0.0, fa1 = 0.0, fb0 = 0.0, fb1 = 0.0;
        if constexpr (PIPE) {
            if (fold_pending) {
                const int lane = threadIdx.x & 63;
                const double* pa = a.partAp + (size_t)rhs * pstride * 2;
                const double* pbp = a.partB + (size_t)rhs * pstride * 2;
                const int i0c = min(lane, pstride - 1), i1c = min(lane + 64, pstride - 1);
                fa0 = pa[2 * i0c];
                fb0 = pbp[2 * i0c];
                fa1 = pa[2 * i1c];
                fb1 = pbp[2 * i1c];  // masked where summed (a use here would wait for them at once)
            }
        }
        {
            const int rc = act ? r : nreg - 1;
            const int ir = rc / WR, jr = 2 * (rc - ir * WR) + ((ir + a.tpar) & 1);
            int t = t0 + ir - 2, x = x0 + jr - 2;
            t = t < 0 ? t + L.Nt : (t >= L.Nt ? t - L.Nt : t);
            x = x < 0 ? x + L.Nx : (x >= L.Nx ? x - L.Nx : x);
            i0 = (long long)h * LH * V4h + t * Nxh + (x >> 1);
            own = act && ir >= 2 && ir < TY + 2 && jr >= 2 && jr < TX + 2;
            const R* src = PIPE ? a.r + rb + i0 * 4 : in + i0 * 4;
#pragma unroll
            for (int k = 0; k < LH; ++k) v[k] = ld(src, (long long)k * V4h);
            if constexpr (PIPE) {
                const R* psrc = (pb ? a.pd1 : a.pd0) + rb + i0 * 4;
#pragma unroll
                for (int k = 0; k < LH; ++k) po[k] = ld(psrc, (long long)k * V4h);
            }
        }
        if constexpr (EPI == EPI_CGUPD) {
            // alpha = rr / p^dag A p: the sum of sum_partials in the same order, per wave
            const int lane = threadIdx.x & 63;
            double v = (lane < pstride ? cpa0 : 0.0) + (lane + 64 < pstride ? cpa1 : 0.0);
            if (pstride > 128) {
                const double* pa = a.partA + (size_t)rhs * pstride * 2;
                for (int i = lane + 128; i < pstride; i += 64) v += pa[2 * i];
            }
            const double pAp = wave_sum(v);
            if (pAp == 0.0) return;  // block-uniform (every wave has the same sum)
            alpha = (R)(crr / pAp);
        }
        if constexpr (PIPE) {
            if (fold_pending) {
                const int lane = threadIdx.x & 63;
                const bool m0 = lane < pstride, m1 = lane + 64 < pstride;
                double sa = (m0 ? fa0 : 0.0) + (m1 ? fa1 : 0.0), sb = (m0 ? fb0 : 0.0) + (m1 ? fb1 : 0.0);
                if (pstride > 128) {
                    const double* pa = a.partAp + (size_t)rhs * pstride * 2;
                    const double* pbp = a.partB + (size_t)rhs * pstride * 2;
                    for (int i = lane + 128; i < pstride; i += 64) {
                        sa += pa[2 * i];
                        sb += pbp[2 * i];
                    }
                }
                const double pAp = wave_sum(sa), rr_new = wave_sum(sb);  // the same in every wave
                int nf = f_prev & ~F_FRESH;
                CgState& W = a.stw[rhs];
                if (pAp == 0.0) {
                    nf |= F_DONE;
                    if (lead) {
                        W.status = 4;
                        W.rr[a.slot] = rr_prev;
                        W.pbuf[a.slot] = pb_prev;
                    }
                } else {
                    const double relres = std::sqrt(rr_new) / std::sqrt(bn2);
                    // the mixed solver's freeze (the pipelined CG runs only there)
                    if (relres <= tol || rr_new <= del2rr) nf |= F_FROZEN;
                    const double beta = rr_new / rr_prev;
                    pbeta = (R)beta;
                    if (lead) {
                        W.rr[a.slot] = rr_new;
                        W.relres = relres;
                        W.iters = it_prev + 1;
                        W.beta = beta;
                        W.pbuf[a.slot] = pb;
                    }
                }
                pflags = nf;
                if (lead) W.flags[a.slot] = nf;
                if (pflags & a.skip) return;  // block-uniform (every wave computed the same flags)
            }
            if (act) {
                if (with_old) {
                    const R beta = pbeta;
#pragma unroll
                    for (int k = 0; k < LH; ++k) {
                        v[k].ar = v[k].ar + beta * po[k].ar;
                        v[k].ai = v[k].ai + beta * po[k].ai;
                        v[k].br = v[k].br + beta * po[k].br;
                        v[k].bi = v[k].bi + beta * po[k].bi;
                    }
                }
                if (own) {
                    R* pdst = (pb ? a.pd0 : a.pd1) + rb + i0 * 4;
#pragma unroll
                    for (int k = 0; k < LH; ++k) st(pdst, (long long)k * V4h, v[k]);
                }
            }
        }
        if (act) {
#pragma unroll
            for (int k = 0; k < LH; ++k) P[(h * LH + k) * nreg + r] = v[k];
        }
    }
    __syncthreads();
    // ---- phase 1b: the hop H psi at the box's other-parity sites, two threads per site
    // (half of the slices each; the hop is local to a slice)
    const int eb = threadIdx.x >> 1, hb = threadIdx.x & 1;
    Sp<R> hv[LH];
    if (eb < nbox) {
        const int ib = eb / WB, xb = 2 * (eb - ib * WB) + ((opar + ib) & 1);
        int t = t0 + ib - 1, x = x0 + xb - 1;
        t = t < 0 ? t + L.Nt : (t >= L.Nt ? t - L.Nt : t);
        x = x < 0 ? x + L.Nx : (x >= L.Nx ? x - L.Nx : x);
        EoSite<R> g = eo_site<R>(L, Nxh, opar, t * Nxh + (x >> 1), links);
        fold_signs(g);
        // the box site's neighbours in the region: +t, -t, +x, -x
        const int rtp = (ib + 2) * WR + ((xb + 1) >> 1), rtm = ib * WR + ((xb + 1) >> 1);
        const int rxp = (ib + 1) * WR + ((xb + 2) >> 1), rxm = (ib + 1) * WR + (xb >> 1);
#pragma unroll
        for (int k = 0; k < LH; ++k) {
            const Sp<R>* Ps = P + (hb * LH + k) * nreg;
            Sp<R> o{0, 0, 0, 0};
            hop_sum<R, DAG, false>(o, (R)1, g, Ps[rtp], Ps[rtm], Ps[rxp], Ps[rxm]);
            hv[k] = o;
        }
    }
    // phase 2's thread: wavefront c walks s-chunk c (SPT slices) of target site k = lane;
    // its window psi(s0 - 1 .. s0 + SPT) comes from the region in LDS
    constexpr int SPT = LS / 4;
    const int TXh = TX / 2, ntarg = TY * TXh;
    const int k2 = threadIdx.x & 63, s0 = (threadIdx.x >> 6) * SPT;
    const int i2 = k2 / TXh, j2 = 2 * (k2 - i2 * TXh) + ((i2 + a.tpar) & 1);
    const bool act2 = k2 < ntarg;
    Sp<R> win[SPT + 2];
    if (act2) {
        const int rt = (i2 + 2) * WR + ((j2 + 2) >> 1);
#pragma unroll
        for (int q = 0; q < SPT + 2; ++q) {
            int sq = s0 - 1 + q;
            sq = sq < 0 ? sq + LS : (sq >= LS ? sq - LS : sq);
            win[q] = P[sq * nreg + rt];
        }
    }
    __syncthreads();  // every read of P is done
    if (eb < nbox) {
#pragma unroll
        for (int k = 0; k < LH; ++k) w[(hb * LH + k) * nbox + eb] = hv[k];
    }
    __syncthreads();
    // ---- phase 1c: w = A^-1 (or A^-dag) of the hop, one thread per site and spin component
    // (the two recurrences are independent); A: spin 0 lower, spin 1 upper, A^dag swapped
    {
        constexpr int kHalf = kBlockS / 2;
        const int sp = threadIdx.x >= kHalf ? 1 : 0;  // wave-uniform
        const int e = threadIdx.x - sp * kHalf;
        if (e < nbox) {
            R* col = reinterpret_cast<R*>(w) + e * 4 + 2 * sp;
            if (sp == (DAG ? 1 : 0))
                wall_solve_col<R, LS, false>(col, nbox * 4, d, mf);
            else
                wall_solve_col<R, LS, true>(col, nbox * 4, d, mf);
        }
    }
    __syncthreads();

    // ---- phase 2: the operator on the tile's target sites (ca = 1)
    double acc0 = 0.0, acc1 = 0.0;
    // the direction of x += alpha p: in the pipelined CG, this iteration's (pd[pbuf ^ 1];
    // the next state update records pbuf ^ 1 for the next slot)
    const R* pdir = a.p;
    if constexpr (EPI == EPI_CGUPD)
        if (a.pd0) pdir = a.st[rhs].pbuf[a.slot] ? a.pd0 : a.pd1;
    if (act2) {
        const int hs = (t0 + i2) * Nxh + ((x0 + j2) >> 1);
        EoSite<R> g = eo_site<R>(L, Nxh, a.tpar, hs, links);
        fold_signs(g);
        const int etp = (i2 + 2) * WB + ((j2 + 1) >> 1), etm = i2 * WB + ((j2 + 1) >> 1);
        const int exp = (i2 + 1) * WB + ((j2 + 2) >> 1), exm = (i2 + 1) * WB + (j2 >> 1);
        // CGUPD: p, x, r of slice q + 1 are loaded before slice q's stores (software
        // pipelined by hand: loads after a store to the same arrays would each wait for a
        // full round trip, as the stores may alias them); eo_epi's arithmetic
        constexpr bool PIPE2 = EPI == EPI_CGUPD;
        Sp<R> ep_p[2], ep_x[2], ep_r[2];
        if constexpr (PIPE2) {
            const long long i0 = (long long)s0 * V4h + hs;
            ep_p[0] = ld(pdir + rb, i0);
            ep_x[0] = ld(a.x + rb, i0);
            ep_r[0] = ld(a.r + rb, i0);
        }
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int s = s0 + q;
            const Sp<R>& cur = win[q + 1];
            const Sp<R>& up = win[q + 2];
            const Sp<R>& dn = win[q];
            // A psi: d psi + D5 (D: spin1 <- s+1, spin0 <- s-1; D^dag swapped)
            const R cup = (s + 1 == LS) ? mf : (R)-1;
            const R cdn = (s == 0) ? mf : (R)-1;
            Sp<R> o{d * cur.ar, d * cur.ai, d * cur.br, d * cur.bi};
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
            const Sp<R>* ws = w + s * nbox;
            hop_sum<R, DAG, false>(o, a.ch, g, ws[etp], ws[etm], ws[exp], ws[exm]);
            if constexpr (PIPE2) {
                if (q + 1 < SPT) {
                    const long long in = (long long)(s + 1) * V4h + hs;
                    ep_p[(q + 1) & 1] = ld(pdir + rb, in);
                    ep_x[(q + 1) & 1] = ld(a.x + rb, in);
                    ep_r[(q + 1) & 1] = ld(a.r + rb, in);
                }
                const long long idx = (long long)s * V4h + hs;
                const Sp<R>& pv = ep_p[q & 1];
                Sp<R> xv = ep_x[q & 1];
                xv.ar += alpha * pv.ar;
                xv.ai += alpha * pv.ai;
                xv.br += alpha * pv.br;
                xv.bi += alpha * pv.bi;
                st(a.x + rb, idx, xv);
                Sp<R> rv = ep_r[q & 1];
                rv.ar -= alpha * o.ar;
                rv.ai -= alpha * o.ai;
                rv.br -= alpha * o.br;
                rv.bi -= alpha * o.bi;
                st(a.r + rb, idx, rv);
                acc0 += n2(rv);
            } else {
                eo_epi<R, EPI>(a, pdir, rb, (long long)s * V4h + hs, o, alpha, acc0, acc1);
            }
        }
    }
    if constexpr (EPI != EPI_STORE) {
        acc0 = wave_sum(acc0);
        acc1 = wave_sum(acc1);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        __syncthreads();  // sh may still hold sum_partials' value being read
        if (lane == 0) {
            sh[wv] = acc0;
            sh[4 + wv] = acc1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s0 = 0.0, s1 = 0.0;
            for (int v = 0; v < kBlockS / 64; ++v) {
                s0 += sh[v];
                s1 += sh[4 + v];
            }
            double* pp = a.part + ((size_t)rhs * pstride + tile) * 2;
            pp[0] = s0;
            pp[1] = s1;
        }
    }
}

// full lexicographic [rhs][s][t][x] <-> halves [rhs][s][t][xh] (FP64)
__global__ void k_eo_split(const VecLayout lay, const double* full, double* ev, double* od)
{
    const int rhs = blockIdx.y;
    const GroupPos gp = group_for_rhs(lay.gr, rhs);
    const int Nx = gp.Nx, V4 = gp.Nt * Nx, V4h = V4 / 2;
    long long V5, off;
    lay.locate(rhs, V5, off);
    const double4* F = reinterpret_cast<const double4*>(full) + off;
    double4* E = reinterpret_cast<double4*>(ev) + off / 2;
    double4* O = reinterpret_cast<double4*>(od) + off / 2;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < V5; i += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(i / V4), site = (int)(i - (long long)s * V4);
        const int t = site / Nx, x = site - t * Nx;
        const long long h = (long long)s * V4h + t * (Nx / 2) + (x >> 1);
        if (((t + x) & 1) == 0) E[h] = F[i];
        else O[h] = F[i];
    }
}

__global__ void k_eo_merge(const VecLayout lay, const double* ev, const double* od, double* full)
{
    const int rhs = blockIdx.y;
    const GroupPos gp = group_for_rhs(lay.gr, rhs);
    const int Nx = gp.Nx, V4 = gp.Nt * Nx, V4h = V4 / 2;
    long long V5, off;
    lay.locate(rhs, V5, off);
    double4* F = reinterpret_cast<double4*>(full) + off;
    const double4* E = reinterpret_cast<const double4*>(ev) + off / 2;
    const double4* O = reinterpret_cast<const double4*>(od) + off / 2;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < V5; i += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(i / V4), site = (int)(i - (long long)s * V4);
        const int t = site / Nx, x = site - t * Nx;
        const long long h = (long long)s * V4h + t * (Nx / 2) + (x >> 1);
        F[i] = (((t + x) & 1) == 0) ? E[h] : O[h];
    }
}

template <typename R, bool DAG, int LS>
void dispatch_a(bool ainv, const EoArgs<R>& a, unsigned grid, hipStream_t s)
{
    if (ainv)
        k_eo_a<R, DAG, LS, true><<<grid, kBlockA, 0, s>>>(a);
    else
        k_eo_a<R, DAG, LS, false><<<grid, kBlockA, 0, s>>>(a);
}

template <typename R, bool DAG, bool AOP>
void dispatch_b(int epi, const EoArgs<R>& a, unsigned grid, hipStream_t s)
{
    switch (epi) {
        case EPI_STORE: k_eo_b<R, DAG, AOP, EPI_STORE><<<grid, kBlock, 0, s>>>(a); break;
        case EPI_NORM: k_eo_b<R, DAG, AOP, EPI_NORM><<<grid, kBlock, 0, s>>>(a); break;
        case EPI_CGUPD: k_eo_b<R, DAG, AOP, EPI_CGUPD><<<grid, kBlock, 0, s>>>(a); break;
        default:
            if constexpr (sizeof(R) == 8) k_eo_b<R, DAG, AOP, EPI_RESID><<<grid, kBlock, 0, s>>>(a);
            break;
    }
}

}  // namespace

bool eo_supported(const Lat& L)
{
    if (!L.d5 || (L.Nt & 1) || (L.Nx & 1) || L.Nt < 2 || L.Nx < 2) return false;
    return L.Ls == 4 || L.Ls == 8 || L.Ls == 12 || L.Ls == 16;
}

namespace {
// Grid size of an even-odd launch: per target sites per block and `chunks` s-chunks per
// right-hand side, over one lattice or the groups of hg.
template <typename R> unsigned eo_grid(const EoArgs<R>& a, int nrhs, int per, int chunks, const Groups* hg)
{
    long long nb = 0;
    if (!a.groups) {
        nb = (long long)((a.V4h + per - 1) / per) * chunks * nrhs;
    } else {
        if (!hg || hg->n < 1 || hg->n > kMaxGroups || hg->rhs0[hg->n] != nrhs)
            throw DwfError(DWF_ERR_ARG, "even-odd launch: bad lattice-group layout");
        for (int g = 0; g < hg->n; ++g)
            nb += (long long)(hg->rhs0[g + 1] - hg->rhs0[g]) * ((hg->Nt[g] * hg->Nx[g] / 2 + per - 1) / per) * chunks;
    }
    if (nb <= 0 || nb > 0x7fffffffLL) throw DwfError(DWF_ERR_ARG, "even-odd launch: grid size out of range");
    return (unsigned)nb;
}
}  // namespace

int eo_pstride(const Groups& G, int chunks)
{
    int m = 0;
    for (int g = 0; g < G.n; ++g) m = std::max(m, (G.Nt[g] * G.Nx[g] / 2 + kBlock - 1) / kBlock * chunks);
    return m;
}

template <typename R>
void launch_eo_a(bool dagger, bool ainv, const EoArgs<R>& a, int nrhs, hipStream_t s, const Groups* hg)
{
    const unsigned grid = eo_grid(a, nrhs, kBlockA, 1, hg);
    const int Ls = a.L.Ls;
#define DWF_EO_A(LSV)                                      \
    if (Ls == LSV) {                                       \
        if (dagger)                                        \
            dispatch_a<R, true, LSV>(ainv, a, grid, s);    \
        else                                               \
            dispatch_a<R, false, LSV>(ainv, a, grid, s);   \
        return;                                            \
    }
    DWF_EO_A(4)
    DWF_EO_A(8)
    DWF_EO_A(12)
    DWF_EO_A(16)
#undef DWF_EO_A
}
template void launch_eo_a<float>(bool, bool, const EoArgs<float>&, int, hipStream_t, const Groups*);
template void launch_eo_a<double>(bool, bool, const EoArgs<double>&, int, hipStream_t, const Groups*);

template <typename R>
void launch_eo_b(bool dagger, bool aop, int epi, const EoArgs<R>& a0, int nrhs, hipStream_t s, const Groups* hg)
{
    EoArgs<R> a = a0;
    if (a.groups) a.pstride = eo_pstride(*hg, a.chunks);
    const unsigned grid = eo_grid(a, nrhs, kBlock, a.chunks, hg);
    if (dagger) {
        if (aop)
            dispatch_b<R, true, true>(epi, a, grid, s);
        else
            dispatch_b<R, true, false>(epi, a, grid, s);
    } else {
        if (aop)
            dispatch_b<R, false, true>(epi, a, grid, s);
        else
            dispatch_b<R, false, false>(epi, a, grid, s);
    }
}
template void launch_eo_b<float>(bool, bool, int, const EoArgs<float>&, int, hipStream_t, const Groups*);
template void launch_eo_b<double>(bool, bool, int, const EoArgs<double>&, int, hipStream_t, const Groups*);

bool eo_fused_tile(const Groups& G, int Ls, EoTile& t)
{
    if (Ls < 4 || Ls % 4) return false;
    for (int g = 0; g < G.n; ++g) {  // every lattice needs a tile of its own (its group's, in the kernel)
        const int ty = eo_tile_dim(G.Nt[g], 8), tx = eo_tile_dim(G.Nx[g], 16);
        if (!ty || !tx || ty * tx / 2 > 64) return false;
    }
    t.ty = eo_tile_dim(G.Nt[0], 8);  // the tile of a one-lattice launch
    t.tx = eo_tile_dim(G.Nx[0], 16);
    t.nch = 4;  // one wavefront per s-chunk, one lane per target site
    t.spt = Ls / 4;
    return true;
}

namespace {
int eo_group_tiles(const Groups& G, int g) { return (G.Nt[g] / eo_tile_dim(G.Nt[g], 8)) * (G.Nx[g] / eo_tile_dim(G.Nx[g], 16)); }
}  // namespace

int eo_fused_pstride(const Groups& G, const EoTile& t)
{
    if (G.n == 1) return G.Nt[0] * G.Nx[0] / (t.ty * t.tx);
    int m = 0;
    for (int g = 0; g < G.n; ++g) m = std::max(m, eo_group_tiles(G, g));
    return m;
}

namespace {
template <typename R, bool DAG, int LS>
void dispatch_schur(int epi, const EoArgs<R>& a, const EoTile& t, unsigned grid, size_t lds, hipStream_t s)
{
    if constexpr (sizeof(R) == 4) {
        if (!a.groups && t.ty == 8 && t.tx == 16) {  // the FP32 CG's kernels on the 8 x 16 tile
            switch (epi) {
                case EPI_STORE: k_eo_schur<R, DAG, LS, EPI_STORE, false, 8, 16><<<grid, kBlockS, lds, s>>>(a, t); return;
                case EPI_NORM: k_eo_schur<R, DAG, LS, EPI_NORM, false, 8, 16><<<grid, kBlockS, lds, s>>>(a, t); return;
                case EPI_CGUPD: k_eo_schur<R, DAG, LS, EPI_CGUPD, false, 8, 16><<<grid, kBlockS, lds, s>>>(a, t); return;
                default: break;
            }
        }
    }
    switch (epi) {
        case EPI_STORE: k_eo_schur<R, DAG, LS, EPI_STORE><<<grid, kBlockS, lds, s>>>(a, t); break;
        case EPI_NORM: k_eo_schur<R, DAG, LS, EPI_NORM><<<grid, kBlockS, lds, s>>>(a, t); break;
        case EPI_CGUPD: k_eo_schur<R, DAG, LS, EPI_CGUPD><<<grid, kBlockS, lds, s>>>(a, t); break;
        default:
            if constexpr (sizeof(R) == 8) k_eo_schur<R, DAG, LS, EPI_RESID><<<grid, kBlockS, lds, s>>>(a, t);
            break;
    }
}
}  // namespace

template <typename R>
void launch_eo_schur(bool dagger, int epi, const EoArgs<R>& a0, const EoTile& t, int nrhs, hipStream_t s,
                     const Groups* hg)
{
    EoArgs<R> a = a0;
    long long nb = 0;
    if (!a.groups) {
        if (a.L.Nt % t.ty || a.L.Nx % t.tx) throw DwfError(DWF_ERR_ARG, "fused even-odd launch: tile does not fit");
        nb = (long long)(a.L.V4 / (t.ty * t.tx)) * nrhs;
    } else {
        if (!hg || hg->n < 1 || hg->n > kMaxGroups || hg->rhs0[hg->n] != nrhs)
            throw DwfError(DWF_ERR_ARG, "fused even-odd launch: bad lattice-group layout");
        EoTile chk;
        if (!eo_fused_tile(*hg, a.L.Ls, chk)) throw DwfError(DWF_ERR_ARG, "fused even-odd launch: no tile fits a lattice");
        for (int g = 0; g < hg->n; ++g) nb += (long long)(hg->rhs0[g + 1] - hg->rhs0[g]) * eo_group_tiles(*hg, g);
        a.pstride = eo_fused_pstride(*hg, t);
    }
    if (nb <= 0 || nb > 0x7fffffffLL) throw DwfError(DWF_ERR_ARG, "fused even-odd launch: grid size out of range");
    if (a.ca != (R)1) throw DwfError(DWF_ERR_ARG, "fused even-odd launch: the A term's coefficient must be 1");
    if (t.ty < 2 || t.ty > 8 || (t.ty & 1) || t.tx < 2 || t.tx > 16 || (t.tx & 1) || t.nch < 1 || t.spt < 1 ||
        t.nch != 4 || t.spt * 4 != a.L.Ls || t.ty * t.tx / 2 > 64 || (t.ty + 4) * ((t.tx + 4) / 2) > kBlockS / 2)
        throw DwfError(DWF_ERR_ARG, "fused even-odd launch: bad tile");
    // psi's region (phase 1), later w's box (phase 2); several lattices: the largest tile's
    const size_t lds = (size_t)(a.groups ? 12 * 10 : (t.ty + 4) * ((t.tx + 4) / 2)) * a.L.Ls * sizeof(Sp<R>);
    const unsigned grid = (unsigned)nb;
    const int Ls = a.L.Ls;
#define DWF_EO_S(LSV)                                                  \
    if (Ls == LSV) {                                                   \
        if (dagger)                                                    \
            dispatch_schur<R, true, LSV>(epi, a, t, grid, lds, s);     \
        else                                                           \
            dispatch_schur<R, false, LSV>(epi, a, t, grid, lds, s);    \
        return;                                                        \
    }
    DWF_EO_S(4)
    DWF_EO_S(8)
    DWF_EO_S(12)
    DWF_EO_S(16)
#undef DWF_EO_S
    throw DwfError(DWF_ERR_ARG, "fused even-odd launch: Ls must be 4, 8, 12 or 16");
}
template void launch_eo_schur<float>(bool, int, const EoArgs<float>&, const EoTile&, int, hipStream_t, const Groups*);

void launch_eo_schur_pipe(const EoArgs<float>& a0, const EoTile& t, int nrhs, hipStream_t s, const Groups* hg)
{
    EoArgs<float> a = a0;
    if (!a.st || !a.r || !a.pd0 || !a.pd1 || !a.out || !a.part)
        throw DwfError(DWF_ERR_ARG, "pipelined even-odd launch: missing vectors or states");
    if (a.fold && (!a.partAp || !a.partB || !a.stw || a.partAp == a.part))
        throw DwfError(DWF_ERR_ARG, "pipelined even-odd launch: the state fold needs the previous partials");
    long long nb = 0;
    if (!a.groups) {
        nb = (long long)(a.L.V4 / (t.ty * t.tx)) * nrhs;
    } else {
        if (!hg || hg->n < 1 || hg->n > kMaxGroups || hg->rhs0[hg->n] != nrhs)
            throw DwfError(DWF_ERR_ARG, "pipelined even-odd launch: bad lattice-group layout");
        EoTile chk;
        if (!eo_fused_tile(*hg, a.L.Ls, chk)) throw DwfError(DWF_ERR_ARG, "pipelined even-odd launch: no tile fits a lattice");
        for (int g = 0; g < hg->n; ++g) nb += (long long)(hg->rhs0[g + 1] - hg->rhs0[g]) * eo_group_tiles(*hg, g);
        a.pstride = eo_fused_pstride(*hg, t);
    }
    if (nb <= 0 || nb > 0x7fffffffLL || t.nch != 4 || t.spt * 4 != a.L.Ls || a.ca != 1.f)
        throw DwfError(DWF_ERR_ARG, "pipelined even-odd launch: bad shape");
    const size_t lds = (size_t)(a.groups ? 12 * 10 : (t.ty + 4) * ((t.tx + 4) / 2)) * a.L.Ls * sizeof(Sp<float>);
    const unsigned grid = (unsigned)nb;
    if (!a.groups && t.ty == 8 && t.tx == 16) {  // compile-time tile
        switch (a.L.Ls) {
            case 4: k_eo_schur<float, false, 4, EPI_NORM, true, 8, 16><<<grid, kBlockS, lds, s>>>(a, t); return;
            case 8: k_eo_schur<float, false, 8, EPI_NORM, true, 8, 16><<<grid, kBlockS, lds, s>>>(a, t); return;
            case 12: k_eo_schur<float, false, 12, EPI_NORM, true, 8, 16><<<grid, kBlockS, lds, s>>>(a, t); return;
            case 16: k_eo_schur<float, false, 16, EPI_NORM, true, 8, 16><<<grid, kBlockS, lds, s>>>(a, t); return;
            default: break;
        }
    }
    switch (a.L.Ls) {
        case 4: k_eo_schur<float, false, 4, EPI_NORM, true><<<grid, kBlockS, lds, s>>>(a, t); break;
        case 8: k_eo_schur<float, false, 8, EPI_NORM, true><<<grid, kBlockS, lds, s>>>(a, t); break;
        case 12: k_eo_schur<float, false, 12, EPI_NORM, true><<<grid, kBlockS, lds, s>>>(a, t); break;
        case 16: k_eo_schur<float, false, 16, EPI_NORM, true><<<grid, kBlockS, lds, s>>>(a, t); break;
        default: throw DwfError(DWF_ERR_ARG, "pipelined even-odd launch: Ls must be 4, 8, 12 or 16");
    }
}
template void launch_eo_schur<double>(bool, int, const EoArgs<double>&, const EoTile&, int, hipStream_t, const Groups*);

namespace {
unsigned split_blocks(const VecLayout& lay)
{
    long long m = 0;
    for (int g = 0; g < lay.gr.n; ++g) m = std::max(m, lay.sites(g));
    return (unsigned)std::min<long long>((m + 255) / 256, 4096);
}
}  // namespace

void launch_eo_split(const VecLayout& lay, const double* full, double* ev, double* od, int nrhs, hipStream_t s)
{
    k_eo_split<<<dim3(split_blocks(lay), nrhs), 256, 0, s>>>(lay, full, ev, od);
}

void launch_eo_merge(const VecLayout& lay, const double* ev, const double* od, double* full, int nrhs, hipStream_t s)
{
    k_eo_merge<<<dim3(split_blocks(lay), nrhs), 256, 0, s>>>(lay, ev, od, full);
}

}  // namespace dwfk
