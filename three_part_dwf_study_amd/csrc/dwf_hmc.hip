// dwf_hmc.hip -- FP64 kernels of the HMC update around the fermion solves.
//
//   * pseudofermion force (domain_wall/include/pseudofermion.hpp:58-157 and
//     wilson/include/gauge.hpp:84-119) in gather form: one thread per link pair
//     (t,x) sums over s both terms that touch U_t(t,x) and both that touch U_x(t,x),
//     so no scatter-add or float atomics are needed and the result is deterministic;
//   * Wilson gauge action/force and plaquette (domain_wall/include/gauge.hpp:15-73);
//   * leapfrog momentum and angle updates, kinetic energy, dot products.
// All reductions: fixed-order wavefront trees, one partial per block, then one block
// folds the partials in index order into a device scalar.

#include "dwf_internal.hpp"

#include <cmath>

namespace dwfk {
namespace {

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block reduction of one value into part[blockIdx.x] (blockDim.x == kBlock).
__device__ __forceinline__ void block_partial(double v, double* part)
{
    __shared__ double sh[kBlock / 64];
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) s += sh[i];
        part[blockIdx.x] = s;
    }
}

// out[slot] = sum of n partials (one 64-thread block, fixed order).
__global__ void k_fold(const double* part, int n, double* out, int slot, double scale)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) v += part[i];
    v = wsum(v);
    if (threadIdx.x == 0) out[slot] = scale * v;
}

__device__ __forceinline__ double plaq(const double* th, int Nt, int Nx, int t, int x)
{
    const int V4 = Nt * Nx;
    const int tp = (t + 1 == Nt) ? 0 : t + 1;
    const int xp = (x + 1 == Nx) ? 0 : x + 1;
    return th[t * Nx + x] + th[V4 + tp * Nx + x] - th[t * Nx + xp] - th[V4 + t * Nx + x];
}

// force[mu][t][x] = beta dS_g/dtheta; partials of beta (1 - cos P) and of cos P.
__global__ __launch_bounds__(kBlock) void k_gauge(const double* th, int Nt, int Nx, double beta,
                                                  double* force, double* part_s, double* part_c)
{
    const int V4 = Nt * Nx;
    double s_act = 0.0, s_cos = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < V4; i += gridDim.x * kBlock) {
        const int t = i / Nx, x = i - t * Nx;
        const int tm = (t == 0) ? Nt - 1 : t - 1;
        const int xm = (x == 0) ? Nx - 1 : x - 1;
        const double p = plaq(th, Nt, Nx, t, x);
        const double sp = sin(p), cp = cos(p);
        const double sp_xm = sin(plaq(th, Nt, Nx, t, xm));
        const double sp_tm = sin(plaq(th, Nt, Nx, tm, x));
        if (force) {
            force[i] = beta * (sp - sp_xm);
            force[V4 + i] = beta * (sp_tm - sp);
        }
        s_act += beta * (1.0 - cp);
        s_cos += cp;
    }
    if (part_s) block_partial(s_act, part_s);
    if (part_c) {
        __syncthreads();
        block_partial(s_cos, part_c);
    }
}

// Pseudofermion force from eta and xi = D^dag eta (FP64), gather form.
//   F_t(t,x) = -2 Re sum_s [ (-i/2) U_t (eta0-eta1)^*(t,x) a   + (i/2) U_t^* (eta0+eta1)^*(t+1,x) b ]
//     a = bc (xi0 - xi1)(t+1,x),  b = bc (xi0 + xi1)(t,x)
//   F_x(t,x) = -2 Re sum_s [ (-i/2) U_x (eta0+i eta1)^*(t,x) c + (i/2) U_x^* (eta0-i eta1)^*(t,x+1) d ]
//     c = bc (xi0 + i xi1)(t,x+1), d = bc (xi0 - i xi1)(t,x)
// bc is the boundary sign of the link crossing t=Nt-1 (resp. x=Nx-1).
__global__ __launch_bounds__(kBlock) void k_pf_force(const Lat L, const double2* links, const double* eta,
                                                     const double* xi, double* force)
{
    const int V4 = L.V4, Nt = L.Nt, Nx = L.Nx;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= V4) return;
    const int t = i / Nx, x = i - t * Nx;
    const int tp = (t + 1 == Nt) ? 0 : t + 1;
    const int xp = (x + 1 == Nx) ? 0 : x + 1;
    const double bt = (t == Nt - 1) ? L.bct : 1.0;
    const double bx = (x == Nx - 1) ? L.bcx : 1.0;
    const int jt = tp * Nx + x, jx = t * Nx + xp;
    const double4* E = reinterpret_cast<const double4*>(eta);
    const double4* X = reinterpret_cast<const double4*>(xi);
    // accumulate sum_s of the complex brackets, then apply the links once
    double fr_t = 0, fi_t = 0, br_t = 0, bi_t = 0, fr_x = 0, fi_x = 0, br_x = 0, bi_x = 0;
    for (int s = 0; s < L.Ls; ++s) {
        const long long o = (long long)s * V4;
        const double4 eh = E[o + i], et = E[o + jt], ex = E[o + jx];
        const double4 xh = X[o + i], xt = X[o + jt], xx = X[o + jx];
        {   // (eta0 - eta1)^*(here) * (xi0 - xi1)(t+1)
            const double ur = eh.x - eh.z, ui = eh.y - eh.w;
            const double ar = xt.x - xt.z, ai = xt.y - xt.w;
            fr_t += ur * ar + ui * ai;
            fi_t += ur * ai - ui * ar;
        }
        {   // (eta0 + eta1)^*(t+1) * (xi0 + xi1)(here)
            const double ur = et.x + et.z, ui = et.y + et.w;
            const double br = xh.x + xh.z, bi = xh.y + xh.w;
            br_t += ur * br + ui * bi;
            bi_t += ur * bi - ui * br;
        }
        {   // (eta0 + i eta1)^*(here) * (xi0 + i xi1)(x+1)
            const double ur = eh.x - eh.w, ui = eh.y + eh.z;
            const double cr = xx.x - xx.w, ci = xx.y + xx.z;
            fr_x += ur * cr + ui * ci;
            fi_x += ur * ci - ui * cr;
        }
        {   // (eta0 - i eta1)^*(x+1) * (xi0 - i xi1)(here)
            const double ur = ex.x + ex.w, ui = ex.y - ex.z;
            const double dr = xh.x + xh.w, di = xh.y - xh.z;
            br_x += ur * dr + ui * di;
            bi_x += ur * di - ui * dr;
        }
    }
    const double2 ut = links[i], ux = links[V4 + i];
    // Re[(-i/2) U z] = (1/2) Im(U z);  Re[(i/2) U^* z] = -(1/2) Im(U^* z)
    const double im_fwd_t = ut.x * fi_t + ut.y * fr_t;   // Im(U_t z_f)
    const double im_bwd_t = ut.x * bi_t - ut.y * br_t;   // Im(U_t^* z_b)
    const double im_fwd_x = ux.x * fi_x + ux.y * fr_x;
    const double im_bwd_x = ux.x * bi_x - ux.y * br_x;
    force[i] = -2.0 * bt * (0.5 * im_fwd_t - 0.5 * im_bwd_t);
    force[V4 + i] = -2.0 * bx * (0.5 * im_fwd_x - 0.5 * im_bwd_x);
}

// Re <a, b> partials over n complex entries (interleaved), one per block.
__global__ __launch_bounds__(kBlock) void k_redot(const double* a, const double* b, long long n, double* part)
{
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        v += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
    block_partial(v, part);
}

// p -= c * (fa + fb); optional partial of 1/2 p^2 (after the update).
__global__ __launch_bounds__(kBlock) void k_mom(double* p, const double* fa, const double* fb, int n, double c,
                                                double* part_k)
{
    double k = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double v = p[i];
        if (c != 0.0) v -= c * (fa[i] + fb[i]);
        p[i] = v;
        k += 0.5 * v * v;
    }
    if (part_k) block_partial(k, part_k);
}

// theta += eps p (optionally wrapped to [-pi, pi)), and repack the links.
__global__ __launch_bounds__(kBlock) void k_theta(double* th, const double* p, int n, double eps, int wrap,
                                                  double2* l64, float2* l32)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double a = th[i] + eps * p[i];
    if (wrap) {  // wilson/include/indexing.hpp:13-18
        const double pi = 3.14159265358979323846;
        a = fmod(a + pi, 2.0 * pi);
        if (a < 0.0) a += 2.0 * pi;
        a = a - pi;
    }
    th[i] = a;
    double s, c;
    sincos(a, &s, &c);
    l64[i] = make_double2(c, s);
    l32[i] = make_float2((float)c, (float)s);
}

}  // namespace

int reduce_blocks(long long n) { return (int)std::min<long long>((n + kBlock - 1) / kBlock, 1024); }

void launch_gauge(const Lat& L, const double* th, double beta, double* force, double* part_s,
                  double* part_c, double* scalars, int slot_s, int slot_c, hipStream_t s)
{
    const int nb = reduce_blocks(L.V4);
    k_gauge<<<nb, kBlock, 0, s>>>(th, L.Nt, L.Nx, beta, force, part_s, part_c);
    if (part_s) k_fold<<<1, 64, 0, s>>>(part_s, nb, scalars, slot_s, 1.0);
    if (part_c) k_fold<<<1, 64, 0, s>>>(part_c, nb, scalars, slot_c, 1.0 / L.V4);
}

void launch_pf_force(const Lat& L, const double2* l64, const double* eta, const double* xi, double* force,
                     hipStream_t s)
{
    k_pf_force<<<(L.V4 + kBlock - 1) / kBlock, kBlock, 0, s>>>(L, l64, eta, xi, force);
}

void launch_redot(const double* a, const double* b, long long ncomplex, double* part, double* scalars, int slot,
                  hipStream_t s)
{
    const int nb = reduce_blocks(ncomplex);
    k_redot<<<nb, kBlock, 0, s>>>(a, b, ncomplex, part);
    k_fold<<<1, 64, 0, s>>>(part, nb, scalars, slot, 1.0);
}

void launch_mom(double* p, const double* fa, const double* fb, int n, double c, double* part_k, double* scalars,
                int slot, hipStream_t s)
{
    const int nb = reduce_blocks(n);
    k_mom<<<nb, kBlock, 0, s>>>(p, fa, fb, n, c, part_k);
    if (part_k) k_fold<<<1, 64, 0, s>>>(part_k, nb, scalars, slot, 1.0);
}

void launch_theta(double* th, const double* p, int n, double eps, int wrap, double2* l64, float2* l32,
                  hipStream_t s)
{
    k_theta<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(th, p, n, eps, wrap, l64, l32);
}

}  // namespace dwfk
