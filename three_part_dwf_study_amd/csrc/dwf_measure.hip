// dwf_measure.hip -- the DWF analyzer's propagator solves and contractions
// (domain_wall/include/correlator.hpp:27-310) on the device.
//
// One configuration's sources -- 4 boundary wall sources (compute_CPP_CJ5), 4 boundary
// point sources (compute_momentum_CPP) and N Z4 noise vectors
// (estimate_pseudoscalar_loop_density) -- go through ONE batched multi-RHS CG on
// D^dag D with right-hand sides D^dag src.  The contractions run in FP64 on the device;
// only Nt-length correlators and the V4 loop density leave it.
//
// Quark field of a 5D solution X: q(x) = P_L X(s=0, x) + P_R X(s=Ls-1, x), i.e.
// q0 = X(Ls-1, x).spin0 and q1 = X(0, x).spin1 (P_R = diag(1,0), P_L = diag(0,1)).

#include "../../include/dwf_hip.h"
#include "dwf_context.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace dwfk;

namespace {

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the block of v (result in every thread).  Every thread must call it.
__device__ __forceinline__ double bsum(double v, double* sh)
{
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
    __syncthreads();
    return r;
}

// Geometry of configuration `cfg` (K right-hand sides each) in a batch over lattice
// groups: its group, lattice, and the offset (in sites) of its first right-hand side.
struct CfgGeo {
    int g, Nt, Nx, V4;
    long long V5, site;
};
__device__ __forceinline__ CfgGeo cfg_geo(const VecLayout& lay, int K, int cfg)
{
    CfgGeo c;
    const int rhs = cfg * K;
    c.g = group_of_rhs(lay.gr, rhs);
    c.Nt = lay.gr.Nt[c.g];
    c.Nx = lay.gr.Nx[c.g];
    c.V4 = c.Nt * c.Nx;
    c.V5 = (long long)c.V4 * lay.Ls;
    c.site = lay.offset(c.g, rhs);
    return c;
}

// Sources.  Wall: all x at t_src; point: the single site (t_src, x_src).  The source
// spin ss fills P_L e_ss at s=0 and P_R e_ss at s=Ls-1 (correlator.hpp:131-138, 199-211).
// The 4 wall / 4 reference point right-hand sides are ordered (t_src=0, ss=0), (0, 1),
// (Nt/2, 0), (Nt/2, 1), point ones at x_src = 0.  With psite ([cfg][2] = (t0, x0)) the
// point block is 2 right-hand sides (ss = 0, 1) at that site.  Noise: src = P_L eta at
// s=0, + P_R eta at s=Ls-1 (:60-75); configuration cfg's noise vectors start at
// noise_all + noff[cfg] (nnoise vectors of V4 x 2 complex).
__global__ void k_sources(const VecLayout lay, const int nwall, const int npoint, const int nnoise,
                          const double* noise_all, const long long* noff, const int* psite, double* src)
{
    const int K = nwall + npoint + nnoise;       // right-hand sides per configuration
    const int cfg = blockIdx.y / K, rhs = blockIdx.y - cfg * K;
    const CfgGeo G = cfg_geo(lay, K, cfg);
    const int Nt = G.Nt, Nx = G.Nx, V4 = G.V4, Ls = lay.Ls;
    const long long V5 = G.V5;
    const double* noise = noise_all ? noise_all + noff[cfg] : nullptr;
    double4* out = reinterpret_cast<double4*>(src) + G.site + (long long)rhs * V5;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V4; i += gridDim.x * blockDim.x) {
        const int t = i / Nx, x = i - t * Nx;
        double4 lo = make_double4(0, 0, 0, 0), hi = make_double4(0, 0, 0, 0);  // s = 0, s = Ls-1
        if (rhs < nwall + npoint) {
            int t_src, x_src = 0, ss;
            if (rhs < nwall || !psite) {
                const int k = rhs < nwall ? rhs : rhs - nwall;
                t_src = (k >> 1) ? Nt / 2 : 0;
                ss = k & 1;
            } else {
                t_src = psite[2 * cfg];
                x_src = psite[2 * cfg + 1];
                ss = rhs - nwall;
            }
            const bool on = (t == t_src) && (rhs < nwall || x == x_src);
            if (on) {
                if (ss == 1) lo.z = 1.0;  // P_L e_1 = (0, 1) at s = 0
                else hi.x = 1.0;          // P_R e_0 = (1, 0) at s = Ls-1
            }
        } else {
            const int n = rhs - nwall - npoint;
            const double* e = noise + ((long long)n * V4 + i) * 4;  // eta(t,x,spin) complex
            lo.z = e[2];
            lo.w = e[3];
            hi.x = e[0];
            hi.y = e[1];
        }
        if (Ls == 1) {
            out[i] = make_double4(hi.x, hi.y, lo.z, lo.w);
        } else {
            out[i] = lo;
            out[(long long)(Ls - 1) * V4 + i] = hi;
        }
        for (int s = 1; s < Ls - 1; ++s) out[(long long)s * V4 + i] = make_double4(0, 0, 0, 0);
    }
}

// C_PP / C_J5 of the wall solutions (correlator.hpp:258-300): block per (t, source time k,
// configuration), outputs pp[cfg][k][t], j5[cfg][k][t] (stride ntmax per source time)
// summed over x and source spin.
__global__ void k_contract_wall(const VecLayout lay, const int K, const int ntmax, const double* X, double* pp_all,
                                double* j5_all)
{
    __shared__ double sh[8];
    const int t = blockIdx.x, k = blockIdx.y, cfg = blockIdx.z;
    const CfgGeo G = cfg_geo(lay, K, cfg);
    if (t >= G.Nt) return;  // block-uniform
    const int Nx = G.Nx, V4 = G.V4, Ls = lay.Ls;
    const long long V5 = G.V5;
    const double4* P = reinterpret_cast<const double4*>(X) + G.site;
    double* pp = pp_all + (long long)cfg * 2 * ntmax;
    double* j5 = j5_all + (long long)cfg * 2 * ntmax;
    double a = 0.0, b = 0.0;
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
        const int i = t * Nx + x;
        for (int ss = 0; ss < 2; ++ss) {
            const double4* Y = P + (2 * k + ss) * V5;
            const double4 l = Y[i], r = Y[(long long)(Ls - 1) * V4 + i], m = Y[(long long)(Ls / 2) * V4 + i];
            const double q0r = r.x, q0i = r.y, q1r = l.z, q1i = l.w;
            a += q0r * q0r + q0i * q0i + q1r * q1r + q1i * q1i;
            b += m.x * q0r + m.y * q0i + m.z * q1r + m.w * q1i;  // Re(mid^* q), mid unprojected
        }
    }
    a = bsum(a, sh);
    b = bsum(b, sh);
    if (threadIdx.x == 0) {
        pp[k * ntmax + t] = a;
        j5[k * ntmax + t] = b;
    }
}

// Point correlators (correlator.hpp:149-174): block per (t, source time k, configuration);
// out[cfg][k][n][t] = sum_x cos(2 pi n (x - x0) / Nx) sum_ss |q|^2 and j5[cfg][k][t] the
// midpoint correlator of k_contract_wall (stride ntmax per row).  nsrc_t source times of
// 2 spins each follow one another in the point block, which starts at right-hand side
// `first` of each configuration; x0 = psite[cfg][1] (0 without psite).
__global__ void k_contract_point(const VecLayout lay, const int nmom, const int K, const int first, const int nsrc_t,
                                 const int ntmax, const int* psite, const double* X, double* out_all, double* j5_all)
{
    __shared__ double sh[8];
    const int t = blockIdx.x, k = blockIdx.y, cfg = blockIdx.z;
    const CfgGeo G = cfg_geo(lay, K, cfg);
    if (t >= G.Nt) return;  // block-uniform
    const int Nx = G.Nx, V4 = G.V4, Ls = lay.Ls;
    const long long V5 = G.V5;
    const double4* P = reinterpret_cast<const double4*>(X) + G.site + (long long)first * V5;
    double* out = out_all + (long long)cfg * nsrc_t * nmom * ntmax;
    double* j5 = j5_all + (long long)cfg * nsrc_t * ntmax;
    const int x0 = psite ? psite[2 * cfg + 1] : 0;
    const double pi = 3.141592653589793238462643383279502884;
    double b = 0.0;
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
        const int i = t * Nx + x;
        for (int ss = 0; ss < 2; ++ss) {
            const double4* Y = P + (2 * k + ss) * V5;
            const double4 l = Y[i], r = Y[(long long)(Ls - 1) * V4 + i], m = Y[(long long)(Ls / 2) * V4 + i];
            b += m.x * r.x + m.y * r.y + m.z * l.z + m.w * l.w;
        }
    }
    b = bsum(b, sh);
    if (threadIdx.x == 0) j5[(long long)k * ntmax + t] = b;
    for (int n = 0; n < nmom; ++n) {
        const double p = 2.0 * pi * n / Nx;
        double v = 0.0;
        for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
            const int i = t * Nx + x;
            double local = 0.0;
            for (int ss = 0; ss < 2; ++ss) {
                const double4* Y = P + (2 * k + ss) * V5;
                const double4 l = Y[i], r = Y[(long long)(Ls - 1) * V4 + i];
                local += r.x * r.x + r.y * r.y + l.z * l.z + l.w * l.w;
            }
            v += cos(p * (x - x0)) * local;
        }
        v = bsum(v, sh);
        if (threadIdx.x == 0) out[((long long)k * nmom + n) * ntmax + t] = v;
    }
}

// Loop density (correlator.hpp:83-107): L(x) = (1/N) sum_noise eta(x)^dag sigma3 q(x);
// the noise solutions start at right-hand side `first` of each configuration, the
// output is loop[cfg][V4max] complex.
__global__ void k_contract_loop(const VecLayout lay, const int nnoise, const int K, const int first, const int v4max,
                                const double* noise_all, const long long* noff, const double* X, double* loop_all)
{
    const int cfg = blockIdx.y;
    const CfgGeo G = cfg_geo(lay, K, cfg);
    const int V4 = G.V4, Ls = lay.Ls;
    const long long V5 = G.V5;
    const double4* P = reinterpret_cast<const double4*>(X) + G.site + (long long)first * V5;
    const double* noise = noise_all + noff[cfg];
    double* loop = loop_all + (long long)cfg * v4max * 2;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V4; i += gridDim.x * blockDim.x) {
        double lr = 0.0, li = 0.0;
        for (int n = 0; n < nnoise; ++n) {
            const double* e = noise + ((long long)n * V4 + i) * 4;
            const double4* Y = P + n * V5;
            const double4 l = Y[i], r = Y[(long long)(Ls - 1) * V4 + i];
            // conj(e0) q0 - conj(e1) q1
            lr += e[0] * r.x + e[1] * r.y - (e[2] * l.z + e[3] * l.w);
            li += e[0] * r.y - e[1] * r.x - (e[2] * l.w - e[3] * l.z);
        }
        loop[2 * i] = lr / nnoise;
        loop[2 * i + 1] = li / nnoise;
    }
}

// ||v_rhs||^2 per right-hand side (block per rhs, fixed order) of a batch laid out as lay.
__global__ void k_norms(const double* v, const VecLayout lay, double* out)
{
    __shared__ double sh[8];
    const int rhs = blockIdx.x;
    long long n, off;
    lay.locate(rhs, n, off);
    const double* a = v + off * 4;
    const long long n_per = n * 4;
    double s = 0.0;
    for (long long i = threadIdx.x; i < n_per; i += blockDim.x) s += a[i] * a[i];
    s = bsum(s, sh);
    if (threadIdx.x == 0) out[rhs] = s;
}

// d = a - b elementwise
__global__ void k_sub(const double* a, const double* b, double* d, const long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] = a[i] - b[i];
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

namespace {

int measure_impl(dwf_ctx* c, const dwf_measure_spec* spec, int ncfg, const unsigned long long* seeds,
                 dwf_measure_out* outs)
{
    if (!c || !spec || !outs) throw DwfError(DWF_ERR_ARG, "dwf_measure: null argument");
    if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "dwf_measure: no gauge field set");
    if (c->kind != 0) throw DwfError(DWF_ERR_ARG, "dwf_measure: needs a domain-wall context");
    if (ncfg != c->ncfg) throw DwfError(DWF_ERR_ARG, "dwf_measure: ncfg differs from the configurations held");
    // per configuration lattice (configuration-major over the lattice groups)
    std::vector<int> cNt(ncfg), cNx(ncfg);
    if (c->multi()) {
        int k = 0;
        for (const auto& d : c->gdefs)
            for (int j = 0; j < d.ncfg; ++j, ++k) {
                cNt[k] = d.Nt;
                cNx[k] = d.Nx;
            }
    } else {
        std::fill(cNt.begin(), cNt.end(), c->L.Nt);
        std::fill(cNx.begin(), cNx.end(), c->L.Nx);
    }
    int ntmax = 0, v4max = 0, nxmin = 1 << 30;
    for (int k = 0; k < ncfg; ++k) {
        ntmax = std::max(ntmax, cNt[k]);
        v4max = std::max(v4max, cNt[k] * cNx[k]);
        nxmin = std::min(nxmin, cNx[k]);
    }
    const int nwall = spec->wall ? 4 : 0;
    const int* hsite = spec->point ? spec->point_site : nullptr;
    const int nsrc_t = hsite ? 1 : 2;  // point source times per configuration
    const int npoint = spec->point ? 2 * nsrc_t : 0;
    const int nnoise = spec->noise_vectors;
    if (nnoise < 0) throw DwfError(DWF_ERR_ARG, "eta.noise_vectors must be positive");
    if (spec->point && (spec->max_momentum < 0 || spec->max_momentum > nxmin / 2))
        throw DwfError(DWF_ERR_ARG, "analysis.max_momentum is outside 0..Nx/2");
    if (hsite)
        for (int k = 0; k < ncfg; ++k)
            if (hsite[2 * k] < 0 || hsite[2 * k] >= cNt[k] || hsite[2 * k + 1] < 0 || hsite[2 * k + 1] >= cNx[k])
                throw DwfError(DWF_ERR_ARG, "dwf_measure: point_site outside the lattice");
    for (int k = 0; k < ncfg; ++k) {
        const dwf_measure_out& o = outs[k];
        if ((spec->wall && (!o.CPP || !o.CJ5)) || (spec->point && !o.momentum_CPP) || (nnoise > 0 && !o.loop))
            throw DwfError(DWF_ERR_ARG, "dwf_measure: missing output buffer");
    }
    if (spec->maxiter < 0 || (spec->method != DWF_SOLVE_CG64 && spec->method != DWF_SOLVE_MIXED &&
                              spec->method != DWF_SOLVE_EO_MIXED && spec->method != DWF_SOLVE_LEARNED))
        throw DwfError(DWF_ERR_ARG, "dwf_measure: bad argument");
    const int K = nwall + npoint + nnoise;  // right-hand sides per configuration
    for (int k = 0; k < ncfg; ++k) outs[k].wall_nonconverged = 0;
    if (K == 0) return DWF_OK;
    const int nrhs = K * ncfg;
    c->bind_rhs(nrhs);
    const VecLayout lay = c->layout(nrhs, 0);
    const size_t nvt = (size_t)lay.total_sites() * 4;  // doubles of the whole batch
    hipStream_t s = c->stream;

    // Z4 noise from the host stream of correlator.hpp:35-58, one engine per configuration;
    // configuration k's vectors start at noff[k] doubles
    double* dnoise = nullptr;
    long long* dnoff = nullptr;
    if (nnoise > 0) {
        std::vector<long long> noff(ncfg);
        long long tot = 0;
        for (int k = 0; k < ncfg; ++k) {
            noff[k] = tot;
            tot += (long long)nnoise * cNt[k] * cNx[k] * 4;
        }
        std::vector<double> h((size_t)tot);
        for (int k = 0; k < ncfg; ++k)
            dwf_z4_noise(seeds ? seeds[k] : spec->noise_seed, nnoise, cNt[k] * cNx[k] * 2, h.data() + noff[k]);
        dnoise = (double*)c->buf(B_AUX3, h.size() * 8);
        dnoff = (long long*)c->buf(B_NOFF, (size_t)ncfg * sizeof(long long));
        c->check(hipMemcpyAsync(dnoise, h.data(), h.size() * 8, hipMemcpyHostToDevice, s), "noise upload");
        c->check(hipMemcpyAsync(dnoff, noff.data(), (size_t)ncfg * sizeof(long long), hipMemcpyHostToDevice, s),
                 "noise upload");
        c->check(hipStreamSynchronize(s), "noise upload");
    }
    int* dsite = nullptr;
    if (hsite) {
        dsite = (int*)c->buf(B_AUX4, (size_t)ncfg * 2 * sizeof(int));
        c->check(hipMemcpyAsync(dsite, hsite, (size_t)ncfg * 2 * sizeof(int), hipMemcpyHostToDevice, s), "site upload");
        c->check(hipStreamSynchronize(s), "site upload");
    }
    double* src = (double*)c->buf(B_AUX0, nvt * 8);
    double* X = (double*)c->buf(B_AUX2, nvt * 8);
    k_sources<<<dim3((v4max + 255) / 256, nrhs), 256, 0, s>>>(lay, nwall, npoint, nnoise, dnoise, dnoff, dsite, src);
    c->check_launch("sources");
    std::vector<dwf_solve_stats> stv(nrhs);
    const int rc = dwf_solve_dirac_impl(c, spec->method, src, nullptr, spec->rtol, spec->maxiter, nrhs, X, stv.data());
    if (rc == DWF_ERR_ZERO_DENOM) throw DwfError(DWF_ERR_ZERO_DENOM, "CG encountered a zero denominator");
    for (int k = 0; k < ncfg; ++k) {
        for (int j = 0; j < K; ++j) {
            const dwf_solve_stats& st = stv[(size_t)k * K + j];
            if (outs[k].iterations) outs[k].iterations[j] = st.iterations + (spec->method == DWF_SOLVE_CG64 ? 0 : st.fp64_iterations);
            if (j >= nwall && j < nwall + npoint && !st.converged)
                throw DwfError(DWF_ERR_NOT_CONVERGED, "Momentum propagator CGNR failed to converge");
            if (j >= nwall + npoint && !st.converged)
                throw DwfError(DWF_ERR_NOT_CONVERGED, "Eta loop CGNR failed to converge");
            if (j < nwall && !st.converged) outs[k].wall_nonconverged += 1;  // "CGNR info = nonzero" (:222-223)
        }
    }
    if (nwall) {
        // Dirac residual ||D x - src|| / ||src|| of the wall solves (:225-227)
        double* dx = (double*)c->buf(B_APPLY_TMP, nvt * 8);
        double* diff = (double*)c->buf(B_T64, nvt * 8);
        dwf_apply_device_impl(c, DWF_OP_D, DWF_PREC_FP64, X, dx, nrhs);
        k_sub<<<1024, 256, 0, s>>>(dx, src, diff, (long long)nvt);
        double* nd = (double*)c->buf(B_AUX1, (size_t)2 * nrhs * 8);
        k_norms<<<nrhs, 256, 0, s>>>(diff, lay, nd);
        k_norms<<<nrhs, 256, 0, s>>>(src, lay, nd + nrhs);
        double* pp = (double*)c->buf(B_PA, (size_t)2 * ntmax * ncfg * 8);
        double* j5 = (double*)c->buf(B_PB, (size_t)2 * ntmax * ncfg * 8);
        k_contract_wall<<<dim3(ntmax, 2, ncfg), 256, 0, s>>>(lay, K, ntmax, X, pp, j5);
        c->check_launch("wall contraction");
        std::vector<double> hpp((size_t)2 * ntmax * ncfg), hj5((size_t)2 * ntmax * ncfg), hn((size_t)2 * nrhs);
        c->check(hipMemcpyAsync(hpp.data(), pp, hpp.size() * 8, hipMemcpyDeviceToHost, s), "readback");
        c->check(hipMemcpyAsync(hj5.data(), j5, hj5.size() * 8, hipMemcpyDeviceToHost, s), "readback");
        c->check(hipMemcpyAsync(hn.data(), nd, hn.size() * 8, hipMemcpyDeviceToHost, s), "readback");
        c->check(hipStreamSynchronize(s), "readback");
        for (int k = 0; k < ncfg; ++k) {
            const int Nt = cNt[k];
            const double* a = hpp.data() + (size_t)k * 2 * ntmax;
            const double* b = hj5.data() + (size_t)k * 2 * ntmax;
            for (int dt = 0; dt < Nt; ++dt) {
                outs[k].CPP[dt] = (a[dt] + a[ntmax + (dt + Nt / 2) % Nt]) / 2.0;
                outs[k].CJ5[dt] = (b[dt] + b[ntmax + (dt + Nt / 2) % Nt]) / 2.0;
            }
            if (outs[k].dirac_residual)
                for (int j = 0; j < nwall; ++j)
                    outs[k].dirac_residual[j] = std::sqrt(hn[(size_t)k * K + j]) / std::sqrt(hn[nrhs + (size_t)k * K + j]);
        }
    }
    if (npoint) {
        const int nmom = spec->max_momentum + 1;
        double* pm = (double*)c->buf(B_PR, (size_t)nsrc_t * nmom * ntmax * ncfg * 8);
        double* pj = (double*)c->buf(B_PA, (size_t)nsrc_t * ntmax * ncfg * 8);
        k_contract_point<<<dim3(ntmax, nsrc_t, ncfg), 256, 0, s>>>(lay, nmom, K, nwall, nsrc_t, ntmax, dsite, X, pm, pj);
        c->check_launch("point contraction");
        std::vector<double> h((size_t)nsrc_t * nmom * ntmax * ncfg), hj((size_t)nsrc_t * ntmax * ncfg);
        c->check(hipMemcpyAsync(h.data(), pm, h.size() * 8, hipMemcpyDeviceToHost, s), "readback");
        c->check(hipMemcpyAsync(hj.data(), pj, hj.size() * 8, hipMemcpyDeviceToHost, s), "readback");
        c->check(hipStreamSynchronize(s), "readback");
        for (int k = 0; k < ncfg; ++k) {
            const int Nt = cNt[k];
            const double* a = h.data() + (size_t)k * nsrc_t * nmom * ntmax;
            const double* b = hj.data() + (size_t)k * nsrc_t * ntmax;
            // relative to the source time: folded over t0 = 0, Nt/2 (:158-174) or a single t0
            for (int dt = 0; dt < Nt; ++dt) {
                if (!hsite) {
                    for (int n = 0; n < nmom; ++n)
                        outs[k].momentum_CPP[n * Nt + dt] =
                            (a[n * ntmax + dt] + a[(nmom + n) * ntmax + (dt + Nt / 2) % Nt]) / 2.0;
                    if (outs[k].point_CJ5) outs[k].point_CJ5[dt] = (b[dt] + b[ntmax + (dt + Nt / 2) % Nt]) / 2.0;
                } else {
                    const int t = (hsite[2 * k] + dt) % Nt;
                    for (int n = 0; n < nmom; ++n) outs[k].momentum_CPP[n * Nt + dt] = a[n * ntmax + t];
                    if (outs[k].point_CJ5) outs[k].point_CJ5[dt] = b[t];
                }
            }
        }
    }
    if (nnoise) {
        double* lp = (double*)c->buf(B_PR, (size_t)v4max * 2 * ncfg * 8);
        k_contract_loop<<<dim3((v4max + 255) / 256, ncfg), 256, 0, s>>>(lay, nnoise, K, nwall + npoint, v4max, dnoise,
                                                                        dnoff, X, lp);
        c->check_launch("loop contraction");
        for (int k = 0; k < ncfg; ++k)
            c->check(hipMemcpyAsync(outs[k].loop, lp + (size_t)k * v4max * 2, (size_t)cNt[k] * cNx[k] * 2 * 8,
                                    hipMemcpyDeviceToHost, s),
                     "readback");
        c->check(hipStreamSynchronize(s), "readback");
    }
    return DWF_OK;
}

}  // namespace

extern "C" int dwf_measure(dwf_ctx* c, const dwf_measure_spec* spec, dwf_measure_out* out)
{
    return guard([&]() -> int {
        if (c && c->ncfg != 1) throw DwfError(DWF_ERR_ARG, "dwf_measure: the context holds several configurations; use dwf_measure_batch");
        return measure_impl(c, spec, 1, nullptr, out);
    });
}

extern "C" int dwf_measure_batch(dwf_ctx* c, const dwf_measure_spec* spec, int ncfg, const unsigned long long* seeds,
                                 dwf_measure_out* outs)
{
    return guard([&]() -> int { return measure_impl(c, spec, ncfg, seeds, outs); });
}
