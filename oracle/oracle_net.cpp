// oracle_net.cpp -- TEST INFRASTRUCTURE, not part of the product.
//
// The oracle of the learned preconditioner: a plain C++17 restatement, in FP64, of the
// gauge-equivariant parallel-transport convolution network that libdwfhip runs with
// BF16 matrix-core kernels, and of its loss gradient.  The reference has no learned
// component (it solves with unpreconditioned CG, domain_wall/include/cg_solver.hpp:17-64);
// the network is this build's own design, so there is no reference file to follow and
// the oracle is pinned by mathematics instead (tests/test_oracle_net.py): gauge
// covariance, linearity, the adjoint identity <u, M v> = <M^dag u, v>, central finite
// differences of the loss against the analytic gradient, and a dense matrix of one layer
// built independently in numpy.
//
// Network.  Three linear layers with channel counts 2 -> 8 -> 8 -> 2 (a 5D spinor has 2
// complex components per site).  Layer l maps a field h with C_in complex channels per
// 5D site to
//     y(s,t,x) = sum_p W_p(s) [U_p(t,x) h(neighbour_p(s,t,x))]
// over the 7 stencil points p = self, +t, -t, +x, -x, +s, -s.  U_p is the parallel
// transporter of the Dirac operator (dwf_operator.hpp:15-92 conventions: U_t(t,x) =
// exp(i theta[0][t][x]), U_x = exp(i theta[1][t][x]), backward hops conj(U) of the
// neighbour, sign -1 on hops across the antiperiodic time boundary); +-s hops are
// periodic in s and carry no link.  W_p(s) is a complex C_out x C_in matrix that depends
// on s (the fifth dimension is not translation invariant: domain walls) but not on
// (t, x).  No bias and no activation: the network is linear, hence exactly
// scale-equivariant, and gauge covariant.
//
// Weights: per layer [Ls][7][C_out][C_in] complex, layers concatenated, (re, im) pairs.
//
// Quantized mode (bf16 = 1) rounds at exactly the points the device kernels round: the
// weights to bf16; stored hidden activations and gradients to bf16; in layers with two
// input channels (and in the weight gradients) the layer input to float and then bf16,
// the transporter to float and then bf16, and the transported value (exact bf16 products,
// one rounding of their sum to float) to bf16; in layers with eight input channels the
// transporter to float, applied without rounding to the product of the bf16 input.  The
// oracle accumulates in FP64 where the device accumulates in FP32, so device and oracle
// agree to accumulation-order error.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using cplx = std::complex<double>;

constexpr int kPoints = 7;
constexpr int kCin[3] = {2, 8, 8};
constexpr int kCout[3] = {8, 8, 2};

struct Geo {
    int Nt, Nx, Ls;
    int V4() const { return Nt * Nx; }
    long long V5() const { return (long long)Nt * Nx * Ls; }
};

long long layer_weights(int l, int Ls) { return (long long)Ls * kPoints * kCout[l] * kCin[l]; }
long long layer_offset(int l, int Ls)
{
    long long o = 0;
    for (int i = 0; i < l; ++i) o += layer_weights(i, Ls);
    return o;
}

// float -> bf16 -> float, round to nearest even (finite inputs)
float bf16_round(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// Neighbour site index and transporter of stencil point p at (s, t, x).
struct Hop {
    long long site;  // s * V4 + t * Nx + x of the neighbour
    cplx u;          // exact transporter (FP64)
    float cf, sf;    // the device's FP32 transporter (cos, sin with the boundary sign)
};

Hop hop(const Geo& g, const double* th, int p, int s, int t, int x)
{
    const int V4 = g.V4();
    int s2 = s, t2 = t, x2 = x;
    double sign = 1.0;
    bool link = false, conj = false;
    int mu = 0, lt = t, lx = x;
    switch (p) {
        case 1:
            t2 = (t + 1) % g.Nt;
            link = true;
            mu = 0;
            sign = t == g.Nt - 1 ? -1.0 : 1.0;
            break;
        case 2:
            t2 = (t - 1 + g.Nt) % g.Nt;
            link = true;
            conj = true;
            mu = 0;
            lt = t2;
            sign = t == 0 ? -1.0 : 1.0;
            break;
        case 3:
            x2 = (x + 1) % g.Nx;
            link = true;
            mu = 1;
            break;
        case 4:
            x2 = (x - 1 + g.Nx) % g.Nx;
            link = true;
            conj = true;
            mu = 1;
            lx = x2;
            break;
        case 5: s2 = (s + 1) % g.Ls; break;
        case 6: s2 = (s - 1 + g.Ls) % g.Ls; break;
        default: break;
    }
    Hop h;
    h.site = (long long)s2 * V4 + (long long)t2 * g.Nx + x2;
    if (!link) {
        h.u = 1.0;
        h.cf = 1.0f;
        h.sf = 0.0f;
        return h;
    }
    const double ang = th[(mu * g.Nt + lt) * g.Nx + lx];
    const double c = std::cos(ang), sn = conj ? -std::sin(ang) : std::sin(ang);
    h.u = sign * cplx(c, sn);
    // the device stores (float)cos, (float)sin and applies sign and conjugation exactly
    h.cf = (float)sign * (float)c;
    h.sf = (float)sign * (conj ? -(float)std::sin(ang) : (float)std::sin(ang));
    return h;
}

// Transported value U v as the device computes it with bf16 dot products: v rounded to
// float then bf16, the transporter's float (cos, sin) rounded to bf16, the two products of
// each component exact, their sum rounded once to float, then to bf16.
cplx transport_q(const Hop& h, cplx v)
{
    const double hr = bf16_round((float)v.real()), hi = bf16_round((float)v.imag());
    const double c = bf16_round(h.cf), sn = bf16_round(h.sf);
    const float ar = (float)(c * hr - sn * hi);
    const float ai = (float)(sn * hr + c * hi);
    return cplx(bf16_round(ar), bf16_round(ai));
}

cplx qw(cplx w) { return cplx(bf16_round((float)w.real()), bf16_round((float)w.imag())); }
cplx q16(cplx v) { return cplx(bf16_round((float)v.real()), bf16_round((float)v.imag())); }

const cplx* wptr(const double* w, int l, int Ls, int s, int p)
{
    return reinterpret_cast<const cplx*>(w) + layer_offset(l, Ls) +
           ((long long)s * kPoints + p) * kCout[l] * kCin[l];
}

// One layer forward (adj = false) or its adjoint (adj = true: maps C_out channels back
// to C_in).  in/out: [nrhs][V5][C] complex.  quantize the output to bf16 if `store16`.
void layer(const Geo& g, const double* th, const double* w, int l, bool adj, const cplx* in, cplx* out, int nrhs,
           bool q, bool store16)
{
    const int ci = adj ? kCout[l] : kCin[l], co = adj ? kCin[l] : kCout[l];
    const long long V5 = g.V5();
    const int neg[kPoints] = {0, 2, 1, 4, 3, 6, 5};
#pragma omp parallel for collapse(2) schedule(static)
    for (int r = 0; r < nrhs; ++r)
        for (long long i = 0; i < V5; ++i) {
            const int s = (int)(i / g.V4()), rem = (int)(i % g.V4());
            const int t = rem / g.Nx, x = rem % g.Nx;
            std::vector<cplx> acc(co, 0.0);
            for (int p = 0; p < kPoints; ++p) {
                const Hop h = hop(g, th, p, s, t, x);
                const cplx* hn = in + ((long long)r * V5 + h.site) * ci;
                // weights: forward W_p(s); adjoint W_{-p}(s of the neighbour)^H
                const int sw = adj ? (int)(h.site / g.V4()) : s;
                const cplx* W = wptr(w, l, g.Ls, sw, adj ? neg[p] : p);
                for (int c = 0; c < ci; ++c) {
                    // quantized: C_in = 2 layers transport the bf16 input in bf16 before the
                    // product; C_in = 8 layers multiply the product by the FP32 transporter
                    const cplx a = !q ? h.u * hn[c] : (ci == 2 ? transport_q(h, hn[c]) : cplx(h.cf, h.sf) * q16(hn[c]));
                    for (int o = 0; o < co; ++o) {
                        cplx wv = adj ? std::conj(W[c * kCin[l] + o]) : W[o * kCin[l] + c];
                        if (q) wv = qw(wv);
                        acc[o] += wv * a;
                    }
                }
            }
            cplx* y = out + ((long long)r * V5 + i) * co;
            for (int o = 0; o < co; ++o) {
                cplx v = acc[o];
                if (q) v = cplx((float)v.real(), (float)v.imag());
                y[o] = (q && store16) ? q16(v) : v;
            }
        }
}

// Weight gradient of layer l: G_p(s)[o][c] += sum_x gy(x)[o] conj(a_p(x)[c]), a_p the
// transported layer input.  Quantized: a and gy rounded to bf16 as the device's matrix
// operands are.
void layer_wgrad(const Geo& g, const double* th, int l, const cplx* in, const cplx* gy, int nrhs, bool q, cplx* G)
{
    const int ci = kCin[l], co = kCout[l];
    const long long V5 = g.V5();
    const long long nper = (long long)kPoints * co * ci;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < g.Ls; ++s) {
        cplx* Gs = G + (long long)s * nper;
        for (int r = 0; r < nrhs; ++r)
            for (int rem = 0; rem < g.V4(); ++rem) {
                const long long i = (long long)s * g.V4() + rem;
                const int t = rem / g.Nx, x = rem % g.Nx;
                const cplx* gv = gy + ((long long)r * V5 + i) * co;
                for (int p = 0; p < kPoints; ++p) {
                    const Hop h = hop(g, th, p, s, t, x);
                    const cplx* hn = in + ((long long)r * V5 + h.site) * ci;
                    for (int c = 0; c < ci; ++c) {
                        const cplx a = q ? transport_q(h, hn[c]) : h.u * hn[c];
                        for (int o = 0; o < co; ++o) {
                            const cplx gq = q ? q16(gv[o]) : gv[o];
                            Gs[(long long)p * co * ci + o * ci + c] += gq * std::conj(a);
                        }
                    }
                }
            }
    }
}

struct Net {
    Geo g;
    const double* th;
    const double* w;
    bool q;
};

// z = M v for nrhs spinors ([nrhs][V5][2]); keeps the hidden activations if asked.
void forward(const Net& n, const cplx* v, cplx* z, int nrhs, std::vector<cplx>* h0, std::vector<cplx>* h1,
             std::vector<cplx>* h2)
{
    const long long V5 = n.g.V5();
    std::vector<cplx> in0(v, v + (size_t)nrhs * V5 * 2);
    if (n.q)  // the device reads its input in FP32 (training data) or casts FP64 to FP32
        for (cplx& c : in0) c = cplx((float)c.real(), (float)c.imag());
    std::vector<cplx> a1((size_t)nrhs * V5 * 8), a2((size_t)nrhs * V5 * 8);
    layer(n.g, n.th, n.w, 0, false, in0.data(), a1.data(), nrhs, n.q, true);
    layer(n.g, n.th, n.w, 1, false, a1.data(), a2.data(), nrhs, n.q, true);
    layer(n.g, n.th, n.w, 2, false, a2.data(), z, nrhs, n.q, false);
    if (h0) *h0 = std::move(in0);
    if (h1) *h1 = std::move(a1);
    if (h2) *h2 = std::move(a2);
}

}  // namespace

extern "C" {

long long orc_net_num_weights(int Ls) { return 2 * layer_offset(3, Ls); }

// z = M v (adj = 0) or M^dag v (adj = 1); theta [2][Nt][Nx]; v, z [nrhs][Ls][Nt][Nx][2]
// complex; w the weights (real pairs).  bf16 = 1 rounds as the device kernels do.
int orc_net_apply(int Nt, int Nx, int Ls, const double* theta, const double* w, const double* v, double* z, int nrhs,
                  int adj, int bf16)
{
    if (Nt < 2 || Nx < 2 || Ls < 2 || nrhs < 1 || !theta || !w || !v || !z) return 1;
    const Net n{Geo{Nt, Nx, Ls}, theta, w, bf16 != 0};
    const cplx* vin = reinterpret_cast<const cplx*>(v);
    cplx* out = reinterpret_cast<cplx*>(z);
    if (!adj) {
        forward(n, vin, out, nrhs, nullptr, nullptr, nullptr);
        return 0;
    }
    const long long V5 = n.g.V5();
    std::vector<cplx> a2((size_t)nrhs * V5 * 8), a1((size_t)nrhs * V5 * 8);
    layer(n.g, theta, w, 2, true, vin, a2.data(), nrhs, n.q, true);
    layer(n.g, theta, w, 1, true, a2.data(), a1.data(), nrhs, n.q, true);
    layer(n.g, theta, w, 0, true, a1.data(), out, nrhs, n.q, false);
    return 0;
}

// Loss and its gradient with respect to the weights (dL/dRe w, dL/dIm w in the weight
// layout).  loss 0: L = (1/npairs) sum_i ||M b_i - x_i||^2 / ||x_i||^2;
// loss 1: L = (1/npairs) sum_i ||A M b_i - b_i||^2 / ||b_i||^2 with A = D^dag D (the
// oracle's domain-wall operator, parameters M5 and mf; x unused).  Quantized mode
// follows the device's training step: y and the output gradient in FP32, bf16 matrix
// operands, bf16 stored activations and activation gradients.
int orc_dwf_apply(int op, int Nt, int Nx, int Ls, double M5, double mf, const double* theta, const double* in,
                  double* out);

int orc_net_loss_grad(int Nt, int Nx, int Ls, double M5, double mf, const double* theta, const double* w,
                      const double* b, const double* x, int npairs, int loss_kind, int bf16, double* loss,
                      double* grad)
{
    if (Nt < 2 || Nx < 2 || Ls < 2 || npairs < 1 || !theta || !w || !b || (loss_kind == 0 && !x) || !loss ||
        !grad || (loss_kind != 0 && loss_kind != 1))
        return 1;
    const Net n{Geo{Nt, Nx, Ls}, theta, w, bf16 != 0};
    const long long V5 = n.g.V5();
    const size_t nv = (size_t)V5 * 2;
    std::vector<cplx> y(nv * npairs), h0, h1, h2;
    forward(n, reinterpret_cast<const cplx*>(b), y.data(), npairs, &h0, &h1, &h2);
    const cplx* bs = reinterpret_cast<const cplx*>(b);
    const cplx* ts = loss_kind == 0 ? reinterpret_cast<const cplx*>(x) : bs;  // the target's norm
    auto rd = [&](cplx v) { return n.q ? cplx((float)v.real(), (float)v.imag()) : v; };
    std::vector<cplx> gy(nv * npairs), u(nv), ad(nv);
    double L = 0.0;
    for (int r = 0; r < npairs; ++r) {
        double tn = 0.0, dn = 0.0;
        for (size_t i = 0; i < nv; ++i) tn += std::norm(rd(ts[r * nv + i]));
        const double scale = 2.0 / ((double)npairs * tn);
        cplx* g = gy.data() + r * nv;
        if (loss_kind == 0) {
            for (size_t i = 0; i < nv; ++i) {
                const cplx d = rd(y[r * nv + i]) - rd(ts[r * nv + i]);
                dn += std::norm(d);
                g[i] = scale * d;
            }
        } else {
            orc_dwf_apply(3, Nt, Nx, Ls, M5, mf, theta, reinterpret_cast<const double*>(y.data() + r * nv),
                          reinterpret_cast<double*>(u.data()));
            for (size_t i = 0; i < nv; ++i) {
                u[i] = rd(u[i]) - rd(bs[r * nv + i]);
                dn += std::norm(u[i]);
            }
            orc_dwf_apply(3, Nt, Nx, Ls, M5, mf, theta, reinterpret_cast<const double*>(u.data()),
                          reinterpret_cast<double*>(ad.data()));
            for (size_t i = 0; i < nv; ++i) g[i] = scale * ad[i];
        }
        if (n.q)
            for (size_t i = 0; i < nv; ++i) g[i] = rd(g[i]);
        L += dn / tn / npairs;
    }
    *loss = L;
    // backward: g2 = L3^dag gy, g1 = L2^dag g2 (stored bf16 in quantized mode)
    std::vector<cplx> g2(nv * 4 * npairs), g1(nv * 4 * npairs);
    layer(n.g, theta, w, 2, true, gy.data(), g2.data(), npairs, n.q, true);
    layer(n.g, theta, w, 1, true, g2.data(), g1.data(), npairs, n.q, true);
    std::vector<cplx> G((size_t)layer_offset(3, Ls), 0.0);
    layer_wgrad(n.g, theta, 2, h2.data(), gy.data(), npairs, n.q, G.data() + layer_offset(2, Ls));
    layer_wgrad(n.g, theta, 1, h1.data(), g2.data(), npairs, n.q, G.data() + layer_offset(1, Ls));
    layer_wgrad(n.g, theta, 0, h0.data(), g1.data(), npairs, n.q, G.data() + layer_offset(0, Ls));
    for (size_t i = 0; i < G.size(); ++i) {
        grad[2 * i] = G[i].real();
        grad[2 * i + 1] = G[i].imag();
    }
    return 0;
}

}  // extern "C"
