// oracle.cpp -- TEST INFRASTRUCTURE, not part of the product.
//
// This is the oracle: a plain C++17 (std::complex<double>) CPU restatement of the
// FP64 fermion path of wurdahl/Three_Part_DWF_Study.  Only the test suite, the smoke
// check and the CPU-baseline leg of bench.py load it, and only as the checker (or as the
// timed CPU baseline).  libdwfhip.so never links, loads or calls it.
//
// Every function cites the reference file:line it restates.  The loop order is the
// reference's: s outermost, std::exp of each link evaluated per slice, results built
// out of place, unpreconditioned CG -- so timing it times the reference's algorithm.
//
// The reference itself cannot be compiled here: it needs Eigen 3 (README.md:23,
// compile.sh:9), which is neither vendored nor installed.  Eigen's role is restated:
//   * VectorXcd arithmetic, .dot (conjugates its first argument), .norm, .squaredNorm;
//   * Eigen::ConjugateGradient<SparseMatrix, Lower|Upper> with its default
//     DiagonalPreconditioner and zero initial guess (Eigen 3 public algorithm; used only
//     by the Wilson solves, wilson/include/cg_solver.hpp:19-78).
// Reduction order inside Eigen is not reproduced bit for bit (parity unpinned at the
// last ulp); everything else is pinned by the tests listed in the oracle test files:
// dense brute-force matrices built independently from sigma-matrix Kronecker products,
// plane-wave closed forms, adjointness, gamma5-R hermiticity, gauge covariance,
// dense linear solves, central finite differences of the actions, and the C++
// standard's fixed mt19937_64 output.
//
// Random numbers come from libstdc++ <random> in the reference's order, compiled by g++
// exactly as the reference is (compile.sh:8-15), so seeded streams match.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <omp.h>
#include <random>
#include <vector>

namespace {

using cplx = std::complex<double>;
using vec = std::vector<cplx>;

constexpr double kPi = 3.141592653589793238462643383279502884;

// ---------------------------------------------------------------- conventions
// domain_wall/include/indexing.hpp:5-23, spin.hpp:5-42

struct Mat2 {
    cplx m[2][2];
    cplx operator()(int a, int b) const { return m[a][b]; }
};

Mat2 mat(cplx a, cplx b, cplx c, cplx d) { return Mat2{{{a, b}, {c, d}}}; }
Mat2 identity2() { return mat(1.0, 0.0, 0.0, 1.0); }
Mat2 sigma1() { return mat(0.0, 1.0, 1.0, 0.0); }
Mat2 sigma2() { return mat(0.0, cplx(0.0, -1.0), cplx(0.0, 1.0), 0.0); }
Mat2 sigma3() { return mat(1.0, 0.0, 0.0, -1.0); }
Mat2 add(const Mat2& A, const Mat2& B, double sb)
{
    Mat2 r;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) r.m[i][j] = A.m[i][j] + sb * B.m[i][j];
    return r;
}
Mat2 scale(const Mat2& A, double f)
{
    Mat2 r;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) r.m[i][j] = f * A.m[i][j];
    return r;
}
// P_R = (1+sigma3)/2, P_L = (1-sigma3)/2  (spin.hpp:32-42)
Mat2 projector_R() { return scale(add(identity2(), sigma3(), 1.0), 0.5); }
Mat2 projector_L() { return scale(add(identity2(), sigma3(), -1.0), 0.5); }

struct V2 {
    cplx v[2];
};
V2 mul(const Mat2& A, const V2& x)
{
    V2 r;
    for (int a = 0; a < 2; ++a) r.v[a] = A.m[a][0] * x.v[0] + A.m[a][1] * x.v[1];
    return r;
}

struct Dwf {
    int Nt, Nx, Ls;
    double M5, mf;
    int ndof() const { return Ls * Nt * Nx * 2; }
    int ngauge() const { return 2 * Nt * Nx; }
    int fidx(int s, int t, int x, int a) const { return ((s * Nt + t) * Nx + x) * 2 + a; }
    int gidx(int mu, int t, int x) const { return (mu * Nt + t) * Nx + x; }
};

inline int plus_periodic(int x, int n) { return (x + 1) % n; }
inline int minus_periodic(int x, int n) { return (x - 1 + n) % n; }

// dwf_operator.hpp:10-13
inline cplx link(const Dwf& g, const double* th, int mu, int t, int x)
{
    return std::exp(cplx(0.0, th[g.gidx(mu, t, x)]));
}

// Eigen-style BLAS-1 on std::vector (dot conjugates its first argument).
cplx dot(const vec& a, const vec& b)
{
    cplx s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
    return s;
}
double squared_norm(const vec& a)
{
    double s = 0.0;
    for (const cplx& v : a) s += std::norm(v);
    return s;
}
double norm(const vec& a) { return std::sqrt(squared_norm(a)); }

// ------------------------------------------------------------- DWF operator
// dwf_operator.hpp:15-92
vec dwf_apply_D(const Dwf& g, const double* th, const vec& psi)
{
    vec out(psi.size());
    for (size_t i = 0; i < psi.size(); ++i) out[i] = (3.0 - g.M5) * psi[i];

    const Mat2 I = identity2(), s1 = sigma1(), s2 = sigma2();
    const Mat2 ImS1 = add(I, s1, -1.0), IpS1 = add(I, s1, 1.0);
    const Mat2 ImS2 = add(I, s2, -1.0), IpS2 = add(I, s2, 1.0);

    for (int s = 0; s < g.Ls; ++s)
        for (int t = 0; t < g.Nt; ++t) {
            const int tp = plus_periodic(t, g.Nt), tm = minus_periodic(t, g.Nt);
            const double sign_tp = (t == g.Nt - 1) ? -1.0 : 1.0;
            const double sign_tm = (t == 0) ? -1.0 : 1.0;
            for (int x = 0; x < g.Nx; ++x) {
                const int xp = plus_periodic(x, g.Nx), xm = minus_periodic(x, g.Nx);
                V2 v_tp, v_tm, v_xp, v_xm;
                for (int a = 0; a < 2; ++a) {
                    v_tp.v[a] = sign_tp * psi[g.fidx(s, tp, x, a)];
                    v_tm.v[a] = sign_tm * psi[g.fidx(s, tm, x, a)];
                    v_xp.v[a] = psi[g.fidx(s, t, xp, a)];
                    v_xm.v[a] = psi[g.fidx(s, t, xm, a)];
                }
                const V2 term_tp = mul(ImS1, v_tp), term_tm = mul(IpS1, v_tm);
                const V2 term_xp = mul(ImS2, v_xp), term_xm = mul(IpS2, v_xm);
                const cplx Ut = link(g, th, 0, t, x);
                const cplx Ut_dag = std::conj(link(g, th, 0, tm, x));
                const cplx Ux = link(g, th, 1, t, x);
                const cplx Ux_dag = std::conj(link(g, th, 1, t, xm));
                for (int a = 0; a < 2; ++a) {
                    const int row = g.fidx(s, t, x, a);
                    out[row] += -0.5 * Ut * term_tp.v[a];
                    out[row] += -0.5 * Ut_dag * term_tm.v[a];
                    out[row] += -0.5 * Ux * term_xp.v[a];
                    out[row] += -0.5 * Ux_dag * term_xm.v[a];
                }
                // fifth-direction terms, dwf_operator.hpp:71-86
                const int row1 = g.fidx(s, t, x, 1);
                if (s < g.Ls - 1)
                    out[row1] += -psi[g.fidx(s + 1, t, x, 1)];
                else
                    out[row1] += g.mf * psi[g.fidx(0, t, x, 1)];
                const int row0 = g.fidx(s, t, x, 0);
                if (s > 0)
                    out[row0] += -psi[g.fidx(s - 1, t, x, 0)];
                else
                    out[row0] += g.mf * psi[g.fidx(g.Ls - 1, t, x, 0)];
            }
        }
    return out;
}

// dwf_operator.hpp:94-171
vec dwf_apply_Ddag(const Dwf& g, const double* th, const vec& psi)
{
    vec out(psi.size());
    for (size_t i = 0; i < psi.size(); ++i) out[i] = (3.0 - g.M5) * psi[i];

    const Mat2 I = identity2(), s1 = sigma1(), s2 = sigma2();
    const Mat2 ImS1 = add(I, s1, -1.0), IpS1 = add(I, s1, 1.0);
    const Mat2 ImS2 = add(I, s2, -1.0), IpS2 = add(I, s2, 1.0);

    for (int s = 0; s < g.Ls; ++s)
        for (int t = 0; t < g.Nt; ++t) {
            const int tp = plus_periodic(t, g.Nt), tm = minus_periodic(t, g.Nt);
            const double sign_tp = (t == g.Nt - 1) ? -1.0 : 1.0;
            const double sign_tm = (t == 0) ? -1.0 : 1.0;
            for (int x = 0; x < g.Nx; ++x) {
                const int xp = plus_periodic(x, g.Nx), xm = minus_periodic(x, g.Nx);
                V2 v_tm, v_tp, v_xm, v_xp;
                for (int a = 0; a < 2; ++a) {
                    v_tm.v[a] = sign_tm * psi[g.fidx(s, tm, x, a)];
                    v_tp.v[a] = sign_tp * psi[g.fidx(s, tp, x, a)];
                    v_xm.v[a] = psi[g.fidx(s, t, xm, a)];
                    v_xp.v[a] = psi[g.fidx(s, t, xp, a)];
                }
                const V2 term_tm = mul(ImS1, v_tm), term_tp = mul(IpS1, v_tp);
                const V2 term_xm = mul(ImS2, v_xm), term_xp = mul(IpS2, v_xp);
                const cplx Ut_dag = std::conj(link(g, th, 0, tm, x));
                const cplx Ut = link(g, th, 0, t, x);
                const cplx Ux_dag = std::conj(link(g, th, 1, t, xm));
                const cplx Ux = link(g, th, 1, t, x);
                for (int a = 0; a < 2; ++a) {
                    const int row = g.fidx(s, t, x, a);
                    out[row] += -0.5 * Ut_dag * term_tm.v[a];
                    out[row] += -0.5 * Ut * term_tp.v[a];
                    out[row] += -0.5 * Ux_dag * term_xm.v[a];
                    out[row] += -0.5 * Ux * term_xp.v[a];
                }
                // dwf_operator.hpp:150-165
                const int row1 = g.fidx(s, t, x, 1);
                if (s > 0)
                    out[row1] += -psi[g.fidx(s - 1, t, x, 1)];
                else
                    out[row1] += g.mf * psi[g.fidx(g.Ls - 1, t, x, 1)];
                const int row0 = g.fidx(s, t, x, 0);
                if (s < g.Ls - 1)
                    out[row0] += -psi[g.fidx(s + 1, t, x, 0)];
                else
                    out[row0] += g.mf * psi[g.fidx(0, t, x, 0)];
            }
        }
    return out;
}

// ------------------------------------------------------------------- DWF CG
// domain_wall/include/cg_solver.hpp:9-64
struct CgResult {
    vec x;
    int iterations;
    double relative_residual;
    bool converged;
    bool zero_denominator;
};

CgResult conjugate_gradient(const std::function<vec(const vec&)>& apply_A, const vec& b,
                            const vec* guess, double rtol, int maxiter)
{
    vec x = guess ? *guess : vec(b.size(), cplx(0.0, 0.0));
    vec Ax = apply_A(x);
    vec r(b.size());
    for (size_t i = 0; i < b.size(); ++i) r[i] = b[i] - Ax[i];
    vec p = r;
    const double bnorm = norm(b);
    if (bnorm == 0.0) return {x, 0, 0.0, true, false};
    double rr_old = std::real(dot(r, r));
    double relres = std::sqrt(rr_old) / bnorm;
    if (relres <= rtol) return {x, 0, relres, true, false};
    for (int it = 1; it <= maxiter; ++it) {
        const vec Ap = apply_A(p);
        const cplx den = dot(p, Ap);
        if (std::abs(den) == 0.0) return {x, it, relres, false, true};
        const double alpha = rr_old / std::real(den);
        for (size_t i = 0; i < x.size(); ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        const double rr_new = std::real(dot(r, r));
        relres = std::sqrt(rr_new) / bnorm;
        if (relres <= rtol) return {x, it, relres, true, false};
        const double beta_cg = rr_new / rr_old;
        for (size_t i = 0; i < p.size(); ++i) p[i] = r[i] + beta_cg * p[i];
        rr_old = rr_new;
    }
    return {x, maxiter, relres, false, false};
}

// ------------------------------------------------------- DWF pseudofermion force
// domain_wall/include/pseudofermion.hpp:22-166 (force part, given eta)
void dwf_force_from_eta(const Dwf& g, const double* th, const vec& eta, double* force)
{
    const vec xi = dwf_apply_Ddag(g, th, eta);
    const int V4 = g.Nt * g.Nx;
    for (int i = 0; i < g.ngauge(); ++i) force[i] = 0.0;
    const Mat2 I = identity2(), s1 = sigma1(), s2 = sigma2();
    const Mat2 ImS1 = add(I, s1, -1.0), IpS1 = add(I, s1, 1.0);
    const Mat2 ImS2 = add(I, s2, -1.0), IpS2 = add(I, s2, 1.0);
    std::vector<cplx> t_fwd(V4), t_bwd_out(V4), x_fwd(V4), x_bwd_out(V4);
    for (int t = 0; t < g.Nt; ++t) {
        const int tp = plus_periodic(t, g.Nt), tm = minus_periodic(t, g.Nt);
        const double sign_tp = (t == g.Nt - 1) ? -1.0 : 1.0;
        const double sign_tm = (t == 0) ? -1.0 : 1.0;
        for (int x = 0; x < g.Nx; ++x) {
            const int xp = plus_periodic(x, g.Nx), xm = minus_periodic(x, g.Nx);
            cplx tf = 0.0, tb = 0.0, xf = 0.0, xb = 0.0;
            for (int s = 0; s < g.Ls; ++s) {
                V2 xi_tp, xi_tm, xi_xp, xi_xm, eta_here;
                for (int a = 0; a < 2; ++a) {
                    xi_tp.v[a] = sign_tp * xi[g.fidx(s, tp, x, a)];
                    xi_tm.v[a] = sign_tm * xi[g.fidx(s, tm, x, a)];
                    xi_xp.v[a] = xi[g.fidx(s, t, xp, a)];
                    xi_xm.v[a] = xi[g.fidx(s, t, xm, a)];
                    eta_here.v[a] = eta[g.fidx(s, t, x, a)];
                }
                const V2 term_tf = mul(ImS1, xi_tp), term_tb = mul(IpS1, xi_tm);
                const V2 term_xf = mul(ImS2, xi_xp), term_xb = mul(IpS2, xi_xm);
                const cplx Ut = std::exp(cplx(0.0, th[g.gidx(0, t, x)]));
                const cplx Ut_dag = std::conj(std::exp(cplx(0.0, th[g.gidx(0, tm, x)])));
                const cplx Ux = std::exp(cplx(0.0, th[g.gidx(1, t, x)]));
                const cplx Ux_dag = std::conj(std::exp(cplx(0.0, th[g.gidx(1, t, xm)])));
                for (int a = 0; a < 2; ++a) {
                    const cplx es = std::conj(eta_here.v[a]);
                    tf += es * (cplx(0.0, -0.5) * Ut * term_tf.v[a]);
                    tb += es * (cplx(0.0, 0.5) * Ut_dag * term_tb.v[a]);
                    xf += es * (cplx(0.0, -0.5) * Ux * term_xf.v[a]);
                    xb += es * (cplx(0.0, 0.5) * Ux_dag * term_xb.v[a]);
                }
            }
            t_fwd[t * g.Nx + x] = tf;
            t_bwd_out[t * g.Nx + x] = tb;
            x_fwd[t * g.Nx + x] = xf;
            x_bwd_out[t * g.Nx + x] = xb;
        }
    }
    // pseudofermion.hpp:139-157 (np.roll shifts)
    for (int t = 0; t < g.Nt; ++t) {
        const int tp = plus_periodic(t, g.Nt);
        for (int x = 0; x < g.Nx; ++x) {
            const int xp = plus_periodic(x, g.Nx);
            const cplx t_bwd = t_bwd_out[tp * g.Nx + x];
            const cplx x_bwd = x_bwd_out[t * g.Nx + xp];
            force[g.gidx(0, t, x)] = -2.0 * std::real(t_fwd[t * g.Nx + x] + t_bwd);
            force[g.gidx(1, t, x)] = -2.0 * std::real(x_fwd[t * g.Nx + x] + x_bwd);
        }
    }
}

struct PfResult {
    double action;
    std::vector<double> force;
    vec eta;
    int iterations;
    double relres;
    bool converged;
};

PfResult dwf_pf_action_force(const Dwf& g, const double* th, const vec& phi, const vec* guess,
                             double rtol, int maxiter)
{
    auto A = [&](const vec& v) { return dwf_apply_D(g, th, dwf_apply_Ddag(g, th, v)); };
    CgResult sol = conjugate_gradient(A, phi, guess, rtol, maxiter);
    PfResult r;
    r.iterations = sol.iterations;
    r.relres = sol.relative_residual;
    r.converged = sol.converged;
    r.eta = sol.x;
    r.force.assign(g.ngauge(), 0.0);
    if (!sol.converged) {
        r.action = 0.0;
        return r;
    }
    r.action = std::real(dot(phi, r.eta));
    dwf_force_from_eta(g, th, r.eta, r.force.data());
    return r;
}

// --------------------------------------------------------------------- gauge
// domain_wall/include/gauge.hpp:15-73 (identical to wilson/include/gauge.hpp:20-62)
double plaquette_angle(int Nt, int Nx, const double* th, int t, int x)
{
    const int tp = plus_periodic(t, Nt), xp = plus_periodic(x, Nx);
    return th[(0 * Nt + t) * Nx + x] + th[(1 * Nt + tp) * Nx + x] - th[(0 * Nt + t) * Nx + xp] -
           th[(1 * Nt + t) * Nx + x];
}

double gauge_action_force(int Nt, int Nx, double beta, const double* th, double* force)
{
    double action = 0.0;
    std::vector<double> sin_p(Nt * Nx);
    for (int t = 0; t < Nt; ++t)
        for (int x = 0; x < Nx; ++x) {
            const double p = plaquette_angle(Nt, Nx, th, t, x);
            action += beta * (1.0 - std::cos(p));
            sin_p[t * Nx + x] = std::sin(p);
        }
    for (int t = 0; t < Nt; ++t) {
        const int tm = minus_periodic(t, Nt);
        for (int x = 0; x < Nx; ++x) {
            const int xm = minus_periodic(x, Nx);
            const double here = sin_p[t * Nx + x];
            force[(0 * Nt + t) * Nx + x] = beta * (here - sin_p[t * Nx + xm]);
            force[(1 * Nt + t) * Nx + x] = beta * (sin_p[tm * Nx + x] - here);
        }
    }
    return action;
}

double average_plaquette(int Nt, int Nx, const double* th)
{
    double sum = 0.0;
    for (int t = 0; t < Nt; ++t)
        for (int x = 0; x < Nx; ++x) sum += std::cos(plaquette_angle(Nt, Nx, th, t, x));
    return sum / static_cast<double>(Nt * Nx);
}

vec to_vec(const double* p, int ndof)
{
    vec v(ndof);
    for (int i = 0; i < ndof; ++i) v[i] = cplx(p[2 * i], p[2 * i + 1]);
    return v;
}
void from_vec(const vec& v, double* p)
{
    for (size_t i = 0; i < v.size(); ++i) {
        p[2 * i] = v[i].real();
        p[2 * i + 1] = v[i].imag();
    }
}

// -------------------------------------------------------------------- Wilson
// wilson/include/wilson_operator.hpp:10-62 applied matrix-free.  Row (t,x,a) of the
// sparse matrix holds the diagonal 2+m0 and four hops -1/2 bc U (1 -/+ gamma); the
// matrix-free sum over the same entries is the same linear map.
struct Wil {
    int Nt, Nx;
    double m0, bct, bcx;
    int ntot() const { return Nt * Nx * 2; }
    int sidx(int t, int x, int a) const { return (t * Nx + x) * 2 + a; }
    int lidx(int mu, int t, int x) const { return (mu * Nt + t) * Nx + x; }
};

// Visit every nonzero D[(t,x,a),(ct,cx,b)] in the order build_wilson_matrix adds them.
template <class F> void wilson_entries(const Wil& w, const double* th, F&& emit)
{
    const Mat2 I = identity2();
    const Mat2 gt = sigma1(), gx = sigma2();
    auto add_block = [&](int rt, int rx, int ct, int cx, const Mat2& A, cplx coeff) {
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const cplx value = coeff * A(a, b);
                if (std::abs(value) > 0.0) emit(w.sidx(rt, rx, a), w.sidx(ct, cx, b), value);
            }
    };
    for (int t = 0; t < w.Nt; ++t)
        for (int x = 0; x < w.Nx; ++x) {
            add_block(t, x, t, x, I, 2.0 + w.m0);
            const int tp = (t + 1) % w.Nt;
            const double bctp = (t == w.Nt - 1) ? w.bct : 1.0;
            const cplx Ut = std::exp(cplx(0, th[w.lidx(0, t, x)]));
            add_block(t, x, tp, x, add(I, gt, -1.0), -0.5 * bctp * Ut);
            const int tm = (t - 1 + w.Nt) % w.Nt;
            const double bctm = (t == 0) ? w.bct : 1.0;
            const cplx Utd = std::exp(cplx(0, -th[w.lidx(0, tm, x)]));
            add_block(t, x, tm, x, add(I, gt, 1.0), -0.5 * bctm * Utd);
            const int xp = (x + 1) % w.Nx;
            const double bcxp = (x == w.Nx - 1) ? w.bcx : 1.0;
            const cplx Ux = std::exp(cplx(0, th[w.lidx(1, t, x)]));
            add_block(t, x, t, xp, add(I, gx, -1.0), -0.5 * bcxp * Ux);
            const int xm = (x - 1 + w.Nx) % w.Nx;
            const double bcxm = (x == 0) ? w.bcx : 1.0;
            const cplx Uxd = std::exp(cplx(0, -th[w.lidx(1, t, xm)]));
            add_block(t, x, t, xm, add(I, gx, 1.0), -0.5 * bcxm * Uxd);
        }
}

vec wilson_apply(const Wil& w, const double* th, const vec& in, bool dagger)
{
    vec out(in.size(), cplx(0.0, 0.0));
    wilson_entries(w, th, [&](int row, int col, cplx v) {
        if (!dagger)
            out[row] += v * in[col];
        else
            out[col] += std::conj(v) * in[row];
    });
    return out;
}

// Diagonal of the normal matrix: column norms of D (D^dag D) or row norms (D D^dag),
// with coincident entries summed first as setFromTriplets does.
std::vector<double> wilson_normal_diag(const Wil& w, const double* th, bool DdagD)
{
    std::map<std::pair<int, int>, cplx> entries;
    wilson_entries(w, th, [&](int row, int col, cplx v) { entries[{row, col}] += v; });
    std::vector<double> d(w.ntot(), 0.0);
    for (const auto& e : entries) d[DdagD ? e.first.second : e.first.first] += std::norm(e.second);
    return d;
}

// Eigen 3 ConjugateGradient<SparseMatrix, Lower|Upper, DiagonalPreconditioner> with a
// zero initial guess (Eigen 3 public algorithm; third-party, not in the reference).
struct EigenCg {
    vec x;
    int iterations;
    double error;
};
EigenCg eigen_cg(const std::function<vec(const vec&)>& A, const std::vector<double>& diag,
                 const vec& rhs, double tol, int maxIters)
{
    const size_t n = rhs.size();
    vec x(n, cplx(0.0, 0.0));
    vec residual = rhs;  // rhs - A*0
    {
        vec Ax = A(x);
        for (size_t i = 0; i < n; ++i) residual[i] = rhs[i] - Ax[i];
    }
    const double rhsNorm2 = squared_norm(rhs);
    if (rhsNorm2 == 0.0) return {vec(n, cplx(0.0, 0.0)), 0, 0.0};
    const double considerAsZero = std::numeric_limits<double>::min();
    const double threshold = std::max(tol * tol * rhsNorm2, considerAsZero);
    double residualNorm2 = squared_norm(residual);
    if (residualNorm2 < threshold) return {x, 0, std::sqrt(residualNorm2 / rhsNorm2)};
    std::vector<double> invdiag(n);
    for (size_t i = 0; i < n; ++i) invdiag[i] = diag[i] != 0.0 ? 1.0 / diag[i] : 1.0;
    vec p(n), z(n);
    for (size_t i = 0; i < n; ++i) p[i] = invdiag[i] * residual[i];
    double absNew = std::real(dot(residual, p));
    int i = 0;
    while (i < maxIters) {
        const vec tmp = A(p);
        const cplx alpha = absNew / dot(p, tmp);
        for (size_t k = 0; k < n; ++k) {
            x[k] += alpha * p[k];
            residual[k] -= alpha * tmp[k];
        }
        residualNorm2 = squared_norm(residual);
        if (residualNorm2 < threshold) break;
        for (size_t k = 0; k < n; ++k) z[k] = invdiag[k] * residual[k];
        const double absOld = absNew;
        absNew = std::real(dot(residual, z));
        const double beta = absNew / absOld;
        for (size_t k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
        ++i;
    }
    return {x, i, std::sqrt(residualNorm2 / rhsNorm2)};
}

struct WilSolve {
    vec x;
    int iterations;
    double normal_residual, dirac_residual;
    bool failed;
};

// wilson/include/cg_solver.hpp:19-51
WilSolve wilson_solve_cgnr(const Wil& w, const double* th, const vec& source, double tol, int maxit)
{
    auto normal = [&](const vec& v) { return wilson_apply(w, th, wilson_apply(w, th, v, false), true); };
    const vec rhs = wilson_apply(w, th, source, true);
    EigenCg cg = eigen_cg(normal, wilson_normal_diag(w, th, true), rhs, tol, maxit);
    vec nx = normal(cg.x);
    for (size_t i = 0; i < nx.size(); ++i) nx[i] -= rhs[i];
    vec dx = wilson_apply(w, th, cg.x, false);
    for (size_t i = 0; i < dx.size(); ++i) dx[i] -= source[i];
    WilSolve r;
    r.x = cg.x;
    r.iterations = cg.iterations;
    r.normal_residual = norm(nx) / std::max(norm(rhs), 1.0e-300);
    r.dirac_residual = norm(dx) / std::max(norm(source), 1.0e-300);
    r.failed = !std::isfinite(r.normal_residual) || r.normal_residual > 20.0 * tol;
    return r;
}

// wilson/include/cg_solver.hpp:53-78
WilSolve wilson_solve_dddag(const Wil& w, const double* th, const vec& phi, double tol, int maxit)
{
    auto A = [&](const vec& v) { return wilson_apply(w, th, wilson_apply(w, th, v, true), false); };
    EigenCg cg = eigen_cg(A, wilson_normal_diag(w, th, false), phi, tol, maxit);
    vec ax = A(cg.x);
    for (size_t i = 0; i < ax.size(); ++i) ax[i] -= phi[i];
    WilSolve r;
    r.x = cg.x;
    r.iterations = cg.iterations;
    r.normal_residual = norm(ax) / std::max(norm(phi), 1.0e-300);
    r.dirac_residual = 0.0;
    r.failed = !std::isfinite(r.normal_residual) || r.normal_residual > 20.0 * tol;
    return r;
}

// wilson/include/gauge.hpp:68-122, force part with scatter-adds, given eta.
void wilson_force_from_eta(const Wil& w, const double* th, const vec& eta, double* force)
{
    const vec xi = wilson_apply(w, th, eta, true);
    for (int i = 0; i < 2 * w.Nt * w.Nx; ++i) force[i] = 0.0;
    const Mat2 I = identity2(), gt = sigma1(), gx = sigma2();
    const Mat2 ImGt = add(I, gt, -1.0), IpGt = add(I, gt, 1.0);
    const Mat2 ImGx = add(I, gx, -1.0), IpGx = add(I, gx, 1.0);
    for (int t = 0; t < w.Nt; ++t)
        for (int x = 0; x < w.Nx; ++x) {
            const int tp = (t + 1) % w.Nt, tm = (t - 1 + w.Nt) % w.Nt;
            const int xp = (x + 1) % w.Nx, xm = (x - 1 + w.Nx) % w.Nx;
            const double bctp = (t == w.Nt - 1) ? w.bct : 1.0;
            const double bctm = (t == 0) ? w.bct : 1.0;
            const double bcxp = (x == w.Nx - 1) ? w.bcx : 1.0;
            const double bcxm = (x == 0) ? w.bcx : 1.0;
            const cplx Ut = std::exp(cplx(0, th[w.lidx(0, t, x)]));
            const cplx Utd = std::exp(cplx(0, -th[w.lidx(0, tm, x)]));
            const cplx Ux = std::exp(cplx(0, th[w.lidx(1, t, x)]));
            const cplx Uxd = std::exp(cplx(0, -th[w.lidx(1, t, xm)]));
            V2 eh, xtp, xtm, xxp, xxm;
            for (int a = 0; a < 2; ++a) {
                eh.v[a] = eta[w.sidx(t, x, a)];
                xtp.v[a] = bctp * xi[w.sidx(tp, x, a)];
                xtm.v[a] = bctm * xi[w.sidx(tm, x, a)];
                xxp.v[a] = bcxp * xi[w.sidx(t, xp, a)];
                xxm.v[a] = bcxm * xi[w.sidx(t, xm, a)];
            }
            auto scaled = [](cplx c, const V2& v) { return V2{{c * v.v[0], c * v.v[1]}}; };
            const V2 vt = scaled(-0.5 * cplx(0, 1) * Ut, mul(ImGt, xtp));
            const V2 vtb = scaled(0.5 * cplx(0, 1) * Utd, mul(IpGt, xtm));
            const V2 vx = scaled(-0.5 * cplx(0, 1) * Ux, mul(ImGx, xxp));
            const V2 vxb = scaled(0.5 * cplx(0, 1) * Uxd, mul(IpGx, xxm));
            auto edot = [&](const V2& v) { return std::conj(eh.v[0]) * v.v[0] + std::conj(eh.v[1]) * v.v[1]; };
            force[w.lidx(0, t, x)] += -2.0 * std::real(edot(vt));
            force[w.lidx(0, tm, x)] += -2.0 * std::real(edot(vtb));
            force[w.lidx(1, t, x)] += -2.0 * std::real(edot(vx));
            force[w.lidx(1, t, xm)] += -2.0 * std::real(edot(vxb));
        }
}

// The leapfrog of domain_wall/generate_configs.cpp:103-169: a half momentum step with the
// force of the starting point (`force`, gauge + pseudofermion), then `steps` times a full
// angle step (no wrap, :113-114) and the new force (pseudofermion solve with the previous
// eta as its guess), with full momentum steps between them and a half step at the end.
// th, momentum, force, eta_guess, gauge_action and pf_action hold the end point's values;
// false if a solve failed to converge.
bool dwf_leapfrog(const Dwf& g, double beta, double epsilon, int steps, std::vector<double>& th,
                  std::vector<double>& momentum, std::vector<double>& force, const vec& phi, vec& eta_guess,
                  double pf_rtol, int pf_maxiter, double& gauge_action, double& pf_action, int& iters)
{
    const int Ngauge = g.ngauge();
    std::vector<double> gforce(Ngauge);
    for (int i = 0; i < Ngauge; ++i) momentum[i] -= 0.5 * epsilon * force[i];
    for (int step = 0; step < steps; ++step) {
        for (int i = 0; i < Ngauge; ++i) th[i] += epsilon * momentum[i];
        const double ga = gauge_action_force(g.Nt, g.Nx, beta, th.data(), gforce.data());
        PfResult pf = dwf_pf_action_force(g, th.data(), phi, &eta_guess, pf_rtol, pf_maxiter);
        if (!pf.converged) return false;
        iters += pf.iterations;
        eta_guess = pf.eta;
        gauge_action = ga;
        pf_action = pf.action;
        for (int i = 0; i < Ngauge; ++i) force[i] = gforce[i] + pf.force[i];
        if (step != steps - 1)
            for (int i = 0; i < Ngauge; ++i) momentum[i] -= epsilon * force[i];
    }
    for (int i = 0; i < Ngauge; ++i) momentum[i] -= 0.5 * epsilon * force[i];
    return true;
}

// wilson/include/indexing.hpp:13-18
double wrap_angle(double a)
{
    a = std::fmod(a + kPi, 2.0 * kPi);
    if (a < 0.0) a += 2.0 * kPi;
    return a - kPi;
}

}  // namespace

// =========================================================================== C ABI
extern "C" {

// Return codes shared with the tests.
enum { ORC_OK = 0, ORC_ARG = 1, ORC_NOT_CONVERGED = 3, ORC_ZERO_DENOM = 4 };

// --- DWF operator, domain_wall/include/dwf_operator.hpp:15-181.  op: 0 D, 1 D^dag,
//     2 D D^dag, 3 D^dag D.  Vectors are interleaved complex in the reference layout.
int orc_dwf_apply(int op, int Nt, int Nx, int Ls, double M5, double mf, const double* theta,
                  const double* in, double* out)
{
    if (Nt < 1 || Nx < 1 || Ls < 1 || op < 0 || op > 3) return ORC_ARG;
    const Dwf g{Nt, Nx, Ls, M5, mf};
    const vec v = to_vec(in, g.ndof());
    vec r;
    switch (op) {
        case 0: r = dwf_apply_D(g, theta, v); break;
        case 1: r = dwf_apply_Ddag(g, theta, v); break;
        case 2: r = dwf_apply_D(g, theta, dwf_apply_Ddag(g, theta, v)); break;
        default: r = dwf_apply_Ddag(g, theta, dwf_apply_D(g, theta, v)); break;
    }
    from_vec(r, out);
    return ORC_OK;
}

// --- conjugate_gradient on D D^dag (normal=2) or D^dag D (normal=3),
//     domain_wall/include/cg_solver.hpp:17-64.  guess may be NULL.
int orc_dwf_cg(int normal, int Nt, int Nx, int Ls, double M5, double mf, const double* theta,
               const double* b, const double* guess, double rtol, int maxiter, double* x,
               int* iterations, double* relres, int* converged)
{
    if (normal != 2 && normal != 3) return ORC_ARG;
    const Dwf g{Nt, Nx, Ls, M5, mf};
    auto A = [&](const vec& v) {
        return normal == 2 ? dwf_apply_D(g, theta, dwf_apply_Ddag(g, theta, v))
                           : dwf_apply_Ddag(g, theta, dwf_apply_D(g, theta, v));
    };
    const vec bv = to_vec(b, g.ndof());
    vec gv;
    if (guess) gv = to_vec(guess, g.ndof());
    CgResult r = conjugate_gradient(A, bv, guess ? &gv : nullptr, rtol, maxiter);
    from_vec(r.x, x);
    *iterations = r.iterations;
    *relres = r.relative_residual;
    *converged = r.converged ? 1 : 0;
    return r.zero_denominator ? ORC_ZERO_DENOM : ORC_OK;
}

// --- The analyzers' parallelism (domain_wall/analyze_configs.cpp:45-76: OpenMP
//     `parallel for schedule(dynamic)` over independent work): n independent
//     conjugate_gradient solves on D^dag D, solve i on the Nt[i] x Nx[i] lattice with links
//     theta[i], right-hand side b[i] and solution x[i], spread over the OpenMP threads
//     (OMP_NUM_THREADS).  Returns the number of threads that ran.
int orc_dwf_cg_many(int n, const int* Nt, const int* Nx, int Ls, double M5, double mf, const double* const* theta,
                    const double* const* b, double rtol, int maxiter, double* const* x, int* iterations,
                    int* converged)
{
    int threads = 1;
#pragma omp parallel
    {
#pragma omp single
        threads = omp_get_num_threads();
#pragma omp for schedule(dynamic)
        for (int i = 0; i < n; ++i) {
            double rr = 0.0;
            orc_dwf_cg(3, Nt[i], Nx[i], Ls, M5, mf, theta[i], b[i], nullptr, rtol, maxiter, x[i], &iterations[i], &rr,
                       &converged[i]);
        }
    }
    return threads;
}

// --- pseudofermion_action_force, domain_wall/include/pseudofermion.hpp:22-166.
int orc_dwf_pf_action_force(int Nt, int Nx, int Ls, double M5, double mf, const double* theta,
                            const double* phi, const double* eta_guess, double rtol, int maxiter,
                            double* action, double* force, double* eta, int* iterations,
                            double* relres)
{
    const Dwf g{Nt, Nx, Ls, M5, mf};
    const vec pv = to_vec(phi, g.ndof());
    vec gv;
    if (eta_guess) gv = to_vec(eta_guess, g.ndof());
    PfResult r = dwf_pf_action_force(g, theta, pv, eta_guess ? &gv : nullptr, rtol, maxiter);
    *iterations = r.iterations;
    *relres = r.relres;
    if (eta) from_vec(r.eta, eta);
    if (!r.converged) return ORC_NOT_CONVERGED;
    *action = r.action;
    for (int i = 0; i < g.ngauge(); ++i) force[i] = r.force[i];
    return ORC_OK;
}

// --- force pass alone for a given eta (pseudofermion.hpp:109-157).
int orc_dwf_force_from_eta(int Nt, int Nx, int Ls, double M5, double mf, const double* theta,
                           const double* eta, double* force)
{
    const Dwf g{Nt, Nx, Ls, M5, mf};
    dwf_force_from_eta(g, theta, to_vec(eta, g.ndof()), force);
    return ORC_OK;
}

// --- gauge.hpp:26-73
double orc_gauge_action_force(int Nt, int Nx, double beta, const double* theta, double* force)
{
    return gauge_action_force(Nt, Nx, beta, theta, force);
}
double orc_average_plaquette(int Nt, int Nx, const double* theta)
{
    return average_plaquette(Nt, Nx, theta);
}

// --- DWF HMC, domain_wall/generate_configs.cpp:23-207.  theta (in/out) holds the
//     start configuration and receives the final theta0.  Per trajectory outputs may be
//     NULL.  Returns ORC_NOT_CONVERGED if a force solve fails (the reference throws).
// --- The leapfrog of one trajectory from (theta, momentum) with pseudofermion phi
//     (domain_wall/generate_configs.cpp:103-169, as dwf_leapfrog): a first solve without a
//     guess gives H at the start, the integration updates theta and momentum in place, and
//     H at the end is the kinetic term plus the actions at the last step.
int orc_dwf_leapfrog(int Nt, int Nx, int Ls, double M5, double mf, double beta, double epsilon, int steps,
                     double* theta, double* momentum, const double* phi, double pf_rtol, int pf_maxiter,
                     double* h_initial, double* h_final)
{
    const Dwf g{Nt, Nx, Ls, M5, mf};
    const int Ngauge = g.ngauge();
    const vec ph = to_vec(phi, g.ndof());
    std::vector<double> th(theta, theta + Ngauge), mom(momentum, momentum + Ngauge), gforce(Ngauge), force(Ngauge);
    double gauge_action = gauge_action_force(Nt, Nx, beta, th.data(), gforce.data());
    PfResult pf0 = dwf_pf_action_force(g, th.data(), ph, nullptr, pf_rtol, pf_maxiter);
    if (!pf0.converged) return ORC_NOT_CONVERGED;
    double kin = 0.0;
    for (double p : mom) kin += 0.5 * p * p;
    *h_initial = kin + gauge_action + pf0.action;
    for (int i = 0; i < Ngauge; ++i) force[i] = gforce[i] + pf0.force[i];
    vec eta_guess = pf0.eta;
    double pf_action = pf0.action;
    int iters = 0;
    if (!dwf_leapfrog(g, beta, epsilon, steps, th, mom, force, ph, eta_guess, pf_rtol, pf_maxiter, gauge_action,
                      pf_action, iters))
        return ORC_NOT_CONVERGED;
    kin = 0.0;
    for (double p : mom) kin += 0.5 * p * p;
    *h_final = kin + gauge_action + pf_action;
    std::copy(th.begin(), th.end(), theta);
    std::copy(mom.begin(), mom.end(), momentum);
    return ORC_OK;
}

int orc_dwf_hmc(int Nt, int Nx, int Ls, double M5, double mf, double beta, double epsilon,
                int leapfrog_steps, int trajectories, unsigned long long seed, double pf_rtol,
                int pf_maxiter, double* theta, int* accepted, double* acc_prob, double* delta_h,
                double* plaquette, int* cg_iterations)
{
    const Dwf g{Nt, Nx, Ls, M5, mf};
    const int Ndof = g.ndof();
    const int Ngauge = g.ngauge();
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> normal(0.0, 1.0);
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    std::vector<double> theta0(theta, theta + Ngauge);
    int total_iters = 0;
    for (int traj = 0; traj < trajectories; ++traj) {
        vec chi(Ndof);
        for (int i = 0; i < Ndof; ++i) chi[i] = cplx(normal(rng), normal(rng)) / std::sqrt(2.0);
        const vec phi = dwf_apply_D(g, theta0.data(), chi);
        std::vector<double> momentum(Ngauge);
        for (double& p : momentum) p = normal(rng);
        std::vector<double> gforce(Ngauge);
        const double g_action0 = gauge_action_force(Nt, Nx, beta, theta0.data(), gforce.data());
        PfResult pf0 = dwf_pf_action_force(g, theta0.data(), phi, nullptr, pf_rtol, pf_maxiter);
        if (!pf0.converged) return ORC_NOT_CONVERGED;
        total_iters += pf0.iterations;
        vec eta_guess = pf0.eta;
        double kin0 = 0.0;
        for (double p : momentum) kin0 += 0.5 * p * p;
        const double H0 = kin0 + g_action0 + pf0.action;
        std::vector<double> th = theta0, force(Ngauge);
        for (int i = 0; i < Ngauge; ++i) force[i] = gforce[i] + pf0.force[i];
        double gauge_action = g_action0, pf_action = pf0.action;
        if (!dwf_leapfrog(g, beta, epsilon, leapfrog_steps, th, momentum, force, phi, eta_guess, pf_rtol, pf_maxiter,
                          gauge_action, pf_action, total_iters))
            return ORC_NOT_CONVERGED;
        double kin = 0.0;
        for (double p : momentum) kin += 0.5 * p * p;
        const double H = kin + gauge_action + pf_action;
        const double dH = H - H0;
        const double pacc = std::min(std::exp(-dH), 1.0);
        const bool acc = uniform(rng) < pacc;
        if (acc) theta0 = th;
        if (accepted) accepted[traj] = acc ? 1 : 0;
        if (acc_prob) acc_prob[traj] = pacc;
        if (delta_h) delta_h[traj] = dH;
        if (plaquette) plaquette[traj] = average_plaquette(Nt, Nx, theta0.data());
    }
    for (int i = 0; i < Ngauge; ++i) theta[i] = theta0[i];
    if (cg_iterations) *cg_iterations = total_iters;
    return ORC_OK;
}

// --- correlator.hpp:183-310 (wall sources, C_PP and C_J5).  nonconverged counts the
//     wall solves that printed "CGNR info = nonzero"; dres receives the 4 D residuals.
int orc_dwf_cpp_cj5(int Nt, int Nx, int Ls, double M5, double mf, const double* theta,
                    double rtol, int maxiter, double* CPP, double* CJ5, int* nonconverged,
                    int* iterations, double* dres)
{
    const Dwf g{Nt, Nx, Ls, M5, mf};
    const Mat2 PL = projector_L(), PR = projector_R();
    auto A = [&](const vec& v) { return dwf_apply_Ddag(g, theta, dwf_apply_D(g, theta, v)); };
    for (int t = 0; t < Nt; ++t) CPP[t] = CJ5[t] = 0.0;
    int nc = 0, k = 0;
    const int source_times[2] = {0, Nt / 2};
    for (int t_src : source_times) {
        vec props[2];
        for (int source_spin = 0; source_spin < 2; ++source_spin) {
            vec source(g.ndof(), cplx(0.0, 0.0));
            for (int x = 0; x < Nx; ++x)
                for (int spin = 0; spin < 2; ++spin) {
                    source[g.fidx(0, t_src, x, spin)] += PL(spin, source_spin);
                    source[g.fidx(Ls - 1, t_src, x, spin)] += PR(spin, source_spin);
                }
            const vec rhs = dwf_apply_Ddag(g, theta, source);
            CgResult sol = conjugate_gradient(A, rhs, nullptr, rtol, maxiter);
            if (!sol.converged) ++nc;
            vec dx = dwf_apply_D(g, theta, sol.x);
            for (size_t i = 0; i < dx.size(); ++i) dx[i] -= source[i];
            if (dres) dres[k] = norm(dx) / norm(source);
            if (iterations) iterations[k] = sol.iterations;
            ++k;
            props[source_spin] = sol.x;
        }
        for (int t = 0; t < Nt; ++t) {
            const int dt = (t - t_src + Nt) % Nt;
            for (int x = 0; x < Nx; ++x)
                for (int ss = 0; ss < 2; ++ss) {
                    V2 l, r, m;
                    for (int spin = 0; spin < 2; ++spin) {
                        l.v[spin] = props[ss][g.fidx(0, t, x, spin)];
                        r.v[spin] = props[ss][g.fidx(Ls - 1, t, x, spin)];
                        m.v[spin] = props[ss][g.fidx(Ls / 2, t, x, spin)];
                    }
                    const V2 pl = mul(PL, l), pr = mul(PR, r);
                    const V2 q{{pl.v[0] + pr.v[0], pl.v[1] + pr.v[1]}};
                    CPP[dt] += std::norm(q.v[0]) + std::norm(q.v[1]);
                    for (int spin = 0; spin < 2; ++spin) CJ5[dt] += std::real(std::conj(m.v[spin]) * q.v[spin]);
                }
        }
    }
    for (int t = 0; t < Nt; ++t) {
        CPP[t] /= 2.0;
        CJ5[t] /= 2.0;
    }
    if (nonconverged) *nonconverged = nc;
    return ORC_OK;
}

// --- Z4 noise of correlator.hpp:35-58: fills eta[nnoise][Nt][Nx][2] (complex).
void orc_z4_noise(int Nt, int Nx, int nnoise, unsigned long long seed, double* eta)
{
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> z4(0, 3);
    const cplx phases[4] = {cplx(1.0, 0.0), cplx(0.0, 1.0), cplx(-1.0, 0.0), cplx(0.0, -1.0)};
    const int n = Nt * Nx * 2;
    for (int k = 0; k < nnoise; ++k)
        for (int i = 0; i < n; ++i) {
            const cplx v = phases[z4(rng)];
            eta[2 * ((size_t)k * n + i)] = v.real();
            eta[2 * ((size_t)k * n + i) + 1] = v.imag();
        }
}

// --- estimate_pseudoscalar_loop_density, correlator.hpp:27-109.  loop: Nt*Nx complex.
int orc_dwf_loop_density(int Nt, int Nx, int Ls, double M5, double mf, const double* theta,
                         unsigned long long seed, int nnoise, double rtol, int maxiter,
                         double* loop_out)
{
    if (nnoise < 1) return ORC_ARG;
    const Dwf g{Nt, Nx, Ls, M5, mf};
    std::vector<cplx> loop(Nt * Nx, cplx(0.0, 0.0));
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> z4(0, 3);
    const cplx phases[4] = {cplx(1.0, 0.0), cplx(0.0, 1.0), cplx(-1.0, 0.0), cplx(0.0, -1.0)};
    const Mat2 PL = projector_L(), PR = projector_R(), g5 = sigma3();
    auto A = [&](const vec& v) { return dwf_apply_Ddag(g, theta, dwf_apply_D(g, theta, v)); };
    for (int noise = 0; noise < nnoise; ++noise) {
        std::vector<cplx> eta(Nt * Nx * 2);
        vec source(g.ndof(), cplx(0.0, 0.0));
        for (int t = 0; t < Nt; ++t)
            for (int x = 0; x < Nx; ++x)
                for (int spin = 0; spin < 2; ++spin) eta[(t * Nx + x) * 2 + spin] = phases[z4(rng)];
        for (int t = 0; t < Nt; ++t)
            for (int x = 0; x < Nx; ++x) {
                V2 value;
                for (int spin = 0; spin < 2; ++spin) value.v[spin] = eta[(t * Nx + x) * 2 + spin];
                const V2 left = mul(PL, value), right = mul(PR, value);
                for (int spin = 0; spin < 2; ++spin) {
                    source[g.fidx(0, t, x, spin)] = left.v[spin];
                    source[g.fidx(Ls - 1, t, x, spin)] += right.v[spin];
                }
            }
        const vec rhs = dwf_apply_Ddag(g, theta, source);
        CgResult sol = conjugate_gradient(A, rhs, nullptr, rtol, maxiter);
        if (!sol.converged) return ORC_NOT_CONVERGED;
        for (int t = 0; t < Nt; ++t)
            for (int x = 0; x < Nx; ++x) {
                V2 l, r, nv;
                for (int spin = 0; spin < 2; ++spin) {
                    l.v[spin] = sol.x[g.fidx(0, t, x, spin)];
                    r.v[spin] = sol.x[g.fidx(Ls - 1, t, x, spin)];
                    nv.v[spin] = eta[(t * Nx + x) * 2 + spin];
                }
                const V2 pl = mul(PL, l), pr = mul(PR, r);
                const V2 q{{pl.v[0] + pr.v[0], pl.v[1] + pr.v[1]}};
                const V2 g5q = mul(g5, q);
                loop[t * Nx + x] += std::conj(nv.v[0]) * g5q.v[0] + std::conj(nv.v[1]) * g5q.v[1];
            }
    }
    for (int i = 0; i < Nt * Nx; ++i) {
        loop[i] /= static_cast<double>(nnoise);
        loop_out[2 * i] = loop[i].real();
        loop_out[2 * i + 1] = loop[i].imag();
    }
    return ORC_OK;
}

// --- compute_momentum_CPP, correlator.hpp:111-181.  result[(max_momentum+1)][Nt].
int orc_dwf_momentum_cpp(int Nt, int Nx, int Ls, double M5, double mf, const double* theta,
                         int max_momentum, double rtol, int maxiter, double* result)
{
    if (max_momentum < 0 || max_momentum > Nx / 2) return ORC_ARG;
    const Dwf g{Nt, Nx, Ls, M5, mf};
    const Mat2 PL = projector_L(), PR = projector_R();
    auto A = [&](const vec& v) { return dwf_apply_Ddag(g, theta, dwf_apply_D(g, theta, v)); };
    for (int i = 0; i < (max_momentum + 1) * Nt; ++i) result[i] = 0.0;
    const int source_times[2] = {0, Nt / 2};
    for (int t_src : source_times) {
        vec props[2];
        for (int ss = 0; ss < 2; ++ss) {
            vec source(g.ndof(), cplx(0.0, 0.0));
            for (int spin = 0; spin < 2; ++spin) {
                source[g.fidx(0, t_src, 0, spin)] += PL(spin, ss);
                source[g.fidx(Ls - 1, t_src, 0, spin)] += PR(spin, ss);
            }
            const vec rhs = dwf_apply_Ddag(g, theta, source);
            CgResult sol = conjugate_gradient(A, rhs, nullptr, rtol, maxiter);
            if (!sol.converged) return ORC_NOT_CONVERGED;
            props[ss] = sol.x;
        }
        for (int t = 0; t < Nt; ++t) {
            const int dt = (t - t_src + Nt) % Nt;
            for (int x = 0; x < Nx; ++x) {
                double local = 0.0;
                for (int ss = 0; ss < 2; ++ss) {
                    V2 l, r;
                    for (int spin = 0; spin < 2; ++spin) {
                        l.v[spin] = props[ss][g.fidx(0, t, x, spin)];
                        r.v[spin] = props[ss][g.fidx(Ls - 1, t, x, spin)];
                    }
                    const V2 pl = mul(PL, l), pr = mul(PR, r);
                    local += std::norm(pl.v[0] + pr.v[0]) + std::norm(pl.v[1] + pr.v[1]);
                }
                for (int n = 0; n <= max_momentum; ++n) {
                    const double p = 2.0 * kPi * n / Nx;
                    result[n * Nt + dt] += std::cos(p * x) * local;
                }
            }
        }
    }
    for (int i = 0; i < (max_momentum + 1) * Nt; ++i) result[i] /= 2.0;
    return ORC_OK;
}

// --- Wilson operator, wilson/include/wilson_operator.hpp:10-62.  op: 0 D, 1 D^dag.
int orc_wilson_apply(int op, int Nt, int Nx, double m0, double bct, double bcx, const double* theta,
                     const double* in, double* out)
{
    const Wil w{Nt, Nx, m0, bct, bcx};
    from_vec(wilson_apply(w, theta, to_vec(in, w.ntot()), op == 1), out);
    return ORC_OK;
}

// --- solve_CGNR, wilson/include/cg_solver.hpp:19-51.  Returns ORC_NOT_CONVERGED where
//     the reference throws "CGNR failed: ...".
int orc_wilson_solve_cgnr(int Nt, int Nx, double m0, double bct, double bcx, const double* theta,
                          const double* source, double tol, int maxit, double* x, int* iterations,
                          double* normal_residual, double* dirac_residual)
{
    const Wil w{Nt, Nx, m0, bct, bcx};
    WilSolve r = wilson_solve_cgnr(w, theta, to_vec(source, w.ntot()), tol, maxit);
    from_vec(r.x, x);
    *iterations = r.iterations;
    *normal_residual = r.normal_residual;
    *dirac_residual = r.dirac_residual;
    return r.failed ? ORC_NOT_CONVERGED : ORC_OK;
}

// --- solve_DDdag, wilson/include/cg_solver.hpp:53-78.
int orc_wilson_solve_dddag(int Nt, int Nx, double m0, double bct, double bcx, const double* theta,
                           const double* phi, double tol, int maxit, double* eta, int* iterations,
                           double* residual)
{
    const Wil w{Nt, Nx, m0, bct, bcx};
    WilSolve r = wilson_solve_dddag(w, theta, to_vec(phi, w.ntot()), tol, maxit);
    from_vec(r.x, eta);
    *iterations = r.iterations;
    *residual = r.normal_residual;
    return r.failed ? ORC_NOT_CONVERGED : ORC_OK;
}

// --- Wilson pseudofermion_action_force, wilson/include/gauge.hpp:68-122.
int orc_wilson_pf_action_force(int Nt, int Nx, double m0, double bct, double bcx,
                               const double* theta, const double* phi, double tol, int maxit,
                               double* action, double* force, double* eta_out)
{
    const Wil w{Nt, Nx, m0, bct, bcx};
    const vec pv = to_vec(phi, w.ntot());
    WilSolve r = wilson_solve_dddag(w, theta, pv, tol, maxit);
    if (r.failed) return ORC_NOT_CONVERGED;
    *action = std::real(dot(pv, r.x));
    wilson_force_from_eta(w, theta, r.x, force);
    if (eta_out) from_vec(r.x, eta_out);
    return ORC_OK;
}

int orc_wilson_force_from_eta(int Nt, int Nx, double m0, double bct, double bcx, const double* theta,
                              const double* eta, double* force)
{
    const Wil w{Nt, Nx, m0, bct, bcx};
    wilson_force_from_eta(w, theta, to_vec(eta, w.ntot()), force);
    return ORC_OK;
}

double orc_wrap_angle(double a) { return wrap_angle(a); }

// --- Wilson HMC: wilson/generate_configs.cpp:20-48 driving hmc_step,
//     wilson/include/gauge.hpp:132-180, with one mt19937_64(seed) for the whole run.
int orc_wilson_hmc(int Nt, int Nx, double m0, double bct, double bcx, double beta,
                   double trajectory_length, int hmc_steps, int trajectories,
                   unsigned long long seed, double tol, int maxit, double* theta, int* accepted,
                   double* acc_prob, double* delta_h, double* plaquette)
{
    const Wil w{Nt, Nx, m0, bct, bcx};
    const int Ng = 2 * Nt * Nx, Ntot = w.ntot();
    const double eps = trajectory_length / hmc_steps;
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> gaussian(0.0, 1.0);
    std::uniform_real_distribution<double> uniform01(0.0, 1.0);
    std::vector<double> old_theta(theta, theta + Ng);
    for (int traj = 0; traj < trajectories; ++traj) {
        std::vector<double> momentum(Ng);
        for (double& p : momentum) p = gaussian(rng);
        vec chi(Ntot);
        for (int i = 0; i < Ntot; ++i) chi[i] = cplx(gaussian(rng), gaussian(rng)) / std::sqrt(2.0);
        const vec phi = wilson_apply(w, old_theta.data(), chi, false);
        std::vector<double> Fg(Ng), Fpf(Ng), total(Ng);
        auto pf = [&](const std::vector<double>& th, std::vector<double>& F, double& S) {
            WilSolve r = wilson_solve_dddag(w, th.data(), phi, tol, maxit);
            if (r.failed) return false;
            S = std::real(dot(phi, r.x));
            wilson_force_from_eta(w, th.data(), r.x, F.data());
            return true;
        };
        double old_Spf = 0.0, new_Spf = 0.0;
        const double old_Sg = gauge_action_force(Nt, Nx, beta, old_theta.data(), Fg.data());
        if (!pf(old_theta, Fpf, old_Spf)) return ORC_NOT_CONVERGED;
        double old_K = 0.0;
        for (double p : momentum) old_K += p * p;
        old_K *= 0.5;
        std::vector<double> th = old_theta;
        for (int i = 0; i < Ng; ++i) total[i] = Fg[i] + Fpf[i];
        for (int i = 0; i < Ng; ++i) momentum[i] -= 0.5 * eps * total[i];
        for (int step = 0; step < hmc_steps; ++step) {
            for (int i = 0; i < Ng; ++i) th[i] = wrap_angle(th[i] + eps * momentum[i]);
            double dummy = 0.0;
            gauge_action_force(Nt, Nx, beta, th.data(), Fg.data());
            if (!pf(th, Fpf, dummy)) return ORC_NOT_CONVERGED;
            for (int i = 0; i < Ng; ++i) total[i] = Fg[i] + Fpf[i];
            const double factor = (step == hmc_steps - 1) ? 0.5 : 1.0;
            for (int i = 0; i < Ng; ++i) momentum[i] -= factor * eps * total[i];
        }
        const double new_Sg = gauge_action_force(Nt, Nx, beta, th.data(), Fg.data());
        if (!pf(th, Fpf, new_Spf)) return ORC_NOT_CONVERGED;
        double new_K = 0.0;
        for (double p : momentum) new_K += p * p;
        new_K *= 0.5;
        const double dH = (new_K + new_Sg + new_Spf) - (old_K + old_Sg + old_Spf);
        const double pacc = std::min(1.0, std::exp(-dH));
        const bool acc = uniform01(rng) < pacc;
        if (acc) old_theta = th;
        if (accepted) accepted[traj] = acc ? 1 : 0;
        if (acc_prob) acc_prob[traj] = pacc;
        if (delta_h) delta_h[traj] = dH;
        if (plaquette) plaquette[traj] = average_plaquette(Nt, Nx, old_theta.data());
    }
    for (int i = 0; i < Ng; ++i) theta[i] = old_theta[i];
    return ORC_OK;
}

// --- pseudoscalar_correlator, wilson/include/correlator.hpp:10-40.  C[Nt].
int orc_wilson_pseudoscalar_correlator(int Nt, int Nx, double m0, double bct, double bcx,
                                       const double* theta, int source_t, int source_x, double tol,
                                       int maxit, double* C)
{
    const Wil w{Nt, Nx, m0, bct, bcx};
    vec prop[2];
    for (int ss = 0; ss < 2; ++ss) {
        vec source(w.ntot(), cplx(0.0, 0.0));
        source[w.sidx(source_t, source_x, ss)] = 1.0;
        WilSolve r = wilson_solve_cgnr(w, theta, source, tol, maxit);
        if (r.failed) return ORC_NOT_CONVERGED;
        prop[ss] = r.x;
    }
    for (int t = 0; t < Nt; ++t) {
        C[t] = 0.0;
        for (int x = 0; x < Nx; ++x)
            for (int sink = 0; sink < 2; ++sink)
                for (int ss = 0; ss < 2; ++ss) C[t] += std::norm(prop[ss][w.sidx(t, x, sink)]);
    }
    return ORC_OK;
}

// --- std::mt19937_64 / libstdc++ distributions, exposed so tests can pin the engine
//     against the C++ standard's fixed value and compare host RNG streams.
unsigned long long orc_mt19937_64_nth(unsigned long long seed, int n)
{
    std::mt19937_64 rng(seed);
    unsigned long long v = 0;
    for (int i = 0; i < n; ++i) v = rng();
    return v;
}

// DWF heat-bath draws of domain_wall/generate_configs.cpp:56-69 for one trajectory
// from a fresh engine: chi (Ndof complex) then momentum (Ngauge), then the accept draw.
void orc_dwf_heatbath_draws(int ndof, int ngauge, unsigned long long seed, double* chi,
                            double* momentum, double* u)
{
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> normal(0.0, 1.0);
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    for (int i = 0; i < ndof; ++i) {
        const cplx c = cplx(normal(rng), normal(rng)) / std::sqrt(2.0);
        chi[2 * i] = c.real();
        chi[2 * i + 1] = c.imag();
    }
    for (int i = 0; i < ngauge; ++i) momentum[i] = normal(rng);
    *u = uniform(rng);
}

}  // extern "C"
