// dwf_rng.cpp -- host random numbers of the hot path's callers, drawn in the reference's
// order.  Compiled by g++ against libstdc++ (not by hipcc/clang): the reference builds
// with g++ (compile.sh:8-15), std::normal_distribution is implementation-defined, and
// the evaluation order of Complex(normal(rng), normal(rng)) is unspecified in C++ -- the
// same expression under the same compiler gives the same stream.

#include "../../include/dwf_hip.h"

#include <cmath>
#include <complex>
#include <random>

struct dwf_rng {
    std::mt19937_64 engine;
    std::normal_distribution<double> normal{0.0, 1.0};
    std::uniform_real_distribution<double> uniform{0.0, 1.0};
    explicit dwf_rng(unsigned long long seed) : engine(seed) {}
};

using Complex = std::complex<double>;

extern "C" {

dwf_rng* dwf_rng_create(unsigned long long seed) { return new dwf_rng(seed); }

void dwf_rng_destroy(dwf_rng* r) { delete r; }

void dwf_rng_dwf_heatbath(dwf_rng* r, int ndof, int ngauge, double* chi, double* momentum)
{
    // domain_wall/generate_configs.cpp:56-69
    for (int i = 0; i < ndof; ++i) {
        const Complex c = Complex(r->normal(r->engine), r->normal(r->engine)) / std::sqrt(2.0);
        chi[2 * i] = c.real();
        chi[2 * i + 1] = c.imag();
    }
    for (int i = 0; i < ngauge; ++i) momentum[i] = r->normal(r->engine);
}

void dwf_rng_wilson_heatbath(dwf_rng* r, int ngauge, int ntot, double* momentum, double* chi)
{
    // wilson/include/gauge.hpp:134-141: the momenta are drawn before chi
    for (int i = 0; i < ngauge; ++i) momentum[i] = r->normal(r->engine);
    for (int i = 0; i < ntot; ++i) {
        const Complex c = Complex(r->normal(r->engine), r->normal(r->engine)) / std::sqrt(2.0);
        chi[2 * i] = c.real();
        chi[2 * i + 1] = c.imag();
    }
}

double dwf_rng_uniform(dwf_rng* r) { return r->uniform(r->engine); }

void dwf_z4_noise(unsigned long long seed, int nnoise, int sites_spin, double* eta)
{
    // correlator.hpp:35-58: one engine per configuration, noise vectors back to back
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> z4(0, 3);
    static const double re[4] = {1.0, 0.0, -1.0, 0.0};
    static const double im[4] = {0.0, 1.0, 0.0, -1.0};
    const long long n = (long long)nnoise * sites_spin;
    for (long long i = 0; i < n; ++i) {
        const int k = z4(rng);
        eta[2 * i] = re[k];
        eta[2 * i + 1] = im[k];
    }
}

}  // extern "C"
