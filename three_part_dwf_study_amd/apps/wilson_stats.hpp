// wilson_stats.hpp -- host statistics of the Wilson analyzer: ensemble mean and error of
// the pseudoscalar correlator, leave-one-out jackknife cosh effective mass, the first
// flat effective-mass window, and the jackknife mass over it.  The rules and the printed
// report follow wilson/include/statistics.hpp:13-154 and wilson/analyze_configs.cpp:36-79
// (drop-in output); the code is this project's.  O(configurations x Nt) on the CPU.
#pragma once

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace app {

using Series = std::vector<double>;

struct PlateauSettings {
    int window = 4;           // analysis.plateau_window_size
    double flatness = 0.08;   // analysis.plateau_flatness_tolerance (relative drift)
    int earliest = 2;         // analysis.earliest_plateau_time
};

struct Plateau {
    int first_t = -1, last_t = -1;
    double mean = std::numeric_limits<double>::quiet_NaN();
    double drift = std::numeric_limits<double>::quiet_NaN();
};

struct WilsonMass {
    Series average, error;        // ensemble mean of C(t) and its standard error
    Series meff, meff_error;      // jackknife cosh effective mass (NaN where a sample is invalid)
    Plateau plateau;
    double mass = 0.0, mass_error = 0.0;
};

// Mean over the configurations, leaving out `omit` (-1: none) (statistics.hpp:13-36).
inline Series ensemble_mean(const std::vector<Series>& c, int omit = -1)
{
    if (c.empty()) throw std::runtime_error("No configurations in mean_correlator");
    Series m(c[0].size(), 0.0);
    int n = 0;
    for (int k = 0; k < (int)c.size(); ++k) {
        if (k == omit) continue;
        for (size_t t = 0; t < m.size(); ++t) m[t] += c[k][t];
        ++n;
    }
    if (n == 0) throw std::runtime_error("No configurations in mean_correlator");
    for (double& v : m) v /= n;
    return m;
}

// m(t) = acosh((C(t-1) + C(t+1)) / 2C(t)) for 0 < t < Nt-1 where C(t) > 0 and the ratio is
// at least 1; NaN elsewhere (statistics.hpp:38-56).
inline Series cosh_mass(const Series& C)
{
    const int Nt = (int)C.size();
    Series m(Nt, std::numeric_limits<double>::quiet_NaN());
    for (int t = 1; t + 1 < Nt; ++t) {
        if (!(std::isfinite(C[t]) && C[t] > 0.0)) continue;
        const double r = (C[t - 1] + C[t + 1]) / (2.0 * C[t]);
        if (std::isfinite(r) && r >= 1.0) m[t] = std::acosh(r);
    }
    return m;
}

// sqrt((n-1)/n sum (x - mean)^2) (statistics.hpp:139-154).
inline double jackknife_error(const std::vector<double>& x)
{
    if (x.size() < 2) throw std::runtime_error("Need at least two jackknife samples");
    double mean = 0.0;
    for (double v : x) mean += v;
    mean /= x.size();
    double ss = 0.0;
    for (double v : x) ss += (v - mean) * (v - mean);
    return std::sqrt((x.size() - 1.0) / x.size() * ss);
}

// The first window of `window` consecutive times, from `earliest` up to t = Nt/2, whose
// effective masses are all finite and positive and whose (max - min) / mean is within
// `flatness`; the candidates are reported as they are tried (statistics.hpp:66-124).
inline Plateau flat_plateau(const Series& m, const PlateauSettings& P, std::ostream* report)
{
    const int Nt = (int)m.size(), last = Nt / 2;
    if (report)
        *report << "\nAutomatic contiguous-window plateau search\n"
                << "-------------------------------------------\n"
                << "Window size: " << P.window << '\n'
                << "Maximum relative drift: " << 100.0 * P.flatness << "%\n\n"
                << "Candidate windows:\n";
    for (int a = P.earliest; a + P.window - 1 <= last; ++a) {
        const int b = a + P.window - 1;
        double sum = 0.0, lo = std::numeric_limits<double>::infinity(), hi = -lo;
        bool ok = true;
        for (int t = a; t <= b && ok; ++t) {
            const double v = m[t];
            ok = std::isfinite(v) && v > 0.0;
            if (ok) {
                sum += v;
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
        }
        if (!ok) {
            if (report) *report << "t = " << a << " through " << b << "   INVALID\n";
            continue;
        }
        const double mean = sum / P.window, drift = (hi - lo) / mean;
        const bool accept = drift <= P.flatness;
        if (report)
            *report << "t = " << a << " through " << b << "   mean = " << mean << "   drift = " << 100.0 * drift << "%"
                    << (accept ? "   ACCEPTED" : "   rejected") << '\n';
        if (accept) return Plateau{a, b, mean, drift};
    }
    throw std::runtime_error("No contiguous effective-mass window passed the flatness test");
}

// Mean of m over the window (statistics.hpp:126-137).
inline double window_mean(const Series& m, const Plateau& w)
{
    double s = 0.0;
    for (int t = w.first_t; t <= w.last_t; ++t) {
        if (!std::isfinite(m[t])) throw std::runtime_error("A jackknife sample is invalid inside the selected plateau");
        s += m[t];
    }
    return s / (w.last_t - w.first_t + 1);
}

// The whole chain of wilson/analyze_configs.cpp:36-79 on the per-configuration
// correlators: mean and error, jackknife effective mass (a time counts only where every
// leave-one-out sample is finite), plateau on the jackknife mean, jackknife mass.
inline WilsonMass wilson_mass(const std::vector<Series>& all, const PlateauSettings& P, std::ostream* report)
{
    if (all.size() < 2) throw std::runtime_error("Need at least two configurations");
    const int Nt = (int)all[0].size();
    const double n = (double)all.size();
    WilsonMass r;
    r.average = ensemble_mean(all);
    r.error.assign(Nt, 0.0);
    for (int t = 0; t < Nt; ++t) {
        double ss = 0.0;
        for (const Series& c : all) ss += (c[t] - r.average[t]) * (c[t] - r.average[t]);
        r.error[t] = std::sqrt(ss / (n - 1.0) / n);
    }
    std::vector<Series> jk;
    for (int k = 0; k < (int)all.size(); ++k) jk.push_back(cosh_mass(ensemble_mean(all, k)));
    const double nan = std::numeric_limits<double>::quiet_NaN();
    r.meff.assign(Nt, nan);
    r.meff_error.assign(Nt, nan);
    for (int t = 0; t < Nt; ++t) {
        std::vector<double> s;
        for (const Series& m : jk)
            if (std::isfinite(m[t])) s.push_back(m[t]);
        if (s.size() != jk.size()) continue;
        double mean = 0.0;
        for (double v : s) mean += v;
        r.meff[t] = mean / s.size();
        r.meff_error[t] = jackknife_error(s);
    }
    r.plateau = flat_plateau(r.meff, P, report);
    std::vector<double> masses;
    for (const Series& m : jk) masses.push_back(window_mean(m, r.plateau));
    double mean = 0.0;
    for (double v : masses) mean += v;
    r.mass = mean / masses.size();
    r.mass_error = jackknife_error(masses);
    if (report)
        *report << std::setprecision(8) << "\nSelected plateau: t=" << r.plateau.first_t << " through "
                << r.plateau.last_t << "\n"
                << "Wilson pseudoscalar mass = " << r.mass << " +/- " << r.mass_error << "\n";
    return r;
}

// t, C(t) mean and error, m_eff(t) and its jackknife error (the series the reference
// plots, analyze_configs.cpp:81-90); NaN where the effective mass is undefined.
inline void write_effective_mass_csv(const char* path, const WilsonMass& m)
{
    std::ofstream f(path);
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    f << std::setprecision(12) << "t,C_average,C_error,m_eff,m_eff_error\n";
    for (size_t t = 0; t < m.average.size(); ++t)
        f << t << "," << m.average[t] << "," << m.error[t] << "," << m.meff[t] << "," << m.meff_error[t] << "\n";
}

}  // namespace app
