// dwf_eo.hpp -- arguments and launchers of the even-odd Schur-complement kernels.
#pragma once

#include "dwf_internal.hpp"

namespace dwfk {

constexpr int kBlockA = 64;  // k_eo_a owns a whole s-column per thread: small blocks, more of them

template <typename R> struct EoArgs {
    Lat L;               // Ls, masses and signs (Nt, Nx per group in gr)
    Groups gr;           // right-hand sides by lattice (n = 0: one group from L, nrhs and rhs_per_cfg)
    GroupBlocks gb;      // set by the launchers: k_eo_b blocks of kBlock target sites per s-chunk,
                         // k_eo_a blocks of kBlockA target sites (whole s-columns)
    int chunks, spt;     // s-chunks (k_eo_b only)
    int remap;           // XCD-aware block order
    int tpar;            // target parity (0 even, 1 odd)
    const void* links;   // [ncfg][2][V4] complex R
    long long link_stride;  // complex entries per configuration (2 V4)
    int rhs_per_cfg;     // right-hand side r uses configuration r / rhs_per_cfg
    const R* in;         // k_eo_b: psi on the target parity (A term / identity term)
    const R* hin;        // the other-parity vector the hop reads
    const R* b;          // k_eo_a: optional target-parity source; k_eo_b EPI_RESID: rhs
    R* out;
    R cb, ca, ch;        // coefficients of b, of the A term, of the hop
    R* x;                // EPI_CGUPD
    R* r;
    const R* p;
    const double* partA;
    double* r64;         // EPI_RESID
    double* p64;
    float* r32;
    float* p32;
    double* part;        // partials, 2 per block
    const CgState* st;
    int slot;
    int skip;
};

bool eo_supported(const Lat& L);
// Block partials per right-hand side of k_eo_b over groups with `chunks` s-chunks.
int eo_pstride(const Groups& G, int chunks);
template <typename R> void launch_eo_a(bool dagger, bool ainv, const EoArgs<R>& a, int nrhs, hipStream_t s);
template <typename R> void launch_eo_b(bool dagger, bool aop, int epi, const EoArgs<R>& a, int nrhs, hipStream_t s);
// full lexicographic vectors <-> even / odd halves, for the batch layout of lay (full-lattice
// offsets: shift 0)
void launch_eo_split(const VecLayout& lay, const double* full, double* ev, double* od, int nrhs, hipStream_t s);
void launch_eo_merge(const VecLayout& lay, const double* ev, const double* od, double* full, int nrhs, hipStream_t s);

}  // namespace dwfk
