// dwf_eo.hpp -- arguments and launchers of the even-odd Schur-complement kernels.
#pragma once

#include "dwf_internal.hpp"

namespace dwfk {

constexpr int kBlockA = 64;  // k_eo_a owns a whole s-column per thread: small blocks, more of them

template <typename R> struct EoArgs {
    Lat L;
    int Nxh, V4h;        // Nx/2, V4/2
    int bx;              // k_eo_b: blocks of kBlock threads over the V4h target sites
    int bxa;             // k_eo_a: blocks of kBlockA threads over the V4h target sites
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
    // Several lattices (null for one): the context's device copy of the call's lattice
    // groups (Nt, Nx, Nxh, V4h, bx and bxa then follow from each group) and the partial
    // arrays' stride (eo_pstride).
    const Groups* groups;
    int pstride;
    // pipelined even-odd CG (k_eo_schur only; null otherwise): the two direction buffers,
    // CgState::pbuf naming the current one
    R* pd0;
    R* pd1;
    // pipelined even-odd CG, the direction-forming launch: with fold set, it also does the
    // previous iteration's state update (k_cg_flags) from that iteration's partials
    // (partAp: its operator's |t|^2, partB: its r.r) into slot `slot` of stw (block 0 of each
    // right-hand side writes; every block computes the same values)
    int fold;
    const double* partAp;
    const double* partB;
    CgState* stw;
};

// Tile of the fused Schur-operator kernel (k_eo_schur): ty x tx 4D sites (both even,
// ty <= 8, tx <= 16, dividing every lattice's Nt and Nx) per block of kBlockS threads,
// its target-parity sites (at most 64, one per lane) walked in nch = 4 s-chunks (one per
// wavefront) of spt = Ls / 4 slices.
struct EoTile {
    int ty, tx, nch, spt;
};
constexpr int kBlockS = 256;

bool eo_supported(const Lat& L);
// The largest tile for the lattices of G (false: none fits).
bool eo_fused_tile(const Groups& G, int Ls, EoTile& t);
// Block partials per right-hand side of k_eo_schur over the groups of G.
int eo_fused_pstride(const Groups& G, const EoTile& t);
// out_o = a.ca [A | A^dag] a.in + a.ch H w, w = [A^-1 | A^-dag] H a.in (the Schur operator
// Dhat or Dhat^dag for ca = 1, ch = -1/4), with the epilogue epi of launch_eo_b; one
// partial pair per tile.
template <typename R>
void launch_eo_schur(bool dagger, int epi, const EoArgs<R>& a, const EoTile& t, int nrhs, hipStream_t s,
                     const Groups* hg = nullptr);
// The pipelined even-odd CG's operator launch (FP32): the direction p = r + beta p_old
// (p = r on an F_FRESH iteration; p_old = a.pd[pbuf], p written to a.pd[pbuf ^ 1]) formed
// on the fly from a.r, then out = Dhat p with |out|^2 partials.  With a.pd0 set, the
// EPI_CGUPD launch (Dhat^dag) takes that new direction for x += alpha p.
void launch_eo_schur_pipe(const EoArgs<float>& a, const EoTile& t, int nrhs, hipStream_t s, const Groups* hg = nullptr);
// Block partials per right-hand side of k_eo_b over groups with `chunks` s-chunks.
int eo_pstride(const Groups& G, int chunks);
// hg: the host copy of a.groups' table for a multi-lattice launch (it sizes the grid), else null.
template <typename R>
void launch_eo_a(bool dagger, bool ainv, const EoArgs<R>& a, int nrhs, hipStream_t s, const Groups* hg = nullptr);
template <typename R>
void launch_eo_b(bool dagger, bool aop, int epi, const EoArgs<R>& a, int nrhs, hipStream_t s,
                 const Groups* hg = nullptr);
// full lexicographic vectors <-> even / odd halves, for the batch layout of lay (full-lattice
// offsets: shift 0)
void launch_eo_split(const VecLayout& lay, const double* full, double* ev, double* od, int nrhs, hipStream_t s);
void launch_eo_merge(const VecLayout& lay, const double* ev, const double* od, double* full, int nrhs, hipStream_t s);

}  // namespace dwfk
