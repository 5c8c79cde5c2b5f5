// dwf_internal.hpp -- types shared by the gfx950 kernels and the host driver of
// libdwfhip.so.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dwfk {

constexpr int kBlock = 256;  // threads per block for every stencil/BLAS kernel

// Per right-hand-side CG state.  Fields a kernel reads are never written by the same
// launch: scalars the next iteration needs live in slot (it+1)&1 while iteration `it`
// reads slot it&1, so blocks of one launch never race.
// F_HOLD marks, during one reliable update, the right-hand sides whose FP32 cycle is still
// running: that update skips them and clears the mark.
// F_FRESH: the first iteration of a pipelined even-odd CG (its direction is the residual).
enum : int { F_DONE = 1, F_FROZEN = 2, F_HOLD = 4, F_FRESH = 8 };
enum : int { MODE_CG64 = 0, MODE_MIXED = 1 };

struct CgState {
    double rr[2];       // r.r of the current residual, by slot
    double bnorm2;      // b.b
    double rr_ref;      // mixed: r.r at the last reliable update
    double relres;      // last recursive ||r||/||b||
    double true_relres; // last FP64 ||b-Ax||/||b||
    double tol;         // relative tolerance
    double delta;       // mixed: reliable update when ||r|| < delta ||r_ref||
    int flags[2];       // F_DONE / F_FROZEN, by slot
    int iters;          // CG iterations performed in this precision
    int status;         // 0, or 4 for a zero denominator
    int reliable;       // reliable updates performed
    int fp64_iters;     // FP64 iterations (fallback)
    // pipelined even-odd CG (the direction update inside the next operator launch):
    double beta;        // r.r / r.r_prev of the last iteration, for the next direction
    int pbuf[2];        // by slot: which of the two direction buffers holds the direction an
                        // iteration of that slot starts from
};

// Lattice and operator constants.  Ls = 1 and d5 = 0 give the Wilson operator.
struct Lat {
    int Nt, Nx, Ls, V4;
    int d5;             // 1: Shamir fifth-dimension hops with wall coupling mf
    double diag;        // 3 - M5 (DWF) or 2 + m0 (Wilson)
    double mf;
    double bct, bcx;    // boundary signs across t = Nt-1 -> 0 and x = Nx-1 -> 0
    __host__ __device__ long long V5() const { return (long long)V4 * Ls; }
};

// ------------------------------------------------------------------ lattice groups
// A batch of right-hand sides on several lattices at once (the batched analyzer over
// configurations of different Nt x Nx: every kernel launch covers all of them).  Group g
// holds right-hand sides [rhs0[g], rhs0[g+1]) on an Nt[g] x Nx[g] x Ls lattice.  Their
// vectors follow one another, after the vectors of all earlier groups: site0[g] 5D sites
// of the full lattice precede the group's first vector (a half-lattice vector batch has
// half of every count).  Right-hand side r of group g uses configuration
// (r - rhs0[g]) / rpc[g] of that group, whose links start link0[g] complex entries into
// the link array.  A context's single lattice is one group with site0 = link0 = 0.
constexpr int kMaxGroups = 8;
struct Groups {
    int n;                       // groups (1..kMaxGroups)
    int rhs0[kMaxGroups + 1];    // first right-hand side of each group; rhs0[n] = nrhs
    int Nt[kMaxGroups], Nx[kMaxGroups];
    int rpc[kMaxGroups];         // right-hand sides per configuration
    long long site0[kMaxGroups]; // full-lattice 5D sites before the group's first vector
    long long link0[kMaxGroups]; // complex link entries before the group's first configuration
};

// Scalar scan (n <= kMaxGroups, uniform over the block).
__host__ __device__ __forceinline__ int group_of_rhs(const Groups& G, int rhs)
{
    int g = 0;
    for (int k = 1; k < G.n; ++k) g = rhs >= G.rhs0[k] ? k : g;
    return g;
}

// One group's fields as scalars.
struct GroupPos {
    int g, Nt, Nx, rhs0, rpc;
    long long site0, link0;
};
__host__ __device__ __forceinline__ GroupPos group_fields(const Groups& G, int g)
{
    return GroupPos{g, G.Nt[g], G.Nx[g], G.rhs0[g], G.rpc[g], G.site0[g], G.link0[g]};
}
__host__ __device__ __forceinline__ GroupPos group_for_rhs(const Groups& G, int rhs)
{
    if (G.n == 1) return group_fields(G, 0);
    return group_fields(G, group_of_rhs(G, rhs));
}

// The group of block lb of a launch over several lattices whose group g gives each of its
// right-hand sides bpr(Nt[g], Nx[g]) consecutive blocks, the groups one after another.
// Returns the group's fields, the group's first block and its blocks per right-hand side.
// The table is the context's device copy (one kernel argument, read through the scalar
// cache), so single-lattice launches keep their small argument structs.
template <typename F>
__device__ __forceinline__ GroupPos block_group(const Groups* G, int lb, F bpr, int& blk0, int& per_rhs)
{
    const int n = G->n;
    int start = 0, g = n - 1, nb = 0;
    for (int k = 0; k < n; ++k) {
        nb = bpr(G->Nt[k], G->Nx[k]);
        const int cnt = (G->rhs0[k + 1] - G->rhs0[k]) * nb;
        if (lb < start + cnt) {
            g = k;
            break;
        }
        start += cnt;
    }
    blk0 = start;
    per_rhs = nb;
    return group_fields(*G, g);
}

// Side of the fused even-odd kernel's tile along a lattice extent n: the largest even
// divisor of n up to maxd (0: none).
__host__ __device__ __forceinline__ int eo_tile_dim(int n, int maxd)
{
    for (int c = maxd; c >= 2; c -= 2)
        if (n % c == 0) return c;
    return 0;
}

// Extent and offset (in sites) of right-hand side rhs's vector in a batch over groups;
// shift 0 for full-lattice vectors, 1 for half-lattice (even-odd) ones.
struct VecLayout {
    Groups gr;
    int Ls;
    int shift;
    const Groups* dev;  // the context's device copy of gr (several groups), for the hot kernels
    __host__ __device__ long long sites(int g) const { return ((long long)Ls * gr.Nt[g] * gr.Nx[g]) >> shift; }
    __host__ __device__ long long offset(int g, int rhs) const
    {
        return (gr.site0[g] + (long long)(rhs - gr.rhs0[g]) * Ls * gr.Nt[g] * gr.Nx[g]) >> shift;
    }
    // (extent, offset) of right-hand side rhs, with the one-group fast path
    __host__ __device__ void locate(int rhs, long long& n, long long& off) const
    {
        const GroupPos p = group_for_rhs(gr, rhs);
        const long long v5 = (long long)Ls * p.Nt * p.Nx;
        n = v5 >> shift;
        off = (p.site0 + (long long)(rhs - p.rhs0) * v5) >> shift;
    }
    __host__ __device__ long long total_sites() const
    {
        long long t = 0;
        for (int g = 0; g < gr.n; ++g) t += (long long)(gr.rhs0[g + 1] - gr.rhs0[g]) * sites(g);
        return t;
    }
};

// Launch shape of the stencil kernels: one thread per 4D site per s-chunk; each thread
// walks spt consecutive slices reusing its four links.
struct HopCfg {
    int bx;      // blocks along the 4D sites
    int chunks;  // s-chunks
    int spt;     // slices per thread
    int remap;   // 1: XCD-aware block order (contiguous tiles per XCD)
    int prefetch;  // 1: software-pipelined slice loop (next slice's loads in flight)
    int tx;        // 0: block = 256 consecutive 4D sites; else a 2D tile of (256/tx) rows x tx columns
    int tiles_x;   // Nx / tx
    int nt_store;  // 1: streaming (non-temporal) stores of the stencil output
    int lds;       // 1..6: LDS-staged slice tiles of 256 / 512 / 64 / 128 / 384 / 768 threads
    __host__ __device__ int nblk() const { return bx * chunks; }
};

// Logical block id for a 1D grid of `total` blocks.  Workgroups are dealt round-robin
// to the 8 XCDs, each with its own L2; this bijection hands every XCD a contiguous run
// of logical blocks, so neighbouring tiles (which share halo rows) share one L2.
__device__ __forceinline__ int xcd_logical_block(int hw, int total)
{
    constexpr int kXcd = 8;
    const int xcd = hw % kXcd, local = hw / kXcd;
    const int per = (total + kXcd - 1) / kXcd;
    int tall = total % kXcd;
    if (tall == 0) tall = kXcd;
    return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}

// Scalars of the learned solver's preconditioned CG, per right-hand side.
struct FcgScal {
    double rz, pq, alpha, beta;
    double rz_prev;  // (z, r) of the previous iteration (Fletcher-Reeves beta)
    double pbuf;     // direction buffer (0 / 1) the last direction update wrote
};

enum Epi : int {
    EPI_STORE = 0,  // out = Op in
    EPI_NORM = 1,   // out = Op in, partial |out|^2
    EPI_CGUPD = 2,  // alpha = rr/sum(partA); x += alpha p (unless x is null); r -= alpha (Op in); partial
                    // |r|^2; optionally out = Op in and aout = alpha
    EPI_RESID = 3,  // r64 = b - Op in (+ copies to p64/r32/p32); partials |r|^2, |b|^2
    EPI_PNORM = 4,  // FP32, learned solver: p = z + beta p_old (beta from zpart, fs), pnew = p,
                    // out = Op p, partial |out|^2; the previous iteration's x += alpha p_old
    EPI_RUPD = 5    // FP32, learned solver: EPI_CGUPD without x and p (r -= alpha (Op in), |r|^2,
                    // aout = alpha)
};

template <typename R> struct HopArgs {
    Lat L;               // the lattice of a single-group launch (Ls, masses and signs of every group)
    HopCfg cfg;          // launch shape; with several groups: linear blocks, bx per group in gb
    const void* links;   // [ncfg][2][V4] complex R: (cos theta, sin theta)
    long long link_stride;  // complex entries per configuration (2 V4)
    int rhs_per_cfg;     // right-hand side r uses configuration r / rhs_per_cfg ...
    const int* cfg_map;  // ... or cfg_map[r] when not null
    const R* in;         // [nrhs][V5] spinors of 4 R
    R* out;              // EPI_STORE / EPI_NORM
    R* x;                // EPI_CGUPD (nullable: no x update, p unused); EPI_PNORM (nullable): see below
    R* r;                // EPI_CGUPD
    const R* p;          // EPI_CGUPD
    const double* partA; // EPI_CGUPD: partials of p^dag A p from the first stencil
    const double* anum;  // EPI_CGUPD (nullable): alpha's numerator is anum[rhs * anum_stride] instead of
    int anum_stride;     //   st[rhs].rr[slot] (the flexible CG's (z, r)); out (nullable) receives Op in
    double* aout;        // EPI_CGUPD (nullable): block 0 stores alpha (0 when p^dag A p is not positive)
                         //   at aout[rhs * anum_stride]
    const double* b;     // EPI_RESID
    double* r64;         // EPI_RESID
    double* p64;         // EPI_RESID (nullable)
    float* r32;          // EPI_RESID (nullable)
    float* p32;          // EPI_RESID (nullable)
    double* part;        // partial sums, 2 per block, [nrhs][nblk][2]
    // EPI_PNORM: in = z, pold -> pnew; beta = (z, r) / rz_prev (fr) or -(z, q) / pq from the
    // partials zpart [nrhs][nz][2] of (z, r), (z, q) and fs; a zero beta takes p = z (a cycle's
    // first direction: rz_prev, pq set to +inf).  (z, r) <= 0 marks a breakdown in stw.
    // With x set, the previous iteration's update x += alpha p_old (alpha = fs[rhs].alpha,
    // stored by that iteration's EPI_CGUPD; 0: none) is applied here, where p_old is read
    // anyway; block 0 records pbuf, the index of pnew's buffer, in fs[rhs].pbuf.
    const R* pold;
    R* pnew;
    const double* zpart;
    int nz;
    int fr;
    FcgScal* fs;
    CgState* stw;
    int pbuf;
    const CgState* st;   // nullable
    int slot;
    int skip;            // skip a right-hand side whose flags[slot] & skip
    // Several lattices (null for one): the context's device copy of the call's lattice
    // groups, and the partial arrays' stride (the largest group's blocks per right-hand
    // side; a group with fewer leaves the rest of its partials at the zero the solvers put
    // there).  A multi-lattice launch uses linear blocks (cfg.tx = 0, no LDS tiles).
    const Groups* groups;
    int pstride;
};

// Host launchers (dwf_device.hip).
void launch_pack_links(long long nlinks, const double* theta, double2* l64, float2* l32, hipStream_t s);
// hg: the host copy of a.groups' table for a multi-lattice launch (it sizes the grid), else null.
template <typename R>
void launch_hop(bool dagger, int epi, const HopArgs<R>& a, int nrhs, hipStream_t s, const Groups* hg = nullptr);
// The single-group layout of nrhs vectors of L.
Groups single_group(const Lat& L, int nrhs, int rhs_per_cfg);
// Block partials per right-hand side of the direct-load stencil over groups (linear
// blocks, cfg.chunks s-chunks): HopArgs::pstride of its multi-lattice launches.
int hop_pstride(const Groups& G, const HopCfg& cfg);
template <typename R>
void launch_cg_p(const VecLayout& lay, int nblk, R* r, R* p, const double* partA, const double* partB, CgState* st,
                 int slot, int mode, int nrhs, hipStream_t s);
void launch_cg_start(int nblk, const double* partR, CgState* st, double tol, double delta, int nrhs, hipStream_t s);
void launch_ru_commit(int nblk, const double* partR, CgState* st, int slot, int nrhs, hipStream_t s);
// k_cg_p's state update alone (the pipelined even-odd CG forms the direction itself), and
// the pipelined CG's start (F_FRESH, direction buffer 0).
void launch_cg_flags(int nblk, const double* partA, const double* partB, CgState* st, int slot, int mode, int nrhs,
                     hipStream_t s);
void launch_set_fresh(CgState* st, int nrhs, hipStream_t s);
void launch_true_resid(int nblk, const double* partR, CgState* st, int nrhs, hipStream_t s);
void launch_accum_x(const VecLayout& lay, double* x64, float* x32, const CgState* st, int slot, int skip, int nrhs,
                    hipStream_t s);
// Marks F_HOLD on the right-hand sides with hold[i] != 0 (device array) that are not done.
void launch_hold(CgState* st, const int* hold, int slot, int nrhs, hipStream_t s);
template <typename R> void launch_zero(R* p, long long n, hipStream_t s);
void launch_set_frozen_all(CgState* st, int slot, int nrhs, hipStream_t s);

HopCfg hop_cfg(const Lat& L, int nrhs, int chunks_override = 0, int remap = 1, int prefetch = 0, int tile_x = 0,
              int nt_store = 0, int lds = 0);

// HMC kernels (dwf_hmc.hip).  `scalars` is a small device array of FP64 results; each
// reduction folds its block partials in index order into scalars[slot].
int reduce_blocks(long long n);
void launch_gauge(const Lat& L, const double* th, double beta, double* force, double* part_s,
                  double* part_c, double* scalars, int slot_s, int slot_c, hipStream_t s);
void launch_pf_force(const Lat& L, const double2* l64, const double* eta, const double* xi, double* force,
                     hipStream_t s);
void launch_redot(const double* a, const double* b, long long ncomplex, double* part, double* scalars, int slot,
                  hipStream_t s);
void launch_mom(double* p, const double* fa, const double* fb, int n, double c, double* part_k, double* scalars,
                int slot, hipStream_t s);
void launch_theta(double* th, const double* p, int n, double eps, int wrap, double2* l64, float2* l32,
                  hipStream_t s);

}  // namespace dwfk
