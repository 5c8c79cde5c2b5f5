// dwf_learned.hip -- the learned preconditioner of libdwfhip (no counterpart in the
// reference, which solves with unpreconditioned CG: domain_wall/include/cg_solver.hpp).
//
// Network.  Three linear parallel-transport convolution layers, 2 -> 8 -> 8 -> 2
// complex channels per 5D site.  Layer output
//     y(s,t,x) = sum_p W_p(s) U_p(t,x) h(neighbour_p(s,t,x)),
// p over the 7 points self, +t, -t, +x, -x, +s, -s; U_p the Dirac operator's transporter
// (antiperiodic time boundary), +-s periodic without link; W_p(s) complex C_out x C_in,
// s-dependent.  Linear and gauge covariant; M ~ (D^dag D)^-1 after training.
//
// Matrix cores.  Per layer the work is a GEMM [sites] x [7 * 2 C_in] x [2 C_out] in the
// real form of the complex products (re/im interleaved per channel: row k = 2c + im,
// column n = 2o + im).  One wavefront owns a tile of 16 consecutive 4D sites at fixed
// (rhs, s) and issues v_mfma_f32_16x16x16_bf16 per stencil point (C_in = 8) or per 4
// stencil points (C_in = 2): the A operand is the neighbour field transported in FP32
// registers and rounded to BF16, the B operand the BF16 weight fragment of that s (kept
// in registers while the wave walks tiles of one slice), FP32 accumulation.  Hidden
// activations are stored as BF16 (32 bytes per site), inputs and outputs are the
// solver's spinors.
//
// Lane layout of v_mfma_f32_16x16x16_bf16 (64 lanes, lane = 16 g + m):
//   A[i][k]: lane supplies A[m][4g + j], j = 0..3;  B[k][n]: lane supplies B[4g + j][m];
//   D[i][n]: lane holds D[4g + j][m].
//
// Training (dwf_net_train_device): forward, loss gradient, adjoint layers (the same
// kernel with conjugate-transposed, point-reversed weight fragments), weight gradients
// as MFMA outer-product sums G = sum_sites a^T g with deterministic block partials,
// Adam on FP32 master weights, re-packing of the BF16 fragments -- all on the device.
//
// Solver (DWF_SOLVE_LEARNED): flexible CG (Notay's FCG(1)) on D^dag D in FP64 with
// z = M r, certified by the FP64 true residual.

#include "../../include/dwf_hip.h"
#include "dwf_context.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

using namespace dwfk;

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

struct dwf_net {
    int Ls = 0;
    int device = 0;
    float2* w[3] = {nullptr, nullptr, nullptr};   // FP32 master weights [Ls][7][Cout][Cin]
    float2* m1[3] = {nullptr, nullptr, nullptr};  // Adam first moments
    float2* m2[3] = {nullptr, nullptr, nullptr};  // Adam second moments
    unsigned short* wf[3] = {nullptr, nullptr, nullptr};  // BF16 B fragments, forward
    unsigned short* wa[3] = {nullptr, nullptr, nullptr};  // BF16 B fragments, adjoint
    unsigned short* wg0 = nullptr;  // layer 1's fragments in the gauge-fixed fused kernel's point order
    int adam_t = 0;
    hipStream_t pack_stream = nullptr;  // for set_weights outside a context
    ~dwf_net();
};

namespace {

constexpr int kPts = 7;
constexpr int kCinL[3] = {2, 8, 8};
constexpr int kCoutL[3] = {8, 8, 2};

__host__ __device__ constexpr int nstep_of(int cin) { return cin == 8 ? 7 : 2; }
long long layer_cw(int l, int Ls) { return (long long)Ls * kPts * kCoutL[l] * kCinL[l]; }  // complex weights
long long frag_shorts(int cin, int Ls) { return (long long)Ls * nstep_of(cin) * 64 * 4; }

__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite values)
    return (unsigned short)(u >> 16);
}

// Two floats to packed bf16 (low, high) with gfx950's v_cvt_pk_bf16_f32 (round to
// nearest even, the same rounding as f2bf for finite values).
typedef __bf16 bf16x2_hw __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pk_bf16(float lo, float hi)
{
    const f32x2_t v = {lo, hi};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2_hw));
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

struct PtcArgs {
    int Nt, Nx, Ls, V4;
    int tps;                   // 16-site tiles per (rhs, s) slice: V4 / 16
    long long ntiles;          // nrhs * Ls * tps
    const void* in;            // layer input
    void* out;                 // layer output (forward kernel)
    const int* in_row;         // nullable: right-hand side -> row of `in`
    const int* cfg;            // right-hand side -> configuration of the links
    const float2* links;       // [cfg][2][V4] (cos, sin)
    long long link_stride;     // entries per configuration
    const unsigned short* wpk; // B fragments [Ls][nstep][64][4]
    const CgState* st;         // nullable: skip right-hand sides with flags[0] & F_DONE
    int tpw;                   // tiles per wavefront
};

// Neighbour (slice, 4D site) and FP32 transporter of stencil point p at (s, t, x).
struct Nb {
    int s, i;
    float c, sn;
};

__device__ __forceinline__ Nb neighbour(const PtcArgs& a, const float2* L, int p, int s, int t, int x)
{
    Nb r{s, t * a.Nx + x, 1.f, 0.f};
    switch (p) {
        case 1: {
            const int t2 = t + 1 == a.Nt ? 0 : t + 1;
            const float2 u = L[t * a.Nx + x];
            const float sg = t == a.Nt - 1 ? -1.f : 1.f;
            r.i = t2 * a.Nx + x;
            r.c = sg * u.x;
            r.sn = sg * u.y;
            break;
        }
        case 2: {
            const int t2 = t == 0 ? a.Nt - 1 : t - 1;
            const float2 u = L[t2 * a.Nx + x];
            const float sg = t == 0 ? -1.f : 1.f;
            r.i = t2 * a.Nx + x;
            r.c = sg * u.x;
            r.sn = sg * -u.y;
            break;
        }
        case 3: {
            const int x2 = x + 1 == a.Nx ? 0 : x + 1;
            const float2 u = L[a.V4 + t * a.Nx + x];
            r.i = t * a.Nx + x2;
            r.c = u.x;
            r.sn = u.y;
            break;
        }
        case 4: {
            const int x2 = x == 0 ? a.Nx - 1 : x - 1;
            const float2 u = L[a.V4 + t * a.Nx + x2];
            r.i = t * a.Nx + x2;
            r.c = u.x;
            r.sn = -u.y;
            break;
        }
        case 5: r.s = s + 1 == a.Ls ? 0 : s + 1; break;
        case 6: r.s = s == 0 ? a.Ls - 1 : s - 1; break;
        default: break;
    }
    return r;
}

// Transport in bf16 dot products (v_dot2_f32_bf16): the field value h = (hr, hi) and the
// transporter u = (c, sn) are bf16, the two products are exact in FP32 and their sum is
// rounded once.  lk.x = (c, -sn), lk.y = (sn, c) as bf16 pairs (low half first), so
// Re(u h) = dot2(h, lk.x) and Im(u h) = dot2(h, lk.y).  A point without a link has
// u = 1, for which both products are exact: the value passes through unchanged.
__device__ __forceinline__ uint2 link_ops(const Nb& n)
{
    return make_uint2(pk_bf16(n.c, -n.sn), pk_bf16(n.sn, n.c));
}
__device__ __forceinline__ uint2 link_ops_f2(float2 u) { return make_uint2(pk_bf16(u.x, -u.y), pk_bf16(u.y, u.x)); }
__device__ __forceinline__ float dot2(unsigned h, unsigned l)
{
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_hw, h), __builtin_bit_cast(bf16x2_hw, l), 0.0f,
                                           false);
}
// u h for one complex channel (bf16 pair in, bf16 pair out)
__device__ __forceinline__ unsigned transport(unsigned h, const uint2& lk) { return pk_bf16(dot2(h, lk.x), dot2(h, lk.y)); }

__device__ __forceinline__ float4 ld_spinor(const double* p, long long site)
{
    const double4 v = reinterpret_cast<const double4*>(p)[site];
    return make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}
__device__ __forceinline__ float4 ld_spinor(const float* p, long long site)
{
    return reinterpret_cast<const float4*>(p)[site];
}
__device__ __forceinline__ void st_spinor(double* p, long long site, float a, float b, float c, float d)
{
    reinterpret_cast<double4*>(p)[site] = make_double4(a, b, c, d);
}
__device__ __forceinline__ void st_spinor(float* p, long long site, float a, float b, float c, float d)
{
    reinterpret_cast<float4*>(p)[site] = make_float4(a, b, c, d);
}

__device__ __forceinline__ bf16x4_t pack4(float a, float b, float c, float d)
{
    const unsigned lo = pk_bf16(a, b), hi = pk_bf16(c, d);
    bf16x4_t v;
    v[0] = (short)(lo & 0xffffu);
    v[1] = (short)(lo >> 16);
    v[2] = (short)(hi & 0xffffu);
    v[3] = (short)(hi >> 16);
    return v;
}

// ------------------------------------------------------------------ forward layer
// Each MFMA computes D' = W^T H^T for one 16-site tile: the A operand is the weight
// fragment (lane (m, g) holds W[in 4g + j][out m], j = 0..3), the B operand the input
// (lane (m, g) holds input components 4g .. 4g + 3 of site m), and the result leaves
// lane (m, g) with output components 4g .. 4g + 3 of site m, i.e. the (re, im) pairs of
// output channels 2g and 2g + 1.
//
// C_in = 8 layers transport after the product: W_p (U_p h) = U_p (W_p h) because U_p is a
// complex scalar per site, so each hop is one MFMA of the stored bf16 values followed by
// the FP32 complex multiply-accumulate by U_p of the output site.  The order is fixed
// (it makes the layer kernel and the fused network bit-identical): the MFMA chain of the
// points without a link (self, +s, -s), then the hops +t, -t, +x, -x.
// C_in = 2 layers pack four stencil points into one MFMA, so they transport before the
// product, in bf16 dot products (transport()).

// bf16 pairs (re, im) of two complex values, as an MFMA operand
__device__ __forceinline__ bf16x4_t pack_pairs(unsigned lo, unsigned hi)
{
    return __builtin_bit_cast(bf16x4_t, make_uint2(lo, hi));
}

// y += u G for the two complex values of a lane (FP32, fixed fma order)
__device__ __forceinline__ void cmac2(f32x4_t& y, const f32x4_t& G, float c, float sn)
{
    y[0] = fmaf(c, G[0], y[0]);
    y[0] = fmaf(-sn, G[1], y[0]);
    y[1] = fmaf(sn, G[0], y[1]);
    y[1] = fmaf(c, G[1], y[1]);
    y[2] = fmaf(c, G[2], y[2]);
    y[2] = fmaf(-sn, G[3], y[2]);
    y[3] = fmaf(sn, G[2], y[3]);
    y[3] = fmaf(c, G[3], y[3]);
}

// One 16-site tile of a C_in = 8 layer: h[q] the stored values at stencil point q's
// neighbour, u[p - 1] the transporter (cos, sin) of hop p = 1..4 at the output site.
__device__ __forceinline__ f32x4_t tile8(const bf16x4_t (&W)[7], const bf16x4_t (&h)[7], const float2 (&u)[4])
{
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    f32x4_t y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[0], h[0], zero, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[5], h[5], y, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[6], h[6], y, 0, 0, 0);
    f32x4_t G[4];
#pragma unroll
    for (int p = 1; p <= 4; ++p) G[p - 1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[p], h[p], zero, 0, 0, 0);
#pragma unroll
    for (int p = 1; p <= 4; ++p) cmac2(y, G[p - 1], u[p - 1].x, u[p - 1].y);
    return y;
}

// tile8 in a gauge where the +-t transporters are 1: those points chain into the
// accumulator like self and +-s; u[0], u[1] are the +x, -x transporters.
__device__ __forceinline__ f32x4_t tile8_tg(const bf16x4_t (&W)[7], const bf16x4_t (&h)[7], const float2 (&u)[2])
{
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    f32x4_t y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[0], h[0], zero, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[5], h[5], y, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[6], h[6], y, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[1], h[1], y, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[2], h[2], y, 0, 0, 0);
    const f32x4_t G3 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[3], h[3], zero, 0, 0, 0);
    const f32x4_t G4 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[4], h[4], zero, 0, 0, 0);
    cmac2(y, G3, u[0].x, u[0].y);
    cmac2(y, G4, u[1].x, u[1].y);
    return y;
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Operands of a C_in = 8 tile at the lane's site (t, x) of slice s (channels 2g, 2g+1).
__device__ __forceinline__ void gather8(const PtcArgs& a, const float2* L, const void* inb, int s, int t, int x, int g,
                                        bf16x4_t (&h)[7], float2 (&u)[4])
{
#pragma unroll
    for (int q = 0; q < kPts; ++q) {
        const Nb n = neighbour(a, L, q, s, t, x);
        const uint2 v = *reinterpret_cast<const uint2*>(static_cast<const unsigned short*>(inb) +
                                                        ((long long)n.s * a.V4 + n.i) * 16 + 4 * g);
        h[q] = pack_pairs(v.x, v.y);
        if (q >= 1 && q <= 4) u[q - 1] = make_float2(n.c, n.sn);
    }
}

// B operand of MFMA step q of a C_in = 2 layer for the lane's site (t, x) of slice s:
// stencil point 4q + g's two channels, transported in bf16.
template <typename TIN>
__device__ __forceinline__ bf16x4_t build_B2(const PtcArgs& a, const float2* L, const void* inb, int q, int s, int t,
                                             int x, int g)
{
    const int p = 4 * q + g;  // this lane group's stencil point
    bf16x4_t B = {0, 0, 0, 0};
    if (p < kPts) {
        const Nb n = neighbour(a, L, p, s, t, x);
        const float4 v = ld_spinor(static_cast<const TIN*>(inb), (long long)n.s * a.V4 + n.i);
        const uint2 lk = link_ops(n);
        B = pack_pairs(transport(pk_bf16(v.x, v.y), lk), transport(pk_bf16(v.z, v.w), lk));
    }
    return B;
}

// Output components 4g .. 4g + 3 of site m of one 16-site tile at obase.
template <int COUT, typename TOUT>
__device__ __forceinline__ void store_tile(const PtcArgs& a, const f32x4_t& acc, long long obase, int m, int g)
{
    if constexpr (COUT == 8) {
        unsigned short* out = static_cast<unsigned short*>(a.out) + (obase + m) * 16 + 4 * g;
        *reinterpret_cast<uint2*>(out) = make_uint2(pk_bf16(acc[0], acc[1]), pk_bf16(acc[2], acc[3]));
    } else {
        if (g == 0) st_spinor(static_cast<TOUT*>(a.out), obase + m, acc[0], acc[1], acc[2], acc[3]);
    }
}

// TIN/TOUT: double or float for spinors (2 channels), unsigned short for 8 BF16 channels.
// Two adjacent tiles of one slice are processed together when they exist, so the loads
// of both are in flight before either MFMA chain starts.
template <int CIN, int COUT, typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void k_ptc_fwd(const PtcArgs a)
{
    constexpr int NSTEP = nstep_of(CIN);
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long long T0 = wave * a.tpw;
    const long long T1 = T0 + a.tpw < a.ntiles ? T0 + a.tpw : a.ntiles;
    const long long per_rhs = (long long)a.Ls * a.tps;
    constexpr int ch = CIN == 8 ? 16 : 4;  // input values per site
    int cur_s = -1, cur_rhs = -1;
    bool skip = false;
    const float2* L = a.links;
    const void* inb = a.in;
    bf16x4_t W[NSTEP];
    for (long long T = T0; T < T1;) {
        const int rhs = (int)(T / per_rhs);
        const int rem = (int)(T - rhs * per_rhs);
        const int s = rem / a.tps;
        const int tl = rem - s * a.tps;
        const int i0 = tl * 16;
        const bool two = T + 1 < T1 && tl + 1 < a.tps;  // the next tile is in the same slice
        if (rhs != cur_rhs) {  // per right-hand side lookups, once per run of its tiles
            skip = a.st && (a.st[rhs].flags[0] & (F_DONE | F_FROZEN));
            const int row = a.in_row ? a.in_row[rhs] : rhs;
            L = a.links + (long long)a.cfg[rhs] * a.link_stride;
            if constexpr (CIN == 8)
                inb = static_cast<const unsigned short*>(a.in) + (long long)row * a.Ls * a.V4 * ch;
            else
                inb = static_cast<const TIN*>(a.in) + (long long)row * a.Ls * a.V4 * ch;
            cur_rhs = rhs;
        }
        T += two ? 2 : 1;
        if (skip) continue;  // uniform per wavefront
        if (s != cur_s) {
            const bf16x4_t* wp = reinterpret_cast<const bf16x4_t*>(a.wpk) + (long long)s * NSTEP * 64 + lane;
#pragma unroll
            for (int q = 0; q < NSTEP; ++q) W[q] = wp[q * 64];
            cur_s = s;
        }
        const long long obase = ((long long)rhs * a.Ls + s) * a.V4 + i0;
        const int ia = i0 + m, ta = ia / a.Nx, xa = ia - ta * a.Nx;
        const int ib = ia + 16, tb = ib / a.Nx, xb = ib - tb * a.Nx;
        if constexpr (CIN == 8) {
            bf16x4_t h0[7], h1[7];
            float2 u0[4], u1[4];
            gather8(a, L, inb, s, ta, xa, g, h0, u0);
            if (two) gather8(a, L, inb, s, tb, xb, g, h1, u1);
            store_tile<COUT, TOUT>(a, tile8(W, h0, u0), obase, m, g);
            if (two) store_tile<COUT, TOUT>(a, tile8(W, h1, u1), obase + 16, m, g);
        } else if (two) {
            f32x4_t acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < NSTEP; ++q) {
                const bf16x4_t H0 = build_B2<TIN>(a, L, inb, q, s, ta, xa, g);
                const bf16x4_t H1 = build_B2<TIN>(a, L, inb, q, s, tb, xb, g);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[q], H0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[q], H1, acc1, 0, 0, 0);
            }
            store_tile<COUT, TOUT>(a, acc0, obase, m, g);
            store_tile<COUT, TOUT>(a, acc1, obase + 16, m, g);
        } else {
            f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < NSTEP; ++q)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W[q], build_B2<TIN>(a, L, inb, q, s, ta, xa, g), acc,
                                                                0, 0, 0);
            store_tile<COUT, TOUT>(a, acc, obase, m, g);
        }
    }
}

// --------------------------------------------------------------- weight gradient
// G_s[step][k][n] = sum over the sites of slice s (all right-hand sides) of
// a_p(site)[k] * g(site)[n], a_p the transported layer input, g the gradient of the
// layer output.  Grid (blocks per slice, Ls); block partials [Ls][gridDim.x][NSTEP][256].
template <int CIN, int COUT, typename TIN, typename TG>
__global__ __launch_bounds__(256) void k_ptc_wgrad(const PtcArgs a, const TG* grad, float* part, int nrhs)
{
    constexpr int NSTEP = nstep_of(CIN);
    __shared__ float red[4][NSTEP * 256];
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4, w = threadIdx.x >> 6;
    const int s = blockIdx.y;
    const long long per_s = (long long)nrhs * a.tps;
    f32x4_t acc[NSTEP];
#pragma unroll
    for (int q = 0; q < NSTEP; ++q) acc[q] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    int cur_rhs = -1, row = 0;
    const float2* L = a.links;
    for (long long q = (long long)blockIdx.x * 4 + w; q < per_s; q += (long long)gridDim.x * 4) {
        const int rhs = (int)(q / a.tps);
        const int i0 = (int)(q - (long long)rhs * a.tps) * 16;
        if (rhs != cur_rhs) {
            row = a.in_row ? a.in_row[rhs] : rhs;
            L = a.links + (long long)a.cfg[rhs] * a.link_stride;
            cur_rhs = rhs;
        }
        const long long gbase = ((long long)rhs * a.Ls + s) * a.V4 + i0;
        // B operand: gradient component n = m at sites 4g + j
        bf16x4_t Bv;
        if constexpr (COUT == 8) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                Bv[j] = (short)reinterpret_cast<const unsigned short*>(grad)[(gbase + 4 * g + j) * 16 + m];
        } else {
            float gv4[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float v = 0.f;
                if (m < 4) {
                    const float4 gv = ld_spinor(grad, gbase + 4 * g + j);
                    v = m == 0 ? gv.x : m == 1 ? gv.y : m == 2 ? gv.z : gv.w;
                }
                gv4[j] = v;
            }
            Bv = pack4(gv4[0], gv4[1], gv4[2], gv4[3]);
        }
#pragma unroll
        for (int st = 0; st < NSTEP; ++st) {
            float av[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = i0 + 4 * g + j, t = i / a.Nx, x = i - t * a.Nx;
                float v = 0.f;
                if constexpr (CIN == 8) {
                    const Nb n = neighbour(a, L, st, s, t, x);
                    const unsigned* hp = reinterpret_cast<const unsigned*>(
                        static_cast<const unsigned short*>(a.in) + (((long long)row * a.Ls + n.s) * a.V4 + n.i) * 16) +
                        (m >> 1);
                    const uint2 lk = link_ops(n);
                    v = dot2(*hp, (m & 1) ? lk.y : lk.x);  // rounded to bf16 by pack4
                } else {
                    const int p = 4 * st + (m >> 2), comp = m & 3;
                    if (p < kPts) {
                        const Nb n = neighbour(a, L, p, s, t, x);
                        const float4 h = ld_spinor(static_cast<const TIN*>(a.in) + (long long)row * a.Ls * a.V4 * 4,
                                                   (long long)n.s * a.V4 + n.i);
                        const unsigned hb = comp < 2 ? pk_bf16(h.x, h.y) : pk_bf16(h.z, h.w);
                        const uint2 lk = link_ops(n);
                        v = dot2(hb, (comp & 1) ? lk.y : lk.x);
                    }
                }
                av[j] = v;
            }
            const bf16x4_t A = pack4(av[0], av[1], av[2], av[3]);
            acc[st] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, Bv, acc[st], 0, 0, 0);
        }
    }
#pragma unroll
    for (int st = 0; st < NSTEP; ++st)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[w][st * 256 + (4 * g + j) * 16 + m] = acc[st][j];
    __syncthreads();
    float* out = part + ((long long)s * gridDim.x + blockIdx.x) * NSTEP * 256;
    for (int e = threadIdx.x; e < NSTEP * 256; e += 256) out[e] = red[0][e] + red[1][e] + red[2][e] + red[3][e];
}

// GB[s][e] = sum_b part[s][b][e] (fixed order)
__global__ void k_wgrad_reduce(const float* part, int nb, int per, int Ls, double* GB)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)Ls * per) return;
    const int s = (int)(idx / per), e = (int)(idx - (long long)s * per);
    double v = 0.0;
    for (int b = 0; b < nb; ++b) v += part[((long long)s * nb + b) * per + e];
    GB[idx] = v;
}

// Complex weight gradient from GB and (optionally) one Adam step on the master weights.
__global__ void k_fold_adam(const double* GB, int cin, int cout, int Ls, float2* w, float2* m1, float2* m2,
                            float* grad_out, int adam, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int n = Ls * kPts * cout * cin;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = idx % cin, o = (idx / cin) % cout, p = (idx / (cin * cout)) % kPts, s = idx / (cin * cout * kPts);
    const int NSTEP = nstep_of(cin);
    const int step = cin == 8 ? p : p / 4;
    const int k0 = cin == 8 ? 2 * c : 4 * (p % 4) + 2 * c;
    const double* G = GB + ((long long)s * NSTEP + step) * 256;
    const int n0 = 2 * o;
    const float gr = (float)(G[k0 * 16 + n0] + G[(k0 + 1) * 16 + n0 + 1]);
    const float gi = (float)(G[k0 * 16 + n0 + 1] - G[(k0 + 1) * 16 + n0]);
    if (grad_out) {
        grad_out[2 * idx] = gr;
        grad_out[2 * idx + 1] = gi;
    }
    if (!adam) return;
    float2 a = m1[idx], v = m2[idx], x = w[idx];
    a.x = b1 * a.x + (1.f - b1) * gr;
    a.y = b1 * a.y + (1.f - b1) * gi;
    v.x = b2 * v.x + (1.f - b2) * gr * gr;
    v.y = b2 * v.y + (1.f - b2) * gi * gi;
    x.x -= lr * (a.x * c1) / (sqrtf(v.x * c2) + eps);
    x.y -= lr * (a.y * c1) / (sqrtf(v.y * c2) + eps);
    m1[idx] = a;
    m2[idx] = v;
    w[idx] = x;
}

// BF16 B fragments of layer (cin_l -> cout_l): forward, or the adjoint kernel's
// (cout_l -> cin_l) with W'_p(s) = W_{-p}(s + ds_p)^H.
// adj: 0 forward, 1 adjoint, 2 forward C_in = 2 fragments in the point order of the
// gauge-fixed fused kernel (MFMA step 0: self, +t, -t, +s; step 1: -s, +x, -x, none).
__global__ void k_pack(const float2* w, int cin_l, int cout_l, int Ls, int mode, unsigned short* out)
{
    const bool adj = mode == 1;
    const int kin = adj ? cout_l : cin_l, kout = adj ? cin_l : cout_l;
    const int NSTEP = nstep_of(kin);
    const long long total = (long long)Ls * NSTEP * 256;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int jj = (int)(idx & 3), lane = (int)((idx >> 2) & 63);
    const int step = (int)((idx >> 8) % NSTEP), s = (int)(idx / (256LL * NSTEP));
    const int g = lane >> 4, n = lane & 15, k = 4 * g + jj;
    int p, c, kr;
    if (kin == 8) {
        p = step;
        c = k >> 1;
        kr = k & 1;
    } else {
        constexpr int gf[2][4] = {{0, 1, 2, 5}, {6, 3, 4, kPts}};
        p = mode == 2 ? gf[step][g] : 4 * step + g;
        c = jj >> 1;
        kr = jj & 1;
    }
    const int o = n >> 1, nr = n & 1;
    float val = 0.f;
    if (p < kPts && o < kout) {
        float wr, wi;
        if (!adj) {
            const float2 W = w[(((long long)s * kPts + p) * cout_l + o) * cin_l + c];
            wr = W.x;
            wi = W.y;
        } else {
            const int neg[kPts] = {0, 2, 1, 4, 3, 6, 5};
            const int ds = p == 5 ? 1 : p == 6 ? -1 : 0;
            const int sn = (s + ds + Ls) % Ls;
            const float2 W = w[(((long long)sn * kPts + neg[p]) * cout_l + c) * cin_l + o];
            wr = W.x;
            wi = -W.y;
        }
        val = kr == 0 ? (nr == 0 ? wr : wi) : (nr == 0 ? -wi : wr);
    }
    out[idx] = f2bf(val);
}

// ----------------------------------------------------------------- reductions
__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sums of a and b (256 threads) into part[0..1] (thread 0 writes).
__device__ __forceinline__ void block_part2(double a, double b, double* part)
{
    __shared__ double sh[8];
    a = wsum(a);
    b = wsum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[2 * w] = a;
        sh[2 * w + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[0] = sh[0] + sh[2] + sh[4] + sh[6];
        part[1] = sh[1] + sh[3] + sh[5] + sh[7];
    }
}

// Sum of element `which` of n partial pairs (64-thread block, every thread gets it).
__device__ __forceinline__ double fold64(const double* p, int n, int which)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) v += p[2 * i + which];
    return wsum(v);
}

// ------------------------------------------------------- fused network forward
// z = M r in one launch: the three layers of one output tile of TY x TX sites of one
// right-hand side, walking the fifth dimension.  A block stages, for the slices it
// needs, the tile's input with a 3-site halo (bf16), the first hidden layer with a
// 2-site halo and the second with a 1-site halo in LDS rings of 4 slices, so HBM sees
// the input once (plus the halo and the wrapped slices, mostly from L2) and the output
// once instead of every layer's activations.  The +-s hops are periodic, so the walk
// covers the extended range: input slices -3 .. Ls+2, layer 1 at -2 .. Ls+1, layer 2
// at -1 .. Ls, output at 0 .. Ls-1 (extended slices are recomputed from the same data
// and weights as their periodic images, so the result equals the layer-by-layer one).
//
// Step j (one barrier each): every wave loads its layers' weight fragments for the
// step, then input slice j-2 into registers (written to LDS at the end of step j+1, so
// its latency spans two steps); layer 1 runs at slice j-5, layer 2 at j-7, the output
// at j-9.  Waves have fixed roles and fixed 16-site tiles (no loop over tiles, so the
// compiler waits for exactly the weights before the first MFMA and never for the
// input in flight): waves 0-3 run layer 2 (3 tiles each), waves 4-7 the output layer
// (2 tiles each), and every wave 2 tiles of layer 1.  The transport operands of a wave's tiles
// and, for layers 2 and 3, the fifth-dimension window (the s-1 and s values at the
// tile's sites) stay in registers across steps.  The arithmetic (bf16 operands, dot2
// transport, MFMA order over stencil points) is that of k_ptc_fwd, so both give
// bit-identical results.  Optionally the flexible-CG dot products (z, r) and (z, q) per
// block, in a pass over the block's output after the walk.
constexpr int kFuTy = 8, kFuTx = 16;  // largest output tile
constexpr int kFuR0 = (kFuTy + 6) * (kFuTx + 6), kFuR1 = (kFuTy + 4) * (kFuTx + 4), kFuR2 = (kFuTy + 2) * (kFuTx + 2);
constexpr int kFuThreads = 512, kFuWaves = kFuThreads / 64;
// tiles per wave: waves 0-3 run 3 layer-2 tiles, waves 4-7 2 output tiles, and every wave
// 2 layer-1 tiles (wave w takes layer-1 tiles 2w, 2w + 1)
constexpr int kFuL2Tiles = 3, kFuL3Tiles = 2, kFuL1Tiles = 2;
static_assert(4 * kFuL2Tiles * 16 >= kFuR2 && 4 * kFuL3Tiles * 16 >= kFuTy * kFuTx &&
                  kFuWaves * kFuL1Tiles * 16 >= kFuR1,
              "the waves' tiles must cover every region of the largest output tile");

// Element of channel group g (channels 2g, 2g+1) of a hidden-layer site in an LDS ring
// slot.  A half-wave of an MFMA operand read covers 16 sites x groups {0, 1} (or {2, 3});
// with 32-byte sites, sites m and m + 8 would share banks, so the group index is XOR-ed
// with 2 * bit 3 of the site: 16 consecutive sites then cover the 64 banks once.
__device__ __forceinline__ int hel(int site, int g) { return site * 4 + (g ^ ((site >> 2) & 2)); }

struct NetFusedArgs {
    int Nt, Nx, Ls, V4;
    int TY, TX;            // output tile, dividing Nt and Nx
    int tiles_x, tiles;    // tiles along x, tiles per lattice
    int nsplit;            // parts of the fifth dimension per tile (one block each), dividing Ls
    const void* in;        // [nrhs][Ls][V4] spinors (TIN)
    void* out;             // [nrhs][Ls][V4] spinors (TOUT)
    const float2* links;   // [cfg][2][V4] (cos, sin)
    long long link_stride;
    const int* cfg;        // right-hand side -> configuration
    const unsigned short* wf[3];
    const CgState* st;     // nullable: skip right-hand sides with flags[0] & (F_DONE | F_FROZEN)
    const float* q;        // DOTS: q = A p of the flexible CG
    double* part;          // DOTS: [nrhs][tiles][2] partials of (z, r), (z, q)
    // DOTS, optional (null hpart): the CG iteration's check before the network (in place of
    // k_fcg_check): (p, A p) and |r|^2 from the stencils' partials hpart, rpart ([nrhs][hnb][2]).
    const double* hpart;
    const double* rpart;
    int hnb;
    CgState* cst;          // the same states as st, updated by the block of tile 0
    FcgScal* fs;
};

// s_waitcnt immediate (gfx9 encoding): vmcnt(0), expcnt and lgkmcnt unconstrained
constexpr int kWaitVmcnt0 = 0x0F70;

// i = k / w for 0 <= k < 2^16 / w (w <= 22 here): one multiply and a shift
__device__ __forceinline__ int div_small(int k, unsigned inv) { return (int)(((unsigned)k * inv) >> 16); }

// Transporter (cos, sin) of hop p (1 +t, 2 -t, 3 +x, 4 -x; anything else: none, 1) at
// lattice site (t, x): the Dirac operator's transporters with the antiperiodic time
// boundary, as neighbour() forms them.
__device__ __forceinline__ float2 hop_u(const float2* L, int Nt, int Nx, int V4, int p, int t, int x)
{
    float c = 1.f, sn = 0.f;
    if (p == 1) {
        const float2 u = L[t * Nx + x];
        const float sg = t == Nt - 1 ? -1.f : 1.f;
        c = sg * u.x;
        sn = sg * u.y;
    } else if (p == 2) {
        const int tm = t == 0 ? Nt - 1 : t - 1;
        const float2 u = L[tm * Nx + x];
        const float sg = t == 0 ? -1.f : 1.f;
        c = sg * u.x;
        sn = sg * -u.y;
    } else if (p == 3) {
        const float2 u = L[V4 + t * Nx + x];
        c = u.x;
        sn = u.y;
    } else if (p == 4) {
        const float2 u = L[V4 + t * Nx + (x == 0 ? Nx - 1 : x - 1)];
        c = u.x;
        sn = -u.y;
    }
    return make_float2(c, sn);
}

// bf16 dot-product operands of the same transporter
__device__ __forceinline__ uint2 hop_link(const float2* L, int Nt, int Nx, int V4, int p, int t, int x)
{
    const float2 u = hop_u(L, Nt, Nx, V4, p, t, x);
    return link_ops(Nb{0, 0, u.x, u.y});
}

template <int V> using IC = std::integral_constant<int, V>;

template <typename TIN, typename TOUT, bool DOTS, bool TG>
__global__ __launch_bounds__(kFuThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_net_fused(const NetFusedArgs a)
{
    __shared__ uint2 ring0[4][kFuR0];     // input, bf16 (comp 0, comp 1)
    __shared__ uint2 ring1[4][kFuR1][4];  // layer 1 output, 8 channels bf16
    __shared__ uint2 ring2[4][kFuR2][4];  // layer 2 output
    __shared__ uint2 wrec[2][2 * kFuThreads]; // weight fragments of a step: layer 1 (2), 2 (7), 3 (7) x 64 lanes
    __shared__ double red[2 * kFuWaves];

    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int nsp = a.nsplit, rt = lb / nsp, part = lb - rt * nsp;
    const int rhs = rt / a.tiles, tile = rt - rhs * a.tiles;
    if (a.st && (a.st[rhs].flags[0] & (F_DONE | F_FROZEN))) return;  // block-uniform
    if constexpr (DOTS) {
        if (a.hpart && tile == 0 && part == 0) {
            // the iteration's check (as k_fcg_check), by the block of tile 0 alone; the
            // other blocks compute z regardless (a right-hand side that converges or breaks
            // down here wastes one network application; the kernels after this one skip it)
            if (threadIdx.x < 64) {
                const double pq = fold64(a.hpart + (long long)rhs * a.hnb * 2, a.hnb, 0);
                const double rr = fold64(a.rpart + (long long)rhs * a.hnb * 2, a.hnb, 0);
                if (threadIdx.x == 0) {
                    CgState& S = a.cst[rhs];
                    if (!(pq > 0.0)) {
                        S.status = 4;
                        S.flags[0] |= F_DONE;
                    } else {
                        a.fs[rhs].pq = pq;
                        a.fs[rhs].rz_prev = a.fs[rhs].rz;  // consumed by this iteration's alpha
                        S.rr[0] = rr;
                        S.iters += 1;
                        S.relres = std::sqrt(rr / S.bnorm2);
                        if (S.relres <= S.tol || rr <= S.delta * S.delta * S.rr_ref) S.flags[0] |= F_FROZEN;
                    }
                }
            }
        }
    }
    const int TY = a.TY, TX = a.TX;
    const int W0 = TX + 6, W1 = TX + 4, W2 = TX + 2;
    const int n0 = (TY + 6) * W0, n1 = (TY + 4) * W1, n2 = (TY + 2) * W2, n3 = TY * TX;
    const unsigned inv0 = (65535u + W0) / W0, inv1 = (65535u + W1) / W1, inv2 = (65535u + W2) / W2,
                   inv3 = (65535u + TX) / TX;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const int t0 = ty * TY, x0 = tx * TX;
    const int Nt = a.Nt, Nx = a.Nx, Ls = a.Ls, V4 = a.V4;
    auto wrap_t = [Nt](int t) { return t < 0 ? t + Nt : (t >= Nt ? t - Nt : t); };
    auto wrap_x = [Nx](int x) { return x < 0 ? x + Nx : (x >= Nx ? x - Nx : x); };
    const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar control
    const float2* L = a.links + (long long)a.cfg[rhs] * a.link_stride;

    // this thread's input site of R0 (origin t0-3, x0-3)
    int gsite = -1;
    if (tid < n0) {
        const int i = div_small(tid, inv0), jx = tid - i * W0;
        gsite = wrap_t(t0 - 3 + i) * Nx + wrap_x(x0 - 3 + jx);
    }
    // TG: the block computes in a gauge of its own (temporal gauge on its input region
    // R0, rows t0-3 .. t0+TY+2, columns x0-3 .. x0+TX+2, in patch coordinates, so a region
    // that wraps around a small lattice is a patch of the covering lattice): Om = 1 on the
    // region's first row and Om(i + 1, jx) = Om(i, jx) U_+t(i, jx), so every +-t transporter
    // inside the region is 1 and no layer transports along t; the +-x transporters become
    // Om(a) U Om(b)^*, the input is multiplied by Om and the output by Om^*.  The network
    // is gauge covariant, so the output is the same up to the bf16 rounding of the
    // transformed values.  Om is kept in the first bytes of ring2 until the walk (ring2 is
    // first written in step 6, after the set-up's reads).
    float2 om_in = make_float2(1.f, 0.f);
    const float2* om = reinterpret_cast<const float2*>(&ring2[0][0][0]);
    if constexpr (TG) {
        float2* omw = reinterpret_cast<float2*>(&ring2[0][0][0]);
        if (tid < W0) {
            float2 o = make_float2(1.f, 0.f);
            const int x = wrap_x(x0 - 3 + tid);
            for (int i = 0; i < TY + 6; ++i) {
                omw[i * W0 + tid] = o;
                o = cmulf(o, hop_u(L, Nt, Nx, V4, 1, wrap_t(t0 - 3 + i), x));
            }
        }
        __syncthreads();
        if (tid < n0) om_in = om[tid];
    }
    // +x (p = 3) or -x (p = 4) transporter at the site (t, x) with R0 index b, in the block's gauge
    auto tg_u = [&](int p, int b, int t, int x) {
        return cmulf(cmulf(om[b], hop_u(L, Nt, Nx, V4, p, t, x)), conjf2(om[p == 3 ? b + 1 : b - 1]));
    };
    const TIN* inr = static_cast<const TIN*>(a.in) + (long long)rhs * Ls * V4 * 4;
    TOUT* outr = static_cast<TOUT*>(a.out) + (long long)rhs * Ls * V4 * 4;
    auto smod = [Ls](int s) { return ((s % Ls) + Ls) % Ls; };
    // The block's output slices s_lo .. s_lo + Lw - 1: the walk below runs in slices
    // relative to s_lo (its ranges end at Lw), global loads and stores at s_lo + s (mod
    // Ls).  Parts of one tile recompute their neighbours' edge slices (the walk's extended
    // range), which buys blocks when a launch has too few to fill the GPU.
    const int Lw = Ls / nsp, s_lo = part * Lw;
    // this thread's 16 bytes of a step's weight record: layer 1 (threads 0..63), layer 2
    // (64..287), layer 3 (288..511); the layers run 5, 7 and 9 steps behind the step index
    const int wl = tid < 64 ? 0 : (tid < 288 ? 1 : 2);
    const int wbase = (tid < 64 ? tid : (tid < 288 ? tid - 64 : tid - 288)) * 8;  // shorts

    // The walk runs steps j = -3 .. Lw + 8 four at a time; the position K = (j + 3) mod 4
    // is a compile-time constant in each copy of the step, and so is every ring slot
    // (slot(s) = s mod 4): slot(j - lag + d) = (K + 13 - lag + d) mod 4.  Nothing runs
    // before step 3, whose layer 1 needs input slices -3 .. -1 (loaded in steps -1 .. 1).
    //
    // Global loads are issued at the end of a step (after its output stores) and stored
    // to LDS at the end of the next one, so their latency spans a step; the two register
    // sets alternate between steps, so no copy waits for a load in flight.  Every thread
    // issues both loads on every step (clamped), so the count in flight never depends on
    // the path.  Step j loads input slice j-2 and the weights of step j+2, and stores
    // input slice j-3 (its ring slot held slice j-7, last read by layer 1 in step j-1) and
    // the weights of step j+1 (their buffer was last read in step j-1).
    const int jend = -3 + 4 * ((Lw + 15) / 4);
    // The slices a step loads, smod(s_lo + j - 2) (input; past relative slice Lw + 2 the
    // loaded value is not used) and smod(s_lo + j - 3), smod(s_lo + j - 5), smod(s_lo + j - 7)
    // (weights of layers 1-3), are
    // carried from step to step as scalar counters: a modulo per step costs ~25 scalar
    // instructions.
    struct Slices {
        int c2, c3, c5, c7;
    };
    auto slices_at = [&](int j) {
        return Slices{smod(s_lo + j - 2), smod(s_lo + j - 3), smod(s_lo + j - 5), smod(s_lo + j - 7)};
    };
    auto loads = [&](Slices& c, float4& in_new, uint4& w_new) {
        in_new = ld_spinor(inr, (long long)c.c2 * V4 + max(gsite, 0));
        // the three layers' slices are wave-uniform; the thread picks its layer's record
        const unsigned short* w0 = a.wf[0] + c.c3 * (2 * 64 * 4);
        const unsigned short* w1 = a.wf[1] + c.c5 * (7 * 64 * 4);
        const unsigned short* w2 = a.wf[2] + c.c7 * (7 * 64 * 4);
        w_new = *reinterpret_cast<const uint4*>((wl == 0 ? w0 : (wl == 1 ? w1 : w2)) + wbase);
        auto inc = [Ls](int& v) { v = v + 1 == Ls ? 0 : v + 1; };
        inc(c.c2);
        inc(c.c3);
        inc(c.c5);
        inc(c.c7);
    };
    auto stores = [&](auto Kc, int j, const float4& in_old, const uint4& w_old) {
        constexpr int K = decltype(Kc)::value;
        if (j - 3 >= -3 && j - 3 <= Lw + 2 && gsite >= 0) {
            float4 v = in_old;
            if constexpr (TG) {
                const float2 c0 = cmulf(om_in, make_float2(v.x, v.y)), c1 = cmulf(om_in, make_float2(v.z, v.w));
                v = make_float4(c0.x, c0.y, c1.x, c1.y);
            }
            ring0[(K + 2) & 3][tid] = make_uint2(pk_bf16(v.x, v.y), pk_bf16(v.z, v.w));
        }
        reinterpret_cast<uint4*>(wrec[K & 1])[tid] = w_old;
        __syncthreads();
    };
    // fragment q of layer l (0, 1, 2) in the weight record of a step at position K
    auto frag = [&](auto Kc, int l, int q) {
        constexpr int K = decltype(Kc)::value;
        const uint2 v = wrec[(K + 1) & 1][(l == 0 ? 0 : (l == 1 ? 128 : 576)) + q * 64 + lane];
        return pack_pairs(v.x, v.y);
    };

    // One 16-site tile of a C_in = 8 layer: rin the input ring, S0 the slot of slice s,
    // etp / etm / exp / exm the lane's +t / -t / +x / -x neighbours (hel) in the input
    // region, u the transporters at the output site; win[0], win[1] the lane's own values
    // at s-1 and s and up its value at s+1.
    //
    // The window is advanced in every step of the walk, also where the layer is outside
    // its range (advance8): at step j the slot of slice s+1 holds the value the layer
    // below wrote in step j-1, so the two steps before a layer's first slice fill its
    // window with the same values a load at that slice would read, and no step branches
    // on a first slice (a branch there left the window in different registers on its two
    // sides, paid for with register moves in every step).  Steps before the layer below
    // has written read slots not yet written; those values leave the window before use.
    auto hop8 = [&](auto Kc, auto Lc, auto S0c, const auto& rin, int etp, int etm, int exp, int exm,
                    const float2 (&u)[4], const uint2 (&win)[2], uint2 up) {
        constexpr int S0 = decltype(S0c)::value;
        const uint2* rs = &rin[S0][0][0];
        bf16x4_t h[7], W[7];
        (void)u;
        const uint2 vtp = rs[etp], vtm = rs[etm], vxp = rs[exp], vxm = rs[exm];
        h[0] = pack_pairs(win[1].x, win[1].y);
        h[1] = pack_pairs(vtp.x, vtp.y);
        h[2] = pack_pairs(vtm.x, vtm.y);
        h[3] = pack_pairs(vxp.x, vxp.y);
        h[4] = pack_pairs(vxm.x, vxm.y);
        h[5] = pack_pairs(up.x, up.y);
        h[6] = pack_pairs(win[0].x, win[0].y);
#pragma unroll
        for (int q = 0; q < 7; ++q) W[q] = frag(Kc, decltype(Lc)::value, q);
        if constexpr (TG) {
            const float2 ux[2] = {u[2], u[3]};
            return tile8_tg(W, h, ux);
        } else {
            return tile8(W, h, u);
        }
    };

    // the values at slice s+1 (slot SP of the ring) of every tile, read in every step
    auto ahead8 = [&](auto SPc, const auto& rin, int e0) {
        return (&rin[decltype(SPc)::value][0][0])[e0];
    };
    auto advance8 = [&](uint2 (&win)[2], uint2 up) {
        win[0] = win[1];
        win[1] = up;
    };

    // ---- layer 1 (every wave): R1 tiles wv*2 + e (origin t0-2, x0-2), input ring0 (width W0).
    // Lane group g supplies stencil point g (MFMA step 0: self, +t, -t, +x, all in slice
    // s) and point 4 + g (step 1: -x in slice s, +s, -s, and none with zero transporter).
    // TG: step 0 self, +t, -t, +s (no transport at all), step 1 -s, +x, -x, none.
    int e1a[kFuL1Tiles], e1b[kFuL1Tiles], o1[kFuL1Tiles];
    uint2 lk1[kFuL1Tiles][2];
    // slice offsets of the lane's points in steps 0 and 1
    const int dsa = TG && g == 3 ? 1 : 0;
    const int dsb = TG ? (g == 0 ? -1 : 0) : (g == 1 ? 1 : (g == 2 ? -1 : 0));
#pragma unroll
    for (int e = 0; e < kFuL1Tiles; ++e) {
        const int k = wv * kFuL1Tiles + e;
        const int sg = min(16 * k + m, n1 - 1);
        const int i = div_small(sg, inv1), jx = sg - i * W1;
        const int b0 = (i + 1) * W0 + jx + 1;
        o1[e] = hel(sg, g);
        const int t = wrap_t(t0 - 2 + i), x = wrap_x(x0 - 2 + jx);
        if constexpr (TG) {
            e1a[e] = g == 1 ? b0 + W0 : (g == 2 ? b0 - W0 : b0);
            e1b[e] = g == 1 ? b0 + 1 : (g == 2 ? b0 - 1 : b0);
            lk1[e][0] = make_uint2(0u, 0u);
            lk1[e][1] = g == 0 ? link_ops(Nb{0, 0, 1.f, 0.f})
                                : (g == 3 ? make_uint2(0u, 0u) : link_ops_f2(tg_u(g == 1 ? 3 : 4, b0, t, x)));
        } else {
            e1a[e] = g == 1 ? b0 + W0 : (g == 2 ? b0 - W0 : (g == 3 ? b0 + 1 : b0));
            e1b[e] = g == 0 ? b0 - 1 : b0;
            lk1[e][0] = hop_link(L, Nt, Nx, V4, g, t, x);  // point g (0: none)
            lk1[e][1] = g == 3 ? make_uint2(0u, 0u) : hop_link(L, Nt, Nx, V4, 4 + g, t, x);  // -x, none, none, zero
        }
    }
    // Every wave runs all its tile slots: a slot past the region's end computes the clamped
    // last site again and stores the same bits there (no branch between the tiles, so the
    // compiler interleaves their MFMA chains).
    auto layer1 = [&](auto Kc, int j) {
        constexpr int K = decltype(Kc)::value;
        const int s1 = j - 5;
        if (s1 < -2 || s1 > Lw + 1) return;
        constexpr int S0 = K & 3;
        const int sa = (K + dsa) & 3, sb = (K + dsb) & 3;  // slots of the lane's step-0 / step-1 points
        const bf16x4_t Wa0 = frag(Kc, 0, 0), Wa1 = frag(Kc, 0, 1);
#pragma unroll
        for (int e = 0; e < kFuL1Tiles; ++e)
            {
                const uint2 va = ring0[sa][e1a[e]];
                const uint2 vb = ring0[sb][e1b[e]];
                bf16x4_t B0;
                if constexpr (TG)
                    B0 = pack_pairs(va.x, va.y);
                else
                    B0 = pack_pairs(transport(va.x, lk1[e][0]), transport(va.y, lk1[e][0]));
                const bf16x4_t B1 = pack_pairs(transport(vb.x, lk1[e][1]), transport(vb.y, lk1[e][1]));
                f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Wa0, B0, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Wa1, B1, acc, 0, 0, 0);
                // a lane past the region's end computed its clamped last site: it stores the
                // same bits there as that site's own lane (no branch on the lane)
                (&ring1[S0][0][0])[o1[e]] = make_uint2(pk_bf16(acc[0], acc[1]), pk_bf16(acc[2], acc[3]));
            }
    };

    if (wv < 4) {
        // ---- layer 2: R2 tiles wv*3 + e (origin t0-1, x0-1), input ring1 (width W1)
        int e0[kFuL2Tiles], etp[kFuL2Tiles], etm[kFuL2Tiles], exp[kFuL2Tiles], exm[kFuL2Tiles], eo[kFuL2Tiles];
        float2 lk[kFuL2Tiles][4];
        uint2 win[kFuL2Tiles][2];
#pragma unroll
        for (int e = 0; e < kFuL2Tiles; ++e) {
            const int k = wv * kFuL2Tiles + e;
            const int sg = min(16 * k + m, n2 - 1);
            const int i = div_small(sg, inv2), jx = sg - i * W2;
            const int b = (i + 1) * W1 + jx + 1;
            e0[e] = hel(b, g);
            etp[e] = hel(b + W1, g);
            etm[e] = hel(b - W1, g);
            exp[e] = hel(b + 1, g);
            exm[e] = hel(b - 1, g);
            eo[e] = hel(sg, g);
            const int t = wrap_t(t0 - 1 + i), x = wrap_x(x0 - 1 + jx);
            if constexpr (TG) {
                const int pb = (i + 2) * W0 + jx + 2;
                lk[e][0] = lk[e][1] = make_float2(1.f, 0.f);
                lk[e][2] = tg_u(3, pb, t, x);
                lk[e][3] = tg_u(4, pb, t, x);
            } else {
#pragma unroll
                for (int p = 1; p <= 4; ++p) lk[e][p - 1] = hop_u(L, Nt, Nx, V4, p, t, x);
            }
            win[e][0] = win[e][1] = make_uint2(0u, 0u);
        }
        auto step = [&](auto Kc, int jv, Slices& sl, float4& in_new, uint4& w_new, const float4& in_old,
                        const uint4& w_old) {
            constexpr int K = decltype(Kc)::value;
            const int j = __builtin_amdgcn_readfirstlane(jv);  // keep the walk's arithmetic scalar
            const int s2 = j - 7;
            layer1(Kc, j);
            uint2 up[kFuL2Tiles];
#pragma unroll
            for (int e = 0; e < kFuL2Tiles; ++e) up[e] = ahead8(IC<(K + 3) & 3>{}, ring1, e0[e]);
            if (s2 >= -1 && s2 <= Lw) {
#pragma unroll
                for (int e = 0; e < kFuL2Tiles; ++e)
                    {
                        const f32x4_t acc = hop8(Kc, IC<1>{}, IC<(K + 2) & 3>{}, ring1, etp[e], etm[e], exp[e], exm[e],
                                                 lk[e], win[e], up[e]);
                        // past the region's end: the clamped last site's value, stored there again
                        (&ring2[(K + 2) & 3][0][0])[eo[e]] =
                            make_uint2(pk_bf16(acc[0], acc[1]), pk_bf16(acc[2], acc[3]));
                    }
            }
#pragma unroll
            for (int e = 0; e < kFuL2Tiles; ++e) advance8(win[e], up[e]);
            loads(sl, in_new, w_new);
            stores(Kc, j, in_old, w_old);
        };
        float4 inX = make_float4(0.f, 0.f, 0.f, 0.f), inY = inX;
        uint4 wX = make_uint4(0u, 0u, 0u, 0u), wY = wX;
        // the set-up's global loads (transporters) are complete before the walk, so no
        // wait inside it is for them (such a wait would also wait for the input in flight)
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        Slices sl = slices_at(-3);
        for (int j = -3; j < jend; j += 4) {
            step(IC<0>{}, j, sl, inX, wX, inY, wY);
            step(IC<1>{}, j + 1, sl, inY, wY, inX, wX);
            step(IC<2>{}, j + 2, sl, inX, wX, inY, wY);
            step(IC<3>{}, j + 3, sl, inY, wY, inX, wX);
        }
    } else {
        // ---- output layer: tiles (wv-4)*2 + e of the TY x TX tile, input ring2 (width W2)
        const int wr = wv - 4;
        int e3[kFuL3Tiles], e3p[kFuL3Tiles], e3m[kFuL3Tiles], e3xp[kFuL3Tiles], e3xm[kFuL3Tiles], o3[kFuL3Tiles];
        bool v3[kFuL3Tiles];
        float2 lk3[kFuL3Tiles][4];
        float2 om3[kFuL3Tiles];  // TG: Om^* at the output site
        uint2 win3[kFuL3Tiles][2];
#pragma unroll
        for (int e = 0; e < kFuL3Tiles; ++e) {
            const int k = wr * kFuL3Tiles + e;
            const int sg = min(16 * k + m, n3 - 1);
            v3[e] = 16 * k + m < n3 && g == 0;
            const int i = div_small(sg, inv3), jx = sg - i * TX;
            const int b = (i + 1) * W2 + jx + 1;
            e3[e] = hel(b, g);
            e3p[e] = hel(b + W2, g);
            e3m[e] = hel(b - W2, g);
            e3xp[e] = hel(b + 1, g);
            e3xm[e] = hel(b - 1, g);
            o3[e] = (t0 + i) * Nx + x0 + jx;
            if constexpr (TG) {
                const int pb = (i + 3) * W0 + jx + 3;
                lk3[e][0] = lk3[e][1] = make_float2(1.f, 0.f);
                lk3[e][2] = tg_u(3, pb, t0 + i, x0 + jx);
                lk3[e][3] = tg_u(4, pb, t0 + i, x0 + jx);
                om3[e] = conjf2(om[pb]);
            } else {
#pragma unroll
                for (int p = 1; p <= 4; ++p) lk3[e][p - 1] = hop_u(L, Nt, Nx, V4, p, t0 + i, x0 + jx);
                om3[e] = make_float2(1.f, 0.f);
            }
            win3[e][0] = win3[e][1] = make_uint2(0u, 0u);
        }
        auto step = [&](auto Kc, int jv, Slices& sl, float4& in_new, uint4& w_new, const float4& in_old,
                        const uint4& w_old) {
            constexpr int K = decltype(Kc)::value;
            const int j = __builtin_amdgcn_readfirstlane(jv);  // keep the walk's arithmetic scalar
            const int s3 = j - 9;
            layer1(Kc, j);
            uint2 up3[kFuL3Tiles];
#pragma unroll
            for (int e = 0; e < kFuL3Tiles; ++e) up3[e] = ahead8(IC<(K + 1) & 3>{}, ring2, e3[e]);
            if (s3 >= 0 && s3 <= Lw - 1) {
#pragma unroll
                for (int e = 0; e < kFuL3Tiles; ++e)
                    {
                        const f32x4_t acc = hop8(Kc, IC<2>{}, IC<K & 3>{}, ring2, e3p[e], e3m[e], e3xp[e], e3xm[e],
                                                 lk3[e], win3[e], up3[e]);
                        if (v3[e]) {
                            if constexpr (TG) {
                                const float2 c0 = cmulf(om3[e], make_float2(acc[0], acc[1]));
                                const float2 c1 = cmulf(om3[e], make_float2(acc[2], acc[3]));
                                st_spinor(outr, (long long)(s_lo + s3) * V4 + o3[e], c0.x, c0.y, c1.x, c1.y);
                            } else {
                                st_spinor(outr, (long long)(s_lo + s3) * V4 + o3[e], acc[0], acc[1], acc[2], acc[3]);
                            }
                        }
                    }
            }
#pragma unroll
            for (int e = 0; e < kFuL3Tiles; ++e) advance8(win3[e], up3[e]);
            loads(sl, in_new, w_new);
            stores(Kc, j, in_old, w_old);
        };
        float4 inX = make_float4(0.f, 0.f, 0.f, 0.f), inY = inX;
        uint4 wX = make_uint4(0u, 0u, 0u, 0u), wY = wX;
        // the set-up's global loads (transporters) are complete before the walk, so no
        // wait inside it is for them (such a wait would also wait for the input in flight)
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        Slices sl = slices_at(-3);
        for (int j = -3; j < jend; j += 4) {
            step(IC<0>{}, j, sl, inX, wX, inY, wY);
            step(IC<1>{}, j + 1, sl, inY, wY, inX, wX);
            step(IC<2>{}, j + 2, sl, inX, wX, inY, wY);
            step(IC<3>{}, j + 3, sl, inY, wY, inX, wX);
        }
    }
    if constexpr (DOTS) {
        // (z, r) and, with q, (z, q) over the block's output (z written above by this block)
        const float4* Z = reinterpret_cast<const float4*>(outr);
        const float4* Rv = reinterpret_cast<const float4*>(inr);
        const float4* Q = a.q ? reinterpret_cast<const float4*>(a.q) + (long long)rhs * Ls * V4 : nullptr;
        double dzr = 0.0, dzq = 0.0;
        for (int e = tid; e < n3 * Lw; e += kFuThreads) {
            const int sr = e / n3, k = e - sr * n3, s = s_lo + sr;
            const int i = div_small(k, inv3), jx = k - i * TX;
            const long long site = (long long)s * V4 + (t0 + i) * Nx + x0 + jx;
            const float4 z = Z[site], r = Rv[site];
            dzr += (double)z.x * r.x + (double)z.y * r.y + (double)z.z * r.z + (double)z.w * r.w;
            if (Q) {
                const float4 qv = Q[site];
                dzq += (double)z.x * qv.x + (double)z.y * qv.y + (double)z.z * qv.z + (double)z.w * qv.w;
            }
        }
        dzr = wsum(dzr);
        dzq = wsum(dzq);
        if (lane == 0) {
            red[2 * wv] = dzr;
            red[2 * wv + 1] = dzq;
        }
        __syncthreads();
        if (tid == 0) {
            double zr = 0.0, zq = 0.0;
            for (int w = 0; w < kFuWaves; ++w) {
                zr += red[2 * w];
                zq += red[2 * w + 1];
            }
            double* pp = a.part + (((long long)rhs * a.tiles + tile) * nsp + part) * 2;
            pp[0] = zr;
            pp[1] = zq;
        }
    }
}

// |x_r|^2 of float spinor rows (one block per row)
__global__ void k_row_norms(const float* x, long long sites, double* out)
{
    double a = 0.0;
    const float4* v = reinterpret_cast<const float4*>(x) + (long long)blockIdx.x * sites;
    for (long long i = threadIdx.x; i < sites; i += 256) {
        const float4 q = v[i];
        a += (double)q.x * q.x + (double)q.y * q.y + (double)q.z * q.z + (double)q.w * q.w;
    }
    __shared__ double tmp[2];
    block_part2(a, 0.0, tmp);
    if (threadIdx.x == 0) out[blockIdx.x] = tmp[0];
}

// Solution loss: d = y - x_row, gy = d * 2 / (npairs |x_row|^2), partial |d|^2.
__global__ __launch_bounds__(256) void k_loss_sol(const float* y, const float* x, const int* row,
                                                  const double* xn2, long long sites, int npairs, float* gy,
                                                  double* part)
{
    const int rhs = blockIdx.y, r = row[rhs];
    const float sc = (float)(2.0 / ((double)npairs * xn2[r]));
    const float4* Y = reinterpret_cast<const float4*>(y) + (long long)rhs * sites;
    const float4* X = reinterpret_cast<const float4*>(x) + (long long)r * sites;
    float4* G = reinterpret_cast<float4*>(gy) + (long long)rhs * sites;
    double a = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < sites; i += (long long)gridDim.x * 256) {
        const float4 u = Y[i], v = X[i];
        const float4 d = make_float4(u.x - v.x, u.y - v.y, u.z - v.z, u.w - v.w);
        a += (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z + (double)d.w * d.w;
        G[i] = make_float4(d.x * sc, d.y * sc, d.z * sc, d.w * sc);
    }
    block_part2(a, 0.0, part + ((long long)rhs * gridDim.x + blockIdx.x) * 2);
}

// Residual loss, first half: d = u - b_row (u = A y), partial |d|^2; d overwrites u.
__global__ __launch_bounds__(256) void k_loss_res_d(float* u, const float* b, const int* row, long long sites,
                                                    double* part)
{
    const int rhs = blockIdx.y, r = row[rhs];
    float4* U = reinterpret_cast<float4*>(u) + (long long)rhs * sites;
    const float4* Bv = reinterpret_cast<const float4*>(b) + (long long)r * sites;
    double a = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < sites; i += (long long)gridDim.x * 256) {
        const float4 p = U[i], q = Bv[i];
        const float4 d = make_float4(p.x - q.x, p.y - q.y, p.z - q.z, p.w - q.w);
        a += (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z + (double)d.w * d.w;
        U[i] = d;
    }
    block_part2(a, 0.0, part + ((long long)rhs * gridDim.x + blockIdx.x) * 2);
}

// gy *= 2 / (npairs |b_row|^2)  (residual loss: gy = A d scaled)
__global__ void k_scale_rows(float* v, const int* row, const double* n2, long long sites, int npairs)
{
    const int rhs = blockIdx.y;
    const float sc = (float)(2.0 / ((double)npairs * n2[row[rhs]]));
    float4* V = reinterpret_cast<float4*>(v) + (long long)rhs * sites;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < sites; i += (long long)gridDim.x * 256) {
        float4 q = V[i];
        q.x *= sc;
        q.y *= sc;
        q.z *= sc;
        q.w *= sc;
        V[i] = q;
    }
}

// loss = (1/npairs) sum_rhs (sum_b part) / n2[row]  -> *out
__global__ void k_loss_fold(const double* part, int nb, const int* row, const double* n2, int nrhs, int npairs,
                            double* out)
{
    double L = 0.0;
    for (int r = 0; r < nrhs; ++r) {
        const double d = fold64(part + (long long)r * nb * 2, nb, 0);
        L += d / n2[row[r]];
    }
    if (threadIdx.x == 0) *out = L / npairs;
}

// ------------------------------------------- flexible CG (FP32 + FP64 reliable updates)
template <typename R> struct V4R;
template <> struct V4R<float> { using T = float4; };
template <> struct V4R<double> { using T = double4; };

template <typename F> __device__ __forceinline__ void for_sites(long long n, F f)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) f(i);
}

template <typename T4> __device__ __forceinline__ double dotr(const T4& a, const T4& b)
{
    return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
}

constexpr int kSkip = F_DONE | F_FROZEN;

// FP64: r = b - q; partials |r|^2, |b|^2
__global__ __launch_bounds__(256) void k_fcg_resid64(const double* b, const double* q, double* r, long long sites,
                                                     double* part)
{
    const int rhs = blockIdx.y;
    const long long o = (long long)rhs * sites;
    const double4* B = reinterpret_cast<const double4*>(b) + o;
    const double4* Q = reinterpret_cast<const double4*>(q) + o;
    double4* R = reinterpret_cast<double4*>(r) + o;
    double a = 0.0, c = 0.0;
    for_sites(sites, [&](long long i) {
        const double4 bb = B[i], qq = Q[i];
        const double4 rr = make_double4(bb.x - qq.x, bb.y - qq.y, bb.z - qq.z, bb.w - qq.w);
        R[i] = rr;
        a += dotr(rr, rr);
        c += dotr(bb, bb);
    });
    block_part2(a, c, part + ((long long)rhs * gridDim.x + blockIdx.x) * 2);
}

// Reliable update: the FP64 true residual certifies a right-hand side or restarts its
// FP32 cycle (first: the start of the solve).
__global__ void k_fcg_start(const double* part, int nb, CgState* st, double tol, double delta, int first)
{
    const int rhs = blockIdx.x;
    const double rr = fold64(part + (long long)rhs * nb * 2, nb, 0);
    const double bb = fold64(part + (long long)rhs * nb * 2, nb, 1);
    if (threadIdx.x == 0) {
        CgState& S = st[rhs];
        if (first) {
            S.iters = 0;
            S.reliable = 0;
            S.fp64_iters = 0;
            S.status = 0;
        } else if (!(S.flags[0] & F_DONE)) {
            S.reliable += 1;
        }
        S.bnorm2 = bb;
        S.tol = tol;
        S.delta = delta;
        S.rr_ref = rr;
        S.rr[0] = rr;
        S.true_relres = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
        S.relres = S.true_relres;
        const bool done = (S.status != 0) || bb == 0.0 || S.true_relres <= tol;
        S.flags[0] = done ? F_DONE : 0;
        S.flags[1] = S.flags[0];
    }
}

// r = (R) r64 and x = 0 for the right-hand sides still running
template <typename R>
__global__ __launch_bounds__(256) void k_fcg_restart(const double* r64, R* r, R* x, long long sites, const CgState* st)
{
    const int rhs = blockIdx.y;
    if (st[rhs].flags[0] & kSkip) return;
    using T4 = typename V4R<R>::T;
    const long long o = (long long)rhs * sites;
    const double4* S = reinterpret_cast<const double4*>(r64) + o;
    T4* Rr = reinterpret_cast<T4*>(r) + o;
    T4* X = reinterpret_cast<T4*>(x) + o;
    for_sites(sites, [&](long long i) {
        const double4 v = S[i];
        T4 w;
        w.x = (R)v.x;
        w.y = (R)v.y;
        w.z = (R)v.z;
        w.w = (R)v.w;
        Rr[i] = w;
        T4 zero;
        zero.x = zero.y = zero.z = zero.w = (R)0;
        X[i] = zero;
    });
}

// p = z; partial (z, r)
template <typename R>
__global__ __launch_bounds__(256) void k_fcg_p0(const R* z, const R* r, R* p, long long sites, const CgState* st,
                                                double* part)
{
    const int rhs = blockIdx.y;
    if (st[rhs].flags[0] & kSkip) return;
    using T4 = typename V4R<R>::T;
    const long long o = (long long)rhs * sites;
    const T4* Z = reinterpret_cast<const T4*>(z) + o;
    const T4* Rr = reinterpret_cast<const T4*>(r) + o;
    T4* P = reinterpret_cast<T4*>(p) + o;
    double a = 0.0;
    for_sites(sites, [&](long long i) {
        const T4 zz = Z[i];
        P[i] = zz;
        a += dotr(zz, Rr[i]);
    });
    block_part2(a, 0.0, part + ((long long)rhs * gridDim.x + blockIdx.x) * 2);
}

__global__ void k_fcg_rz(const double* part, int nb, CgState* st, FcgScal* fs)
{
    const int rhs = blockIdx.x;
    if (st[rhs].flags[0] & kSkip) return;
    const double zr = fold64(part + (long long)rhs * nb * 2, nb, 0);
    if (threadIdx.x == 0) {
        fs[rhs].rz = zr;
        if (!(zr > 0.0)) {  // preconditioner not positive on this residual
            st[rhs].status = 5;
            st[rhs].flags[0] |= F_DONE;
        }
    }
}

// Start of a cycle of the fused iteration: +inf for the previous (z, r) and (p, A p), so the
// first direction update's beta is 0 and p = z; no x update pending (alpha 0).
__global__ void k_fcg_first(CgState* st, FcgScal* fs, int nrhs)
{
    const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
    if (rhs >= nrhs) return;
    fs[rhs].alpha = 0.0;
    if (st[rhs].flags[0] & kSkip) return;
    fs[rhs].rz_prev = INFINITY;
    fs[rhs].pq = INFINITY;
}

// End of a cycle of the fused iteration: each right-hand side's last update x += alpha p,
// which the next iteration's D stencil would have applied (alpha 0: nothing pending).
__global__ __launch_bounds__(256) void k_fcg_flush_x(float* x, const float* p0, const float* p1,
                                                     const FcgScal* fs, long long sites)
{
    const int rhs = blockIdx.y;
    const double al = fs[rhs].alpha;
    if (al == 0.0) return;
    const float a = (float)al;
    const long long o = (long long)rhs * sites;
    const float4* P = reinterpret_cast<const float4*>(fs[rhs].pbuf != 0.0 ? p1 : p0) + o;
    float4* X = reinterpret_cast<float4*>(x) + o;
    for_sites(sites, [&](long long i) {
        const float4 pv = P[i];
        float4 xv = X[i];
        xv.x += a * pv.x;
        xv.y += a * pv.y;
        xv.z += a * pv.z;
        xv.w += a * pv.w;
        X[i] = xv;
    });
}

// After the update stencil: (p, A p) from the first stencil's partials (a breakdown if
// not positive: the update stencil skipped the right-hand side), then |r|^2 from the
// update stencil's partials: freeze the right-hand side for a reliable update when it
// met the tolerance or fell by delta since the last reliable update.
__global__ void k_fcg_check(const double* hpart, const double* rpart, int nb, CgState* st, FcgScal* fs)
{
    const int rhs = blockIdx.x;
    if (st[rhs].flags[0] & kSkip) return;
    const double pq = fold64(hpart + (long long)rhs * nb * 2, nb, 0);
    if (!(pq > 0.0)) {
        if (threadIdx.x == 0) {
            st[rhs].status = 4;
            st[rhs].flags[0] |= F_DONE;
        }
        return;
    }
    const double rr = fold64(rpart + (long long)rhs * nb * 2, nb, 0);
    if (threadIdx.x == 0) {
        fs[rhs].pq = pq;
        fs[rhs].rz_prev = fs[rhs].rz;  // consumed by this iteration's alpha
        CgState& S = st[rhs];
        S.rr[0] = rr;
        S.iters += 1;
        S.relres = std::sqrt(rr / S.bnorm2);
        if (S.relres <= S.tol || rr <= S.delta * S.delta * S.rr_ref) S.flags[0] |= F_FROZEN;
    }
}

// partials (z, r), (z, q)
template <typename R>
__global__ __launch_bounds__(256) void k_fcg_dots(const R* z, const R* r, const R* q, long long sites,
                                                  const CgState* st, double* part)
{
    const int rhs = blockIdx.y;
    if (st[rhs].flags[0] & kSkip) return;
    using T4 = typename V4R<R>::T;
    const long long o = (long long)rhs * sites;
    const T4* Z = reinterpret_cast<const T4*>(z) + o;
    const T4* Rr = reinterpret_cast<const T4*>(r) + o;
    const T4* Q = reinterpret_cast<const T4*>(q) + o;
    double a = 0.0, c = 0.0;
    for_sites(sites, [&](long long i) {
        const T4 zz = Z[i];
        a += dotr(zz, Rr[i]);
        c += dotr(zz, Q[i]);
    });
    block_part2(a, c, part + ((long long)rhs * gridDim.x + blockIdx.x) * 2);
}

// beta = -(z, A p) / (p, A p) (flexible CG, Notay's FCG(1)) or, Fletcher-Reeves, the
// ratio of the new and previous (z, r), from the (z, r), (z, q) partials folded by every
// block in the same order; then p = z + beta p.  Block 0 records rz = (z, r) for the next alpha
// (read by the next update stencil) or the breakdown when (z, r) is not positive.
template <typename R>
__global__ __launch_bounds__(256) void k_fcg_pbeta(const R* z, R* p, long long sites, CgState* st, FcgScal* fs,
                                                   const double* part, int nb, int fletcher_reeves)
{
    const int rhs = blockIdx.y;
    if (st[rhs].flags[0] & kSkip) return;
    __shared__ double sh[2];
    if (threadIdx.x < 64) {
        const double zr = fold64(part + (long long)rhs * nb * 2, nb, 0);
        const double zq = fold64(part + (long long)rhs * nb * 2, nb, 1);
        if (threadIdx.x == 0) {
            sh[0] = zr;
            sh[1] = zq;
        }
    }
    __syncthreads();
    const double zr = sh[0], zq = sh[1];
    if (!(zr > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st[rhs].status = 5;
            st[rhs].flags[0] |= F_DONE;
        }
        return;
    }
    const double beta = fletcher_reeves ? zr / fs[rhs].rz_prev : -zq / fs[rhs].pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        fs[rhs].rz = zr;
        fs[rhs].beta = beta;
    }
    using T4 = typename V4R<R>::T;
    const R be = (R)beta;
    const long long o = (long long)rhs * sites;
    const T4* Z = reinterpret_cast<const T4*>(z) + o;
    T4* P = reinterpret_cast<T4*>(p) + o;
    for_sites(sites, [&](long long i) {
        const T4 zz = Z[i];
        T4 pp = P[i];
        pp.x = zz.x + be * pp.x;
        pp.y = zz.y + be * pp.y;
        pp.z = zz.z + be * pp.z;
        pp.w = zz.w + be * pp.w;
        P[i] = pp;
    });
}

__global__ void k_iota_cfg(int* cfg, int nrhs, int rhs_per_cfg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrhs) cfg[i] = i / rhs_per_cfg;
}

// ------------------------------------------------------------------- host side
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

void hchk(hipError_t e, const char* what) { dwf_ctx::check(e, what); }

int nb_vec(long long sites) { return (int)std::max(1LL, std::min(256LL, (sites + 1023) / 1024)); }

// Launch one layer: forward weights (adj = 0) or the adjoint kernel of layer l.
struct LayerIO {
    const void* in;
    const int* in_row;
    void* out;
    int in_kind;   // 0 double spinor, 1 float spinor, 2 bf16 channels
    int out_kind;  // 0 double spinor, 1 float spinor, 2 bf16 channels
};

PtcArgs ptc_args(const dwf_ctx* c, const dwf_net* net, int l, bool adj, const LayerIO& io, const int* cfg, int nrhs,
                 const CgState* st)
{
    const Lat& L = c->L;
    PtcArgs a{};
    a.Nt = L.Nt;
    a.Nx = L.Nx;
    a.Ls = L.Ls;
    a.V4 = L.V4;
    a.tps = L.V4 / 16;
    a.ntiles = (long long)nrhs * L.Ls * a.tps;
    a.in = io.in;
    a.out = io.out;
    a.in_row = io.in_row;
    a.cfg = cfg;
    a.links = c->l32;
    a.link_stride = 2LL * L.V4;
    a.wpk = adj ? net->wa[l] : net->wf[l];
    a.st = st;
    a.tpw = 2 * (int)std::max(1LL, std::min(8LL, a.ntiles / 8192));  // even: tiles go in pairs
    return a;
}

void launch_layer(dwf_ctx* c, const dwf_net* net, int l, bool adj, const LayerIO& io, const int* cfg, int nrhs,
                  const CgState* st)
{
    const PtcArgs a = ptc_args(c, net, l, adj, io, cfg, nrhs, st);
    const long long waves = (a.ntiles + a.tpw - 1) / a.tpw;
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    const int cin = adj ? kCoutL[l] : kCinL[l], cout = adj ? kCinL[l] : kCoutL[l];
    hipStream_t s = c->stream;
    if (cin == 2 && cout == 8) {
        if (io.in_kind == 0)
            k_ptc_fwd<2, 8, double, unsigned short><<<blocks, 256, 0, s>>>(a);
        else
            k_ptc_fwd<2, 8, float, unsigned short><<<blocks, 256, 0, s>>>(a);
    } else if (cin == 8 && cout == 8) {
        k_ptc_fwd<8, 8, unsigned short, unsigned short><<<blocks, 256, 0, s>>>(a);
    } else {  // 8 -> 2
        if (io.out_kind == 0)
            k_ptc_fwd<8, 2, unsigned short, double><<<blocks, 256, 0, s>>>(a);
        else
            k_ptc_fwd<8, 2, unsigned short, float><<<blocks, 256, 0, s>>>(a);
    }
}

void require_net(const dwf_ctx* c, const dwf_net* net)
{
    if (!c || !net) throw DwfError(DWF_ERR_ARG, "learned preconditioner: null argument");
    if (c->kind != 0) throw DwfError(DWF_ERR_ARG, "learned preconditioner: needs a domain-wall context");
    if (!c->have_gauge) throw DwfError(DWF_ERR_ARG, "learned preconditioner: no gauge field set");
    c->require_single_lattice("the learned preconditioner");
    if (net->Ls != c->L.Ls) throw DwfError(DWF_ERR_ARG, "learned preconditioner: network Ls differs from the lattice's");
    if (c->L.V4 % 16 != 0)
        throw DwfError(DWF_ERR_ARG, "learned preconditioner: Nt*Nx must be a multiple of 16");
}

void repack(const dwf_net* net, hipStream_t s)
{
    for (int l = 0; l < 3; ++l) {
        const long long nf = frag_shorts(kCinL[l], net->Ls), na = frag_shorts(kCoutL[l], net->Ls);
        k_pack<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(net->w[l], kCinL[l], kCoutL[l], net->Ls, 0, net->wf[l]);
        k_pack<<<(unsigned)((na + 255) / 256), 256, 0, s>>>(net->w[l], kCinL[l], kCoutL[l], net->Ls, 1, net->wa[l]);
    }
    const long long ng = frag_shorts(kCinL[0], net->Ls);
    k_pack<<<(unsigned)((ng + 255) / 256), 256, 0, s>>>(net->w[0], kCinL[0], kCoutL[0], net->Ls, 2, net->wg0);
    dwf_ctx::check_launch("pack weights");
}

// Output tile of the fused network kernel: the largest divisors of Nt (<= 8) and Nx (<= 16).
bool fused_shape(const dwf_ctx* c, int& TY, int& TX)
{
    const Lat& L = c->L;
    if (L.Nt < 3 || L.Nx < 3) return false;
    if (c->net_tile > 0) {  // set by hand (DWF_TUNE_NET_TILE), validated there
        TY = c->net_tile >> 8;
        TX = c->net_tile & 255;
        return true;
    }
    TY = 1;
    TX = 1;
    for (int d = 1; d <= kFuTy; ++d)
        if (L.Nt % d == 0) TY = d;
    for (int d = 1; d <= kFuTx; ++d)
        if (L.Nx % d == 0) TX = d;
    return true;
}

// Parts of the fifth dimension per output tile of the fused network (one block each): a
// launch of few blocks (few right-hand sides) is bound by the latency of its walk, one
// barrier-separated step per slice, so splitting the walk buys blocks and shortens each
// block's walk at the price of the parts' recomputed edge slices.  DWF_TUNE_NET_SPLIT sets
// it; 0 picks the most parts (up to 4, each at least 4 slices) that keep the launch within
// kNetSplitBlocks blocks.
constexpr int kNetSplitBlocks = 256;
int net_split(const dwf_ctx* c, int tiles, int nrhs)
{
    const int Ls = c->L.Ls;
    if (c->net_split > 0) return c->net_split;
    int n = 1;
    for (int d = 2; d <= 4; d *= 2)
        if (Ls % d == 0 && Ls / d >= 4 && (long long)tiles * nrhs * d <= kNetSplitBlocks) n = d;
    return n;
}

// Fused z = M in (k_net_fused); with part also the (z, in) and, with q, (z, q) partials,
// part [nrhs][tiles * net_split][2].  Returns the number of partials per right-hand side
// (tiles per lattice times parts), or 0 if the shape does not allow the fused kernel (then
// nothing is launched).
// The FCG iteration's check folded into the fused network launch (see NetFusedArgs).
struct FusedCheck {
    const double* hpart;
    const double* rpart;
    int hnb;
    CgState* st;
    FcgScal* fs;
};

int launch_net_fused(dwf_ctx* c, const dwf_net* net, const void* in, int in_kind, void* out, int out_kind,
                     const int* cfg, int nrhs, const CgState* st, const float* q, double* part,
                     const FusedCheck* chk = nullptr)
{
    const Lat& L = c->L;
    int TY, TX;
    if (!c->net_fused || !fused_shape(c, TY, TX)) return 0;
    NetFusedArgs a{};
    a.Nt = L.Nt;
    a.Nx = L.Nx;
    a.Ls = L.Ls;
    a.V4 = L.V4;
    a.TY = TY;
    a.TX = TX;
    a.tiles_x = L.Nx / TX;
    a.tiles = (L.Nt / TY) * a.tiles_x;
    a.nsplit = net_split(c, a.tiles, nrhs);
    a.in = in;
    a.out = out;
    a.links = c->l32;
    a.link_stride = 2LL * L.V4;
    a.cfg = cfg;
    for (int l = 0; l < 3; ++l) a.wf[l] = net->wf[l];
    const bool tg = c->net_fused == 2;
    if (tg) a.wf[0] = net->wg0;
    a.st = st;
    a.q = q;
    a.part = part;
    if (chk) {
        if (!part) throw DwfError(DWF_ERR_ARG, "fused network: the check needs the dot-product epilogue");
        a.hpart = chk->hpart;
        a.rpart = chk->rpart;
        a.hnb = chk->hnb;
        a.cst = chk->st;
        a.fs = chk->fs;
    }
    const unsigned grid = (unsigned)(a.tiles * nrhs * a.nsplit);
    hipStream_t s = c->stream;
    if (part) {
        if (in_kind != 1 || out_kind != 1) throw DwfError(DWF_ERR_ARG, "fused network: dot products need FP32 spinors");
        if (tg)
            k_net_fused<float, float, true, true><<<grid, kFuThreads, 0, s>>>(a);
        else
            k_net_fused<float, float, true, false><<<grid, kFuThreads, 0, s>>>(a);
    } else if (in_kind == 0 && out_kind == 0) {
        if (tg)
            k_net_fused<double, double, false, true><<<grid, kFuThreads, 0, s>>>(a);
        else
            k_net_fused<double, double, false, false><<<grid, kFuThreads, 0, s>>>(a);
    } else if (in_kind == 1 && out_kind == 1) {
        if (tg)
            k_net_fused<float, float, false, true><<<grid, kFuThreads, 0, s>>>(a);
        else
            k_net_fused<float, float, false, false><<<grid, kFuThreads, 0, s>>>(a);
    } else {
        throw DwfError(DWF_ERR_ARG, "fused network: unsupported spinor types");
    }
    return a.tiles * a.nsplit;
}

// z = M in for nrhs right-hand sides, configuration per cfg[rhs]: the fused kernel when the
// shape allows it and the input is not gathered by rows, else layer by layer through h1, h2.
void net_forward(dwf_ctx* c, const dwf_net* net, const void* in, int in_kind, const int* in_row, void* out,
                 int out_kind, const int* cfg, int nrhs, const CgState* st, unsigned short* h1, unsigned short* h2)
{
    if (!in_row && launch_net_fused(c, net, in, in_kind, out, out_kind, cfg, nrhs, st, nullptr, nullptr)) return;
    launch_layer(c, net, 0, false, LayerIO{in, in_row, h1, in_kind, 2}, cfg, nrhs, st);
    launch_layer(c, net, 1, false, LayerIO{h1, nullptr, h2, 2, 2}, cfg, nrhs, st);
    launch_layer(c, net, 2, false, LayerIO{h2, nullptr, out, 2, out_kind}, cfg, nrhs, st);
}

void net_adjoint(dwf_ctx* c, const dwf_net* net, const void* in, int in_kind, void* out, int out_kind,
                 const int* cfg, int nrhs, unsigned short* h1, unsigned short* h2)
{
    launch_layer(c, net, 2, true, LayerIO{in, nullptr, h2, in_kind, 2}, cfg, nrhs, nullptr);
    launch_layer(c, net, 1, true, LayerIO{h2, nullptr, h1, 2, 2}, cfg, nrhs, nullptr);
    launch_layer(c, net, 0, true, LayerIO{h1, nullptr, out, 2, out_kind}, cfg, nrhs, nullptr);
}

int* cfg_identity_map(dwf_ctx* c, int nrhs)
{
    int* cfg = (int*)c->buf(B_NET_CFG, (size_t)nrhs * sizeof(int));
    k_iota_cfg<<<(nrhs + 63) / 64, 64, 0, c->stream>>>(cfg, nrhs, c->ncfg <= 1 ? (1 << 30) : c->rhs_per_cfg);
    return cfg;
}

template <typename R> HopArgs<R> hop_args(const dwf_ctx* c, int nrhs, const int* cfg_map)
{
    HopArgs<R> a{};
    a.L = c->L;
    a.cfg = hop_cfg(c->L, nrhs, c->chunks_override, c->xcd_remap, c->prefetch, c->tile_x, c->nt_store, c->lds);
    a.links = sizeof(R) == 8 ? (const void*)c->l64 : (const void*)c->l32;
    a.link_stride = 2LL * c->L.V4;
    a.rhs_per_cfg = c->ncfg <= 1 ? (1 << 30) : c->rhs_per_cfg;
    a.cfg_map = cfg_map;
    return a;
}

// Weight gradients of all layers from the stored activations (training step core).
// Buffers: h0 the network input rows, h1/h2 the hidden activations (overwritten by the
// activation gradients), gy the output gradient (float spinors).
void backward(dwf_ctx* c, dwf_net* net, const float* h0, const int* row, const int* cfg, int nrhs, unsigned short* h1,
              unsigned short* h2, const float* gy, double* GB[3])
{
    const Lat& L = c->L;
    hipStream_t s = c->stream;
    const long long per_s = (long long)nrhs * (L.V4 / 16);
    const int nbx = (int)std::max(1LL, std::min(128LL, per_s / 8));
    float* part = (float*)c->buf(B_NET_WPART, (size_t)L.Ls * nbx * 7 * 256 * sizeof(float));
    auto reduce = [&](int cin, double* out) {
        const int per = nstep_of(cin) * 256;
        const long long n = (long long)L.Ls * per;
        k_wgrad_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(part, nbx, per, L.Ls, out);
    };
    const dim3 grid(nbx, L.Ls);
    // layer 3 (8 -> 2): input h2, output gradient gy
    {
        const PtcArgs a = ptc_args(c, net, 2, false, LayerIO{h2, nullptr, nullptr, 2, 1}, cfg, nrhs, nullptr);
        k_ptc_wgrad<8, 2, unsigned short, float><<<grid, 256, 0, s>>>(a, gy, part, nrhs);
        reduce(8, GB[2]);
    }
    // g2 = L3^dag gy  (into h2)
    launch_layer(c, net, 2, true, LayerIO{gy, nullptr, h2, 1, 2}, cfg, nrhs, nullptr);
    // layer 2 (8 -> 8): input h1, output gradient g2
    {
        const PtcArgs a = ptc_args(c, net, 1, false, LayerIO{h1, nullptr, nullptr, 2, 2}, cfg, nrhs, nullptr);
        k_ptc_wgrad<8, 8, unsigned short, unsigned short><<<grid, 256, 0, s>>>(a, h2, part, nrhs);
        reduce(8, GB[1]);
    }
    // g1 = L2^dag g2  (into h1)
    launch_layer(c, net, 1, true, LayerIO{h2, nullptr, h1, 2, 2}, cfg, nrhs, nullptr);
    // layer 1 (2 -> 8): input h0 rows, output gradient g1
    {
        const PtcArgs a = ptc_args(c, net, 0, false, LayerIO{h0, row, nullptr, 1, 2}, cfg, nrhs, nullptr);
        k_ptc_wgrad<2, 8, float, unsigned short><<<grid, 256, 0, s>>>(a, h1, part, nrhs);
        reduce(2, GB[0]);
    }
    dwf_ctx::check_launch("network backward");
}

// Loss and output gradient for one batch; leaves gy and the loss value in *loss_dev.
void loss_and_grad(dwf_ctx* c, int loss, const float* y, const float* b, const float* x, const int* row,
                   const int* cfg, const double* n2, int nrhs, int npairs, float* gy, double* loss_dev)
{
    const Lat& L = c->L;
    const long long sites = L.V5();
    hipStream_t s = c->stream;
    const int nb = nb_vec(sites);
    double* part = (double*)c->buf(B_NET_PART, (size_t)nrhs * nb * 2 * 8);
    if (loss == 0) {
        k_loss_sol<<<dim3(nb, nrhs), 256, 0, s>>>(y, x, row, n2, sites, npairs, gy, part);
    } else {
        // u = A y (A = D^dag D, FP32 stencils), d = u - b, gy = 2 A d / (npairs |b|^2)
        float* t = (float*)c->buf(B_NET_T, (size_t)nrhs * sites * 16);
        HopArgs<float> a = hop_args<float>(c, nrhs, cfg);
        a.in = y;
        a.out = t;
        launch_hop<float>(false, EPI_STORE, a, nrhs, s);
        a.in = t;
        a.out = gy;
        launch_hop<float>(true, EPI_STORE, a, nrhs, s);
        k_loss_res_d<<<dim3(nb, nrhs), 256, 0, s>>>(gy, b, row, sites, part);
        a.in = gy;
        a.out = t;
        launch_hop<float>(false, EPI_STORE, a, nrhs, s);
        a.in = t;
        a.out = gy;
        launch_hop<float>(true, EPI_STORE, a, nrhs, s);
        k_scale_rows<<<dim3(nb, nrhs), 256, 0, s>>>(gy, row, n2, sites, npairs);
    }
    k_loss_fold<<<1, 64, 0, s>>>(part, nb, row, n2, nrhs, npairs, loss_dev);
    dwf_ctx::check_launch("loss");
}

struct TrainBufs {
    unsigned short *h1, *h2;
    float *y, *gy;
    int *row, *cfg;
    double* GB[3];
};

TrainBufs train_bufs(dwf_ctx* c, int nrhs)
{
    const long long sites = c->L.V5();
    TrainBufs t;
    t.h1 = (unsigned short*)c->buf(B_NET_H1, (size_t)nrhs * sites * 32);
    t.h2 = (unsigned short*)c->buf(B_NET_H2, (size_t)nrhs * sites * 32);
    t.y = (float*)c->buf(B_NET_Y, (size_t)nrhs * sites * 16);
    t.gy = (float*)c->buf(B_NET_GY, (size_t)nrhs * sites * 16);
    t.row = (int*)c->buf(B_NET_ROW, (size_t)nrhs * sizeof(int));
    t.cfg = (int*)c->buf(B_NET_CFG, (size_t)nrhs * sizeof(int));
    double* gb = (double*)c->buf(B_NET_GB, (size_t)c->L.Ls * (2 + 7 + 7) * 256 * 8);
    t.GB[0] = gb;
    t.GB[1] = gb + (size_t)c->L.Ls * 2 * 256;
    t.GB[2] = gb + (size_t)c->L.Ls * 9 * 256;
    return t;
}

void check_pairs(const dwf_ctx* c, const float* b, const float* x, const int* pair_cfg, int npairs, int loss)
{
    if (!b || (loss == 0 && !x) || !pair_cfg || npairs < 1)
        throw DwfError(DWF_ERR_ARG, "learned preconditioner: bad training pairs");
    if (loss != 0 && loss != 1) throw DwfError(DWF_ERR_ARG, "learned preconditioner: loss must be 0 or 1");
    for (int i = 0; i < npairs; ++i)
        if (pair_cfg[i] < 0 || pair_cfg[i] >= c->ncfg)
            throw DwfError(DWF_ERR_ARG, "learned preconditioner: pair configuration outside the configurations held");
}

}  // namespace

dwf_net::~dwf_net()
{
    for (int l = 0; l < 3; ++l) {
        (void)hipFree(w[l]);
        (void)hipFree(m1[l]);
        (void)hipFree(m2[l]);
        (void)hipFree(wf[l]);
        (void)hipFree(wa[l]);
        if (l == 0) (void)hipFree(wg0);
    }
    if (pack_stream) (void)hipStreamDestroy(pack_stream);
}

// ---------------------------------------------------------------- C ABI
extern "C" long long dwf_net_num_weights(int Ls)
{
    if (Ls < 1) return 0;
    long long n = 0;
    for (int l = 0; l < 3; ++l) n += layer_cw(l, Ls);
    return 2 * n;
}

extern "C" int dwf_net_create(dwf_net** out, int Ls, unsigned long long seed)
{
    return guard([&]() -> int {
        if (!out || Ls < 2 || Ls > 64) throw DwfError(DWF_ERR_ARG, "dwf_net_create: Ls must be 2..64");
        *out = nullptr;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            throw DwfError(DWF_ERR_HIP, "no HIP device available: libdwfhip needs an MI355X (gfx950) GPU");
        std::unique_ptr<dwf_net> net(new dwf_net);
        net->Ls = Ls;
        hchk(hipGetDevice(&net->device), "hipGetDevice");
        hchk(hipStreamCreateWithFlags(&net->pack_stream, hipStreamNonBlocking), "hipStreamCreate");
        for (int l = 0; l < 3; ++l) {
            const size_t nw = (size_t)layer_cw(l, Ls) * sizeof(float2);
            hchk(hipMalloc(&net->w[l], nw), "hipMalloc weights");
            hchk(hipMalloc(&net->m1[l], nw), "hipMalloc weights");
            hchk(hipMalloc(&net->m2[l], nw), "hipMalloc weights");
            hchk(hipMalloc(&net->wf[l], (size_t)frag_shorts(kCinL[l], Ls) * 2), "hipMalloc weights");
            hchk(hipMalloc(&net->wa[l], (size_t)frag_shorts(kCoutL[l], Ls) * 2), "hipMalloc weights");
            if (l == 0) hchk(hipMalloc(&net->wg0, (size_t)frag_shorts(kCinL[0], Ls) * 2), "hipMalloc weights");
        }
        // Initial weights: identity embeddings on the self point (layer 3 scaled by 1/4,
        // about the inverse of D^dag D's typical eigenvalue), small random elsewhere.
        std::vector<float> w((size_t)dwf_net_num_weights(Ls));
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.0f, 1e-3f);
        size_t off = 0;
        for (int l = 0; l < 3; ++l) {
            const int ci = kCinL[l], co = kCoutL[l];
            for (int s = 0; s < Ls; ++s)
                for (int p = 0; p < kPts; ++p)
                    for (int o = 0; o < co; ++o)
                        for (int cc = 0; cc < ci; ++cc) {
                            float re = nd(rng), im = nd(rng);
                            if (p == 0 && o == cc && o < 2) re += (l == 2) ? 0.25f : 1.0f;
                            if (p == 0 && l == 1 && o == cc && o >= 2) re += 1.0f;
                            w[off++] = re;
                            w[off++] = im;
                        }
        }
        dwf_net* raw = net.release();
        const int rc = dwf_net_set_weights(raw, w.data());
        if (rc != DWF_OK) {
            delete raw;
            return rc;
        }
        *out = raw;
        return DWF_OK;
    });
}

extern "C" int dwf_net_destroy(dwf_net* net)
{
    if (net) {
        (void)hipDeviceSynchronize();
        delete net;
    }
    return DWF_OK;
}

extern "C" int dwf_net_set_weights(dwf_net* net, const float* w)
{
    return guard([&]() -> int {
        if (!net || !w) throw DwfError(DWF_ERR_ARG, "dwf_net_set_weights: null argument");
        size_t off = 0;
        for (int l = 0; l < 3; ++l) {
            const size_t n = (size_t)layer_cw(l, net->Ls);
            hchk(hipMemcpyAsync(net->w[l], w + off, n * 8, hipMemcpyHostToDevice, net->pack_stream), "weights upload");
            hchk(hipMemsetAsync(net->m1[l], 0, n * 8, net->pack_stream), "weights");
            hchk(hipMemsetAsync(net->m2[l], 0, n * 8, net->pack_stream), "weights");
            off += 2 * n;
        }
        net->adam_t = 0;
        repack(net, net->pack_stream);
        hchk(hipStreamSynchronize(net->pack_stream), "weights upload");
        return DWF_OK;
    });
}

extern "C" int dwf_net_get_weights(const dwf_net* net, float* w)
{
    return guard([&]() -> int {
        if (!net || !w) throw DwfError(DWF_ERR_ARG, "dwf_net_get_weights: null argument");
        hchk(hipDeviceSynchronize(), "weights download");
        size_t off = 0;
        for (int l = 0; l < 3; ++l) {
            const size_t n = (size_t)layer_cw(l, net->Ls);
            hchk(hipMemcpy(w + off, net->w[l], n * 8, hipMemcpyDeviceToHost), "weights download");
            off += 2 * n;
        }
        return DWF_OK;
    });
}

extern "C" int dwf_net_apply_device(dwf_ctx* c, const dwf_net* net, int adjoint, const double* in, double* out,
                                    int nrhs)
{
    return guard([&]() -> int {
        require_net(c, net);
        if (!in || !out || nrhs < 1) throw DwfError(DWF_ERR_ARG, "dwf_net_apply_device: bad argument");
        c->bind_rhs(nrhs);
        const long long sites = c->L.V5();
        unsigned short* h1 = (unsigned short*)c->buf(B_NET_H1, (size_t)nrhs * sites * 32);
        unsigned short* h2 = (unsigned short*)c->buf(B_NET_H2, (size_t)nrhs * sites * 32);
        const int* cfg = cfg_identity_map(c, nrhs);
        if (adjoint)
            net_adjoint(c, net, in, 0, out, 0, cfg, nrhs, h1, h2);
        else
            net_forward(c, net, in, 0, nullptr, out, 0, cfg, nrhs, nullptr, h1, h2);
        c->check_launch("network apply");
        return DWF_OK;
    });
}

extern "C" int dwf_net_loss_grad_device(dwf_ctx* c, dwf_net* net, int loss, const float* b, const float* x,
                                        const int* pair_cfg, int npairs, double* loss_out, float* grad)
{
    return guard([&]() -> int {
        require_net(c, net);
        check_pairs(c, b, x, pair_cfg, npairs, loss);
        if (!loss_out || !grad) throw DwfError(DWF_ERR_ARG, "dwf_net_loss_grad_device: null output");
        const long long sites = c->L.V5();
        hipStream_t s = c->stream;
        TrainBufs tb = train_bufs(c, npairs);
        std::vector<int> rows(npairs);
        std::iota(rows.begin(), rows.end(), 0);
        hchk(hipMemcpyAsync(tb.row, rows.data(), npairs * sizeof(int), hipMemcpyHostToDevice, s), "upload");
        hchk(hipMemcpyAsync(tb.cfg, pair_cfg, npairs * sizeof(int), hipMemcpyHostToDevice, s), "upload");
        double* n2 = (double*)c->buf(B_NET_N2, (size_t)npairs * 8);
        k_row_norms<<<npairs, 256, 0, s>>>(loss == 0 ? x : b, sites, n2);
        net_forward(c, net, b, 1, tb.row, tb.y, 1, tb.cfg, npairs, nullptr, tb.h1, tb.h2);
        double* ld = (double*)c->buf(B_NET_LOSS, 8);
        loss_and_grad(c, loss, tb.y, b, x, tb.row, tb.cfg, n2, npairs, npairs, tb.gy, ld);
        backward(c, net, b, tb.row, tb.cfg, npairs, tb.h1, tb.h2, tb.gy, tb.GB);
        float* gdev = (float*)c->buf(B_NET_GRAD, (size_t)dwf_net_num_weights(net->Ls) * 4);
        size_t off = 0;
        for (int l = 0; l < 3; ++l) {
            const int n = (int)layer_cw(l, net->Ls);
            k_fold_adam<<<(n + 255) / 256, 256, 0, s>>>(tb.GB[l], kCinL[l], kCoutL[l], net->Ls, net->w[l], net->m1[l],
                                                         net->m2[l], gdev + off, 0, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
            off += 2 * (size_t)n;
        }
        c->check_launch("loss gradient");
        hchk(hipMemcpyAsync(grad, gdev, off * 4, hipMemcpyDeviceToHost, s), "gradient download");
        hchk(hipMemcpyAsync(loss_out, ld, 8, hipMemcpyDeviceToHost, s), "loss download");
        hchk(hipStreamSynchronize(s), "loss gradient");
        return DWF_OK;
    });
}

extern "C" int dwf_net_train_device(dwf_ctx* c, dwf_net* net, const float* b, const float* x, const int* pair_cfg,
                                    int npairs, const dwf_train_params* p, double* loss_history)
{
    return guard([&]() -> int {
        require_net(c, net);
        if (!p || p->steps < 1 || p->batch < 1 || !(p->lr > 0.0))
            throw DwfError(DWF_ERR_ARG, "dwf_net_train_device: bad training parameters");
        check_pairs(c, b, x, pair_cfg, npairs, p->loss);
        const int B = std::min(p->batch, npairs);
        const long long sites = c->L.V5();
        hipStream_t s = c->stream;
        TrainBufs tb = train_bufs(c, B);
        double* n2 = (double*)c->buf(B_NET_N2, (size_t)npairs * 8);
        k_row_norms<<<npairs, 256, 0, s>>>(p->loss == 0 ? x : b, sites, n2);
        double* hist = (double*)c->buf(B_NET_LOSS, (size_t)p->steps * 8);
        // the batch tables (rows, configurations) of every step, uploaded once
        int* stage = (int*)c->buf(B_NET_STAGE, (size_t)p->steps * 2 * B * sizeof(int));
        std::vector<int> order(npairs), tab((size_t)p->steps * 2 * B);
        std::iota(order.begin(), order.end(), 0);
        std::mt19937_64 rng(p->seed);
        int pos = npairs;
        for (int st = 0; st < p->steps; ++st)
            for (int j = 0; j < B; ++j) {
                if (pos == npairs) {
                    std::shuffle(order.begin(), order.end(), rng);
                    pos = 0;
                }
                const int r = order[pos++];
                tab[(size_t)st * 2 * B + j] = r;
                tab[(size_t)st * 2 * B + B + j] = pair_cfg[r];
            }
        hchk(hipMemcpyAsync(stage, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, s), "upload");
        const float b1 = (float)p->beta1, b2 = (float)p->beta2;
        for (int st = 0; st < p->steps; ++st) {
            const int* row = stage + (size_t)st * 2 * B;
            const int* cfg = row + B;
            net_forward(c, net, b, 1, row, tb.y, 1, cfg, B, nullptr, tb.h1, tb.h2);
            loss_and_grad(c, p->loss, tb.y, b, x, row, cfg, n2, B, B, tb.gy, hist + st);
            backward(c, net, b, row, cfg, B, tb.h1, tb.h2, tb.gy, tb.GB);
            net->adam_t += 1;
            const float c1 = 1.f / (1.f - std::pow(b1, (float)net->adam_t));
            const float c2 = 1.f / (1.f - std::pow(b2, (float)net->adam_t));
            for (int l = 0; l < 3; ++l) {
                const int n = (int)layer_cw(l, net->Ls);
                k_fold_adam<<<(n + 255) / 256, 256, 0, s>>>(tb.GB[l], kCinL[l], kCoutL[l], net->Ls, net->w[l],
                                                             net->m1[l], net->m2[l], nullptr, 1, (float)p->lr, b1, b2,
                                                             (float)p->eps, c1, c2);
            }
            repack(net, s);
            c->check_launch("training step");
        }
        if (loss_history)
            hchk(hipMemcpyAsync(loss_history, hist, (size_t)p->steps * 8, hipMemcpyDeviceToHost, s), "loss download");
        hchk(hipStreamSynchronize(s), "training");
        return DWF_OK;
    });
}

extern "C" int dwf_ctx_set_preconditioner(dwf_ctx* c, const dwf_net* net)
{
    return guard([&]() -> int {
        if (!c) throw DwfError(DWF_ERR_ARG, "dwf_ctx_set_preconditioner: null context");
        if (net && net->Ls != c->L.Ls)
            throw DwfError(DWF_ERR_ARG, "learned preconditioner: network Ls differs from the lattice's");
        if (net != c->net) c->drop_graphs();  // cached iterations captured the old network's weights
        c->net = net;
        return DWF_OK;
    });
}

// Flexible CG on D^dag D x = b with z = M r: FP32 cycles (stencils, vectors and the
// network's FP32 inputs/outputs) between FP64 reliable updates, which accumulate the
// FP32 correction into x, recompute the FP64 true residual and either certify the
// right-hand side or restart its cycle from that residual.  Leaves states in st.
int dwf_fcg_solve(dwf_ctx* c, const double* b, double* x, bool have_guess, double rtol, int maxiter, int nrhs,
                  dwf_solve_stats* stats)
{
    const dwf_net* net = c->net;
    if (!net) throw DwfError(DWF_ERR_ARG, "DWF_SOLVE_LEARNED: no preconditioner attached (dwf_ctx_set_preconditioner)");
    require_net(c, net);
    const long long sites = c->L.V5();
    const size_t nv = (size_t)sites * 4 * nrhs;
    hipStream_t s = c->stream;
    double* r64 = (double*)c->buf(B_FCG_R, nv * 8);
    double* q64 = (double*)c->buf(B_FCG_Q, nv * 8);
    double* t64 = (double*)c->buf(B_T64, nv * 8);
    float* r32 = (float*)c->buf(B_FCG_R32, nv * 4);
    float* p32 = (float*)c->buf(B_FCG_P32, nv * 4);
    float* p32b = (float*)c->buf(B_FCG_P32B, nv * 4);
    float* q32 = (float*)c->buf(B_FCG_Q32, nv * 4);
    float* z32 = (float*)c->buf(B_FCG_Z32, nv * 4);
    float* t32 = (float*)c->buf(B_FCG_T32, nv * 4);
    float* x32 = (float*)c->buf(B_FCG_X32, nv * 4);
    unsigned short* h1 = (unsigned short*)c->buf(B_NET_H1, (size_t)nrhs * sites * 32);
    unsigned short* h2 = (unsigned short*)c->buf(B_NET_H2, (size_t)nrhs * sites * 32);
    const int nb = nb_vec(sites);
    double* part = (double*)c->buf(B_FCG_PART, (size_t)nrhs * nb * 2 * 8);
    CgState* st = (CgState*)c->buf(B_STATE, sizeof(CgState) * nrhs);
    FcgScal* fs = (FcgScal*)c->buf(B_FCG_SCAL, sizeof(FcgScal) * nrhs);
    const int* cfg = cfg_identity_map(c, nrhs);
    HopArgs<float> ha = hop_args<float>(c, nrhs, nullptr);
    const int hnb = ha.cfg.nblk();
    double* hpart = (double*)c->buf(B_PA, (size_t)nrhs * hnb * 2 * 8);
    double* rpart = (double*)c->buf(B_PB, (size_t)nrhs * hnb * 2 * 8);
    ha.st = st;
    ha.slot = 0;
    ha.skip = kSkip;
    const dim3 vgrid(nb, nrhs);
    const int fr = c->fcg_beta;
    int fty = 0, ftx = 0;
    const int ftiles = fused_shape(c, fty, ftx) ? (c->L.Nt / fty) * (c->L.Nx / ftx) : 1;
    const int ntiles = ftiles * net_split(c, ftiles, nrhs);  // the fused network's partials
    double* npart = (double*)c->buf(B_FCG_NPART, (size_t)nrhs * ntiles * 2 * 8);

    if (!have_guess) launch_zero<double>(x, (long long)nv, s);
    std::vector<CgState> h(nrhs);
    auto read = [&]() {
        hchk(hipMemcpyAsync(h.data(), st, sizeof(CgState) * nrhs, hipMemcpyDeviceToHost, s), "state readback");
        hchk(hipStreamSynchronize(s), "state readback");
    };
    auto reliable_update = [&](bool first) {  // r64 = b - A x (FP64), certify or restart
        HopArgs<double> a = hop_args<double>(c, nrhs, nullptr);
        a.in = x;
        a.out = t64;
        launch_hop<double>(false, EPI_STORE, a, nrhs, s);
        a.in = t64;
        a.out = q64;
        launch_hop<double>(true, EPI_STORE, a, nrhs, s);
        k_fcg_resid64<<<vgrid, 256, 0, s>>>(b, q64, r64, sites, part);
        k_fcg_start<<<nrhs, 64, 0, s>>>(part, nb, st, rtol, c->reliable_delta, first ? 1 : 0);
    };
    // The fused iteration (fused network kernel, LDS-staged stencils): the D stencil forms the
    // direction p = z + beta p_old on the fly (EPI_PNORM, beta from the network's partials),
    // writes it to the other direction buffer, computes t = D p with |t|^2 and applies the
    // previous iteration's x += alpha p_old (it reads p_old anyway); the D^dag stencil updates
    // r and records alpha (EPI_RUPD); the network checks the iteration and computes z = M r with its
    // dot products.  Three kernels; directions alternate between p32 and p32b by parity; the
    // last update of each right-hand side is applied at the end of the cycle
    // (k_fcg_flush_x).
    const bool fused = c->net_fused && ha.cfg.lds && fused_shape(c, fty, ftx);
    auto iteration_fused = [&](int par) {
        float* pin = par ? p32b : p32;
        float* pout = par ? p32 : p32b;
        HopArgs<float> a = ha;
        a.in = z32;
        a.pold = pin;
        a.pnew = pout;
        a.out = t32;
        a.part = hpart;
        a.zpart = npart;
        a.nz = ntiles;
        a.fr = fr;
        a.fs = fs;
        a.stw = st;
        a.x = x32;
        a.pbuf = pout == p32b ? 1 : 0;
        launch_hop<float>(false, EPI_PNORM, a, nrhs, s);
        HopArgs<float> u = ha;
        u.in = t32;
        u.out = fr ? nullptr : q32;  // Fletcher-Reeves beta needs no q
        u.r = r32;
        u.partA = hpart;
        u.anum = &fs[0].rz;
        u.anum_stride = (int)(sizeof(FcgScal) / sizeof(double));
        u.aout = &fs[0].alpha;
        u.part = rpart;
        launch_hop<float>(true, EPI_RUPD, u, nrhs, s);
        const FusedCheck chk{hpart, rpart, hnb, st, fs};
        launch_net_fused(c, net, r32, 1, z32, 1, cfg, nrhs, st, fr ? nullptr : q32, npart, &chk);
    };
    // Otherwise (layer kernels or direct-load stencils): t = D p, the D^dag update, the check,
    // the network, a dot-product pass and the direction update, in separate kernels.
    auto iteration_layers = [&]() {
        HopArgs<float> a = ha;
        a.in = p32;
        a.out = t32;
        a.part = hpart;
        launch_hop<float>(false, EPI_NORM, a, nrhs, s);
        // q = D^dag t with x += alpha p, r -= alpha q and |r|^2 in the stencil's epilogue
        // (alpha = (z, r) / (p, A p))
        a.in = t32;
        a.out = fr ? nullptr : q32;  // Fletcher-Reeves beta needs no q
        a.x = x32;
        a.r = r32;
        a.p = p32;
        a.partA = hpart;
        a.anum = &fs[0].rz;
        a.anum_stride = (int)(sizeof(FcgScal) / sizeof(double));
        a.part = rpart;
        launch_hop<float>(true, EPI_CGUPD, a, nrhs, s);
        k_fcg_check<<<nrhs, 64, 0, s>>>(hpart, rpart, hnb, st, fs);
        net_forward(c, net, r32, 1, nullptr, z32, 1, cfg, nrhs, st, h1, h2);
        k_fcg_dots<float><<<vgrid, 256, 0, s>>>(z32, r32, q32, sites, st, part);
        k_fcg_pbeta<float><<<vgrid, 256, 0, s>>>(z32, p32, sites, st, fs, part, nb, fr);
    };
    constexpr int kBlockIters = 8;
    hipGraphExec_t graph = nullptr;
    if (c->use_graphs) {
        std::vector<unsigned long long> key{0xFC7ULL, (unsigned long long)nrhs};
        for (const void* ptr : {(const void*)r32, (const void*)p32, (const void*)p32b, (const void*)q32, (const void*)z32,
                                (const void*)t32, (const void*)x32, (const void*)h1, (const void*)h2,
                                (const void*)part, (const void*)hpart, (const void*)st, (const void*)fs,
                                (const void*)npart, (const void*)rpart,
                                (const void*)cfg, (const void*)net, (const void*)net->wf[0], (const void*)net->wf[1],
                                (const void*)net->wf[2], (const void*)net->wa[0], (const void*)net->wa[1],
                                (const void*)net->wa[2], (const void*)net->wg0, (const void*)c->l32})
            key.push_back((unsigned long long)(uintptr_t)ptr);
        key.push_back((unsigned long long)ha.rhs_per_cfg);
        key.push_back((unsigned long long)c->net_fused);
        key.push_back((unsigned long long)ntiles);
        key.push_back((unsigned long long)fr);
        const int v[] = {ha.cfg.bx, ha.cfg.chunks, ha.cfg.spt, ha.cfg.remap, ha.cfg.prefetch, ha.cfg.tx,
                         ha.cfg.tiles_x, ha.cfg.nt_store, ha.cfg.lds};
        for (int e : v) key.push_back((unsigned long long)(unsigned)e);
        key.push_back((unsigned long long)fused);
        graph = c->cached_graph(key, [&] {  // starts on an even iteration of the cycle
            for (int i = 0; i < kBlockIters; ++i) {
                if (fused)
                    iteration_fused(i & 1);
                else
                    iteration_layers();
            }
        });
    }
    auto max_iters = [&]() {
        int m = 0;
        for (const CgState& S : h) m = std::max(m, S.iters);
        return m;
    };
    auto running = [&]() {
        for (const CgState& S : h)
            if (!(S.flags[0] & kSkip)) return true;
        return false;
    };
    reliable_update(true);
    read();
    int cycles = 0;
    while (true) {
        bool all_done = true;
        for (const CgState& S : h)
            if (!(S.flags[0] & F_DONE)) all_done = false;
        if (all_done || max_iters() >= maxiter || cycles >= 1000) break;
        ++cycles;
        k_fcg_restart<float><<<vgrid, 256, 0, s>>>(r64, r32, x32, sites, st);
        if (fused) {
            // z = M r with its (z, r) partials; the first iteration's beta is 0 (p = z)
            launch_net_fused(c, net, r32, 1, z32, 1, cfg, nrhs, st, nullptr, npart);
            k_fcg_first<<<(nrhs + 63) / 64, 64, 0, s>>>(st, fs, nrhs);
        } else {
            net_forward(c, net, r32, 1, nullptr, z32, 1, cfg, nrhs, st, h1, h2);
            k_fcg_p0<float><<<vgrid, 256, 0, s>>>(z32, r32, p32, sites, st, part);
            k_fcg_rz<<<nrhs, 64, 0, s>>>(part, nb, st, fs);
        }
        c->check_launch("FCG cycle start");
        read();
        int it = 0;  // iterations of this cycle: the fused iteration's direction buffers go by parity
        auto iteration = [&]() {
            if (fused)
                iteration_fused(it & 1);
            else
                iteration_layers();
            ++it;
        };
        while (running() && max_iters() < maxiter) {
            const int n = std::min(32, maxiter - max_iters());
            int j = 0;
            if (graph && (it & 1) == 0)
                for (; j + kBlockIters <= n; j += kBlockIters, it += kBlockIters)
                    hchk(hipGraphLaunch(graph, s), "hipGraphLaunch");
            for (; j < n; ++j) iteration();
            c->check_launch("FCG iteration");
            read();
        }
        if (fused) k_fcg_flush_x<<<vgrid, 256, 0, s>>>(x32, p32, p32b, fs, sites);
        launch_accum_x(c->layout(nrhs, 0), x, x32, st, 0, F_DONE, nrhs, s);  // x += x32 (all not done)
        reliable_update(false);
        c->check_launch("FCG reliable update");
        read();
    }
    int rc = DWF_OK;
    for (int i = 0; i < nrhs; ++i) {
        dwf_solve_stats& S = stats[i];
        S = dwf_solve_stats{};
        S.iterations = h[i].iters;
        S.fp64_iterations = 0;
        S.reliable_updates = h[i].reliable;
        S.passes = 1;
        S.relres = h[i].relres;
        S.true_relres = h[i].true_relres;
        S.converged = (h[i].status == 0 && (h[i].flags[0] & F_DONE) && h[i].true_relres <= rtol) ? 1 : 0;
        if (!S.converged && rc == DWF_OK) rc = h[i].status ? DWF_ERR_ZERO_DENOM : DWF_ERR_NOT_CONVERGED;
    }
    if (rc == DWF_ERR_ZERO_DENOM)
        dwf_set_error("flexible CG broke down: the preconditioner is not positive on a residual");
    else if (rc == DWF_ERR_NOT_CONVERGED)
        dwf_set_error("CG did not converge within maxiter");
    return rc;
}
