// dwf_context.hpp -- the opaque dwf_ctx of the C ABI (host side).
#pragma once

#include "dwf_internal.hpp"
#include "../../include/dwf_hip.h"

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

// Error carrying a C ABI status code; caught at the ABI boundary.
struct DwfError : std::runtime_error {
    int code;
    DwfError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void dwf_set_error(const std::string& msg);

// Workspace slots (grow-only device buffers).
enum BufId : int {
    B_HOST_IN = 0,
    B_HOST_OUT,
    B_APPLY_TMP,
    B_R64,
    B_T64,
    B_P64,
    B_R32,
    B_T32,
    B_P32,
    B_P32B,      // pipelined even-odd CG: the second direction buffer
    B_X32,
    B_PA,
    B_PA2,       // pipelined even-odd CG: the operator's partials of odd iterations
    B_PB,
    B_PR,
    B_STATE,
    B_AUX0,
    B_AUX1,
    B_AUX2,
    B_AUX3,
    B_AUX4,
    B_PHI,     // HMC: pseudofermion field
    B_ETA,     // HMC: (D D^dag)^-1 phi (also the chronological guess)
    B_XI,      // HMC: D^dag eta
    B_MOM,     // HMC: momenta
    B_FG,      // HMC: gauge force
    B_FPF,     // HMC: pseudofermion force
    B_THETA0,  // HMC: configuration at the start of the trajectory
    B_SCAL,    // FP64 scalar results
    B_RPART,   // reduction partials
    B_TQ64,    // CG64 direction image
    B_STATE2,  // certification states
    B_PR2,     // certification partials
    B_EO_B,    // D^dag src
    B_EO_RSRC, // Dirac residual of the current iterate
    B_EO_BE,
    B_EO_BO,
    B_EO_UE,
    B_EO_BH,
    B_EO_BHAT,
    B_EO_XO,
    B_EO_XE,
    B_EO_Q,
    B_EO_TE32,
    B_EO_DELTA,
    B_NET_H1,
    B_NET_H2,
    B_NET_Y,
    B_NET_GY,
    B_NET_T,
    B_NET_ROW,
    B_NET_CFG,
    B_NET_GB,
    B_NET_WPART,
    B_NET_PART,
    B_NET_N2,
    B_NET_LOSS,
    B_NET_STAGE,
    B_NET_GRAD,
    B_FCG_R,
    B_FCG_P,
    B_FCG_Q,
    B_FCG_Z,
    B_FCG_PART,
    B_FCG_SCAL,
    B_FCG_R32,
    B_FCG_P32,
    B_FCG_Q32,
    B_FCG_Z32,
    B_FCG_T32,
    B_FCG_X32,
    B_FCG_NPART,
    B_GAMMA_B,   // learned D D^dag solves: Gamma b
    B_FCG_P32B,  // learned solver: the second direction buffer of the fused iteration
    B_NOFF,      // analyzer: offsets of each configuration's noise vectors
    B_HOLD,      // mixed CG: right-hand sides a reliable update skips
    B_NEO_R,     // even-odd normal solves: residual of the normal system, then the second solution
    B_NEO_Y,     // even-odd normal solves: the first Dirac solution
    B_NEO_PART,  // even-odd normal solves: certification partials
    B_NEO_STATE, // even-odd normal solves: certification states
    B_COUNT
};

struct dwf_ctx {
    dwfk::Lat L{};
    dwfk::Lat L0{};  // the lattice the context was created with
    int kind = 0;  // 0 DWF, 1 Wilson
    int device = 0;
    hipStream_t stream = nullptr;
    double* theta = nullptr;  // [2][V4] FP64 angles on the device
    double2* l64 = nullptr;   // [2][V4] (cos, sin) FP64
    float2* l32 = nullptr;    // [2][V4] (cos, sin) FP32
    bool have_gauge = false;
    int ncfg = 1;                 // configurations held (all lattice groups together)
    int cfg_cap = 1;              // configurations the link buffers can hold
    int rhs_per_cfg = 1 << 30;    // set per call: right-hand side r uses configuration r / rhs_per_cfg
    void bind_rhs(int nrhs);      // checks nrhs against ncfg and sets rhs_per_cfg (and gr)
    // Lattice groups (dwf_set_gauge_groups): per group (Nt, Nx, configurations); empty for
    // the context's own lattice L.  With two or more groups, L holds the largest lattice
    // (Ls, masses and signs are shared) and every supported entry point runs all groups in
    // each launch.
    struct GroupDef {
        int Nt, Nx, ncfg;
    };
    std::vector<GroupDef> gdefs;
    size_t link_cap = 0;          // complex link entries the link buffers can hold
    dwfk::Groups gr{};            // the current call's layout (set by bind_rhs)
    bool multi() const { return gdefs.size() > 1; }
    // Device copy of gr for multi-lattice launches (uploaded when the layout changes; the
    // kernels of one call all read the call's layout, and cached graphs are keyed by it).
    const dwfk::Groups* groups_device();
    dwfk::Groups* gr_dev = nullptr;
    dwfk::Groups gr_dev_copy{};
    bool gr_dev_valid = false;
    // gr for a multi-lattice launcher (its grid), null for one lattice
    const dwfk::Groups* groups_host() const { return multi() ? &gr : nullptr; }
    void require_single_lattice(const char* what) const;
    // The layout of nrhs right-hand sides (after bind_rhs(nrhs)): gr for several groups,
    // else one group of L.
    dwfk::Groups groups_for(int nrhs) const;
    dwfk::VecLayout layout(int nrhs, int shift) const;
    long long vec_sites(int nrhs) const { return layout(nrhs, 0).total_sites(); }  // full-lattice 5D sites
    void require_single_config(const char* what) const;
    void reserve_configs(int n);  // grow the gauge buffers to n configurations of L
    void reserve_links(size_t nl);  // grow the gauge buffers to nl complex links
    int chunks_override = 0;      // s-chunks per 4D site for the stencils (0: heuristic)
    int xcd_remap = 1;            // XCD-aware block order for the stencils
    int prefetch = 0;             // software-pipelined stencil slice loop
    int tile_x = 0;               // stencil 2D tile width (0: linear blocks)
    int nt_store = 0;             // streaming stores of stencil outputs
    int lds = 0;                  // LDS-staged stencil tiles
    int net_fused = 2;            // learned preconditioner: 0 layer by layer, 1 fused, 2 fused in a gauge per block
    int fcg_beta = 0;             // learned solver: 0 flexible CG (Notay's FCG(1)), 1 Fletcher-Reeves
    int eo_fused = 1;             // Schur operator: 1 one kernel per application, 0 two
    int eo_pipe = 2;              // fused Schur CG: 1 direction update inside the operator launch,
                                  // 2 also the state update inside the next one (graph blocks)
    int net_tile = 0;             // fused network's output tile ty * 256 + tx (0: the largest that divides)
    int net_split = 0;            // fused network's parts of the fifth dimension per tile (0: by block count)
    double reliable_delta = 0.01; // mixed CG: reliable update after a 100x residual drop
    bool delta_set = false;       // reliable_delta set by hand (else each operator's default)
    int use_graphs = 1;           // replay blocks of CG iterations as HIP graphs
    const dwf_net* net = nullptr; // learned preconditioner of DWF_SOLVE_LEARNED (not owned)
    // Concurrent streams of a batched Dirac solve (DWF_TUNE_STREAMS): a one-lattice solve of
    // at least 2 * streams right-hand sides runs as `streams` sub-batches, each on a sub-context
    // of its own (stream, workspace, graphs; this context's gauge buffers borrowed), from as
    // many host threads, so one sub-batch's kernels fill the others' launch tails and host
    // round trips.
    int streams = 2;
    std::vector<std::unique_ptr<dwf_ctx>> kids;
    bool borrowed = false;        // a sub-context: theta, l64, l32 belong to its parent
    hipEvent_t ev = nullptr;      // orders a sub-context's stream after its parent's, and back
    // Sub-context i, created on first use, with this context's lattice, settings and
    // configurations cfg0 .. cfg0 + ncfg - 1 (borrowed link buffers).
    dwf_ctx* kid(int i, int cfg0, int kid_ncfg);

    // Instantiated CG iteration blocks, keyed by every argument their kernels capture.
    std::map<std::vector<unsigned long long>, hipGraphExec_t> graphs;
    void drop_graphs();
    // The instantiated graph of `body` (kernel launches on `stream` only), cached under
    // `key`, which must hold every argument the captured kernels receive.
    hipGraphExec_t cached_graph(const std::vector<unsigned long long>& key, const std::function<void()>& body);
    // The instantiated graph of `body`, not cached: the caller destroys it.
    hipGraphExec_t capture_graph(const std::function<void()>& body);

    struct Buf {
        void* p = nullptr;
        size_t n = 0;
    } bufs[B_COUNT];

    dwfk::CgState* st_host = nullptr;  // pinned mirror of the per-RHS solver states
    int st_host_n = 0;
    int last_slot = 0;

    ~dwf_ctx();
    void* buf(int id, size_t bytes);
    dwfk::CgState* host_states(int nrhs);
    static void check(hipError_t e, const char* what);
    static void check_launch(const char* what);
};

void dwf_apply_device_impl(dwf_ctx* c, int op, int prec, const void* in, void* out, int nrhs);
// The context's stencil launch shape for nrhs right-hand sides (tuned knobs; linear blocks
// without LDS tiles when it holds several lattice groups).
dwfk::HopCfg ctx_hop_cfg(const dwf_ctx* c, int nrhs);
// Stencil arguments with the context's links and the call's lattice groups.
template <typename R> dwfk::HopArgs<R> ctx_hop_args(const dwf_ctx* c, const dwfk::HopCfg& cfg, int nrhs);
// Scalar slots in the B_SCAL buffer.
enum ScalarSlot : int { S_GAUGE = 0, S_PLAQ, S_PF, S_KIN, S_COUNT = 8 };

int dwf_solve_dirac_impl(dwf_ctx* c, int method, const double* src, const double* x0, double rtol, int maxiter,
                         int nrhs, double* x, dwf_solve_stats* stats);
int dwf_fcg_solve(dwf_ctx* c, const double* b, double* x, bool have_guess, double rtol, int maxiter, int nrhs,
                  dwf_solve_stats* stats);
// Learned-preconditioner solve of A x = b, A = D^dag D or D D^dag (device buffers), with
// the plain FP64 CG fallback; D D^dag = Gamma D^dag D Gamma (Gamma = sigma3 R).
int dwf_solve_learned_impl(dwf_ctx* c, int normal, const double* b, const double* x0, double rtol, int maxiter,
                           int nrhs, double* x, dwf_solve_stats* stats);
// dwf_solve_device without the error guard: every method, with its certified fallbacks.
int dwf_solve_normal_impl(dwf_ctx* c, int normal, int method, const double* b, const double* x0, double rtol,
                          int maxiter, int nrhs, double* x, dwf_solve_stats* stats);
int dwf_solve_device_impl(dwf_ctx* c, int normal, int method, const double* b, const double* x0,
                          double rtol, int maxiter, int nrhs, double* x);
