/* dwf_hip.h -- C ABI of libdwfhip.so, the MI355X (gfx950) fermion hot path for 2D U(1)
 * Wilson and Shamir domain-wall HMC and measurements.
 *
 * The reference (wurdahl/Three_Part_DWF_Study) has no FFI: its hot path is a set of
 * header-only C++ functions over Eigen types.  Each entry point below replaces one of
 * them (cited file:line); include/dwf_hip.hpp wraps these entry points back into the
 * reference's own function names, result structs and exception messages.
 *
 * Conventions
 *   - Host buffers use the reference layout, interleaved complex double:
 *       DWF spinor     psi[s][t][x][spin]    ((s*Nt+t)*Nx+x)*2+spin  (domain_wall/include/indexing.hpp:5-8)
 *       Wilson spinor  psi[t][x][spin]       (t*Nx+x)*2+spin         (wilson/include/indexing.hpp:5-8)
 *       gauge angles   theta[mu][t][x]       (mu*Nt+t)*Nx+x, mu=0 is t (indexing.hpp:10-13)
 *     A batch of nrhs vectors is nrhs such vectors back to back.
 *   - *_device entry points take device pointers in the same layout (FP64: double,
 *     FP32: float) and run asynchronously on the context's stream.
 *   - The caller owns every buffer; the library never keeps a host pointer.  The
 *     context owns its device memory.
 *   - Every function returns DWF_OK (0) or an error code; no C++ exception crosses the
 *     ABI.  dwf_last_error() gives the message of the calling thread's last failure.
 *   - A context is used by one host thread at a time.  Contexts are independent, so
 *     each OpenMP thread of an analyzer holds its own (each has its own HIP stream).
 */
#ifndef DWF_HIP_H
#define DWF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum dwf_status {
    DWF_OK = 0,
    DWF_ERR_ARG = 1,             /* bad argument (sizes, op, null pointer)            */
    DWF_ERR_HIP = 2,             /* HIP runtime error or no GPU                       */
    DWF_ERR_NOT_CONVERGED = 3,   /* CG reached maxiter                                */
    DWF_ERR_ZERO_DENOM = 4,      /* CG encountered a zero denominator                 */
    DWF_ERR_CERTIFY = 5          /* FP64 true residual above tolerance after fallback */
};

enum dwf_op {
    DWF_OP_D = 0,      /* D          dwf_operator.hpp:15-92 / wilson_operator.hpp:10-62 */
    DWF_OP_DDAG = 1,   /* D^dagger   dwf_operator.hpp:94-171                             */
    DWF_OP_DDDAG = 2,  /* D D^dagger dwf_operator.hpp:173-176                            */
    DWF_OP_DDAGD = 3   /* D^dagger D dwf_operator.hpp:178-181                            */
};

enum dwf_precision {
    DWF_PREC_FP64 = 0,
    DWF_PREC_FP32 = 1
};

enum dwf_method {
    /* FP64 CG, the reference's algorithm (domain_wall/include/cg_solver.hpp:17-64):
     * stops on the recursive residual, as the reference does. */
    DWF_SOLVE_CG64 = 0,
    /* FP32 CG with FP64 reliable updates; stops on the FP64 true residual. */
    DWF_SOLVE_MIXED = 1,
    /* dwf_solve_dirac only: even-odd (4D checkerboard) Schur complement of the DWF
     * operator, FP32 CGNR with FP64 reliable updates on the odd half-lattice and FP64
     * defect correction; certified on the reference's full system.  Needs even Nt, Nx
     * and Ls in {4, 8, 12, 16}; otherwise it runs as DWF_SOLVE_MIXED. */
    DWF_SOLVE_EO_MIXED = 2,
    /* Domain-wall contexts only: CG preconditioned by the learned network attached with
     * dwf_ctx_set_preconditioner (DWF_TUNE_FCG_BETA picks flexible CG or
     * Fletcher-Reeves), FP32 cycles between FP64 reliable updates, certified on the FP64
     * true residual; plain FP64 CG continues if the network breaks down (stats:
     * fp64_iterations > 0).  D^dag D directly; D D^dag (the pseudofermion force and HMC
     * solves) as the same solve of Gamma b, Gamma = sigma3 R, since D D^dag =
     * Gamma D^dag D Gamma.  No counterpart in the reference. */
    DWF_SOLVE_LEARNED = 3
};

typedef struct dwf_ctx dwf_ctx;

/* Per right-hand side solve statistics. */
typedef struct dwf_solve_stats {
    int iterations;        /* CG iterations (FP32 ones for DWF_SOLVE_MIXED)            */
    int converged;         /* 1 if the stopping rule was met                           */
    int reliable_updates;  /* FP64 true-residual recomputations (mixed only)           */
    int fp64_iterations;   /* FP64 iterations (CG64, or the mixed solver's fallback)   */
    int passes;            /* even-odd defect-correction passes (0 for the other solvers) */
    int pad;
    double relres;         /* the solver's own ||r||/||b|| at exit (recursive for CG64) */
    double true_relres;    /* FP64 ||b - A x|| / ||b|| recomputed at exit              */
} dwf_solve_stats;

/* ---- context --------------------------------------------------------------------- */

/* Shamir domain-wall operator on an Nt x Nx x Ls lattice: diagonal 3-M5, wall coupling
 * +mf, antiperiodic in t and periodic in x (dwf_operator.hpp:15-92).  Replaces the
 * process-global sizes of domain_wall/include/parameters.hpp:7-12 with arguments. */
int dwf_ctx_create_dwf(dwf_ctx** out, int Nt, int Nx, int Ls, double M5, double mf);

/* Wilson operator: diagonal 2+m0, boundary signs bc_t at t=Nt-1 and bc_x at x=Nx-1
 * (wilson/include/wilson_operator.hpp:10-62, wilson/include/parameters.hpp:6-38). */
int dwf_ctx_create_wilson(dwf_ctx** out, int Nt, int Nx, double m0, double bc_t, double bc_x);

int dwf_ctx_destroy(dwf_ctx* ctx);

/* The context's HIP stream (hipStream_t), for callers that order their own work
 * against the *_device entry points. */
void* dwf_ctx_stream(dwf_ctx* ctx);

/* Launch-shape and solver knobs (no counterpart in the reference; set by the
 * autotuner or by hand).  DWF_TUNE_SCHUNKS: s-chunks per 4D site of the stencils
 * (0 = heuristic, 1..Ls); DWF_TUNE_XCD_REMAP: 0/1 XCD-aware block order;
 * DWF_TUNE_RELIABLE_DELTA: residual drop (0 < delta < 1) that triggers a reliable update
 * of the mixed solvers (unset, or set to 0: 1e-2 for DWF_SOLVE_MIXED and the learned
 * solver, 3e-4 for the even-odd Schur CG of DWF_SOLVE_EO_MIXED); DWF_TUNE_PREFETCH: 0/1 software-pipelined stencil slice loop;
 * DWF_TUNE_TILE_X: 0 for blocks of consecutive 4D sites, or the width (dividing Nx and
 * 256) of 2D stencil tiles of 256/width rows; DWF_TUNE_NT_STORE: 0/1 streaming
 * (non-temporal) stores of stencil outputs (the LDS-staged tiles always stream); DWF_TUNE_GRAPHS: 0/1 (default 1) replay
 * blocks of 8 CG iterations as instantiated HIP graphs instead of launching their
 * kernels one by one (same kernels, same order, same results); DWF_TUNE_LDS: 0, or
 * 1 / 2 / 3 / 4 / 5 / 6 to stage each slice of a 2D stencil tile of 256 / 512 / 64 /
 * 128 / 384 / 768 sites and its halo in LDS (with DWF_TUNE_TILE_X between 8 and 128 dividing
 * Nx and the tile size, sites/width dividing Nt, and 2 (width + sites/width) <= sites;
 * ignored otherwise); DWF_TUNE_NET_FUSED: the learned preconditioner's kernels, 0
 * layer by layer, 1 the three layers as one fused kernel (LDS-resident hidden layers;
 * bit-identical to 0), 2 (default) the fused kernel with each block computing in a
 * temporal gauge of its own (no transport along t; the same network up to the bf16
 * rounding of the gauge-transformed values); DWF_TUNE_FCG_BETA: the direction
 * update of DWF_SOLVE_LEARNED, 0 (default) flexible CG's beta = -(z, A p) / (p, A p)
 * (Notay's FCG(1), for any preconditioner), 1 Fletcher-Reeves beta = (z, r) / (z, r)_prev
 * (exact for a symmetric preconditioner; one vector pass less per iteration);
 * DWF_TUNE_EO_FUSED: the Schur operator of DWF_SOLVE_EO_MIXED, 0 two kernels per
 * application with A^-1 H psi through device memory, 1 one kernel (A^-1 H psi kept in
 * LDS; same arithmetic), 2 one kernel with the FP32 CG's direction update
 * p = r + beta p inside the next operator launch (one vector pass and one launch less per
 * iteration), 3 (default) as 2 with the CG's per-iteration state update also done inside
 * the next operator launch in the replayed blocks of 8 iterations (7 launches less per
 * block); DWF_TUNE_NET_TILE: the fused network's output tile,
 * ty * 256 + tx sites (ty <= 8, tx <= 16, dividing Nt and Nx), 0 (default) the largest
 * that divides; DWF_TUNE_STREAMS: 1 to 8 (default 2) concurrent sub-batches of a batched
 * one-lattice dwf_solve_dirac* of at least twice as many right-hand sides (contiguous
 * ranges, whole configurations when several are held), each on a stream, workspace and
 * host thread of its own: the same per-right-hand-side solves and certification, the
 * batch schedule (chunk lengths, reliable updates) per sub-batch; 1 solves the batch as one;
 * DWF_TUNE_NET_SPLIT: the fused network's parts of the fifth dimension per output tile,
 * one block each (n dividing Ls; the parts recompute their neighbours' edge slices, the
 * same values bit for bit), 0 (default) the most parts up to 4 of at least 4 slices each
 * that keep a launch within 256 blocks, one per CU (one right-hand side at 96 x 96: 2). */
enum dwf_tune_param {
    DWF_TUNE_SCHUNKS = 0,
    DWF_TUNE_XCD_REMAP = 1,
    DWF_TUNE_RELIABLE_DELTA = 2,
    DWF_TUNE_PREFETCH = 3,
    DWF_TUNE_TILE_X = 4,
    DWF_TUNE_NT_STORE = 5,
    DWF_TUNE_GRAPHS = 6,
    DWF_TUNE_LDS = 7,
    DWF_TUNE_NET_FUSED = 8,
    DWF_TUNE_FCG_BETA = 9,
    DWF_TUNE_EO_FUSED = 10,
    DWF_TUNE_NET_TILE = 11,
    DWF_TUNE_STREAMS = 12,
    DWF_TUNE_NET_SPLIT = 13
};
int dwf_ctx_tune(dwf_ctx* ctx, int param, double value);

/* Autotuner (no counterpart in the reference): times the D stencil over its launch
 * shapes (block mapping, software pipelining, s-chunks, streaming stores) for nrhs right-hand sides in
 * the given precision, applies the fastest to the context and reports it.  cache_path
 * (nullable) is a text file of earlier results: a cached key is reused without timing,
 * a new one is appended.  Needs the gauge field set. */
typedef struct dwf_tune_result {
    int tile_x, prefetch, chunks, nt_store;
    int from_cache;
    double microseconds;  /* one D apply of nrhs right-hand sides (kernel time: the applies
                             are timed as a HIP graph, without host launch overhead, unless
                             DWF_TUNE_GRAPHS is 0) */
    int lds;              /* 1..6: the LDS-staged tile kernel of 256 / 512 / 64 / 128 / 384 /
                             768 threads won (0: the direct-load kernel) */
    int pad;
} dwf_tune_result;
int dwf_autotune(dwf_ctx* ctx, int nrhs, int precision, const char* cache_path, dwf_tune_result* result);

/* Right-hand-side blocking (no counterpart in the reference, whose analyzer solves one
 * source at a time per OpenMP thread, domain_wall/analyze_configs.cpp:45-76): the batch
 * size of dwf_solve_dirac* with `method` that gives the most solves per second on this
 * context's lattice and (single) gauge field.  It times whole solves to rtol of nrhs
 * boundary Z4-noise sources (the analyzer's source shape) for nrhs = 8, 16, 32, ... up to
 * max_nrhs and what half the free device memory holds, stopping once a doubling gains
 * less than 5%, and reports the best nrhs and its solves/s.  Nothing is applied to the
 * context.  cache_path (nullable) works as for dwf_autotune. */
typedef struct dwf_batch_tune_result {
    int nrhs;
    int from_cache;
    double solves_per_s;
} dwf_batch_tune_result;
int dwf_autotune_batch(dwf_ctx* ctx, int method, double rtol, int max_nrhs, const char* cache_path,
                       dwf_batch_tune_result* result);

/* Upload gauge angles theta[2][Nt][Nx] (host).  The library keeps them on the device in
 * FP64 and as cos/sin link arrays in FP64 and FP32; U_mu(x) = exp(i theta_mu(x))
 * (dwf_operator.hpp:10-13). */
int dwf_set_gauge(dwf_ctx* ctx, const double* theta);
int dwf_set_gauge_device(dwf_ctx* ctx, const double* theta_device);

/* Hold ncfg configurations theta[ncfg][2][Nt][Nx] of the same lattice at once (no
 * counterpart in the reference, which measures one configuration per OpenMP thread,
 * domain_wall/analyze_configs.cpp:45-76).  Afterwards the operator, solver and
 * measurement entry points take nrhs = ncfg * k right-hand sides, and right-hand side r
 * belongs to configuration r / k: one batched launch serves many configurations.
 * The force, gauge and HMC entry points need a single configuration. */
int dwf_set_gauge_batch(dwf_ctx* ctx, int ncfg, const double* theta);

/* One lattice size of a multi-lattice batch: ncfg configurations of Nt x Nx. */
typedef struct dwf_lattice_group {
    int Nt, Nx, ncfg, pad;
} dwf_lattice_group;

/* Hold configurations of several lattice sizes at once (no counterpart in the reference,
 * whose lattice sizes are process-global constants, domain_wall/include/parameters.hpp:7-12,
 * and whose analyzer runs one configuration per OpenMP thread,
 * domain_wall/analyze_configs.cpp:45-76).  groups[0..ngroups) (ngroups <= 8) list the
 * sizes (the context's Ls, M5 and mf apply to all), theta is the concatenation over the
 * groups of ncfg x [2][Nt][Nx] angles.  Afterwards dwf_apply*, dwf_solve* (CG64, MIXED),
 * dwf_solve_dirac* (CG64, MIXED, EO_MIXED) and dwf_measure_batch take nrhs = k * (total
 * configurations) right-hand sides, k per configuration, configuration-major in group
 * order; each right-hand side's vector has its own lattice's size and layout and the
 * vectors follow one another (the batch is the concatenation).  Every kernel launch of
 * those entry points covers all the lattice sizes at once.  The learned preconditioner,
 * force, gauge, HMC and autotuning entry points need a single lattice; dwf_set_gauge and
 * dwf_set_gauge_batch return the context to its own lattice. */
int dwf_set_gauge_groups(dwf_ctx* ctx, int ngroups, const dwf_lattice_group* groups, const double* theta);

/* ---- operators --------------------------------------------------------------------- */

/* out = op(in) for nrhs vectors in FP64 (host buffers).  Replaces apply_D,
 * apply_D_dagger, apply_DD_dagger, apply_D_dagger_D (dwf_operator.hpp:15-181) and the
 * products with build_wilson_matrix (wilson_operator.hpp:10-62). */
int dwf_apply(dwf_ctx* ctx, int op, const double* in, double* out, int nrhs);

/* Same on device buffers in the given precision (in != out). */
int dwf_apply_device(dwf_ctx* ctx, int op, int precision, const void* in, void* out, int nrhs);

/* Allocation-free form of dwf_apply_device, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same kernels and results as
 * dwf_apply_device; the normal operators' intermediate vector lives in a fixed device
 * array of DWF_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of
 * the context's workspace, so the call makes no allocation.  Limit: for DWF_OP_DDDAG and
 * DWF_OP_DDAGD, nrhs * Ls * Nt * Nx * 4 * (8 for FP64, 4 for FP32) bytes must not exceed
 * DWF_NOALLOC_SCRATCH_BYTES (DWF_ERR_ARG otherwise); D and D^dagger need no scratch.
 * Every call on one GPU shares that GPU's array: calls that overlap in time on one GPU
 * (from several threads or streams) race. */
#define DWF_NOALLOC_SCRATCH_BYTES (64ULL << 20)
int dwf_apply_device_noalloc(dwf_ctx* ctx, int op, int precision, const void* in, void* out, int nrhs);

/* ---- solvers ----------------------------------------------------------------------- */

/* Solve A x = b for nrhs right-hand sides with A = D D^dagger (DWF_OP_DDDAG) or
 * D^dagger D (DWF_OP_DDAGD).  x0 (nullable) is the initial guess for every right-hand
 * side's own slot.  stats (nullable) receives nrhs records.  DWF_SOLVE_EO_MIXED solves
 * (D D^dagger)^-1 = Gamma D^-1 Gamma D^-1 (D^dagger = Gamma D Gamma, Gamma = sigma3 R)
 * with two even-odd Dirac solves per correction pass and certifies every pass on the FP64
 * residual of A itself (up to four passes, then the mixed CG continues and decides; stats
 * sum the Dirac solves' iterations, passes = correction passes; maxiter bounds each Dirac
 * solve and the continuing CG on its own, as it bounds the mixed CG's FP32 cycle and its
 * FP64 fallback on their own); on a Wilson context or a
 * lattice without an even-odd kernel it is the mixed CG.  DWF_SOLVE_EO_MIXED and
 * DWF_SOLVE_LEARNED need a context holding one lattice size.  Replaces
 * conjugate_gradient (domain_wall/include/cg_solver.hpp:17-64); the Wilson solves use
 * the same entry point (wilson/include/cg_solver.hpp:19-78).  Returns
 * DWF_ERR_NOT_CONVERGED if any right-hand side missed the tolerance (x still holds
 * the last iterate) and DWF_ERR_ZERO_DENOM as cg_solver.hpp:45-46. */
int dwf_solve(dwf_ctx* ctx, int normal_op, int method, const double* b, const double* x0,
              double rtol, int maxiter, int nrhs, double* x, dwf_solve_stats* stats);

/* Device-buffer form (FP64 b, x0 nullable, x).  Synchronises the context stream before
 * returning because convergence is decided on the host.  x may not overlap b, and x0 is
 * either x itself or disjoint from it (DWF_ERR_ARG otherwise); the same holds for src
 * in dwf_solve_dirac_device. */
int dwf_solve_device(dwf_ctx* ctx, int normal_op, int method, const double* b_device,
                     const double* x0_device, double rtol, int maxiter, int nrhs,
                     double* x_device, dwf_solve_stats* stats);

/* x = D^-1 src for nrhs sources, the propagator pattern of the analyzers: rhs =
 * D^dagger src, conjugate_gradient(D^dagger D, rhs) (correlator.hpp:77-79, 139-142,
 * 213-220; wilson/include/cg_solver.hpp:19-51).  CG64 stops on the recursive residual as
 * the reference; MIXED and EO_MIXED certify on the FP64 true residual of that same
 * system, ||D^dagger src - D^dagger D x|| / ||D^dagger src|| <= rtol.  stats[i].relres
 * is the solver's own residual, stats[i].true_relres the certified one. */
int dwf_solve_dirac(dwf_ctx* ctx, int method, const double* src, const double* x0, double rtol,
                    int maxiter, int nrhs, double* x, dwf_solve_stats* stats);
int dwf_solve_dirac_device(dwf_ctx* ctx, int method, const double* src_device,
                           const double* x0_device, double rtol, int maxiter, int nrhs,
                           double* x_device, dwf_solve_stats* stats);

/* ---- pseudofermion force, gauge action, HMC ---------------------------------------- */

/* eta = (D D^dagger)^-1 phi by the method's solve of dwf_solve (initial guess eta_guess nullable),
 * action = Re phi^dagger eta, force[2][Nt][Nx] = dS/dtheta with xi = D^dagger eta.
 * Replaces pseudofermion_action_force (domain_wall/include/pseudofermion.hpp:22-166)
 * and, on a Wilson context, pseudofermion_action_force + solve_DDdag
 * (wilson/include/gauge.hpp:68-122, wilson/include/cg_solver.hpp:53-78).  eta and stats
 * may be NULL.  A DWF solve that misses rtol returns DWF_ERR_NOT_CONVERGED with the
 * message "Pseudofermion CG failed to converge"; a Wilson one fails, as the reference
 * does, only when its true residual exceeds 20 x rtol. */
int dwf_pf_action_force(dwf_ctx* ctx, int method, const double* phi, const double* eta_guess,
                        double rtol, int maxiter, double* action, double* force, double* eta,
                        dwf_solve_stats* stats);

/* The force pass alone for a given eta (pseudofermion.hpp:109-157). */
int dwf_force_from_eta(dwf_ctx* ctx, const double* eta, double* force);

/* Wilson plaquette gauge action beta sum(1 - cos P) and its force dS/dtheta
 * (domain_wall/include/gauge.hpp:26-62, wilson/include/gauge.hpp:38-62).  force may be
 * NULL. */
int dwf_gauge_action_force(dwf_ctx* ctx, double beta, double* action, double* force);

/* average_plaquette (domain_wall/include/gauge.hpp:64-73). */
int dwf_average_plaquette(dwf_ctx* ctx, double* plaquette);

/* Download the context's current gauge angles theta[2][Nt][Nx]. */
int dwf_get_gauge(dwf_ctx* ctx, double* theta);

enum dwf_integrator {
    /* domain_wall/generate_configs.cpp:52-169: no angle wrap, chronological eta guess */
    DWF_HMC_DOMAIN_WALL = 0,
    /* wilson/include/gauge.hpp:132-180: wrap_angle after each drift, zero initial guess */
    DWF_HMC_WILSON = 1
};

typedef struct dwf_hmc_params {
    double beta;      /* gauge coupling                                    */
    double epsilon;   /* leapfrog step                                     */
    int steps;        /* leapfrog steps                                    */
    int integrator;   /* enum dwf_integrator                               */
    int method;       /* force solver: DWF_SOLVE_CG64, _MIXED, _EO_MIXED or _LEARNED */
    int maxiter;      /* force CG iteration limit                          */
    double rtol;      /* force CG tolerance                                */
} dwf_hmc_params;

typedef struct dwf_hmc_result {
    int accepted;
    int solves;              /* pseudofermion solves in the trajectory            */
    int cg_iterations;       /* their CG iterations (FP32 ones for the mixed solver) */
    int pad;
    double acceptance_probability;
    double delta_h;
    double h_initial;
    double h_final;
    double plaquette;        /* average plaquette of the configuration after the accept step */
} dwf_hmc_result;

/* One HMC trajectory from the context's current gauge field, entirely on the device in
 * FP64 (angles, momenta, forces, actions, Delta H).  The caller draws the random
 * numbers on the host in the reference's order and passes them in: chi (the
 * pseudofermion heat-bath vector, phi = D chi), the momenta, and u, the uniform draw
 * of the accept step (accepted = u < min(1, exp(-dH))).  On acceptance the context's
 * gauge field becomes the new configuration, otherwise it is left unchanged.
 * Replaces the trajectory body of domain_wall/generate_configs.cpp:52-169 and
 * hmc_step (wilson/include/gauge.hpp:132-180). */
int dwf_hmc_trajectory(dwf_ctx* ctx, const dwf_hmc_params* params, const double* chi,
                       const double* momentum, double u, dwf_hmc_result* result);

/* ---- analyzer measurements ---------------------------------------------------------- */

/* Which of the DWF analyzer's per-configuration measurements to run.  All requested
 * sources go through one batched CG on D^dagger D (rhs D^dagger src). */
typedef struct dwf_measure_spec {
    int method;                    /* DWF_SOLVE_CG64 (reference), _MIXED, _EO_MIXED, or
                                      _LEARNED (a network attached, one lattice size)     */
    int maxiter;                   /* dwf.propagator_maxiter                             */
    double rtol;                   /* dwf.propagator_rtol                                */
    int wall;                      /* 1: compute_CPP_CJ5, 4 wall-source solves          */
    int point;                     /* 1: compute_momentum_CPP, 4 point-source solves     */
    int max_momentum;              /* analysis.max_momentum (0..Nx/2)                    */
    int noise_vectors;             /* eta.noise_vectors (0 skips the loop density)       */
    unsigned long long noise_seed; /* eta.random_seed + 104729 * config                  */
    const int* point_site;         /* nullable: the reference's 4 point sources at
                                      (t0 in {0, Nt/2}, x=0).  Otherwise [ncfg][2] = (t0, x0):
                                      2 point sources (one per spin) at that site per
                                      configuration; correlators are relative to it.  No
                                      reference counterpart (the batched-analyzer workload) */
} dwf_measure_spec;

typedef struct dwf_measure_out {
    double* CPP;             /* [Nt]  C_PP(dt)   (correlator.hpp:243-310), if wall       */
    double* CJ5;             /* [Nt]  C_J5(dt)                                          */
    double* momentum_CPP;    /* [max_momentum+1][Nt] (correlator.hpp:111-181), if point   */
    double* loop;            /* [Nt][Nx] complex L(t,x) (correlator.hpp:27-109), if noise */
    int* iterations;         /* nullable: [4 wall][4 or 2 point][noise] CG iterations      */
    double* dirac_residual;  /* nullable: [4] ||D x - src|| / ||src|| of the wall solves  */
    int wall_nonconverged;   /* wall solves that missed rtol ("CGNR info = nonzero")      */
    int pad;
    double* point_CJ5;       /* nullable: [Nt] C_J5(dt) of the point sources (midpoint
                                s = Ls/2 as for the wall sources), averaged like
                                momentum_CPP; no reference counterpart                  */
} dwf_measure_out;

/* One configuration's measurements (the loop body of domain_wall/analyze_configs.cpp:57-75:
 * compute_CPP_CJ5, estimate_pseudoscalar_loop_density, compute_momentum_CPP).  Errors
 * carry the reference's messages: "Momentum propagator CGNR failed to converge",
 * "Eta loop CGNR failed to converge", "analysis.max_momentum is outside 0..Nx/2".  Wall
 * solves that miss rtol are counted, not fatal, as in the reference. */
int dwf_measure(dwf_ctx* ctx, const dwf_measure_spec* spec, dwf_measure_out* out);

/* The same for the ncfg configurations a context holds (dwf_set_gauge_batch): all
 * ncfg x (8 + noise) sources go through one batched solve.  seeds[k] (nullable: every
 * configuration uses spec->noise_seed) is configuration k's Z4 seed, outs[k] its outputs.
 * Replaces the OpenMP loop over configurations of domain_wall/analyze_configs.cpp:45-76. */
int dwf_measure_batch(dwf_ctx* ctx, const dwf_measure_spec* spec, int ncfg,
                      const unsigned long long* seeds, dwf_measure_out* outs);

/* ---- learned preconditioner (no counterpart in the reference) -------------------- */

/* A gauge-equivariant linear convolutional network M ~ (D^dag D)^-1: three layers of
 * parallel-transported 7-point stencils in (t, x, s) with s-dependent complex weights,
 * 2 -> 8 -> 8 -> 2 complex channels, run in BF16 on the matrix cores with FP32
 * accumulation.  Weights (host arrays of float): per layer [Ls][7][C_out][C_in]
 * (re, im), layers concatenated; points self, +t, -t, +x, -x, +s, -s.  The network
 * lives on the current device and can serve any context with the same Ls and Nt*Nx a
 * multiple of 16. */
typedef struct dwf_net dwf_net;
long long dwf_net_num_weights(int Ls);
int dwf_net_create(dwf_net** out, int Ls, unsigned long long seed);
int dwf_net_destroy(dwf_net* net);
int dwf_net_set_weights(dwf_net* net, const float* w);
int dwf_net_get_weights(const dwf_net* net, float* w);

/* out = M in (adjoint = 0) or M^dag in (adjoint = 1) for nrhs FP64 spinors on the
 * device, with the context's configurations. */
int dwf_net_apply_device(dwf_ctx* ctx, const dwf_net* net, int adjoint, const double* in, double* out, int nrhs);

/* Training pairs: b[npairs] (sources of D^dag D x = b) and x[npairs] (their solutions),
 * FP32 spinors on the device; pair i uses configuration pair_cfg[i] (host array) of the
 * configurations the context holds.  loss 0: mean ||M b - x||^2 / ||x||^2;
 * loss 1: mean ||D^dag D M b - b||^2 / ||b||^2 (x unused, may be null). */
typedef struct dwf_train_params {
    int steps;        /* Adam steps                                   */
    int batch;        /* pairs per step (reshuffled every epoch)      */
    int loss;         /* 0 or 1, as above                              */
    int pad;
    double lr, beta1, beta2, eps;
    unsigned long long seed;  /* batch order                          */
} dwf_train_params;

/* Loss over all npairs and its gradient (host float array in the weight layout). */
int dwf_net_loss_grad_device(dwf_ctx* ctx, dwf_net* net, int loss, const float* b, const float* x,
                             const int* pair_cfg, int npairs, double* loss_out, float* grad);

/* p->steps Adam steps; loss_history (nullable) receives each step's batch loss. */
int dwf_net_train_device(dwf_ctx* ctx, dwf_net* net, const float* b, const float* x, const int* pair_cfg,
                         int npairs, const dwf_train_params* p, double* loss_history);

/* The network DWF_SOLVE_LEARNED uses (null detaches).  The context keeps the pointer:
 * destroy the network only after detaching it or destroying the context. */
int dwf_ctx_set_preconditioner(dwf_ctx* ctx, const dwf_net* net);

/* ---- host random numbers of the callers (reference draw order) --------------------- */

/* A std::mt19937_64 engine with the reference's normal(0,1) and uniform[0,1)
 * distributions (libstdc++), for the HMC heat-bath and accept draws. */
typedef struct dwf_rng dwf_rng;
dwf_rng* dwf_rng_create(unsigned long long seed);
void dwf_rng_destroy(dwf_rng* rng);

/* chi[ndof] complex = (N + iN)/sqrt(2), then momentum[ngauge] = N
 * (domain_wall/generate_configs.cpp:56-69). */
void dwf_rng_dwf_heatbath(dwf_rng* rng, int ndof, int ngauge, double* chi, double* momentum);

/* momentum[ngauge] = N first, then chi[ntot] complex (wilson/include/gauge.hpp:134-141). */
void dwf_rng_wilson_heatbath(dwf_rng* rng, int ngauge, int ntot, double* momentum, double* chi);

/* The accept draw uniform(rng) (generate_configs.cpp:162-163, gauge.hpp:178). */
double dwf_rng_uniform(dwf_rng* rng);

/* Z4 noise {1, i, -1, -i} of estimate_pseudoscalar_loop_density
 * (domain_wall/include/correlator.hpp:35-58): nnoise vectors of sites_spin = Nt*Nx*2
 * complex entries from one mt19937_64(seed). */
void dwf_z4_noise(unsigned long long seed, int nnoise, int sites_spin, double* eta);

/* ---- device memory and timing helpers -----------------------------------------------
 * No counterpart in the reference (it has no device).  Plumbing for callers that keep
 * their fields in device memory and time the *_device entry points, so they need no
 * second GPU runtime in their process. */
void* dwf_device_alloc(unsigned long long bytes);              /* NULL on failure */
int dwf_device_free(void* ptr);
int dwf_memcpy_htod(void* dst_device, const void* src_host, unsigned long long bytes);
int dwf_memcpy_dtoh(void* dst_host, const void* src_device, unsigned long long bytes);
int dwf_ctx_synchronize(dwf_ctx* ctx);

/* HIP events recorded on a context's stream. */
typedef struct dwf_timer dwf_timer;
dwf_timer* dwf_timer_create(void);
void dwf_timer_destroy(dwf_timer* timer);
int dwf_timer_start(dwf_timer* timer, dwf_ctx* ctx);
int dwf_timer_stop(dwf_timer* timer, dwf_ctx* ctx);
double dwf_timer_elapsed_ms(dwf_timer* timer);  /* waits for the stop event; < 0 on error */

/* dst = src (bytes a multiple of 16) with a plain vector copy kernel on the context's
 * stream: the achievable-bandwidth reference for the stencils' roofline. */
int dwf_copy_device(dwf_ctx* ctx, void* dst, const void* src, unsigned long long bytes);

/* dst[i] = (float)src[i] for n doubles on the device, on the context's stream (training
 * pairs for the learned preconditioner are stored in FP32). */
int dwf_convert_f64_f32_device(dwf_ctx* ctx, const double* src, float* dst, unsigned long long n);

/* Message of the calling thread's last failure ("" if none). */
const char* dwf_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DWF_HIP_H */
