#!/usr/bin/env python3
"""Benchmark of the DWF fermion hot path on one MI355X.

Workload (one "step"): solve D^dag D x = D^dag b for a batch of 32 boundary sources on a
96x96 lattice with Ls=16, M5=1.0, m_f=0.05 and links theta ~ N(0, 1/beta), beta=4 (seed
1234), to an FP64 true residual ||D^dag b - D^dag D x|| / ||D^dag b|| <= 1e-10.  Sources
are complex Gaussian eta(t,x,spin) placed as P_L eta at s=0 and P_R eta at s=Ls-1, as
domain_wall/include/correlator.hpp:60-75 places its noise.

``value`` is the solves/s of the learned path: preconditioned CG in FP32 with FP64
reliable updates whose preconditioner is the gauge-equivariant network, certified by the
FP64 true residual.  The network is trained first, outside the timed region, on the device
(--learn-configs seeded configurations held out from the test one, seeds 1234 + 104729 i
for i = 1 .. n, with --learn-pairs full-volume Gaussian sources each solved to 1e-10).  The
plain FP64 CG (the reference's algorithm) is timed on the same batch beside it, and so are
the even-odd Schur-complement mixed solver and the full-lattice mixed solver.  The
roofline object covers the dominant stencil, the FP32 Dirac operator, measured with HIP
events on the library's stream over a batch of 128 right-hand sides (604 MB, larger than
the 256 MB MALL).  The CPU baseline is the oracle's FP64 CG (the reference's algorithm and
loop order) on one host core, timed over --cpu-solves whole solves.

With --workload analyzer a step is instead the batched analyzer pass: 128 seeded
configurations, 32 each at Nx = Nt in {32, 56, 80, 96} with Ls=16, each with 2 point
sources (one per spin) at a seeded site and 32 Z4 noise vectors -- 4,352 propagator
solves (FP32 CG + FP64 reliable updates on the even-odd Schur complement, FP64 true
residual <= 1e-10) and their FP64 contractions (point C_PP, C_J5, loop density).  All
configurations of all four sizes sit in one context (dwf_set_gauge_groups), so the 4,352
right-hand sides are one batched solve and every kernel launch covers the four sizes
(--analyzer-split: one context and stream per size instead, for comparison).  Its CPU
baseline solves the 34 sources of one configuration per size with the oracle, spread
over every OpenMP thread.

With --workload learned: only the learned path and its comparisons (FP64 CG, even-odd
mixed, full-lattice mixed) on held-out configurations, with the training statistics and
the FP32 D apply's bandwidth.  --lattice reference-default runs it on the reference's own
default lattice (parameters.txt:25-30: 32x50, N5 = 12, M5 = 1, mf = 0.2, beta = 3), where
no LDS tile of the D stencil fits (a tile's width divides Nx and the block size), so the
stencils take the direct-load kernel and the learned iteration its unfused form.

With --workload hmc a step is one Nf=2 domain-wall HMC trajectory at 96x96, Ls=16,
beta=4, mf=0.05, M5=1, tau=1 with 20 leapfrog steps (epsilon=0.05), from a hot start
theta ~ U[-pi, pi) (seed 1234), the host random stream seeded 1234, force solves to 1e-10
with chronological initial guesses; the gauge update, forces and Delta H stay in FP64 on
the device.  --hmc-solver picks the force solve: cg64, mixed (default), eo_mixed (two
even-odd Dirac solves per solve, certified on D D^dag) or learned.

With --workload eo: the even-odd solver alone on the default batch (for kernel traces and
counter runs of its kernels), with its fraction of the Schur CG's HBM ceiling.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload solves|analyzer|learned|hmc|eo]
                      [--lattice bench|reference-default]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DWF solves/s 96²,Ls=16 @1e-10: learned+FP32 vs plain FP64 CG; Dslash HBM GB/s"
HBM_PEAK_GBS = 8000.0
NT = NX = 96
LS = 16
M5, MF, BETA = 1.0, 0.05, 4.0
TOL = 1e-10


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def boundary_sources(nrhs, seed):
    rng = np.random.default_rng(seed)
    eta = rng.normal(size=(nrhs, NT, NX, 2)) + 1j * rng.normal(size=(nrhs, NT, NX, 2))
    src = np.zeros((nrhs, LS, NT, NX, 2), np.complex128)
    src[:, 0, :, :, 1] = eta[..., 1]        # P_L = diag(0, 1) at s = 0
    src[:, LS - 1, :, :, 0] = eta[..., 0]   # P_R = diag(1, 0) at s = Ls-1
    return src


def dslash_bytes(prec_bytes, nrhs):
    """Algorithmic bytes of one D apply: read psi, write out, read cos/sin links."""
    V4 = NT * NX
    Bv = V4 * LS * 4 * prec_bytes
    BL = V4 * 4 * prec_bytes
    return nrhs * 2 * Bv + BL


def cg_iteration_bytes(prec_bytes):
    """Algorithmic bytes of one CG iteration on D^dag D for one right-hand side: two
    stencils (2 B_v + B_L each) and nine vector passes (x, r update with |r|^2: 4 reads,
    2 writes; p update: 2 reads, 1 write)."""
    V4 = NT * NX
    Bv = V4 * LS * 4 * prec_bytes
    BL = V4 * 4 * prec_bytes
    return 2 * (2 * Bv + BL) + 9 * Bv


def eo_iteration_bytes():
    """Algorithmic bytes of one FP32 iteration of the even-odd Schur CG on Dhat^dag Dhat
    for one right-hand side, vectors on the odd half lattice (B_h = B_v / 2): the Dhat
    launch reads r and the old direction and writes the new direction and t = Dhat p
    (4 B_h), the Dhat^dag launch reads t, p, x, r and writes x, r (6 B_h); A^-1 H p stays
    on the chip; each launch reads the links (B_L)."""
    V4 = NT * NX
    Bh = V4 * LS * 4 * 4 // 2
    BL = V4 * 4 * 4
    return 10 * Bh + 2 * BL


def time_kernel(ctx, capi, launch, reps):
    """Average duration (ms) of `launch` over `reps` back-to-back launches, from HIP
    events recorded on the library's stream around the timed region."""
    for _ in range(3):
        launch()
    t = capi.Timer()
    t.start(ctx)
    for _ in range(reps):
        launch()
    t.stop(ctx)
    return t.elapsed_ms() / reps


def apply_shape(ctx, dwf, shape):
    """Sets the context's stencil launch shape to an autotuner result."""
    ctx.tune(dwf.DWF_TUNE_TILE_X, shape["tile_x"])
    ctx.tune(dwf.DWF_TUNE_PREFETCH, shape["prefetch"])
    ctx.tune(dwf.DWF_TUNE_SCHUNKS, shape["chunks"])
    ctx.tune(dwf.DWF_TUNE_NT_STORE, shape["nt_store"])
    ctx.tune(dwf.DWF_TUNE_LDS, shape["lds"])


# Threads per block of each LDS tile mode of the stencil (DWF_TUNE_LDS 1..6).
LDS_THREADS = {1: 256, 2: 512, 3: 64, 4: 128, 5: 384, 6: 768}

# LDS tile shapes (mode, tile width) that cover a 96 x 96 lattice, tried on whole solves.
SOLVE_SHAPES = ((3, 8), (3, 16), (4, 16), (4, 32), (1, 16), (1, 32), (2, 32), (5, 96), (6, 96))
TUNE_SOURCE_SEED = 5550123  # sources of the shape-picking solves (not the timed ones)


def pick_solve_shape(ctx, dwf, solve_once, tuned, log):
    """The stencil launch shape for one solver, chosen on whole solves.  The autotuner
    times the plain D apply, but the solvers' stencils carry epilogues (norms, the CG
    update, the direction update) that move other vectors, and the fastest plain shape
    is not always the fastest solve.  Candidates: the autotuned shape and every LDS tile
    shape of SOLVE_SHAPES with one and two s-chunks.  Each runs a warm-up solve (which
    captures its iteration graphs) and a timed one; the three fastest are timed twice
    more and the best time wins."""
    keys = ("tile_x", "prefetch", "chunks", "nt_store", "lds")
    cands = [{k: tuned[k] for k in keys}]
    cands += [dict(tile_x=tx, prefetch=0, chunks=ch, nt_store=1, lds=m) for m, tx in SOLVE_SHAPES for ch in (1, 2)]

    def timed(shape):
        apply_shape(ctx, dwf, shape)
        ctx.synchronize()
        t0 = time.perf_counter()
        solve_once()
        ctx.synchronize()
        return time.perf_counter() - t0

    first = []
    for c in cands:
        apply_shape(ctx, dwf, c)
        solve_once()
        first.append(timed(c))
    top = sorted(range(len(cands)), key=first.__getitem__)[:3]
    best = {i: min(first[i], timed(cands[i]), timed(cands[i])) for i in top}
    i = min(best, key=best.__getitem__)
    log(f"solve shape {cands[i]}: {best[i] * 1e3:.2f} ms per solve (autotuned D-apply shape "
        f"{cands[0]}: {first[0] * 1e3:.2f} ms)")
    apply_shape(ctx, dwf, cands[i])
    return cands[i]


def solve_shape(args, ctx, dwf, solve_once, tuned, log):
    """--solve-shape: 'auto' picks on whole solves, 'tuned' keeps the autotuner's D-apply
    shape, 'lds,tile_x' fixes an LDS tile shape (FP32 solvers; FP64 CG keeps its own)."""
    if args.solve_shape == "auto":
        return pick_solve_shape(ctx, dwf, solve_once, tuned, log)
    shape = tuned if args.solve_shape == "tuned" or tuned.get("fp64") else parse_shape(args.solve_shape)
    shape = {k: shape[k] for k in ("tile_x", "prefetch", "chunks", "nt_store", "lds")}
    apply_shape(ctx, dwf, shape)
    return shape


def pick_streams(args, ctx, dwf, solve_once, log):
    """Concurrent sub-batch streams of a batched solve (DWF_TUNE_STREAMS), chosen on whole
    solves like the stencil shape: 1, 2, 4 and 8, each after a warm-up solve, timed twice;
    --streams fixes it."""
    if args.streams != "auto":
        ctx.tune(dwf.DWF_TUNE_STREAMS, int(args.streams))
        return int(args.streams)
    best = None
    for k in (1, 2, 4, 8):
        ctx.tune(dwf.DWF_TUNE_STREAMS, k)
        solve_once()
        ctx.synchronize()
        t = []
        for _ in range(2):
            t0 = time.perf_counter()
            solve_once()
            ctx.synchronize()
            t.append(time.perf_counter() - t0)
        log(f"streams {k}: {min(t) * 1e3:.2f} ms per solve")
        if best is None or min(t) < best[1]:
            best = (k, min(t))
    ctx.tune(dwf.DWF_TUNE_STREAMS, best[0])
    return best[0]


def time_dslash(ctx, dwf, capi, prec, nrhs, reps):
    n = nrhs * LS * NT * NX * 4
    dt = np.float32 if prec == dwf.DWF_PREC_FP32 else np.float64
    host = np.random.default_rng(77).normal(size=n).astype(dt)
    src = capi.DeviceArray.from_host(host)
    dst = capi.DeviceArray(host.nbytes)
    ms = time_kernel(ctx, capi, lambda: ctx.apply_device(dwf.DWF_OP_D, prec, src.ptr, dst.ptr, nrhs), reps)
    nbytes = dslash_bytes(4 if prec == dwf.DWF_PREC_FP32 else 8, nrhs)
    return ms, nbytes / (ms * 1e-3) / 1e9


def time_copy(ctx, capi, nbytes, reps):
    """Streaming copy of random data (HBM moves zeros measurably faster than random
    bits, and the stencil is timed on random data too)."""
    src = capi.DeviceArray.from_host(np.random.default_rng(78).normal(size=nbytes // 4).astype(np.float32))
    dst = capi.DeviceArray(nbytes)
    ms = time_kernel(ctx, capi, lambda: ctx.copy_device(dst.ptr, src.ptr, nbytes), reps)
    return ms, 2 * nbytes / (ms * 1e-3) / 1e9


def measured_traffic(shape):
    """HBM bytes per launch of the roofline kernel in this launch shape, from the
    committed counter summary (two rocprofv3 --pmc runs calibrated on the copy kernel),
    or None when that shape was not measured."""
    path = os.path.join(ROOT, "profiles", "k_hop_fp32_dwf96_batch128_traffic.md")
    key = f"traffic_bytes_per_launch lds={shape['lds']} tile_x={shape['tile_x']} ="
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line[len(key):].strip())
    except OSError:
        pass
    return None


def parse_shape(text):
    """'lds,tile_x' -> a launch shape of the LDS stencil with one s-chunk."""
    lds, tx = (int(v) for v in text.split(","))
    return {"tile_x": tx, "prefetch": 0, "chunks": 1, "nt_store": 1, "lds": lds}


def host_info():
    """CPU model, logical CPU count and OpenMP thread count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(theta, bs, nsolves):
    """The oracle's FP64 CG (the reference's algorithm and loop order, single-threaded as
    the reference's solves are) on `nsolves` whole 96x96x16 solves to 1e-10."""
    import oracle

    t0 = time.perf_counter()
    iters = []
    for b in bs[:nsolves]:
        _, it, _, cv, _ = oracle.dwf_cg(3, theta, b, NT, NX, LS, M5, MF, TOL, 100000)
        iters.append(it)
    dt = time.perf_counter() - t0
    return {
        "value": len(iters) / dt,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle FP64 CG (reference algorithm, 1 core) on {len(iters)} whole 96x96x16 boundary-source "
                   f"solve(s) to 1e-10: {iters} iterations in {dt:.1f} s"),
        "host": host_info(),
    }


ANALYZER_SIZES = (32, 56, 80, 96)


def analyzer_inputs(per_size):
    """Per lattice size: thetas [per][2*N*N] ~ N(0, 1/4) with seed 1234 + 104729 i, point
    sites drawn from seed 555 + i, Z4 seeds 918273 + 104729 i (i = global config index)."""
    out = []
    for g, N in enumerate(ANALYZER_SIZES):
        idx = [g * per_size + k for k in range(per_size)]
        thetas = np.stack([np.random.default_rng(1234 + 104729 * i).normal(0.0, 1.0 / np.sqrt(BETA), 2 * N * N)
                           for i in idx])
        sites = np.array([np.random.default_rng(555 + i).integers(0, N, size=2) for i in idx], np.int32)
        seeds = [918273 + 104729 * i for i in idx]
        out.append((N, thetas, sites, seeds))
    return out


def analyzer_sources(N, site, seed, nnoise):
    """The 2 point sources (spin 0, 1) at site = (t0, x0) and nnoise Z4 noise sources of one
    configuration, projected as correlator.hpp:60-75, 131-138 place them; [2 + nnoise][ndof]."""
    import oracle

    out = np.zeros((2 + nnoise, LS, N, N, 2), np.complex128)
    t0, x0 = int(site[0]), int(site[1])
    out[0, LS - 1, t0, x0, 0] = 1.0  # P_R e_0 at s = Ls-1
    out[1, 0, t0, x0, 1] = 1.0       # P_L e_1 at s = 0
    eta = oracle.z4_noise(N, N, nnoise, seed)
    out[2:, 0, :, :, 1] = eta[:, :, :, 1]
    out[2:, LS - 1, :, :, 0] = eta[:, :, :, 0]
    return out.reshape(2 + nnoise, -1)


def analyzer_cpu_baseline(inputs, nrhs):
    """The reference's analyzer on the CPU: the oracle's FP64 CGNR (the reference's
    algorithm) on every source of one configuration per lattice size, the solves spread
    over all OpenMP threads as domain_wall/analyze_configs.cpp:45 spreads its work.
    Contractions (O(V4) per solve) are left out."""
    import oracle

    thetas, bs, lats = [], [], []
    for N, th, sites, seeds in inputs:
        src = analyzer_sources(N, sites[0], seeds[0], 32)[:nrhs]
        for v in src:
            thetas.append(th[0])
            bs.append(oracle.dwf_apply(1, th[0], v, N, N, LS, M5, MF))
            lats.append((N, N))
    t0 = time.perf_counter()
    _, iters, conv, threads = oracle.dwf_cg_many(thetas, bs, lats, LS, M5, MF, TOL, 100000)
    dt = time.perf_counter() - t0
    per = {f"N{N}": int(np.mean(iters[i * nrhs:(i + 1) * nrhs])) for i, (N, _, _, _) in enumerate(inputs)}
    return {
        "value": len(bs) / dt,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP64 CGNR (reference algorithm) on {nrhs} sources (2 point + Z4) of one configuration "
                   f"per size, {len(bs)} whole solves to 1e-10 spread over {threads} OpenMP threads in {dt:.1f} s; "
                   f"mean iterations {per}; all converged: {bool(conv.all())}"),
        "host": host_info(),
    }


def run_analyzer(args, dwf):
    per = args.analyzer_configs
    inputs = analyzer_inputs(per)
    method = dwf.DWF_SOLVE_EO_MIXED
    ncfg = per * len(inputs)
    nsolves = ncfg * 34
    groups = [(N, N, thetas) for N, thetas, _, _ in inputs]
    sites = np.concatenate([s for _, _, s, _ in inputs])
    seeds = [sd for _, _, _, ss in inputs for sd in ss]
    if args.analyzer_split:
        step, close = analyzer_split_step(args, dwf, inputs, method)
    else:
        # every configuration of every size in one context: one batched solve whose
        # kernel launches each cover all four lattice sizes
        ctx = dwf.DwfContext.dwf(NT, NX, LS, M5, MF)
        if args.no_graphs:
            ctx.tune(dwf.DWF_TUNE_GRAPHS, 0)

        def step():
            t0 = time.perf_counter()
            ctx.set_gauge_groups(groups)
            m = ctx.measure_batch(ncfg, TOL, 20000, wall=False, point=True, max_momentum=0, noise_vectors=32,
                                  seeds=seeds, method=method, point_sites=sites)
            sec = time.perf_counter() - t0
            return [(sec, m[g * per:(g + 1) * per]) for g in range(len(inputs))]

        close = ctx.close

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    dt = time.perf_counter() - t0
    groups_out = {}
    for (N, _, _, _), (sec, m) in zip(inputs, out):
        its = np.concatenate([d["iterations"] for d in m])
        groups_out[f"N{N}"] = {"solves": per * 34, "mean_iterations": round(float(its.mean()), 1),
                               "max_iterations": int(its.max())}
        if args.analyzer_split:
            groups_out[f"N{N}"]["seconds"] = round(sec, 4)
        log(f"N={N}: {groups_out[f'N{N}']}")
    close()
    cpu = None if args.no_cpu else analyzer_cpu_baseline(inputs, args.cpu_rhs)
    if cpu:
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "DWF analyzer propagator solves/s (batched, 4 lattice sizes) @1e-10",
        "value": round(nsolves * args.steps / dt, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched analyzer: {ncfg} configurations, {per} each at Nx=Nt in "
                        f"{list(ANALYZER_SIZES)}, Ls=16, theta~N(0,1/4); per configuration 2 point sources at a "
                        "seeded site + 32 Z4 noise vectors; point C_PP, C_J5 and loop density in FP64",
            "solves_per_step": nsolves,
            "solver": "even-odd Schur complement, FP32 CGNR + FP64 reliable updates, certified at 1e-10",
            "launch": ("one context per lattice size on concurrent streams" if args.analyzer_split else
                       "one context holding all four lattice sizes: one batched solve, every kernel launch "
                       "covers all sizes"),
        },
        "groups": groups_out,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def analyzer_split_step(args, dwf, inputs, method):
    """The same work with one context and stream per lattice size (for comparison)."""
    import threading

    per = args.analyzer_configs
    ctxs = [dwf.DwfContext.dwf(N, N, LS, M5, MF) for N, _, _, _ in inputs]
    if args.no_graphs:
        for c in ctxs:
            c.tune(dwf.DWF_TUNE_GRAPHS, 0)

    def one(g, out):
        N, thetas, sites, seeds = inputs[g]
        t0 = time.perf_counter()
        ctxs[g].set_gauge_batch(thetas)
        m = ctxs[g].measure_batch(per, TOL, 20000, wall=False, point=True, max_momentum=0, noise_vectors=32,
                                  seeds=seeds, method=method, point_sites=sites)
        out[g] = (time.perf_counter() - t0, m)

    def step():
        out = [None] * len(inputs)
        th = [threading.Thread(target=one, args=(g, out)) for g in range(len(inputs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if any(o is None for o in out):
            raise SystemExit("analyzer step failed")
        return out

    def close():
        for c in ctxs:
            c.close()

    return step, close


def train_network(args, dwf, capi, hot=False):
    """Training pairs (D^dag eta, solution of D^dag D x = D^dag eta to 1e-10) from
    --learn-configs seeded configurations (seeds 1234 + 104729 i, i = 1 .. n: the test
    configuration, seed 1234, is held out; theta ~ N(0, 1/beta), or uniform on [-pi, pi)
    with hot=True, the HMC's hot start) with --learn-pairs full-volume Gaussian sources
    each, then device training.  Returns (network, statistics)."""
    V5 = NT * NX * LS
    ntrain, per = args.learn_configs, args.learn_pairs
    npairs = ntrain * per

    def config(i):
        g = np.random.default_rng(1234 + 104729 * i)
        return g.uniform(-np.pi, np.pi, 2 * NT * NX) if hot else g.normal(0.0, 1.0 / np.sqrt(BETA), 2 * NT * NX)

    thetas = np.stack([config(i) for i in range(1, ntrain + 1)])
    ctx = dwf.DwfContext.dwf(NT, NX, LS, M5, MF)
    if args.no_graphs:
        ctx.tune(dwf.DWF_TUNE_GRAPHS, 0)
    ctx.set_gauge_batch(thetas)
    # pairs: chunk j holds pair j of every configuration (right-hand side c uses config c)
    d_b32 = capi.DeviceArray(npairs * V5 * 16)
    d_x32 = capi.DeviceArray(npairs * V5 * 16)
    d_src = capi.DeviceArray(ntrain * V5 * 32)
    d_b = capi.DeviceArray(ntrain * V5 * 32)
    d_x = capi.DeviceArray(ntrain * V5 * 32)
    rng = np.random.default_rng(918273)
    t0 = time.perf_counter()
    for j in range(per):
        eta = rng.standard_normal(2 * ntrain * V5 * 2, dtype=np.float32).astype(np.float64)
        d_src.upload(eta)
        ctx.apply_device(dwf.DWF_OP_DDAG, dwf.DWF_PREC_FP64, d_src.ptr, d_b.ptr, ntrain)
        ctx.solve_dirac_device(dwf.DWF_SOLVE_EO_MIXED, d_src.ptr, None, TOL, 20000, ntrain, d_x.ptr)
        n = ntrain * V5 * 4
        ctx.convert_f64_f32(d_b.ptr, d_b32.ptr + j * n * 4, n)
        ctx.convert_f64_f32(d_x.ptr, d_x32.ptr + j * n * 4, n)
    ctx.synchronize()
    t_pairs = time.perf_counter() - t0
    log(f"{npairs} training pairs in {t_pairs:.1f} s")
    pair_cfg = np.tile(np.arange(ntrain), per)
    net = dwf.Network(LS, seed=11)
    t0 = time.perf_counter()
    hist = net.train_device(ctx, d_b32.ptr, d_x32.ptr, pair_cfg, npairs, steps=args.learn_steps,
                            batch=args.learn_batch, loss=args.learn_loss, lr=args.learn_lr, seed=5)
    t_train = time.perf_counter() - t0
    log(f"training: {args.learn_steps} steps in {t_train:.1f} s, loss {hist[0]:.4f} -> "
        f"{float(np.mean(hist[-20:])):.4f}")
    for d in (d_b32, d_x32, d_src, d_b, d_x):
        d.free()
    ctx.close()
    stats = {
        "training_pairs": npairs, "training_configs": ntrain, "pairs_seconds": round(t_pairs, 2),
        "training_seconds": round(t_train, 2), "training_steps": args.learn_steps, "batch": args.learn_batch,
        "lr": args.learn_lr, "loss": ["||M b - x||^2/||x||^2", "||A M b - b||^2/||b||^2"][args.learn_loss],
        "loss_first": float(hist[0]), "loss_last": float(np.mean(hist[-20:])),
    }
    return net, stats


LIB_RELIABLE_DELTA = 0.0  # the library's defaults: each solver's own residual drop between reliable updates


def learned_solver_config(args, dwf, ctx, net):
    """Attach the network to ctx with the learned solver's direction update; the learned
    solver's reliable-update threshold is set around its solves (solve_with)."""
    ctx.tune(dwf.DWF_TUNE_FCG_BETA, args.fcg_beta)
    if args.net_fused is not None:
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, args.net_fused)
    ctx.set_preconditioner(net)
    return ["flexible CG (FCG(1))", "preconditioned CG (Fletcher-Reeves beta)"][args.fcg_beta] + \
        f" in FP32 with FP64 reliable updates (residual drop {args.reliable_delta:g}), certified at 1e-10"


def run_learned(args, dwf, capi):
    V5 = NT * NX * LS
    net, tstats = train_network(args, dwf, capi)
    nrhs = args.nrhs
    # --nrhs held-out configurations (seeds 1234 + 104729 i, i = ntrain + 1 .., which the
    # training never saw), one boundary source each, solved as one batch
    i0 = args.learn_configs + 1
    thetas = np.stack([np.random.default_rng(1234 + 104729 * i).normal(0.0, 1.0 / np.sqrt(BETA), 2 * NT * NX)
                       for i in range(i0, i0 + nrhs)])
    ctx = dwf.DwfContext.dwf(NT, NX, LS, M5, MF)
    ctx.set_gauge_batch(thetas)
    if args.no_graphs:
        ctx.tune(dwf.DWF_TUNE_GRAPHS, 0)
    tuned64 = dict(ctx.autotune(nrhs, dwf.DWF_PREC_FP64, args.tune_cache), fp64=1)
    tuned = ctx.autotune(nrhs, dwf.DWF_PREC_FP32, args.tune_cache)
    log(f"autotuned stencil shapes for {nrhs} RHS: FP32 {tuned}, FP64 {tuned64}")
    solver = learned_solver_config(args, dwf, ctx, net)
    src = boundary_sources(nrhs, 918273)
    d_src = capi.DeviceArray.from_host(src.reshape(-1).view(np.float64))
    d_tsrc = capi.DeviceArray.from_host(boundary_sources(nrhs, TUNE_SOURCE_SEED).reshape(-1).view(np.float64))
    d_x = capi.DeviceArray(nrhs * V5 * 32)
    results, shapes = {}, {}
    for name, method in (("learned", dwf.DWF_SOLVE_LEARNED), ("fp64_cg", dwf.DWF_SOLVE_CG64),
                         ("eo_mixed", dwf.DWF_SOLVE_EO_MIXED), ("mixed", dwf.DWF_SOLVE_MIXED)):
        ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA,
                 args.reliable_delta if method == dwf.DWF_SOLVE_LEARNED else LIB_RELIABLE_DELTA)
        shapes[name] = solve_shape(
            args, ctx, dwf, lambda: ctx.solve_dirac_device(method, d_tsrc.ptr, None, TOL, 20000, nrhs, d_x.ptr),
            tuned64 if method == dwf.DWF_SOLVE_CG64 else tuned, log)
        shapes[name]["streams"] = pick_streams(
            args, ctx, dwf, lambda: ctx.solve_dirac_device(method, d_tsrc.ptr, None, TOL, 20000, nrhs, d_x.ptr), log)
        for _ in range(args.warmup):
            ctx.solve_dirac_device(method, d_src.ptr, None, TOL, 20000, nrhs, d_x.ptr)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st, _ = ctx.solve_dirac_device(method, d_src.ptr, None, TOL, 20000, nrhs, d_x.ptr)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        tr = max(q["true_relres"] for q in st)
        if tr > TOL or not all(q["converged"] for q in st):
            raise SystemExit(f"{name} solve failed certification: max true residual {tr:.3e}")
        results[name] = dict(solves_per_s=args.steps * nrhs / dt, ms_per_step=dt / args.steps * 1e3,
                             iterations=float(np.mean([q["iterations"] for q in st])),
                             fp64_iterations=float(np.mean([q["fp64_iterations"] for q in st])),
                             max_true_relres=tr)
        log(f"{name}: {results[name]}")
    # the FP32 D apply at 128 right-hand sides in the autotuner's shape (on a lattice no LDS
    # tile covers, the direct-load kernel: "lds" 0)
    d128 = ctx.autotune(args.dslash_rhs, dwf.DWF_PREC_FP32, args.tune_cache)
    ms_d, gbs_d = time_dslash(ctx, dwf, capi, dwf.DWF_PREC_FP32, args.dslash_rhs, 20)
    log(f"dslash fp32 x{args.dslash_rhs}: {ms_d:.4f} ms {gbs_d:.0f} GB/s, shape {d128}")
    fast, judge = results["learned"], results["fp64_cg"]
    line = {
        "metric": METRIC,
        "value": round(fast["solves_per_s"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fast["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"learned preconditioner: trained on {tstats['training_pairs']} (source, solution) pairs "
                        f"from {tstats['training_configs']} configurations (full-volume Gaussian sources), tested "
                        f"on {nrhs} held-out configurations (seeds 1234 + 104729 i, i = {i0} .. {i0 + nrhs - 1}), "
                        f"one boundary source each, solved as one batch per step, FP64 true residual <= 1e-10",
            "Nt": NT, "Nx": NX, "Ls": LS, "mf": MF, "beta": BETA, "rhs_per_step": nrhs,
            "lattice": args.lattice,
            "network": "3 linear parallel-transport conv layers 2-8-8-2 complex channels, BF16 MFMA",
            "solver": solver,
        },
        "training": tstats,
        "fp64_cg_solves_per_s": round(judge["solves_per_s"], 3),
        "eo_mixed_solves_per_s": round(results["eo_mixed"]["solves_per_s"], 3),
        "mixed_solves_per_s": round(results["mixed"]["solves_per_s"], 3),
        "speedup_vs_fp64_cg": round(fast["solves_per_s"] / judge["solves_per_s"], 3),
        "dslash_fp32_gbs": {"batch%d" % args.dslash_rhs: round(gbs_d, 1),
                            "frac_of_8TBs": round(gbs_d / HBM_PEAK_GBS, 4),
                            "shape": {k: d128[k] for k in ("tile_x", "prefetch", "chunks", "nt_store", "lds")}},
        "iterations": {k: v["iterations"] for k, v in results.items()},
        "iteration_ratio_vs_fp64_cg": round(fast["iterations"] / judge["iterations"], 3),
        "max_true_relres": {k: v["max_true_relres"] for k, v in results.items()},
        "solve_shapes": shapes,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    ctx.set_preconditioner(None)
    ctx.close()
    net.close()


def run_eo(args, dwf, capi):
    """EO_MIXED alone on the bench's batch (32 boundary sources, 96x96, Ls=16, seed-1234
    configuration): solves/s, iterations and the fraction of the Schur CG's HBM ceiling
    (eo_iteration_bytes), for kernel traces and counter runs of the even-odd kernels."""
    theta = np.random.default_rng(1234).normal(0.0, 1.0 / np.sqrt(BETA), size=2 * NT * NX)
    nrhs = args.nrhs
    ctx = dwf.DwfContext.dwf(NT, NX, LS, M5, MF)
    ctx.set_gauge(theta)
    if args.no_graphs:
        ctx.tune(dwf.DWF_TUNE_GRAPHS, 0)
    if args.streams != "auto":  # else the library's default
        ctx.tune(dwf.DWF_TUNE_STREAMS, int(args.streams))
    src = boundary_sources(nrhs, 918273)
    d_src = capi.DeviceArray.from_host(src.reshape(-1).view(np.float64))
    d_x = capi.DeviceArray(src.size * 16)
    for _ in range(args.warmup):
        ctx.solve_dirac_device(dwf.DWF_SOLVE_EO_MIXED, d_src.ptr, None, TOL, 20000, nrhs, d_x.ptr)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st, _ = ctx.solve_dirac_device(dwf.DWF_SOLVE_EO_MIXED, d_src.ptr, None, TOL, 20000, nrhs, d_x.ptr)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    tr = max(q["true_relres"] for q in st)
    if tr > TOL or not all(q["converged"] for q in st):
        raise SystemExit(f"eo_mixed solve failed certification: max true residual {tr:.3e}")
    its = float(np.mean([q["iterations"] for q in st]))
    sps = args.steps * nrhs / dt
    ceiling = 8e12 / (its * eo_iteration_bytes())
    line = {
        "metric": "DWF EO_MIXED solves/s 96^2,Ls=16 @1e-10", "value": round(sps, 3), "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "even-odd Schur complement solves (fused Schur operator, pipelined FP32 CG, FP64 "
                               "reliable updates), 32 boundary sources, 96x96 Ls=16, certified at 1e-10",
                   "Nt": NT, "Nx": NX, "Ls": LS, "rhs_per_step": nrhs},
        "iterations": its, "max_true_relres": tr,
        "hbm_ceiling_solves_per_s": round(ceiling, 1), "fraction_of_ceiling": round(sps / ceiling, 4),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def run_hmc(args, dwf, capi):
    # hot start theta ~ U[-pi, pi) (seed 1234), or the reference's cold start theta = 1
    # (domain_wall/generate_configs.cpp:27-28)
    theta0 = (np.random.default_rng(1234).uniform(-np.pi, np.pi, 2 * NT * NX) if args.hmc_start == "hot"
              else np.ones(2 * NT * NX))
    ctx = dwf.DwfContext.dwf(NT, NX, LS, M5, MF)
    ctx.set_gauge(theta0)
    if args.no_graphs:
        ctx.tune(dwf.DWF_TUNE_GRAPHS, 0)
    rng = capi.HostRng(1234)
    ndof, ngauge = LS * NT * NX * 2, 2 * NT * NX
    method = {"cg64": dwf.DWF_SOLVE_CG64, "mixed": dwf.DWF_SOLVE_MIXED, "eo_mixed": dwf.DWF_SOLVE_EO_MIXED,
              "learned": dwf.DWF_SOLVE_LEARNED}[args.hmc_solver]
    # force solves have one right-hand side: stencil shape tuned for that (graph-timed)
    shape = ctx.autotune(1, dwf.DWF_PREC_FP64 if method == dwf.DWF_SOLVE_CG64 else dwf.DWF_PREC_FP32,
                         args.tune_cache)
    log(f"autotuned stencil shape for 1 RHS: {shape}")
    eps, nsteps = 0.05, 20
    net, tstats = None, None
    if method == dwf.DWF_SOLVE_LEARNED:
        # trained on configurations like the chain's (hot: uniform angles; cold: theta ~
        # N(0, 1/beta), what the cold start relaxes to), the trajectory's own seed held out
        net, tstats = train_network(args, dwf, capi, hot=args.hmc_start == "hot")
        learned_solver_config(args, dwf, ctx, net)
        ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA, args.reliable_delta)

    def trajectory():
        chi, mom = rng.dwf_heatbath(ndof, ngauge)
        u = rng.uniform()
        return ctx.hmc_trajectory(chi, mom, u, BETA, eps, nsteps, TOL, 20000, method=method)

    for _ in range(args.warmup):
        trajectory()
    ctx.synchronize()
    t0 = time.perf_counter()
    hist = [trajectory() for _ in range(args.steps)]
    ctx.synchronize()
    dt = time.perf_counter() - t0
    for i, h in enumerate(hist):
        log(f"trajectory {i}: {h}")
    cg_per_traj = float(np.mean([h["cg_iterations"] for h in hist]))
    cpu = None
    if not args.no_cpu:
        import oracle

        # the reference's FP64 CG iterations per trajectory: the timed ones for cg64, else
        # the chain's first trajectory re-run (untimed) with FP64 CG force solves
        ref_iters = cg_per_traj
        if method != dwf.DWF_SOLVE_CG64:
            with dwf.DwfContext.dwf(NT, NX, LS, M5, MF) as c64:
                c64.set_gauge(theta0)
                chi, mom = capi.HostRng(1234).dwf_heatbath(ndof, ngauge)
                ref_iters = float(c64.hmc_trajectory(chi, mom, 0.5, BETA, eps, nsteps, TOL, 20000,
                                                     method=dwf.DWF_SOLVE_CG64)["cg_iterations"])
            log(f"FP64 CG iterations of the first trajectory: {ref_iters:.0f}")
        src = boundary_sources(1, 918273)[0].reshape(-1)
        t1 = time.perf_counter()
        _, it, _, _, _ = oracle.dwf_cg(2, theta0, src, NT, NX, LS, M5, MF, TOL, args.cpu_iters)
        per_iter = (time.perf_counter() - t1) / max(1, it)
        cpu = {"value": 1.0 / (per_iter * ref_iters), "unit": "trajectories/s", "cores": 1, "kind": "port",
               "sample": (f"oracle FP64 CG on D D^dag (reference algorithm, 1 core): {it} iterations at 96x96x16 "
                          f"timed ({per_iter * 1e3:.1f} ms/iteration), scaled by the {ref_iters:.0f} FP64 CG "
                          "iterations of one trajectory" + ("" if method == dwf.DWF_SOLVE_CG64 else
                                                            " (the chain's first, re-run with FP64 CG)")
                          + "; forces and gauge update excluded")}
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "DWF HMC trajectories/s 96²,Ls=16 (tau=1, 20 leapfrog steps)",
        "value": round(args.steps / dt, 4),
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"cg64": "f64", "mixed": "f32+f64", "eo_mixed": "f32+f64",
                  "learned": "bf16+f32+f64"}[args.hmc_solver],
        "data": "synthetic",
        "config": {
            "workload": "Nf=2 Shamir DWF pseudofermion HMC, 96x96 Ls=16 beta=4 mf=0.05 M5=1, tau=1, 20 leapfrog "
                        + ("steps, hot start seed 1234" if args.hmc_start == "hot" else "steps, cold start theta=1")
                        + ", force solves to 1e-10 with chronological guesses",
            "force_solver": args.hmc_solver,
            "start": args.hmc_start,
        },
        "acceptance_rate": float(np.mean([h["accepted"] for h in hist])),
        "cg_iterations_per_trajectory": cg_per_traj,
        "stencil_shape": {k: shape[k] for k in ("tile_x", "prefetch", "chunks", "nt_store", "lds")},
        "training": tstats,
        "delta_h": [h["delta_h"] for h in hist],
        "plaquette": [h["plaquette"] for h in hist],
        "accepted": [h["accepted"] for h in hist],
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    if net is not None:
        net.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nrhs", type=int, default=32, help="right-hand sides per step")
    ap.add_argument("--dslash-rhs", type=int, default=128)
    ap.add_argument("--solve-shape", default="auto",
                    help="stencil shape of the solvers: 'auto' (picked on whole solves), 'tuned' (the "
                         "autotuner's D-apply shape) or 'lds,tile_x' (FP32 solvers)")
    ap.add_argument("--streams", default="auto", choices=("auto", "1", "2", "3", "4", "6", "8"),
                    help="concurrent sub-batch streams of the batched solves (DWF_TUNE_STREAMS): 'auto' "
                         "picks 1, 2, 4 or 8 per solver on whole solves")
    ap.add_argument("--dslash-shape", default=None,
                    help="'lds,tile_x': time the FP32 D apply in this LDS tile shape instead of the autotuned one")
    ap.add_argument("--cpu-iters", type=int, default=60,
                    help="oracle CG iterations timed for the HMC workload's CPU baseline")
    ap.add_argument("--cpu-solves", type=int, default=1,
                    help="whole oracle solves timed for the solves workload's CPU baseline")
    ap.add_argument("--cpu-rhs", type=int, default=8,
                    help="analyzer CPU baseline: sources per configuration solved (34 = all; 8 x 4 sizes "
                         "= two solves per thread of 16, about 30 s)")
    ap.add_argument("--analyzer-split", action="store_true",
                    help="analyzer: one context per lattice size on concurrent streams instead of one launch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("solves", "analyzer", "learned", "hmc", "eo"), default="solves")
    ap.add_argument("--hmc-solver", choices=("cg64", "mixed", "eo_mixed", "learned"), default="mixed")
    ap.add_argument("--hmc-start", choices=("hot", "cold"), default="hot",
                    help="HMC start: theta ~ U[-pi, pi) (seed 1234) or the reference's theta = 1")
    ap.add_argument("--learn-configs", type=int, default=64)
    ap.add_argument("--learn-pairs", type=int, default=64, help="training pairs per configuration")
    ap.add_argument("--learn-steps", type=int, default=1000)
    ap.add_argument("--learn-batch", type=int, default=16)
    ap.add_argument("--learn-loss", type=int, default=1, choices=(0, 1))
    ap.add_argument("--learn-lr", type=float, default=3e-3)
    ap.add_argument("--reliable-delta", type=float, default=1e-4,
                    help="learned solver: residual drop between FP64 reliable updates (library default 0.01)")
    ap.add_argument("--net-fused", type=int, default=None, choices=(0, 1, 2),
                    help="learned network kernel: 0 layer by layer, 1 fused, 2 fused in a gauge per block "
                         "(default: the library's)")
    ap.add_argument("--fcg-beta", type=int, default=1, choices=(0, 1),
                    help="learned solver direction update: 0 flexible CG (FCG(1)), 1 Fletcher-Reeves")
    ap.add_argument("--analyzer-configs", type=int, default=32, help="configurations per lattice size")
    ap.add_argument("--lattice", choices=("bench", "reference-default"), default="bench",
                    help="learned workload's lattice: 96x96 Ls=16, mf 0.05, beta 4, or the reference's own default "
                         "(parameters.txt:25-30: 32x50, N5 = 12, mf 0.2, beta 3), which no LDS stencil tile covers")
    ap.add_argument("--tune-cache", default=None, help="autotuner result file (reused when present)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch CG iterations one by one instead of replaying HIP graphs (for tracing tools)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the roofline launches (batched FP32 D apply and the copy kernel), for counter runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (independent "
              "configurations shard without communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import three_part_dwf_study_amd as dwf
    from three_part_dwf_study_amd import capi

    if args.workload == "analyzer":
        run_analyzer(args, dwf)
        return
    if args.lattice == "reference-default":
        if args.workload != "learned":
            raise SystemExit("--lattice reference-default applies to --workload learned")
        global NT, NX, LS, MF, BETA
        NT, NX, LS, MF, BETA = 32, 50, 12, 0.2, 3.0
    if args.workload == "learned":
        run_learned(args, dwf, capi)
        return
    if args.workload == "hmc":
        run_hmc(args, dwf, capi)
        return
    if args.workload == "eo":
        run_eo(args, dwf, capi)
        return
    theta = np.random.default_rng(1234).normal(0.0, 1.0 / np.sqrt(BETA), size=2 * NT * NX)
    if args.kernels_only:
        ctx = dwf.DwfContext.dwf(NT, NX, LS, M5, MF)
        ctx.set_gauge(theta)
        if args.dslash_shape:
            shape = parse_shape(args.dslash_shape)
            apply_shape(ctx, dwf, shape)
        else:
            shape = ctx.autotune(args.dslash_rhs, dwf.DWF_PREC_FP32, args.tune_cache)
        ms32, gbs32 = time_dslash(ctx, dwf, capi, dwf.DWF_PREC_FP32, args.dslash_rhs, 5)
        cb = dslash_bytes(4, args.dslash_rhs) // 2 // 16 * 16
        msc, gbsc = time_copy(ctx, capi, cb, 5)
        print(json.dumps({"dslash_fp32_us": ms32 * 1e3, "dslash_fp32_gbs": gbs32,
                          "dslash_bytes": dslash_bytes(4, args.dslash_rhs), "copy_us": msc * 1e3,
                          "copy_gbs": gbsc, "copy_bytes_each_way": cb,
                          "shape": {k: shape[k] for k in ("lds", "tile_x", "chunks")}}), flush=True)
        ctx.close()
        return
    nrhs = args.nrhs
    src = boundary_sources(nrhs, 918273)
    net, tstats = train_network(args, dwf, capi)  # outside the timed region
    ctx = dwf.DwfContext.dwf(NT, NX, LS, M5, MF)
    ctx.set_gauge(theta)
    if args.no_graphs:
        ctx.tune(dwf.DWF_TUNE_GRAPHS, 0)

    # each solver runs with the stencil shape tuned for the precision its iterations use
    tuned = {"solves_fp64": dict(ctx.autotune(nrhs, dwf.DWF_PREC_FP64, args.tune_cache), fp64=1),
             "solves": ctx.autotune(nrhs, dwf.DWF_PREC_FP32, args.tune_cache)}
    log(f"autotuned stencil shapes for {nrhs} RHS: FP32 {tuned['solves']}, FP64 {tuned['solves_fp64']}")
    learned_solver = learned_solver_config(args, dwf, ctx, net)
    n = nrhs * LS * NT * NX * 4
    d_src = capi.DeviceArray.from_host(src.reshape(-1).view(np.float64))
    d_tsrc = capi.DeviceArray.from_host(boundary_sources(nrhs, TUNE_SOURCE_SEED).reshape(-1).view(np.float64))
    d_x = capi.DeviceArray(n * 8)

    def solve(method):
        st, rc = ctx.solve_dirac_device(method, d_src.ptr, None, TOL, 20000, nrhs, d_x.ptr)
        return st

    results = {}
    for name, method in (("learned", dwf.DWF_SOLVE_LEARNED), ("eo_mixed", dwf.DWF_SOLVE_EO_MIXED),
                         ("mixed", dwf.DWF_SOLVE_MIXED), ("fp64_cg", dwf.DWF_SOLVE_CG64)):
        ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA,
                 args.reliable_delta if method == dwf.DWF_SOLVE_LEARNED else LIB_RELIABLE_DELTA)
        tuned["solve_" + name] = solve_shape(
            args, ctx, dwf, lambda: ctx.solve_dirac_device(method, d_tsrc.ptr, None, TOL, 20000, nrhs, d_x.ptr),
            tuned["solves_fp64" if method == dwf.DWF_SOLVE_CG64 else "solves"], log)
        tuned["solve_" + name]["streams"] = pick_streams(
            args, ctx, dwf, lambda: ctx.solve_dirac_device(method, d_tsrc.ptr, None, TOL, 20000, nrhs, d_x.ptr), log)
        steps = args.steps
        for _ in range(args.warmup):
            solve(method)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = solve(method)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        tr = max(s["true_relres"] for s in st)
        if tr > TOL or not all(s["converged"] for s in st):
            raise SystemExit(f"{name} solve failed certification: max true residual {tr:.3e}")
        results[name] = dict(seconds=dt, steps=steps, solves_per_s=steps * nrhs / dt,
                             ms_per_step=dt / steps * 1e3,
                             iterations=float(np.mean([s["iterations"] for s in st])),
                             reliable_updates=float(np.mean([s["reliable_updates"] for s in st])),
                             passes=float(np.mean([s["passes"] for s in st])),
                             max_true_relres=tr)
        log(f"{name}: {results[name]}")

    # the same solves through the host-array entry point (sources in and solutions out over
    # PCIe, as the reference's call sites hand over host vectors)
    host_src = src.reshape(nrhs, -1)
    pcie = {}
    for name, method in (("learned", dwf.DWF_SOLVE_LEARNED), ("eo_mixed", dwf.DWF_SOLVE_EO_MIXED)):
        ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA,
                 args.reliable_delta if method == dwf.DWF_SOLVE_LEARNED else LIB_RELIABLE_DELTA)
        apply_shape(ctx, dwf, tuned["solve_" + name])
        ctx.tune(dwf.DWF_TUNE_STREAMS, tuned["solve_" + name]["streams"])
        ctx.solve_dirac(host_src, rtol=TOL, maxiter=20000, method=method)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.solve_dirac(host_src, rtol=TOL, maxiter=20000, method=method)
        pcie[name] = args.steps * nrhs / (time.perf_counter() - t0)
        log(f"{name} with host arrays (PCIe-inclusive): {pcie[name]:.1f} solves/s")

    if args.dslash_shape:
        tuned["fp32_batch"] = parse_shape(args.dslash_shape)
        apply_shape(ctx, dwf, tuned["fp32_batch"])
    else:
        tuned["fp32_batch"] = ctx.autotune(args.dslash_rhs, dwf.DWF_PREC_FP32, args.tune_cache)
    ms32, gbs32 = time_dslash(ctx, dwf, capi, dwf.DWF_PREC_FP32, args.dslash_rhs, 20)
    tuned["fp64_batch"] = ctx.autotune(args.dslash_rhs // 2, dwf.DWF_PREC_FP64, args.tune_cache)
    ms64, gbs64 = time_dslash(ctx, dwf, capi, dwf.DWF_PREC_FP64, args.dslash_rhs // 2, 20)
    tuned["fp32_single"] = ctx.autotune(1, dwf.DWF_PREC_FP32, args.tune_cache)
    ms1e, gbs1e = time_dslash(ctx, dwf, capi, dwf.DWF_PREC_FP32, 1, 200)
    # one right-hand side is a few microseconds of kernel time: eager launches from here
    # measure the host's launch rate, the autotuner's graph-timed figure the kernel
    ms1 = tuned["fp32_single"]["microseconds"] * 1e-3
    gbs1 = dslash_bytes(4, 1) / (ms1 * 1e-3) / 1e9
    msc, gbsc = time_copy(ctx, capi, dslash_bytes(4, args.dslash_rhs) // 2 // 16 * 16, 20)
    log(f"copy kernel ({dslash_bytes(4, args.dslash_rhs) // 2 >> 20} MiB each way): {msc:.4f} ms {gbsc:.0f} GB/s")
    log(f"dslash fp32 x{args.dslash_rhs}: {ms32:.4f} ms {gbs32:.0f} GB/s; fp64 x{args.dslash_rhs // 2}: "
        f"{ms64:.4f} ms {gbs64:.0f} GB/s; fp32 x1 (cache-resident): {ms1 * 1e3:.1f} us {gbs1:.0f} GB/s "
        f"(eager launches from Python: {ms1e * 1e3:.1f} us)")

    cpu = None
    if not args.no_cpu:
        import oracle

        bs = [oracle.dwf_apply(1, theta, src[i].reshape(-1), NT, NX, LS, M5, MF)
              for i in range(min(args.cpu_solves, nrhs))]
        cpu = cpu_baseline(theta, bs, args.cpu_solves)
        log(f"cpu baseline: {cpu}")

    fast, judge, eo = results["learned"], results["fp64_cg"], results["eo_mixed"]
    line = {
        "metric": METRIC,
        "value": round(fast["solves_per_s"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": fast["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(fast["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "DWF CGNR solves, 96x96 Ls=16 M5=1 mf=0.05, theta~N(0,1/4) seed 1234, "
                        "Gaussian boundary sources, FP64 true residual <= 1e-10",
            "Nt": NT, "Nx": NX, "Ls": LS, "rhs_per_step": nrhs,
            "solver": "learned gauge-equivariant preconditioner (BF16 MFMA network) + " + learned_solver,
            "network": "3 linear parallel-transport conv layers 2-8-8-2 complex channels, trained on the device",
        },
        "training": tstats,
        "fp64_cg_solves_per_s": round(judge["solves_per_s"], 3),
        "speedup_vs_fp64_cg": round(fast["solves_per_s"] / judge["solves_per_s"], 3),
        "eo_mixed_solves_per_s": round(eo["solves_per_s"], 3),
        "eo_mixed_speedup_vs_fp64_cg": round(eo["solves_per_s"] / judge["solves_per_s"], 3),
        "mixed_solves_per_s": round(results["mixed"]["solves_per_s"], 3),
        "pcie_inclusive_solves_per_s": {k: round(v, 3) for k, v in pcie.items()},
        "iterations": {k: v["iterations"] for k, v in results.items()},
        # HBM-roofline ceiling of the CG solvers: 8 TB/s / (iterations x algorithmic bytes per iteration)
        "roofline_ceiling_solves_per_s": {
            "fp64_cg": round(8e12 / (results["fp64_cg"]["iterations"] * cg_iteration_bytes(8)), 1),
            "mixed": round(8e12 / (results["mixed"]["iterations"] * cg_iteration_bytes(4)), 1),
            "eo_mixed": round(8e12 / (results["eo_mixed"]["iterations"] * eo_iteration_bytes()), 1)},
        "reliable_updates": {k: v["reliable_updates"] for k, v in results.items()},
        "max_true_relres": {k: v["max_true_relres"] for k, v in results.items()},
        "dslash_gbs": {"fp32_batch%d" % args.dslash_rhs: round(gbs32, 1),
                       "fp64_batch%d" % (args.dslash_rhs // 2): round(gbs64, 1),
                       "fp32_single_rhs_cache_resident": round(gbs1, 1),
                       "fp32_single_rhs_eager_launches": round(gbs1e, 1),
                       "copy_kernel_achievable": round(gbsc, 1)},
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs32, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs32 / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(tuned["fp32_batch"]) if args.dslash_rhs == 128 else None,
            "kernel": (f"k_hop_lds<float,D,DWF,EPI_STORE,{LDS_THREADS[tuned['fp32_batch']['lds']]}>"
                       if tuned["fp32_batch"]["lds"] else "k_hop<float,D,DWF,EPI_STORE>"),
            "avg_launch_us": round(ms32 * 1e3, 2),
            "bytes_per_launch": dslash_bytes(4, args.dslash_rhs),
        },
        "cpu_baseline": cpu,
        "autotuned_shapes": {k: {kk: v[kk] for kk in ("tile_x", "prefetch", "chunks", "nt_store", "lds", "streams")
                                 if kk in v}
                             for k, v in tuned.items()},
    }
    print(json.dumps(line), flush=True)
    ctx.set_preconditioner(None)
    ctx.close()
    net.close()


if __name__ == "__main__":
    main()
