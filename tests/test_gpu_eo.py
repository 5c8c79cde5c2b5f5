"""GPU tests of the even-odd Schur-complement solver (dwf_solve_dirac, EO_MIXED):
x = D^-1 src certified by the oracle's own FP64 operator on the reference's system
||D^dag src - D^dag D x|| / ||D^dag src||, without falling back to the full-lattice
solver (passes >= 1), and equal to the oracle's CG solution."""

import numpy as np
import pytest

import oracle
from conftest import print_case
from dense_ref import rand_spinor, rand_theta

import three_part_dwf_study_amd as dwf

pytestmark = pytest.mark.gpu

M5, MF = 1.0, 0.05


def normal_relres(th, x, src, Nt, Nx, Ls, mf=MF):
    B = oracle.dwf_apply(1, th, src, Nt, Nx, Ls, M5, mf)
    r = B - oracle.dwf_apply(3, th, x, Nt, Nx, Ls, M5, mf)
    return np.linalg.norm(r) / np.linalg.norm(B)


@pytest.mark.parametrize("Nt,Nx,Ls,nrhs", [(8, 12, 4, 2), (16, 16, 8, 1), (12, 8, 12, 3), (96, 96, 16, 2)])
def test_eo_solve_certified(Nt, Nx, Ls, nrhs):
    print_case(test="eo_solve", Nt=Nt, Nx=Nx, Ls=Ls, nrhs=nrhs)
    rng = np.random.default_rng(Nt * 100 + Ls)
    th = rand_theta(rng, Nt, Nx)
    src = rand_spinor(rng, nrhs * Ls * Nt * Nx * 2).reshape(nrhs, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        x, st, rc = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_EO_MIXED)
        x2, st2, rc2 = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_MIXED)
    assert rc == 0
    for i in range(nrhs):
        assert st[i]["converged"] and 1 <= st[i]["passes"] <= 3, st[i]
        tr = normal_relres(th, x[i], src[i], Nt, Nx, Ls)
        assert tr <= 1e-10, tr
        assert abs(tr - st[i]["true_relres"]) <= 1e-12
        assert np.linalg.norm(x[i] - x2[i]) / np.linalg.norm(x2[i]) < 1e-6
        # the even-odd system needs fewer iterations than the full normal equations
        assert st[i]["iterations"] < st2[i]["iterations"], (st[i]["iterations"], st2[i]["iterations"])


@pytest.mark.parametrize("Nt,Nx,Ls,nrhs", [(8, 12, 4, 2), (12, 10, 8, 2), (96, 96, 16, 3)])
def test_eo_fused_matches_two_kernel(Nt, Nx, Ls, nrhs):
    """The one-kernel Schur operator (DWF_TUNE_EO_FUSED 1: A^-1 H psi kept in LDS, tiles of
    8 x 16, 6 x 10 ...) against the two-kernel form (0, A^-1 H psi through device memory):
    the same arithmetic, so the same iterations and, both certified, the same solutions.
    The pipelined form (2, the default: p = r + beta p inside the next operator launch,
    directions double-buffered per right-hand side) continues a cycle after a reliable
    update from the new residual, so its iterations may differ by a few: certified, and
    its solution equal to the others within the tolerance's reach (1e-6)."""
    print_case(test="eo_fused", Nt=Nt, Nx=Nx, Ls=Ls, nrhs=nrhs)
    rng = np.random.default_rng(Nt * 7 + Nx)
    th = rand_theta(rng, Nt, Nx)
    src = rand_spinor(rng, nrhs * Ls * Nt * Nx * 2).reshape(nrhs, -1)
    out = []
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        for fused in (0, 1, 2):
            ctx.tune(dwf.DWF_TUNE_EO_FUSED, fused)
            out.append(ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_EO_MIXED))
    (x0, st0, rc0), (x1, st1, rc1), (x2, st2, rc2) = out
    assert rc0 == 0 and rc1 == 0 and rc2 == 0
    for i in range(nrhs):
        for st in (st1, st2):
            assert st[i]["converged"] and st[i]["passes"] >= 1, st[i]
        assert abs(st1[i]["iterations"] - st0[i]["iterations"]) <= 2, (st0[i], st1[i])
        assert abs(st2[i]["iterations"] - st0[i]["iterations"]) <= 0.05 * st0[i]["iterations"] + 4, (st0[i], st2[i])
        if Nt * Nx <= 144:
            assert normal_relres(th, x1[i], src[i], Nt, Nx, Ls) <= 1e-10
            assert normal_relres(th, x2[i], src[i], Nt, Nx, Ls) <= 1e-10
        assert np.linalg.norm(x1[i] - x0[i]) / np.linalg.norm(x0[i]) < 1e-8
        assert np.linalg.norm(x2[i] - x0[i]) / np.linalg.norm(x0[i]) < 1e-6


@pytest.mark.parametrize("graphs", [1, 0])
@pytest.mark.parametrize("Nt,Nx,Ls,nrhs,groups", [(12, 10, 8, 3, False), (96, 96, 16, 4, False), (8, 8, 4, 0, True)])
def test_eo_state_fold_matches_state_kernel(Nt, Nx, Ls, nrhs, groups, graphs):
    """DWF_TUNE_EO_FUSED 3 (the default): in the replayed blocks of 8 iterations the
    operator launch that forms the direction also does the previous iteration's state
    update (convergence, freezing, beta, the direction buffer), which 2 launches as a kernel
    of its own.  The same sums in the same order, so the same iterations and bit-identical
    solutions, with graphs (the fold) and without (no fold: the state kernel every
    iteration), on one lattice and on lattice groups of several sizes."""
    print_case(test="eo_state_fold", Nt=Nt, Nx=Nx, Ls=Ls, nrhs=nrhs, groups=groups, graphs=graphs)
    out = []
    for var in (2, 3):
        if groups:
            # two configurations of 8 x 8 and one of 12 x 12, two right-hand sides each
            r2 = np.random.default_rng(Nt * 5 + Nx + nrhs)
            grp = [(8, 8, np.stack([rand_theta(r2, 8, 8) for _ in range(2)])),
                   (12, 12, rand_theta(r2, 12, 12)[None, :])]
            src = np.concatenate([rand_spinor(r2, Ls * a * b * 2) for a, b, th in grp for _ in range(2 * len(th))])
            with dwf.DwfContext.dwf(8, 8, Ls, M5, MF) as ctx:
                ctx.set_gauge_groups(grp)
                ctx.tune(dwf.DWF_TUNE_GRAPHS, graphs)
                ctx.tune(dwf.DWF_TUNE_EO_FUSED, var)
                out.append(ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_EO_MIXED))
        else:
            r2 = np.random.default_rng(Nt * 5 + Nx + nrhs)
            th = rand_theta(r2, Nt, Nx)
            src = rand_spinor(r2, nrhs * Ls * Nt * Nx * 2).reshape(nrhs, -1)
            with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
                ctx.set_gauge(th)
                ctx.tune(dwf.DWF_TUNE_GRAPHS, graphs)
                ctx.tune(dwf.DWF_TUNE_EO_FUSED, var)
                out.append(ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_EO_MIXED))
    (xa, sta, rca), (xb, stb, rcb) = out
    assert rca == 0 and rcb == 0
    assert np.array_equal(np.asarray(xa), np.asarray(xb))
    for a, b in zip(sta, stb):
        assert a["converged"] and b["converged"] and b["true_relres"] <= 1e-10
        assert a["iterations"] == b["iterations"] and a["true_relres"] == b["true_relres"], (a, b)


def test_eo_reliable_delta_default():
    """The even-odd CG's reliable updates follow its own residual drop (3e-4) unless one
    is set; setting 0 restores it.  Every solve is certified either way."""
    Nt, Nx, Ls, nrhs = 32, 32, 16, 4
    print_case(test="eo_delta")
    rng = np.random.default_rng(515)
    th = rand_theta(rng, Nt, Nx)
    src = rand_spinor(rng, nrhs * Ls * Nt * Nx * 2).reshape(nrhs, -1)
    res = {}
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_STREAMS, 1)
        for key, delta in (("default", None), ("1e-2", 1e-2), ("3e-4", 3e-4), ("reset", 0.0)):
            if delta is not None:
                ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA, delta)
            res[key] = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_EO_MIXED)
        with pytest.raises(dwf.DwfError):
            ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA, 1.0)
    for key, (x, st, rc) in res.items():
        assert rc == 0, key
        for r in range(nrhs):
            assert st[r]["converged"] and st[r]["true_relres"] <= 1e-10, (key, r)
            assert normal_relres(th, x[r], src[r], Nt, Nx, Ls) <= 1.01e-10, (key, r)
    rel = lambda k: sum(q["reliable_updates"] for q in res[k][1])
    assert rel("default") < rel("1e-2"), (rel("default"), rel("1e-2"))
    # the default is 3e-4, and 0 restores it: the same solves bit for bit
    assert np.array_equal(res["default"][0], res["3e-4"][0])
    assert np.array_equal(res["default"][0], res["reset"][0])


def test_eo_matches_oracle_cg():
    Nt, Nx, Ls = 8, 8, 8
    print_case(test="eo_vs_oracle")
    rng = np.random.default_rng(42)
    th = rand_theta(rng, Nt, Nx)
    src = rand_spinor(rng, Ls * Nt * Nx * 2)
    B = oracle.dwf_apply(1, th, src, Nt, Nx, Ls, M5, MF)
    xo, it, rr, conv, rc = oracle.dwf_cg(3, th, B, Nt, Nx, Ls, M5, MF, 1e-12, 20000)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        x, st, rc = ctx.solve_dirac(src, rtol=1e-12, maxiter=20000)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-9


def test_eo_falls_back_on_odd_extent_and_guess():
    Nt, Nx, Ls = 7, 6, 4
    print_case(test="eo_fallback")
    rng = np.random.default_rng(3)
    th = rand_theta(rng, Nt, Nx)
    src = rand_spinor(rng, Ls * Nt * Nx * 2)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        x, st, rc = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000)
        assert rc == 0 and st[0]["passes"] == 0 and normal_relres(th, x, src, Nt, Nx, Ls) <= 1e-10
    with dwf.DwfContext.dwf(8, 6, 4, M5, MF) as ctx:
        th = rand_theta(rng, 8, 6)
        src = rand_spinor(rng, 4 * 8 * 6 * 2)
        ctx.set_gauge(th)
        x, st, rc = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000)
        g = x + 1e-4 * rand_spinor(rng, x.size)
        x2, st2, rc = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, x0=g)
        assert rc == 0 and normal_relres(th, x2, src, 8, 6, 4) <= 1e-10
        assert st2[0]["iterations"] < st[0]["iterations"]
        # a guess that already solves the system is certified first: no pass for a batch of
        # such guesses, no iteration for such a right-hand side beside one that needs them
        x3, st3, rc = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, x0=x)
        assert rc == 0 and st3[0]["converged"] and st3[0]["passes"] == 0 and st3[0]["iterations"] == 0
        assert np.array_equal(x3, x)
        x4, st4, rc = ctx.solve_dirac(np.stack([src, src]), rtol=1e-10, maxiter=20000, x0=np.stack([x, g]))
        assert rc == 0 and st4[0]["converged"] and st4[0]["iterations"] == 0 and np.array_equal(x4[0], x)
        assert st4[1]["converged"] and st4[1]["passes"] >= 1 and normal_relres(th, x4[1], src, 8, 6, 4) <= 1e-10


def test_autotune_batch(tmp_path):
    """Right-hand-side blocking: the tuner times whole solves over doubling batch sizes and
    reports the best; a second call with the same cache file reuses the result."""
    Nt, Nx, Ls = 16, 16, 8
    print_case(test="autotune_batch")
    rng = np.random.default_rng(12)
    cache = str(tmp_path / "tune.txt")
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(rand_theta(rng, Nt, Nx))
        r = ctx.autotune_batch(dwf.DWF_SOLVE_EO_MIXED, 1e-10, 64, cache)
        assert r["nrhs"] in (8, 16, 32, 64) and r["solves_per_s"] > 0 and r["from_cache"] == 0
        r2 = ctx.autotune_batch(dwf.DWF_SOLVE_EO_MIXED, 1e-10, 64, cache)
        assert r2["from_cache"] == 1 and r2["nrhs"] == r["nrhs"]
        # the context is left as it was: a single-configuration solve still works
        src = rand_spinor(rng, Ls * Nt * Nx * 2)
        _, st, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_EO_MIXED)
        assert st[0]["converged"]
