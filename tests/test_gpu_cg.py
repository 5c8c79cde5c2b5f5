"""GPU CG parity (through the C ABI) against the oracle's restatement of
domain_wall/include/cg_solver.hpp:17-64, and certification of the mixed-precision
solver by an independent FP64 true residual computed with the oracle's operator."""

import numpy as np
import pytest

import oracle
from conftest import print_case
from dense_ref import rand_spinor, rand_theta

import three_part_dwf_study_amd as dwf
from three_part_dwf_study_amd import capi

pytestmark = pytest.mark.gpu

M5, MF = 1.0, 0.05


def true_relres(th, x, b, normal, Nt, Nx, Ls, mf=MF):
    Ax = oracle.dwf_apply(normal, th, x, Nt, Nx, Ls, M5, mf)
    return np.linalg.norm(b - Ax) / np.linalg.norm(b)


@pytest.mark.parametrize("normal", [dwf.DWF_OP_DDDAG, dwf.DWF_OP_DDAGD])
def test_cg64_matches_oracle(normal):
    Nt, Nx, Ls = 8, 6, 4
    print_case(test="cg64_vs_oracle", normal=normal)
    rng = np.random.default_rng(5)
    th = rand_theta(rng, Nt, Nx)
    b = rand_spinor(rng, Ls * Nt * Nx * 2)
    xo, it_o, rr_o, conv_o, rc = oracle.dwf_cg(normal, th, b, Nt, Nx, Ls, M5, MF, 1e-10, 5000)
    assert conv_o
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        x, st, rc = ctx.solve(normal, b, rtol=1e-10, maxiter=5000)
    assert rc == 0 and st[0]["converged"]
    assert abs(st[0]["iterations"] - it_o) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    assert st[0]["relres"] <= 1e-10
    assert abs(st[0]["true_relres"] - true_relres(th, x, b, normal, Nt, Nx, Ls)) < 1e-12


def test_cg64_guess_and_special_cases():
    Nt, Nx, Ls = 8, 6, 4
    print_case(test="cg64_special")
    rng = np.random.default_rng(6)
    th = rand_theta(rng, Nt, Nx)
    n = Ls * Nt * Nx * 2
    b = rand_spinor(rng, n)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        x, st, rc = ctx.solve(3, np.zeros(n, complex), rtol=1e-8, maxiter=100)
        assert rc == 0 and st[0]["iterations"] == 0 and not np.any(x)
        xs, st, rc = ctx.solve(3, b, rtol=1e-12, maxiter=5000)
        x, st, rc = ctx.solve(3, b, x0=xs, rtol=1e-8, maxiter=100)
        assert st[0]["iterations"] == 0 and st[0]["converged"]
        # with a guess, CG agrees with the oracle's guessed solve
        g = xs + 1e-3 * rand_spinor(rng, n)
        x, st, rc = ctx.solve(3, b, x0=g, rtol=1e-10, maxiter=5000)
        xo, it_o, *_ = oracle.dwf_cg(3, th, b, Nt, Nx, Ls, M5, MF, 1e-10, 5000, guess=g)
        assert abs(st[0]["iterations"] - it_o) <= 1
        x, st, rc = ctx.solve(3, b, rtol=1e-14, maxiter=3, raise_on_fail=False)
        assert rc == capi.DWF_ERR_NOT_CONVERGED and st[0]["iterations"] == 3 and not st[0]["converged"]


@pytest.mark.parametrize("Nt,Nx,Ls,nrhs", [(16, 16, 8, 1), (32, 32, 8, 3), (96, 96, 16, 1)])
def test_mixed_certified(Nt, Nx, Ls, nrhs):
    print_case(test="mixed_certified", Nt=Nt, Nx=Nx, Ls=Ls, nrhs=nrhs)
    rng = np.random.default_rng(Nt + nrhs)
    th = rand_theta(rng, Nt, Nx)
    b = rand_spinor(rng, nrhs * Ls * Nt * Nx * 2).reshape(nrhs, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        x, st, rc = ctx.solve(3, b, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_MIXED)
        x64, st64, rc64 = ctx.solve(3, b, rtol=1e-10, maxiter=20000)
    assert rc == 0
    for i in range(nrhs):
        assert st[i]["converged"] and st[i]["reliable_updates"] >= 1
        tr = true_relres(th, x[i], b[i], 3, Nt, Nx, Ls)
        assert tr <= 1e-10, tr
        assert abs(tr - st[i]["true_relres"]) < 1e-12
        # same solution as FP64 CG to the accuracy the tolerance allows
        assert np.linalg.norm(x[i] - x64[i]) / np.linalg.norm(x64[i]) < 1e-6


def test_wilson_cgnr_matches_oracle():
    Nt, Nx, m0 = 16, 16, 0.1
    print_case(test="wilson_cgnr")
    rng = np.random.default_rng(9)
    th = rng.uniform(-np.pi, np.pi, size=2 * Nt * Nx)
    src = np.zeros(Nt * Nx * 2, complex)
    src[0] = 1.0
    xo, it, nres, dres, rc = oracle.wilson_solve_cgnr(th, src, Nt, Nx, m0, -1.0, 1.0, 1e-12, 5000)
    assert rc == 0
    with dwf.DwfContext.wilson(Nt, Nx, m0, -1.0, 1.0) as ctx:
        ctx.set_gauge(th)
        rhs = ctx.apply(dwf.DWF_OP_DDAG, src)
        x, st, rc = ctx.solve(dwf.DWF_OP_DDAGD, rhs, rtol=1e-12, maxiter=5000)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-9


@pytest.mark.parametrize("method", [dwf.DWF_SOLVE_CG64, dwf.DWF_SOLVE_MIXED, dwf.DWF_SOLVE_EO_MIXED])
def test_graph_replay_matches_eager(method):
    """CG iteration blocks replayed as HIP graphs give the eager launches' results bit
    for bit, over repeated calls (cached graphs) and a change of batch size."""
    Nt, Nx, Ls = 16, 12, 8
    print_case(test="graph_replay", method=method)
    rng = np.random.default_rng(123)
    th = rand_theta(rng, Nt, Nx)
    src = rand_spinor(rng, 3 * Ls * Nt * Nx * 2).reshape(3, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_GRAPHS, 0)
        eager3, st3, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=method)
        eager1, st1, _ = ctx.solve_dirac(src[:1], rtol=1e-10, maxiter=20000, method=method)
        ctx.tune(dwf.DWF_TUNE_GRAPHS, 1)
        for rep in range(3):
            g3, gs3, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=method)
            assert np.array_equal(g3, eager3), rep
            assert [s["iterations"] for s in gs3] == [s["iterations"] for s in st3]
            g1, gs1, _ = ctx.solve_dirac(src[:1], rtol=1e-10, maxiter=20000, method=method)
            assert np.array_equal(g1, eager1), rep
            assert gs1[0]["iterations"] == st1[0]["iterations"]


def test_device_solve_rejects_overlapping_operands():
    """The device solve entry points refuse an x that shares memory with the right-hand
    side, or a guess that partly overlaps x, before any device work; a guess passed as x
    itself stays allowed.  Every pointer tried lies inside one buffer three vectors long,
    so no candidate reaches outside it."""
    Nt, Nx, Ls = 8, 6, 4
    print_case(test="device_solve_overlap")
    rng = np.random.default_rng(8)
    th = rand_theta(rng, Nt, Nx)
    n = Ls * Nt * Nx * 2  # complex values of one vector
    b = rand_spinor(rng, n)
    host = np.zeros(3 * n, complex)
    host[n:2 * n] = b
    vb = n * 16  # bytes of one vector
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        big = capi.DeviceArray.from_host(host.view(np.float64))
        pb, px = big.ptr + vb, big.ptr  # b in the middle third, x in the first
        calls = (
            lambda x, x0: ctx.solve_dirac_device(dwf.DWF_SOLVE_MIXED, pb, x0, 1e-10, 5000, 1, x),
            lambda x, x0: ctx.solve_device(dwf.DWF_OP_DDDAG, dwf.DWF_SOLVE_CG64, pb, x0, 1e-10, 5000, 1, x),
        )
        for call in calls:
            for x, x0 in ((pb, None), (pb + 64, None), (px + 64, None), (px, px + 64)):
                with pytest.raises(capi.DwfError) as e:
                    call(x, x0)
                assert e.value.code == capi.DWF_ERR_ARG and "overlap" in str(e.value)
            st, rc = call(px, None)
            assert rc == 0 and st[0]["converged"]
            st2, rc = call(px, px)  # the previous solution as the guess, in place
            assert rc == 0 and st2[0]["converged"] and st2[0]["iterations"] < st[0]["iterations"] // 4
        assert np.array_equal(big.download(np.complex128)[n:2 * n], b)  # b never written
        big.free()
