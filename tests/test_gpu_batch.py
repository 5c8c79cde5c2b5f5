"""Several gauge configurations on one context (dwf_set_gauge_batch): right-hand side r
uses configuration r / (nrhs / ncfg).  Batched applies, solves and analyzer measurements
must agree with the same work done one configuration at a time, and the apply with the
oracle on each configuration."""

import numpy as np
import pytest

import oracle
from conftest import print_case
from dense_ref import rand_spinor, rand_theta

import three_part_dwf_study_amd as dwf
from three_part_dwf_study_amd import capi

pytestmark = pytest.mark.gpu

M5, MF = 1.0, 0.05


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _thetas(rng, ncfg, Nt, Nx):
    return np.stack([rand_theta(rng, Nt, Nx) for _ in range(ncfg)])


@pytest.mark.parametrize("ncfg,per", [(3, 2), (4, 1)])
def test_dwf_apply_batch_per_config(ncfg, per):
    Nt, Nx, Ls = 8, 12, 6
    print_case(test="dwf_apply_batch", ncfg=ncfg, per=per)
    rng = np.random.default_rng(31 + ncfg)
    ths = _thetas(rng, ncfg, Nt, Nx)
    psis = rand_spinor(rng, ncfg * per * Ls * Nt * Nx * 2).reshape(ncfg * per, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths)
        for op in (0, 1, 2, 3):
            got = ctx.apply(op, psis)
            for r in range(ncfg * per):
                want = oracle.dwf_apply(op, ths[r // per], psis[r], Nt, Nx, Ls, M5, MF)
                assert rel(got[r], want) <= 1e-12, (op, r)
        # FP32 device path: same configuration selection
        a = capi.DeviceArray.from_host(psis.astype(np.complex128).view(np.float64).astype(np.float32))
        b = capi.DeviceArray(a.nbytes)
        ctx.apply_device(0, dwf.DWF_PREC_FP32, a.ptr, b.ptr, ncfg * per)
        got32 = b.download(np.float32).astype(np.float64).view(np.complex128).reshape(ncfg * per, -1)
        for r in range(ncfg * per):
            want = oracle.dwf_apply(0, ths[r // per], psis[r], Nt, Nx, Ls, M5, MF)
            assert rel(got32[r], want) <= 2e-6, r
        # back to one configuration
        ctx.set_gauge(ths[1])
        assert rel(ctx.apply(0, psis[0]), oracle.dwf_apply(0, ths[1], psis[0], Nt, Nx, Ls, M5, MF)) <= 1e-12


def test_wilson_apply_batch_per_config():
    Nt, Nx, ncfg = 16, 12, 3
    print_case(test="wilson_apply_batch", ncfg=ncfg)
    rng = np.random.default_rng(5)
    ths = _thetas(rng, ncfg, Nt, Nx)
    psis = rand_spinor(rng, ncfg * Nt * Nx * 2).reshape(ncfg, -1)
    with dwf.DwfContext.wilson(Nt, Nx, 0.1, -1.0, 1.0) as ctx:
        ctx.set_gauge_batch(ths)
        got = ctx.apply(0, psis)
        for r in range(ncfg):
            want = oracle.wilson_apply(0, ths[r], psis[r], Nt, Nx, 0.1, -1.0, 1.0)
            assert rel(got[r], want) <= 1e-12, r


@pytest.mark.parametrize("method", [dwf.DWF_SOLVE_CG64, dwf.DWF_SOLVE_MIXED, dwf.DWF_SOLVE_EO_MIXED])
def test_solve_dirac_batch_matches_single(method):
    Nt, Nx, Ls, ncfg, per = 16, 12, 8, 3, 2
    print_case(test="solve_dirac_batch", method=method)
    rng = np.random.default_rng(77)
    ths = _thetas(rng, ncfg, Nt, Nx)
    src = rand_spinor(rng, ncfg * per * Ls * Nt * Nx * 2).reshape(ncfg * per, -1)
    rtol = 1e-10
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths)
        xb, stb, _ = ctx.solve_dirac(src, rtol=rtol, maxiter=20000, method=method)
        single = []
        for k in range(ncfg):
            ctx.set_gauge(ths[k])
            xs, sts, _ = ctx.solve_dirac(src[k * per:(k + 1) * per], rtol=rtol, maxiter=20000, method=method)
            single.append(xs)
    for r in range(ncfg * per):
        k = r // per
        assert stb[r]["converged"], r
        Dx = oracle.dwf_apply(0, ths[k], xb[r], Nt, Nx, Ls, M5, MF)
        lhs = oracle.dwf_apply(1, ths[k], src[r] - Dx, Nt, Nx, Ls, M5, MF)
        rhs = oracle.dwf_apply(1, ths[k], src[r], Nt, Nx, Ls, M5, MF)
        assert np.linalg.norm(lhs) / np.linalg.norm(rhs) <= rtol * 1.01, r
        assert rel(xb[r], single[k][r % per]) <= 1e-8, r


def test_measure_batch_matches_single():
    Nt, Nx, Ls, ncfg, nn, mm = 8, 6, 4, 3, 2, 3
    print_case(test="measure_batch", ncfg=ncfg)
    rng = np.random.default_rng(404)
    ths = _thetas(rng, ncfg, Nt, Nx)
    seeds = [918273 + 104729 * k for k in range(ncfg)]
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths)
        mb = ctx.measure_batch(ncfg, 1e-12, 20000, max_momentum=mm, noise_vectors=nn, seeds=seeds)
        with pytest.raises(dwf.DwfError, match="use dwf_measure_batch"):
            ctx.measure(1e-12, 20000)
        for k in range(ncfg):
            ctx.set_gauge(ths[k])
            m = ctx.measure(1e-12, 20000, max_momentum=mm, noise_vectors=nn, noise_seed=seeds[k])
            for key in ("CPP", "CJ5", "momentum_CPP", "loop", "dirac_residual"):
                scale = np.max(np.abs(m[key]))
                assert np.allclose(mb[k][key], m[key], rtol=1e-9, atol=1e-12 * scale), (k, key)
            assert mb[k]["wall_nonconverged"] == 0
            cpp, cj5, _, _, _ = oracle.dwf_cpp_cj5(ths[k], Nt, Nx, Ls, M5, MF, 1e-12, 20000)
            assert np.allclose(mb[k]["CPP"], cpp, rtol=1e-9, atol=0)


def test_batch_argument_errors():
    Nt, Nx, Ls = 8, 6, 4
    print_case(test="batch_errors")
    rng = np.random.default_rng(2)
    ths = _thetas(rng, 2, Nt, Nx)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths)
        with pytest.raises(dwf.DwfError, match="multiple"):
            ctx.apply(0, rand_spinor(rng, 3 * Ls * Nt * Nx * 2).reshape(3, -1))
        with pytest.raises(dwf.DwfError, match="one configuration"):
            ctx.average_plaquette()
        with pytest.raises(dwf.DwfError, match="ncfg differs"):
            ctx.measure_batch(3, 1e-8, 100)


def test_point_site_sources():
    """Point sources at chosen sites (the batched-analyzer workload).  Checks: the two
    single-time-slice results at t0 = 0 and Nt/2 average to the reference's folded
    momentum correlator (pinned by the oracle); translating the gauge field in x by x0
    and moving the source with it leaves every correlator unchanged (x is periodic)."""
    Nt, Nx, Ls, mm, x0 = 8, 6, 4, 3, 2
    print_case(test="point_site", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(9)
    th = rand_theta(rng, Nt, Nx)
    th_shift = np.roll(th.reshape(2, Nt, Nx), x0, axis=2).reshape(-1)
    rtol, maxit = 1e-12, 20000
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(np.stack([th, th, th_shift]))
        m = ctx.measure_batch(3, rtol, maxit, wall=False, point=True, max_momentum=mm,
                              point_sites=[[0, 0], [Nt // 2, 0], [Nt // 2, x0]])
        with pytest.raises(dwf.DwfError, match="outside the lattice"):
            ctx.measure_batch(3, rtol, maxit, wall=False, point=True, point_sites=[[0, 0], [Nt, 0], [0, 0]])
        ctx.set_gauge(th)
        ref = ctx.measure(rtol, maxit, wall=False, point=True, max_momentum=mm)
    want = oracle.dwf_momentum_cpp(th, Nt, Nx, Ls, M5, MF, mm, rtol, maxit)
    assert np.allclose(ref["momentum_CPP"], want, rtol=1e-8, atol=1e-11 * np.max(np.abs(want)))
    avg = (m[0]["momentum_CPP"] + m[1]["momentum_CPP"]) / 2
    assert np.allclose(avg, want, rtol=1e-8, atol=1e-11 * np.max(np.abs(want)))
    pj = (m[0]["point_CJ5"] + m[1]["point_CJ5"]) / 2
    assert np.allclose(pj, ref["point_CJ5"], rtol=1e-9, atol=1e-12 * np.max(np.abs(ref["point_CJ5"])))
    for key in ("momentum_CPP", "point_CJ5"):
        scale = np.max(np.abs(m[1][key]))
        assert np.allclose(m[2][key], m[1][key], rtol=1e-8, atol=1e-10 * scale), key
