"""Batched Dirac solves split over concurrent streams (DWF_TUNE_STREAMS): every right-hand
side is solved and certified as in one batch; only the batch schedule (chunk lengths,
reliable-update points) differs, so the solutions agree to the tolerance's reach."""

import numpy as np
import pytest

import oracle
import three_part_dwf_study_amd as dwf
from three_part_dwf_study_amd import capi
from conftest import print_case
from dense_ref import rand_spinor, rand_theta

pytestmark = pytest.mark.gpu

M5, MF = 1.0, 0.05


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def normal_relres(th, x, src, Nt, Nx, Ls):
    b = oracle.dwf_apply(1, th, src, Nt, Nx, Ls, M5, MF)
    r = b - oracle.dwf_apply(3, th, x, Nt, Nx, Ls, M5, MF)
    return np.linalg.norm(r) / np.linalg.norm(b)


@pytest.mark.parametrize("method", [dwf.DWF_SOLVE_EO_MIXED, dwf.DWF_SOLVE_MIXED, dwf.DWF_SOLVE_CG64,
                                    dwf.DWF_SOLVE_LEARNED])
@pytest.mark.parametrize("ncfg", [1, 3, 6])
def test_split_solve_matches_one_batch(method, ncfg):
    Nt, Nx, Ls, nrhs = 16, 12, 8, 6
    print_case(test="split_solve", method=method, ncfg=ncfg)
    rng = np.random.default_rng(40 + 7 * ncfg + method)
    ths = [rand_theta(rng, Nt, Nx) for _ in range(ncfg)]
    src = rand_spinor(rng, nrhs * Ls * Nt * Nx * 2).reshape(nrhs, -1)
    net = dwf.Network(Ls, seed=2) if method == dwf.DWF_SOLVE_LEARNED else None
    out = {}
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        if ncfg == 1:
            ctx.set_gauge(ths[0])
        else:
            ctx.set_gauge_batch(np.stack(ths))
        if net is not None:
            ctx.set_preconditioner(net)
        for k in (1, 2, 3, 4):
            ctx.tune(dwf.DWF_TUNE_STREAMS, k)
            out[k] = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=method)
        ctx.set_preconditioner(None)
    if net is not None:
        net.close()
    x1, st1, rc1 = out[1]
    assert rc1 == 0
    for k in (2, 3, 4):
        xk, stk, rck = out[k]
        assert rck == 0
        for r in range(nrhs):
            assert stk[r]["converged"] and stk[r]["true_relres"] <= 1e-10, (k, r, stk[r])
            assert rel(xk[r], x1[r]) <= 1e-7, (k, r)
            # iteration counts of an unsplit right-hand side are unchanged up to the schedule
            assert abs(stk[r]["iterations"] - st1[r]["iterations"]) <= 0.1 * st1[r]["iterations"] + 8
    # certified against the oracle's operator, configuration r // (nrhs // ncfg)
    xs = out[2][0]
    for r in range(nrhs):
        th = ths[r // (nrhs // ncfg)]
        assert normal_relres(th, xs[r], src[r], Nt, Nx, Ls) <= 1.01e-10, r


def test_split_solve_device_buffers_and_small_batches():
    """The device entry point splits in place (sub-ranges of the caller's buffers); a batch
    smaller than twice the streams is solved as one."""
    Nt, Nx, Ls = 8, 8, 4
    print_case(test="split_solve_device")
    rng = np.random.default_rng(77)
    th = rand_theta(rng, Nt, Nx)
    for nrhs in (3, 8):
        src = rand_spinor(rng, nrhs * Ls * Nt * Nx * 2).reshape(nrhs, -1)
        with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
            ctx.set_gauge(th)
            res = []
            for k in (1, 2):
                ctx.tune(dwf.DWF_TUNE_STREAMS, k)
                d_src = capi.DeviceArray.from_host(src.reshape(-1).view(np.float64))
                d_x = capi.DeviceArray(src.size * 16)
                st, rc = ctx.solve_dirac_device(dwf.DWF_SOLVE_EO_MIXED, d_src.ptr, None, 1e-10, 20000, nrhs, d_x.ptr)
                ctx.synchronize()
                res.append((d_x.download(np.float64).view(np.complex128).reshape(nrhs, -1), st, rc))
                d_src.free()
                d_x.free()
        for x, st, rc in res:
            assert rc == 0
            for r in range(nrhs):
                assert st[r]["converged"] and st[r]["true_relres"] <= 1e-10
                assert normal_relres(th, x[r], src[r], Nt, Nx, Ls) <= 1.01e-10
        if nrhs == 3:  # 3 < 2 * 2: one batch either way, bit for bit
            assert np.array_equal(res[0][0], res[1][0])


def test_streams_tune_range():
    with dwf.DwfContext.dwf(8, 8, 4, M5, MF) as ctx:
        for bad in (0, 9, 2.5):
            with pytest.raises(dwf.DwfError):
                ctx.tune(dwf.DWF_TUNE_STREAMS, bad)
        ctx.tune(dwf.DWF_TUNE_STREAMS, 8)
