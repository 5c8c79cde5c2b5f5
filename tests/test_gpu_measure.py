"""GPU parity of the analyzer measurements (correlator.hpp:27-310) through dwf_measure
against the oracle, on the same configuration and Z4 noise seed."""

import numpy as np
import pytest

import oracle
from conftest import print_case
from dense_ref import rand_theta

import three_part_dwf_study_amd as dwf

pytestmark = pytest.mark.gpu

M5, MF = 1.0, 0.05


@pytest.mark.parametrize("Nt,Nx,Ls", [(8, 6, 4), (16, 12, 8)])
@pytest.mark.parametrize("method", [dwf.DWF_SOLVE_CG64, dwf.DWF_SOLVE_MIXED])
def test_measure_matches_oracle(Nt, Nx, Ls, method):
    print_case(test="measure", Nt=Nt, Nx=Nx, Ls=Ls, method=method)
    rng = np.random.default_rng(Nt * Ls)
    th = rand_theta(rng, Nt, Nx)
    rtol, maxit, nn, seed, mm = 1e-12, 20000, 3, 918273 + 104729 * 5, Nx // 2
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        m = ctx.measure(rtol, maxit, wall=True, point=True, max_momentum=mm, noise_vectors=nn, noise_seed=seed,
                        method=method)
    cpp, cj5, nc, iters, dres = oracle.dwf_cpp_cj5(th, Nt, Nx, Ls, M5, MF, rtol, maxit)
    mom = oracle.dwf_momentum_cpp(th, Nt, Nx, Ls, M5, MF, mm, rtol, maxit)
    loop = oracle.dwf_loop_density(th, Nt, Nx, Ls, M5, MF, seed, nn, rtol, maxit)
    assert m["wall_nonconverged"] == 0
    assert np.allclose(m["CPP"], cpp, rtol=1e-9, atol=0)
    assert np.allclose(m["CJ5"], cj5, rtol=1e-8, atol=1e-12 * np.max(np.abs(cpp)))
    assert np.allclose(m["momentum_CPP"], mom, rtol=1e-8, atol=1e-11 * np.max(np.abs(mom)))
    assert np.allclose(m["loop"].reshape(-1), loop, rtol=1e-8, atol=1e-11)
    assert np.all(m["dirac_residual"] < 1e-9)
    if method == dwf.DWF_SOLVE_CG64:
        assert np.all(np.abs(m["iterations"][:4] - iters) <= 1)


def test_measure_errors():
    Nt, Nx, Ls = 8, 6, 4
    print_case(test="measure_errors")
    rng = np.random.default_rng(1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(rand_theta(rng, Nt, Nx))
        with pytest.raises(dwf.DwfError, match="max_momentum is outside"):
            ctx.measure(1e-8, 100, wall=False, point=True, max_momentum=Nx // 2 + 1)
        with pytest.raises(dwf.DwfError, match="Momentum propagator CGNR failed to converge"):
            ctx.measure(1e-14, 2, wall=False, point=True, max_momentum=1)
        with pytest.raises(dwf.DwfError, match="Eta loop CGNR failed to converge"):
            ctx.measure(1e-14, 2, wall=False, point=False, noise_vectors=2, noise_seed=5)
        m = ctx.measure(1e-14, 2, wall=True, point=False)
        assert m["wall_nonconverged"] == 4
