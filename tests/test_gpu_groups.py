"""Configurations of several lattice sizes on one context (dwf_set_gauge_groups): every
kernel launch of an apply, a solve or an analyzer pass covers all the sizes at once (the
batched analyzer over uneven Nx).  Each right-hand side must match the oracle on its own
lattice, and the same work done one lattice size at a time."""

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


def _groups(rng, spec):
    """spec: [(Nt, Nx, ncfg)] -> [(Nt, Nx, thetas)] with seeded links."""
    return [(Nt, Nx, np.stack([rand_theta(rng, Nt, Nx) for _ in range(n)])) for Nt, Nx, n in spec]


def _split(v, groups, Ls, per):
    """Concatenated batch -> list over right-hand sides (configuration-major)."""
    out, o = [], 0
    for Nt, Nx, th in groups:
        n = Ls * Nt * Nx * 2
        for _ in range(len(th) * per):
            out.append(v[o:o + n])
            o += n
    assert o == v.size
    return out


def _rhs_lattices(groups, per):
    return [(Nt, Nx, th[k]) for Nt, Nx, th in groups for k in range(len(th)) for _ in range(per)]


def test_group_apply_matches_oracle():
    Ls, per = 4, 2
    rng = np.random.default_rng(61)
    groups = _groups(rng, [(6, 6, 2), (8, 8, 1), (12, 10, 2)])
    lats = _rhs_lattices(groups, per)
    psi = np.concatenate([rand_spinor(rng, Ls * Nt * Nx * 2) for Nt, Nx, _ in lats])
    with dwf.DwfContext.dwf(8, 8, Ls, M5, MF) as ctx:
        ctx.set_gauge_groups(groups)
        for op in (dwf.DWF_OP_D, dwf.DWF_OP_DDAG, dwf.DWF_OP_DDDAG, dwf.DWF_OP_DDAGD):
            print_case(test="group_apply", op=op)
            got = _split(ctx.apply(op, psi), groups, Ls, per)
            for r, ((Nt, Nx, th), v) in enumerate(zip(lats, _split(psi, groups, Ls, per))):
                want = oracle.dwf_apply(op, th, v, Nt, Nx, Ls, M5, MF)
                assert rel(got[r], want) <= 1e-12, (op, r)
        # FP32 device path, D
        print_case(test="group_apply_fp32")
        a = capi.DeviceArray.from_host(psi.view(np.float64).astype(np.float32))
        b = capi.DeviceArray(a.nbytes)
        ctx.apply_device(dwf.DWF_OP_D, dwf.DWF_PREC_FP32, a.ptr, b.ptr, len(lats))
        got32 = b.download(np.float32).astype(np.float64).view(np.complex128)
        for r, ((Nt, Nx, th), v, g) in enumerate(zip(lats, _split(psi, groups, Ls, per), _split(got32, groups, Ls, per))):
            assert rel(g, oracle.dwf_apply(0, th, v, Nt, Nx, Ls, M5, MF)) <= 2e-6, r
        a.free()
        b.free()
        # the context's own lattice again
        th8 = groups[1][2][0]
        ctx.set_gauge(th8)
        v = rand_spinor(rng, Ls * 8 * 8 * 2)
        assert rel(ctx.apply(0, v), oracle.dwf_apply(0, th8, v, 8, 8, Ls, M5, MF)) <= 1e-12


@pytest.mark.parametrize("method", [dwf.DWF_SOLVE_CG64, dwf.DWF_SOLVE_MIXED, dwf.DWF_SOLVE_EO_MIXED])
def test_group_solve_dirac_matches_single(method):
    Ls, per, rtol = 8, 2, 1e-10
    rng = np.random.default_rng(83)
    groups = _groups(rng, [(8, 6, 2), (12, 12, 1), (16, 10, 1)])
    lats = _rhs_lattices(groups, per)
    src = np.concatenate([rand_spinor(rng, Ls * Nt * Nx * 2) for Nt, Nx, _ in lats])
    with dwf.DwfContext.dwf(8, 6, Ls, M5, MF) as ctx:
        print_case(test="group_solve_dirac", method=method)
        ctx.set_gauge_groups(groups)
        xg, stg, _ = ctx.solve_dirac(src, rtol=rtol, maxiter=20000, method=method)
        xs = _split(xg, groups, Ls, per)
        srcs = _split(src, groups, Ls, per)
        single, r = [], 0
        for Nt, Nx, th in groups:
            with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as one:
                one.set_gauge_batch(th)
                n = len(th) * per
                x1, _, _ = one.solve_dirac(np.concatenate(srcs[r:r + n]), rtol=rtol, maxiter=20000, method=method)
                single += list(x1.reshape(n, -1))
                r += n
    for r, (Nt, Nx, th) in enumerate(lats):
        assert stg[r]["converged"], r
        Dx = oracle.dwf_apply(0, th, xs[r], Nt, Nx, Ls, M5, MF)
        lhs = oracle.dwf_apply(1, th, srcs[r] - Dx, Nt, Nx, Ls, M5, MF)
        rhs = oracle.dwf_apply(1, th, srcs[r], Nt, Nx, Ls, M5, MF)
        assert np.linalg.norm(lhs) / np.linalg.norm(rhs) <= rtol * 1.01, r
        assert rel(xs[r], single[r]) <= 1e-8, r


def test_group_measure_matches_oracle():
    """The cfg3 source shape at small sizes: 2 point sources at a chosen site plus Z4
    noise vectors per configuration, three lattice sizes in one batch.  Each size holds a
    configuration twice with point sites (0, 0) and (Nt/2, 0), whose average is the
    reference's folded momentum correlator (the oracle's); the loop density is the
    oracle's for the same Z4 seed; a third site (t0, x0) is checked against a context
    holding that lattice alone."""
    Ls, nn, mm, rtol = 4, 4, 2, 1e-12
    rng = np.random.default_rng(97)
    sizes = [(6, 6), (8, 8), (12, 12)]
    groups, sites, seeds, i = [], [], [], 0
    for Nt, Nx in sizes:
        th = rand_theta(rng, Nt, Nx)
        th2 = rand_theta(rng, Nt, Nx)
        groups.append((Nt, Nx, np.stack([th, th, th2])))
        sites += [[0, 0], [Nt // 2, 0], [Nt - 1, Nx // 2 + 1]]
        seeds += [918273 + 104729 * (i + k) for k in range(3)]
        i += 3
    ncfg = len(seeds)
    with dwf.DwfContext.dwf(6, 6, Ls, M5, MF) as ctx:
        print_case(test="group_measure", sizes=sizes)
        ctx.set_gauge_groups(groups)
        m = ctx.measure_batch(ncfg, rtol, 20000, wall=False, point=True, max_momentum=mm, noise_vectors=nn,
                              seeds=seeds, method=dwf.DWF_SOLVE_EO_MIXED, point_sites=sites)
        # wall sources (C_PP, C_J5) over the same groups
        mw = ctx.measure_batch(ncfg, rtol, 20000, wall=True, point=False, method=dwf.DWF_SOLVE_MIXED)
    for g, (Nt, Nx, th) in enumerate(groups):
        k = 3 * g
        print_case(test="group_measure_oracle", Nt=Nt, Nx=Nx)
        want = oracle.dwf_momentum_cpp(th[0], Nt, Nx, Ls, M5, MF, mm, rtol, 20000)
        avg = (m[k]["momentum_CPP"] + m[k + 1]["momentum_CPP"]) / 2
        assert np.allclose(avg, want, rtol=1e-8, atol=1e-11 * np.max(np.abs(want))), (Nt, Nx)
        for j in range(3):
            loop = oracle.dwf_loop_density(th[j], Nt, Nx, Ls, M5, MF, seeds[k + j], nn, rtol, 20000)
            assert np.allclose(m[k + j]["loop"].reshape(-1), np.asarray(loop).reshape(-1), rtol=1e-8,
                               atol=1e-10 * np.max(np.abs(loop))), (Nt, Nx, j)
            cpp, cj5, _, _, _ = oracle.dwf_cpp_cj5(th[j], Nt, Nx, Ls, M5, MF, rtol, 20000)
            assert np.allclose(mw[k + j]["CPP"], cpp, rtol=1e-8, atol=0), (Nt, Nx, j)
            assert np.allclose(mw[k + j]["CJ5"], cj5, rtol=1e-7, atol=1e-10 * np.max(np.abs(cj5))), (Nt, Nx, j)
            assert mw[k + j]["wall_nonconverged"] == 0
        # the third configuration against this lattice size alone
        with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as one:
            one.set_gauge_batch(th[2:3])
            s1 = one.measure_batch(1, rtol, 20000, wall=False, point=True, max_momentum=mm, noise_vectors=nn,
                                   seeds=seeds[k + 2:k + 3], method=dwf.DWF_SOLVE_EO_MIXED,
                                   point_sites=[sites[k + 2]])
        for key in ("momentum_CPP", "point_CJ5", "loop"):
            scale = np.max(np.abs(s1[0][key]))
            assert np.allclose(m[k + 2][key], s1[0][key], rtol=1e-9, atol=1e-11 * scale), (Nt, Nx, key)


def test_group_argument_errors():
    Ls = 4
    rng = np.random.default_rng(3)
    groups = _groups(rng, [(6, 6, 1), (8, 8, 1)])
    with dwf.DwfContext.dwf(6, 6, Ls, M5, MF) as ctx:
        print_case(test="group_errors")
        ctx.set_gauge_groups(groups)
        with pytest.raises(dwf.DwfError, match="single lattice|one configuration"):
            ctx.average_plaquette()
        with pytest.raises(dwf.DwfError, match="single lattice"):
            ctx.autotune(1)
        net = dwf.Network(Ls, seed=1)
        ctx.set_preconditioner(net)
        src = np.concatenate([rand_spinor(rng, Ls * 36 * 2), rand_spinor(rng, Ls * 64 * 2)])
        with pytest.raises(dwf.DwfError, match="single lattice"):
            ctx.solve_dirac(src, method=dwf.DWF_SOLVE_LEARNED)
        ctx.set_preconditioner(None)
        net.close()
        with pytest.raises(ValueError):
            ctx.apply(0, src[:-2])
    with dwf.DwfContext.dwf(6, 6, Ls, M5, MF) as ctx:
        bad = [(6, 6, np.zeros((1, 72)))] * 9
        with pytest.raises(dwf.DwfError, match="lattice groups"):
            ctx.set_gauge_groups(bad)


def test_group_analyzer_full_sizes():
    """The batched-analyzer workload's shape at its real sizes: one configuration at each of
    Nx = Nt in {32, 56, 80, 96}, Ls = 16, 2 point sources + 32 Z4 vectors each, all in
    one batch (every launch covers the four sizes), solved by the even-odd mixed solver
    and certified on the FP64 true residual at 1e-10 (a solve that misses it raises).
    The correlators must equal those of each size measured alone."""
    Ls, nn, rtol = 16, 32, 1e-10
    sizes = [32, 56, 80, 96]
    groups, sites, seeds = [], [], []
    for i, N in enumerate(sizes):
        th = np.random.default_rng(1234 + 104729 * i).normal(0.0, 0.5, size=2 * N * N)
        groups.append((N, N, th[None, :]))
        sites.append([int(v) for v in np.random.default_rng(555 + i).integers(0, N, size=2)])
        seeds.append(918273 + 104729 * i)
    with dwf.DwfContext.dwf(96, 96, Ls, M5, MF) as ctx:
        print_case(test="group_analyzer_full", sizes=sizes)
        ctx.set_gauge_groups(groups)
        m = ctx.measure_batch(len(sizes), rtol, 20000, wall=False, point=True, max_momentum=0, noise_vectors=nn,
                              seeds=seeds, method=dwf.DWF_SOLVE_EO_MIXED, point_sites=sites)
    for i, N in enumerate(sizes):
        print_case(test="group_analyzer_full_single", N=N)
        with dwf.DwfContext.dwf(N, N, Ls, M5, MF) as one:
            one.set_gauge_batch(groups[i][2])
            s1 = one.measure_batch(1, rtol, 20000, wall=False, point=True, max_momentum=0, noise_vectors=nn,
                                   seeds=seeds[i:i + 1], method=dwf.DWF_SOLVE_EO_MIXED, point_sites=sites[i:i + 1])
        assert len(m[i]["iterations"]) == 2 + nn
        for key in ("momentum_CPP", "point_CJ5", "loop"):
            scale = np.max(np.abs(s1[0][key]))
            # both certified at 1e-10: solutions agree to about that, so do the contractions
            assert np.allclose(m[i][key], s1[0][key], rtol=1e-7, atol=1e-9 * scale), (N, key)
