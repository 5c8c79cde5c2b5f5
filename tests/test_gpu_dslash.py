"""GPU parity of the Dirac operators (through the C ABI) against the oracle.

Bar: FP64 <= 1e-12 relative (the north-star Dirac-apply tolerance), FP32 <= 2e-6
relative (single-precision rounding of O(10) terms per component)."""

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


@pytest.mark.parametrize("Nt,Nx,Ls", [(16, 16, 8), (6, 10, 4), (4, 4, 1), (96, 96, 16)])
def test_dwf_ops_fp64(Nt, Nx, Ls):
    print_case(test="dwf_ops_fp64", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(Nt * 1000 + Nx * 10 + Ls)
    th = rand_theta(rng, Nt, Nx, 0.5)
    psi = rand_spinor(rng, Ls * Nt * Nx * 2)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        for op in (0, 1, 2, 3):
            got = ctx.apply(op, psi)
            want = oracle.dwf_apply(op, th, psi, Nt, Nx, Ls, M5, MF)
            assert rel(got, want) <= 1e-12, (op, rel(got, want))


def test_dwf_batch_matches_single():
    Nt, Nx, Ls, n = 8, 12, 6, 3
    print_case(test="dwf_batch", Nt=Nt, Nx=Nx, Ls=Ls, nrhs=n)
    rng = np.random.default_rng(7)
    th = rand_theta(rng, Nt, Nx)
    psis = rand_spinor(rng, n * Ls * Nt * Nx * 2).reshape(n, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        batch = ctx.apply(0, psis)
        for i in range(n):
            assert rel(batch[i], oracle.dwf_apply(0, th, psis[i], Nt, Nx, Ls, M5, MF)) <= 1e-12


@pytest.mark.parametrize("Nt,Nx,bct,bcx", [(16, 16, -1.0, 1.0), (5, 7, 1.0, -1.0), (48, 128, -1.0, 1.0)])
def test_wilson_ops_fp64(Nt, Nx, bct, bcx):
    print_case(test="wilson_ops_fp64", Nt=Nt, Nx=Nx)
    rng = np.random.default_rng(11)
    th = rng.uniform(-np.pi, np.pi, size=2 * Nt * Nx)
    psi = rand_spinor(rng, Nt * Nx * 2)
    m0 = 0.15
    with dwf.DwfContext.wilson(Nt, Nx, m0, bct, bcx) as ctx:
        ctx.set_gauge(th)
        for op in (0, 1):
            got = ctx.apply(op, psi)
            want = oracle.wilson_apply(op, th, psi, Nt, Nx, m0, bct, bcx)
            assert rel(got, want) <= 1e-12


@pytest.mark.parametrize("Nt,Nx,Ls", [(16, 16, 8), (96, 96, 16)])
def test_dwf_ops_fp32_device(Nt, Nx, Ls):
    print_case(test="dwf_ops_fp32", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(3)
    th = rand_theta(rng, Nt, Nx)
    psi = rand_spinor(rng, Ls * Nt * Nx * 2)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        for prec, dt, tol in ((dwf.DWF_PREC_FP32, np.float32, 2e-6), (dwf.DWF_PREC_FP64, np.float64, 1e-12)):
            host = psi.view(np.float64).astype(dt)
            src = capi.DeviceArray.from_host(host)
            dst = capi.DeviceArray(host.nbytes)
            for op in (0, 1, 3):
                ctx.apply_device(op, prec, src.ptr, dst.ptr, 1)
                ctx.synchronize()
                got = dst.download(dt).astype(np.float64).view(np.complex128)
                want = oracle.dwf_apply(op, th, psi, Nt, Nx, Ls, M5, MF)
                assert rel(got, want) <= tol, (prec, op, rel(got, want))


def test_copy_kernel_and_timer():
    print_case(test="copy_kernel")
    n = 1 << 20
    a = np.arange(n, dtype=np.float32)
    with dwf.DwfContext.dwf(4, 4, 4, M5, MF) as ctx:
        src = capi.DeviceArray.from_host(a)
        dst = capi.DeviceArray(a.nbytes)
        t = capi.Timer()
        t.start(ctx)
        ctx.copy_device(dst.ptr, src.ptr, a.nbytes)
        t.stop(ctx)
        assert t.elapsed_ms() >= 0.0
        assert np.array_equal(dst.download(np.float32), a)


@pytest.mark.parametrize("tile_x,prefetch,chunks", [(0, 1, 0), (32, 0, 0), (32, 1, 2), (16, 1, 3), (48, 0, 16)])
def test_tuned_launch_shapes_keep_parity(tile_x, prefetch, chunks):
    Nt, Nx, Ls = 12, 96, 6
    print_case(test="tuned_shapes", tile_x=tile_x, prefetch=prefetch, chunks=chunks)
    rng = np.random.default_rng(tile_x + prefetch + chunks)
    th = rand_theta(rng, Nt, Nx)
    psi = rand_spinor(rng, 2 * Ls * Nt * Nx * 2).reshape(2, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_TILE_X, tile_x)
        ctx.tune(dwf.DWF_TUNE_PREFETCH, prefetch)
        ctx.tune(dwf.DWF_TUNE_NT_STORE, prefetch)
        ctx.tune(dwf.DWF_TUNE_SCHUNKS, min(chunks, Ls))
        for op in (0, 1, 3):
            got = ctx.apply(op, psi)
            for i in range(2):
                want = oracle.dwf_apply(op, th, psi[i], Nt, Nx, Ls, M5, MF)
                assert rel(got[i], want) <= 1e-12, (op, rel(got[i], want))
        b = ctx.apply(1, psi[0])
        x, st, rc = ctx.solve(3, b, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_MIXED)
        Ax = oracle.dwf_apply(3, th, x, Nt, Nx, Ls, M5, MF)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= 1e-10


@pytest.mark.parametrize("tile_x,chunks,mode", [(8, 0, 1), (16, 2, 1), (32, 6, 1), (64, 1, 1), (32, 1, 2),
                                                (16, 3, 2), (64, 2, 2), (16, 0, 3), (8, 2, 3), (32, 0, 4),
                                                (16, 1, 4)])
def test_lds_tiles_keep_parity(tile_x, chunks, mode):
    """The LDS-staged stencil (full 2D tiles, one halo per slice) against the oracle for
    every operator, and inside the FP64 and mixed solvers (its NORM/CGUPD/RESID paths)."""
    lds_parity(32, 64, 6, tile_x, chunks, mode)


@pytest.mark.parametrize("tile_x,chunks,mode", [(96, 0, 5), (96, 2, 6), (32, 1, 5), (48, 0, 5)])
def test_lds_full_row_tiles_keep_parity(tile_x, chunks, mode):
    """Tiles of 384 / 768 threads, among them tiles as wide as the lattice (96): the x
    halo is the tile's own wrapped edge columns."""
    lds_parity(24, 96, 4, tile_x, chunks, mode)


def lds_parity(Nt, Nx, Ls, tile_x, chunks, mode):
    print_case(test="lds_tiles", Nt=Nt, Nx=Nx, Ls=Ls, tile_x=tile_x, chunks=chunks, mode=mode)
    rng = np.random.default_rng(100 + tile_x + chunks + mode)
    th = rand_theta(rng, Nt, Nx)
    psi = rand_spinor(rng, 2 * Ls * Nt * Nx * 2).reshape(2, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_LDS, mode)
        ctx.tune(dwf.DWF_TUNE_TILE_X, tile_x)
        ctx.tune(dwf.DWF_TUNE_NT_STORE, tile_x % 16 == 0)
        ctx.tune(dwf.DWF_TUNE_SCHUNKS, chunks)
        for op in (0, 1, 2, 3):
            got = ctx.apply(op, psi)
            for i in range(2):
                want = oracle.dwf_apply(op, th, psi[i], Nt, Nx, Ls, M5, MF)
                assert rel(got[i], want) <= 1e-12, (op, rel(got[i], want))
        b = ctx.apply(1, psi[0])
        for method in (dwf.DWF_SOLVE_CG64, dwf.DWF_SOLVE_MIXED):
            x, st, rc = ctx.solve(3, b, rtol=1e-10, maxiter=5000, method=method)
            Ax = oracle.dwf_apply(3, th, x, Nt, Nx, Ls, M5, MF)
            assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= 1.01e-10
    with dwf.DwfContext.wilson(Nt, Nx, 0.1, -1.0, 1.0) as ctx:
        w = rand_spinor(rng, Nt * Nx * 2)
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_LDS, mode)
        ctx.tune(dwf.DWF_TUNE_TILE_X, tile_x)
        for op in (0, 1):
            assert rel(ctx.apply(op, w), oracle.wilson_apply(op, th, w, Nt, Nx, 0.1, -1.0, 1.0)) <= 1e-12


def test_autotune_and_cache(tmp_path):
    Nt, Nx, Ls = 16, 32, 8
    print_case(test="autotune")
    rng = np.random.default_rng(8)
    th = rand_theta(rng, Nt, Nx)
    psi = rand_spinor(rng, Ls * Nt * Nx * 2)
    cache = str(tmp_path / "tune.txt")
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        r1 = ctx.autotune(4, dwf.DWF_PREC_FP32, cache)
        r2 = ctx.autotune(4, dwf.DWF_PREC_FP32, cache)
        assert r1["from_cache"] == 0 and r2["from_cache"] == 1 and r1["microseconds"] > 0
        assert all(r1[k] == r2[k] for k in ("tile_x", "prefetch", "chunks", "nt_store", "lds"))
        got = ctx.apply(0, psi)
        assert rel(got, oracle.dwf_apply(0, th, psi, Nt, Nx, Ls, M5, MF)) <= 1e-12


def test_autotune_binds_its_batch_to_the_configurations():
    """The autotuner times applies of its own batch size: with several configurations held
    it maps its right-hand sides onto them (nrhs / ncfg each), whatever batch the context's
    previous call used (a 4-RHS solve on 4 configurations before a 16-RHS autotune: 4 per
    configuration, not 1), and rejects a batch that is not a multiple of them.  Afterwards
    a 16-RHS apply matches the oracle configuration by configuration (right-hand side r
    on configuration r // 4)."""
    Nt, Nx, Ls, ncfg = 8, 10, 4, 4
    print_case(test="autotune_batch_map")
    rng = np.random.default_rng(81)
    ths = np.stack([rand_theta(rng, Nt, Nx) for _ in range(ncfg)])
    src = rand_spinor(rng, ncfg * Ls * Nt * Nx * 2).reshape(ncfg, -1)
    psi = rand_spinor(rng, 4 * ncfg * Ls * Nt * Nx * 2).reshape(4 * ncfg, -1)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths)
        ctx.solve_dirac(src, rtol=1e-8, maxiter=5000, method=dwf.DWF_SOLVE_MIXED)  # one RHS per configuration
        ctx.autotune(4 * ncfg, dwf.DWF_PREC_FP32)
        with pytest.raises(capi.DwfError):
            ctx.autotune(4 * ncfg + 1, dwf.DWF_PREC_FP32)
        got = ctx.apply(0, psi)
    for r in range(4 * ncfg):
        assert rel(got[r], oracle.dwf_apply(0, ths[r // 4], psi[r], Nt, Nx, Ls, M5, MF)) <= 1e-12
