"""The allocation-free operator apply (dwf_apply_device_noalloc) against dwf_apply_device.

Both run the same stencil kernels in the same order; the only difference is where the
normal operators' intermediate vector lives (a fixed per-GPU device array instead of the
context's workspace), so the outputs must agree bit for bit.  Each case makes two calls
in a row, the second with a smaller input, so the second call reuses the scratch array
the first one filled.  Sizes past the array are rejected on the host before any launch.
"""

import ctypes

import numpy as np
import pytest

from conftest import print_case

import three_part_dwf_study_amd as dwf
from three_part_dwf_study_amd import capi

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20  # DWF_NOALLOC_SCRATCH_BYTES in include/dwf_hip.h


def _spinors(rng, n, prec):
    a = rng.normal(size=2 * n)
    return a.astype(np.float64 if prec == dwf.DWF_PREC_FP64 else np.float32)


def _both(ctx, lib, op, prec, host, nrhs):
    """(dwf_apply_device output, dwf_apply_device_noalloc output) for the same input."""
    dtype = host.dtype
    src = capi.DeviceArray.from_host(host)
    ref = capi.DeviceArray(host.nbytes)
    got = capi.DeviceArray(host.nbytes)
    try:
        ctx.apply_device(op, prec, src.ptr, ref.ptr, nrhs)
        rc = lib.dwf_apply_device_noalloc(ctx.handle, op, prec, src.ptr, got.ptr, nrhs)
        assert rc == capi.DWF_OK, lib.dwf_last_error().decode()
        ctx.synchronize()
        return ref.download(dtype), got.download(dtype)
    finally:
        src.free()
        ref.free()
        got.free()


@pytest.mark.parametrize("prec", [dwf.DWF_PREC_FP64, dwf.DWF_PREC_FP32])
@pytest.mark.parametrize("lds", [0, 2])
def test_noalloc_matches_apply_device(prec, lds, rng):
    lib = dwf.load_library()
    # the first call at 96 x 96 x 16 (several right-hand sides), then a smaller call of
    # fewer right-hand sides, all four operators each
    Nt, Nx, Ls = 96, 96, 16
    th = rng.normal(0.0, 0.5, size=2 * Nt * Nx)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, 1.0, 0.05) as ctx:
        ctx.set_gauge(th)
        if lds:
            ctx.tune(dwf.DWF_TUNE_LDS, lds)
            ctx.tune(dwf.DWF_TUNE_TILE_X, 16)
        real = 8 if prec == dwf.DWF_PREC_FP64 else 4
        v5 = Ls * Nt * Nx
        big = min(6, SCRATCH_BYTES // (v5 * 4 * real))
        for op in (dwf.DWF_OP_D, dwf.DWF_OP_DDAG, dwf.DWF_OP_DDDAG, dwf.DWF_OP_DDAGD):
            for nrhs in (big, 1):
                print_case(test="noalloc", prec=prec, lds=lds, op=op, nrhs=nrhs)
                host = _spinors(rng, v5 * 2 * nrhs, prec)
                ref, got = _both(ctx, lib, op, prec, host, nrhs)
                assert np.all(np.isfinite(ref))
                assert np.array_equal(ref.view(np.uint8), got.view(np.uint8)), f"op={op} nrhs={nrhs}"


def test_noalloc_smaller_lattice_after_larger(rng):
    """The second call on a smaller lattice (another context on the same GPU) reads none
    of what the first call left in the scratch array."""
    lib = dwf.load_library()
    outs = []
    for (Nt, Nx, Ls, nrhs) in ((32, 48, 8, 3), (8, 6, 4, 2)):
        th = rng.normal(0.0, 0.5, size=2 * Nt * Nx)
        with dwf.DwfContext.dwf(Nt, Nx, Ls, 1.0, 0.05) as ctx:
            ctx.set_gauge(th)
            for op in (dwf.DWF_OP_DDDAG, dwf.DWF_OP_DDAGD):
                print_case(test="noalloc_lattices", Nt=Nt, Nx=Nx, Ls=Ls, op=op, nrhs=nrhs)
                host = _spinors(rng, Ls * Nt * Nx * 2 * nrhs, dwf.DWF_PREC_FP64)
                ref, got = _both(ctx, lib, op, dwf.DWF_PREC_FP64, host, nrhs)
                assert np.array_equal(ref.view(np.uint8), got.view(np.uint8))
                outs.append(got)
    assert len(outs) == 4


def test_noalloc_rejects_sizes_past_the_scratch(rng):
    lib = dwf.load_library()
    Nt, Nx, Ls = 96, 96, 16
    with dwf.DwfContext.dwf(Nt, Nx, Ls, 1.0, 0.05) as ctx:
        ctx.set_gauge(rng.normal(0.0, 0.5, size=2 * Nt * Nx))
        v5 = Ls * Nt * Nx
        nrhs = SCRATCH_BYTES // (v5 * 4 * 8) + 1  # one FP64 right-hand side too many
        # the size check runs on the host before any launch, so the pointers are never used
        dummy = capi.DeviceArray(256)
        try:
            print_case(test="noalloc_reject", nrhs=nrhs)
            rc = lib.dwf_apply_device_noalloc(ctx.handle, dwf.DWF_OP_DDAGD, dwf.DWF_PREC_FP64, dummy.ptr,
                                              ctypes.c_void_p(dummy.ptr + 128), nrhs)
            assert rc == capi.DWF_ERR_ARG
            assert b"scratch" in lib.dwf_last_error()
            # a bad operator is refused on the host as well
            rc = lib.dwf_apply_device_noalloc(ctx.handle, 7, dwf.DWF_PREC_FP64, dummy.ptr,
                                              ctypes.c_void_p(dummy.ptr + 128), nrhs)
            assert rc == capi.DWF_ERR_ARG
        finally:
            dummy.free()
