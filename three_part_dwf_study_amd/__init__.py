"""MI355X (gfx950) fermion hot path for 2D U(1) Wilson and Shamir domain-wall HMC.

The product is ``libdwfhip.so`` (HIP kernels + C ABI, declared in ``include/dwf_hip.h``)
with C++ drop-in wrappers in ``include/dwf_hip.hpp``.  This package holds its sources
(``csrc/``) and a thin ctypes mirror of the C ABI (:mod:`.capi`) used by the tests, the
smoke check and ``bench.py``.  There is no CPU fallback: every call goes to the GPU and
fails loudly (``DwfError``) when the library or the device is missing.
"""

from .capi import (  # noqa: F401
    DWF_OP_D,
    DWF_OP_DDAG,
    DWF_OP_DDDAG,
    DWF_OP_DDAGD,
    DWF_PREC_FP32,
    DWF_PREC_FP64,
    DWF_SOLVE_CG64,
    DWF_SOLVE_EO_MIXED,
    DWF_SOLVE_LEARNED,
    DWF_SOLVE_MIXED,
    DWF_TUNE_NT_STORE,
    DWF_TUNE_GRAPHS,
    DWF_TUNE_LDS,
    DWF_TUNE_NET_FUSED,
    DWF_TUNE_FCG_BETA,
    DWF_TUNE_EO_FUSED,
    DWF_TUNE_NET_TILE,
    DWF_TUNE_STREAMS,
    DWF_TUNE_NET_SPLIT,
    DWF_TUNE_PREFETCH,
    DWF_TUNE_RELIABLE_DELTA,
    DWF_TUNE_SCHUNKS,
    DWF_TUNE_TILE_X,
    DWF_TUNE_XCD_REMAP,
    DwfContext,
    DwfError,
    Network,
    build_library,
    library_path,
    load_library,
)
