"""pargemslr_amd: MI355X GeMSLR-preconditioned FGMRES solve path.

The compute path is the in-tree shared library libpargemslr_amd.so (host
setup in C++, preconditioner apply and FGMRES in HIP kernels for gfx950)
behind the C ABI of include/pargemslr_amd.h; this package binds it with
ctypes.  Importing does not touch the GPU.
"""
from .capi import (  # noqa: F401
    PGS_ORTH_CGS2, PGS_ORTH_DCGS2, PGS_ORTH_MGS, PGS_PREC_FP32, PGS_PREC_FP64, CsrHost, DeviceArray, DevOp, Gemslr, PgsError,
    Server, Stream, Workspace, config_params, convdiff, default_params, device_count, laplacian, lib, read_mm,
    read_mm_vector, read_inputs, HypreGemslr, PGS_INPUTS_SEQUENTIAL, PGS_INPUTS_HYPRE, rng_normal, rng_uniform, synchronize,
    PGS_ERR_BREAKDOWN, PGS_ERR_NOT_CONVERGED, PGS_ERR_NONFINITE,
)
