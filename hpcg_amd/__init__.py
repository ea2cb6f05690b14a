"""MI355X-native HPCG PCG iteration phase.

The product is the shared library hpcg_amd/libhpcgx.so (HIP kernels for
gfx950 behind the C ABI of include/hpcgx.h) plus the C++ drop-in hooks in
hpcg_amd/dropin/.  This package binds the C ABI with ctypes (hpcg_amd._lib)
and carries the reference's FLOP and byte models for reporting.
"""
from ._lib import (HPCGX_CG_FP64, HPCGX_CG_MIXED_FP16, HPCGX_CG_MIXED_FP32, HPCGX_COLS_COMPACT,
                   HPCGX_COLS_COMPACT_ALWAYS,
                   HPCGX_COLS_EXPLICIT, HPCGX_PREC_FP16, HPCGX_PREC_FP32, HPCGX_PREC_FP64, Context,
                   DeviceVector, HpcgxError, lib)
from .models import (algorithmic_bytes_per_iteration, hpcg_flops, level_shapes, kernel_bytes)

__all__ = [
    "Context", "DeviceVector", "HpcgxError", "lib", "hpcg_flops", "level_shapes",
    "kernel_bytes", "algorithmic_bytes_per_iteration",
    "HPCGX_CG_FP64", "HPCGX_CG_MIXED_FP32", "HPCGX_CG_MIXED_FP16",
    "HPCGX_PREC_FP64", "HPCGX_PREC_FP32", "HPCGX_PREC_FP16",
    "HPCGX_COLS_EXPLICIT", "HPCGX_COLS_COMPACT", "HPCGX_COLS_COMPACT_ALWAYS",
]
