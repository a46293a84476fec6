"""neighbor_amd -- MI355X-native LBVH neighbor search (Python host side).

The compute path is the HIP library ``libneighbor_amd.so`` (built by
``neighbor_amd/build.py`` for gfx950) behind the C ABI of ``include/neighbor_amd.h``.
This module mirrors the reference's objects (``LBVH``, ``LBVHTraverser``, the
count / neighbor-list ops; mphowardlab/neighbor ``include/neighbor/LBVH.h:47``,
``LBVHTraverser.h:60``, ``OutputOps.h:28-176``) over that ABI, with torch
tensors as device memory and torch streams as HIP streams.  There is no CPU
fallback: if the library is missing, importing the bindings raises.
"""
from __future__ import annotations

from ._capi import (  # noqa: F401
    NB_SENTINEL,
    NB_VARIANT_FLOAT_ROPE,
    NB_VARIANT_FLOAT_ROPE_EXACT,
    NB_VARIANT_WAVE,
    NB_VARIANT_WAVE_QUANTIZED,
    NB_VARIANT_QUANTIZED,
    NeighborError,
    lib,
    library_path,
)
from .api import LBVH, LBVHTraverser, exclusive_scan  # noqa: F401

__all__ = [
    "LBVH",
    "LBVHTraverser",
    "exclusive_scan",
    "NeighborError",
    "lib",
    "library_path",
    "NB_SENTINEL",
    "NB_VARIANT_QUANTIZED",
    "NB_VARIANT_FLOAT_ROPE",
    "NB_VARIANT_FLOAT_ROPE_EXACT",
    "NB_VARIANT_WAVE",
    "NB_VARIANT_WAVE_QUANTIZED",
]
