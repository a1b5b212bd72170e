"""cpabdiffeo_amd: MI355X-native CPA-velocity integration (drop-in for the
TransformWrapper path of freifeld/cpabDiffeo).

Host geometry (tessellations, bases, priors) is numpy; every integration,
velocity, cell-lookup, remap and likelihood call runs in hand-written HIP
kernels for gfx950 (libcpab_hip.so, C ABI in include/cpab_hip.h).  There is
no CPU fallback: without the built library or a GPU, those calls raise.
"""
from .tessellation import TESS_I, TESS_II, Tessellation, tess_code
from .space import CpaSpace, PAT
from .transform_wrapper import (TransformWrapper, TransformWrapper1D, TransformWrapper3D,
                                default_params_flow_int, Bunch)
from .gpuarray import CpuGpuArray
from ._lib import CpabError

__all__ = ["TESS_I", "TESS_II", "Tessellation", "tess_code", "CpaSpace", "PAT",
           "TransformWrapper", "TransformWrapper1D", "TransformWrapper3D",
           "default_params_flow_int", "Bunch", "CpuGpuArray", "CpabError"]
