"""MI355X (gfx950) implementation of exact Khatri-Rao leverage-score sampling
and sampled CP-ALS, with the surface of fast_tensor_leverage's
``cpp_ext.als_module``.

The compute path is libftl_hip.so (hand-written HIP kernels, C ABI in
include/ftl_hip.h); this package is its host-side mirror of the reference's
Python-visible classes.
"""
from .als import ALS, LowRankTensor, SparseTensor  # noqa: F401
from .als_driver import als_prod  # noqa: F401
from .sampler import LarsenKoldaSampler, Sampler  # noqa: F401
from .serve import Server  # noqa: F401
from .tensors import PyLowRank, PySparseTensor  # noqa: F401
from ._lib import (cached_memory_bytes, cached_memory_limit, eigensolver_capped, release_cached_memory,  # noqa: F401
                   set_cached_memory_limit, tuning_lookup)

__all__ = ["Sampler", "LarsenKoldaSampler", "Server", "ALS", "LowRankTensor", "SparseTensor", "PyLowRank", "PySparseTensor",
           "als_prod", "eigensolver_capped", "tuning_lookup", "release_cached_memory", "cached_memory_bytes",
           "set_cached_memory_limit", "cached_memory_limit"]
__version__ = "0.1.0"
