"""Sampler: the drop-in for ``cpp_ext.als_module.Sampler`` (als_module.cpp:30-82,
efficient_krp_sampler.hpp) backed by libftl_hip.so on one MI355X.

Host arrays follow the reference's conventions (NumPy, FP64 factors I_k x R,
uint64 samples N x J).  Unlike the reference, arguments are validated: a wrong
dtype or a non-contiguous array raises TypeError instead of being silently
copied (pybind11 forcecast) or misindexed (common.h:121-124), and an unknown
sampler type raises ValueError instead of exit(1) (als_module.cpp:57-61).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

PRECISIONS = {"fp64": _lib.FTL_FP64, "fp32": _lib.FTL_FP32}


def _ptr(a) -> int:
    return a.ctypes.data


def _require_array(a, dtype, ndim, name):
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name} must be a numpy.ndarray of {np.dtype(dtype).name}")
    if a.dtype != dtype:
        raise TypeError(f"{name} must have dtype {np.dtype(dtype).name}, got {a.dtype}")
    if a.ndim != ndim:
        raise TypeError(f"{name} must be {ndim}-D, got shape {a.shape}")
    if not a.flags.c_contiguous:
        raise TypeError(f"{name} must be C-contiguous")


def _dev_ptr(t, name) -> int:
    """Device address of a torch (ROCm) tensor or of an int address."""
    if t is None:
        return 0
    if isinstance(t, int):
        return t
    if hasattr(t, "data_ptr"):
        if hasattr(t, "is_cuda") and not t.is_cuda:
            raise TypeError(f"{name} must be a device tensor")
        if hasattr(t, "is_contiguous") and not t.is_contiguous():
            raise TypeError(f"{name} must be contiguous")
        return int(t.data_ptr())
    raise TypeError(f"{name} must be a device tensor or a device address")


class Sampler:
    """Exact leverage-score sampler over the Khatri-Rao product of ``matrices``.

    ``Sampler(matrices, J, R, "efficient")`` as in the reference
    (als_module.cpp:34-44).  ``precision`` selects the tree storage:
    "fp64" (every level FP64) or "fp32" (FP64 top levels, FP32 bottom levels
    and leaf scan).
    """

    def __new__(cls, matrices=None, J=None, R=None, type: str = "efficient", precision: str = "fp64"):
        # Sampler(U, J, R, "larsen_kolda") builds the Larsen-Kolda sampler
        # (als_module.cpp:45-52)
        if cls is Sampler and type == "larsen_kolda":
            return LarsenKoldaSampler(matrices, J, R)
        if cls is Sampler and type == "larsen_kolda_hybrid":
            return LarsenKoldaSampler(matrices, J, R, hybrid=True)
        return super().__new__(cls)

    def __init__(self, matrices, J: int, R: int, type: str = "efficient", precision: str = "fp64"):
        if type != "efficient":
            raise ValueError(f"Unsupported argument {type!r} passed for sampler construction "
                             "(provided: 'efficient', 'larsen_kolda', 'larsen_kolda_hybrid')")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        if not isinstance(matrices, (list, tuple)) or len(matrices) == 0:
            raise TypeError("matrices must be a non-empty list of I_k x R float64 arrays")
        J, R = int(J), int(R)
        if J < 1 or R < 1:
            raise ValueError("J and R must be positive")
        for k, U in enumerate(matrices):
            _require_array(U, np.float64, 2, f"matrices[{k}]")
            if U.shape[1] != R:
                raise ValueError(f"matrices[{k}] has {U.shape[1]} columns, expected R={R}")
        L = _lib.load()
        _lib.require_device()
        self.N = len(matrices)
        self.J = J
        self.R = R
        self.precision = precision
        self.dims = [int(U.shape[0]) for U in matrices]
        self._matrices = list(matrices)
        dims = (ctypes.c_uint64 * self.N)(*self.dims)
        ptrs = (ctypes.c_void_p * self.N)(*[_ptr(U) for U in matrices])
        h = ctypes.c_void_p()
        _lib.check(L.ftl_sampler_create(self.N, dims, J, R, ptrs, PRECISIONS[precision], ctypes.byref(h)),
                   "Sampler")
        self._h = h

    @classmethod
    def from_device(cls, tensors, J: int, R: int, precision: str = "fp64"):
        """Sampler over device-resident factors (torch tensors of shape I_k x R,
        float64, contiguous, on the GPU); they are copied into the sampler."""
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        for k, t in enumerate(tensors):
            if tuple(t.shape[1:]) != (int(R),) or t.dim() != 2:
                raise ValueError(f"tensors[{k}] must be I_k x R")
            if str(t.dtype) != "torch.float64":
                raise TypeError(f"tensors[{k}] must be float64")
        _lib.require_device()
        obj = cls.__new__(cls)
        obj.N, obj.J, obj.R, obj.precision = len(tensors), int(J), int(R), precision
        obj.dims = [int(t.shape[0]) for t in tensors]
        obj._matrices = None
        dims = (ctypes.c_uint64 * obj.N)(*obj.dims)
        ptrs = (ctypes.c_void_p * obj.N)(*[_dev_ptr(t, f"tensors[{k}]") for k, t in enumerate(tensors)])
        h = ctypes.c_void_p()
        _lib.check(_lib.load().ftl_sampler_create_device(obj.N, dims, obj.J, obj.R, ptrs, PRECISIONS[precision],
                                                         ctypes.byref(h)), "Sampler")
        obj._h = h
        return obj

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().ftl_sampler_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("sampler is closed")
        return self._h

    # -- reference surface --------------------------------------------------------
    def _check_samples(self, samples):
        _require_array(samples, np.uint64, 2, "samples")
        if samples.shape != (self.N, self.J):
            raise ValueError(f"samples must have shape (N, J) = {(self.N, self.J)}, got {samples.shape}")

    def _draws(self, random_draws):
        if random_draws is None:
            return None
        _require_array(random_draws, np.float64, 1, "random_draws")
        if random_draws.size != 2 * self.N * self.J:
            raise ValueError("random_draws must hold 2*N*J uniforms")
        return _ptr(random_draws)

    def KRPDrawSamples(self, j: int, samples: np.ndarray, random_draws: np.ndarray = None):
        """Fill rows k != j of ``samples`` (N x J uint64) with leverage-score draws
        (als_module.cpp:64-68).  ``random_draws`` (2*N*J uniforms, layout of
        efficient_krp_sampler.hpp:221-229) replaces the built-in generator."""
        self._check_samples(samples)
        if int(j) < 0:
            raise ValueError("j must be non-negative")
        _lib.check(_lib.load().ftl_sampler_draw(self.handle, int(j), _ptr(samples), self._draws(random_draws)),
                   "KRPDrawSamples")

    def KRPDrawSamples_materialize(self, j: int, samples: np.ndarray, h_out: np.ndarray,
                                   weights_out: np.ndarray, random_draws: np.ndarray = None):
        """KRPDrawSamples plus the sampled KRP rows and weights (als_module.cpp:70-81)."""
        self._check_samples(samples)
        _require_array(h_out, np.float64, 2, "h_out")
        _require_array(weights_out, np.float64, 1, "weights_out")
        if h_out.shape != (self.J, self.R) or weights_out.shape != (self.J,):
            raise ValueError("h_out must be J x R and weights_out J")
        _lib.check(_lib.load().ftl_sampler_draw_materialize(self.handle, int(j), _ptr(samples), _ptr(h_out),
                                                            _ptr(weights_out), self._draws(random_draws)),
                   "KRPDrawSamples_materialize")

    def update_sampler(self, j: int, Uj: np.ndarray = None):
        """Rebuild the tree of mode j (efficient_krp_sampler.hpp:73-75) from
        ``Uj`` or from the array the sampler was built on."""
        if not 0 <= int(j) < self.N:
            raise ValueError("j out of range")
        U = self._matrices[j] if Uj is None else Uj
        _require_array(U, np.float64, 2, "Uj")
        if U.shape != (self.dims[j], self.R):
            raise ValueError("Uj has the wrong shape")
        if Uj is not None:
            self._matrices[j] = Uj
        _lib.check(_lib.load().ftl_sampler_update(self.handle, int(j), _ptr(U)), "update_sampler")

    # -- device-resident path ---------------------------------------------------------
    def draw_device(self, j: int, samples, h=None, weights=None, draws=None, stream=None):
        """Enqueue a draw on device buffers (torch tensors or addresses); no sync."""
        st = 0 if stream is None else (stream if isinstance(stream, int) else int(stream.cuda_stream))
        _lib.check(_lib.load().ftl_sampler_draw_device(
            self.handle, int(j), _dev_ptr(samples, "samples"), _dev_ptr(h, "h") or None,
            _dev_ptr(weights, "weights") or None, _dev_ptr(draws, "draws") or None, st or None),
            "draw_device")

    def set_seed(self, seed: int, counter: int = 0):
        _lib.check(_lib.load().ftl_sampler_set_seed(self.handle, int(seed) & (2**64 - 1), int(counter)),
                   "set_seed")

    def set_option(self, key: str, value: int):
        """Tuning knob of the draw kernel (see ftl_sampler_set_option)."""
        _lib.check(_lib.load().ftl_sampler_set_option(self.handle, key.encode(), int(value)), "set_option")

    def get_option(self, key: str) -> int:
        out = ctypes.c_int64()
        _lib.check(_lib.load().ftl_sampler_get_option(self.handle, key.encode(), ctypes.byref(out)), "get_option")
        return int(out.value)

    def set_tuning(self, group_lanes: int = 0, grid_blocks: int = 0):
        self.set_option("group_lanes", group_lanes)
        self.set_option("grid_blocks", grid_blocks)

    def enable_timing(self, on: bool = True):
        _lib.check(_lib.load().ftl_sampler_enable_timing(self.handle, 1 if on else 0), "enable_timing")

    def last_timing(self):
        """Average kernel durations (ms, HIP events on the launch stream) over
        the draws since timing was enabled or last read."""
        out = np.zeros(4)
        _lib.check(_lib.load().ftl_sampler_last_timing(self.handle, _ptr(out)), "last_timing")
        return {"gram_pinv_ms": out[0], "eigen_modes_ms": out[1], "draw_ms": out[2], "calls": int(out[3])}

    def stream(self) -> int:
        return int(_lib.load().ftl_sampler_stream(self.handle) or 0)

    def info(self) -> dict:
        out = np.zeros(8, dtype=np.uint64)
        _lib.check(_lib.load().ftl_sampler_info(self.handle, _ptr(out)), "info")
        keys = ["N", "J", "R", "precision", "fp64_levels", "device_bytes", "node_count0", "lfill_level0"]
        return {k: int(v) for k, v in zip(keys, out)}

    # -- test surface -----------------------------------------------------------------------
    def get_tree(self, k: int) -> np.ndarray:
        n = self.dims[k]
        F = min(self.R, n)
        leaves = -(-n // F)
        G = np.empty((2 * leaves - 1, self.R, self.R))
        _lib.check(_lib.load().ftl_sampler_get_tree(self.handle, int(k), _ptr(G), G.shape[0]), "get_tree")
        return G

    def compute_m(self, j: int):
        pinv = np.empty((self.R, self.R))
        E = np.empty((self.N, self.R, self.R))
        _lib.check(_lib.load().ftl_sampler_compute_m(self.handle, int(j), _ptr(pinv), _ptr(E)), "compute_m")
        return pinv, E


class LarsenKoldaSampler:
    """Larsen-Kolda leverage sampler (larsen_kolda_sampler.hpp:18-136):
    ``Sampler(matrices, J, R, "larsen_kolda")``.  Every mode is sampled
    independently in proportion to its factor's row leverage scores; the
    weights are 1 / (J prod_k p_k).  ``random_draws`` uses its first N*J
    uniforms (mode k's at [k*J, (k+1)*J)); without it the built-in Philox
    stream advances per call (the reference draws from mt19937).

    ``hybrid=True`` (``Sampler(matrices, J, R, "larsen_kolda_hybrid")``,
    larsen_kolda_hybrid.hpp:20-248): the tuples of probability >= 1/J come
    first with weight 1, the rest by rejection from the product distribution
    with weight (1 - p_det) / (J p); ``last_info()`` reports the
    deterministic count and mass of the last draw."""

    def __init__(self, matrices, J: int, R: int, hybrid: bool = False):
        if not isinstance(matrices, (list, tuple)) or len(matrices) == 0:
            raise TypeError("matrices must be a non-empty list of I_k x R float64 arrays")
        J, R = int(J), int(R)
        if J < 1 or R < 1:
            raise ValueError("J and R must be positive")
        for k, U in enumerate(matrices):
            _require_array(U, np.float64, 2, f"matrices[{k}]")
            if U.shape[1] != R:
                raise ValueError(f"matrices[{k}] has {U.shape[1]} columns, expected R={R}")
        L = _lib.load()
        _lib.require_device()
        self.N, self.J, self.R = len(matrices), J, R
        self.dims = [int(U.shape[0]) for U in matrices]
        self._matrices = list(matrices)
        dims = (ctypes.c_uint64 * self.N)(*self.dims)
        ptrs = (ctypes.c_void_p * self.N)(*[_ptr(U) for U in matrices])
        h = ctypes.c_void_p()
        self.hybrid = bool(hybrid)
        create = L.ftl_lkh_sampler_create if self.hybrid else L.ftl_lk_sampler_create
        _lib.check(create(self.N, dims, J, R, ptrs, ctypes.byref(h)), "Sampler(larsen_kolda)")
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().ftl_lk_sampler_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("sampler is closed")
        return self._h

    def _check_samples(self, samples):
        _require_array(samples, np.uint64, 2, "samples")
        if samples.shape != (self.N, self.J):
            raise ValueError(f"samples must have shape (N, J) = {(self.N, self.J)}, got {samples.shape}")

    def _draws(self, random_draws):
        if random_draws is None:
            return None
        _require_array(random_draws, np.float64, 1, "random_draws")
        if random_draws.size != 2 * self.N * self.J:
            raise ValueError("random_draws must hold 2*N*J uniforms")
        return _ptr(random_draws)

    def KRPDrawSamples(self, j: int, samples: np.ndarray, random_draws: np.ndarray = None):
        self._check_samples(samples)
        if int(j) < 0:
            raise ValueError("j must be non-negative")
        _lib.check(_lib.load().ftl_lk_sampler_draw(self.handle, int(j), _ptr(samples), None, None,
                                                   self._draws(random_draws)), "KRPDrawSamples")

    def KRPDrawSamples_materialize(self, j: int, samples: np.ndarray, h_out: np.ndarray,
                                   weights_out: np.ndarray, random_draws: np.ndarray = None):
        self._check_samples(samples)
        _require_array(h_out, np.float64, 2, "h_out")
        _require_array(weights_out, np.float64, 1, "weights_out")
        if h_out.shape != (self.J, self.R) or weights_out.shape != (self.J,):
            raise ValueError("h_out must be J x R and weights_out J")
        _lib.check(_lib.load().ftl_lk_sampler_draw(self.handle, int(j), _ptr(samples), _ptr(h_out),
                                                   _ptr(weights_out), self._draws(random_draws)),
                   "KRPDrawSamples_materialize")

    def update_sampler(self, j: int, Uj: np.ndarray = None):
        """Recompute factor j's leverage scores (larsen_kolda_sampler.hpp:45-100)
        from ``Uj`` (copied in) or, without it, from the factor resident on
        the device."""
        if not 0 <= int(j) < self.N:
            raise ValueError("j out of range")
        if Uj is not None:
            _require_array(Uj, np.float64, 2, "Uj")
            if Uj.shape != (self.dims[j], self.R):
                raise ValueError("Uj has the wrong shape")
            self._matrices[j] = Uj
        _lib.check(_lib.load().ftl_lk_sampler_update(self.handle, int(j), None if Uj is None else _ptr(Uj)),
                   "update_sampler")

    def factor_device_ptr(self, k: int) -> int:
        out = ctypes.c_void_p()
        _lib.check(_lib.load().ftl_lk_sampler_factor_device(self.handle, int(k), ctypes.byref(out)), "factor")
        return int(out.value)

    def stream(self) -> int:
        """hipStream_t of the handle (update_sampler and the host draws run on it)."""
        return int(_lib.load().ftl_lk_sampler_stream(self.handle) or 0)

    def draw_device(self, j: int, samples, h=None, weights=None, draws=None, stream=None):
        st = 0 if stream is None else (stream if isinstance(stream, int) else int(stream.cuda_stream))
        _lib.check(_lib.load().ftl_lk_sampler_draw_device(
            self.handle, int(j), _dev_ptr(samples, "samples"), _dev_ptr(h, "h") or None,
            _dev_ptr(weights, "weights") or None, _dev_ptr(draws, "draws") or None, st or None), "draw_device")

    def set_seed(self, seed: int, counter: int = 0):
        _lib.check(_lib.load().ftl_lk_sampler_set_seed(self.handle, int(seed) & (2**64 - 1), int(counter)),
                   "set_seed")

    def last_info(self) -> dict:
        """Hybrid sampler, last draw: deterministic samples, their probability
        mass, candidate tuples, whether the enumeration was skipped, samples
        given weight 0 after their rejection attempts ran out."""
        out = np.zeros(5)
        _lib.check(_lib.load().ftl_lk_sampler_info(self.handle, _ptr(out)), "last_info")
        return {"deterministic": int(out[0]), "p_det": float(out[1]), "candidates": float(out[2]),
                "skipped": bool(out[3]), "exhausted": int(out[4])}

    def leverage(self, k: int):
        """(leverage scores of factor k, their total)."""
        out = np.empty(self.dims[k])
        tot = np.zeros(1)
        _lib.check(_lib.load().ftl_lk_sampler_leverage(self.handle, int(k), _ptr(out), _ptr(tot)), "leverage")
        return out, float(tot[0])
