"""SparseTensor, LowRankTensor and ALS: the drop-ins for the sampled CP-ALS
surface of ``cpp_ext.als_module`` (als_module.cpp:85-115; sparse_tensor.hpp,
low_rank_tensor.hpp, als.hpp) backed by libftl_hip.so on one MI355X.

Ownership differs from the reference's zero-copy aliasing (common.h:70-82):
the sparse tensor and the ALS factors live in HBM.  ``ALS`` copies the
factors of its ``LowRankTensor`` to the device at construction and, after
every update, writes the updated factor and sigma back into the same NumPy
arrays (``host_sync=True``, the reference's visible behaviour) -- or leaves
them on the device (``host_sync=False``) until ``sync_to_host()``.  A factor
changed on the host after ``ALS`` was built is pushed with
``sync_from_host(k)``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .sampler import PRECISIONS, _dev_ptr, _ptr, _require_array


def _handle_out():
    return ctypes.c_void_p()


def _ptr_list(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class SparseTensor:
    """``SparseTensor(idx, values, "sort")`` (sparse_tensor.hpp:46-77).

    idx: nnz x N uint32 (row-major), values: nnz float64.  The per-mode sorted
    lookups are built in HBM.  ``method`` "sort" (binary search of the sorted
    keys, SortIdxLookup) or "hash" (a device open-addressing coordinate hash
    per mode over the same sorted nonzeros, the device form of the
    reference's HashIdxLookup, hash_lookup.hpp:13-162; the same matches).
    ``dims`` optionally gives the mode sizes (default max coordinate + 1).
    """

    def __init__(self, idx, values, method: str = "sort", dims=None):
        if method not in ("sort", "hash"):
            raise ValueError(f"unknown lookup method '{method}' (expected 'sort' or 'hash')")
        _require_array(idx, np.uint32, 2, "idx")
        _require_array(values, np.float64, 1, "values")
        nnz, N = idx.shape
        if values.shape[0] != nnz:
            raise ValueError("values must hold one entry per row of idx")
        if nnz == 0:
            raise ValueError("the tensor needs at least one nonzero")
        _lib.require_device()
        L = _lib.load()
        self._h = None
        dims_arr = None
        if dims is not None:
            dims_arr = np.ascontiguousarray(dims, dtype=np.uint64)
            if dims_arr.shape != (N,):
                raise ValueError("dims must hold one size per mode")
        h = _handle_out()
        _lib.check(L.ftl_sparse_create(N, nnz, _ptr(idx), _ptr(values),
                                       None if dims_arr is None else _ptr(dims_arr), ctypes.byref(h)),
                   "SparseTensor")
        self._h = h
        # held by the object: module globals may be gone when it is collected at exit
        self._destroy = L.ftl_sparse_destroy
        info = np.zeros(2 + N, dtype=np.uint64)
        _lib.check(L.ftl_sparse_info(self._h, _ptr(info)))
        self.N = N
        self.nnz = nnz
        self.dims = [int(d) for d in info[2:]]
        self.last_found = 0
        self.method = method
        if method == "hash":
            _lib.check(L.ftl_sparse_set_lookup(self._h, 1), "SparseTensor (hash lookup)")

    @classmethod
    def from_device(cls, idx, values, dims=None, method: str = "sort"):
        """The same from device tensors (torch, on the GPU, contiguous):
        idx nnz x N int32/uint32 coordinates, values nnz float64.  They are
        read during the call and not retained."""
        if method not in ("sort", "hash"):
            raise ValueError(f"unknown lookup method '{method}' (expected 'sort' or 'hash')")
        if idx.dim() != 2 or values.dim() != 1 or values.shape[0] != idx.shape[0]:
            raise ValueError("idx must be nnz x N and values nnz")
        if str(idx.dtype) not in ("torch.int32", "torch.uint32") or str(values.dtype) != "torch.float64":
            raise TypeError("idx must be int32/uint32 and values float64")
        nnz, N = int(idx.shape[0]), int(idx.shape[1])
        if nnz == 0:
            raise ValueError("the tensor needs at least one nonzero")
        _lib.require_device()
        obj = cls.__new__(cls)
        obj._h = None
        dims_arr = None if dims is None else np.ascontiguousarray(dims, dtype=np.uint64)
        if dims_arr is not None and dims_arr.shape != (N,):
            raise ValueError("dims must hold one size per mode")
        h = _handle_out()
        _lib.check(_lib.load().ftl_sparse_create_device(N, nnz, _dev_ptr(idx, "idx"), _dev_ptr(values, "values"),
                                                        None if dims_arr is None else _ptr(dims_arr),
                                                        ctypes.byref(h)), "SparseTensor")
        obj._h = h
        obj._destroy = _lib.load().ftl_sparse_destroy
        info = np.zeros(2 + N, dtype=np.uint64)
        _lib.check(_lib.load().ftl_sparse_info(obj._h, _ptr(info)))
        obj.N, obj.nnz, obj.dims, obj.last_found = N, nnz, [int(d) for d in info[2:]], 0
        obj.method = method
        if method == "hash":
            _lib.check(_lib.load().ftl_sparse_set_lookup(obj._h, 1), "SparseTensor (hash lookup)")
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            destroy = getattr(self, "_destroy", None) or _lib.load().ftl_sparse_destroy
            destroy(h)
            self._h = None

    def get_normsq(self) -> float:
        out = ctypes.c_double()
        _lib.check(_lib.load().ftl_sparse_normsq(self._h, ctypes.byref(out)))
        return out.value

    def _check_samples(self, samples, j):
        _require_array(samples, np.uint64, 2, "samples")
        if samples.shape[1] != self.N:
            raise ValueError(f"samples must be J x N (N = {self.N})")
        if not 0 <= j < self.N:
            raise ValueError(f"mode index {j} out of range")

    def execute_downsampled_mttkrp_py(self, samples, lhs, j: int, result):
        """result (I_j x R) = sum over samples s and matching nonzeros z of
        value_z * lhs[s] at row z_j (sparse_tensor.hpp:95-115)."""
        self._check_samples(samples, j)
        _require_array(lhs, np.float64, 2, "lhs")
        _require_array(result, np.float64, 2, "result")
        if lhs.shape[0] != samples.shape[0] or result.shape[1] != lhs.shape[1]:
            raise ValueError("lhs must be J x R and result I_j x R")
        found = ctypes.c_uint64()
        _lib.check(_lib.load().ftl_sparse_mttkrp(self._h, j, samples.shape[0], _ptr(samples), _ptr(lhs),
                                                 lhs.shape[1], _ptr(result), result.shape[0],
                                                 ctypes.byref(found)), "execute_downsampled_mttkrp")
        self.last_found = int(found.value)

    def lookup(self, samples, j: int, counts_only: bool = False):
        """Matched nonzeros of every sample on mode j's lookup: (counts per
        sample, input positions of the matches in sample order; None with
        ``counts_only``)."""
        self._check_samples(samples, j)
        Jp = samples.shape[0]
        counts = np.zeros(Jp, dtype=np.uint64)
        found = ctypes.c_uint64()
        L = _lib.load()
        _lib.check(L.ftl_sparse_lookup(self._h, j, Jp, _ptr(samples), _ptr(counts), None, 0, ctypes.byref(found)))
        if counts_only:
            return counts, None
        ids = np.zeros(max(int(found.value), 1), dtype=np.uint64)
        _lib.check(L.ftl_sparse_lookup(self._h, j, Jp, _ptr(samples), None, _ptr(ids), ids.size,
                                       ctypes.byref(found)))
        return counts, ids[: int(found.value)]

    def compute_residual_normsq(self, sigma, U_list) -> float:
        """||X - [[sigma; U]]||^2 expanded as ||X||^2 - 2<X, M> + ||M||^2
        (sort_lookup.hpp:134-197), host factors."""
        _require_array(sigma, np.float64, 1, "sigma")
        if len(U_list) != self.N:
            raise ValueError("one factor per mode expected")
        R = sigma.shape[0]
        for k, U in enumerate(U_list):
            _require_array(U, np.float64, 2, f"U[{k}]")
            if U.shape[1] != R:
                raise ValueError("factors must have R columns")
        I = np.array([U.shape[0] for U in U_list], dtype=np.uint64)
        out = ctypes.c_double()
        _lib.check(_lib.load().ftl_sparse_residual_normsq(self._h, R, _ptr(I), _ptr(sigma), _ptr_list(U_list),
                                                          ctypes.byref(out)), "compute_residual_normsq")
        return out.value


class _FactorHandle:
    """Device copy of a list of factors (an ftl_als handle, optionally bound
    to a sparse tensor)."""

    def __init__(self, U_list, R, sigma, X=None):
        I = np.array([U.shape[0] for U in U_list], dtype=np.uint64)
        h = _handle_out()
        _lib.check(_lib.load().ftl_als_create(X._h if X is not None else None, len(U_list), _ptr(I), R,
                                              _ptr_list(U_list), ctypes.byref(h)), "ALS")
        self._h = h
        self._destroy = _lib.load().ftl_als_destroy
        _lib.check(_lib.load().ftl_als_set_sigma(self._h, _ptr(sigma)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None


class LowRankTensor:
    """``LowRankTensor(R, U_list)`` or ``LowRankTensor(R, max_rhs_rows, U_list)``
    (low_rank_tensor.hpp:26-68): factors U_k (I_k x R float64), sigma = 1."""

    def __init__(self, R: int, *args):
        if len(args) == 1:
            U_list = args[0]
        elif len(args) == 2:
            U_list = args[1]
        else:
            raise TypeError("LowRankTensor(R, U_list) or LowRankTensor(R, max_rhs_rows, U_list)")
        self.R = int(R)
        self.U = list(U_list)
        for k, U in enumerate(self.U):
            _require_array(U, np.float64, 2, f"U[{k}]")
            if U.shape[1] != self.R:
                raise ValueError(f"U[{k}] must have R = {self.R} columns")
        self.N = len(self.U)
        self.dims = [U.shape[0] for U in self.U]
        self.sigma = np.ones(self.R)
        self._als = None

    # sigma -----------------------------------------------------------------
    def get_sigma(self, out, j: int = -1):
        """get_sigma (low_rank_tensor.hpp:181-183); j is ignored as there."""
        if self._als is not None and not self._als.host_sync:
            self._als._pull_sigma()
        out[:] = self.sigma

    def get_sigma_py(self, out, j: int = -1):
        self.get_sigma(out, j)

    def set_sigma(self, new_sigma):
        self.sigma[:] = new_sigma
        if self._als is not None:
            _lib.check(_lib.load().ftl_als_set_sigma(self._als._fh._h, _ptr(self.sigma)))

    def set_sigma_py(self, new_sigma):
        self.set_sigma(new_sigma)

    # device work -------------------------------------------------------------
    def renormalize_columns(self, j: int):
        """Column norms of U_j (every mode for j = -1) divided out; sigma =
        those norms (their product for j = -1) (low_rank_tensor.hpp:203-260)."""
        if not -1 <= j < self.N:
            raise ValueError(f"mode index {j} out of range")
        L = _lib.load()
        if self._als is not None:
            fh = self._als._fh
            _lib.check(L.ftl_als_renormalize_columns(fh._h, j), "renormalize_columns")
            for k in range(self.N):
                if j == -1 or k == j:
                    _lib.check(L.ftl_als_get_factor(fh._h, k, _ptr(self.U[k])))
            _lib.check(L.ftl_als_get_sigma(fh._h, _ptr(self.sigma)))
            return
        _lib.require_device()
        fh = _FactorHandle(self.U, self.R, self.sigma)
        _lib.check(L.ftl_als_renormalize_columns(fh._h, j), "renormalize_columns")
        for k in range(self.N):
            if j == -1 or k == j:
                _lib.check(L.ftl_als_get_factor(fh._h, k, _ptr(self.U[k])))
        _lib.check(L.ftl_als_get_sigma(fh._h, _ptr(self.sigma)))

    def get_normsq(self) -> float:
        """||[[sigma; U]]||^2 = sum(sigma sigma^T (*) prod_k U_k^T U_k)
        (low_rank_tensor.hpp:71-78, common.h:215-226)."""
        L = _lib.load()
        out = ctypes.c_double()
        if self._als is not None:
            if self._als.host_sync:
                _lib.check(L.ftl_als_set_sigma(self._als._fh._h, _ptr(self.sigma)))
            _lib.check(L.ftl_als_lowrank_normsq(self._als._fh._h, ctypes.byref(out)))
            return out.value
        _lib.require_device()
        fh = _FactorHandle(self.U, self.R, self.sigma)
        _lib.check(L.ftl_als_lowrank_normsq(fh._h, ctypes.byref(out)))
        return out.value

    def compute_residual_normsq(self, sigma_other, U_other):
        raise NotImplementedError("the low-rank-vs-low-rank residual is outside the sparse sampled-ALS path; "
                                  "use SparseTensor.compute_residual_normsq")


class ALS:
    """``ALS(cp_decomp, ground_truth)`` (als.hpp:22-37) for a sparse ground
    truth, with ``initialize_ds_als`` and ``execute_ds_als_update``."""

    def __init__(self, cp_decomp: LowRankTensor, ground_truth: SparseTensor, host_sync: bool = True):
        if not isinstance(cp_decomp, LowRankTensor):
            raise TypeError("cp_decomp must be a LowRankTensor")
        if not isinstance(ground_truth, SparseTensor):
            raise TypeError("the ground truth must be a SparseTensor (dense tensors are not on this path)")
        if cp_decomp.N != ground_truth.N:
            raise ValueError("the decomposition and the tensor have different mode counts")
        _lib.require_device()
        self.cp_decomp = cp_decomp
        self.ground_truth = ground_truth
        self.host_sync = bool(host_sync)
        self._fh = _FactorHandle(cp_decomp.U, cp_decomp.R, cp_decomp.sigma, ground_truth)
        cp_decomp._als = self
        self.J = None

    def initialize_ds_als(self, J: int, method: str = "efficient", precision: str = "fp64", seed=None):
        """initialize_ds_als(J, method) (als.hpp:39-55): method "efficient" (the
        exact KRP leverage sampler; precision applies), "larsen_kolda" or
        "larsen_kolda_hybrid"."""
        kinds = {"efficient": _lib.FTL_SAMPLER_EFFICIENT, "larsen_kolda": _lib.FTL_SAMPLER_LARSEN_KOLDA,
                 "larsen_kolda_hybrid": _lib.FTL_SAMPLER_LARSEN_KOLDA_HYBRID}
        if method not in kinds:
            raise ValueError(f"sampler type '{method}' is not available (provided: {sorted(kinds)})")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        L = _lib.load()
        _lib.check(L.ftl_als_initialize_ds_sampler(self._fh._h, int(J), kinds[method], PRECISIONS[precision]),
                   "initialize_ds_als")
        self.J = int(J)
        if seed is not None:
            _lib.check(L.ftl_als_set_seed(self._fh._h, int(seed), 0))

    def set_seed(self, seed: int, counter: int = 0):
        _lib.check(_lib.load().ftl_als_set_seed(self._fh._h, int(seed), int(counter)))

    def execute_ds_als_update(self, j: int, renormalize: bool = True, update_sampler: bool = True,
                              random_draws=None):
        """One sampled least-squares update of factor j (als.hpp:57-220).
        ``random_draws``: optional 2*N*J uniforms (layout of
        efficient_krp_sampler.hpp:221-236) instead of the built-in ones."""
        if self.J is None:
            raise ValueError("initialize_ds_als has not been called")
        if not 0 <= j < self.cp_decomp.N:
            raise ValueError(f"mode index {j} out of range")
        draws = None
        if random_draws is not None:
            _require_array(random_draws, np.float64, 1, "random_draws")
            if random_draws.size != 2 * self.cp_decomp.N * self.J:
                raise ValueError("random_draws must hold 2*N*J uniforms")
            draws = _ptr(random_draws)
        L = _lib.load()
        _lib.check(L.ftl_als_ds_update(self._fh._h, j, int(bool(renormalize)), int(bool(update_sampler)), draws),
                   "execute_ds_als_update")
        if self.host_sync:
            _lib.check(L.ftl_als_get_factor(self._fh._h, j, _ptr(self.cp_decomp.U[j])))
            self._pull_sigma()

    def execute_exact_als_update(self, j: int, renormalize: bool = True, update_sampler: bool = True):
        """ALS::execute_exact_als_update (als.hpp:222-270): U_j = (exact
        MTTKRP) pinv((*)_{k != j} U_k^T U_k), then renormalize_columns(j).
        update_sampler is accepted and unused, as in the reference."""
        if not 0 <= int(j) < self.cp_decomp.N:
            raise ValueError("j out of range")
        L = _lib.load()
        _lib.check(L.ftl_als_exact_update(self._fh._h, int(j), int(bool(renormalize)), int(bool(update_sampler))),
                   "execute_exact_als_update")
        if self.host_sync:
            _lib.check(L.ftl_als_get_factor(self._fh._h, j, _ptr(self.cp_decomp.U[j])))
            self._pull_sigma()

    def _pull_sigma(self):
        _lib.check(_lib.load().ftl_als_get_sigma(self._fh._h, _ptr(self.cp_decomp.sigma)))

    def sync_to_host(self):
        """Copy every device factor and sigma into the LowRankTensor's arrays."""
        L = _lib.load()
        for k in range(self.cp_decomp.N):
            _lib.check(L.ftl_als_get_factor(self._fh._h, k, _ptr(self.cp_decomp.U[k])))
        self._pull_sigma()

    def sync_from_host(self, k: int, update_sampler: bool = True):
        """Push host factor k (changed in place by the caller) to the device."""
        L = _lib.load()
        _lib.check(L.ftl_als_set_factor(self._fh._h, k, _ptr(self.cp_decomp.U[k])))
        if update_sampler and self.J is not None:
            _lib.check(L.ftl_als_update_sampler(self._fh._h, k))

    def residual_normsq(self) -> float:
        """ground_truth.compute_residual_normsq(sigma, U) on the device factors."""
        L = _lib.load()
        if self.host_sync:
            _lib.check(L.ftl_als_set_sigma(self._fh._h, _ptr(self.cp_decomp.sigma)))
        out = ctypes.c_double()
        _lib.check(L.ftl_als_residual_normsq(self._fh._h, ctypes.byref(out)), "compute_residual_normsq")
        return out.value

    def enable_timing(self, enable: bool = True):
        _lib.check(_lib.load().ftl_als_enable_timing(self._fh._h, int(bool(enable))))

    def last_stats(self) -> dict:
        out = np.zeros(8)
        _lib.check(_lib.load().ftl_als_last_stats(self._fh._h, _ptr(out)))
        keys = ["unique_samples", "found", "touched_rows", "pieces", "draw_ms", "dedup_pinv_probe_ms",
                "mttkrp_solve_ms", "renormalize_rebuild_ms"]
        return {k: (int(v) if i < 4 else float(v)) for i, (k, v) in enumerate(zip(keys, out))}

    def last_draw(self):
        """(samples N x J uint64, h J x R, weights J) of the last sampled update."""
        if self.J is None:
            raise RuntimeError("initialize_ds_als has not been called")
        N, J, R = self.cp_decomp.N, self.J, self.cp_decomp.R
        samples = np.zeros((N, J), dtype=np.uint64)
        h = np.empty((J, R))
        w = np.empty(J)
        _lib.check(_lib.load().ftl_als_last_draw(self._fh._h, _ptr(samples), _ptr(h), _ptr(w)), "last_draw")
        return samples, h, w

    def factor_device_ptr(self, k: int) -> int:
        p = ctypes.c_void_p()
        _lib.check(_lib.load().ftl_als_factor_device(self._fh._h, k, ctypes.byref(p)))
        return int(p.value)

    def stream(self) -> int:
        return int(_lib.load().ftl_als_stream(self._fh._h) or 0)
