"""CPU FP64 oracle for the exact Khatri-Rao leverage sampler and sampled CP-ALS.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline; the product (``fast_tensor_leverage_amd``)
never imports it.

The compiled part (``ftl_oracle.cpp``, ``ftl_oracle_als.cpp``) restates the
reference C++ (cpp_ext/*.hpp of fast_tensor_leverage).  How each function is
pinned (none of these are pinned by running the reference, which cannot be
built here: it needs cblas.h/lapacke.h):

  tree_shape / leaf_idx   closed form + brute-force in-order leaf walk for
                          leaf counts 1..4096 (tests/test_oracle_tree.py)
  build_tree              every node = U[rows]^T U[rows] in NumPy; lower
                          triangle zero (tests/test_oracle_tree.py)
  jacobi_eigh             numpy.linalg.eigh (LAPACK syevd): eigenvalues,
                          |eigenvector| columns, reconstruction
  pt_sample               brute-force global inverse CDF over rows with
                          p_row ~ (U_row . x)^2, index equality wherever the
                          margin exceeds 1e-9 (tests/test_oracle_sampler.py)
  compute_m               numpy pinv with the 1e-8 cut; sum_r E[r]^T E[r] = P_k;
                          last mode rows = eigenvectors / sqrt(lambda) descending
  sampler_draw            exact leverage scores by QR of the materialised KRP:
                          TV distance within the multinomial noise band, and
                          per-draw weights = R / (J * h^T pinv(G) h) in NumPy
  als (ftl_oracle_als)    see the header of ftl_oracle_als.cpp
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libftl_oracle.so")
_lib = None

_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ftlo_tree_shape.argtypes = [_u64, _u64, _vp]
        L.ftlo_tree_shape.restype = ctypes.c_int
        L.ftlo_leaf_idx.argtypes = [_u64, _u64, _u64]
        L.ftlo_leaf_idx.restype = _i64
        L.ftlo_build_tree.argtypes = [_vp, _u64, _u64, _u64, _vp]
        L.ftlo_pt_sample.argtypes = [_vp, _u64, _u64, _u64, _vp, _vp, _vp, _vp, _vp, _u64, _vp]
        L.ftlo_jacobi_eigh.argtypes = [_u64, _vp, _vp, _vp]
        L.ftlo_sampler_create.argtypes = [_vp, _vp, _u64, _u64, _u64]
        L.ftlo_sampler_create.restype = _vp
        L.ftlo_sampler_update.argtypes = [_vp, _u64, _vp]
        L.ftlo_sampler_destroy.argtypes = [_vp]
        L.ftlo_sampler_root.argtypes = [_vp, _u64, _vp]
        L.ftlo_sampler_tree.argtypes = [_vp, _u64, _vp]
        L.ftlo_sampler_compute_m.argtypes = [_vp, _u64, _vp, _vp]
        L.ftlo_sampler_draw.argtypes = [_vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp]
        L.ftlo_sampler_draw.restype = ctypes.c_int
        L.ftlo_num_threads.restype = ctypes.c_int
        _als_bind(L)
        _lib = L
    return _lib


def _als_bind(L):
    if not hasattr(L, "ftlo_als_dedup"):
        return
    L.ftlo_als_dedup.argtypes = [_vp, _u64, _u64, _u64, _vp, _vp, _vp, _vp, _vp, _vp]
    L.ftlo_als_dedup.restype = _u64
    L.ftlo_compute_pinv.argtypes = [_vp, _u64, _u64, _vp]
    L.ftlo_compute_pinv_square.argtypes = [_vp, _u64, _u64, _vp]
    L.ftlo_sparse_create.argtypes = [_vp, _vp, _u64, _u64]
    L.ftlo_sparse_create.restype = _vp
    L.ftlo_sparse_destroy.argtypes = [_vp]
    L.ftlo_sparse_normsq.argtypes = [_vp]
    L.ftlo_sparse_normsq.restype = ctypes.c_double
    L.ftlo_sparse_mttkrp.argtypes = [_vp, _u64, _vp, _u64, _vp, _u64, _u64, _vp, _vp]
    L.ftlo_sparse_mttkrp.restype = _u64
    L.ftlo_sparse_residual_normsq.argtypes = [_vp, _vp, _vp, _u64]
    L.ftlo_sparse_residual_normsq.restype = ctypes.c_double
    L.ftlo_renormalize_columns.argtypes = [_vp, _u64, _u64, _vp]
    L.ftlo_als_update.argtypes = [_vp, _vp, _vp, _vp, _u64, _u64, _vp, _vp, _vp]
    L.ftlo_als_update.restype = ctypes.c_int


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    return a


def num_threads() -> int:
    return int(lib().ftlo_num_threads())


# ----------------------------------------------------------------------------
# tree level (partition_tree.hpp / partition_tree.cpp surface)
# ----------------------------------------------------------------------------
def tree_shape(n: int, F: int) -> dict:
    out = np.zeros(7, dtype=np.uint64)
    rc = lib().ftlo_tree_shape(n, F, _p(out))
    if rc != 0:
        raise ValueError("n and F must be positive")
    keys = ["leaf_count", "node_count", "lfill_level", "lfill_count", "nodes_upto_lfill",
            "nodes_before_lfill", "complete_level_offset"]
    return {k: int(v) for k, v in zip(keys, out)}


def leaf_idx(n: int, F: int, c: int) -> int:
    return int(lib().ftlo_leaf_idx(n, F, c))


def build_tree(U: np.ndarray, F: int) -> np.ndarray:
    U = _c(U, np.float64)
    n, R = U.shape
    sh = tree_shape(n, F)
    G = np.empty((sh["node_count"], R, R), dtype=np.float64)
    lib().ftlo_build_tree(_p(U), n, R, F, _p(G))
    return G


def pt_sample(G, U, F, x, h, draws, margins=None):
    """Descent over a prebuilt tree; h is updated in place, returns samples."""
    U = _c(U, np.float64)
    n, R = U.shape
    x = _c(x, np.float64)
    assert h.dtype == np.float64 and h.flags.c_contiguous
    J = x.shape[0]
    draws = _c(draws, np.float64)
    G = _c(G, np.float64)
    samples = np.zeros(J, dtype=np.uint64)
    if margins is not None:
        margins[:] = np.inf
    lib().ftlo_pt_sample(_p(G), n, F, R, _p(U), _p(x), _p(h), _p(samples), _p(draws), J,
                         _p(margins))
    return samples


def jacobi_eigh(A):
    A = _c(A, np.float64)
    R = A.shape[0]
    lam = np.empty(R)
    V = np.empty((R, R))
    lib().ftlo_jacobi_eigh(R, _p(A), _p(lam), _p(V))
    return lam, V


# ----------------------------------------------------------------------------
# sampler level (EfficientKRPSampler surface)
# ----------------------------------------------------------------------------
class OracleSampler:
    """EfficientKRPSampler restated on the CPU (efficient_krp_sampler.hpp:16-275)."""

    def __init__(self, U_list, J: int, R: int):
        self.U = [_c(U, np.float64) for U in U_list]
        for U in self.U:
            if U.ndim != 2 or U.shape[1] != R:
                raise ValueError("factor matrices must be I_k x R")
        self.N = len(self.U)
        self.J = int(J)
        self.R = int(R)
        ptrs = (ctypes.c_void_p * self.N)(*[U.ctypes.data for U in self.U])
        dims = np.array([U.shape[0] for U in self.U], dtype=np.uint64)
        self._ptrs = ptrs
        self._h = lib().ftlo_sampler_create(ptrs, _p(dims), self.N, self.J, self.R)
        self.h = np.ones((self.J, self.R))
        self.weights = np.zeros(self.J)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ftlo_sampler_destroy(h)
            self._h = None

    def update_sampler(self, j: int, Uj=None):
        if Uj is not None:
            self.U[j] = _c(Uj, np.float64)
        lib().ftlo_sampler_update(self._h, j, _p(self.U[j]))

    def root(self, k):
        out = np.empty((self.R, self.R))
        lib().ftlo_sampler_root(self._h, k, _p(out))
        return out

    def tree(self, k):
        n = self.U[k].shape[0]
        F = min(self.R, n)
        sh = tree_shape(n, F)
        out = np.empty((sh["node_count"], self.R, self.R))
        lib().ftlo_sampler_tree(self._h, k, _p(out))
        return out

    def compute_m(self, j):
        pinv = np.empty((self.R, self.R))
        E = np.empty((self.N, self.R, self.R))
        lib().ftlo_sampler_compute_m(self._h, j, _p(pinv), _p(E))
        return pinv, E

    def draw(self, j, draws, samples=None, want_margins=True, want_eig=False):
        """KRPDrawSamples(j, samples, draws) plus h, weights and margins."""
        N, J = self.N, self.J
        draws = _c(draws, np.float64)
        if draws.size != 2 * N * J:
            raise ValueError("draws must hold 2*N*J uniforms")
        if samples is None:
            samples = np.zeros((N, J), dtype=np.uint64)
        margins = np.empty(J) if want_margins else None
        eig = np.zeros((N, J), dtype=np.uint64) if want_eig else None
        lib().ftlo_sampler_draw(self._h, j, _p(draws), _p(samples), _p(self.h), _p(self.weights),
                                _p(margins), _p(eig))
        out = {"samples": samples, "h": self.h.copy(), "weights": self.weights.copy(),
               "margins": margins}
        if want_eig:
            out["eig"] = eig
        return out
