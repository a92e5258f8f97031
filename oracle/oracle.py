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
  sparse mttkrp           brute-force scan of every nonzero: matched nonzero
                          ids equal as integers, result bitwise equal
                          (tests/test_oracle_als.py)
  residual / lowrank      dense reconstruction [[sigma; U]] on small tensors
  normsq
  compute_pinv(_square)   numpy pinv; explicit eigenbasis with the 1e-11 cut
                          and the target-rank truncation
  renormalize_columns     numpy column norms with the 1e-7 floor
  als_dedup               numpy.unique(axis=0) with first index and counts
  als_update              (1) exact recovery of a noise-free low-rank factor,
                          (2) the weighted sampled least-squares solution
                          computed densely from the same samples, h, weights,
                          (3) sweeps drive the fit of a noisy tensor to ~1
  exact_mttkrp            the dense MTTKRP X_(j) (KRP of the other factors)
                          of the densified tensor (numpy einsum)
  lk_update / leverage    leverage = squared row norms of the thin QR factor
  (Larsen-Kolda)          (numpy.linalg.qr) for full-rank factors; the rank
                          deficient case against the projector of the
                          numpy pinv; total = rank (tests/test_oracle_lk.py)
  lk_draw                 row = searchsorted(cumsum(lev), u * total,
                          'right') in NumPy; weights = 1 / (J prod_k p_k) with
                          p_k = lev_k[row] / total_k; h = product of the
                          sampled rows; empirical row frequencies against
                          lev / total (chi-square)
  als_update_drawn        same checks as als_update, fed by either sampler
  als_exact_update        the dense least-squares solution argmin ||X_(j) -
                          U_j KRP^T|| (numpy lstsq); exact recovery of a
                          noise-free factor; the fit never decreases over a
                          sweep (ALS is monotone on the full problem)
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
    L.ftlo_sparse_mttkrp.argtypes = [_vp, _u64, _vp, _u64, _vp, _u64, _u64, _vp, _vp, _vp, _u64]
    L.ftlo_sparse_mttkrp.restype = _u64
    L.ftlo_sparse_residual_normsq.argtypes = [_vp, _vp, _vp, _vp, _u64]
    L.ftlo_sparse_residual_normsq.restype = ctypes.c_double
    L.ftlo_lowrank_normsq.argtypes = [_vp, _vp, _u64, _u64, _vp]
    L.ftlo_lowrank_normsq.restype = ctypes.c_double
    L.ftlo_renormalize_columns.argtypes = [_vp, _u64, _u64, _vp]
    L.ftlo_als_update.argtypes = [_vp, _vp, _vp, _vp, _u64, _u64, _u64, _u64, _vp, ctypes.c_int, ctypes.c_int,
                                  _vp, _vp]
    L.ftlo_als_update.restype = ctypes.c_int
    L.ftlo_sparse_exact_mttkrp.argtypes = [_vp, _u64, _vp, _u64, _u64, _vp]
    L.ftlo_als_exact_update.argtypes = [_vp, _vp, _vp, _u64, _u64, _u64, ctypes.c_int, _vp]
    L.ftlo_als_exact_update.restype = ctypes.c_int
    L.ftlo_als_update_drawn.argtypes = [_vp, _vp, _vp, _u64, _u64, _u64, _u64, _vp, _vp, _vp, ctypes.c_int, _vp,
                                        _vp]
    L.ftlo_als_update_drawn.restype = ctypes.c_int
    L.ftlo_lk_create.argtypes = [_vp, _vp, _u64, _u64, _u64]
    L.ftlo_lk_create.restype = _vp
    L.ftlo_lk_create2.argtypes = [_vp, _vp, _u64, _u64, _u64, ctypes.c_int]
    L.ftlo_lk_create2.restype = _vp
    L.ftlo_lkh_draw.argtypes = [_vp, _u64, _vp, _u64, _vp, _vp, _vp, _vp]
    L.ftlo_lkh_draw.restype = ctypes.c_int
    L.ftlo_lk_alpha.argtypes = [_vp, _vp]
    L.ftlo_lk_destroy.argtypes = [_vp]
    L.ftlo_lk_update.argtypes = [_vp, _u64, _vp]
    L.ftlo_lk_leverage.argtypes = [_vp, _u64, _vp, _vp]
    L.ftlo_lk_draw.argtypes = [_vp, _u64, _vp, _vp, _vp, _vp, _vp]
    L.ftlo_lk_draw.restype = ctypes.c_int


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
        # held by the object: module globals may be gone when it is collected at exit
        self._destroy = lib().ftlo_sampler_destroy
        self.h = np.ones((self.J, self.R))
        self.weights = np.zeros(self.J)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
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
        """computeM for draw j: (pinv, E); pinv holds the upper triangle
        (the lower one is zero), E the N eigen-row matrices."""
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


# ----------------------------------------------------------------------------
# sampled CP-ALS level (SparseTensor / LowRankTensor / ALS surface)
# ----------------------------------------------------------------------------
def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class OracleSparse:
    """SparseTensor(idx, values, "sort") restated (sparse_tensor.hpp:46-119,
    sort_lookup.hpp:33-197).  idx is nnz x N uint32, values nnz float64."""

    def __init__(self, idx, values):
        self.idx = _c(idx, np.uint32)
        self.values = _c(values, np.float64)
        if self.idx.ndim != 2 or self.values.shape != (self.idx.shape[0],):
            raise ValueError("idx must be nnz x N and values nnz")
        self.nnz, self.N = self.idx.shape
        self.dims = [int(self.idx[:, k].max()) + 1 if self.nnz else 1 for k in range(self.N)]
        self._h = lib().ftlo_sparse_create(_p(self.idx), _p(self.values), self.nnz, self.N)
        # held by the object: module globals may be gone when it is collected at exit
        self._destroy = lib().ftlo_sparse_destroy

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def normsq(self) -> float:
        return float(lib().ftlo_sparse_normsq(self._h))

    def mttkrp(self, samples, lhs, j, Ij, want_matches=False):
        """execute_downsampled_mttkrp: samples Jp x N, lhs Jp x R -> Ij x R."""
        samples = _c(samples, np.uint64)
        lhs = _c(lhs, np.float64)
        Jp, R = lhs.shape
        if samples.shape != (Jp, self.N):
            raise ValueError("samples must be Jp x N")
        out = np.empty((Ij, R))
        counts = np.zeros(Jp, dtype=np.uint64) if want_matches else None
        cap = 0
        ids = None
        if want_matches:
            cap = int(lib().ftlo_sparse_mttkrp(self._h, j, _p(samples), Jp, _p(lhs), R, Ij, _p(out),
                                               None, None, 0))
            ids = np.zeros(max(cap, 1), dtype=np.uint64)
        found = int(lib().ftlo_sparse_mttkrp(self._h, j, _p(samples), Jp, _p(lhs), R, Ij, _p(out),
                                             _p(counts), _p(ids), cap))
        if want_matches:
            return out, found, counts, ids[:found]
        return out, found

    def residual_normsq(self, sigma, U_list) -> float:
        U = [_c(u, np.float64) for u in U_list]
        sigma = _c(sigma, np.float64)
        dims = np.array([u.shape[0] for u in U], dtype=np.uint64)
        return float(lib().ftlo_sparse_residual_normsq(self._h, _p(sigma), _ptrs(U), _p(dims), U[0].shape[1]))

    def exact_mttkrp(self, j, U_list):
        """execute_exact_mttkrp(U, j) (sparse_tensor.hpp:79-82): Ij x R with
        Ij = U_list[j].shape[0]."""
        U = [_c(u, np.float64) for u in U_list]
        R = U[0].shape[1]
        Ij = U[j].shape[0]
        out = np.empty((Ij, R))
        lib().ftlo_sparse_exact_mttkrp(self._h, j, _ptrs(U), R, Ij, _p(out))
        return out


def lowrank_normsq(U_list, sigma) -> float:
    U = [_c(u, np.float64) for u in U_list]
    sigma = _c(sigma, np.float64)
    dims = np.array([u.shape[0] for u in U], dtype=np.uint64)
    return float(lib().ftlo_lowrank_normsq(_ptrs(U), _p(dims), len(U), U[0].shape[1], _p(sigma)))


def compute_pinv(A):
    """compute_pinv (common.h:271-323): pinv(A^T A), target rank min(I, R)."""
    A = _c(A, np.float64)
    I, R = A.shape
    out = np.empty((R, R))
    lib().ftlo_compute_pinv(_p(A), I, R, _p(out))
    return out


def compute_pinv_square(M, target_rank):
    M = _c(M, np.float64)
    R = M.shape[0]
    out = np.empty((R, R))
    lib().ftlo_compute_pinv_square(_p(M), R, target_rank, _p(out))
    return out


def renormalize_columns(U):
    """In place; returns the column norms (renormalize_columns(j))."""
    assert U.dtype == np.float64 and U.flags.c_contiguous
    norms = np.empty(U.shape[1])
    lib().ftlo_renormalize_columns(_p(U), U.shape[0], U.shape[1], _p(norms))
    return norms


def als_dedup(samples, j):
    """Sort/unique of execute_ds_als_update: samples N x J -> (first, count,
    tuples J' x N with column j zero) in ascending lexicographic order."""
    samples = _c(samples, np.uint64)
    N, J = samples.shape
    first = np.empty(J, dtype=np.uint64)
    count = np.empty(J, dtype=np.uint64)
    tuples = np.empty((J, N), dtype=np.uint64)
    u = int(lib().ftlo_als_dedup(_p(samples), N, J, j, _p(first), _p(count), _p(tuples), None, None, None))
    return first[:u], count[:u], tuples[:u]


class OracleLKSampler:
    """LarsenKoldaSampler (larsen_kolda_sampler.hpp:18-136) with injected
    uniforms: mode k of sample i uses draws[k*J + i].  The factors are
    copied into ``self.U`` and aliased by the C restatement."""

    def __init__(self, U_list, J: int, R: int, hybrid: bool = False):
        self.U = [_c(u, np.float64).copy() for u in U_list]
        self.N, self.J, self.R = len(self.U), int(J), int(R)
        self.hybrid = bool(hybrid)
        self._I = np.array([u.shape[0] for u in self.U], dtype=np.uint64)
        self._h = lib().ftlo_lk_create2(_ptrs(self.U), _p(self._I), self.N, self.J, self.R, int(self.hybrid))
        # held by the object: module globals may be gone when it is collected at exit
        self._destroy = lib().ftlo_lk_destroy

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def update_sampler(self, j: int, Uj=None):
        if Uj is not None:
            self.U[j] = _c(Uj, np.float64).copy()
        lib().ftlo_lk_update(self._h, j, _p(self.U[j]))

    def leverage(self, k: int):
        lev = np.empty(self.U[k].shape[0])
        tot = np.zeros(1)
        lib().ftlo_lk_leverage(self._h, k, _p(lev), _p(tot))
        return lev, float(tot[0])

    def alpha(self):
        out = np.empty(self.N)
        lib().ftlo_lk_alpha(self._h, _p(out))
        return out

    def draw_hybrid(self, j, utab):
        """LarsenKoldaHybrid::KRPDrawSamples (larsen_kolda_hybrid.hpp:139-246):
        utab (A x N x J) holds the uniform of attempt a, mode k, sample i."""
        if not self.hybrid:
            raise ValueError("not a hybrid sampler")
        utab = _c(utab, np.float64)
        A = utab.shape[0]
        assert utab.shape == (A, self.N, self.J)
        samples = np.zeros((self.N, self.J), dtype=np.uint64)
        h = np.empty((self.J, self.R))
        w = np.empty(self.J)
        info = np.zeros(8)
        rc = lib().ftlo_lkh_draw(self._h, j, _p(utab), A, _p(samples), _p(h), _p(w), _p(info))
        if rc != 0:
            raise RuntimeError(f"oracle hybrid draw failed ({rc})")
        return {"samples": samples, "h": h, "weights": w, "deterministic": int(info[0]), "p_det": info[1],
                "candidates": info[2], "skipped": bool(info[3]), "cut_margin": info[4], "keep_margin": info[5],
                "draw_margin": info[6], "max_attempts": int(info[7])}

    def draw(self, j, draws, samples=None):
        draws = _c(draws, np.float64)
        if draws.size < self.N * self.J:
            raise ValueError("draws must hold at least N*J uniforms")
        samples = np.zeros((self.N, self.J), dtype=np.uint64) if samples is None else samples
        h = np.empty((self.J, self.R))
        w = np.empty(self.J)
        m = np.empty(self.J)
        lib().ftlo_lk_draw(self._h, j, _p(draws), _p(samples), _p(h), _p(w), _p(m))
        return {"samples": samples, "h": h, "weights": w, "margins": m}


class OracleLKALS:
    """ALS(LowRankTensor(R, U), SparseTensor) + initialize_ds_als(J,
    "larsen_kolda") restated: the Larsen-Kolda draw followed by the same
    update as the efficient sampler's (als.hpp:57-220)."""

    def __init__(self, sparse: OracleSparse, U_list, J: int, sigma=None, hybrid: bool = False):
        R = U_list[0].shape[1]
        self.X = sparse
        self.sampler = OracleLKSampler(U_list, J, R, hybrid=hybrid)
        self.U = self.sampler.U
        self.N, self.J, self.R = len(self.U), int(J), R
        self.sigma = np.ones(R) if sigma is None else _c(sigma, np.float64).copy()

    def update(self, j, draws, renormalize=True, update_sampler=True, utab=None):
        """draws: the Larsen-Kolda uniforms; for the hybrid sampler utab
        (A x N x J) gives the uniforms of every attempt instead."""
        d = self.sampler.draw_hybrid(j, utab) if self.sampler.hybrid else self.sampler.draw(j, draws)
        dims = np.array([u.shape[0] for u in self.U], dtype=np.uint64)
        diag = np.zeros(4)
        rc = lib().ftlo_als_update_drawn(self.X._h, _ptrs(self.U), _p(dims), self.N, self.J, self.R, j,
                                         _p(d["samples"]), _p(d["h"]), _p(d["weights"]), int(renormalize),
                                         _p(self.sigma), _p(diag))
        if rc != 0:
            raise RuntimeError("oracle ALS update failed")
        if update_sampler:
            self.sampler.update_sampler(j)
        return {"unique_samples": int(diag[0]), "found": int(diag[1]), "margins": d.get("margins"), "draw": d}

    def residual_normsq(self) -> float:
        return self.X.residual_normsq(self.sigma, self.U)

    def fit(self) -> float:
        return 1.0 - np.sqrt(self.residual_normsq()) / np.sqrt(self.X.normsq())


class OracleALS:
    """ALS(LowRankTensor(R, U), SparseTensor) + initialize_ds_als(J,
    "efficient") restated (als.hpp:31-220).  The factors live in
    ``self.U`` (updated in place), ``self.sigma`` holds the column norms of
    the last renormalized mode."""

    def __init__(self, sparse: OracleSparse, U_list, J: int, sigma=None):
        R = U_list[0].shape[1]
        self.X = sparse
        self.sampler = OracleSampler(U_list, J, R)
        self.U = self.sampler.U
        self.N, self.J, self.R = len(self.U), int(J), R
        self.sigma = np.ones(R) if sigma is None else _c(sigma, np.float64).copy()

    def update(self, j, draws, renormalize=True, update_sampler=True):
        draws = _c(draws, np.float64)
        if draws.size != 2 * self.N * self.J:
            raise ValueError("draws must hold 2*N*J uniforms")
        dims = np.array([u.shape[0] for u in self.U], dtype=np.uint64)
        diag = np.zeros(4)
        rc = lib().ftlo_als_update(self.sampler._h, self.X._h, _ptrs(self.U), _p(dims), self.N, self.J, self.R, j,
                                   _p(draws), int(renormalize), int(update_sampler), _p(self.sigma), _p(diag))
        if rc != 0:
            raise RuntimeError("oracle ALS update failed")
        return {"unique_samples": int(diag[0]), "found": int(diag[1])}

    def update_with_samples(self, j, samples, renormalize=True, update_sampler=True):
        """The update of execute_ds_als_update (als.hpp:76-220) on given
        sample indices (N x J; row j ignored) instead of this sampler's draw:
        h = the product of the sampled rows of the other factors and weights
        R / (J h^T pinv h) with pinv from this sampler's computeM
        (efficient_krp_sampler.hpp:90-150, 257-267), then dedup, pinv, MTTKRP,
        solve and renormalize as in ``update``."""
        samples = _c(samples, np.uint64)
        N, J, R = self.N, self.J, self.R
        if samples.shape != (N, J):
            raise ValueError("samples must be N x J")
        h = np.ones((J, R))
        for k in range(N):
            if k != j:
                h *= self.U[k][samples[k].astype(np.int64)]
        up, _ = self.sampler.compute_m(j)  # upper triangle of pinv
        pinv = np.triu(up) + np.triu(up, 1).T
        w = R / (J * np.einsum("ir,rs,is->i", h, pinv, h))
        dims = np.array([u.shape[0] for u in self.U], dtype=np.uint64)
        diag = np.zeros(4)
        rc = lib().ftlo_als_update_drawn(self.X._h, _ptrs(self.U), _p(dims), N, J, R, j, _p(samples), _p(h),
                                         _p(w), int(renormalize), _p(self.sigma), _p(diag))
        if rc != 0:
            raise RuntimeError("oracle ALS update failed")
        if update_sampler:
            self.sampler.update_sampler(j)
        return {"unique_samples": int(diag[0]), "found": int(diag[1])}

    def exact_update(self, j, renormalize=True, update_sampler=True):
        """execute_exact_als_update (als.hpp:222-270).  update_sampler is
        accepted and unused, as in the reference."""
        dims = np.array([u.shape[0] for u in self.U], dtype=np.uint64)
        rc = lib().ftlo_als_exact_update(self.X._h, _ptrs(self.U), _p(dims), self.N, self.R, j, int(renormalize),
                                         _p(self.sigma))
        if rc != 0:
            raise RuntimeError("oracle exact ALS update failed")

    def residual_normsq(self) -> float:
        return self.X.residual_normsq(self.sigma, self.U)

    def fit(self) -> float:
        """compute_exact_fit (tensors.py:41-53): 1 - sqrt(res) / ||X||."""
        return 1.0 - np.sqrt(self.residual_normsq()) / np.sqrt(self.X.normsq())
