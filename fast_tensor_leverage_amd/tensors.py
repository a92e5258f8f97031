"""PyLowRank and PySparseTensor: the Python-level tensor wrappers of the
reference (tensors.py:16-120) over the MI355X classes of this package."""
from __future__ import annotations

import numpy as np

from .als import LowRankTensor, SparseTensor


class PyLowRank:
    """Random Gaussian CP factors ``U_k = rng.normal(size=(I_k, R))`` and their
    LowRankTensor (tensors.py:16-34)."""

    def __init__(self, dims, R, allow_rhs_mttkrp=False, init_method="gaussian", seed=None):
        if init_method != "gaussian":
            raise ValueError(f"unknown init_method '{init_method}'")
        rng = np.random.default_rng() if seed is None else np.random.default_rng(seed)
        self.dims = [int(i) for i in dims]
        self.N = len(self.dims)
        self.R = int(R)
        self.U = [rng.normal(size=(i, self.R)) for i in self.dims]
        self.ten = LowRankTensor(self.R, 10000, self.U) if allow_rhs_mttkrp else LowRankTensor(self.R, self.U)

    def set_sigma(self, new_sigma):
        self.ten.set_sigma(new_sigma)

    def compute_diff_resid(self, rhs) -> float:
        """||rhs - [[sigma; U]]|| (tensors.py:36-44).  When an ALS on rhs holds
        this decomposition, the residual is taken from its device factors."""
        als = self.ten._als
        if als is not None and als.ground_truth is rhs.ten:
            return float(np.sqrt(als.residual_normsq()))
        sigma_lhs = np.zeros(self.R, dtype=np.double)
        self.ten.get_sigma(sigma_lhs, -1)
        return float(np.sqrt(rhs.ten.compute_residual_normsq(sigma_lhs, self.U)))

    def compute_exact_fit(self, rhs) -> float:
        """1 - ||rhs - M|| / ||rhs|| (tensors.py:46-49).  As in the reference
        the residual is not clamped at zero (a rounding-negative residual
        gives NaN)."""
        diff_norm = self.compute_diff_resid(rhs)
        rhs_norm = np.sqrt(rhs.ten.get_normsq())
        return 1.0 - diff_norm / rhs_norm


class PySparseTensor:
    """Sparse tensor from coordinate arrays (the in-memory form of
    tensors.py:79-113; the reference reads MODE_i / VALUES / MIN_MODE_SET /
    MAX_MODE_SET from HDF5, which needs h5py)."""

    def __init__(self, filename, lookup="sort", preprocessing=None):
        try:
            import h5py  # noqa: F401
        except ImportError as e:
            raise ImportError("reading HDF5 tensors needs h5py; use PySparseTensor.from_arrays") from e
        with h5py.File(filename, "r") as f:
            max_idxs = f["MAX_MODE_SET"][:]
            min_idxs = f["MIN_MODE_SET"][:]
            N = len(max_idxs)
            nnz = len(f["MODE_0"])
            idx = np.zeros((nnz, N), dtype=np.uint32)
            for i in range(N):
                idx[:, i] = f[f"MODE_{i}"][:] - min_idxs[i]
            values = np.ascontiguousarray(f["VALUES"][:], dtype=np.float64)
        self._init(idx, values, lookup, preprocessing, dims=[int(d) for d in max_idxs])

    @classmethod
    def from_npz(cls, filename, lookup="sort", preprocessing=None):
        """The datasets of the reference's HDF5 tensor files (MODE_0 ..
        MODE_{N-1}, VALUES, MIN_MODE_SET, MAX_MODE_SET, tensors.py:85-100) stored
        in a NumPy .npz (e.g. ``numpy.savez(out, **{k: h5[k][:] for k in h5})``
        where h5py is available)."""
        idx, values, dims = read_npz(filename)
        return cls.from_arrays(idx, values, lookup, preprocessing, dims)

    @classmethod
    def from_tns(cls, filename, lookup="sort", preprocessing=None, one_based=True):
        """A FROSTT .tns text file: one nonzero per line, N coordinates then
        the value, whitespace separated, 1-based coordinates ('#' comments)."""
        idx, values, dims = read_tns(filename, one_based)
        return cls.from_arrays(idx, values, lookup, preprocessing, dims)

    @classmethod
    def from_arrays(cls, idx, values, lookup="sort", preprocessing=None, dims=None):
        obj = cls.__new__(cls)
        obj._init(np.ascontiguousarray(idx, dtype=np.uint32), np.ascontiguousarray(values, dtype=np.float64),
                  lookup, preprocessing, dims)
        return obj

    def _init(self, idx, values, lookup, preprocessing, dims):
        if preprocessing is not None:
            if preprocessing == "log_count":
                values = np.log(values + 1.0)
            else:
                raise ValueError(f"unknown preprocessing option '{preprocessing}'")
        self.tensor_idxs = idx
        self.values = values
        self.nnz, self.N = idx.shape
        if dims is None:
            dims = [int(idx[:, k].max()) + 1 for k in range(self.N)]
        self.dims = list(dims)
        self.max_idxs = np.array(self.dims)
        self.min_idxs = np.zeros(self.N, dtype=np.int64)
        self.ten = SparseTensor(idx, values, lookup, dims=[max(int(d), int(idx[:, k].max()) + 1)
                                                           for k, d in enumerate(self.dims)])


def read_npz(filename):
    """(idx nnz x N uint32, values float64, dims) from an .npz with the
    reference's HDF5 dataset names: coordinates minus MIN_MODE_SET, dims =
    MAX_MODE_SET (tensors.py:88-100)."""
    with np.load(filename, allow_pickle=False) as f:
        max_idxs = np.asarray(f["MAX_MODE_SET"])
        min_idxs = np.asarray(f["MIN_MODE_SET"])
        N = len(max_idxs)
        nnz = len(f["MODE_0"])
        idx = np.zeros((nnz, N), dtype=np.uint32)
        for i in range(N):
            idx[:, i] = np.asarray(f[f"MODE_{i}"]) - min_idxs[i]
        values = np.ascontiguousarray(f["VALUES"], dtype=np.float64)
    return idx, values, [int(d) for d in max_idxs]


def read_tns(filename, one_based=True):
    """(idx nnz x N uint32, values float64, dims) from FROSTT .tns text; dims =
    the largest coordinate + 1 of each mode."""
    import pandas as pd
    df = pd.read_csv(filename, sep=r"\s+", header=None, comment="#", engine="c", float_precision="round_trip")
    a = df.to_numpy()
    if a.ndim != 2 or a.shape[1] < 2:
        raise ValueError(f"{filename}: expected N coordinates and a value per line")
    coords = a[:, :-1].astype(np.int64) - (1 if one_based else 0)
    if coords.size and coords.min() < 0:
        raise ValueError(f"{filename}: coordinate below {1 if one_based else 0}")
    idx = np.ascontiguousarray(coords, dtype=np.uint32)
    values = np.ascontiguousarray(a[:, -1], dtype=np.float64)
    dims = [int(idx[:, k].max()) + 1 if idx.shape[0] else 1 for k in range(idx.shape[1])]
    return idx, values, dims
