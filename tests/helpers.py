"""Shared NumPy helpers for the tests: closed forms the oracle and the HIP path
are checked against."""
import numpy as np


def krp(U):
    """Materialised Khatri-Rao product, mode 0 most significant
    (compare_distributions.py:13-19)."""
    out = U[0]
    for M in U[1:]:
        out = np.einsum("ir,jr->ijr", out, M).reshape(out.shape[0] * M.shape[0], -1)
    return out


def exact_leverage(U, exclude=None):
    mats = [M for k, M in enumerate(U) if k != exclude]
    A = krp(mats)
    q, _ = np.linalg.qr(A)
    lev = np.sum(q * q, axis=1)
    return lev / lev.sum()


def flat_index(samples, dims, exclude=None):
    """Row index in the KRP of the sampled modes (mode 0 most significant)."""
    modes = [k for k in range(len(dims)) if k != exclude]
    idx = np.zeros(samples.shape[1], dtype=np.int64)
    for k in modes:
        idx = idx * dims[k] + samples[k].astype(np.int64)
    return idx


def tv_distance(p, q):
    return 0.5 * np.abs(np.asarray(p) - np.asarray(q)).sum()


def multinomial_tv_band(p, J, trials=200, seed=0):
    rng = np.random.default_rng(seed)
    tvs = [tv_distance(rng.multinomial(J, p) / J, p) for _ in range(trials)]
    return float(np.mean(tvs)), float(np.std(tvs))


def gram_pinv(U, exclude=None, tol=1e-8):
    R = U[0].shape[1]
    G = np.ones((R, R))
    for k, M in enumerate(U):
        if k != exclude:
            G *= M.T @ M
    lam, V = np.linalg.eigh(G)
    inv = np.where(lam > tol, 1.0 / np.where(lam > tol, lam, 1.0), 0.0)
    return (V * inv) @ V.T, G


def row_cdf_sample(U, x, u):
    """Brute-force inverse CDF over rows with p_row ~ (U_row . x)^2."""
    q = (U @ x) ** 2
    cdf = np.cumsum(q) / q.sum()
    return int(np.searchsorted(cdf, u, side="right"))
