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


# ----------------------------------------------------------------------------
# Philox4x32-10 restated in NumPy (Salmon et al., SC'11) to reproduce the
# library's built-in uniforms on the host: uniform(seed, call, sample, stage).
# ----------------------------------------------------------------------------
def philox_uniforms(seed, call, J, stage):
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
    mask = np.uint64(0xFFFFFFFF)
    i = np.arange(J, dtype=np.uint64)
    c0 = (i & mask).astype(np.uint32)
    c1 = ((i >> np.uint64(32)).astype(np.uint32)) ^ np.uint32((stage << 16) & 0xFFFFFFFF)
    c2 = np.full(J, call & 0xFFFFFFFF, dtype=np.uint32)
    c3 = np.full(J, (call >> 32) & 0xFFFFFFFF, dtype=np.uint32)
    k0 = np.uint32(seed & 0xFFFFFFFF)
    k1 = np.uint32((seed >> 32) & 0xFFFFFFFF)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & mask).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & mask).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
    a = (c0 >> np.uint32(5)).astype(np.uint64)
    b = (c1 >> np.uint32(6)).astype(np.uint64)
    return (a * np.uint64(67108864) + b).astype(np.float64) * (1.0 / 9007199254740992.0)


def philox_draws(seed, call, N, J):
    """2*N*J uniforms in the layout of efficient_krp_sampler.hpp:221-229
    (eigen draws of mode k at [kJ,(k+1)J), Gram draws at [NJ + kJ, ...))."""
    out = np.empty(2 * N * J)
    for k in range(N):
        out[k * J:(k + 1) * J] = philox_uniforms(seed, call, J, 2 * k)
        out[N * J + k * J:N * J + (k + 1) * J] = philox_uniforms(seed, call, J, 2 * k + 1)
    return out


# ---------------------------------------------------------------------------
# sampled CP-ALS closed forms
# ---------------------------------------------------------------------------
def cp_dense(U, sigma=None):
    """Dense [[sigma; U_0, ..., U_{N-1}]]."""
    R = U[0].shape[1]
    sigma = np.ones(R) if sigma is None else sigma
    out = krp(U) @ sigma
    return out.reshape([M.shape[0] for M in U])


def random_sparse(dims, nnz, rng, unique=True):
    """nnz x N uint32 coordinates (unique unless asked otherwise) and values."""
    idx = np.stack([rng.integers(0, d, size=4 * nnz) for d in dims], axis=1).astype(np.uint32)
    if unique:
        idx = np.unique(idx, axis=0)
        idx = idx[rng.permutation(idx.shape[0])[:nnz]]
    else:
        idx = idx[:nnz]
    return np.ascontiguousarray(idx), rng.standard_normal(idx.shape[0])


def dense_from_sparse(idx, vals, dims):
    X = np.zeros(dims)
    np.add.at(X, tuple(idx[:, k].astype(np.int64) for k in range(idx.shape[1])), vals)
    return X


def brute_mttkrp(idx, vals, samples, lhs, j, Ij):
    """execute_downsampled_mttkrp by brute force: for each sample (in order)
    every nonzero (ascending position) equal on all modes but j."""
    N = idx.shape[1]
    other = [k for k in range(N) if k != j]
    out = np.zeros((Ij, lhs.shape[1]))
    ids, counts = [], []
    for s in range(samples.shape[0]):
        hit = np.nonzero(np.all(idx[:, other] == samples[s, other].astype(np.uint32), axis=1))[0]
        counts.append(len(hit))
        for z in hit:
            out[idx[z, j]] += lhs[s] * vals[z]
            ids.append(z)
    return out, np.array(counts, dtype=np.uint64), np.array(ids, dtype=np.uint64)


def sampled_ls_solution(X, U, j, samples, h, w):
    """U_j = (sum_s w_s X[.., :, ..](tuple_s) h_s^T) pinv(sum_s w_s h_s h_s^T)
    over the J drawn samples (duplicates included), dense X."""
    N = X.ndim
    J = samples.shape[1]
    Xj = np.moveaxis(X, j, 0)
    other = [k for k in range(N) if k != j]
    cols = Xj[(slice(None),) + tuple(samples[k].astype(np.int64) for k in other)]  # I_j x J
    mt = (cols * w[None, :]) @ h
    G = (h * w[:, None]).T @ h
    return mt @ np.linalg.pinv(G, rcond=1e-13, hermitian=True)


def lowrank_as_sparse(U):
    """The dense CP tensor [[U]] (X) and all its entries as a sparse tensor
    (idx uint32 nnz x N, values)."""
    dims = [M.shape[0] for M in U]
    X = cp_dense(U)
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), axis=-1).reshape(-1, len(dims))
    return X, np.ascontiguousarray(idx.astype(np.uint32)), X.reshape(-1).copy()
