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
