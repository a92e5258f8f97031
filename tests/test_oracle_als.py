"""Pins the oracle's sampled CP-ALS step (als.hpp:57-220, sparse_tensor.hpp:46-119,
sort_lookup.hpp:33-197, common.h:152-323, low_rank_tensor.hpp:203-260) against
brute-force lookups, dense NumPy closed forms and exact recovery of a
noise-free low-rank tensor."""
import numpy as np
import pytest

from tests.helpers import (brute_mttkrp, cp_dense, dense_from_sparse, lowrank_as_sparse, random_sparse,
                           sampled_ls_solution)


@pytest.mark.parametrize("dims,nnz,j,unique", [((5, 6, 7), 80, 1, False), ((9, 4, 11), 150, 0, True),
                                               ((3, 4, 5, 6), 200, 3, False), ((40, 30), 300, 1, True)])
def test_mttkrp_matches_brute_force(oracle, dims, nnz, j, unique):
    rng = np.random.default_rng(sum(dims) + nnz)
    idx, vals = random_sparse(dims, nnz, rng, unique)
    S = oracle.OracleSparse(idx, vals)
    Jp, R = 40, 5
    # half the samples are coordinates of nonzeros (guaranteed hits)
    samples = np.stack([rng.integers(0, d, size=Jp) for d in dims], axis=1).astype(np.uint64)
    samples[: Jp // 2] = idx[rng.integers(0, idx.shape[0], size=Jp // 2)]
    lhs = rng.standard_normal((Jp, R))
    out, found, counts, ids = S.mttkrp(samples, lhs, j, dims[j], want_matches=True)
    ref, rc, rids = brute_mttkrp(idx, vals, samples, lhs, j, dims[j])
    assert found == len(rids) and found > 0
    np.testing.assert_array_equal(counts, rc)
    np.testing.assert_array_equal(ids, rids)  # exact integer equality of matched nonzero ids
    np.testing.assert_array_equal(out, ref)   # same summation order -> bitwise equal


@pytest.mark.parametrize("j", [0, 2])
def test_lookup_sort_of_large_tensors_is_stable(oracle, j):
    """Above 2^20 nonzeros the lookup sort runs in parallel chunks merged
    pairwise; the matched ids of every sample must still be the ascending
    positions of the nonzeros with its coordinates (the stable order), with
    chunk boundaries falling inside runs of equal keys."""
    rng = np.random.default_rng(11 + j)
    dims = (6, 5, 7)
    nnz = 3 * 2 ** 20 + 7
    idx = np.stack([rng.integers(0, d, size=nnz) for d in dims], axis=1).astype(np.uint32)
    vals = rng.standard_normal(nnz)
    S = oracle.OracleSparse(idx, vals)
    samples = np.stack([rng.integers(0, d, size=12) for d in dims], axis=1).astype(np.uint64)
    _, found, counts, ids = S.mttkrp(samples, rng.standard_normal((12, 2)), j, dims[j], want_matches=True)
    other = [k for k in range(3) if k != j]
    pos = 0
    for u in range(12):
        mask = np.ones(nnz, dtype=bool)
        for k in other:
            mask &= idx[:, k] == samples[u, k]
        ref = np.nonzero(mask)[0]
        assert int(counts[u]) == ref.size
        np.testing.assert_array_equal(ids[pos:pos + ref.size], ref)
        pos += ref.size
    assert pos == found


def test_mttkrp_no_matches_and_normsq(oracle):
    rng = np.random.default_rng(3)
    idx = np.array([[0, 0, 0], [1, 2, 3]], dtype=np.uint32)
    vals = np.array([2.0, -3.0])
    S = oracle.OracleSparse(idx, vals)
    assert S.normsq() == 13.0
    out, found = S.mttkrp(np.array([[0, 1, 1]], dtype=np.uint64), rng.standard_normal((1, 4)), 0, 2)
    assert found == 0 and not out.any()


@pytest.mark.parametrize("dims,R", [((6, 5, 4), 3), ((4, 3, 5, 2), 6), ((7, 8), 2)])
def test_residual_and_lowrank_normsq_dense(oracle, dims, R):
    rng = np.random.default_rng(R * 7 + len(dims))
    idx, vals = random_sparse(dims, int(np.prod(dims)) // 3, rng, unique=True)
    U = [rng.standard_normal((d, R)) for d in dims]
    sigma = rng.random(R) + 0.5
    M = cp_dense(U, sigma)
    X = dense_from_sparse(idx, vals, dims)
    S = oracle.OracleSparse(idx, vals)
    np.testing.assert_allclose(oracle.lowrank_normsq(U, sigma), np.sum(M * M), rtol=1e-12)
    np.testing.assert_allclose(S.residual_normsq(sigma, U), np.sum((X - M) ** 2), rtol=1e-11)


@pytest.mark.parametrize("I,R", [(50, 8), (8, 8), (5, 8), (200, 25)])
def test_compute_pinv(oracle, I, R):
    rng = np.random.default_rng(I * R)
    A = rng.standard_normal((I, R))
    P = oracle.compute_pinv(A)
    ref = np.linalg.pinv(A.T @ A, rcond=1e-10, hermitian=True)
    np.testing.assert_allclose(P, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
    np.testing.assert_array_equal(P, P.T)


def test_compute_pinv_square_rank_cut(oracle):
    """Only the top target_rank eigenvalues above 1e-11 are inverted."""
    rng = np.random.default_rng(11)
    Q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    lam = np.array([1e-13, 0.5, 1.0, 2.0, 3.0, 4.0])
    M = (Q * lam) @ Q.T
    P = oracle.compute_pinv_square(np.triu(M), 4)
    keep = [2, 3, 4, 5]
    ref = (Q[:, keep] / lam[keep]) @ Q[:, keep].T
    np.testing.assert_allclose(P, ref, atol=1e-10)
    P5 = oracle.compute_pinv_square(np.triu(M), 6)  # 1e-13 is below the 1e-11 cut
    ref5 = (Q[:, 1:] / lam[1:]) @ Q[:, 1:].T
    np.testing.assert_allclose(P5, ref5, atol=1e-9)


def test_renormalize_columns(oracle):
    rng = np.random.default_rng(2)
    U = rng.standard_normal((30, 5))
    U[:, 2] = 0.0
    U[:, 4] = 1e-9  # norm <= 1e-7 -> kept as is, norm reported 1
    U0 = U.copy()
    norms = oracle.renormalize_columns(U)
    ref = np.linalg.norm(U0, axis=0)
    ref[ref <= 1e-7] = 1.0
    np.testing.assert_allclose(norms, ref, rtol=1e-14)
    np.testing.assert_allclose(U, U0 / ref, rtol=1e-14)


@pytest.mark.parametrize("j", [0, 1, 2])
def test_dedup_matches_numpy_unique(oracle, j):
    rng = np.random.default_rng(j)
    N, J = 3, 500
    samples = rng.integers(0, 4, size=(N, J)).astype(np.uint64)
    samples[j] = rng.integers(0, 1000, size=J)  # mode j is ignored
    first, count, tuples = oracle.als_dedup(samples, j)
    key = samples.T.copy()
    key[:, j] = 0
    u, ifirst, cnt = np.unique(key, axis=0, return_index=True, return_counts=True)
    np.testing.assert_array_equal(tuples, u)
    np.testing.assert_array_equal(count, cnt)
    np.testing.assert_array_equal(first, ifirst)


@pytest.mark.parametrize("j", [0, 1, 2])
def test_als_update_recovers_exact_lowrank(oracle, j):
    """X = [[U_true]] exactly; with the other factors at their true values one
    sampled update returns U_true[j] (sampled design has full column rank)."""
    rng = np.random.default_rng(40 + j)
    dims, R, J = (12, 10, 14), 4, 300
    Ut = [rng.standard_normal((d, R)) for d in dims]
    X, idx, vals = lowrank_as_sparse(Ut)
    S = oracle.OracleSparse(idx, vals)
    U0 = [u.copy() for u in Ut]
    U0[j] = rng.standard_normal((dims[j], R))
    als = oracle.OracleALS(S, U0, J)
    d = als.update(j, rng.random(2 * 3 * J), renormalize=False, update_sampler=True)
    assert d["unique_samples"] >= R and d["found"] == d["unique_samples"] * dims[j]
    np.testing.assert_allclose(als.U[j], Ut[j], rtol=1e-8, atol=1e-9)
    # renormalized: unit columns, sigma = norms, residual 0 (to rounding; the
    # reference's fit takes sqrt of it unclamped, so test the residual)
    als.U[j][:] = U0[j]
    als.sampler.update_sampler(j)
    als.update(j, rng.random(2 * 3 * J), renormalize=True, update_sampler=True)
    norms = np.linalg.norm(Ut[j], axis=0)
    np.testing.assert_allclose(als.sigma, norms, rtol=1e-8)
    np.testing.assert_allclose(als.U[j], Ut[j] / norms, rtol=1e-8, atol=1e-9)
    assert abs(als.residual_normsq()) < 1e-10 * S.normsq()


@pytest.mark.parametrize("dims,R,J,j", [((15, 12, 10), 5, 200, 1), ((8, 9, 7, 6), 3, 150, 2), ((30, 25), 4, 100, 0)])
def test_als_update_is_sampled_least_squares(oracle, dims, R, J, j):
    """U_j after one update equals the weighted sampled least-squares
    solution computed densely from the same samples, h and weights."""
    rng = np.random.default_rng(R * J + j)
    N = len(dims)
    idx, vals = random_sparse(dims, int(np.prod(dims)) // 4, rng, unique=False)
    X = dense_from_sparse(idx, vals, dims)
    U = [rng.standard_normal((d, R)) for d in dims]
    draws = rng.random(2 * N * J)
    ref_s = oracle.OracleSampler([u.copy() for u in U], J, R)
    out = ref_s.draw(j, draws)
    expect = sampled_ls_solution(X, U, j, out["samples"], out["h"], out["weights"])
    als = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U], J)
    als.update(j, draws, renormalize=False, update_sampler=False)
    np.testing.assert_allclose(als.U[j], expect, rtol=1e-8, atol=1e-10 * np.abs(expect).max())


@pytest.mark.parametrize("dims,R,J", [((15, 12, 10), 5, 200), ((8, 9, 7, 6), 3, 150)])
def test_update_with_injected_samples_equals_update(oracle, dims, R, J):
    """update_with_samples on the samples the oracle's own sampler draws
    (h and weights recomputed from the factors and computeM's pinv) gives the
    factors and sigma of update() with the uniforms, over a sweep; a planted
    change of one sample changes the result."""
    rng = np.random.default_rng(R + J)
    N = len(dims)
    idx, vals = random_sparse(dims, int(np.prod(dims)) // 3, rng, unique=True)
    U = [rng.standard_normal((d, R)) for d in dims]
    S = oracle.OracleSparse(idx, vals)
    a = oracle.OracleALS(S, [u.copy() for u in U], J)
    b = oracle.OracleALS(S, [u.copy() for u in U], J)
    for j in range(N):
        draws = rng.random(2 * N * J)
        samples = oracle.OracleSampler([u.copy() for u in b.U], J, R).draw(j, draws)["samples"]
        a.update(j, draws)
        b.update_with_samples(j, samples)
        for k in range(N):
            np.testing.assert_allclose(b.U[k], a.U[k], rtol=1e-10, atol=1e-12 * np.abs(a.U[k]).max())
        np.testing.assert_allclose(b.sigma, a.sigma, rtol=1e-10)
    c = oracle.OracleALS(S, [u.copy() for u in a.U], J)
    d = oracle.OracleALS(S, [u.copy() for u in a.U], J)
    draws = rng.random(2 * N * J)
    samples = oracle.OracleSampler([u.copy() for u in c.U], J, R).draw(0, draws)["samples"]
    c.update_with_samples(0, samples)
    samples[1, 0] = (samples[1, 0] + 1) % dims[1]
    d.update_with_samples(0, samples)
    assert np.abs(c.U[0] - d.U[0]).max() > 1e-8


def test_als_sweeps_raise_fit(oracle):
    """A few sampled sweeps on a noisy low-rank tensor bring the fit near 1."""
    rng = np.random.default_rng(9)
    dims, R, J = (14, 13, 12), 3, 400
    Ut = [rng.standard_normal((d, R)) for d in dims]
    X, idx, vals = lowrank_as_sparse(Ut)
    vals = vals + 0.01 * np.linalg.norm(vals) / np.sqrt(vals.size) * rng.standard_normal(vals.size)
    als = oracle.OracleALS(oracle.OracleSparse(idx, vals), [rng.standard_normal((d, R)) for d in dims], J)
    f0 = als.fit()
    for sweep in range(8):
        for j in range(3):
            als.update(j, rng.random(2 * 3 * J))
    assert als.fit() > 0.95 > f0


# ---- exact (unsampled) update: execute_exact_mttkrp / execute_exact_als_update ----

def _dense_mttkrp(X, U, j):
    """X_(j) times the Khatri-Rao product of the other factors, by einsum."""
    N = X.ndim
    letters = "abcdefgh"
    sub_x = letters[:N]
    subs = [letters[k] + "r" for k in range(N) if k != j]
    ops = [U[k] for k in range(N) if k != j]
    return np.einsum(sub_x + "," + ",".join(subs) + "->" + letters[j] + "r", X, *ops)


@pytest.mark.parametrize("dims,nnz,j", [((12, 9, 7), 300, 0), ((12, 9, 7), 300, 2), ((5, 6, 4, 3), 200, 1),
                                        ((30, 25), 250, 1)])
def test_exact_mttkrp_matches_dense(oracle, dims, nnz, j):
    rng = np.random.default_rng(sum(dims) * 7 + j)
    idx, vals = random_sparse(dims, nnz, rng)
    R = 6
    U = [rng.standard_normal((d, R)) for d in dims]
    S = oracle.OracleSparse(idx, vals)
    out = S.exact_mttkrp(j, U)
    ref = _dense_mttkrp(dense_from_sparse(idx, vals, dims), U, j)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # a row with no nonzeros stays zero
    empty = np.setdiff1d(np.arange(dims[j]), idx[:, j])
    assert not out[empty].any()


def _dense_exact_solution(X, U, j):
    """argmin_V ||X_(j) - V KRP^T||_F by least squares on the unfolding."""
    N = X.ndim
    other = [k for k in range(N) if k != j]
    K = np.ones((1, U[0].shape[1]))
    for k in other:  # rows of the KRP in the C order of the remaining axes
        K = (K[:, None, :] * U[k][None, :, :]).reshape(-1, U[0].shape[1])
    Xj = np.moveaxis(X, j, 0).reshape(X.shape[j], -1)
    sol, *_ = np.linalg.lstsq(K, Xj.T, rcond=None)
    return sol.T


@pytest.mark.parametrize("dims,R,j", [((10, 8, 9), 4, 0), ((10, 8, 9), 4, 2), ((6, 5, 4, 5), 3, 1)])
def test_exact_update_is_dense_least_squares(oracle, dims, R, j):
    rng = np.random.default_rng(R * 31 + j)
    nnz = int(0.6 * np.prod(dims))
    idx, vals = random_sparse(dims, nnz, rng)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    als = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U0], J=16)
    als.exact_update(j, renormalize=False)
    ref = _dense_exact_solution(dense_from_sparse(idx, vals, dims), U0, j)
    np.testing.assert_allclose(als.U[j], ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())
    # with renormalize: unit columns, sigma = the norms
    als2 = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U0], J=16)
    als2.exact_update(j, renormalize=True)
    norms = np.linalg.norm(ref, axis=0)
    np.testing.assert_allclose(als2.sigma, norms, rtol=1e-9)
    np.testing.assert_allclose(als2.U[j], ref / norms, rtol=1e-9, atol=1e-11)
    for k in range(len(dims)):
        if k != j:
            np.testing.assert_array_equal(als2.U[k], U0[k])


def test_exact_update_recovers_lowrank_and_fit_is_monotone(oracle):
    rng = np.random.default_rng(5)
    dims, R = (9, 8, 7), 3
    Ut = [rng.standard_normal((d, R)) for d in dims]
    X = cp_dense(Ut)
    idx = np.argwhere(np.ones(dims, dtype=bool)).astype(np.uint32)
    vals = X[tuple(idx.T)]
    S = oracle.OracleSparse(idx, vals)
    U0 = [u.copy() for u in Ut]
    U0[1] = rng.standard_normal((dims[1], R))
    als = oracle.OracleALS(S, U0, J=16)
    als.exact_update(1, renormalize=True)
    # the fully observed noise-free tensor is reproduced exactly
    M = cp_dense(als.U, als.sigma)
    np.testing.assert_allclose(M, X, rtol=1e-9, atol=1e-10 * np.abs(X).max())
    # noisy data: exact ALS never lowers the fit, and reaches the noise
    # floor 1 - ||noise|| / ||X + noise|| (the rank-3 fit can only be better)
    noise = 0.1 * rng.standard_normal(vals.shape)
    noisy = vals + noise
    als = oracle.OracleALS(oracle.OracleSparse(idx, noisy), [rng.standard_normal((d, R)) for d in dims], J=16)
    fits = []
    for sweep in range(15):
        for j in range(3):
            als.exact_update(j, renormalize=True)
            fits.append(als.fit())
    assert all(b >= a - 1e-12 for a, b in zip(fits, fits[1:])), fits
    floor = 1 - np.linalg.norm(noise) / np.linalg.norm(noisy)
    assert floor < fits[-1] < floor + 0.02
