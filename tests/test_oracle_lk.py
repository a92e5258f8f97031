"""Pins the oracle's Larsen-Kolda sampler (larsen_kolda_sampler.hpp:18-136,
compute_DAGAT sampler.hpp:94-122, fill_h_by_samples sampler.hpp:74-89):
leverage scores against the thin QR factor and the pinv projector, the draw
against a NumPy inverse CDF, the weights against 1 / (J prod p), and the
ALS step it drives against the dense sampled least-squares solution."""
import numpy as np
import pytest

from tests.helpers import dense_from_sparse, lowrank_as_sparse, random_sparse, sampled_ls_solution


@pytest.mark.parametrize("I,R", [(50, 8), (200, 25), (8, 8), (1000, 3)])
def test_leverage_is_squared_qr_row_norms(oracle, I, R):
    rng = np.random.default_rng(I + R)
    U = [rng.standard_normal((I, R)), rng.standard_normal((I + 3, R))]
    s = oracle.OracleLKSampler(U, 10, R)
    for k in range(2):
        lev, tot = s.leverage(k)
        Q, _ = np.linalg.qr(U[k])
        ref = np.sum(Q * Q, axis=1)
        np.testing.assert_allclose(lev, ref, rtol=1e-9, atol=1e-12)
        assert abs(tot - R) < 1e-9 * R  # the trace of a rank-R projector
        assert np.all(lev >= -1e-14) and np.all(lev <= 1 + 1e-12)


def test_leverage_rank_deficient_and_short_factors(oracle):
    rng = np.random.default_rng(5)
    # rank 3 factor with R = 6: leverage of the 3-dim column space
    A = rng.standard_normal((40, 3)) @ rng.standard_normal((3, 6))
    # fewer rows than columns: target rank min(I, R) = I, every row leverage 1
    B = rng.standard_normal((4, 6))
    s = oracle.OracleLKSampler([A, B], 10, 6)
    lev, tot = s.leverage(0)
    Ua, sv, _ = np.linalg.svd(A, full_matrices=False)
    basis = Ua[:, sv > 1e-8 * sv[0]]
    np.testing.assert_allclose(lev, np.sum(basis * basis, axis=1), rtol=1e-7, atol=1e-10)
    assert abs(tot - 3) < 1e-7
    lev, tot = s.leverage(1)
    np.testing.assert_allclose(lev, np.ones(4), rtol=1e-9)
    assert abs(tot - 4) < 1e-9


@pytest.mark.parametrize("dims,R,J,j", [((30, 20, 25), 5, 2000, 1), ((12, 9, 7, 5), 3, 1000, 3),
                                        ((40, 30), 6, 500, 0), ((15, 16, 17), 4, 800, 4)])
def test_draw_is_inverse_cdf_with_product_weights(oracle, dims, R, J, j):
    rng = np.random.default_rng(sum(dims) + J)
    N = len(dims)
    U = [rng.standard_normal((d, R)) for d in dims]
    s = oracle.OracleLKSampler(U, J, R)
    draws = rng.random(2 * N * J)
    out = s.draw(j, draws)
    logp = np.zeros(J)
    h = np.ones((J, R))
    for k in range(N):
        if k == j:
            assert not out["samples"][k].any()  # mode j is not sampled
            continue
        lev, tot = s.leverage(k)
        cdf = np.cumsum(lev)
        row = np.minimum(np.searchsorted(cdf, draws[k * J:(k + 1) * J] * cdf[-1], side="right"), dims[k] - 1)
        sure = out["margins"] > 1e-12
        np.testing.assert_array_equal(out["samples"][k][sure], row[sure])
        logp += np.log(lev[out["samples"][k].astype(np.int64)] / tot)
        h = h * U[k][out["samples"][k].astype(np.int64)]
    np.testing.assert_allclose(out["weights"], 1.0 / (J * np.exp(logp)), rtol=1e-11)
    np.testing.assert_array_equal(out["h"], h)  # same product order -> bitwise


def test_row_frequencies_follow_leverage(oracle):
    """Empirical row counts of 200k draws against lev / total (chi-square,
    one mode at a time; the modes are drawn independently)."""
    rng = np.random.default_rng(11)
    dims, R, J = (60, 45), 5, 200000
    U = [rng.standard_normal((d, R)) * rng.random((d, 1)) ** 2 for d in dims]  # skewed leverage
    s = oracle.OracleLKSampler(U, J, R)
    out = s.draw(2, rng.random(2 * 2 * J))
    for k in range(2):
        lev, tot = s.leverage(k)
        p = lev / tot
        obs = np.bincount(out["samples"][k].astype(np.int64), minlength=dims[k])
        exp = J * p
        chi2 = np.sum((obs - exp) ** 2 / exp)
        # dof = I - 1; mean dof, sd sqrt(2 dof): allow 6 sd
        dof = dims[k] - 1
        assert chi2 < dof + 6 * np.sqrt(2 * dof), f"mode {k}: chi2 {chi2:.1f} for {dof} dof"


@pytest.mark.parametrize("dims,R,J,j", [((15, 12, 10), 5, 300, 1), ((8, 9, 7, 6), 3, 200, 2)])
def test_lk_als_update_is_sampled_least_squares(oracle, dims, R, J, j):
    rng = np.random.default_rng(R * J + j + 1)
    N = len(dims)
    idx, vals = random_sparse(dims, int(np.prod(dims)) // 4, rng, unique=False)
    X = dense_from_sparse(idx, vals, dims)
    U = [rng.standard_normal((d, R)) for d in dims]
    draws = rng.random(2 * N * J)
    out = oracle.OracleLKSampler([u.copy() for u in U], J, R).draw(j, draws)
    expect = sampled_ls_solution(X, U, j, out["samples"], out["h"], out["weights"])
    als = oracle.OracleLKALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U], J)
    als.update(j, draws, renormalize=False, update_sampler=False)
    np.testing.assert_allclose(als.U[j], expect, rtol=1e-8, atol=1e-10 * np.abs(expect).max())


def test_lk_als_recovers_lowrank_and_raises_fit(oracle):
    rng = np.random.default_rng(21)
    dims, R, J = (12, 10, 14), 4, 400
    Ut = [rng.standard_normal((d, R)) for d in dims]
    _, idx, vals = lowrank_as_sparse(Ut)
    U0 = [u.copy() for u in Ut]
    U0[0] = rng.standard_normal((dims[0], R))
    als = oracle.OracleLKALS(oracle.OracleSparse(idx, vals), U0, J)
    als.update(0, rng.random(2 * 3 * J), renormalize=False)
    np.testing.assert_allclose(als.U[0], Ut[0], rtol=1e-8, atol=1e-9)
    # sweeps on a noisy tensor from factors near the truth (ALS from random
    # factors can stall in a swamp for many sweeps, with either sampler)
    vals = vals + 0.01 * np.linalg.norm(vals) / np.sqrt(vals.size) * rng.standard_normal(vals.size)
    U1 = [u + 0.5 * rng.standard_normal(u.shape) for u in Ut]
    als = oracle.OracleLKALS(oracle.OracleSparse(idx, vals), U1, J)
    f0 = als.fit()
    for _ in range(8):
        for j in range(3):
            als.update(j, rng.random(2 * 3 * J))
    assert als.fit() > 0.98 > f0


# ---- hybrid sampler (larsen_kolda_hybrid.hpp) ----

def _tuple_probs(s, dims, j):
    """p of every tuple over modes != j (lowest mode fastest) by brute force."""
    modes = [k for k in range(len(dims)) if k != j]
    ps = [s.leverage(k)[0] / s.leverage(k)[1] for k in modes]
    grids = np.meshgrid(*[np.arange(dims[k]) for k in modes], indexing="ij")
    P = np.ones(grids[0].shape)
    for g, p in zip(grids, ps):
        P = P * p[g]
    return modes, grids, P


@pytest.mark.parametrize("dims,R,J,j,skew", [((20, 15, 30), 3, 300, 2, 1), ((40, 50, 10), 4, 1000, 2, 0),
                                             ((12, 9, 7, 6), 3, 400, 1, 2), ((30, 60), 3, 30, 0, 1)])
def test_hybrid_deterministic_set_is_every_tuple_above_tau(oracle, dims, R, J, j, skew):
    """The kept tuples are exactly those with p >= 1/J (brute force over all
    tuples), each once, weight 1; p_det is their mass; the rejected samples
    all have p < 1/J and weight (1 - p_det) / (J p)."""
    rng = np.random.default_rng(sum(dims) + J + skew)
    U = [rng.standard_normal((d, R)) * rng.random((d, 1)) ** skew for d in dims]
    s = oracle.OracleLKSampler(U, J, R, hybrid=True)
    out = s.draw_hybrid(j, rng.random((256, len(dims), J)))
    modes, grids, P = _tuple_probs(s, dims, j)
    tau = 1.0 / J
    above = P >= tau
    nd = out["deterministic"]
    assert nd == int(above.sum()) and nd > 0
    np.testing.assert_allclose(out["p_det"], P[above].sum(), rtol=1e-12)
    det = {tuple(int(out["samples"][k][i]) for k in modes) for i in range(nd)}
    ref = {tuple(int(g[idx]) for g in grids) for idx in zip(*np.nonzero(above))}
    assert det == ref and len(det) == nd
    assert np.all(out["weights"][:nd] == 1.0)
    rest = np.ones(J - nd)
    for k in modes:
        lev, tot = s.leverage(k)
        rest = rest * (lev[out["samples"][k][nd:].astype(np.int64)] / tot)
    assert np.all(rest < tau)
    np.testing.assert_allclose(out["weights"][nd:], (1 - out["p_det"]) / (J * rest), rtol=1e-11)
    # listed in mixed-radix order of the row-sorted candidate lists, lowest
    # mode fastest: lexicographic in (row of the highest mode, ..., lowest)
    keys = [tuple(int(out["samples"][k][i]) for k in modes[::-1]) for i in range(nd)]
    assert keys == sorted(keys)
    # alpha = max leverage / total
    for k in modes:
        lev, tot = s.leverage(k)
        np.testing.assert_allclose(s.alpha()[k], lev.max() / tot, rtol=1e-15)


def test_hybrid_rejected_samples_follow_the_conditional_distribution(oracle):
    """Over 60 draws, the rejection samples' tuple frequencies match
    p / (1 - p_det) on the tuples below tau (chi-square)."""
    rng = np.random.default_rng(31)
    dims, R, J, j = (30, 40, 5), 3, 1000, 2
    U = [rng.standard_normal((d, R)) * rng.random((d, 1)) for d in dims]
    s = oracle.OracleLKSampler(U, J, R, hybrid=True)
    modes, grids, P = _tuple_probs(s, dims, j)
    below = P < 1.0 / J
    counts = np.zeros(P.shape)
    total = 0
    for _ in range(60):
        out = s.draw_hybrid(j, rng.random((256, 3, J)))
        nd = out["deterministic"]
        a = out["samples"][modes[0]][nd:].astype(np.int64)
        b = out["samples"][modes[1]][nd:].astype(np.int64)
        np.add.at(counts, (a, b), 1)
        total += J - nd
    assert not np.any(counts[~below])
    q = P[below] / P[below].sum()
    exp = total * q
    ok = exp > 5
    chi2 = np.sum((counts[below][ok] - exp[ok]) ** 2 / exp[ok])
    dof = int(ok.sum()) - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)


def test_hybrid_als_step_is_sampled_least_squares(oracle):
    rng = np.random.default_rng(5)
    dims, R, J, j = (15, 12, 10), 4, 300, 1
    idx, vals = random_sparse(dims, int(np.prod(dims)) // 3, rng, unique=False)
    X = dense_from_sparse(idx, vals, dims)
    U = [rng.standard_normal((d, R)) for d in dims]
    ut = rng.random((256, 3, J))
    out = oracle.OracleLKSampler([u.copy() for u in U], J, R, hybrid=True).draw_hybrid(j, ut)
    expect = sampled_ls_solution(X, U, j, out["samples"], out["h"], out["weights"])
    als = oracle.OracleLKALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U], J, hybrid=True)
    als.update(j, None, renormalize=False, update_sampler=False, utab=ut)
    np.testing.assert_allclose(als.U[j], expect, rtol=1e-8, atol=1e-10 * np.abs(expect).max())
