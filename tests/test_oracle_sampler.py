"""Pins the oracle's descent, computeM, draw and weights
(partition_tree.hpp:232-405, efficient_krp_sampler.hpp:90-268, sampler.hpp:94-122)
against brute force, NumPy linear algebra and exact leverage scores."""
import numpy as np
import pytest

from tests.helpers import (exact_leverage, flat_index, gram_pinv, multinomial_tv_band,
                           row_cdf_sample, tv_distance)


@pytest.mark.parametrize("n,R", [(100, 8), (37, 4), (1000, 8), (8, 8), (257, 16)])
def test_pt_sample_is_row_inverse_cdf(oracle, n, R):
    rng = np.random.default_rng(n + R)
    U = rng.standard_normal((n, R))
    F = min(R, n)
    G = oracle.build_tree(U, F)
    J = 2000
    x = rng.standard_normal((J, R))
    h = rng.standard_normal((J, R))
    h0 = h.copy()
    u = rng.random(J)
    margins = np.empty(J)
    s = oracle.pt_sample(G, U, F, x, h, u, margins)
    ok = margins > 1e-9
    assert ok.mean() > 0.99
    brute = np.array([row_cdf_sample(U, x[i], u[i]) for i in range(J)])
    assert np.array_equal(s[ok].astype(np.int64), brute[ok])
    np.testing.assert_array_equal(h, h0 * U[s.astype(np.int64)])


def test_pt_sample_uniform_edges(oracle):
    """u = 0 takes the first row with mass; u just below 1 the last."""
    rng = np.random.default_rng(5)
    U = rng.standard_normal((50, 4))
    G = oracle.build_tree(U, 4)
    x = np.tile(rng.standard_normal(4), (2, 1))
    h = np.ones((2, 4))
    s = oracle.pt_sample(G, U, 4, x, h, np.array([0.0, np.nextafter(1.0, 0.0)]))
    assert s[0] == 0 and s[1] == 49


@pytest.mark.parametrize("N,I,R,j", [(3, 8, 8, 4), (3, 8, 8, 1), (4, 6, 5, 0), (2, 40, 25, 5),
                                     (3, 10, 7, 2)])
def test_compute_m_closed_form(oracle, N, I, R, j):
    rng = np.random.default_rng(N * 100 + I + R + j)
    U = [rng.standard_normal((I, R)) for _ in range(N)]
    S = oracle.OracleSampler(U, 16, R)
    pinv, E = S.compute_m(j)
    excl = j if j < N else None
    ref_pinv, G = gram_pinv(U, excl)
    iu = np.triu_indices(R)
    np.testing.assert_allclose(pinv[iu], ref_pinv[iu], rtol=1e-8, atol=1e-10 * np.abs(ref_pinv).max())
    modes = [k for k in range(N) if k != excl]
    grams = {k: U[k].T @ U[k] for k in range(N)}
    for t, k in enumerate(modes):
        if t + 1 < len(modes):
            Pk = ref_pinv.copy()
            for k2 in modes[t + 1:]:
                Pk *= grams[k2]
            np.testing.assert_allclose(E[k].T @ E[k], Pk, rtol=1e-8, atol=1e-10 * np.abs(Pk).max())
            # rows are sqrt(mu) v^T in ascending mu order
            norms = np.sum(E[k] ** 2, axis=1)
            assert np.all(np.diff(norms) >= -1e-12 * norms.max())
        else:
            lam, V = np.linalg.eigh(G)
            ref = (V / np.sqrt(lam)).T[::-1]
            np.testing.assert_allclose(np.abs(E[k]), np.abs(ref), rtol=1e-7, atol=1e-10)
    if excl is not None:
        assert np.all(E[excl] == 0.0)


@pytest.mark.parametrize("N,I,R,j", [(3, 8, 8, 4), (3, 8, 8, 1), (3, 8, 8, 0), (2, 12, 6, 2),
                                     (3, 6, 4, 3)])
def test_draw_matches_exact_leverage(oracle, N, I, R, j):
    """TV(empirical, exact leverage) inside the multinomial noise band; the
    same test at the reference's distribution_comparison shape (N=3, I=8, R=8)."""
    rng = np.random.default_rng(1000 + N * 10 + j)
    U = [rng.standard_normal((I, R)) for _ in range(N)]
    J = 200_000
    S = oracle.OracleSampler(U, J, R)
    draws = np.random.default_rng(7).random(2 * N * J)
    out = S.draw(j, draws)
    excl = j if j < N else None
    p = exact_leverage(U, excl)
    idx = flat_index(out["samples"], [I] * N, excl)
    emp = np.bincount(idx, minlength=p.size) / J
    tv = tv_distance(emp, p)
    mean, std = multinomial_tv_band(p, J)
    assert abs(tv - mean) < 5 * std + 1e-3, (tv, mean, std)
    if excl is not None:
        assert np.all(out["samples"][excl] == 0)


def test_draw_c1_subcase_tv(oracle):
    """C1's TV subcase: N=4 modes of I=16, R=8 (65,536 KRP rows)."""
    rng = np.random.default_rng(42)
    N, I, R = 4, 16, 8
    U = [rng.standard_normal((I, R)) for _ in range(N)]
    J = 400_000
    S = oracle.OracleSampler(U, J, R)
    out = S.draw(N + 1, np.random.default_rng(8).random(2 * N * J), want_margins=False)
    p = exact_leverage(U)
    emp = np.bincount(flat_index(out["samples"], [I] * N), minlength=p.size) / J
    tv = tv_distance(emp, p)
    mean, std = multinomial_tv_band(p, J, trials=40)
    assert abs(tv - mean) < 5 * std + 1e-3, (tv, mean, std)


def test_tv_check_has_power(oracle):
    """Planted mistake on the CPU: squared-row-norm sampling instead of leverage
    must land far outside the band the checks above use."""
    rng = np.random.default_rng(9)
    N, I, R = 3, 8, 8
    U = [rng.standard_normal((I, R)) for _ in range(N)]
    U[0][0] *= 10.0
    p = exact_leverage(U)
    from tests.helpers import krp
    A = krp(U)
    wrong = np.sum(A * A, axis=1)
    wrong /= wrong.sum()
    J = 200_000
    emp = np.random.default_rng(1).multinomial(J, wrong) / J
    mean, std = multinomial_tv_band(p, J)
    assert tv_distance(emp, p) > mean + 20 * std


@pytest.mark.parametrize("j", [3, 1])
def test_weights_and_h_closed_form(oracle, j):
    rng = np.random.default_rng(77 + j)
    N, I, R, J = 3, 30, 6, 5000
    U = [rng.standard_normal((I, R)) for _ in range(N)]
    S = oracle.OracleSampler(U, J, R)
    out = S.draw(j, rng.random(2 * N * J))
    excl = j if j < N else None
    h = np.ones((J, R))
    for k in range(N):
        if k != excl:
            h *= U[k][out["samples"][k].astype(np.int64)]
    np.testing.assert_allclose(out["h"], h, rtol=1e-13)
    pinv, _ = gram_pinv(U, excl)
    lev = np.einsum("ir,rs,is->i", h, pinv, h)
    np.testing.assert_allclose(out["weights"], R / (J * lev), rtol=1e-8)


def test_update_sampler_rebuilds_one_tree(oracle):
    rng = np.random.default_rng(3)
    U = [rng.standard_normal((20, 4)) for _ in range(3)]
    S = oracle.OracleSampler(U, 8, 4)
    new = rng.standard_normal((20, 4))
    S.update_sampler(1, new)
    iu = np.triu_indices(4)
    np.testing.assert_allclose(S.root(1)[iu], (new.T @ new)[iu], rtol=1e-12)
    np.testing.assert_allclose(S.root(0)[iu], (U[0].T @ U[0])[iu], rtol=1e-12)
