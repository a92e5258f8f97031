"""Parity of the HIP sparse tensor and sampled CP-ALS step (libftl_hip.so) with
the CPU FP64 oracle on identical inputs and identical uniforms.

Bar: matched nonzero ids and counts equal as integers; MTTKRP rows bitwise
equal when a row has at most 256 matches (same summation order), else within
1e-12 relative; after an ALS update U_j within 1e-9 relative (the pinv's
eigensolver rounds differently), sigma within 1e-10, J' and found_count
equal; fit and norms within 1e-10 relative.
"""
import zlib

import numpy as np
import pytest

from tests.helpers import cp_dense, dense_from_sparse, random_sparse

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


def _seed(name):
    return zlib.crc32(name.encode())


def _skewed_sparse(dims, nnz, rng):
    """Random nonzeros plus a full slice at mode-0 index 0 (a heavy row)."""
    idx, vals = random_sparse(dims, nnz, rng, unique=False)
    grid = np.stack(np.meshgrid(*[np.arange(d) for d in dims[1:]], indexing="ij"), axis=-1).reshape(-1, len(dims) - 1)
    slab = np.concatenate([np.zeros((grid.shape[0], 1), dtype=np.int64), grid], axis=1).astype(np.uint32)
    idx = np.ascontiguousarray(np.concatenate([idx, slab]))
    vals = np.concatenate([vals, rng.standard_normal(slab.shape[0])])
    return idx, vals


LOOKUP_CASES = [
    ("dup3", (5, 6, 7), 80, 1, False),
    ("uniq3", (9, 4, 11), 150, 0, True),
    ("n4", (3, 4, 5, 6), 200, 3, False),
    ("n2", (40, 30), 300, 1, True),
    ("wide", (1000, 2000, 300), 20000, 2, False),
]


@pytest.mark.parametrize("method", ["sort", "hash"])
@pytest.mark.parametrize("cid,dims,nnz,j,unique", LOOKUP_CASES, ids=[c[0] for c in LOOKUP_CASES])
def test_lookup_and_mttkrp_match_oracle(ftl, oracle, cid, dims, nnz, j, unique, method):
    """Matched nonzero ids and counts bit-exact against the oracle's sorted
    lookup, with the binary search and with the device coordinate hash."""
    print(f"[gpu] sparse lookup/mttkrp case={cid} method={method}", flush=True)
    rng = np.random.default_rng(_seed(cid))
    idx, vals = random_sparse(dims, nnz, rng, unique)
    X = ftl.SparseTensor(idx, vals, method)
    S = oracle.OracleSparse(idx, vals)
    assert X.dims == S.dims
    np.testing.assert_allclose(X.get_normsq(), S.normsq(), rtol=1e-13)
    Jp, R = 400, 7
    samples = np.stack([rng.integers(0, d, size=Jp) for d in dims], axis=1).astype(np.uint64)
    samples[: Jp // 2] = idx[rng.integers(0, idx.shape[0], size=Jp // 2)]
    samples[-3:, (j + 1) % len(dims)] = dims[(j + 1) % len(dims)] + 5  # out of range: no match
    lhs = rng.standard_normal((Jp, R))
    counts, ids = X.lookup(samples, j)
    ref, found, rcounts, rids = S.mttkrp(samples, lhs, j, dims[j], want_matches=True)
    np.testing.assert_array_equal(counts, rcounts)
    np.testing.assert_array_equal(ids, rids)
    out = np.zeros((dims[j], R))
    X.execute_downsampled_mttkrp_py(samples, lhs, j, out)
    assert X.last_found == found
    # rows with at most 256 matches are summed in the oracle's order: bitwise
    per_row = np.bincount(idx[rids.astype(np.int64), j], minlength=dims[j])
    light = per_row <= 256
    assert light.any()
    np.testing.assert_array_equal(out[light], ref[light])
    np.testing.assert_allclose(out[~light], ref[~light], rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_mttkrp_heavy_rows(ftl, oracle):
    """Rows with more than 256 matches are reduced in pieces."""
    rng = np.random.default_rng(3)
    dims = (20, 40, 30)
    idx, vals = _skewed_sparse(dims, 2000, rng)
    X = ftl.SparseTensor(idx, vals)
    S = oracle.OracleSparse(idx, vals)
    Jp, R = 3000, 25
    samples = np.stack([rng.integers(0, d, size=Jp) for d in dims], axis=1).astype(np.uint64)
    lhs = rng.standard_normal((Jp, R))
    ref, found, counts, _ = S.mttkrp(samples, lhs, 0, dims[0], want_matches=True)
    out = np.zeros((dims[0], R))
    X.execute_downsampled_mttkrp_py(samples, lhs, 0, out)
    assert X.last_found == found and found > 1000
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("R", [3, 25, 40, 70])
def test_residual_and_norms(ftl, oracle, R):
    rng = np.random.default_rng(R)
    dims = (30, 20, 25)
    idx, vals = random_sparse(dims, 4000, rng, unique=True)
    X = ftl.SparseTensor(idx, vals)
    S = oracle.OracleSparse(idx, vals)
    U = [rng.standard_normal((d, R)) for d in dims]
    sigma = rng.random(R) + 0.5
    np.testing.assert_allclose(X.compute_residual_normsq(sigma, U), S.residual_normsq(sigma, U), rtol=1e-10)
    M = cp_dense(U, sigma)
    Xd = dense_from_sparse(idx, vals, dims)
    np.testing.assert_allclose(X.compute_residual_normsq(sigma, U), np.sum((Xd - M) ** 2), rtol=1e-10)
    lr = ftl.LowRankTensor(R, [u.copy() for u in U])
    lr.set_sigma(sigma)
    np.testing.assert_allclose(lr.get_normsq(), oracle.lowrank_normsq(U, sigma), rtol=1e-11)
    # renormalize_columns(-1): unit columns, sigma = product of the norms
    lr2 = ftl.LowRankTensor(R, [u.copy() for u in U])
    lr2.renormalize_columns(-1)
    prod = np.ones(R)
    for k, u in enumerate(U):
        nrm = np.linalg.norm(u, axis=0)
        prod *= nrm
        np.testing.assert_allclose(lr2.U[k], u / nrm, rtol=1e-13)
    np.testing.assert_allclose(lr2.sigma, prod, rtol=1e-12)
    lr3 = ftl.LowRankTensor(R, [u.copy() for u in U])
    lr3.renormalize_columns(1)
    np.testing.assert_allclose(lr3.sigma, np.linalg.norm(U[1], axis=0), rtol=1e-13)
    np.testing.assert_array_equal(lr3.U[0], U[0])


ALS_CASES = [
    # id, dims, R, J, nnz, skewed
    ("r5", (40, 30, 20), 5, 500, 3000, False),
    ("r25_heavy", (30, 40, 35), 25, 2000, 4000, True),
    ("n4_r3", (12, 10, 9, 8), 3, 400, 2500, False),
    ("r40", (50, 45, 40), 40, 1500, 6000, False),
    ("r70", (60, 50, 80), 70, 1200, 8000, False),
    ("n2_r6", (300, 250), 6, 800, 5000, False),
]


@pytest.mark.parametrize("cid,dims,R,J,nnz,skewed", ALS_CASES, ids=[c[0] for c in ALS_CASES])
def test_als_updates_match_oracle(ftl, oracle, cid, dims, R, J, nnz, skewed):
    print(f"[gpu] ALS parity case={cid}", flush=True)
    rng = np.random.default_rng(_seed(cid))
    N = len(dims)
    idx, vals = _skewed_sparse(dims, nnz, rng) if skewed else random_sparse(dims, nnz, rng, unique=False)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    ref = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U0], J)
    lr = ftl.LowRankTensor(R, [u.copy() for u in U0])
    als = ftl.ALS(lr, ftl.SparseTensor(idx, vals))
    als.initialize_ds_als(J, "efficient")
    for step, j in enumerate([0, 1, N - 1, 0]):
        draws = rng.random(2 * N * J)
        # identical samples need every uniform away from a CDF boundary
        probe = ref.sampler.draw(j, draws)
        if probe["margins"].min() <= 1e-10:
            pytest.skip(f"uniform within 1e-10 of a boundary at step {step} (seed choice)")
        d = ref.update(j, draws, renormalize=True, update_sampler=True)
        als.execute_ds_als_update(j, True, True, random_draws=draws)
        st = als.last_stats()
        assert st["unique_samples"] == d["unique_samples"], step
        assert st["found"] == d["found"], step
        scale = np.abs(ref.U[j]).max()
        np.testing.assert_allclose(lr.U[j], ref.U[j], rtol=1e-9, atol=1e-11 * scale, err_msg=f"step {step}")
        np.testing.assert_allclose(lr.sigma, ref.sigma, rtol=1e-10, err_msg=f"step {step}")
        # the next draw uses the rebuilt tree: pull the GPU's factor into the
        # oracle so both continue from identical factors
        ref.U[j][:] = lr.U[j]
        ref.sampler.update_sampler(j)
    np.testing.assert_allclose(als.residual_normsq(), ref.residual_normsq(), rtol=1e-10)


def test_als_without_renormalize_or_tree_update(ftl, oracle):
    rng = np.random.default_rng(77)
    dims, R, J = (25, 20, 15), 4, 300
    idx, vals = random_sparse(dims, 1500, rng, unique=True)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    ref = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U0], J)
    lr = ftl.LowRankTensor(R, [u.copy() for u in U0])
    als = ftl.ALS(lr, ftl.SparseTensor(idx, vals))
    als.initialize_ds_als(J)
    draws = rng.random(2 * 3 * J)
    ref.update(1, draws, renormalize=False, update_sampler=False)
    als.execute_ds_als_update(1, False, False, random_draws=draws)
    np.testing.assert_allclose(lr.U[1], ref.U[1], rtol=1e-9, atol=1e-11 * np.abs(ref.U[1]).max())
    np.testing.assert_array_equal(lr.sigma, np.ones(R))
    # the tree of mode 1 was not rebuilt: the next draw still samples the old U_1
    draws2 = rng.random(2 * 3 * J)
    ref.update(0, draws2, renormalize=False, update_sampler=False)
    als.execute_ds_als_update(0, False, False, random_draws=draws2)
    np.testing.assert_allclose(lr.U[0], ref.U[0], rtol=1e-9, atol=1e-11 * np.abs(ref.U[0]).max())


def test_als_prod_fits_noisy_lowrank(ftl):
    """Sampled sweeps (built-in uniforms, both precisions) recover a noisy
    rank-4 tensor stored as sparse (the FP64 oracle reaches 0.99 after 10
    sweeps at J = 4000 on this problem; at J = 1000 it stalls near 0.65 for
    20 sweeps, a swamp of ALS itself)."""
    rng = np.random.default_rng(5)
    dims, R = (30, 25, 20), 4
    Ut = [rng.standard_normal((d, R)) for d in dims]
    X = cp_dense(Ut)
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), axis=-1).reshape(-1, 3)
    vals = X.reshape(-1) + 0.01 * X.std() * rng.standard_normal(X.size)
    rhs = ftl.PySparseTensor.from_arrays(idx.astype(np.uint32), vals)
    for prec in ("fp64", "fp32"):
        lhs = ftl.PyLowRank(dims, R, seed=11)
        lhs.ten.renormalize_columns(-1)
        rec = ftl.als_prod(lhs, rhs, 4000, "efficient", 10, 1e-5, epoch_length=5, precision=prec, seed=3)
        assert rec["fits"][-1] > 0.98, (prec, rec["fits"])
        assert len(rec["update_times"]) >= 15


def test_als_errors(ftl):
    rng = np.random.default_rng(1)
    idx, vals = random_sparse((5, 6, 7), 50, rng)
    X = ftl.SparseTensor(idx, vals)
    with pytest.raises(TypeError):
        ftl.SparseTensor(idx.astype(np.int64), vals)
    with pytest.raises(ValueError):
        ftl.SparseTensor(idx, vals, "btree")
    U = [rng.standard_normal((d, 3)) for d in (5, 6, 7)]
    lr = ftl.LowRankTensor(3, U)
    als = ftl.ALS(lr, X)
    with pytest.raises(ValueError):
        als.execute_ds_als_update(0, True, True)
    with pytest.raises(ValueError):
        als.initialize_ds_als(100, "uniform")
    als.initialize_ds_als(100)
    with pytest.raises(ValueError):
        als.initialize_ds_als(100, "larsen_kolda")  # already initialised
    with pytest.raises(ValueError):
        als.execute_ds_als_update(3, True, True)
    with pytest.raises(ValueError):
        als.execute_ds_als_update(0, True, True, random_draws=np.zeros(5))
    # a factor with fewer rows than the tensor's mode size is refused
    with pytest.raises(ValueError):
        ftl.ALS(ftl.LowRankTensor(3, [rng.standard_normal((d, 3)) for d in (4, 6, 7)]), X)


# ---- exact (unsampled) update: execute_exact_als_update (als.hpp:222-270) ----

EXACT_CASES = [
    # id, dims, R, nnz, skewed (a full slice at mode-0 index 0: a row of 320 nonzeros)
    ("x_r4", (30, 25, 20), 4, 2000, False),
    ("x_skew_r8", (40, 20, 16), 8, 1500, True),
    ("x_n4_r3", (12, 10, 9, 8), 3, 2500, False),
    ("x_r40", (50, 45, 40), 40, 6000, False),
    ("x_r70", (60, 50, 80), 70, 8000, False),
    ("x_n2_r6", (300, 250), 6, 5000, False),
]


@pytest.mark.parametrize("cid,dims,R,nnz,skewed", EXACT_CASES, ids=[c[0] for c in EXACT_CASES])
def test_exact_updates_match_oracle(ftl, oracle, cid, dims, R, nnz, skewed):
    """U_j within 1e-9 relative of the oracle (the MTTKRP is summed per row in
    pieces and the pinv's eigensolver rounds differently), sigma within 1e-10,
    rows of mode j without nonzeros exactly zero."""
    print(f"[gpu] exact ALS parity case={cid}", flush=True)
    rng = np.random.default_rng(_seed(cid))
    N = len(dims)
    idx, vals = _skewed_sparse(dims, nnz, rng) if skewed else random_sparse(dims, nnz, rng, unique=False)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    ref = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U0], 16)
    lr = ftl.LowRankTensor(R, [u.copy() for u in U0])
    als = ftl.ALS(lr, ftl.SparseTensor(idx, vals))
    for step, j in enumerate([0, 1, N - 1, 0]):
        renorm = step != 1
        ref.exact_update(j, renormalize=renorm)
        als.execute_exact_als_update(j, renorm, True)
        scale = np.abs(ref.U[j]).max()
        np.testing.assert_allclose(lr.U[j], ref.U[j], rtol=1e-9, atol=1e-11 * scale, err_msg=f"step {step}")
        np.testing.assert_allclose(lr.sigma, ref.sigma, rtol=1e-10, err_msg=f"step {step}")
        empty = np.setdiff1d(np.arange(dims[j]), idx[:, j])
        assert not lr.U[j][empty].any()
        ref.U[j][:] = lr.U[j]
    np.testing.assert_allclose(als.residual_normsq(), ref.residual_normsq(), rtol=1e-10)


def test_exact_then_sampled_update_matches_oracle(ftl, oracle):
    """The exact-vs-sampled comparison loop of the reference (als.py:16-40):
    execute_exact_als_update(j) then execute_ds_als_update(j) on one handle."""
    rng = np.random.default_rng(91)
    dims, R, J = (40, 30, 25), 5, 600
    N = len(dims)
    idx, vals = random_sparse(dims, 4000, rng, unique=True)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    ref = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U0], J)
    lr = ftl.LowRankTensor(R, [u.copy() for u in U0])
    als = ftl.ALS(lr, ftl.SparseTensor(idx, vals))
    als.initialize_ds_als(J)
    for j in range(N):
        ref.exact_update(j)
        als.execute_exact_als_update(j, True, True)
        np.testing.assert_allclose(lr.U[j], ref.U[j], rtol=1e-9, atol=1e-11 * np.abs(ref.U[j]).max())
        ref.U[j][:] = lr.U[j]
        draws = rng.random(2 * N * J)
        if ref.sampler.draw(j, draws)["margins"].min() <= 1e-10:
            pytest.skip("uniform within 1e-10 of a boundary (seed choice)")
        d = ref.update(j, draws)
        als.execute_ds_als_update(j, True, True, random_draws=draws)
        assert als.last_stats()["found"] == d["found"]
        np.testing.assert_allclose(lr.U[j], ref.U[j], rtol=1e-9, atol=1e-11 * np.abs(ref.U[j]).max())
        ref.U[j][:] = lr.U[j]
        ref.sampler.update_sampler(j)


def test_als_prod_exact_is_monotone(ftl):
    """als_prod(method="exact") (als.py:84-85): the fit never decreases and
    reaches the noise floor of a fully observed noisy rank-4 tensor."""
    rng = np.random.default_rng(8)
    dims, R = (30, 25, 20), 4
    X = cp_dense([rng.standard_normal((d, R)) for d in dims])
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), axis=-1).reshape(-1, 3)
    noise = 0.01 * X.std() * rng.standard_normal(X.size)
    vals = X.reshape(-1) + noise
    rhs = ftl.PySparseTensor.from_arrays(idx.astype(np.uint32), vals)
    lhs = ftl.PyLowRank(dims, R, seed=11)
    lhs.ten.renormalize_columns(-1)
    rec = ftl.als_prod(lhs, rhs, 0, "exact", 20, 1e-9, epoch_length=1)
    fits = rec["fits"]
    assert all(b >= a - 1e-12 for a, b in zip(fits, fits[1:])), fits
    floor = 1 - np.linalg.norm(noise) / np.linalg.norm(vals)
    assert fits[-1] > floor - 1e-3, (fits[-1], floor)


def test_hash_lookup_als_updates_equal_sort(ftl):
    """Sampled ALS updates with the coordinate hash give bitwise the factors
    and counts of the sorted lookup (same matches, same order)."""
    rng = np.random.default_rng(44)
    dims, R, J = (300, 200, 150), 8, 3000
    idx, vals = _skewed_sparse(dims, 20000, rng)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    out = {}
    for method in ("sort", "hash"):
        lr = ftl.LowRankTensor(R, [u.copy() for u in U0])
        als = ftl.ALS(lr, ftl.SparseTensor(idx, vals, method))
        als.initialize_ds_als(J, seed=9)
        stats = []
        for j in (0, 1, 2, 0):
            als.execute_ds_als_update(j, True, True)
            stats.append(als.last_stats()["found"])
        out[method] = ([u.copy() for u in lr.U], stats)
    assert out["sort"][1] == out["hash"][1]
    for a, b in zip(out["sort"][0], out["hash"][0]):
        np.testing.assert_array_equal(a, b)


def test_last_draw_is_the_update_draw(ftl, oracle):
    """ALS.last_draw() returns the draw the last sampled update consumed: the
    samples, h and weights of a separate FP64 sampler on the same factors and
    uniforms, bitwise; and the oracle's update on those samples reproduces the
    GPU update (the comparison the C5 benchmark makes at full size)."""
    rng = np.random.default_rng(5151)
    dims, R, J = (60, 50, 40), 8, 800
    idx, vals = random_sparse(dims, 5000, rng, unique=False)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    lr = ftl.LowRankTensor(R, [u.copy() for u in U0])
    als = ftl.ALS(lr, ftl.SparseTensor(idx, vals))
    als.initialize_ds_als(J, "efficient")
    with pytest.raises(Exception, match="no sampled update"):
        als.last_draw()
    for j in (1, 0, 2):
        before = [u.copy() for u in lr.U]
        sigma_before = lr.sigma.copy()
        draws = rng.random(2 * len(dims) * J)
        als.execute_ds_als_update(j, True, True, random_draws=draws)
        samples, h, w = als.last_draw()
        ref = ftl.Sampler([u.copy() for u in before], J, R, "efficient", precision="fp64")
        s2 = np.zeros((len(dims), J), dtype=np.uint64)
        h2 = np.zeros((J, R))
        w2 = np.zeros(J)
        ref.KRPDrawSamples_materialize(j, s2, h2, w2, draws)
        np.testing.assert_array_equal(samples, s2)
        np.testing.assert_array_equal(h, h2)
        np.testing.assert_array_equal(w, w2)
        assert np.all(samples[j] == 0)
        ref.close()
        o = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in before], J, sigma=sigma_before)
        d = o.update_with_samples(j, samples)
        assert d["unique_samples"] == als.last_stats()["unique_samples"]
        scale = np.abs(o.U[j]).max()
        np.testing.assert_allclose(lr.U[j], o.U[j], rtol=1e-9, atol=1e-11 * scale)
        np.testing.assert_allclose(lr.sigma, o.sigma, rtol=1e-10)
