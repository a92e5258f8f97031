"""The samplers at the headline shape: N=3, I=2^23, R=32, J=50,000, every mode
sampled (j = N+1, as in runtime_benchmark.py:56), checked through properties
that hold at any size.

* Efficient sampler, FP64: indices equal to the oracle's on the first 2,000
  samples (the oracle draws those from its own trees with the same uniforms)
  wherever the oracle's margin exceeds 1e-10; h equal, bitwise, to the product
  of the sampled factor rows in mode order (efficient_krp_sampler.hpp:238-254);
  weights R / (J h^T G^-1 h) with G the Hadamard product of the factor Grams
  from NumPy (sampler.hpp:94-122), within 1e-9 relative; the weighted sampled
  Gram sum_i w_i h_i h_i^T a subspace embedding of G (spectral deviation of
  G^-1/2 S G^-1/2 from I below 0.15; about 2 sqrt(R/J) = 0.05 is expected).
* Efficient sampler, FP32: index mismatch rate against the FP64 sampler on
  identical uniforms at most 1e-4, and every mismatch where the FP64
  oracle's margin is below 1e-8 (the oracle also draws every sample the two
  precisions disagree on); h and weights by the same closed forms.
* Larsen-Kolda: leverage scores within 1e-10 relative of NumPy's
  u_i^T (U^T U)^-1 u_i; each sampled row is the inverse CDF of its uniform over
  the scores (larsen_kolda_sampler.hpp:102-136; checked against NumPy's cumsum
  with a 1e-10 tolerance relative to the total); h bitwise the row product;
  weights (1/J) prod_k total_k / lev_k within 1e-12 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, I, R, J = 3, 2 ** 23, 32, 50_000
J_ORACLE = 2_000


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


@pytest.fixture(scope="module")
def factors():
    print(f"[gpu] full size: generating N={N} factors of {I} x {R}", flush=True)
    U = [np.random.default_rng(1000 + k).standard_normal((I, R)) for k in range(N)]
    grams = [M.T @ M for M in U]
    return U, grams


def _row_product(U, samples):
    h = np.ones((samples.shape[1], R))
    for k in range(N):
        h = h * U[k][samples[k].astype(np.int64)]
    return h


def _check_efficient_closed_forms(U, grams, samples, h, w):
    assert samples.max() < I
    np.testing.assert_array_equal(h, _row_product(U, samples))
    G = grams[0] * grams[1] * grams[2]
    Ginv = np.linalg.inv(G)
    lev = np.einsum("ir,rs,is->i", h, Ginv, h)
    np.testing.assert_allclose(w, R / (J * lev), rtol=1e-9)
    evals, evecs = np.linalg.eigh(G)
    g_half_inv = (evecs / np.sqrt(evals)) @ evecs.T
    S = (h * w[:, None]).T @ h
    dev = np.linalg.norm(g_half_inv @ S @ g_half_inv - np.eye(R), 2)
    print(f"[gpu] full size: embedding deviation {dev:.4f}", flush=True)
    assert dev < 0.15


def test_efficient_sampler_full_size(ftl, oracle, factors):
    U, grams = factors
    draws = np.random.default_rng(7).random(2 * N * J)
    out = {}
    for prec in ("fp64", "fp32"):
        print(f"[gpu] full size: efficient sampler {prec}", flush=True)
        s = ftl.Sampler(U, J, R, "efficient", precision=prec)
        samples = np.zeros((N, J), dtype=np.uint64)
        h = np.zeros((J, R))
        w = np.zeros(J)
        s.KRPDrawSamples_materialize(N + 1, samples, h, w, draws)
        s.close()
        _check_efficient_closed_forms(U, grams, samples, h, w)
        out[prec] = samples
    mism = np.any(out["fp32"] != out["fp64"], axis=0)
    rate = mism.mean()
    print(f"[gpu] full size: fp32 vs fp64 index mismatch rate {rate:.2e}", flush=True)
    assert rate <= 1e-4

    pick = np.union1d(np.arange(J_ORACLE), np.nonzero(mism)[0])
    print(f"[gpu] full size: oracle on the first {J_ORACLE} samples and the {int(mism.sum())} mismatches",
          flush=True)
    ref = oracle.OracleSampler(U, pick.size, R)
    sub = np.concatenate([draws[t * J + pick] for t in range(2 * N)])
    r = ref.draw(N + 1, sub)
    del ref
    sure = r["margins"] > 1e-10
    gpu = out["fp64"][:, pick]
    np.testing.assert_array_equal(gpu[:, sure], r["samples"][:, sure])
    assert sure[:J_ORACLE].mean() > 0.99
    m = r["margins"][np.isin(pick, np.nonzero(mism)[0])]
    print(f"[gpu] full size: oracle margins of the fp32 mismatches {np.array2string(m, precision=2)}", flush=True)
    assert np.all(m < 1e-8)


def test_larsen_kolda_full_size(ftl, factors):
    U, grams = factors
    print("[gpu] full size: larsen-kolda sampler", flush=True)
    s = ftl.Sampler(U, J, R, "larsen_kolda")
    levs, tots = [], []
    for k in range(N):
        lev, tot = s.leverage(k)
        ref = np.einsum("ir,ir->i", U[k] @ np.linalg.inv(grams[k]), U[k])
        np.testing.assert_allclose(lev, ref, rtol=1e-10)
        np.testing.assert_allclose(tot, ref.sum(), rtol=1e-12)
        levs.append(lev)
        tots.append(tot)
    draws = np.random.default_rng(8).random(2 * N * J)
    samples = np.zeros((N, J), dtype=np.uint64)
    h = np.zeros((J, R))
    w = np.zeros(J)
    s.KRPDrawSamples_materialize(N + 1, samples, h, w, draws)
    s.close()
    logw = np.full(J, -np.log(J))
    for k in range(N):
        rows = samples[k].astype(np.int64)
        assert rows.max() < I
        cum = np.cumsum(levs[k])
        t = draws[k * J:(k + 1) * J] * tots[k]
        tol = 1e-10 * tots[k]
        below = np.where(rows > 0, cum[np.maximum(rows - 1, 0)], 0.0)
        assert np.all(below <= t + tol), f"mode {k}: a row below its uniform's interval"
        assert np.all(cum[rows] > t - tol), f"mode {k}: a row above its uniform's interval"
        logw += np.log(tots[k]) - np.log(levs[k][rows])
    np.testing.assert_array_equal(h, _row_product(U, samples))
    np.testing.assert_allclose(w, np.exp(logw), rtol=1e-12)
