"""Degenerate shapes on the GPU against the oracle: single-row factors,
R = 1, R > I (rank-deficient Grams), J = 1, N = 2 and every mode of a
3-mode tensor sampled.  FP64 efficient and Larsen-Kolda samplers: rows equal
where the oracle's draw is more than 1e-10 from a boundary, weights within
1e-9; FP32 and hybrid samplers: rows in range, weights finite and positive."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE = [
    ((1, 5, 3), 1, 1, 3),
    ((1, 5, 3), 3, 7, 0),
    ((2, 1), 2, 5, 0),
    ((4, 4), 4, 3, 2),
    ((1, 1, 1), 2, 4, 3),
    ((3, 70, 2), 5, 64, 1),
]


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


def _case(dims, R, J):
    rng = np.random.default_rng(sum(dims) * 31 + R * 7 + J)
    U = [rng.standard_normal((d, R)) for d in dims]
    return U, rng.random(2 * len(dims) * J)


@pytest.mark.parametrize("dims,R,J,j", EDGE, ids=[f"{d}-R{R}-J{J}-j{j}" for d, R, J, j in EDGE])
def test_edge_shapes_efficient_and_larsen_kolda(ftl, oracle, dims, R, J, j):
    print(f"[gpu] edge dims={dims} R={R} J={J} j={j}", flush=True)
    U, draws = _case(dims, R, J)
    N = len(dims)
    sampled = [k for k in range(N) if k != j]
    ref = oracle.OracleSampler(U, J, R).draw(j, draws)
    s = ftl.Sampler(U, J, R, "efficient", precision="fp64")
    samples = np.zeros((N, J), dtype=np.uint64)
    h = np.zeros((J, R))
    w = np.zeros(J)
    s.KRPDrawSamples_materialize(j, samples, h, w, draws)
    sure = ref["margins"] > 1e-10
    for k in sampled:
        np.testing.assert_array_equal(samples[k][sure], ref["samples"][k][sure])
    same = np.all(samples[sampled] == ref["samples"][sampled], axis=0)
    np.testing.assert_allclose(w[same], ref["weights"][same], rtol=1e-9)
    # Larsen-Kolda
    lref = oracle.OracleLKSampler(U, J, R).draw(j, draws)
    lk = ftl.Sampler(U, J, R, "larsen_kolda")
    ls = np.zeros((N, J), dtype=np.uint64)
    lw = np.zeros(J)
    lk.KRPDrawSamples_materialize(j, ls, np.zeros((J, R)), lw, draws)
    lsure = lref["margins"] > 1e-10
    np.testing.assert_array_equal(ls[:, lsure], lref["samples"][:, lsure])
    lsame = np.all(ls == lref["samples"], axis=0)
    np.testing.assert_allclose(lw[lsame], lref["weights"][lsame], rtol=1e-9)
    # FP32 efficient: valid draws
    f32 = ftl.Sampler(U, J, R, "efficient", precision="fp32")
    out = np.zeros((N, J), dtype=np.uint64)
    ww = np.zeros(J)
    f32.KRPDrawSamples_materialize(j, out, np.zeros((J, R)), ww, draws)
    for k in sampled:
        assert np.all(out[k] < dims[k])
    assert np.all(np.isfinite(ww)) and np.all(ww > 0)
    # hybrid Larsen-Kolda: valid draws; when the deterministic tuples hold all
    # the mass the other samples exhaust their attempts and get weight 0
    hy = ftl.Sampler(U, J, R, "larsen_kolda_hybrid")
    out = np.zeros((N, J), dtype=np.uint64)
    ww = np.zeros(J)
    hy.KRPDrawSamples_materialize(j, out, np.zeros((J, R)), ww, draws)
    info = hy.last_info()
    nd = info["deterministic"]
    for k in sampled:
        assert np.all(out[k] < dims[k])
    assert np.all(ww[:nd] == 1.0) and np.all(np.isfinite(ww)) and np.all(ww >= 0)
    assert info["exhausted"] == int(np.sum(ww[nd:] == 0.0))
    if info["p_det"] >= 1.0 - 1e-12:
        assert info["exhausted"] == J - nd
