"""The HIP path against the committed golden fixtures (tests/golden/, the
CPU oracle's outputs on small seeded cases).  Bar: sampled rows equal as
integers wherever the fixture's draw is more than 1e-10 (efficient) / 1e-9
(Larsen-Kolda, relative to the total) from a CDF boundary; weights within
1e-10 / 1e-9 relative and h bitwise equal on samples whose rows agree; after
the ALS update U_j within 1e-9 relative, sigma within 1e-10, J' and the
matched-nonzero count equal."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _factors(d):
    return [d[f"U{k}"] for k in range(len([f for f in d.files if f.startswith("U") and f[1:].isdigit()]))]


@pytest.mark.parametrize("name", ["efficient_N3_I60-45-50_R6_J400_j1.npz", "efficient_N4_I20-17-12-9_R5_J300_all.npz"])
def test_efficient_sampler_matches_fixture(ftl, name):
    print(f"[gpu] golden {name}", flush=True)
    d = _load(name)
    U = _factors(d)
    N, J, R, j = len(U), int(d["J"]), int(d["R"]), int(d["j"])
    s = ftl.Sampler(U, J, R, "efficient", precision="fp64")
    samples = np.zeros((N, J), dtype=np.uint64)
    h = np.zeros((J, R))
    w = np.zeros(J)
    s.KRPDrawSamples_materialize(j, samples, h, w, d["draws"])
    sure = d["margins"] > 1e-10
    assert sure.mean() > 0.99
    for k in range(N):
        if j < N and k == j:
            continue
        np.testing.assert_array_equal(samples[k][sure], d["samples"][k][sure])
    same = np.all(samples == d["samples"], axis=0)
    np.testing.assert_allclose(w[same], d["weights"][same], rtol=1e-10)
    np.testing.assert_array_equal(h[same], d["h"][same])


def test_larsen_kolda_matches_fixture(ftl):
    d = _load("lk_N3_I70-50-40_R5_J500_j0.npz")
    U = _factors(d)
    J, R, j = int(d["J"]), int(d["R"]), int(d["j"])
    s = ftl.Sampler(U, J, R, "larsen_kolda")
    samples = np.zeros((3, J), dtype=np.uint64)
    h = np.zeros((J, R))
    w = np.zeros(J)
    s.KRPDrawSamples_materialize(j, samples, h, w, d["draws"])
    sure = d["margins"] > 1e-9
    np.testing.assert_array_equal(samples[:, sure], d["samples"][:, sure])
    same = np.all(samples == d["samples"], axis=0)
    np.testing.assert_allclose(w[same], d["weights"][same], rtol=1e-9)
    np.testing.assert_array_equal(h[same], d["h"][same])


def test_larsen_kolda_hybrid_matches_fixture(ftl):
    d = _load("lkh_N3_I30-25-40_R4_J400_j2.npz")
    assert float(d["margin"]) > 1e-9
    U = _factors(d)
    J, R, j = int(d["J"]), int(d["R"]), int(d["j"])
    s = ftl.Sampler(U, J, R, "larsen_kolda_hybrid")
    s.set_seed(int(d["seed"]), int(d["call"]))
    samples = np.zeros((3, J), dtype=np.uint64)
    h = np.zeros((J, R))
    w = np.zeros(J)
    s.KRPDrawSamples_materialize(j, samples, h, w, d["draws"])
    info = s.last_info()
    assert info["deterministic"] == int(d["deterministic"])
    np.testing.assert_allclose(info["p_det"], float(d["p_det"]), rtol=1e-12)
    np.testing.assert_array_equal(samples, d["samples"])
    np.testing.assert_allclose(w, d["weights"], rtol=1e-9)
    np.testing.assert_array_equal(h, d["h"])


def test_als_update_matches_fixture(ftl):
    d = _load("als_N3_I30-25-20_R4_J300_j1.npz")
    assert float(d["margin"]) > 1e-10
    U = _factors(d)
    J, R, j = int(d["J"]), int(d["R"]), int(d["j"])
    lr = ftl.LowRankTensor(R, [u.copy() for u in U])
    als = ftl.ALS(lr, ftl.SparseTensor(d["idx"], d["vals"]))
    als.initialize_ds_als(J)
    als.execute_ds_als_update(j, True, True, random_draws=d["draws"])
    st = als.last_stats()
    assert st["unique_samples"] == int(d["unique"]) and st["found"] == int(d["found"])
    np.testing.assert_allclose(lr.U[j], d["Uj"], rtol=1e-9, atol=1e-11 * np.abs(d["Uj"]).max())
    np.testing.assert_allclose(lr.sigma, d["sigma"], rtol=1e-10)
