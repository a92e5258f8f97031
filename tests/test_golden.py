"""Committed golden fixtures (tests/golden/, written by make_golden.py):

* the reference's statistical fixture (submission_data/oct_30/
  distribution_comparison.json: exact leverage of the 512 KRP rows at N=3,
  I=8, R=8 and the 50,000 rows its sampler drew) pins the total-variation
  noise band the distribution tests use: the reference's own draws must sit
  inside it, and a planted wrong distribution must not;
* the oracle's outputs on small seeded cases must be reproduced bitwise by
  the oracle (a change to the restatement shows here first).
"""
import os

import numpy as np
import pytest

from tests.helpers import multinomial_tv_band, philox_uniforms, tv_distance

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_reference_distribution_fixture_within_noise_band():
    d = _load("reference_distribution_comparison.npz")
    R, J = int(d["R"]), int(d["J"])
    lev = d["true_distribution"]
    assert lev.shape == (512,) and abs(lev.sum() - R) < 1e-9  # leverage scores sum to the rank
    p = lev / lev.sum()
    draws = d["sts_sampler_draws"].astype(np.int64)
    assert draws.shape == (J,) and draws.max() < 512
    emp = np.bincount(draws, minlength=512) / J
    tv = tv_distance(emp, p)
    assert abs(tv - 0.0337) < 5e-4  # the TV the reference reports for this run
    mean, std = multinomial_tv_band(p, J, trials=400)
    assert abs(tv - mean) < 5 * std, (tv, mean, std)
    # a planted mistake: the same draws against a distribution with two
    # rows' probabilities swapped fall far outside the band
    q = p.copy()
    a, b = int(np.argmax(p)), int(np.argmin(p))
    q[a], q[b] = q[b], q[a]
    assert tv_distance(emp, q) - mean > 20 * std


@pytest.mark.parametrize("name", ["efficient_N3_I60-45-50_R6_J400_j1.npz", "efficient_N4_I20-17-12-9_R5_J300_all.npz"])
def test_oracle_reproduces_efficient_fixture(oracle, name):
    d = _load(name)
    U = [d[f"U{k}"] for k in range(len([f for f in d.files if f.startswith("U")]))]
    out = oracle.OracleSampler(U, int(d["J"]), int(d["R"])).draw(int(d["j"]), d["draws"])
    np.testing.assert_array_equal(out["samples"], d["samples"])
    np.testing.assert_array_equal(out["h"], d["h"])
    np.testing.assert_array_equal(out["weights"], d["weights"])
    np.testing.assert_array_equal(out["margins"], d["margins"])


def _utab(draws, seed, call, N, J, A=128):
    ut = np.empty((A, N, J))
    ut[0] = draws[:N * J].reshape(N, J)
    for a in range(1, A):
        for k in range(N):
            ut[a, k] = philox_uniforms(seed, call ^ (a << 40), J, 2 * k)
    return ut


@pytest.mark.parametrize("name,hybrid", [("lk_N3_I70-50-40_R5_J500_j0.npz", False),
                                         ("lkh_N3_I30-25-40_R4_J400_j2.npz", True)])
def test_oracle_reproduces_larsen_kolda_fixture(oracle, name, hybrid):
    d = _load(name)
    U = [d[f"U{k}"] for k in range(3)]
    J, R, j = int(d["J"]), int(d["R"]), int(d["j"])
    s = oracle.OracleLKSampler(U, J, R, hybrid=hybrid)
    out = s.draw_hybrid(j, _utab(d["draws"], int(d["seed"]), int(d["call"]), 3, J)) if hybrid else s.draw(j, d["draws"])
    np.testing.assert_array_equal(out["samples"], d["samples"])
    np.testing.assert_array_equal(out["h"], d["h"])
    np.testing.assert_array_equal(out["weights"], d["weights"])
    if hybrid:
        assert out["deterministic"] == int(d["deterministic"]) and out["p_det"] == float(d["p_det"])


def test_oracle_reproduces_als_fixture(oracle):
    d = _load("als_N3_I30-25-20_R4_J300_j1.npz")
    U = [d[f"U{k}"].copy() for k in range(3)]
    ref = oracle.OracleALS(oracle.OracleSparse(d["idx"], d["vals"]), U, int(d["J"]))
    st = ref.update(int(d["j"]), d["draws"], renormalize=True, update_sampler=True)
    assert st["unique_samples"] == int(d["unique"]) and st["found"] == int(d["found"])
    np.testing.assert_array_equal(ref.U[int(d["j"])], d["Uj"])
    np.testing.assert_array_equal(ref.sigma, d["sigma"])
