"""Writes the golden fixtures of this directory (run from the repository root:
``python tests/golden/make_golden.py``).

* ``reference_distribution_comparison.npz``: the data of the reference's
  statistical fixture submission_data/oct_30/distribution_comparison.json
  (N=3, I=8, R=8, J=50,000: the exact leverage scores of the 512 KRP rows,
  summing to R, and the 50,000 row indices its sampler drew), as arrays.
* ``*.npz``: small seeded cases with the CPU oracle's outputs (factors,
  uniforms, sampled rows, margins, h, weights; one sampled ALS update) --
  parity fixtures for the HIP path and regression fixtures for the oracle.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from tests.helpers import philox_uniforms, random_sparse  # noqa: E402

REF_FIXTURE = "/root/reference/submission_data/oct_30/distribution_comparison.json"


def reference_distribution():
    with open(REF_FIXTURE) as f:
        d = json.load(f)
    draws = np.asarray(d["sts_sampler_draws"], dtype=np.float64)
    assert np.all(draws == np.round(draws)) and draws.min() >= 0 and draws.max() < 2 ** 16
    np.savez_compressed(os.path.join(HERE, "reference_distribution_comparison.npz"), N=d["N"], I=d["I"],
                        R=d["R"], J=d["J"], true_distribution=np.asarray(d["true_distribution"], dtype=np.float64),
                        sts_sampler_draws=draws.astype(np.uint16))


def efficient(name, dims, R, J, j, seed):
    rng = np.random.default_rng(seed)
    U = [rng.standard_normal((d, R)) for d in dims]
    draws = rng.random(2 * len(dims) * J)
    out = oracle.OracleSampler(U, J, R).draw(j, draws)
    np.savez_compressed(os.path.join(HERE, name), j=j, R=R, J=J, draws=draws, samples=out["samples"],
                        margins=out["margins"], h=out["h"], weights=out["weights"],
                        **{f"U{k}": U[k] for k in range(len(U))})


def larsen_kolda(name, dims, R, J, j, seed, hybrid=False, skew=0):
    rng = np.random.default_rng(seed)
    U = [rng.standard_normal((d, R)) * rng.random((d, 1)) ** skew for d in dims]
    N = len(dims)
    draws = rng.random(2 * N * J)
    s = oracle.OracleLKSampler(U, J, R, hybrid=hybrid)
    extra = {}
    if hybrid:
        sseed, call, A = 77, 3, 128
        ut = np.empty((A, N, J))
        ut[0] = draws[:N * J].reshape(N, J)
        for a in range(1, A):
            for k in range(N):
                ut[a, k] = philox_uniforms(sseed, call ^ (a << 40), J, 2 * k)
        out = s.draw_hybrid(j, ut)
        extra = {"seed": sseed, "call": call, "deterministic": out["deterministic"], "p_det": out["p_det"],
                 "margin": min(out["cut_margin"], out["keep_margin"], out["draw_margin"])}
    else:
        out = s.draw(j, draws)
        extra = {"margins": out["margins"]}
    np.savez_compressed(os.path.join(HERE, name), j=j, R=R, J=J, draws=draws, samples=out["samples"],
                        h=out["h"], weights=out["weights"], **extra,
                        **{f"U{k}": U[k] for k in range(len(U))})


def als(name, dims, R, J, nnz, j, seed):
    rng = np.random.default_rng(seed)
    idx, vals = random_sparse(dims, nnz, rng, unique=False)
    U = [rng.standard_normal((d, R)) for d in dims]
    draws = rng.random(2 * len(dims) * J)
    ref = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U], J)
    margins = ref.sampler.draw(j, draws)["margins"]
    st = ref.update(j, draws, renormalize=True, update_sampler=True)
    np.savez_compressed(os.path.join(HERE, name), j=j, R=R, J=J, idx=idx, vals=vals, draws=draws,
                        margin=margins.min(), Uj=ref.U[j], sigma=ref.sigma, unique=st["unique_samples"],
                        found=st["found"], **{f"U{k}": U[k] for k in range(len(U))})


if __name__ == "__main__":
    oracle.build()
    reference_distribution()
    efficient("efficient_N3_I60-45-50_R6_J400_j1.npz", (60, 45, 50), 6, 400, 1, 101)
    efficient("efficient_N4_I20-17-12-9_R5_J300_all.npz", (20, 17, 12, 9), 5, 300, 5, 102)
    larsen_kolda("lk_N3_I70-50-40_R5_J500_j0.npz", (70, 50, 40), 5, 500, 0, 103)
    larsen_kolda("lkh_N3_I30-25-40_R4_J400_j2.npz", (30, 25, 40), 4, 400, 2, 104, hybrid=True, skew=1)
    als("als_N3_I30-25-20_R4_J300_j1.npz", (30, 25, 20), 4, 300, 1500, 1, 105)
    print("fixtures written to", HERE)
