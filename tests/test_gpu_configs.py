"""The GPU path on the benchmark configurations that have no fixture of their
own: C1's total-variation subcase, C3 (ill-conditioned factors, FP32 vs FP64
at I = 2^23) and C4 (the 512-request serving stream).

C1: N=4, I=16, R=8 (65,536 KRP rows), J=400,000 samples of every mode: the
empirical row distribution's TV distance to the exact leverage scores (QR of
the materialised KRP) lies inside the exact-multinomial noise band
(compare_distributions.py:21-47), for both precisions.

C3: N=3, I=2^23, J=50,000, R in {16, 64, 128}, U = Q_U diag(s) Q_V^T with s
geometric from 1 to 1e-6.  FP32 and FP64 samplers draw on identical uniforms.
The FP64 sampler's indices equal the oracle's wherever the oracle's margin
exceeds 1e-10 (the oracle draws the first 2,000 samples and every sample the
two GPU samplers disagree on); every FP32/FP64 mismatch lies where the
oracle's margin is below EPS32; the mismatch rate is at most 1e-4
(partition_tree.hpp:273-391).  The leverage probability the FP32 path gives
each row it drew, 1 / (J w), is compared with the FP64 probability of the
same row, h^T pinv(G) h / R (h the FP64 product of the drawn rows, pinv of
the Hadamard product of the FP64 Grams with the 1e-8 cut); its relative error
is printed and bounded.

C4: 512 requests of random_requests(512) replayed with Poisson arrivals at
500/s (N in {3,4,5}, I log-uniform in [2^10, 2^22], R in {8,16,32,64}, J
log-uniform in [2^10, 2^16], FP32): each result is bitwise equal to the same
request served alone, the arena never exceeds its budget, and p50/p99 are
printed.
"""
import numpy as np
import pytest

from tests.helpers import exact_leverage, flat_index, multinomial_tv_band, tv_distance

pytestmark = pytest.mark.gpu

J_ORACLE = 2_000
# FP32 mismatches must sit within this distance (in uniform units) of a CDF
# boundary of the FP64 oracle's descent: the FP32 levels shift a cutoff by
# about 2^-24 of the level's mass (2^-d at depth d >= 13 here) and the leaf
# scan's FP32 rows shift frac by about 2^-24 of a leaf interval.
EPS32 = 1e-8


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c1_subcase_tv_on_gpu(ftl, precision):
    print(f"[gpu] C1 TV subcase N=4 I=16 R=8 J=400000 {precision}", flush=True)
    rng = np.random.default_rng(42)
    N, I, R, J = 4, 16, 8, 400_000
    U = [rng.standard_normal((I, R)) for _ in range(N)]
    s = ftl.Sampler(U, J, R, "efficient", precision=precision)
    samples = np.zeros((N, J), dtype=np.uint64)
    s.KRPDrawSamples(N + 1, samples, np.random.default_rng(8).random(2 * N * J))
    p = exact_leverage(U)
    emp = np.bincount(flat_index(samples, [I] * N), minlength=p.size) / J
    tv = tv_distance(emp, p)
    mean, std = multinomial_tv_band(p, J, trials=40)
    print(f"[gpu] C1 TV {tv:.5f} (noise band {mean:.5f} +- {std:.5f})", flush=True)
    assert abs(tv - mean) < 5 * std + 1e-3, (tv, mean, std)


def _oracle_check(oracle, U_host, R, draws, J, s64, s32, eps32):
    """Oracle on the first J_ORACLE samples and every FP32/FP64 mismatch."""
    N = len(U_host)
    mism = np.any(s32 != s64, axis=0)
    pick = np.union1d(np.arange(J_ORACLE), np.nonzero(mism)[0])
    ref = oracle.OracleSampler(U_host, pick.size, R)
    sub = np.concatenate([draws[t * J + pick] for t in range(2 * N)])
    r = ref.draw(N + 1, sub)
    del ref
    sure = r["margins"] > 1e-10
    np.testing.assert_array_equal(s64[:, pick][:, sure], r["samples"][:, sure])
    assert sure[:J_ORACLE].mean() > 0.99
    where = np.isin(pick, np.nonzero(mism)[0])
    m = r["margins"][where]
    print(f"[gpu] fp32 mismatches {int(mism.sum())} of {J} ({mism.mean():.2e}); oracle margins "
          f"{np.array2string(m, precision=2)}", flush=True)
    assert np.all(m < eps32), f"FP32 mismatches with oracle margin >= {eps32}: {m[m >= eps32]}"
    return mism.mean()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("R", [16, 64, 128])
def test_c3_fp32_vs_fp64_full_size(ftl, oracle, R):
    import torch
    from fast_tensor_leverage_amd.synth import ill_conditioned_factors
    N, I, J = 3, 2 ** 23, 50_000
    dev = torch.device("cuda", 0)
    print(f"[gpu] C3 R={R}: ill-conditioned factors N={N} I={I}", flush=True)
    U, gen = ill_conditioned_factors(torch, N, I, R, dev, seed=3000 + R)
    draws = torch.rand(2 * N * J, generator=gen, device=dev, dtype=torch.float64)
    out = {}
    for prec in ("fp64", "fp32"):
        print(f"[gpu] C3 R={R}: {prec} sampler", flush=True)
        smp = ftl.Sampler.from_device(U, J, R, precision=prec)
        d_s = torch.zeros((N, J), dtype=torch.int64, device=dev)
        d_w = torch.empty((J,), dtype=torch.float64, device=dev)
        smp.draw_device(N + 1, d_s, None, d_w, draws)
        torch.cuda.synchronize()
        out[prec] = (d_s.cpu().numpy().view(np.uint64).copy(), d_w.cpu().numpy().copy())
        smp.close()
        del smp, d_s, d_w
    s64, w64 = out["fp64"]
    s32, w32 = out["fp32"]
    assert s32.max() < I and s64.max() < I

    # same-row leverage probability of the FP32 path against FP64
    G = torch.ones((R, R), dtype=torch.float64, device=dev)
    for M in U:
        G *= M.T @ M
    lam, V = torch.linalg.eigh(G)
    inv = torch.where(lam > 1e-8, 1.0 / torch.where(lam > 1e-8, lam, torch.ones_like(lam)), torch.zeros_like(lam))
    pinv = (V * inv) @ V.T
    rows = torch.from_numpy(s32.view(np.int64)).to(dev)
    h = torch.ones((J, R), dtype=torch.float64, device=dev)
    for k in range(N):
        h *= U[k][rows[k]]
    p64 = (torch.einsum("ir,rs,is->i", h, pinv, h) / R).cpu().numpy()
    p32 = 1.0 / (J * w32)
    ok = (p64 > 0) & np.isfinite(p32)
    rel = np.abs(p32[ok] - p64[ok]) / p64[ok]
    print(f"[gpu] C3 R={R}: same-row leverage probability rel. error mean {rel.mean():.2e} "
          f"max {rel.max():.2e} (rows with zero FP64 leverage: {int((~ok).sum())})", flush=True)
    assert ok.mean() > 0.999
    assert np.median(rel) < 1e-5 and rel.mean() < 1e-3

    print(f"[gpu] C3 R={R}: oracle on {J_ORACLE} samples plus the mismatches", flush=True)
    U_host = [M.cpu().numpy() for M in U]
    del U, h, rows
    torch.cuda.empty_cache()
    rate = _oracle_check(oracle, U_host, R, draws.cpu().numpy(), J, s64, s32, EPS32)
    assert rate <= 1e-4


def test_c4_serving_stream_512(ftl):
    from fast_tensor_leverage_amd.serve import random_requests, serve_stream
    reqs = random_requests(512, seed=0)
    budget = 96 << 30
    srv = ftl.Server(budget, 8)
    srv.wait(srv.submit([1024, 1024, 1024], 8, 1024, None, 1, "fp32"))  # kernel loading
    print("[gpu] C4: 512 requests at 500/s", flush=True)
    rec = serve_stream(srv, reqs, 500.0, seed=0, precision="fp32", keep_results=True)
    lat = rec["latency_s"]
    st = srv.stats()
    print(f"[gpu] C4: p50 {np.percentile(lat, 50) * 1e3:.2f} ms p99 {np.percentile(lat, 99) * 1e3:.2f} ms, "
          f"makespan {rec['makespan_s']:.2f} s, peak arena {st['peak_bytes'] / 2 ** 30:.1f} GiB", flush=True)
    assert np.all(np.isfinite(lat)) and len(rec["results"]) == len(reqs)
    assert st["peak_bytes"] <= budget and st["completed"] == len(reqs) + 1 and st["bytes_in_use"] == 0
    print("[gpu] C4: every request again, alone", flush=True)
    alone = ftl.Server(budget, 1)
    for i, q in enumerate(reqs):
        got = alone.wait(alone.submit(q["dims"], q["R"], q["J"], None, q["seed"], "fp32"))
        np.testing.assert_array_equal(rec["results"][i], got, err_msg=f"request {i} ({q})")
