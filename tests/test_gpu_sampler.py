"""Parity of the HIP sampler (libftl_hip.so) with the CPU FP64 oracle on
identical inputs and identical uniforms.

Bar: FP64 trees and computeM outputs bit-exact; sampled indices identical
wherever the oracle's margin (distance of a uniform to the CDF boundary it was
compared with) exceeds 1e-10; h identical and weights within 1e-10 relative.
FP32 mode: index mismatch rate vs the FP64 oracle reported and bounded.
"""
import zlib

import numpy as np
import pytest

from tests.helpers import exact_leverage, flat_index, multinomial_tv_band, philox_draws, tv_distance

pytestmark = pytest.mark.gpu

MARGIN = 1e-10

CASES = [
    # id, dims, R, J, j
    ("c1_readme", [1000] * 4, 8, 10_000, 5),
    ("dist_cmp", [8, 8, 8], 8, 50_000, 4),
    ("r32_i4096", [4096] * 3, 32, 20_000, 4),
    ("ragged_r25_excl1", [3000, 2500, 1777], 25, 10_000, 1),
    ("r16_excl_last", [5000, 4000], 16, 8_000, 1),
    ("r64_excl0", [2048, 1000, 3000], 64, 4_000, 0),
    ("r128", [700, 300], 128, 1_000, 2),
    ("n5_r8", [50, 60, 70, 80, 90], 8, 5_000, 2),
    ("n1", [1000], 8, 1_000, 1),
    ("tiny_rows", [1, 5, 7], 4, 1_000, 3),
    ("r1", [200, 300, 100], 1, 2_000, 4),
    ("r33_excl2", [600, 500, 400], 33, 3_000, 2),
    ("r63", [700, 600, 500], 63, 2_000, 4),
    ("r65_excl1", [700, 600, 500], 65, 2_000, 1),
]


def _factors(dims, R, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((I, R)) for I in dims]


def _compact_expected(oracle, I, R, fp32_levels):
    """Whether an FP32 sampler holds mode I's FP32 levels in FP32 only: it has
    FP32 levels, R <= 64, at least 64 nodes at the lowest FP64 level, the
    subtrees below it at most 5 levels deep (4 above R = 32), and a block's
    LDS (a node per leaf pair -- 16 x 16 tiles, or the packed triangle at
    R = 49-64 -- or four leaves and one node per depth) within 150 KB."""
    sh = oracle.tree_shape(I, min(R, I))
    depth = sh["lfill_level"] + (1 if sh["node_count"] > sh["nodes_upto_lfill"] else 0)
    d64 = max(1, depth + 1 - fp32_levels)
    if R > 64 or d64 > depth or d64 < 7:
        return False
    L = depth - (d64 - 1)
    nb = -(-R // 16)
    T = R * (R + 1) // 2
    ts64 = -(-T // 16) * 16
    # a slot per wave: 16 x 16 tiles and the (row, column) table, or the packed triangle at R = 49-64
    pairs = 2 ** (L - 1) * -(-T // 2) * 2 * 8 if nb == 4 else 2 ** (L - 1) * nb * (nb + 1) // 2 * 2048 + 2 * T
    return (L <= (5 if R <= 32 else 4) and pairs <= 150 * 1024 and (4 + L) * ts64 * 8 <= 150 * 1024)


def _fp64_nodes(oracle, I, R, fp32_levels):
    """Nodes held in FP64 by an FP32 sampler: the levels above its bottom
    fp32_levels (at least the root)."""
    sh = oracle.tree_shape(I, min(R, I))
    depth = sh["lfill_level"] + (1 if sh["node_count"] > sh["nodes_upto_lfill"] else 0)
    d64 = max(1, depth + 1 - fp32_levels)
    return min(2 ** d64 - 1, sh["node_count"])


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


def _ill_conditioned(dims, R, seed):
    """U = Q_U diag(geomspace(1, 1e-6, R)) Q_V^T with seeded QR factors (the
    shape of the mixed-precision study at small I)."""
    rng = np.random.default_rng(seed)
    s = np.geomspace(1.0, 1e-6, R)
    out = []
    for I in dims:
        qu, _ = np.linalg.qr(rng.standard_normal((I, R)))
        qv, _ = np.linalg.qr(rng.standard_normal((R, R)))
        out.append(np.ascontiguousarray((qu * s) @ qv.T))
    return out


ILL_CASES = [
    ("ill_r16", [3000, 2500, 2000], 16, 8_000, 4),
    ("ill_r32_excl1", [2048, 3000, 1500], 32, 6_000, 1),
]


@pytest.mark.parametrize("cid,dims,R,J,j", CASES, ids=[c[0] for c in CASES])
def test_fp64_parity(ftl, oracle, cid, dims, R, J, j):
    print(f"[gpu] fp64 parity case={cid} dims={dims} R={R} J={J} j={j}", flush=True)
    U = _factors(dims, R, seed=zlib.crc32(cid.encode()))
    _check_fp64_parity(ftl, oracle, cid, U, R, J, j, check_e_rows=True)


@pytest.mark.parametrize("cid,dims,R,J,j", ILL_CASES, ids=[c[0] for c in ILL_CASES])
def test_fp64_parity_ill_conditioned(ftl, oracle, cid, dims, R, J, j):
    """Ill-conditioned factors: same bar on trees, pinv and indices; the
    eigen rows of near-degenerate eigenvalues are not unique, so only their
    sum of rank-one terms is compared."""
    print(f"[gpu] fp64 parity case={cid} dims={dims} R={R} J={J} j={j}", flush=True)
    U = _ill_conditioned(dims, R, seed=zlib.crc32(cid.encode()))
    _check_fp64_parity(ftl, oracle, cid, U, R, J, j, check_e_rows=False)


# Eigensolver block sizes on either side of the rank where a thread's
# rotation items stop fitting one 2x2 block and two V entries (R <= 32 at 256
# threads, R <= 44 at 512, R <= 64 at 1024): (R, threads, j)
JACOBI_CASES = [(32, 256, 4), (33, 256, 0), (44, 512, 4), (45, 512, 2), (63, 1024, 1), (64, 1024, 4), (65, 1024, 4)]


@pytest.mark.parametrize("R,bs,j", JACOBI_CASES, ids=[f"r{c[0]}_bs{c[1]}" for c in JACOBI_CASES])
def test_fp64_parity_jacobi_block_sizes(ftl, oracle, R, bs, j):
    cid = f"jacobi_r{R}_bs{bs}"
    print(f"[gpu] fp64 parity case={cid} R={R} jacobi_threads={bs} j={j}", flush=True)
    U = _factors([600, 500, 400], R, seed=zlib.crc32(cid.encode()))
    _check_fp64_parity(ftl, oracle, cid, U, R, 2_000, j, check_e_rows=True, options={"jacobi_threads": bs})


# one-sample-per-block draw kernel (draw_kernel=2) at R > 32: id, dims, R, J,
# j, threads per block
BLOCK_CASES = [
    ("blk_r64_excl0", [2048, 1000, 3000], 64, 4_000, 0, 256),
    ("blk_r64_bs128", [2048, 1000, 3000], 64, 4_000, 4, 128),
    ("blk_r48_short_mode", [1500, 900, 47], 48, 3_000, 4, 256),
    ("blk_r128", [700, 300], 128, 1_000, 2, 256),
    ("blk_r100_bs512", [900, 400, 333], 100, 1_500, 1, 512),
    ("blk_r128_bs128", [1200, 500], 128, 1_000, 2, 128),
]


@pytest.mark.parametrize("cid,dims,R,J,j,bs", BLOCK_CASES, ids=[c[0] for c in BLOCK_CASES])
def test_fp64_parity_block_kernel(ftl, oracle, cid, dims, R, J, j, bs):
    print(f"[gpu] fp64 parity block kernel case={cid} dims={dims} R={R} J={J} j={j} bs={bs}", flush=True)
    U = _factors(dims, R, seed=zlib.crc32(cid.encode()))
    _check_fp64_parity(ftl, oracle, cid, U, R, J, j, check_e_rows=True,
                       options={"draw_kernel": 2, "block_threads": bs})


@pytest.mark.parametrize("cid,dims,R,J,j,bs", BLOCK_CASES, ids=[c[0] for c in BLOCK_CASES])
def test_fp32_mismatch_rate_block_kernel(ftl, oracle, cid, dims, R, J, j, bs):
    print(f"[gpu] fp32 block kernel case={cid} bs={bs}", flush=True)
    U = _factors(dims, R, seed=zlib.crc32(cid.encode()))
    N = len(dims)
    ref = oracle.OracleSampler(U, J, R)
    gpu = ftl.Sampler(U, J, R, "efficient", precision="fp32")
    gpu.set_option("draw_kernel", 2)
    gpu.set_option("block_threads", bs)
    draws = np.random.default_rng(11).random(2 * N * J)
    out = ref.draw(j, draws)
    samples = np.zeros((N, J), dtype=np.uint64)
    gpu.KRPDrawSamples(j, samples, draws)
    diff = np.zeros(J, dtype=bool)
    for k in range(N):
        if not (j < N and k == j):
            diff |= samples[k] != out["samples"][k]
    print(f"[gpu] {cid}: fp32 mismatch rate {diff.mean():.2e}", flush=True)
    assert diff.mean() <= 2e-3


def _check_fp64_parity(ftl, oracle, cid, U, R, J, j, check_e_rows, options=None):
    N = len(U)
    ref = oracle.OracleSampler(U, J, R)
    gpu = ftl.Sampler(U, J, R, "efficient", precision="fp64")
    for key, val in (options or {}).items():
        gpu.set_option(key, val)
    # trees: bit-exact
    for k in range(N):
        np.testing.assert_array_equal(gpu.get_tree(k), ref.tree(k))
    # computeM: pinv and eigen matrices
    p_ref, E_ref = ref.compute_m(j)
    p_gpu, E_gpu = gpu.compute_m(j)
    iu = np.triu_indices(R)
    np.testing.assert_allclose(p_gpu[iu], p_ref[iu], rtol=1e-12, atol=1e-14 * np.abs(p_ref).max())
    for k in range(N):
        if j < N and k == j:
            continue
        # eigenvector rows agree up to sign and rounding (~eps/gap), and both
        # reproduce the same sum of rank-one terms
        if check_e_rows:
            np.testing.assert_allclose(np.abs(E_gpu[k]), np.abs(E_ref[k]), rtol=1e-9,
                                       atol=1e-10 * np.abs(E_ref[k]).max())
        ref_sum = E_ref[k].T @ E_ref[k]
        np.testing.assert_allclose(E_gpu[k].T @ E_gpu[k], ref_sum, rtol=1e-10,
                                   atol=1e-12 * np.abs(ref_sum).max())
    draws = np.random.default_rng(7).random(2 * N * J)
    out = ref.draw(j, draws)
    samples = np.zeros((N, J), dtype=np.uint64)
    h = np.zeros((J, R))
    w = np.zeros(J)
    gpu.KRPDrawSamples_materialize(j, samples, h, w, draws)
    modes = [k for k in range(N) if not (j < N and k == j)]
    sure = out["margins"] > MARGIN
    diff = np.zeros(J, dtype=bool)
    for k in modes:
        diff |= samples[k] != out["samples"][k]
    print(f"[gpu] {cid}: mismatches {int(diff.sum())} / {J}, near-boundary {int((~sure).sum())}", flush=True)
    assert not np.any(diff & sure), np.flatnonzero(diff & sure)[:10]
    same = ~diff
    np.testing.assert_array_equal(h[same], out["h"][same])
    np.testing.assert_allclose(w[same], out["weights"][same], rtol=1e-10)
    if j < N:
        assert np.all(samples[j] == 0)


@pytest.mark.parametrize("cid,dims,R,J,j", CASES[:7], ids=[c[0] for c in CASES[:7]])
def test_fp32_mismatch_rate(ftl, oracle, cid, dims, R, J, j):
    print(f"[gpu] fp32 case={cid}", flush=True)
    U = _factors(dims, R, seed=zlib.crc32(cid.encode()))
    N = len(dims)
    ref = oracle.OracleSampler(U, J, R)
    gpu = ftl.Sampler(U, J, R, "efficient", precision="fp32")
    draws = np.random.default_rng(11).random(2 * N * J)
    out = ref.draw(j, draws)
    samples = np.zeros((N, J), dtype=np.uint64)
    gpu.KRPDrawSamples(j, samples, draws)
    modes = [k for k in range(N) if not (j < N and k == j)]
    diff = np.zeros(J, dtype=bool)
    for k in modes:
        diff |= samples[k] != out["samples"][k]
    rate = diff.mean()
    print(f"[gpu] {cid}: fp32 mismatch rate {rate:.2e}, largest oracle margin of a mismatch "
          f"{out['margins'][diff].max() if diff.any() else 0:.2e}", flush=True)
    assert rate <= 2e-3
    # FP32 storage moves a cutoff by about 2^-24 of the mass below it: every
    # mismatch lies that close to a boundary of the FP64 oracle's descent
    assert np.all(out["margins"][diff] < 1e-6)


def test_builtin_generator_matches_philox(ftl, oracle):
    """draws=None uses Philox4x32-10; the NumPy restatement reproduces it."""
    dims, R, J = [300, 200, 100], 8, 4000
    U = _factors(dims, R, 5)
    gpu = ftl.Sampler(U, J, R)
    gpu.set_seed(1234, 5)
    samples = np.zeros((3, J), dtype=np.uint64)
    gpu.KRPDrawSamples(4, samples)
    draws = philox_draws(1234, 5, 3, J)
    out = oracle.OracleSampler(U, J, R).draw(4, draws)
    sure = out["margins"] > MARGIN
    assert np.array_equal(samples[:, sure], out["samples"][:, sure])
    # the counter advances: a second call draws different uniforms
    s2 = np.zeros_like(samples)
    gpu.KRPDrawSamples(4, s2)
    assert not np.array_equal(s2, samples)


def test_distribution_tv(ftl):
    """Empirical TV to exact leverage within the multinomial noise band
    (the check of compare_distributions.py, made quantitative)."""
    N, I, R, J = 3, 8, 8, 200_000
    U = _factors([I] * N, R, 2024)
    p = exact_leverage(U)
    mean, std = multinomial_tv_band(p, J)
    for prec in ("fp64", "fp32"):
        gpu = ftl.Sampler(U, J, R, precision=prec)
        samples = np.zeros((N, J), dtype=np.uint64)
        gpu.KRPDrawSamples(N + 1, samples)
        emp = np.bincount(flat_index(samples, [I] * N), minlength=p.size) / J
        tv = tv_distance(emp, p)
        print(f"[gpu] TV {prec}: {tv:.4f} (band {mean:.4f} +- {std:.4f})", flush=True)
        assert abs(tv - mean) < 5 * std + 1e-3


def test_update_sampler(ftl, oracle):
    dims, R, J = [500, 400, 300], 16, 2000
    U = _factors(dims, R, 9)
    gpu = ftl.Sampler(U, J, R)
    new = np.random.default_rng(10).standard_normal((400, R))
    gpu.update_sampler(1, new)
    ref = oracle.OracleSampler([U[0], new, U[2]], J, R)
    for k in range(3):
        np.testing.assert_array_equal(gpu.get_tree(k), ref.tree(k))
    draws = np.random.default_rng(1).random(6 * J)
    out = ref.draw(4, draws)
    samples = np.zeros((3, J), dtype=np.uint64)
    gpu.KRPDrawSamples(4, samples, draws)
    sure = out["margins"] > MARGIN
    assert np.array_equal(samples[:, sure], out["samples"][:, sure])


def test_device_draws_and_updates_are_ordered(ftl, oracle):
    """Device draws on a stream of the caller's, then update_sampler with no
    host synchronisation between them, then another device draw: each draw
    sees the factors and trees current at its call."""
    import torch
    rng = np.random.default_rng(61)
    dims, R, J = (400000, 3000, 2000), 16, 20000
    U = [rng.standard_normal((d, R)) for d in dims]
    U1 = rng.standard_normal((dims[1], R)) * rng.random((dims[1], 1))
    s = ftl.Sampler(U, J, R, "efficient")
    dev = torch.device("cuda:0")
    draws = rng.random(2 * 3 * J)
    td = torch.from_numpy(draws).to(dev)
    outs = [torch.zeros((3, J), dtype=torch.int64, device=dev) for _ in range(4)]
    st = torch.cuda.Stream()
    assert st.cuda_stream != 0
    st.wait_stream(torch.cuda.current_stream())
    for o in outs[:3]:
        s.draw_device(4, o, None, None, td, st)
    s.update_sampler(1, U1)  # must not overwrite factor 1 or its tree under the running draws
    s.draw_device(4, outs[3], None, None, td, st)
    st.synchronize()
    before = oracle.OracleSampler(U, J, R).draw(4, draws)
    after = oracle.OracleSampler([U[0], U1, U[2]], J, R).draw(4, draws)
    for o, ref in zip(outs, [before] * 3 + [after]):
        got = o.cpu().numpy().view(np.uint64)
        sure = ref["margins"] > MARGIN
        np.testing.assert_array_equal(got[:, sure], ref["samples"][:, sure])


def test_argument_errors(ftl):
    U = _factors([100, 100], 4, 1)
    gpu = ftl.Sampler(U, 10, 4)
    with pytest.raises(TypeError):
        gpu.KRPDrawSamples(2, np.zeros((2, 10), dtype=np.int64))
    with pytest.raises(ValueError):
        gpu.KRPDrawSamples(2, np.zeros((3, 10), dtype=np.uint64))
    with pytest.raises(ValueError):
        ftl.Sampler([np.zeros((10, 4))], 10, 4).KRPDrawSamples(0, np.zeros((1, 10), dtype=np.uint64))
    with pytest.raises(ValueError):
        gpu.update_sampler(5)


def test_device_factors_ordered_after_default_stream(ftl):
    """Factors produced by torch on the default stream and handed over
    without a host synchronisation are complete when the sampler reads them
    (regression: a non-blocking sampler stream read them too early)."""
    import torch
    dev = torch.device("cuda", 0)
    N, I, R, J = 3, 2 ** 18, 16, 4000
    gen = torch.Generator(device=dev)
    gen.manual_seed(77)
    U = []
    for _ in range(N):
        q, _ = torch.linalg.qr(torch.randn((I, R), generator=gen, device=dev, dtype=torch.float64))
        U.append((q * torch.logspace(0, -3, R, dtype=torch.float64, device=dev)).contiguous())
    s_dev = ftl.Sampler.from_device(U, J, R)  # no torch.cuda.synchronize() before this
    U_host = [u.cpu().numpy() for u in U]
    s_host = ftl.Sampler(U_host, J, R)
    draws = np.random.default_rng(3).random(2 * N * J)
    a = np.zeros((N, J), dtype=np.uint64)
    b = np.zeros((N, J), dtype=np.uint64)
    s_dev.KRPDrawSamples(N + 1, a, draws)
    s_host.KRPDrawSamples(N + 1, b, draws)
    np.testing.assert_array_equal(a, b)


def test_overlapped_compute_m_matches_serial(ftl):
    """computeM of a call runs on a side stream into one of two buffer sets
    while the previous call's draw kernels run; back-to-back device calls
    (changing j, a factor updated in between) give the same samples and
    weights as the serial order."""
    import torch
    dev = torch.device("cuda", 0)
    dims, R, J = [3000, 2000, 1500], 16, 4000
    N = len(dims)
    U = _factors(dims, R, 5)
    new = np.random.default_rng(6).standard_normal((dims[1], R))
    js = [4, 0, 1, 4, 2, 4, 1]
    draws = [torch.from_numpy(np.random.default_rng(100 + t).random(2 * N * J)).to(dev) for t in range(len(js))]
    outs = {}
    for overlap in (1, 0):
        print(f"[gpu] overlapped computeM: overlap_compute_m={overlap}", flush=True)
        s = ftl.Sampler(U, J, R)
        s.set_option("overlap_compute_m", overlap)
        assert s.get_option("overlap_compute_m") == overlap
        res = [torch.zeros((N, J), dtype=torch.int64, device=dev) for _ in js]
        wts = [torch.zeros(J, dtype=torch.float64, device=dev) for _ in js]
        for t, j in enumerate(js):
            s.draw_device(j, res[t], None, wts[t], draws[t])
            if t == 3:
                s.update_sampler(1, new)
        torch.cuda.synchronize()
        outs[overlap] = ([r.cpu().numpy() for r in res], [w.cpu().numpy() for w in wts])
        s.close()
    for a, b in zip(outs[1][0], outs[0][0]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(outs[1][1], outs[0][1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("R", [8, 32, 64])
def test_fp32_copies_written_by_the_build_equal_a_separate_conversion(ftl, R):
    """The tree build writes the FP32 factor rows and bottom levels itself;
    rebuilding the FP32 levels afterwards by a separate conversion
    (fp32_levels reset) gives bitwise the same draws."""
    print(f"[gpu] fused fp32 copies R={R}", flush=True)
    dims = [3000 + R, 2000, 900]
    U = _factors(dims, R, 200 + R)
    J = 4000
    s = ftl.Sampler(U, J, R, "efficient", precision="fp32")
    draws = np.random.default_rng(R).random(2 * 3 * J)
    a = np.zeros((3, J), dtype=np.uint64)
    wa = np.zeros(J)
    ha = np.zeros((J, R))
    s.KRPDrawSamples_materialize(4, a, ha, wa, draws)
    s.set_option("fp32_levels", 0)  # the default again: the FP32 levels are re-converted from FP64
    b = np.zeros((3, J), dtype=np.uint64)
    wb = np.zeros(J)
    hb = np.zeros((J, R))
    s.KRPDrawSamples_materialize(4, b, hb, wb, draws)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(wa, wb)
    np.testing.assert_array_equal(ha, hb)
    # an updated factor goes through the same fused build
    s.update_sampler(1, U[1] * 0.5)
    s.KRPDrawSamples_materialize(4, a, ha, wa, draws)
    s.set_option("fp32_levels", 0)
    s.KRPDrawSamples_materialize(4, b, hb, wb, draws)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(wa, wb)


@pytest.mark.parametrize("R", [16, 32, 64])
def test_fp32_sampler_matrix_core_leaf_grams(ftl, oracle, R):
    """FP32 samplers build their leaf Grams on the matrix cores: the FP64 tree
    equals the oracle's to rounding; with leaf_mfma = 0 the sequential sums
    give it bitwise; the draws of the two trees agree on all but a few
    boundary samples."""
    print(f"[gpu] matrix-core leaf grams R={R}", flush=True)
    dims = [5000 + R, 700, 300]
    U = _factors(dims, R, 300 + R)
    J = 4000
    ref = oracle.OracleSampler(U, J, R)
    s = ftl.Sampler(U, J, R, "efficient", precision="fp32")
    assert s.get_option("leaf_mfma") == 1
    for k in range(3):
        t_gpu, t_ref = s.get_tree(k), ref.tree(k)
        n64 = _fp64_nodes(oracle, dims[k], R, s.get_option("fp32_levels"))
        # FP64 levels to rounding; the FP32 levels (held in FP32 only) to FP32 rounding
        np.testing.assert_allclose(t_gpu[:n64], t_ref[:n64], rtol=1e-12, atol=1e-13 * np.abs(t_ref).max())
        np.testing.assert_allclose(t_gpu[n64:], t_ref[n64:], rtol=2e-7, atol=1e-7 * np.abs(t_ref).max())
    draws = np.random.default_rng(R).random(6 * J)
    a = np.zeros((3, J), dtype=np.uint64)
    s.KRPDrawSamples(4, a, draws)
    s.set_option("leaf_mfma", 0)
    for k in range(3):
        s.update_sampler(k, U[k])
        np.testing.assert_array_equal(s.get_tree(k), ref.tree(k))
    b = np.zeros((3, J), dtype=np.uint64)
    s.KRPDrawSamples(4, b, draws)
    assert np.mean(np.any(a != b, axis=0)) < 1e-3



@pytest.mark.parametrize("dims,R,levels", [
    ([70001, 40000, 300], 32, 5),             # ragged trees, the headline split; a small tree built level-wise
    ([2 ** 18, 1024 * 80, 257], 8, 1),        # power-of-two and ragged, one FP32 level
    ([30000, 20000, 900], 16, 2),
    ([2 ** 17 + 3, 60000, 64], 25, 5),        # R = 25 (padded tiles)
    ([2 ** 15 + 3, 900, 64], 25, 7),          # seven FP32 levels: too deep for one block, full build
    ([60000, 15000, 130], 48, 4),
    ([90000, 70000, 65], 64, 3),              # and a factor of fewer rows than two leaves' worth
    ([9000, 700, 65], 64, 4),                 # R = 64, trees too small to hold compact: full build
    ([2 ** 16 + 5, 40000, 64], 64, 4),        # R = 64 four levels deep: packed-triangle slots
    ([70001, 2 ** 16, 300], 49, 4),           # R = 49: packed slots, partial last tile
])
def test_compact_fp32_trees_equal_the_full_build(ftl, oracle, dims, R, levels):
    """FP32 samplers hold their FP32 levels only in FP32 (the subtree build):
    the draws, h and weights are bitwise those of the build that also keeps
    FP64 copies of those levels (compact_fp32 = 0), the FP64 levels are
    bitwise equal, the FP32 levels are the FP32 rounding of the full build's
    FP64 nodes, and the device footprint is smaller by the FP64 copies."""
    print(f"[gpu] compact fp32 trees dims={dims} R={R} fp32_levels={levels}", flush=True)
    U = _factors(dims, R, 7 * R + levels)
    J = 3000
    a = ftl.Sampler(U, J, R, "efficient", precision="fp32")
    a.set_option("fp32_levels", levels)
    b = ftl.Sampler(U, J, R, "efficient", precision="fp32")
    b.set_option("compact_fp32", 0)
    b.set_option("fp32_levels", levels)
    assert a.get_option("compact_modes") == sum(1 for I in dims if _compact_expected(oracle, I, R, levels))
    assert b.get_option("compact_modes") == 0
    for k in range(3):
        ta, tb = a.get_tree(k), b.get_tree(k)
        n64 = _fp64_nodes(oracle, dims[k], R, levels)
        np.testing.assert_array_equal(ta[:n64], tb[:n64])
        if _compact_expected(oracle, dims[k], R, levels):  # FP32 levels held in FP32 only
            np.testing.assert_array_equal(ta[n64:], tb[n64:].astype(np.float32).astype(np.float64))
        else:
            np.testing.assert_array_equal(ta[n64:], tb[n64:])
    draws = np.random.default_rng(R).random(6 * J)
    for j in (0, 4):
        out = []
        for s in (a, b):
            smp = np.zeros((3, J), dtype=np.uint64)
            h = np.zeros((J, R))
            w = np.zeros(J)
            s.KRPDrawSamples_materialize(j, smp, h, w, draws)
            out.append((smp, h, w))
        for x, y in zip(out[0], out[1]):
            np.testing.assert_array_equal(x, y)
    # an update goes through the same subtree build
    Un = U[1] * 0.5 + 0.1
    for s in (a, b):
        s.update_sampler(1, Un)
    np.testing.assert_array_equal(a.get_tree(1)[:_fp64_nodes(oracle, dims[1], R, levels)],
                                  b.get_tree(1)[:_fp64_nodes(oracle, dims[1], R, levels)])
    sa, sb = np.zeros((3, J), dtype=np.uint64), np.zeros((3, J), dtype=np.uint64)
    a.KRPDrawSamples(4, sa, draws)
    b.KRPDrawSamples(4, sb, draws)
    np.testing.assert_array_equal(sa, sb)
    saved = sum(8 * (oracle.tree_shape(I, min(R, I))["node_count"] - _fp64_nodes(oracle, I, R, levels))
                * (-(-(R * (R + 1) // 2) // 16) * 16) for I in dims if _compact_expected(oracle, I, R, levels))
    assert b.info()["device_bytes"] - a.info()["device_bytes"] >= saved - 4096


def test_compact_fp32_trees_random_shapes(ftl, oracle):
    """Seeded random shapes (ragged and power-of-two trees, R from 5 to 64,
    one to five FP32 levels): draws of the compact FP32 trees equal those of
    the full build bitwise, whether or not a mode is held compact."""
    rng = np.random.default_rng(2027)
    J = 2000
    held = 0
    for case in range(16):
        R = int(rng.choice([5, 8, 13, 16, 24, 32, 40, 48, 57, 64]))
        levels = int(rng.integers(1, 6))
        dims = [int(R * 2 ** rng.uniform(7, 12.5)) + int(rng.integers(0, R)) for _ in range(int(rng.integers(2, 5)))]
        print(f"[gpu] random compact case {case}: dims={dims} R={R} fp32_levels={levels}", flush=True)
        U = _factors(dims, R, 5000 + case)
        a = ftl.Sampler(U, J, R, "efficient", precision="fp32")
        a.set_option("fp32_levels", levels)
        b = ftl.Sampler(U, J, R, "efficient", precision="fp32")
        b.set_option("compact_fp32", 0)
        b.set_option("fp32_levels", levels)
        held += a.get_option("compact_modes")
        assert a.get_option("compact_modes") == sum(1 for I in dims if _compact_expected(oracle, I, R, levels))
        draws = rng.random(2 * len(dims) * J)
        sa = np.zeros((len(dims), J), dtype=np.uint64)
        sb = np.zeros((len(dims), J), dtype=np.uint64)
        wa, wb = np.zeros(J), np.zeros(J)
        ha, hb = np.zeros((J, R)), np.zeros((J, R))
        a.KRPDrawSamples_materialize(len(dims) + 1, sa, ha, wa, draws)
        b.KRPDrawSamples_materialize(len(dims) + 1, sb, hb, wb, draws)
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(wa, wb)
        a.close()
        b.close()
    assert held > 10  # most cases exercise the subtree build
