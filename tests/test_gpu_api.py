"""Host-side surfaces of the HIP library that the parity tests do not reach:
kernel timing and tuning knobs of the sampler, the device-resident ALS
(host_sync=False: sigma and factors pulled and pushed on request), the
residual of PyLowRank against a PySparseTensor, and the file readers feeding
the device format."""
import numpy as np
import pytest

from tests.helpers import cp_dense, dense_from_sparse, random_sparse

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


def test_sampler_timing_and_tuning_knobs(ftl):
    """last_timing averages the kernels of the draws since timing was
    enabled; set_tuning sets the lane count and the grid cap, and a draw with
    them gives the samples of the default launch shape."""
    print("[gpu] sampler timing and tuning knobs", flush=True)
    rng = np.random.default_rng(3)
    R, J = 16, 3000
    U = [rng.standard_normal((I, R)) for I in (5000, 4000, 3000)]
    s = ftl.Sampler(U, J, R, "efficient")
    s.set_option("draw_kernel", 0)  # the lane-group kernel whatever the tuning table picks
    draws = rng.random(6 * J)
    base = np.zeros((3, J), dtype=np.uint64)
    s.KRPDrawSamples(3, base, draws)
    s.enable_timing(True)
    for _ in range(3):
        s.KRPDrawSamples(3, np.zeros((3, J), dtype=np.uint64), draws)
    t = s.last_timing()
    assert t["calls"] == 3
    assert t["draw_ms"] > 0 and t["gram_pinv_ms"] > 0 and t["eigen_modes_ms"] > 0
    s.enable_timing(False)
    s.set_tuning(group_lanes=16, grid_blocks=7)
    assert s.get_option("group_lanes") == 16 and s.get_option("grid_blocks") == 7
    out = np.zeros((3, J), dtype=np.uint64)
    s.KRPDrawSamples(3, out, draws)
    np.testing.assert_array_equal(out, base)
    with pytest.raises(Exception):
        s.set_tuning(group_lanes=12)
    s.close()


def test_device_resident_als_syncs_on_request(ftl):
    """ALS(host_sync=False) keeps factors and sigma on the device: get_sigma
    pulls sigma, sync_to_host copies every factor, and sync_from_host pushes
    a factor changed in place and rebuilds its tree, after which the next
    update draws what a new sampler on the same factors draws."""
    print("[gpu] device-resident ALS", flush=True)
    rng = np.random.default_rng(21)
    dims, R, J = (40, 35, 30), 6, 500
    idx, vals = random_sparse(dims, 3000, rng, unique=False)
    U0 = [rng.standard_normal((d, R)) for d in dims]
    lr = ftl.LowRankTensor(R, [u.copy() for u in U0])
    als = ftl.ALS(lr, ftl.SparseTensor(idx, vals), host_sync=False)
    als.initialize_ds_als(J, "efficient")
    als.execute_ds_als_update(0, True, True, random_draws=rng.random(6 * J))
    sigma = np.zeros(R)
    lr.get_sigma(sigma)
    assert np.all(sigma > 0) and not np.allclose(sigma, 1.0)
    als.sync_to_host()
    # mode 0 was renormalized: unit columns on the touched rows
    norms = np.linalg.norm(lr.U[0], axis=0)
    np.testing.assert_allclose(norms, 1.0, rtol=1e-12)
    lr.U[2][:] = rng.standard_normal(lr.U[2].shape)
    als.sync_from_host(2)
    draws = rng.random(6 * J)
    current = [u.copy() for u in lr.U]
    als.execute_ds_als_update(1, True, True, random_draws=draws)
    samples, h, w = als.last_draw()
    fresh = ftl.Sampler(current, J, R, "efficient")
    s2, h2, w2 = np.zeros((3, J), dtype=np.uint64), np.zeros((J, R)), np.zeros(J)
    fresh.KRPDrawSamples_materialize(1, s2, h2, w2, draws)
    np.testing.assert_array_equal(samples, s2)
    np.testing.assert_array_equal(h, h2)
    fresh.close()


def test_compute_diff_resid_against_dense(ftl):
    """PyLowRank.compute_diff_resid and compute_exact_fit against the dense
    residual ||X - [[sigma; U]]|| of a small tensor, with and without an ALS
    holding the decomposition."""
    print("[gpu] compute_diff_resid", flush=True)
    rng = np.random.default_rng(8)
    dims, R = (12, 10, 9), 3
    idx, vals = random_sparse(dims, 400, rng, unique=True)
    rhs = ftl.PySparseTensor.from_arrays(idx, vals, dims=list(dims))
    lhs = ftl.PyLowRank(dims, R, seed=4)
    sigma = np.array([1.5, -0.5, 2.0])
    lhs.set_sigma(sigma)
    X = dense_from_sparse(idx, vals, dims)
    M = cp_dense(lhs.U, sigma)
    want = np.linalg.norm(X - M)
    np.testing.assert_allclose(lhs.compute_diff_resid(rhs), want, rtol=1e-10)
    np.testing.assert_allclose(lhs.compute_exact_fit(rhs), 1.0 - want / np.linalg.norm(X), rtol=1e-10)
    als = ftl.ALS(lhs.ten, rhs.ten)
    np.testing.assert_allclose(lhs.compute_diff_resid(rhs), want, rtol=1e-10)
    del als


def test_file_readers_feed_the_device_tensor(ftl, tmp_path):
    """The .npz (the reference's HDF5 dataset names) and FROSTT .tns readers
    build device tensors whose norm, dims and MTTKRP-visible nonzeros are
    those of the arrays written."""
    print("[gpu] file readers", flush=True)
    rng = np.random.default_rng(12)
    dims = (7, 9, 5)
    idx, vals = random_sparse(dims, 120, rng, unique=True)
    npz = tmp_path / "t.npz"
    np.savez(npz, MODE_0=idx[:, 0] + 1, MODE_1=idx[:, 1] + 1, MODE_2=idx[:, 2] + 1, VALUES=vals,
             MIN_MODE_SET=np.array([1, 1, 1]), MAX_MODE_SET=np.array(dims))
    tns = tmp_path / "t.tns"
    with open(tns, "w") as f:
        f.write("# comment\n")
        for row, v in zip(idx, vals):
            f.write(" ".join(str(int(c) + 1) for c in row) + f" {float(v)!r}\n")
    seen = [int(idx[:, k].max()) + 1 for k in range(3)]  # .tns sizes: the largest coordinate
    for t, want in ((ftl.PySparseTensor.from_npz(str(npz)), list(dims)),
                    (ftl.PySparseTensor.from_tns(str(tns)), seen)):
        assert t.nnz == len(vals) and list(t.dims) == want
        np.testing.assert_allclose(t.ten.get_normsq(), float(np.sum(vals ** 2)), rtol=1e-13)


def test_destroyed_sampler_memory_is_reused(ftl):
    """A destroyed sampler's device blocks are kept and handed to the next
    sampler of the same shape, whose draws are those of the first; releasing
    the kept blocks empties the pool."""
    print("[gpu] block pool", flush=True)
    ftl.release_cached_memory()
    assert ftl.cached_memory_bytes() == 0
    rng = np.random.default_rng(44)
    R, J = 24, 4000
    U = [rng.standard_normal((I, R)) for I in (30000, 20000, 9000)]
    draws = rng.random(6 * J)
    out = []
    for prec in ("fp32", "fp32", "fp64"):
        s = ftl.Sampler(U, J, R, "efficient", precision=prec)
        held = s.info()["device_bytes"]
        kept_before = ftl.cached_memory_bytes()
        smp = np.zeros((3, J), dtype=np.uint64)
        s.KRPDrawSamples(3, smp, draws)
        out.append(smp)
        s.close()
        assert ftl.cached_memory_bytes() >= kept_before + held
    np.testing.assert_array_equal(out[0], out[1])
    ftl.release_cached_memory()
    assert ftl.cached_memory_bytes() == 0


def test_block_pool_limit_evicts_the_oldest(ftl):
    """Under a lowered limit the pool keeps at most that many bytes, the
    blocks of the most recently destroyed sampler first; a limit of 0 keeps
    nothing."""
    print("[gpu] block pool limit", flush=True)
    old = ftl.cached_memory_limit()
    try:
        ftl.release_cached_memory()
        limit = 48 << 20
        ftl.set_cached_memory_limit(limit)
        assert ftl.cached_memory_limit() == limit
        rng = np.random.default_rng(45)
        for I in (40000, 60000, 80000):  # about 10, 15 and 20 MB of blocks each
            U = [rng.standard_normal((I, 16)) for _ in range(3)]
            s = ftl.Sampler(U, 2000, 16, "efficient", precision="fp64")
            s.close()
            assert 0 < ftl.cached_memory_bytes() <= limit
        ftl.set_cached_memory_limit(0)
        assert ftl.cached_memory_bytes() == 0
        s = ftl.Sampler(U, 2000, 16, "efficient", precision="fp64")
        s.close()
        assert ftl.cached_memory_bytes() == 0
    finally:
        ftl.set_cached_memory_limit(old)
