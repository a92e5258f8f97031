"""Sampled CP-ALS at scale on Zipf(1.1) sparse tensors.

* als_prod fit trajectory: a 3-way tensor of 2x10^6 nonzeros (dims 2^16 x 2^15
  x 2^12, rank-25 CP values + 1% noise), R=25, J=2^16, 10 sweeps with a fit
  after every sweep.  The oracle's sweep loop (als.py:59-107: every mode
  once per sweep, the exact fit after each epoch, the stop rule on the last
  three fits) runs on identical initial factors and identical uniforms (the
  library's Philox stream restated in NumPy); the GPU trajectory's fits agree
  with it within 1e-6 relative at every fit point and both stop after the
  same sweep.  Also for method "exact" (execute_exact_als_update).
* C5 at full size: dims 2^22 x 2^21 x 2^17, 2x10^8 nonzeros generated on the
  device, R=25, J=2^16.  Checked through properties that hold at any size:
  ||X||^2; matched counts and ids of sampled tuples against a stable sort +
  binary search of the packed coordinates (torch, sort_lookup.hpp:33-132);
  one execute_ds_als_update's U_j rows against the dense sampled
  least-squares solution (sum_s w_s X(r, tuple_s) h_s) pinv(sum_s w_s h_s
  h_s^T) on 256 touched rows (als.hpp:57-220); the device fit against a
  chunked FP64 residual ||X||^2 - 2 <X, M> + ||M||^2 (sort_lookup.hpp:134-197).
* The host-sample MTTKRP refuses a match total of 2^32 or more.
"""
import numpy as np
import pytest

from tests.helpers import philox_draws

pytestmark = pytest.mark.gpu

SMALL = ((2 ** 16, 2 ** 15, 2 ** 12), 2_000_000, 25, 2 ** 16)


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


@pytest.fixture(scope="module")
def small_tensor():
    import torch
    from fast_tensor_leverage_amd.synth import zipf_sparse
    dims, nnz, R, _ = SMALL
    print(f"[gpu] als trajectory: Zipf tensor dims={dims} nnz={nnz}", flush=True)
    idx, vals = zipf_sparse(torch, dims, nnz, R, torch.device("cuda", 0), seed=77)
    return idx.cpu().numpy().view(np.uint32).copy(), vals.cpu().numpy().copy()


def _oracle_trajectory(oracle, idx, vals, U0, sigma0, J, method, seed, sweeps, epoch, tol):
    """The sweep loop of als.py:59-107 over the oracle's ALS step."""
    N = len(U0)
    ref = oracle.OracleALS(oracle.OracleSparse(idx, vals), [u.copy() for u in U0], J, sigma=sigma0)
    iterations, fits = [0], [ref.fit()]
    call = 0
    for i in range(sweeps):
        for j in range(N):
            if method == "exact":
                ref.exact_update(j)
            else:
                ref.update(j, philox_draws(seed, call, N, J))
                call += 1
        if (i + 1) % epoch == 0:
            iterations.append(i + 1)
            fits.append(ref.fit())
            if len(fits) > 4 and np.max(fits[-3:]) < np.max(fits[:-3]) + tol:
                break
    return iterations, fits


@pytest.mark.timeout(600)
@pytest.mark.parametrize("method", ["efficient", "exact"])
def test_als_prod_trajectory_matches_oracle(ftl, oracle, small_tensor, method):
    idx, vals = small_tensor
    dims, nnz, R, J = SMALL
    sweeps, epoch, tol, seed = 10, 1, 1e-4, 123
    lhs = ftl.PyLowRank(dims, R, seed=5)
    lhs.ten.renormalize_columns(-1)
    U0 = [u.copy() for u in lhs.U]
    sigma0 = lhs.ten.sigma.copy()
    rhs = ftl.PySparseTensor.from_arrays(idx, vals, dims=list(dims))
    print(f"[gpu] als_prod {method}: {sweeps} sweeps on the GPU", flush=True)
    rec = ftl.als_prod(lhs, rhs, J, method, sweeps, tol, epoch_length=epoch, seed=seed)
    print(f"[gpu] als_prod {method}: the oracle's sweeps", flush=True)
    it, fits = _oracle_trajectory(oracle, idx, vals, U0, sigma0, J, method, seed, sweeps, epoch, tol)
    rel = np.abs(np.array(rec["fits"]) - np.array(fits[: len(rec["fits"])])) / np.abs(fits[: len(rec["fits"])])
    print(f"[gpu] als_prod {method}: fits {np.array2string(np.array(rec['fits']), precision=6)}; "
          f"max rel diff vs oracle {rel.max():.2e}", flush=True)
    assert rec["iterations"] == it
    assert len(rec["fits"]) == len(fits)
    assert rel.max() <= 1e-6


def _pack(torch, idx, dims, modes):
    """int64 key of the coordinates of `modes` (first mode most significant)."""
    key = torch.zeros(idx.shape[0], dtype=torch.int64, device=idx.device)
    for k in modes:
        key = key * int(dims[k]) + idx[:, k].long()
    return key


@pytest.mark.timeout(900)
def test_c5_full_size_properties(ftl):
    import torch
    from fast_tensor_leverage_amd.synth import zipf_sparse
    dev = torch.device("cuda", 0)
    dims, nnz, R, J = (2 ** 22, 2 ** 21, 2 ** 17), 200_000_000, 25, 2 ** 16
    N = 3
    print(f"[gpu] C5: generating dims={dims} nnz={nnz} on the device", flush=True)
    idx, vals = zipf_sparse(torch, dims, nnz, R, dev, seed=2024)
    X = ftl.SparseTensor.from_device(idx, vals, dims=list(dims))
    assert X.dims == list(dims) and X.nnz == nnz
    ref_normsq = float((vals * vals).sum())
    assert abs(X.get_normsq() - ref_normsq) <= 1e-12 * ref_normsq

    # lookups: counts of 4096 tuples (half taken from nonzeros) and the ids of
    # a subset whose total stays host-sized, against a stable torch sort
    rng = np.random.default_rng(5)
    for j in range(N):
        other = [k for k in range(N) if k != j]
        print(f"[gpu] C5: lookups on mode {j}", flush=True)
        keys = _pack(torch, idx, dims, other)
        skeys, perm = torch.sort(keys, stable=True)
        del keys
        q = np.stack([rng.integers(0, d, size=4096) for d in dims], axis=1).astype(np.uint64)
        q[:2048] = idx[torch.from_numpy(rng.integers(0, nnz, size=2048)).to(dev)].cpu().numpy().astype(np.uint64)
        qk = _pack(torch, torch.from_numpy(q.astype(np.int64)).to(dev), dims, other)
        lo = torch.searchsorted(skeys, qk, right=False)
        hi = torch.searchsorted(skeys, qk, right=True)
        ref_counts = (hi - lo).cpu().numpy()
        all_counts, _ = X.lookup(q, j, counts_only=True)
        np.testing.assert_array_equal(all_counts, ref_counts.astype(np.uint64))
        print(f"[gpu] C5: mode {j}: {int(ref_counts.sum())} matches of 4096 tuples, max {int(ref_counts.max())}",
              flush=True)
        small = np.nonzero(ref_counts <= 2000)[0]
        sel = np.concatenate([small[:1500], np.nonzero(ref_counts > 2000)[0][:2]])
        counts, ids = X.lookup(np.ascontiguousarray(q[sel]), j)
        np.testing.assert_array_equal(counts, ref_counts[sel].astype(np.uint64))
        lo_h, hi_h = lo.cpu().numpy()[sel], hi.cpu().numpy()[sel]
        ref_ids = np.concatenate([perm[a:b].cpu().numpy() for a, b in zip(lo_h, hi_h)]) if sel.size else []
        np.testing.assert_array_equal(ids, np.asarray(ref_ids, dtype=np.uint64))
        del skeys, perm, qk, lo, hi

    # one sampled update of mode 0 against the dense sampled least squares,
    # after a sweep that moves the factors onto the tensor's support (from
    # the random start almost no sampled tuple matches a nonzero)
    print("[gpu] C5: one execute_ds_als_update(0) against the sampled least squares", flush=True)
    lhs = ftl.PyLowRank(dims, R, seed=5)
    lhs.ten.renormalize_columns(-1)
    als = ftl.ALS(lhs.ten, X, host_sync=True)
    als.initialize_ds_als(J, "efficient", seed=17)
    for j in range(N):
        als.execute_ds_als_update(j, True, True)
    U_init = [u.copy() for u in lhs.U]
    draws = np.random.default_rng(99).random(2 * N * J)
    smp = ftl.Sampler(U_init, J, R, "efficient")
    samples = np.zeros((N, J), dtype=np.uint64)
    h = np.zeros((J, R))
    w = np.zeros(J)
    smp.KRPDrawSamples_materialize(0, samples, h, w, draws)
    smp.close()
    als.execute_ds_als_update(0, False, False, random_draws=draws)
    Uj = lhs.U[0]
    touched = np.nonzero(np.any(Uj != 0, axis=1))[0]
    print(f"[gpu] C5: {touched.size} touched rows of mode 0, found {als.last_stats()['found']}", flush=True)
    assert touched.size > 100_000
    rows = rng.choice(touched, size=min(256, touched.size), replace=False)
    fkeys = _pack(torch, idx, dims, [0, 1, 2])
    sfk, sperm = torch.sort(fkeys)
    del fkeys
    t1 = torch.from_numpy(samples[1].astype(np.int64)).to(dev)
    t2 = torch.from_numpy(samples[2].astype(np.int64)).to(dev)
    hd = torch.from_numpy(h).to(dev)
    wd = torch.from_numpy(w).to(dev)
    M = torch.zeros((rows.size, R), dtype=torch.float64, device=dev)
    for a, r in enumerate(rows):
        qk = (int(r) * dims[1] + t1) * dims[2] + t2
        pos = torch.searchsorted(sfk, qk).clamp_(max=nnz - 1)
        hit = sfk[pos] == qk
        xv = torch.where(hit, vals[sperm[pos]], torch.zeros_like(wd))
        M[a] = (xv * wd) @ hd
    Gs = (hd * wd[:, None]).T @ hd
    ref_rows = (M @ torch.linalg.pinv(Gs, hermitian=True)).cpu().numpy()
    scale = np.abs(ref_rows).max()
    np.testing.assert_allclose(Uj[rows], ref_rows, rtol=1e-9, atol=1e-10 * scale)
    del sfk, sperm, t1, t2, hd, wd, M

    # device fit against a chunked FP64 residual
    print("[gpu] C5: device residual against a chunked FP64 residual", flush=True)
    dev_res = als.residual_normsq()
    sigma = torch.from_numpy(lhs.ten.sigma.copy()).to(dev)
    Ud = [torch.from_numpy(u).to(dev) for u in lhs.U]
    inner = torch.zeros((), dtype=torch.float64, device=dev)
    for s0 in range(0, nnz, 4_000_000):
        s1 = min(nnz, s0 + 4_000_000)
        p = Ud[0][idx[s0:s1, 0].long()] * sigma
        for k in range(1, N):
            p *= Ud[k][idx[s0:s1, k].long()]
        inner += (vals[s0:s1] * p.sum(dim=1)).sum()
    Gh = torch.ones((R, R), dtype=torch.float64, device=dev)
    for u in Ud:
        Gh *= u.T @ u
    mnorm = float(sigma @ Gh @ sigma)
    ref_res = ref_normsq - 2 * float(inner) + mnorm
    print(f"[gpu] C5: residual device {dev_res:.12e} reference {ref_res:.12e}", flush=True)
    assert abs(dev_res - ref_res) <= 1e-10 * abs(ref_res)


def test_mttkrp_refuses_2p32_matches(ftl):
    """65,537 samples each matching a fiber of 65,537 nonzeros: 2^32 + 2^17 + 1
    matches overflow the 32-bit offsets and are refused before any write."""
    n = 65_537
    idx = np.zeros((n, 3), dtype=np.uint32)
    idx[:, 0] = np.arange(n)
    X = ftl.SparseTensor(idx, np.ones(n))
    samples = np.zeros((n, 3), dtype=np.uint64)
    lhs = np.ones((n, 2))
    out = np.full((n, 2), 7.0)
    with pytest.raises(NotImplementedError, match="2\\^32"):
        X.execute_downsampled_mttkrp_py(samples, lhs, 0, out)
    assert np.all(out == 7.0)
    # a smaller total on the same tensor still works
    X.execute_downsampled_mttkrp_py(samples[:3], lhs[:3], 0, out)
    assert X.last_found == 3 * n and np.all(out == 3.0)
