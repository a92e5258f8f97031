"""ftl_sparse_residual_normsq_noalloc (the allocation-free residual) against
ftl_sparse_residual_normsq on the same tensor, factors and sigma: equal
bitwise, on two calls in a row (the second on a smaller tensor and rank, so
a stale intermediate from the first call would show), and a shape whose
intermediates exceed the library's scratch array is refused before any work
is queued."""
import ctypes

import numpy as np
import pytest

from tests.helpers import random_sparse

pytestmark = pytest.mark.gpu


def _noalloc(L, X, U_dev, sigma_dev, out_dev, stream=None, I=None):
    import torch
    N = len(U_dev)
    dims = np.ascontiguousarray(I if I is not None else [u.shape[0] for u in U_dev], dtype=np.uint64)
    ptrs = (ctypes.c_void_p * N)(*[u.data_ptr() for u in U_dev])
    return L.ftl_sparse_residual_normsq_noalloc(X._h, int(U_dev[0].shape[1]), dims.ctypes.data, sigma_dev.data_ptr(),
                                                ptrs, out_dev.data_ptr(),
                                                None if stream is None else ctypes.c_void_p(stream.cuda_stream))


@pytest.mark.parametrize("dims,nnz,R,dims2,nnz2,R2", [
    ((3000, 2000, 1500), 200_000, 25, (300, 200, 100), 5_000, 7),
    ((70000, 900, 4000), 300_000, 64, (50, 40, 30, 20), 2_000, 3),
])
def test_noalloc_residual_matches_allocating_entry_point(dims, nnz, R, dims2, nnz2, R2):
    import torch
    import fast_tensor_leverage_amd as ftl
    from fast_tensor_leverage_amd import _lib
    L = _lib.load()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(R)
    cases = []
    for d, n, r in ((dims, nnz, R), (dims2, nnz2, R2)):
        idx, vals = random_sparse(d, n, rng)
        X = ftl.SparseTensor(idx, vals, dims=list(d))
        U = [rng.standard_normal((k, r)) for k in d]
        sigma = rng.random(r) + 0.5
        ref = X.compute_residual_normsq(sigma, U)
        cases.append((X, U, sigma, ref))
    stream = torch.cuda.Stream()
    for X, U, sigma, ref in cases:  # two calls in a row, the second smaller
        print(f"[gpu] noalloc residual dims={X.dims} R={U[0].shape[1]}", flush=True)
        U_dev = [torch.from_numpy(u).to(dev) for u in U]
        sig = torch.from_numpy(sigma).to(dev)
        out = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
        stream.wait_stream(torch.cuda.current_stream())
        _lib.check(_noalloc(L, X, U_dev, sig, out, stream), "residual_normsq_noalloc")
        stream.synchronize()
        got = float(out.cpu().item())
        assert got == ref, (got, ref)


def test_noalloc_residual_refuses_oversized_shapes():
    import torch
    import fast_tensor_leverage_amd as ftl
    from fast_tensor_leverage_amd import _lib
    L = _lib.load()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(1)
    idx, vals = random_sparse((40, 30, 20), 500, rng)
    X = ftl.SparseTensor(idx, vals, dims=[40, 30, 20])
    R, rows = 128, 1_000_000  # 512 Gram blocks of 8256 entries: 4.2e6 doubles
    U_dev = [torch.zeros((rows, R), dtype=torch.float64, device=dev) for _ in range(3)]
    sig = torch.ones(R, dtype=torch.float64, device=dev)
    out = torch.full((1,), 7.0, dtype=torch.float64, device=dev)
    rc = _noalloc(L, X, U_dev, sig, out)
    assert rc == _lib.FTL_EUNSUPPORTED
    assert "scratch" in _lib.last_error()
    torch.cuda.synchronize()
    assert float(out.cpu().item()) == 7.0
