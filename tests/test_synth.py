"""The synthetic workload generators (host torch here; the benchmarks and the
GPU tests run them on the device): distinct in-range coordinates of any mode
count, values of the CP model, and ill-conditioned factors with the stated
singular values."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from fast_tensor_leverage_amd.synth import ill_conditioned_factors, zipf_coords, zipf_sparse  # noqa: E402


@pytest.mark.parametrize("dims", [(50, 40, 30), (9, 5, 13, 7), (183, 24, 1140, 1717), (1, 6, 4)])
def test_zipf_sparse_coordinates_are_distinct_and_in_range(dims):
    nnz = min(500, int(np.prod(dims)) // 3)
    idx, vals = zipf_sparse(torch, list(dims), nnz, 3, torch.device("cpu"), seed=5, noise=0.0)
    assert idx.shape == (nnz, len(dims)) and idx.dtype == torch.int32
    a = idx.numpy().astype(np.int64)
    assert np.all(a >= 0) and np.all(a < np.array(dims))
    assert len(np.unique(a, axis=0)) == nnz
    assert vals.shape == (nnz,) and torch.isfinite(vals).all()


def test_zipf_sparse_values_follow_the_cp_model():
    """Without noise every value is sum_r prod_k A_k[i_k, r] for the factors
    drawn after the coordinates from the same generator."""
    dims, nnz, R = [20, 15, 10, 8], 300, 4
    idx, vals = zipf_sparse(torch, dims, nnz, R, torch.device("cpu"), seed=9, noise=0.0)
    # replay the generator: the coordinate draws, the permutation, then the factors
    gen = torch.Generator()
    gen.manual_seed(9)
    bits = [(d - 1).bit_length() for d in dims]
    shift = [sum(bits[k + 1:]) for k in range(len(dims))]
    keys = torch.empty(0, dtype=torch.int64)
    while keys.numel() < nnz:
        m = max(int(1.25 * (nnz - keys.numel())), 1_000_000)
        k = zipf_coords(torch, dims[0], m, 1.1, gen, "cpu") << shift[0]
        for q in range(1, len(dims)):
            k |= zipf_coords(torch, dims[q], m, 1.1, gen, "cpu") << shift[q]
        keys = torch.unique(torch.cat([keys, k]))
    torch.randperm(keys.numel(), generator=gen)
    A = [torch.randn((d, R), generator=gen, dtype=torch.float64) for d in dims]
    want = torch.ones((nnz, R), dtype=torch.float64)
    for q in range(len(dims)):
        want *= A[q][idx[:, q].long()]
    np.testing.assert_allclose(vals.numpy(), want.sum(dim=1).numpy(), rtol=1e-12, atol=1e-14)


def test_zipf_sparse_refuses_keys_over_63_bits():
    with pytest.raises(ValueError):
        zipf_sparse(torch, [4_821_207, 1_774_269, 1_805_187], 10, 2, torch.device("cpu"))


def test_zipf_coords_are_skewed_and_in_range():
    gen = torch.Generator()
    gen.manual_seed(1)
    c = zipf_coords(torch, 1000, 200_000, 1.1, gen, "cpu").numpy()
    assert c.min() >= 0 and c.max() < 1000
    counts = np.bincount(c, minlength=1000)
    assert counts[0] > 10 * counts[100] > 0  # heavy head


def test_ill_conditioned_factors_have_the_stated_singular_values():
    U, _ = ill_conditioned_factors(torch, 2, 300, 6, torch.device("cpu"), seed=3, smin=1e-6)
    for M in U:
        s = np.linalg.svd(M.numpy(), compute_uv=False)
        np.testing.assert_allclose(s, np.geomspace(1.0, 1e-6, 6), rtol=1e-8)
