"""Sparse tensor file readers (the reference's HDF5 layout as .npz, FROSTT
.tns) -- host-side parsing, no GPU."""
import numpy as np
import pytest

from fast_tensor_leverage_amd.tensors import read_npz, read_tns


def _random_coo(rng, dims, nnz):
    idx = np.stack([rng.integers(0, d, size=nnz) for d in dims], axis=1).astype(np.uint32)
    return idx, rng.standard_normal(nnz)


def test_npz_with_reference_dataset_names(tmp_path):
    rng = np.random.default_rng(1)
    dims = (7, 5, 9)
    idx, vals = _random_coo(rng, dims, 60)
    # the reference stores 1-based coordinates with MIN_MODE_SET = 1
    path = tmp_path / "t.npz"
    np.savez(path, MODE_0=idx[:, 0] + 1, MODE_1=idx[:, 1] + 1, MODE_2=idx[:, 2] + 1, VALUES=vals,
             MIN_MODE_SET=np.ones(3, dtype=np.int64), MAX_MODE_SET=np.array(dims, dtype=np.int64))
    i2, v2, d2 = read_npz(path)
    np.testing.assert_array_equal(i2, idx)
    np.testing.assert_array_equal(v2, vals)
    assert d2 == list(dims)


def test_tns_one_based_with_comments(tmp_path):
    rng = np.random.default_rng(2)
    idx, vals = _random_coo(rng, (4, 6, 3, 5), 40)
    path = tmp_path / "t.tns"
    with open(path, "w") as f:
        f.write("# a FROSTT tensor\n")
        for row, v in zip(idx, vals):
            f.write(" ".join(str(int(c) + 1) for c in row) + "\t" + repr(float(v)) + "\n")
    i2, v2, d2 = read_tns(path)
    np.testing.assert_array_equal(i2, idx)
    np.testing.assert_array_equal(v2, vals)
    assert d2 == [int(idx[:, k].max()) + 1 for k in range(4)]


def test_tns_zero_based_and_errors(tmp_path):
    path = tmp_path / "z.tns"
    path.write_text("0 0 1.5\n2 1 -2\n")
    i2, v2, d2 = read_tns(path, one_based=False)
    np.testing.assert_array_equal(i2, [[0, 0], [2, 1]])
    np.testing.assert_array_equal(v2, [1.5, -2.0])
    assert d2 == [3, 2]
    with pytest.raises(ValueError):
        read_tns(path, one_based=True)  # coordinate 0 in a 1-based file


def test_als_prod_stop_rule():
    """The stop rule of als.py:99-101: more than four fits, and the best of
    the newest three below the best before them plus the tolerance."""
    from fast_tensor_leverage_amd.als_driver import stalled
    assert not stalled([0.1, 0.2, 0.3, 0.4], 0.0)  # four fits: never
    assert not stalled([0.1, 0.2, 0.3, 0.4, 0.5], 0.0)  # improving
    assert stalled([0.1, 0.5, 0.4, 0.45, 0.49], 0.0)  # no new best
    assert stalled([0.1, 0.2, 0.40, 0.401, 0.402, 0.403], 0.01)  # gain below the tolerance
    assert not stalled([0.1, 0.2, 0.40, 0.401, 0.402, 0.411], 0.01)
    assert stalled([0.5, 0.4, 0.3, 0.2, 0.1, 0.6], 0.2)  # only the newest three count
