"""Pins the oracle's tree shape, leaf map and tree build (partition_tree.hpp:83-170)
against closed forms and brute force."""
import numpy as np
import pytest


def _inorder_leaves(node_count):
    """Left-to-right leaves of the heap-ordered tree, by an explicit DFS."""
    out, stack, c = [], [], 0
    while stack or c < node_count:
        while c < node_count:
            stack.append(c)
            c = 2 * c + 1
        c = stack.pop()
        if 2 * c + 1 >= node_count:
            out.append(c)
        c = 2 * c + 2
    return out


def _leaf_map(sh, c):
    if c >= sh["nodes_upto_lfill"]:
        return c - sh["nodes_upto_lfill"]
    return c - sh["complete_level_offset"]


@pytest.mark.parametrize("leaves", list(range(1, 300)) + [511, 512, 513, 1000, 1023, 1024, 1025,
                                                          2047, 2048, 2049, 4095, 4096])
def test_leaf_order_is_inorder(oracle, leaves):
    sh = oracle.tree_shape(leaves, 1)
    assert sh["leaf_count"] == leaves
    assert sh["node_count"] == 2 * leaves - 1
    assert sh["lfill_count"] == 1 << (leaves.bit_length() - 1)
    order = _inorder_leaves(sh["node_count"])
    assert len(order) == leaves
    # the leaves are exactly the last L nodes of the heap
    assert sorted(order) == list(range(leaves - 1, 2 * leaves - 1))
    assert [_leaf_map(sh, c) for c in order] == list(range(leaves))
    if leaves in (1, 7, 100, 1025):
        assert [oracle.leaf_idx(leaves, 1, c) for c in order] == list(range(leaves))
        assert oracle.leaf_idx(leaves, 1, 0) == (0 if leaves == 1 else -1)


@pytest.mark.parametrize("n,F", [(1000, 8), (37, 4), (64, 8), (8, 8), (5, 8), (26, 1), (33, 32)])
def test_tree_shape_ragged(oracle, n, F):
    sh = oracle.tree_shape(n, F)
    assert sh["leaf_count"] == -(-n // F)
    # every row is covered exactly once by the leaf blocks
    covered = np.zeros(n, dtype=int)
    for c in range(sh["leaf_count"] - 1, sh["node_count"]):
        idx = oracle.leaf_idx(n, F, c)
        covered[idx * F: min((idx + 1) * F, n)] += 1
    assert np.all(covered == 1)


def _subtree_rows(n, F, node_count, c):
    leaves = []
    stack = [c]
    while stack:
        v = stack.pop()
        if 2 * v + 1 >= node_count:
            leaves.append(v)
        else:
            stack += [2 * v + 1, 2 * v + 2]
    return leaves


@pytest.mark.parametrize("n,R,F", [(37, 4, 4), (1000, 8, 8), (64, 8, 8), (25, 25, 1), (200, 5, 5)])
def test_build_tree_closed_form(oracle, n, R, F):
    rng = np.random.default_rng(n * 31 + R)
    U = rng.standard_normal((n, R))
    G = oracle.build_tree(U, F)
    sh = oracle.tree_shape(n, F)
    iu = np.triu_indices(R)
    il = np.tril_indices(R, -1)
    for c in range(sh["node_count"]):
        rows = []
        for leaf in _subtree_rows(n, F, sh["node_count"], c):
            idx = oracle.leaf_idx(n, F, leaf)
            rows += list(range(idx * F, min((idx + 1) * F, n)))
        rows = sorted(rows)
        # a node covers a contiguous row range
        assert rows == list(range(rows[0], rows[-1] + 1))
        ref = U[rows].T @ U[rows]
        np.testing.assert_allclose(G[c][iu], ref[iu], rtol=1e-12, atol=1e-12)
        assert np.all(G[c][il] == 0.0)
    np.testing.assert_allclose(G[0][iu], (U.T @ U)[iu], rtol=1e-12, atol=1e-12)


def test_planted_mistake_is_caught(oracle):
    """A tree whose leaf Gram skips the last row of each block must fail the
    closed-form check used above (the check has the power to see it)."""
    rng = np.random.default_rng(3)
    U = rng.standard_normal((64, 8))
    G = oracle.build_tree(U, 8)
    bad = U.copy()
    bad[7::8] = 0.0
    ref = bad.T @ bad
    iu = np.triu_indices(8)
    assert not np.allclose(G[0][iu], ref[iu], rtol=1e-6)


@pytest.mark.parametrize("R", [2, 3, 8, 25, 32, 64])
def test_jacobi_eigh_matches_lapack(oracle, R):
    rng = np.random.default_rng(R)
    X = rng.standard_normal((3 * R, R))
    A = X.T @ X
    lam, V = oracle.jacobi_eigh(np.triu(A))
    ref_lam, ref_V = np.linalg.eigh(A)
    np.testing.assert_allclose(lam, ref_lam, rtol=1e-11, atol=1e-12 * ref_lam.max())
    assert np.all(np.diff(lam) >= 0)
    np.testing.assert_allclose(V.T @ V, np.eye(R), atol=1e-12)
    np.testing.assert_allclose((V * lam) @ V.T, A, atol=1e-11 * np.abs(A).max())
    dots = np.abs(np.sum(V * ref_V, axis=0))
    np.testing.assert_allclose(dots, 1.0, atol=1e-8)


def test_jacobi_rank_deficient_and_graded(oracle):
    rng = np.random.default_rng(11)
    R = 16
    Q, _ = np.linalg.qr(rng.standard_normal((R, R)))
    s = np.geomspace(1.0, 1e-12, R)
    s[-3:] = 0.0
    A = (Q * s) @ Q.T
    lam, V = oracle.jacobi_eigh(A)
    np.testing.assert_allclose(lam[::-1][:R - 3], s[:R - 3], rtol=1e-6, atol=1e-15)
    assert np.all(np.abs(lam[:3]) < 1e-14)
    np.testing.assert_allclose((V * lam) @ V.T, A, atol=1e-14)
