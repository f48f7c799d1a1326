"""Host-side graph construction (no GPU): the generators' invariants and
the device form of the largest-component step, run here on torch's CPU
backend against the scipy one."""
import numpy as np
import pytest
import scipy.sparse as sp

from ipdps_2021_graph_coarsening_amd import graphs


def _check_csr(rp, col, w):
    n = len(rp) - 1
    assert rp[0] == 0 and np.all(np.diff(rp.astype(np.int64)) >= 0) and rp[-1] == len(col)
    src = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(src != col)                                  # no self-loops
    A = sp.csr_matrix((w.astype(np.int64), col, rp), shape=(n, n))
    assert (A != A.T).nnz == 0                                 # symmetric, weights too
    for r in range(0, n, max(1, n // 50)):                     # rows sorted, no duplicates
        c = col[rp[r]:rp[r + 1]]
        assert np.all(np.diff(c.astype(np.int64)) > 0)


def test_from_edges_merges_and_sums():
    u = np.array([0, 1, 0, 2, 2, 3, 1])
    v = np.array([1, 0, 1, 2, 3, 2, 3])
    w = np.array([2, 3, 4, 9, 1, 5, 7])
    rp, col, ww = graphs.from_edges(5, u, v, w, lcc=False)
    A = sp.csr_matrix((ww, col, rp), shape=(5, 5)).toarray()
    exp = np.zeros((5, 5), np.int64)
    for a, b, c in zip(u, v, w):
        if a != b:
            exp[a, b] += c
            exp[b, a] += c
    assert np.array_equal(A, exp)
    rp2, col2, ww2 = graphs.from_edges(5, u, v, None, lcc=False)
    A2 = sp.csr_matrix((ww2, col2, rp2), shape=(5, 5)).toarray()
    exp1 = np.zeros((5, 5), np.int64)
    for a, b in zip(u, v):
        if a != b:
            exp1[a, b] += 1
            exp1[b, a] += 1
    assert np.array_equal(A2, exp1)
    assert A2[0, 1] == 3 and A2[2, 3] == 2  # unit weights: the number of copies


@pytest.mark.parametrize("gen", ["rmat", "rgg", "grid3d"])
def test_generators_are_valid_csr(gen):
    if gen == "rmat":
        rp, col, w = graphs.rmat(12, 16, seed=2)
    elif gen == "rgg":
        rp, col, w = graphs.rgg(5000, seed=4)
    else:
        rp, col, w = graphs.grid3d(10)
    _check_csr(rp, col, w)
    from scipy.sparse.csgraph import connected_components
    n = len(rp) - 1
    nc, _ = connected_components(sp.csr_matrix((np.ones(len(col)), col, rp), shape=(n, n)), directed=False)
    assert nc == 1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_largest_component_device_form_equals_host(seed):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(seed)
    n = 3000
    # several components of different sizes, a tie between two of them,
    # isolated vertices, and labels scattered so components interleave
    blocks = [700, 700, 400, 3, 1]
    perm = rng.permutation(n)
    u, v, start = [], [], 0
    for b in blocks:
        ids = perm[start:start + b]
        k = 3 * b
        u.append(ids[rng.integers(0, b, k)])
        v.append(ids[rng.integers(0, b, k)])
        u.append(ids[:-1])
        v.append(ids[1:])  # a path through the block keeps it connected
        start += b
    u, v = np.concatenate(u), np.concatenate(v)
    keep = u != v
    u, v = u[keep], v[keep]
    key = np.unique(np.concatenate([u * n + v, v * n + u]))
    rows, cols = key // n, key % n
    w = np.ones(len(rows), np.int64)
    hr, hc, hw, hn = graphs._largest_component(n, rows, cols, w)
    dr, dc, dw, dn = graphs._largest_component_device(n, torch.from_numpy(rows), torch.from_numpy(cols),
                                                      torch.from_numpy(w))
    assert dn == hn == 700
    assert np.array_equal(dr.numpy(), hr) and np.array_equal(dc.numpy(), hc) and np.array_equal(dw.numpy(), hw)
