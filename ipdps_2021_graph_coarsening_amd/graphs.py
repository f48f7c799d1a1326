"""CSR graphs on the host: the reference's binary CSR format and the
synthetic inputs of the benchmark configurations.

Binary CSR (written by the reference's mtx2csr.cpp:276-288, read by
sgp_load_graph, sgpar.h:2262-2289): six int64 header words
``n, m(stored), undirected=1, graph_type, one_indexed=1, verification=0``,
then ``uint32 offsets[n+1]`` and ``uint32 adj[m]``.  The reference reader
ignores header words 3-6 and stops after ``adj``; a weighted graph is stored
here with ``graph_type = 1`` and a trailing ``uint32 w[m]`` section, which
the reference reader therefore loads as the same topology.

Every generator returns ``(rowptr, col, w)`` as uint32 arrays, symmetric,
without self-loops, duplicates merged with summed weights, rows sorted and
(unless disabled) restricted to the largest connected component with the
original vertex order kept, as mtx2csr does (mtx2csr.cpp:133-263).
"""
from __future__ import annotations

import os

import numpy as np

HEADER_WORDS = 6


# ---------------------------------------------------------------- file format
def write_csr(path: str, rowptr, col, w=None) -> None:
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    n = len(rowptr) - 1
    m = int(rowptr[-1])
    hdr = np.array([n, m, 1, 0 if w is None else 1, 1, 0], dtype=np.int64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(rowptr.tobytes())
        f.write(col[:m].tobytes())
        if w is not None:
            f.write(np.ascontiguousarray(w, dtype=np.uint32)[:m].tobytes())


def read_csr(path: str):
    """(rowptr, col, w or None); raises ValueError on a short or malformed file."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < 8 * HEADER_WORDS:
        raise ValueError(f"{path}: truncated header")
    hdr = np.frombuffer(raw[:48], dtype=np.int64)
    n, m, gtype = int(hdr[0]), int(hdr[1]), int(hdr[3])
    if n < 0 or m < 0:
        raise ValueError(f"{path}: negative sizes in header")
    need = 48 + 4 * (n + 1) + 4 * m
    if len(raw) < need:
        raise ValueError(f"{path}: file shorter than its header says")
    rowptr = np.frombuffer(raw, dtype=np.uint32, count=n + 1, offset=48).copy()
    col = np.frombuffer(raw, dtype=np.uint32, count=m, offset=48 + 4 * (n + 1)).copy()
    w = None
    if gtype == 1 and len(raw) >= need + 4 * m:
        w = np.frombuffer(raw, dtype=np.uint32, count=m, offset=need).copy()
    if rowptr[0] != 0 or int(rowptr[-1]) != m or np.any(np.diff(rowptr.astype(np.int64)) < 0):
        raise ValueError(f"{path}: malformed offsets")
    if m and int(col.max()) >= n:
        raise ValueError(f"{path}: column index out of range")
    return rowptr, col, w


def read_config(path: str) -> dict:
    """The six whitespace-separated config scalars (sgpar.h:2754-2787);
    trailing tokens are ignored."""
    with open(path) as f:
        tok = f.read().split()
    if len(tok) < 6:
        raise ValueError(f"{path}: config needs 6 values")
    return dict(coarsening_alg=int(tok[0]), refine_alg=int(tok[1]), local_search_alg=int(tok[2]),
                num_partitions=int(tok[3]), num_iter=int(tok[4]), tol=float(tok[5]))


def write_partition(path: str, part) -> None:
    """One integer per line, the format sgp_load_partition reads (sgpar.h:2321-2341)."""
    np.savetxt(path, np.asarray(part, dtype=np.int64), fmt="%d")


def read_partition(path: str, n: int) -> np.ndarray:
    p = np.loadtxt(path, dtype=np.int64, ndmin=1)
    if len(p) < n:
        raise ValueError(f"{path}: {len(p)} entries, expected {n}")
    return p[:n].astype(np.uint32)


def partition_edit_distance(p1, p2) -> int:
    """compute_partition_edit_distance, sgpar.h:2343-2361: the smaller of the
    disagreement counts under both labellings, halved."""
    p1 = np.asarray(p1)
    p2 = np.asarray(p2)
    d = int(np.count_nonzero(p1 != p2))
    d2 = len(p1) - d
    return (d if d < d2 else d2) // 2


def use_partition(rowptr, col, part):
    """Induced subgraphs of parts 0 and 1 with vertices relabelled in order,
    topology only: sgp_use_partition, sgpar.h:2666-2752."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    part = np.asarray(part)
    n = len(rowptr) - 1
    src = np.repeat(np.arange(n), np.diff(rowptr))
    keep = part[src] == part[col]
    out = []
    for p in (0, 1):
        sel = part == p
        newid = np.cumsum(sel) - 1
        e = keep & (part[src] == p)
        cnt = np.bincount(newid[src[e]], minlength=int(sel.sum()))
        rp = np.zeros(int(sel.sum()) + 1, dtype=np.uint32)
        np.cumsum(cnt, out=rp[1:])
        out.append((rp, newid[col[e]].astype(np.uint32)))
    return out[0], out[1]


# ---------------------------------------------------------------- construction
def from_edges(n: int, u, v, w=None, lcc: bool = True):
    """Symmetrise, drop self-loops, merge duplicates (summing weights), sort
    rows, keep the largest connected component (order-preserving relabel)."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    keep = u != v
    u, v = u[keep], v[keep]
    key = np.concatenate([u * n + v, v * n + u])
    del u, v
    # integer sums do not depend on the order of equal keys, so no stable
    # sort is needed; unit weights are the run lengths
    if w is None:
        key.sort()
        ww = None
    else:
        w = np.asarray(w, dtype=np.int64)[keep]
        order = np.argsort(key)
        key = key[order]
        ww = np.concatenate([w, w])[order]
        del order
    heads = np.empty(len(key), dtype=bool)
    if len(key):
        heads[0] = True
        np.not_equal(key[1:], key[:-1], out=heads[1:])
    start = np.flatnonzero(heads)
    if ww is None:
        wsum = np.diff(np.append(start, len(key))).astype(np.int64)
    else:
        wsum = np.add.reduceat(ww, start) if len(start) else np.zeros(0, np.int64)
    ukey = key[start]
    del key, heads
    rows = ukey // n
    cols = ukey % n
    if np.any(wsum > np.iinfo(np.uint32).max):
        raise ValueError("merged weight exceeds uint32")
    if lcc and n > 0:
        rows, cols, wsum, n = _largest_component(n, rows, cols, wsum)
    cnt = np.bincount(rows, minlength=n)
    rowptr = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(cnt, out=rowptr[1:])
    return rowptr, cols.astype(np.uint32), wsum.astype(np.uint32)


def _largest_component(n, rows, cols, wsum):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    A = sp.csr_matrix((np.ones(len(rows), dtype=np.int8), (rows, cols)), shape=(n, n))
    ncomp, lab = connected_components(A, directed=False)
    if ncomp <= 1:
        return rows, cols, wsum, n
    sizes = np.bincount(lab)
    big = int(np.argmax(sizes))
    inside = lab == big
    newid = np.cumsum(inside) - 1
    e = inside[rows]
    return newid[rows[e]], newid[cols[e]], wsum[e], int(inside.sum())


# ---------------------------------------------------------------- generators
def grid2d(k: int, weights: str | None = None, seed: int = 0):
    """k x k 5-point grid, row-major ids; unit weights unless `weights` is
    'uniform16' (U{1..16})."""
    idx = np.arange(k * k, dtype=np.int64).reshape(k, k)
    u = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    v = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    w = None
    if weights == "uniform16":
        w = np.random.default_rng(seed).integers(1, 17, size=len(u))
    return from_edges(k * k, u, v, w, lcc=False)


def grid3d(k: int, weights: str | None = "uniform16", seed: int = 2):
    """k^3 7-point grid, ids x-fastest; weights U{1..16} by default."""
    idx = np.arange(k ** 3, dtype=np.int64).reshape(k, k, k)
    u = np.concatenate([idx[:, :, :-1].ravel(), idx[:, :-1, :].ravel(), idx[:-1, :, :].ravel()])
    v = np.concatenate([idx[:, :, 1:].ravel(), idx[:, 1:, :].ravel(), idx[1:, :, :].ravel()])
    w = None
    if weights == "uniform16":
        w = np.random.default_rng(seed).integers(1, 17, size=len(u))
    return from_edges(k ** 3, u, v, w, lcc=False)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1, lcc: bool = True):
    """R-MAT (Chakrabarti et al.) with 2^scale vertices and edge_factor*2^scale
    edges of weight 1; duplicates merge with summed weights."""
    if scale > 32:
        raise ValueError("scale above 32 does not fit 32-bit vertex ids")
    rng = np.random.default_rng(seed)
    m = edge_factor << scale
    u = np.zeros(m, dtype=np.uint32)
    v = np.zeros(m, dtype=np.uint32)
    ab, abc = a + b, a + b + c
    chunk = 1 << 22
    for s in range(0, m, chunk):
        e = min(m, s + chunk)
        uu = u[s:e]
        vv = v[s:e]
        for bit in range(scale):
            r = rng.random(e - s)
            uu |= np.left_shift((r >= ab).view(np.uint8), bit, dtype=np.uint32)
            vb = (r >= a) & (r < ab)
            vb |= r >= abc
            vv |= np.left_shift(vb.view(np.uint8), bit, dtype=np.uint32)
    return from_edges(1 << scale, u, v, None, lcc=lcc)


def rmat_device(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                seed: int = 4, lcc: bool = True, cache_dir: str | None = None):
    """R-MAT as `rmat`, with the edges drawn, symmetrised and merged and the
    largest component found on the GPU through torch (scale 24: seconds
    instead of minutes).  Another random stream than `rmat`, so another
    instance of the same distribution.

    The torch work runs in a child process.  torch ships its own HIP and HSA
    runtimes while libsgpar_mi.so links the system ones; with both started
    in one process the second finds no device.  The child writes the three
    arrays as .npy files, which are read back here: into `cache_dir` when
    given (a later call with the same arguments reads them without drawing
    again), else into a temporary directory removed afterwards."""
    import shutil
    import subprocess
    import sys
    import tempfile

    tag = f"rmat{scale}_ef{edge_factor}_{a}_{b}_{c}_seed{seed}_lcc{int(bool(lcc))}"
    d = cache_dir if cache_dir else tempfile.mkdtemp(prefix="rmat_device_")
    files = [os.path.join(d, f"{tag}_{x}.npy") for x in ("rowptr", "col", "w")]
    try:
        if not all(os.path.exists(f) for f in files):
            os.makedirs(d, exist_ok=True)
            # the module file runs as a script: the package (and with it
            # libsgpar_mi.so's runtime) is not loaded in the child
            subprocess.run([sys.executable, os.path.abspath(__file__), "rmat-device", str(scale), str(edge_factor),
                            repr(a), repr(b), repr(c), str(seed), str(int(bool(lcc))), d, tag], check=True)
        return tuple(np.load(f) for f in files)
    finally:
        if not cache_dir:
            shutil.rmtree(d, ignore_errors=True)


def _rmat_device_inproc(scale: int, edge_factor: int, a: float, b: float, c: float, seed: int, lcc: bool):
    """The torch body of `rmat_device`; runs in the child process."""
    import torch

    n, m = 1 << scale, edge_factor << scale
    ab, abc = a + b, a + b + c
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    u = torch.zeros(m, dtype=torch.int64, device=dev)
    v = torch.zeros(m, dtype=torch.int64, device=dev)
    for bit in range(scale):
        r = torch.rand(m, generator=gen, device=dev)
        u |= (r >= ab).long() << bit
        v |= (((r >= a) & (r < ab)) | (r >= abc)).long() << bit
        del r
    keep = u != v
    u, v = u[keep], v[keep]
    key = torch.cat([u * n + v, v * n + u])
    del u, v, keep
    key, cnt = torch.unique(key, sorted=True, return_counts=True)
    if int(cnt.max()) > np.iinfo(np.uint32).max:
        raise ValueError("merged weight exceeds uint32")
    rows = key // n
    cols = key % n
    del key
    if lcc:
        rows, cols, cnt, n = _largest_component_device(n, rows, cols, cnt)
    rowptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    rowptr[1:] = torch.cumsum(torch.bincount(rows, minlength=n), 0)
    del rows
    out = (rowptr.to(torch.int32).cpu().numpy().view(np.uint32),
           cols.to(torch.int32).cpu().numpy().view(np.uint32),
           cnt.to(torch.int32).cpu().numpy().view(np.uint32))
    del rowptr, cols, cnt
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def _largest_component_device(n, rows, cols, w):
    """Largest connected component of a symmetric edge list on the GPU
    (torch tensors, rows sorted): every vertex takes the smallest label among
    its neighbours and the label of its label, until nothing changes (the
    label is then the smallest vertex id of the component); the component
    with the most vertices wins, ties to the smallest id, which is the one
    `_largest_component` picks (scipy numbers components in order of their
    smallest vertex).  Vertices keep their order when relabelled."""
    import torch

    lab = torch.arange(n, dtype=torch.int64, device=rows.device)
    while True:
        new = lab.scatter_reduce(0, rows, lab[cols], reduce="amin", include_self=True)
        new = new[new]
        if torch.equal(new, lab):
            break
        lab = new
    sizes = torch.bincount(lab, minlength=n)
    big = int(torch.argmax(sizes))
    inside = lab == big
    del lab, sizes
    if bool(inside.all()):
        return rows, cols, w, n
    newid = torch.cumsum(inside.to(torch.int64), 0) - 1
    e = inside[rows]
    return newid[rows[e]], newid[cols[e]], w[e], int(inside.sum())


def rgg(n: int, radius: float | None = None, seed: int = 3, lcc: bool = True):
    """Random geometric graph in the unit square, unit weights; default
    radius gives mean degree 10."""
    from scipy.spatial import cKDTree

    if radius is None:
        radius = float(np.sqrt(10.0 / (np.pi * n)))
    pts = np.random.default_rng(seed).random((n, 2))
    pairs = cKDTree(pts).query_pairs(radius, output_type="ndarray")
    return from_edges(n, pairs[:, 0], pairs[:, 1], None, lcc=lcc)


def path(n: int):
    i = np.arange(n - 1, dtype=np.int64)
    return from_edges(n, i, i + 1, None, lcc=False)


if __name__ == "__main__":
    import sys

    if len(sys.argv) == 11 and sys.argv[1] == "rmat-device":
        _, _, sc, ef, fa, fb, fc, sd, lc, outdir, tag = sys.argv
        arrays = _rmat_device_inproc(int(sc), int(ef), float(fa), float(fb), float(fc), int(sd), lc == "1")
        os.makedirs(outdir, exist_ok=True)
        for name, arr in zip(("rowptr", "col", "w"), arrays):
            tmp = os.path.join(outdir, f"{tag}_{name}.tmp.npy")
            np.save(tmp, arr)
            os.replace(tmp, os.path.join(outdir, f"{tag}_{name}.npy"))
    else:
        sys.exit("usage: graphs.py rmat-device scale edge_factor a b c seed lcc outdir tag")
