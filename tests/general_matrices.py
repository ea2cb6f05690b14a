"""Test matrices for the general (unstructured) layout, as hierarchies of
(nnz_row[n], cols[n,27], vals[n,27], f2c) levels in the reference's row form
(SparseMatrix.hpp:44-52).  The reference requires optimisations to work for
general sparse matrices (GenerateProblem.cpp:47-49); these are the cases the
tests hold the device to:

* permuted_hpcg: the HPCG hierarchy renumbered by a random permutation per
  level (P A P^T, f2c mapped along), entries kept in their original order, so
  rows are no longer ascending and no grid describes them;
* random_spd: random symmetric patterns with up to 26 off-diagonal entries per
  row, strictly diagonally dominant values (SPD), the diagonal at a random
  position in each row, and a random one-to-one f2c onto the next level;
* unsymmetric: random_spd with some a_ji removed while a_ij stays (the
  colouring must still keep coupled rows apart).
"""
import numpy as np

from oracle import pyoracle as po

SLOTS = 27


def _pack(rows, n):
    nr = np.zeros(n, np.uint8)
    cols = np.zeros((n, SLOTS), np.int32)
    vals = np.zeros((n, SLOTS))
    for i, ents in enumerate(rows):
        nr[i] = len(ents)
        for j, (c, v) in enumerate(ents):
            cols[i, j], vals[i, j] = c, v
        cols[i, len(ents):] = -1
    return nr, cols, vals


def permuted_hpcg(nx, levels=4, seed=0):
    """nx: the edge of a cubic grid, or its (nx, ny, nz)."""
    dims = tuple(nx) if isinstance(nx, (tuple, list)) else (nx, nx, nx)
    P = po.Problem(*dims, levels=levels)
    g = np.random.default_rng(seed)
    arrs = [P.level(k).arrays() for k in range(levels)]
    perms = [g.permutation(len(a[0])) for a in arrs]  # new index of old row
    out = []
    for k, (nr, cols, vals, f2c) in enumerate(arrs):
        n, p = len(nr), perms[k]
        inv = np.empty(n, np.int64)
        inv[p] = np.arange(n)
        nr2 = nr[inv]
        cols2 = np.full((n, SLOTS), -1, np.int32)
        vals2 = np.zeros((n, SLOTS))
        for new in range(n):
            old = inv[new]
            m = nr[old]
            cols2[new, :m] = p[cols[old, :m]]
            vals2[new, :m] = vals[old, :m]
        f2c2 = None
        if f2c is not None:
            pc = perms[k + 1]
            f2c2 = np.empty(len(f2c), np.int32)
            f2c2[pc] = p[f2c]
        out.append((nr2, cols2, vals2, f2c2))
    return out


def _random_level(n, g, max_off=26, density=8, symmetric=True):
    nbrs = [dict() for _ in range(n)]
    target = g.integers(1, density + 1, size=n)
    for i in range(n):
        tries = 0
        while len(nbrs[i]) < target[i] and tries < 4 * density:
            tries += 1
            j = int(g.integers(0, n))
            if j == i or j in nbrs[i] or len(nbrs[j]) >= max_off or len(nbrs[i]) >= max_off:
                continue
            a = -float(g.uniform(0.1, 1.0))
            nbrs[i][j] = a
            nbrs[j][i] = a
    if not symmetric:
        for i in range(n):
            for j in list(nbrs[i]):
                if j > i and i in nbrs[j] and g.random() < 0.3:
                    del nbrs[j][i]  # keep a_ij, drop a_ji
    rows = []
    for i in range(n):
        ents = list(nbrs[i].items())
        d = sum(abs(v) for _, v in ents) + float(g.uniform(0.5, 2.0))
        ents.append((i, d))
        order = g.permutation(len(ents))  # arbitrary entry order, diagonal anywhere
        rows.append([ents[t] for t in order])
    return _pack(rows, n)


def random_spd(sizes=(4096, 512, 64, 8), seed=0, symmetric=True):
    g = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        nr, cols, vals = _random_level(n, g, symmetric=symmetric)
        f2c = None
        if k + 1 < len(sizes):
            f2c = np.sort(g.choice(n, size=sizes[k + 1], replace=False)).astype(np.int32)
            g.shuffle(f2c)
        out.append((nr, cols, vals, f2c))
    return out


def unsymmetric(sizes=(4096, 512, 64, 8), seed=0):
    return random_spd(sizes, seed, symmetric=False)


def dense(level_arrays):
    nr, cols, vals = level_arrays[:3]
    n = len(nr)
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(nr[i]):
            A[i, cols[i, j]] += vals[i, j]
    return A
