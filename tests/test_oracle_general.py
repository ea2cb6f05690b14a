"""Pins the oracle's general (unstructured) levels: rows in any order and
any pattern, coloured by orc_colour_general.

* natural-order kernels, the V-cycle and CG bit for bit against the
  reference compiled from its own sources, after its levels' rows are
  replaced in place by the same rows (ref_replace_rows);
* the colouring: free of same-colour couplings (also for unsymmetric
  patterns), a first fit in its own order (every smaller colour is held by an
  earlier neighbour), that order a breadth-first sequence, and 8 colours -- the
  minimum, a 2x2x2 cell is a clique -- on a randomly renumbered 27-point
  stencil, where the reference's natural-order greedy example
  (OptimizeProblem.cpp:43-79) needs 17;
* coloured symmetric Gauss-Seidel against a dense textbook statement in the
  colour order.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import general_matrices as gm

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference parity build not present")

CASES = {"permuted_hpcg": lambda: gm.permuted_hpcg(16, seed=1),
         "random_spd": lambda: gm.random_spd(seed=2)}


@needs_ref
@pytest.mark.parametrize("case", sorted(CASES))
def test_general_oracle_matches_reference_build(case):
    levels = CASES[case]()
    P = po.Problem.from_rows(levels)
    R = po.RefProblem(16, 16, 16)
    for k, (nr, cols, vals, f2c) in enumerate(levels):
        R.replace_rows(k, nr, cols, vals, f2c)
    g = np.random.default_rng(3)
    for k in range(4):
        lv = P.level(k)
        x = g.random(lv.n)
        assert np.array_equal(lv.spmv(x), R.spmv(x, level=k))
        r, x0 = g.random(lv.n), g.random(lv.n)
        assert np.array_equal(lv.symgs(r, x0, mode=0), R.symgs(r, x0, level=k))
        assert np.array_equal(lv.mg(r, mode=0), R.mg(r, level=k))
    b = g.random(P.n)
    want = R.cg(b=b, max_iter=30, tol=0.0)
    got = P.cg(b, max_iter=30, tol=0.0, mode=0)
    assert got["niters"] == want["niters"]
    assert np.array_equal(got["x"], want["x"])


def _sym_adjacency(lv):
    nr, cols, _, _ = lv.arrays()
    adj = [set() for _ in range(lv.n)]
    for i in range(lv.n):
        for c in cols[i, :nr[i]]:
            if c != i:
                adj[i].add(int(c))
                adj[int(c)].add(i)
    return adj


def _assert_first_fit_bfs(lv, col, order):
    adj = _sym_adjacency(lv)
    n = lv.n
    assert sorted(order) == list(range(n))
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    last_parent = -1
    for q, i in enumerate(order):
        earlier = [j for j in adj[i] if pos[j] < q]
        used = {col[j] for j in earlier}
        assert col[i] not in used
        assert all(c in used for c in range(col[i]))       # first fit
        if earlier:                                          # breadth-first:
            parent = min(pos[j] for j in earlier)            # parents come in
            assert parent >= last_parent                     # order
            last_parent = parent
        else:
            last_parent = q  # a new component starts here


@pytest.mark.parametrize("case", sorted(CASES))
def test_general_colouring_is_first_fit_in_breadth_first_order(case):
    P = po.Problem.from_rows(CASES[case]())
    for k in range(4):
        lv = P.level(k)
        col, nc, order = lv.general_colours(with_order=True)
        assert lv.ncolours() == nc == col.max() + 1
        _assert_independent(lv, col)
        _assert_first_fit_bfs(lv, col, order)


@pytest.mark.parametrize("case", sorted(CASES))
def test_injected_colours_are_swept_last(case):
    """orc_link orders a general level's colours by how many injected rows
    they hold (ascending): the colour with most injected rows is swept last
    forward and first backward.  The ordering only relabels the first-fit
    colouring (same classes)."""
    levels = CASES[case]()
    P = po.Problem.from_rows(levels)
    for k in range(3):
        lv = P.level(k)
        col = lv.colours()
        f2c = levels[k][3]
        counts = np.bincount(col[f2c], minlength=lv.ncolours())
        assert np.all(np.diff(counts) >= 0)
        raw, nc = lv.general_colours()
        pairs = set(zip(raw.tolist(), col.tolist()))
        assert len(pairs) == nc   # a relabelling: classes unchanged
        _assert_independent(lv, col)


def test_general_colouring_of_renumbered_stencil_is_minimal():
    P = po.Problem.from_rows(gm.permuted_hpcg(16, levels=2, seed=3))
    assert P.general_colours()[1] == 8
    assert P.level(1).general_colours()[1] == 8
    assert P.greedy_colours()[1] > 8   # the natural-order example does worse here


def _assert_independent(lv, col):
    nr, cols, _, _ = lv.arrays()
    for i in range(lv.n):
        for j in cols[i, :nr[i]]:
            assert j == i or col[j] != col[i]


def test_general_colouring_unsymmetric_pattern():
    levels = gm.unsymmetric(sizes=(600, 100), seed=4)
    P = po.Problem.from_rows(levels)
    col, nc, order = P.general_colours(with_order=True)
    _assert_independent(P, col)
    _assert_first_fit_bfs(P, col, order)
    # the pattern really is unsymmetric
    A = gm.dense(levels[0])
    assert np.any((A != 0) != (A.T != 0))


def test_set_colouring_rejects_coupled_rows():
    P = po.Problem.from_rows(gm.random_spd(sizes=(200,), seed=5))
    with pytest.raises(ValueError):
        P.set_colouring(np.zeros(P.n, np.int32))


@pytest.mark.parametrize("case", ["permuted", "random", "unsymmetric"])
def test_general_coloured_symgs_is_textbook_gs(case):
    levels = {"permuted": lambda: gm.permuted_hpcg(8, levels=1, seed=6),
              "random": lambda: gm.random_spd(sizes=(300,), seed=7),
              "unsymmetric": lambda: gm.unsymmetric(sizes=(300,), seed=8)}[case]()
    P = po.Problem.from_rows(levels)
    A = gm.dense(levels[0])
    col, nc = P.general_colours()
    order = np.concatenate([np.flatnonzero(col == k) for k in range(nc)])
    g = np.random.default_rng(9)
    r, x0 = g.random(P.n), g.random(P.n)
    Ap = A[np.ix_(order, order)]
    D, L, U = np.diag(np.diag(Ap)), np.tril(Ap, -1), np.triu(Ap, 1)
    xp = np.linalg.solve(D + L, r[order] - U @ x0[order])
    xp = np.linalg.solve(D + U, r[order] - L @ xp)
    want = np.empty(P.n)
    want[order] = xp
    for mode in (1, 2):
        assert np.allclose(P.symgs(r, x0, mode=mode), want, rtol=1e-12, atol=1e-13)
