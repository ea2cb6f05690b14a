"""Pins the oracle's general (unstructured) levels: rows in any order and
any pattern, coloured by orc_colour_general.

* natural-order kernels, the V-cycle and CG bit for bit against the
  reference compiled from its own sources, after its levels' rows are
  replaced in place by the same rows (ref_replace_rows);
* the colouring: equal to the reference's greedy example
  (OptimizeProblem.cpp:43-79) on structurally symmetric patterns, and free of
  same-colour couplings also for unsymmetric ones;
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


@pytest.mark.parametrize("case", sorted(CASES))
def test_general_colouring_is_reference_greedy_on_symmetric_patterns(case):
    P = po.Problem.from_rows(CASES[case]())
    for k in range(4):
        lv = P.level(k)
        col, nc = lv.general_colours()
        ref_col, ref_nc = lv.greedy_colours()
        assert nc == ref_nc and np.array_equal(col, ref_col)
        assert lv.ncolours() == nc
        _assert_independent(lv, col)


def _assert_independent(lv, col):
    nr, cols, _, _ = lv.arrays()
    for i in range(lv.n):
        for j in cols[i, :nr[i]]:
            assert j == i or col[j] != col[i]


def test_general_colouring_unsymmetric_pattern():
    levels = gm.unsymmetric(sizes=(600, 100), seed=4)
    P = po.Problem.from_rows(levels)
    col, nc = P.general_colours()
    _assert_independent(P, col)
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
