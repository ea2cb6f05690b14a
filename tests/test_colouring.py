"""The general layout's colouring, computed by the library on the host
(hpcgx_colour_rows: no device needed), is the oracle's (orc_colour_general,
pinned in test_oracle_general.py) colour for colour and row for row."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import general_matrices as gm

CASES = {"permuted_hpcg": lambda: gm.permuted_hpcg(16, seed=1),
         "random_spd": lambda: gm.random_spd(seed=2),
         "unsymmetric": lambda: gm.unsymmetric(sizes=(3000, 400), seed=3)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_library_colouring_equals_oracle(hip, case):
    levels = CASES[case]()
    P = po.Problem.from_rows(levels)
    for k, (nr, cols, _, f2c) in enumerate(levels):
        col, order, nc = hip._lib.colour_rows(nr, cols, f2c)
        lv = P.level(k)
        _, wnc, word = lv.general_colours(with_order=True)
        assert nc == wnc == lv.ncolours()
        assert np.array_equal(col, lv.colours())   # after the injected-rows ordering
        assert np.array_equal(order, word)


def test_library_colouring_rejects_bad_rows(hip):
    nr = np.array([1, 2], np.uint8)
    cols = np.zeros((2, 27), np.int32)
    cols[1, :2] = [1, 5]  # column out of range
    with pytest.raises(hip.HpcgxError):
        hip._lib.colour_rows(nr, cols)


@pytest.mark.parametrize("g,seed", [(16, 1), (16, 5), (24, 7), (32, 11), (32, 0), (32, 1), (48, 0),
                                    ((16, 32, 8), 5), ((40, 24, 16), 6), ((8, 8, 64), 7),
                                    ((64, 16, 16), 8)])
def test_renumbered_stencil_gets_eight_colours(hip, g, seed):
    """8 is the minimum for the 27-point stencil (a 2x2x2 cell is a clique);
    first fit in natural order of a random numbering needs about 17.  A start
    on a face instead of a corner of the grid (the last row of a
    breadth-first pass rather than its least coupled one) needed 11-12 for
    several of these numberings."""
    nr, cols, _, _ = gm.permuted_hpcg(g, levels=1, seed=seed)[0]
    assert hip._lib.colour_rows(nr, cols)[2] == 8
