"""The column index formats give the same bits.

The compact format (the default) stores, per 64-row slice and slot, one
32-bit base, plus an int8 delta per entry where the slice's entries do not
all sit at base + lane, or keeps the explicit 32-bit indices where a delta
does not fit in int8.  Zero-valued entries (a neighbour outside the grid)
may point at any row, their product being an exact zero.  The explicit
format reads the 32-bit index of every entry, as the reference's mtxIndL.
A level on which at least half the slices keep explicit indices reads the
explicit indices under the default; HPCGX_COLS_COMPACT_ALWAYS sends those
levels through the compact kernels too.  Every kernel must give the same
bits in all three settings, on stencil levels (all three slice forms occur),
renumbered and random general levels (delta and wide slices) and the
variable-coefficient operator, and all are held to the oracle elsewhere."""
import numpy as np
import pytest

import general_matrices as gm
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _stencil(hip, dims, kappa=None):
    ctx = hip.Context(0)
    ctx.generate(*dims, levels=4, kappa=kappa)
    return ctx


def _rows(hip, levels):
    ctx = hip.Context(0)
    for k, (nr, cols, vals, f2c) in enumerate(levels):
        nc = 0 if f2c is None else len(f2c)
        ctx.upload_rows(k, nr, cols, vals, f2c, nc)
    return ctx


CASES = {
    "hpcg_16": lambda hip: _stencil(hip, (16, 16, 16)),
    "hpcg_24x8x16": lambda hip: _stencil(hip, (24, 8, 16)),
    "hpcg_256x256x8_strips": lambda hip: _stencil(hip, (256, 256, 8)),
    "kappa_16": lambda hip: _stencil(hip, (16, 16, 16), np.exp(0.5 * np.random.default_rng(1).standard_normal(4096))),
    "permuted_hpcg_16": lambda hip: _rows(hip, gm.permuted_hpcg(16, seed=3)),
    "random_spd": lambda hip: _rows(hip, gm.random_spd(seed=4)),
}


def _run_all(hip, ctx, fmt):
    """SpMV, SYMGS, restriction, V-cycle (fp64, fp32, fp16) and 20 CG
    iterations with history, in format fmt."""
    ctx.set_column_format(fmt)
    n = ctx.info(0)["n"]
    g = np.random.default_rng(5)
    x, r = g.random(n) - 0.5, g.random(n)
    dx, dr, dy = ctx.vec(0, x), ctx.vec(0, r), ctx.vec(0)
    out = {}
    ctx.spmv(0, dx, dy)
    out["spmv"] = dy.download()
    ctx.symgs(0, dr, dx)
    out["symgs"] = dx.download()
    dc = ctx.vec(1)
    ctx.restrict(0, dr, dx, dc)
    out["restrict"] = dc.download()
    ctx.mg(0, dr, dy)
    out["mg"] = dy.download()
    for name, prec in (("mg32", hip.HPCGX_PREC_FP32), ("mg16", hip.HPCGX_PREC_FP16)):
        ctx.prepare(prec)
        ctx.mg_mixed(prec, dr, dy)
        out[name] = dy.download()
    db, dz = ctx.vec(0, r), ctx.vec(0)
    o = ctx.cg(db, dz, max_iter=20, tol=0.0, history=True)
    out["cg_hist"], out["cg_x"] = o["hist"], dz.download()
    for v in (dx, dr, dy, dc, db, dz):
        v.free()
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_compact_and_explicit_columns_agree(hip, case):
    ctx = CASES[case](hip)
    try:
        a = _run_all(hip, ctx, hip.HPCGX_COLS_EXPLICIT)
        b = _run_all(hip, ctx, hip.HPCGX_COLS_COMPACT)
        c = _run_all(hip, ctx, hip.HPCGX_COLS_COMPACT_ALWAYS)   # wide slices through the compact kernels
        for k in a:
            assert np.array_equal(a[k], b[k]), (case, k)
            assert np.array_equal(a[k], c[k]), (case, k)
        forms = ctx.column_forms(0)
        assert sum(forms.values()) == ctx.info(0)["npad"] // 64
        # the default reads the compact copy unless half the slices are wide
        ctx.set_column_format(hip.HPCGX_COLS_COMPACT)
        want = "compact" if 2 * forms["explicit"] < sum(forms.values()) else "explicit"
        assert ctx.column_reads(0) == want, (forms, ctx.column_reads(0))
        if case.startswith("hpcg") or case.startswith("kappa"):
            # stencil levels: no slice needs explicit indices, and most need
            # no per-entry index.  The others are the slices whose base + lane
            # would leave [0, npad) on a lane without a stored entry (their
            # neighbours lie in the first or last rows of the level): 18 of
            # the 64 slices at 16^3, a vanishing share on large grids
            assert forms["explicit"] == 0, forms
            assert forms["uniform"] >= 0.7 * sum(forms.values()), forms
    finally:
        ctx.close()


def test_compact_columns_match_oracle_on_general_levels(hip):
    """The default (compact) format on a renumbered hierarchy: SpMV against
    the oracle's natural-order arithmetic (bit-exact against the reference's)."""
    levels = gm.permuted_hpcg(16, seed=8)
    ctx = _rows(hip, levels)
    try:
        P = po.Problem.from_rows(levels)
        r = np.random.default_rng(9).random(P.n)
        dr, dz = ctx.vec(0, r), ctx.vec(0)
        ctx.spmv(0, dr, dz)
        assert np.array_equal(dz.download(), P.spmv(r))
        forms = ctx.column_forms(0)
        assert forms["uniform"] < sum(forms.values())   # renumbered: per-entry indices needed
    finally:
        ctx.close()
