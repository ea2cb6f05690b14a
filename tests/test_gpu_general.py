"""General (unstructured) matrices through the C ABI, against the oracle's
general levels (pinned bit for bit against the reference build in
tests/test_oracle_general.py).

The reference requires every optimisation to work for general sparse
matrices (GenerateProblem.cpp:47-49).  Rows that are not the 27-point stencil
of their grid take the general layout: first-fit colouring in breadth-first
order, rows' own entry order, the diagonal wherever the row holds it, and transfers over a list of
injected rows.  Bars as for the stencil layout: bit-exact matrix, colouring,
SpMV, symmetric Gauss-Seidel, restriction, prolongation and V-cycles (fp64,
fp32, scaled fp16); the CG residual history within 1e-8 relative.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import general_matrices as gm

pytestmark = pytest.mark.gpu

CASES = {
    # upload with the grid: the stencil check fails, the general layout is taken
    "permuted_hpcg_grid": (lambda: gm.permuted_hpcg(16, seed=1), (16, 16, 16)),
    # upload without a grid
    "random_spd_rows": (lambda: gm.random_spd(seed=2), None),
    "unsymmetric_rows": (lambda: gm.unsymmetric(seed=3), None),
    # larger and non-cubic renumbered hierarchies (more slices per colour,
    # list transfers over thousands of injected rows)
    "permuted_hpcg_32_rows": (lambda: gm.permuted_hpcg(32, seed=5), None),
    "permuted_hpcg_48x32x16_grid": (lambda: gm.permuted_hpcg((48, 32, 16), seed=6), (48, 32, 16)),
}


def rng(seed):
    return np.random.default_rng(seed)


def _upload(hip, levels, grid):
    ctx = hip.Context(0)
    for k, (nr, cols, vals, f2c) in enumerate(levels):
        if grid is None:
            nc = len(levels[k + 1][0]) if f2c is not None else 0
            ctx.upload_rows(k, nr, cols, vals, f2c, nc)
        else:
            dims = tuple(d >> k for d in grid)
            ctx.upload_level(k, dims, nr, cols, vals, f2c)
    return ctx


@pytest.fixture(scope="module")
def cases(hip):
    out = {}
    for name, (make, grid) in CASES.items():
        levels = make()
        out[name] = (_upload(hip, levels, grid), po.Problem.from_rows(levels), levels)
    yield out
    for ctx, _, _ in out.values():
        ctx.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_general_layout_matrix_and_colouring(cases, name):
    ctx, P, levels = cases[name]
    for k, (nr, cols, vals, _) in enumerate(levels):
        col, nc, general = ctx.colouring(k)
        assert general
        lv = P.level(k)
        assert nc == lv.ncolours() and np.array_equal(col, lv.colours())
        enr, ecols, evals = ctx.export_level(k)
        assert np.array_equal(enr, nr)
        mask = np.arange(27)[None, :] < nr[:, None]
        assert np.array_equal(ecols[mask], cols[mask])
        assert np.array_equal(evals[mask], vals[mask])


@pytest.mark.parametrize("name", sorted(CASES))
def test_general_spmv_symgs_bitexact(cases, name):
    ctx, P, _ = cases[name]
    for k in range(4):
        lv = P.level(k)
        x = rng(10 + k).random(lv.n)
        dx, dy = ctx.vec(k, x), ctx.vec(k)
        ctx.spmv(k, dx, dy)
        assert np.array_equal(dy.download(), lv.spmv(x)), f"spmv level {k}"
        r, x0 = rng(20 + k).random(lv.n), rng(30 + k).random(lv.n)
        dr, dx0 = ctx.vec(k, r), ctx.vec(k, x0)
        ctx.symgs(k, dr, dx0)
        assert np.array_equal(dx0.download(), lv.symgs(r, x0, mode=2)), f"symgs level {k}"
        for v in (dx, dy, dr, dx0):
            v.free()


@pytest.mark.parametrize("name", sorted(CASES))
def test_general_restrict_prolong_mg_bitexact(cases, name):
    ctx, P, _ = cases[name]
    for k in range(3):
        lv, cv = P.level(k), P.level(k + 1)
        r, x, xc = rng(40 + k).random(lv.n), rng(50 + k).random(lv.n), rng(60 + k).random(cv.n)
        dr, dx, drc, dxc = ctx.vec(k, r), ctx.vec(k, x), ctx.vec(k + 1), ctx.vec(k + 1, xc)
        ctx.restrict(k, dr, dx, drc)
        assert np.array_equal(drc.download(), lv.restrict(r, lv.spmv(x)))
        ctx.prolong(k, dxc, dx)
        assert np.array_equal(dx.download(), lv.prolong(xc, x))
        for v in (dr, dx, drc, dxc):
            v.free()
    r = rng(70).random(P.n) - 0.5
    dr, dz = ctx.vec(0, r), ctx.vec(0)
    ctx.mg(0, dr, dz)
    assert np.array_equal(dz.download(), P.mg(r, mode=2))


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
@pytest.mark.parametrize("name", ["permuted_hpcg_grid", "random_spd_rows"])
def test_general_mixed_mg_bitexact(hip, cases, name, prec):
    _, _, levels = cases[name]
    ctx = _upload(hip, levels, CASES[name][1])
    P = po.Problem.from_rows(levels)
    code = hip.HPCGX_PREC_FP32 if prec == "fp32" else hip.HPCGX_PREC_FP16
    ctx.prepare(code)
    r = rng(80).random(P.n) - 0.5
    dr, dz = ctx.vec(0, r), ctx.vec(0)
    ctx.mg_mixed(code, dr, dz)
    r32 = r.astype(np.float32)
    want = (P.mg_f32(r32) if prec == "fp32" else P.mg_f16(r32)).astype(np.float64)
    assert np.array_equal(dz.download(), want)
    ctx.close()


@pytest.mark.parametrize("name", ["permuted_hpcg_grid", "random_spd_rows"])
def test_general_cg_history(cases, name):
    ctx, P, _ = cases[name]
    b = rng(90).random(P.n)
    db, dx = ctx.vec(0, b), ctx.vec(0)
    got = ctx.cg(db, dx, max_iter=30, tol=0.0, history=True)
    want = P.cg(b, max_iter=30, mode=2, history=True)
    assert got["niters"] == want["niters"]
    h, hw = got["hist"] / got["normr0"], want["hist"] / want["normr0"]
    live = hw > 1e-10
    assert np.allclose(h[live], hw[live], rtol=1e-8, atol=0)


def test_general_diagonal_roundtrip(hip):
    levels = gm.random_spd(sizes=(700, 90), seed=11)
    ctx = _upload(hip, levels, None)
    P = po.Problem.from_rows(levels)
    d = P.diagonal()
    assert np.array_equal(ctx.get_diagonal(0), d)
    d2 = d * 1.5
    ctx.set_diagonal(0, d2)
    P.set_diagonal(d2)
    assert np.array_equal(ctx.get_diagonal(0), d2)
    r, x0 = rng(12).random(P.n), rng(13).random(P.n)
    dr, dx = ctx.vec(0, r), ctx.vec(0, x0)
    ctx.symgs(0, dr, dx)
    assert np.array_equal(dx.download(), P.symgs(r, x0, mode=2))
    ctx.close()
