"""Kernel-level parity of the HIP path (through the C ABI) against the oracle.

Bars: bit-exact for the matrix, SpMV, coloured SYMGS, restriction,
prolongation, the V-cycle and WAXPBY (same per-row summation order, no FMA on
either side); dot products within 1e-13 relative (only the summation order
differs).  The oracle's natural-order functions are themselves bit-exact
against the reference build (tests/test_oracle.py).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# (16, 256, 8), (32, 256, 16) and (256, 256, 8) have blocks with >= 128 rows
# of Y: their finest rows are stored in Y strips (block_index; strips of 16
# rows of Y for the last), the others in natural order.  (8, 8, 8) is the
# smallest grid with 4 levels (the coarsest is 1 x 1 x 1).
GRIDS = [(8, 8, 8), (16, 16, 16), (24, 24, 24), (32, 16, 24), (64, 48, 40), (16, 256, 8), (32, 256, 16),
         (256, 256, 8)]


@pytest.fixture(scope="module")
def problems(hip):
    out = {}
    for g in GRIDS:
        ctx = hip.Context(0)
        ctx.generate(*g, levels=4)
        out[g] = (ctx, po.Problem(*g, levels=4))
    yield out
    for ctx, _ in out.values():
        ctx.close()


def rng(seed):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("g", GRIDS)
def test_generated_matrix_equals_reference_layout(problems, g):
    ctx, P = problems[g]
    for l in range(4):
        lv = P.level(l)
        onr, ocols, ovals, _ = lv.arrays()
        nr, cols, vals = ctx.export_level(l)
        assert ctx.info(l)["nnz"] == lv.nnz
        assert np.array_equal(nr, onr)
        mask = np.arange(27)[None, :] < onr[:, None]
        assert np.array_equal(cols[mask], ocols[mask])
        assert np.array_equal(vals[mask], ovals[mask])


def test_upload_equals_generate(hip):
    P = po.Problem(16, 16, 16)
    up = hip.Context(0)
    for l in range(4):
        lv = P.level(l)
        nr, cols, vals, f2c = lv.arrays()
        up.upload_level(l, lv.dims, nr, cols, vals, f2c)
    gen = hip.Context(0).generate(16, 16, 16)
    for l in range(4):
        a, b = up.export_level(l), gen.export_level(l)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
    x = rng(5).random(P.n) + 1.0
    ys = []
    for ctx in (up, gen):
        dx, dy = ctx.vec(0, x), ctx.vec(0)
        ctx.spmv(0, dx, dy)
        ys.append(dy.download())
    assert np.array_equal(ys[0], ys[1])


@pytest.mark.parametrize("dims", [(24, 16, 8), (8, 16, 40), (17, 9, 12)])
def test_upload_equals_generate_on_box_grids(hip, dims):
    """HPCG rows of grids with unequal (and odd) extents, uploaded: the host
    recognises them and the device writes them from the closed form, the same
    bits as generating them on the device, and SpMV and the V-cycle agree
    with the oracle."""
    levels = 3 if min(dims) >= 4 and all(d % 4 == 0 for d in dims) else 1
    P = po.Problem(*dims, levels=levels)
    up = hip.Context(0)
    for l in range(levels):
        lv = P.level(l)
        nr, cols, vals, f2c = lv.arrays()
        up.upload_level(l, lv.dims, nr, cols, vals, f2c if l + 1 < levels else None)
    gen = hip.Context(0).generate(*dims, levels=levels)
    for l in range(levels):
        for u, v in zip(up.export_level(l), gen.export_level(l)):
            assert np.array_equal(u, v), (dims, l)
    r = rng(13).random(P.n)
    dr, dz = up.vec(0, r), up.vec(0)
    up.spmv(0, dr, dz)
    assert np.array_equal(dz.download(), P.spmv(r))
    up.mg(0, dr, dz)
    assert np.array_equal(dz.download(), P.mg(r, mode=2))
    up.close()
    gen.close()


def test_upload_stencil_rows_with_other_values(hip):
    """Rows with the stencil's pattern but values other than the HPCG
    matrix's are copied and checked on the device (the HPCG matrix itself is
    recognised on the host and written from the closed form): the kappa
    operator uploaded equals the kappa operator generated, and a HPCG level
    with one diagonal entry changed keeps that entry, in the stencil layout."""
    kappa = np.exp(0.5 * rng(11).standard_normal(16 ** 3))
    P = po.Problem(16, 16, 16, kappa=kappa)
    up = hip.Context(0)
    for l in range(4):
        lv = P.level(l)
        nr, cols, vals, f2c = lv.arrays()
        up.upload_level(l, lv.dims, nr, cols, vals, f2c)
    gen = hip.Context(0).generate(16, 16, 16, kappa=kappa)
    for l in range(4):
        assert not up.colouring(l)[2]          # stencil layout
        for u, v in zip(up.export_level(l), gen.export_level(l)):
            assert np.array_equal(u, v), l
    r = rng(12).random(P.n)
    zs = []
    for ctx in (up, gen):
        dr, dz = ctx.vec(0, r), ctx.vec(0)
        ctx.mg(0, dr, dz)
        zs.append(dz.download())
    assert np.array_equal(zs[0], zs[1])
    assert np.array_equal(zs[0], P.mg(r, mode=2))
    up.close()
    gen.close()
    H = po.Problem(8, 8, 8, levels=1)
    nr, cols, vals, _ = H.arrays()
    vals = vals.copy()
    i = 100
    j = int(np.nonzero(cols[i, :nr[i]] == i)[0][0])
    vals[i, j] = 27.5
    ctx = hip.Context(0)
    ctx.upload_level(0, H.dims, nr, cols, vals, None)
    assert not ctx.colouring(0)[2]
    enr, ecols, evals = ctx.export_level(0)
    mask = np.arange(27)[None, :] < nr[:, None]
    assert np.array_equal(evals[mask], vals[mask]) and np.array_equal(ecols[mask], cols[mask])
    ctx.close()


def test_upload_non_stencil_takes_general_layout(hip):
    """A row that is not the stencil of the stated grid: the level is
    uploaded in the general layout (not rejected) and matches the oracle's
    general level."""
    P = po.Problem(8, 8, 8, levels=1)
    nr, cols, vals, _ = P.arrays()
    nr, cols, vals = nr.copy(), cols.copy(), vals.copy()
    a, b = 0, P.n - 1                      # opposite corners (8 entries each)
    cols[a, nr[a]], vals[a, nr[a]] = b, -0.5
    cols[b, nr[b]], vals[b, nr[b]] = a, -0.5
    nr[a] += 1
    nr[b] += 1
    ctx = hip.Context(0)
    ctx.upload_level(0, P.dims, nr, cols, vals, None)
    assert ctx.colouring(0)[2]
    G = po.Problem.from_rows([(nr, cols, vals, None)])
    x = rng(6).random(P.n)
    dx, dy = ctx.vec(0, x), ctx.vec(0)
    ctx.spmv(0, dx, dy)
    assert np.array_equal(dy.download(), G.spmv(x))
    r = rng(7).random(P.n)
    dr, dz = ctx.vec(0, r), ctx.vec(0)
    ctx.mg(0, dr, dz)
    assert np.array_equal(dz.download(), G.mg(r, mode=2))
    ctx.close()


def test_upload_rejects_rows_without_diagonal(hip):
    P = po.Problem(8, 8, 8, levels=1)
    nr, cols, vals, _ = P.arrays()
    cols = cols.copy()
    cols[5, :nr[5]] = np.where(cols[5, :nr[5]] == 5, 6, cols[5, :nr[5]])
    ctx = hip.Context(0)
    with pytest.raises(hip.HpcgxError, match="no diagonal"):
        ctx.upload_level(0, P.dims, nr, cols, vals, None)


@pytest.mark.parametrize("g", GRIDS)
def test_spmv_bitexact(problems, g):
    ctx, P = problems[g]
    for l in range(4):
        lv = P.level(l)
        x = rng(l + 1).random(lv.n) + 1.0
        dx, dy = ctx.vec(l, x), ctx.vec(l)
        ctx.spmv(l, dx, dy)
        assert np.array_equal(dy.download(), lv.spmv(x)), f"level {l}"
        dx.free(), dy.free()


def test_spmv_times_ones_is_b(problems):
    ctx, P = problems[(24, 24, 24)]
    b, _, xe = P.rhs()
    dx, dy = ctx.vec(0, xe), ctx.vec(0)
    ctx.spmv(0, dx, dy)
    assert np.array_equal(dy.download(), b)


@pytest.mark.parametrize("g", GRIDS)
def test_symgs_bitexact(problems, g):
    ctx, P = problems[g]
    for l in range(4):
        lv = P.level(l)
        r = rng(10 + l).random(lv.n)
        x0 = rng(20 + l).random(lv.n)
        dr, dx = ctx.vec(l, r), ctx.vec(l, x0)
        ctx.symgs(l, dr, dx)
        got = dx.download()
        assert np.array_equal(got, lv.symgs(r, x0, mode=2)), f"level {l}"
        # against the full coloured sweep (backward step of the last colour
        # performed): a few ulp apart
        full = lv.symgs(r, x0, mode=1)
        assert np.max(np.abs(got - full) / np.abs(full)) < 1e-14
        dr.free(), dx.free()


@pytest.mark.parametrize("g", GRIDS)
def test_restrict_prolong_bitexact(problems, g):
    ctx, P = problems[g]
    for l in range(3):
        lv, cv = P.level(l), P.level(l + 1)
        r = rng(30 + l).random(lv.n)
        x = rng(40 + l).random(lv.n)
        xc = rng(50 + l).random(cv.n)
        dr, dx, drc, dxc = ctx.vec(l, r), ctx.vec(l, x), ctx.vec(l + 1), ctx.vec(l + 1, xc)
        ctx.restrict(l, dr, dx, drc)
        assert np.array_equal(drc.download(), lv.restrict(r, lv.spmv(x)))
        ctx.prolong(l, dxc, dx)
        assert np.array_equal(dx.download(), lv.prolong(xc, x))
        for v in (dr, dx, drc, dxc):
            v.free()


@pytest.mark.parametrize("g", GRIDS)
def test_mg_bitexact(problems, g):
    ctx, P = problems[g]
    r = rng(60).random(P.n)
    dr, dz = ctx.vec(0, r), ctx.vec(0)
    want = P.mg(r, mode=2)
    for _ in range(2):  # the V-cycle must not depend on what the buffers held before
        ctx.mg(0, dr, dz)
        assert np.array_equal(dz.download(), want)
    dr.free(), dz.free()


def test_dot_and_waxpby(problems):
    ctx, P = problems[(32, 16, 24)]
    x = rng(70).random(P.n) + 1.0
    y = rng(71).random(P.n) - 0.5
    dx, dy, dw = ctx.vec(0, x), ctx.vec(0, y), ctx.vec(0)
    d = ctx.dot(dx, dy)
    ref = po.dot(x, y)
    assert abs(d - ref) <= 1e-13 * abs(ref)
    assert abs(ctx.dot(dx, dx) - po.dot(x, x)) <= 1e-13 * po.dot(x, x)
    assert ctx.dot(dx, dy) == d  # fixed-order reduction: identical bits every run
    for alpha, beta in ((1.0, -1.0), (0.75, 1.0), (-0.3, 2.5), (1.0, 0.0)):
        ctx.waxpby(alpha, dx, beta, dy, dw)
        assert np.array_equal(dw.download(), po.waxpby(alpha, x, beta, y))
        nrm = ctx.waxpby_dot(alpha, dx, beta, dy, dw)
        w = po.waxpby(alpha, x, beta, y)
        assert np.array_equal(dw.download(), w)
        assert abs(nrm - po.dot(w, w)) <= 1e-13 * po.dot(w, w)
    if po.ref_available():
        assert abs(d - po.ref_dot(x, y)) <= 1e-13 * abs(ref)
        assert np.array_equal(po.ref_waxpby(0.75, x, 1.0, y), po.waxpby(0.75, x, 1.0, y))


def test_vector_roundtrip_and_fill(problems):
    ctx, P = problems[(32, 16, 24)]
    x = rng(80).random(P.n)
    dx = ctx.vec(0, x)
    assert np.array_equal(dx.download(), x)
    ctx.fill(dx, 0.0)
    assert not dx.download().any()
    # padding stays zero: dot over the padded vector equals the true dot
    dx.upload(np.ones(P.n))
    assert ctx.dot(dx, dx) == float(P.n)


def test_other_colour_order_matches_oracle(hip):
    """A non-default sweep order (hpcgx_set_colour_order) against the oracle
    with the same order: the V-cycle stays bit-exact."""
    order = [6, 7, 3, 5, 4, 1, 2, 0]
    ctx = hip.Context(0)
    ctx.set_colour_order(order)
    ctx.generate(16, 16, 16)
    P = po.Problem(16, 16, 16)
    P.set_colour_order(order)
    r = rng(90).random(P.n)
    x0 = rng(91).random(P.n)
    dr, dx, dz = ctx.vec(0, r), ctx.vec(0, x0), ctx.vec(0)
    ctx.symgs(0, dr, dx)
    assert np.array_equal(dx.download(), P.symgs(r, x0, mode=2))
    ctx.mg(0, dr, dz)
    assert np.array_equal(dz.download(), P.mg(r, mode=2))


@pytest.mark.parametrize("grid,levels", [((15, 9, 7), 1), ((1, 1, 1), 1), ((2, 3, 1), 1), ((30, 18, 14), 2)])
def test_ragged_grids(hip, grid, levels):
    """Odd extents give colour blocks of different sizes (and empty ones on
    1-wide grids); the single-row grid is the smallest problem.  SpMV, the
    symmetric Gauss-Seidel and the V-cycle stay bit-exact, and the CG
    history follows the oracle's."""
    ctx = hip.Context(0)
    ctx.generate(*grid, levels=levels)
    P = po.Problem(*grid, levels=levels)
    x = rng(100).random(P.n)
    dx, dy = ctx.vec(0, x), ctx.vec(0)
    ctx.spmv(0, dx, dy)
    assert np.array_equal(dy.download(), P.spmv(x))
    r, x0 = rng(101).random(P.n), rng(102).random(P.n)
    dr, dx0, dz = ctx.vec(0, r), ctx.vec(0, x0), ctx.vec(0)
    ctx.symgs(0, dr, dx0)
    assert np.array_equal(dx0.download(), P.symgs(r, x0, mode=2))
    ctx.mg(0, dr, dz)
    assert np.array_equal(dz.download(), P.mg(r, mode=2))
    b = P.rhs()[0]
    db, dxx = ctx.vec(0, b), ctx.vec(0)
    got = ctx.cg(db, dxx, max_iter=10, tol=0.0, history=True)
    want = P.cg(b, max_iter=10, mode=2, history=True)
    assert got["niters"] == want["niters"]
    live = want["hist"] / want["normr0"] > 1e-10
    assert np.allclose(got["hist"][live] / got["normr0"], want["hist"][live] / want["normr0"], rtol=1e-8)
    ctx.close()


@pytest.mark.parametrize("g", [(16, 16, 16), (16, 256, 8)])
def test_mg_ignores_output_contents(hip, g):
    """ComputeMG_ref.cpp:43 zeroes x before the V-cycle: whatever the output
    vector holds on entry (here NaN) must not reach the result."""
    ctx = hip.Context(0)
    ctx.generate(*g)
    P = po.Problem(*g)
    r = rng(110).random(P.n)
    dr, dz = ctx.vec(0, r), ctx.vec(0, np.full(P.n, np.nan))
    ctx.mg(0, dr, dz)
    assert np.array_equal(dz.download(), P.mg(r, mode=2))
    ctx.close()


def test_wrong_size_vectors_are_refused(hip):
    """A vector made for a coarser level, or for a smaller problem that was
    since replaced, is refused before any launch."""
    ctx = hip.Context(0)
    ctx.generate(16, 16, 16)
    small = ctx.vec(1)
    big = ctx.vec(0)
    with pytest.raises(hip.HpcgxError, match="vector holds"):
        ctx.spmv(0, small, big)
    with pytest.raises(hip.HpcgxError, match="vector holds"):
        ctx.mg(0, big, small)
    ctx.generate(32, 32, 32)
    fresh = ctx.vec(0)
    with pytest.raises(hip.HpcgxError, match="vector holds"):
        ctx.cg(big, fresh, max_iter=1)
    ctx.close()


def test_vector_hooks_take_the_row_count_only(hip):
    """ComputeDotProduct / ComputeWAXPBY receive n = nrow (the whole vector).
    Device vectors are in colour order, so a shorter n would select other
    rows than the reference's first n: it is refused, before any launch."""
    ctx = hip.Context(0)
    ctx.generate(16, 16, 16)
    n, npad = ctx.info(0)["n"], ctx.info(0)["npad"]
    x, y, w = ctx.vec(0, np.ones(n)), ctx.vec(0, np.full(n, 2.0)), ctx.vec(0)
    assert ctx.dot(x, y) == 2.0 * n
    assert ctx.dot(x, y, n=npad) == 2.0 * n
    for bad in (n - 1, 5, npad + 1, 0):
        with pytest.raises(hip.HpcgxError, match="row count"):
            ctx.dot(x, y, n=bad)
        with pytest.raises(hip.HpcgxError, match="row count"):
            ctx.waxpby(1.0, x, 1.0, y, w, n=bad)
        with pytest.raises(hip.HpcgxError, match="row count"):
            ctx.waxpby_dot(1.0, x, 1.0, y, w, n=bad)
    ctx.close()
