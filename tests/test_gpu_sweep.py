"""Seeded sweep over grid shapes: 20 grids with random extents (1 to 40 per
dimension, divisible as the level count needs) and 1 to 4 levels, each
generated on the device and checked against the oracle: SpMV, the coloured
symmetric Gauss-Seidel and the V-cycle (fp64, fp32, scaled fp16) bit for
bit, the CG residual history within 1e-8 relative.  Extents that are odd, 1
or 2 exercise empty and single-row colour blocks, padded slices and small
coarse levels."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _cases(count=20, seed=12345):
    g = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        levels = int(g.integers(1, 5))
        m = 1 << (levels - 1)
        dims = tuple(int(m * g.integers(1, max(2, 40 // m) + 1)) for _ in range(3))
        out.append((dims, levels))
    return out


@pytest.mark.parametrize("dims,levels", _cases())
def test_random_grid_matches_oracle(hip, dims, levels):
    g = np.random.default_rng(sum(dims) * 10 + levels)
    ctx = hip.Context(0)
    ctx.generate(*dims, levels=levels)
    P = po.Problem(*dims, levels=levels)
    n = P.n
    x, r, b = g.random(n), g.random(n) - 0.5, g.random(n)
    dx, dy = ctx.vec(0, x), ctx.vec(0)
    ctx.spmv(0, dx, dy)
    assert np.array_equal(dy.download(), P.level(0).spmv(x)), "spmv"
    dr, dx2 = ctx.vec(0, r), ctx.vec(0, x)
    ctx.symgs(0, dr, dx2)
    assert np.array_equal(dx2.download(), P.level(0).symgs(r, x, mode=2)), "symgs"
    dz = ctx.vec(0)
    ctx.mg(0, dr, dz)
    assert np.array_equal(dz.download(), P.mg(r, mode=2)), "mg"
    db, dxx = ctx.vec(0, b), ctx.vec(0)
    got = ctx.cg(db, dxx, max_iter=10, tol=0.0, history=True)
    want = P.cg(b, max_iter=10, mode=2, history=True)
    assert got["niters"] == want["niters"]
    h, hw = got["hist"] / got["normr0"], want["hist"] / want["normr0"]
    live = hw > 1e-10
    assert np.allclose(h[live], hw[live], rtol=1e-8, atol=0)
    # the fp32 and scaled fp16 V-cycles, bit for bit
    r32 = r.astype(np.float32)
    for code, want_mg in ((hip.HPCGX_PREC_FP32, P.mg_f32), (hip.HPCGX_PREC_FP16, P.mg_f16)):
        ctx.prepare(code)
        dm = ctx.vec(0)
        ctx.mg_mixed(code, dr, dm)
        assert np.array_equal(dm.download(), want_mg(r32).astype(np.float64)), f"mg precision {code}"
    ctx.close()


@pytest.mark.parametrize("dims,levels", _cases(count=8, seed=777))
def test_random_grid_kappa_operator_matches_oracle(hip, dims, levels):
    """The variable-coefficient operator on random grids: every level's matrix
    (built on the device from kappa injected level by level), the kappa
    gradient kernel and the V-cycle, bit for bit."""
    g = np.random.default_rng(sum(dims) * 7 + levels)
    n = int(np.prod(dims))
    kappa = np.exp(0.5 * g.standard_normal(n))
    ctx = hip.Context(0)
    ctx.generate(*dims, levels=levels, kappa=kappa)
    P = po.Problem(*dims, levels=levels, kappa=kappa)
    for lv in range(levels):
        onr, ocols, ovals, _ = P.level(lv).arrays()
        nr, cols, vals = ctx.export_level(lv)
        mask = np.arange(27)[None, :] < onr[:, None]
        assert np.array_equal(nr, onr)
        assert np.array_equal(cols[mask], ocols[mask])
        assert np.array_equal(vals[mask], ovals[mask]), f"level {lv}"
    lam, x, r = g.standard_normal(n), g.standard_normal(n), g.random(n)
    dl, dx, dg = ctx.vec(0, lam), ctx.vec(0, x), ctx.vec(0)
    ctx.kappa_grad(dl, dx, dg)
    assert np.array_equal(dg.download(), P.kappa_grad(lam, x))
    dr, dz = ctx.vec(0, r), ctx.vec(0)
    ctx.mg(0, dr, dz)
    assert np.array_equal(dz.download(), P.mg(r, mode=2))
    ctx.close()
