"""hpcgx_dot_noalloc, the dot product that allocates nothing (its partial
sums live in the library's static device scratch), against hpcgx_dot through
the C ABI: the same kernels in the same order, so the same bits.  Two calls
in a row on two problems, the second smaller, share the scratch; each must
give what hpcgx_dot gives on it, and both stay within 1e-13 of the oracle's
dot (ComputeDotProduct_ref.cpp:54-64)."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _dots(hip, ctx, x, y):
    L = hip.lib()
    n = ctx.info(0)["n"]
    dx, dy = ctx.vec(0, x), ctx.vec(0, y)
    a, b = C.c_double(), C.c_double()
    assert L.hpcgx_dot(ctx.h, n, dx.ptr, dy.ptr, C.byref(a)) == 0
    assert L.hpcgx_dot_noalloc(ctx.h, n, dx.ptr, dy.ptr, C.byref(b)) == 0
    # x . x takes the same path (one vector read twice)
    c = C.c_double()
    assert L.hpcgx_dot_noalloc(ctx.h, n, dx.ptr, dx.ptr, C.byref(c)) == 0
    xx = ctx.dot(dx, dx)
    dx.free()
    dy.free()
    return a.value, b.value, c.value, xx


@pytest.mark.parametrize("dims", [((48, 40, 32), (16, 16, 16)), ((24, 24, 24), (8, 8, 8))])
def test_dot_noalloc_matches_dot_twice_in_a_row(hip, dims):
    g = np.random.default_rng(11)
    for d in dims:   # the second problem is smaller: the scratch keeps the first call's contents
        ctx = hip.Context(0)
        try:
            ctx.generate(*d, levels=4)
            n = ctx.info(0)["n"]
            x, y = g.random(n) - 0.5, g.random(n) - 0.5
            a, b, c, xx = _dots(hip, ctx, x, y)
            assert b == a, (d, a, b)       # bit for bit
            assert c == xx, (d, c, xx)
            want = po.dot(x, y)
            assert abs(b - want) <= 1e-13 * abs(want), (d, b, want)
        finally:
            ctx.close()


def test_dot_noalloc_at_512_cubed_folds(hip):
    """Above 32768 producer blocks the partials are folded first: at 256^3
    (16384 blocks) they are not, at 512^3 (131072) they are."""
    for d in ((512, 512, 512), (256, 256, 256)):
        ctx = hip.Context(0)
        try:
            ctx.generate(*d, levels=2)
            n = ctx.info(0)["n"]
            g = np.random.default_rng(d[0])
            x, y = g.random(n), g.random(n)
            a, b, c, xx = _dots(hip, ctx, x, y)
            assert b == a and c == xx, (d, a, b, c, xx)
            want = float(np.dot(x, y))
            assert abs(b - want) <= 1e-12 * abs(want)
        finally:
            ctx.close()


def test_dot_noalloc_refuses_bad_arguments(hip):
    L = hip.lib()
    ctx = hip.Context(0)
    try:
        ctx.generate(16, 16, 16, levels=2)
        n = ctx.info(0)["n"]
        dx = ctx.vec(0, np.ones(n))
        out = C.c_double()
        assert L.hpcgx_dot_noalloc(ctx.h, n - 1, dx.ptr, dx.ptr, C.byref(out)) == -1
        assert L.hpcgx_dot_noalloc(ctx.h, n, None, dx.ptr, C.byref(out)) == -1
        assert L.hpcgx_dot_noalloc(ctx.h, n, dx.ptr, dx.ptr, None) == -1
        assert L.hpcgx_dot_noalloc(None, n, dx.ptr, dx.ptr, C.byref(out)) == -1
        assert L.hpcgx_dot_noalloc(ctx.h, n, dx.ptr, dx.ptr, C.byref(out)) == 0
        assert out.value == float(n)
        dx.free()
    finally:
        ctx.close()
    fresh = hip.Context(0)   # no problem yet: it cannot set one up without allocating
    try:
        out = C.c_double()
        assert L.hpcgx_dot_noalloc(fresh.h, 8, None, None, C.byref(out)) == -5
    finally:
        fresh.close()
