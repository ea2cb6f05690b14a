"""Writes the golden fixtures from the reference's own code.

Run in a checkout where oracle/_ref/libhpcg_ref.so is built from the
reference sources (`make -C oracle ref`):

    python tests/golden/make_golden.py

For each grid the fixture holds seeded inputs and the reference's outputs:
y = ComputeSPMV_ref(x), ComputeSYMGS_ref(r, x0), ComputeMG_ref(r), the
scaled-residual history of CG_ref (50 iterations from x = 0; entry k is
CG_ref run for k iterations), the final x, and the iteration counts of the
TestCG procedure (TestCG.cpp:68-111) driven through CG_ref.  Inputs use
numpy's PCG64 with fixed seeds (FillRandomVector's glibc rand() is not
portable), in FillRandomVector's range [1, 2).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402

GRIDS = [(16, 16, 16), (24, 24, 24)]


def testcg_counts(R, n):
    b, _, _ = R.rhs()
    orig = np.array([R.export(0)[2][i, R.export(0)[3][i]] for i in range(n)])
    scale = np.full(n, 1.0e6)
    gid = np.arange(n)
    scale[gid < 9] = (gid[gid < 9] + 2) * 1.0e6
    R.set_diagonal(0, orig * scale)
    counts = []
    for prec in (False, True):
        out = R.cg(max_iter=50, tol=1e-12, precond=prec, b=b * scale)
        counts.append(out["niters"])
    R.set_diagonal(0, orig)
    return np.array(counts, np.int32)


def make(grid):
    R = po.RefProblem(*grid)
    n = R.n
    g = np.random.default_rng(1234 + grid[0])
    x = g.random(n) + 1.0
    r = g.random(n) + 1.0
    x0 = g.random(n) + 1.0
    y = R.spmv(x)
    sg = R.symgs(r, x0)
    mg = R.mg(r)[:n]
    hist = [1.0]
    xf = None
    for k in range(1, 51):
        out = R.cg(max_iter=k)
        hist.append(out["normr"] / out["normr0"])
        xf = out["x"]
    tc = testcg_counts(R, n)
    name = os.path.join(HERE, "hpcg_%dx%dx%d.npz" % grid)
    np.savez_compressed(name, grid=np.array(grid, np.int32), x=x, y=y, r=r, x0=x0, symgs=sg, mg=mg,
                        cg_hist=np.array(hist), cg_x=xf, testcg=tc)
    print("wrote", name, os.path.getsize(name), "bytes; TestCG iterations", tc.tolist())


if __name__ == "__main__":
    for grid in GRIDS:
        make(grid)
