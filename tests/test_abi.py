"""CPU-side checks of the boundary: the HIP library loads and exports every
entry point declared in include/hpcgx.h; the C++ drop-in library exports the
reference's hook symbols; the FLOP / byte models reproduce the reference's
closed forms.  No device calls."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(hpcgx_[a-z0-9_]+)\s*\(", text)))


def exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True)
    return {l.split()[-1] for l in out.stdout.splitlines() if l.strip()}


def test_library_loads_and_exports_header(hip):
    from hpcg_amd import _lib
    syms = exported(_lib.LIB_PATH)
    decl = declared("hpcgx.h")
    assert len(decl) >= 25
    missing = [s for s in decl if s not in syms]
    assert not missing, missing
    L = ctypes.CDLL(_lib.LIB_PATH)
    for s in decl:
        assert hasattr(L, s)
    assert set(_lib.SIGNATURES) <= set(decl)


def test_null_context_is_an_error(hip):
    L = hip.lib()
    assert L.hpcgx_error_string(None) == b"null context"
    assert L.hpcgx_destroy(None) == -1
    assert L.hpcgx_sync(None) == -1


def test_flop_model_matches_reference_closed_form():
    from hpcg_amd.models import hpcg_flops, level_shapes
    sh = level_shapes(192, 192, 192)
    assert [s[1] for s in sh] == [189119224, 23393656, 2863288, 343000]
    n, m = sh[0][0], sh[0][1]
    per_it = 12 * n + 2 * m + 10 * sum(s[1] for s in sh[:3]) + 4 * sh[3][1]
    f = hpcg_flops(sh, iters=50, sets=1)
    assert f["total"] == pytest.approx(50 * per_it + 4 * n + 2 * m, rel=1e-15)
    assert per_it / 1e9 == pytest.approx(2.618, abs=5e-4)


def test_memory_model_matches_the_reference_report():
    """hpcg_bytes restates the reference's read/write model
    (ReportResults.cpp:103-135).  Pinned against a report the reference's own
    ReportResults wrote (the reference driver with the drop-in hooks at 192^3,
    one set of 50 iterations): Raw Read B/W=7940.07, Raw Write B/W=1834.94
    GB/s over Benchmark Time Summary::Total=0.101896 s, and Floating Point
    Operations Summary::Total=1.31322e+11 for the FLOP model beside it; each
    printed to 6 significant digits."""
    from hpcg_amd.models import hpcg_bytes, hpcg_flops, level_shapes
    sh = level_shapes(192, 192, 192)
    b = hpcg_bytes(sh, iters=50, sets=1)
    t = 0.101896
    assert b["reads"] / t / 1e9 == pytest.approx(7940.07, rel=2e-5)
    assert b["writes"] / t / 1e9 == pytest.approx(1834.94, rel=2e-5)
    assert hpcg_flops(sh, iters=50, sets=1)["total"] == pytest.approx(1.31322e11, rel=1e-5)
    # a mistake the pin catches: post-smoother writes per row, not per nonzero
    wrong = b["writes"] - 50 * sum(s[1] - s[0] for s in sh[:-1]) * 8.0
    assert wrong / t / 1e9 != pytest.approx(1834.94, rel=1e-3)


def test_mxp_rate_counts_the_solves_iterations():
    """The mixed solvers' HPCG-MxP-style rate credits the HPCG model's
    operations for exactly the iterations the solve took (plus the set's
    preamble: one dot, one WAXPBY, one SpMV), whatever their precision."""
    from hpcg_amd.models import level_shapes, mxp_rate
    sh = level_shapes(512, 512, 512)
    n, m = sh[0][0], sh[0][1]
    per_it = 12 * n + 2 * m + 10 * sum(s[1] for s in sh[:3]) + 4 * sh[3][1]
    r = mxp_rate(sh, 232, 6.26)
    assert r["iterations"] == 232
    assert r["flops"] == pytest.approx(232 * per_it + 4 * n + 2 * m, rel=1e-15)
    assert r["gflops"] == pytest.approx(r["flops"] / 6.26 / 1e9, rel=1e-15)
    assert mxp_rate(sh, 233, 1.0)["flops"] - mxp_rate(sh, 232, 1.0)["flops"] == pytest.approx(per_it, rel=1e-12)


def test_dropin_library_exports_reference_hooks():
    """The C++ hooks keep the reference signatures (CG.hpp:22-24,
    ComputeSPMV.hpp:20, ComputeSYMGS.hpp:20, ComputeMG.hpp:20,
    ComputeDotProduct.hpp:18-19, ComputeWAXPBY.hpp:18-19,
    OptimizeProblem.hpp:22,28, ExchangeHalo.hpp:19)."""
    so = os.path.join(ROOT, "hpcg_amd", "libhpcg_dropin.so")
    if not os.path.exists(so):
        pytest.skip("drop-in library not built")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    want = [
        "CG(SparseMatrix_STRUCT const&, CGData_STRUCT&, Vector_STRUCT const&, Vector_STRUCT&, int, double, "
        "int&, double&, double&, double*, bool)",
        "ComputeSPMV(SparseMatrix_STRUCT const&, Vector_STRUCT&, Vector_STRUCT&)",
        "ComputeSYMGS(SparseMatrix_STRUCT const&, Vector_STRUCT const&, Vector_STRUCT&)",
        "ComputeMG(SparseMatrix_STRUCT const&, Vector_STRUCT const&, Vector_STRUCT&)",
        "ComputeDotProduct(int, Vector_STRUCT const&, Vector_STRUCT const&, double&, double&, bool&)",
        "ComputeWAXPBY(int, double, Vector_STRUCT const&, double, Vector_STRUCT const&, Vector_STRUCT&, bool&)",
        "OptimizeProblem(SparseMatrix_STRUCT&, CGData_STRUCT&, Vector_STRUCT&, Vector_STRUCT&, Vector_STRUCT&)",
        "OptimizeProblemMemoryUse(SparseMatrix_STRUCT const&)",
        "ExchangeHalo(SparseMatrix_STRUCT const&, Vector_STRUCT&)",
        "HPCG_VectorHostModified", "HPCG_VectorHostZeroed", "HPCG_VectorRelease",
        "HPCG_MatrixDiagonalReplaced", "HPCG_MatrixRelease",
    ]
    for w in want:
        assert w in out, w
