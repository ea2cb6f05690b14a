"""HPCG's FLOP model (ReportResults.cpp:75-99) and the algorithmic-byte model
used for roofline fractions.

Algorithmic bytes count each array element once per pass: matrix values
(v bytes) and column indices (i bytes: 4 for the fp64 matrices, 1 for the
int8 deltas of the compact fp32/fp16 copies) per stored nonzero of the rows a
kernel touches, plus the vector elements it reads or writes (s bytes each).
ELL padding, the gathered re-reads of neighbour entries and metadata (the
compact copies' per-slice bases) are not counted, so achieved/peak is a lower
bound on how busy HBM is.
"""


def level_shapes(nx, ny, nz, levels=4):
    """[(rows, nonzeros, (nx, ny, nz))] per level; nnz = prod(3 d - 2)."""
    out = []
    for l in range(levels):
        d = (nx >> l, ny >> l, nz >> l)
        nnz = 1
        for v in d:
            nnz *= 3 * v - 2
        out.append((d[0] * d[1] * d[2], nnz, d))
    return out


def hpcg_flops(shapes, iters, sets=1, ref_iters=None, pre=1, post=1):
    """Floating-point operations credited by ReportResults.cpp:81-99 for
    `sets` CG sets of `iters` iterations.  ref_iters: the reference's 50 when
    the convergence-overhead correction (:99) applies, else None."""
    fsets, fit = float(sets), float(sets) * iters
    nrow, nnz = float(shapes[0][0]), float(shapes[0][1])
    ddot = (3.0 * fit + fsets) * 2.0 * nrow
    waxpby = (3.0 * fit + fsets) * 2.0 * nrow
    spmv = (fit + fsets) * 2.0 * nnz
    mg = 0.0
    for n_l, m_l, _ in shapes[:-1]:
        mg += pre * fit * 4.0 * m_l + fit * 2.0 * m_l + post * fit * 4.0 * m_l
    mg += fit * 4.0 * shapes[-1][1]
    total = ddot + waxpby + spmv + mg
    if ref_iters is not None:
        total *= float(ref_iters) / float(iters)
    return dict(ddot=ddot, waxpby=waxpby, spmv=spmv, mg=mg, total=total)


def colour_nnz(nx, ny, nz):
    """Stored nonzeros of the rows of each parity class c = (x&1)|(y&1)<<1|(z&1)<<2."""
    def per_dim(n):
        s = [0, 0]
        for i in range(n):
            s[i & 1] += 3 - (i == 0) - (i == n - 1)
        return s
    sx, sy, sz = per_dim(nx), per_dim(ny), per_dim(nz)
    return [sx[c & 1] * sy[(c >> 1) & 1] * sz[(c >> 2) & 1] for c in range(8)]


def colour_rows(nx, ny, nz):
    def per_dim(n):
        return [(n + 1) // 2, n // 2]
    rx, ry, rz = per_dim(nx), per_dim(ny), per_dim(nz)
    return [rx[c & 1] * ry[(c >> 1) & 1] * rz[(c >> 2) & 1] for c in range(8)]


def kernel_bytes(kind, n, m, v=8, s=8, nc=0, mc=0, i=4):
    """Algorithmic bytes of one call (i: column-index bytes per nonzero).
    spmv: m(v+i) + 2ns            symgs (fwd+bwd): 2[m(v+i) + 3ns]
    gs_step (one colour, full):  m(v+i) + 3ns  with n, m of that colour
    restrict (fused):            mc(v+i) + ns_c(read r) + n_c(4 + 2s)... see below
    prolong: nc(4 + 3s)           dot: 16n   waxpby / waxpby_dot: 24n"""
    if kind == "spmv":
        return m * (v + i) + 2 * n * s
    if kind == "symgs":
        return 2 * (m * (v + i) + 3 * n * s)
    if kind == "gs_step":
        return m * (v + i) + 3 * n * s
    if kind == "restrict":
        # injected rows only: their nonzeros (mc = m/8 for the HPCG grid), the
        # fine x once, r at those rows, inj index and the coarse write
        return mc * (v + i) + n * s + nc * (4 + 2 * s)
    if kind == "prolong":
        return nc * (4 + 3 * s)
    if kind == "dot":
        return 16 * n
    if kind in ("waxpby", "waxpby_dot"):
        return 24 * n
    raise ValueError(kind)


def algorithmic_bytes_per_iteration(nx, ny, nz, levels=4, v_mg=8, s_mg=8, fused=True, i_mg=None):
    """Bytes of one CG iteration: MG V-cycle + SpMV + 3 dots + 3 WAXPBYs
    (fused: rz dot 16n, p update 24n, SpMV+dot, update+dot 48n).  i_mg:
    column-index bytes of the MG matrices (default 4 for fp64 values, 1 for
    the compact reduced-precision copies)."""
    if i_mg is None:
        i_mg = 4 if v_mg == 8 else 1
    shapes = level_shapes(nx, ny, nz, levels)
    n0, m0, _ = shapes[0]
    total = 0.0
    for l, (n, m, d) in enumerate(shapes):
        total += kernel_bytes("symgs", n, m, v_mg, s_mg, i=i_mg)          # pre (or coarsest)
        if l + 1 < levels:
            nc, mcl, _ = shapes[l + 1]
            total += kernel_bytes("symgs", n, m, v_mg, s_mg, i=i_mg)      # post
            if fused:
                total += kernel_bytes("restrict", n, m, v_mg, s_mg, nc=nc, mc=m // 8, i=i_mg)
            else:
                total += kernel_bytes("spmv", n, m, v_mg, s_mg, i=i_mg) + nc * (4 + 3 * s_mg)
            total += kernel_bytes("prolong", n, m, v_mg, s_mg, nc=nc)
    if v_mg != 8 or s_mg != 8:
        total += 24 * n0  # casts at the MG boundary
    total += kernel_bytes("spmv", n0, m0)
    if fused:
        total += 16 * n0 + 24 * n0 + 48 * n0
    else:
        total += 3 * 16 * n0 + 3 * 24 * n0
    return total


def hpcg_rhs(nx, ny, nz):
    """b_i = 26 - (nnz_i - 1) in natural order (GenerateProblem_ref.cpp:178),
    nnz_i the product of in-grid neighbour counts per dimension."""
    import numpy as np

    def cnt(n):
        c = np.full(n, 3.0)
        c[0] -= 1
        c[-1] -= 1
        if n == 1:
            c[0] = 1.0
        return c
    nnz = cnt(nz)[:, None, None] * cnt(ny)[None, :, None] * cnt(nx)[None, None, :]
    return (26.0 - (nnz - 1.0)).reshape(-1)
