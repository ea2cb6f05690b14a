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


def hpcg_bytes(shapes, iters, sets=1, pre=1, post=1, int_bytes=4):
    """Bytes read and written by the reference's memory model
    (ReportResults.cpp:103-135) for `sets` CG sets of `iters` iterations;
    int_bytes is sizeof(local_int_t).  Its "Raw Read / Write / Total B/W"
    (:339-341) divide these by the CG time.  The model counts the fine-grid
    residual's and the post-smoother's writes once per stored nonzero
    (:124,128) and every sweep in full, so it exceeds what any implementation
    moves: it is reported for comparison with the reference's report, never
    as a roofline."""
    d = 8.0
    vi = d + int_bytes
    fsets, fit = float(sets), float(sets) * iters
    nrow, nnz = float(shapes[0][0]), float(shapes[0][1])
    reads = (3.0 * fit + fsets) * 2.0 * nrow * d          # dots: two vectors
    writes = (3.0 * fit + fsets) * d                       # dots: the scalar
    reads += (3.0 * fit + fsets) * 2.0 * nrow * d         # WAXPBY
    writes += (3.0 * fit + fsets) * nrow * d
    reads += (fit + fsets) * (nnz * vi + nrow * d)         # SpMV
    writes += (fit + fsets) * nrow * d
    for n_l, m_l, _ in shapes[:-1]:
        n_l, m_l = float(n_l), float(m_l)
        reads += pre * fit * (2.0 * m_l * vi + n_l * d)    # pre-smoother
        writes += pre * fit * n_l * d
        reads += fit * (m_l * vi + n_l * d)                # residual for the restriction
        writes += fit * m_l * d
        reads += post * fit * (2.0 * m_l * vi + n_l * d)   # post-smoother
        writes += post * fit * m_l * d
    n_c, m_c = float(shapes[-1][0]), float(shapes[-1][1])
    reads += fit * (2.0 * m_c * vi + n_c * d)              # the coarsest level's sweep
    writes += fit * n_c * d
    return dict(reads=reads, writes=writes, total=reads + writes)


def mxp_rate(shapes, iters, seconds):
    """HPCG-MxP-style rate of one mixed-precision solve: the operations the
    HPCG model (ReportResults.cpp:81-99) credits for one set of `iters`
    iterations (preamble included), counted as if done in FP64, over the
    solve's wall time.  It is not an HPCG rating: the reference forbids
    reduced precision, so it is reported apart from the FP64 rating."""
    flops = hpcg_flops(shapes, iters, sets=1)["total"]
    return {"iterations": int(iters), "flops": flops, "gflops": flops / seconds / 1e9}


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


DEFAULT_ORDER = (7, 3, 4, 5, 2, 6, 1, 0)  # parity class swept k-th forward (the library's default)


def zero_start_slots(order=DEFAULT_ORDER):
    """Slots a forward step from x = 0 reads for colour block k (27-point
    stencil, blocks in `order`): the neighbours whose class was swept
    before, i.e. whose x is no longer zero.  Block 0 reads none."""
    pos = {c: k for k, c in enumerate(order)}
    out = []
    for k, c in enumerate(order):
        cnt = 0
        for s in range(27):
            odd = int(s % 3 != 1) | (int((s // 3) % 3 != 1) << 1) | (int(s // 9 != 1) << 2)
            if pos[c ^ odd] < k:
                cnt += 1
        out.append(cnt)
    return out


def index_bytes_per_slot(forms):
    """Column-index bytes a kernel loads per ELL slot, from a level's slice
    counts per column form ({'uniform', 'delta', 'explicit'}, hpcgx_level_column_forms;
    None: explicit 32-bit indices): 0 / 1 / 4 bytes per slot, plus the
    per-slice 32-bit base of each slot (4/64 byte per slot) for the compact
    forms."""
    if forms is None:
        return 4.0
    tot = float(sum(forms.values()))
    if tot == 0:
        return 4.0
    fu, fd, fw = forms["uniform"] / tot, forms["delta"] / tot, forms["explicit"] / tot
    return fd * 1.0 + fw * 4.0 + (fu + fd) * 4.0 / 64.0


def implementation_bytes_per_iteration(nx, ny, nz, levels=4, v=8, s=8, forms=None, order=DEFAULT_ORDER):
    """Bytes the FP64 iteration's kernels move as built, in the convention
    of algorithmic_bytes_per_iteration (each array element once per pass,
    neighbour gathers not counted), but counting matrix slots as loaded: all
    27 slots of a row (the zeros of rows at the grid boundary included), only
    the slots of blocks swept before in the zero-start forward sweep (plus
    the diagonal), 15 colour steps per symmetric sweep (the last block's
    backward step is elided), column-index bytes per the level's column
    forms (forms[l], see index_bytes_per_slot), no separate r.z pass (its
    partials come from the post-smoother).  Even grids (every colour block
    holds n/8 rows)."""
    shapes = level_shapes(nx, ny, nz, levels)
    zs = zero_start_slots(order)
    total = 0.0
    for l, (n, m, d) in enumerate(shapes):
        ib = index_bytes_per_slot(None if forms is None else forms[l])
        nb = n / 8.0
        full = 27 * (v + ib) + 3 * s
        # zero-start forward sweep: block k reads zs[k] neighbour slots, the diagonal, r; writes x
        pre = sum(nb * (zs[k] * (v + ib) + v + 2 * s) for k in range(8)) + 7 * nb * full
        total += pre
        if l + 1 < levels:
            nc = shapes[l + 1][0]
            total += 15 * nb * full                                   # post-smoother
            total += nb * 27 * (v + ib) + n * s + nc * (4 + 2 * s)    # residual + restriction
            total += nc * (4 + 3 * s)                                 # prolongation
    n0 = shapes[0][0]
    ib0 = index_bytes_per_slot(None if forms is None else forms[0])
    total += n0 * 27 * (8 + ib0) + 2 * n0 * 8                         # CG SpMV + p.Ap
    total += 40 * n0 + 24 * n0                                         # x and p updates; r update + r.r
    return total
