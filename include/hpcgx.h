/*
 * hpcgx.h -- C ABI of the MI355X-native HPCG PCG iteration phase.
 *
 * One shared library (hpcg_amd/libhpcgx.so) exports these symbols with plain
 * C types: pointers, sizes, int64_t, double.  Host buffers belong to the
 * caller; device buffers made by hpcgx_vec_alloc and all matrix storage
 * belong to the context.  Every call returns HPCGX_OK (0) or a negative
 * code; hpcgx_error_string() describes the last failure of a context.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the HPCG 3.1 reference's src/ directory).  The C++ hooks in
 * hpcg_amd/dropin/ implement the reference signatures on top of them.
 *
 * Device vectors live in the context's row order: rows grouped by colour
 * (8 colours, the parity classes of the grid coordinates, for a 27-point
 * stencil level; greedy colours for a general level) and padded to 64 rows
 * per colour.  hpcgx_vec_upload / hpcgx_vec_download convert from/to the
 * reference's natural row order.  Padding entries are zero after an upload.
 *
 * Calls are ordered on the context's stream (hpcgx_set_stream); calls that
 * return a scalar to the host (dot, cg) synchronise that stream.  A context
 * is used by one host thread at a time.
 */
#ifndef HPCGX_H
#define HPCGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPCGX_OK 0
#define HPCGX_EINVAL (-1)      /* bad argument */
#define HPCGX_ENOTSTENCIL (-2) /* reserved: no longer returned (any rows upload, see hpcgx_level_upload) */
#define HPCGX_EHIP (-3)        /* a HIP runtime call failed */
#define HPCGX_ENOMEM (-4)      /* device or host allocation failed */
#define HPCGX_ESTATE (-5)      /* call out of order (e.g. level missing) */

/* storage precision of a multigrid copy */
#define HPCGX_PREC_FP64 0
#define HPCGX_PREC_FP32 1
#define HPCGX_PREC_FP16 2 /* values D^-1/2 A D^-1/2 in fp16, fp32 arithmetic */

/* column index formats (hpcgx_set_column_format) */
#define HPCGX_COLS_EXPLICIT 0 /* 32-bit column per stored entry, the reference's mtxIndL */
#define HPCGX_COLS_COMPACT 1  /* per 64-row slice: base + lane, + an int8 delta where needed */
#define HPCGX_COLS_COMPACT_ALWAYS 2 /* the compact form on every level, even where most slices are wide */

/* hpcgx_cg variants */
#define HPCGX_CG_FP64 0       /* CG.cpp recurrence, fp64 MG */
#define HPCGX_CG_MIXED_FP32 1 /* flexible PCG, fp64 outer, fp32 MG */
#define HPCGX_CG_MIXED_FP16 2 /* flexible PCG, fp64 outer, scaled fp16 MG */

typedef struct hpcgx_ctx hpcgx_ctx;

/* ---- context ---------------------------------------------------------- */

/* Opens a context on a HIP device.  Holds the device state that the
 * reference keeps behind SparseMatrix::optimizationData
 * (SparseMatrix.hpp:58-64, OptimizeProblem.cpp:37). */
int hpcgx_create(int device, hpcgx_ctx **out);
/* Starts the HIP runtime and the device's context ahead of hpcgx_create (no
 * reference counterpart: the reference has no device).  Safe to call from
 * another thread while the caller does host work; the drop-in's
 * GenerateProblem hook runs it that way, so the runtime comes up while the
 * host builds the matrix instead of at the start of OptimizeProblem. */
int hpcgx_device_init(int device);
/* Frees every device buffer of the context (the reference's DeleteMatrix,
 * SparseMatrix.hpp:151-183, for the device copies). */
int hpcgx_destroy(hpcgx_ctx *ctx);
const char *hpcgx_error_string(const hpcgx_ctx *ctx);
/* Orders all later work on `stream` (a hipStream_t; NULL is the null stream).
 * A new context works on a non-blocking stream of its own. */
int hpcgx_set_stream(hpcgx_ctx *ctx, void *stream);
int hpcgx_sync(hpcgx_ctx *ctx);

/* ---- problem (OptimizeProblem.cpp:37-100) ----------------------------- */

/* Sweep order of the problems uploaded or generated afterwards: order[k] is the
 * parity class ((x&1)|(y&1)<<1|(z&1)<<2) of the k-th colour of the forward
 * sweep (the backward sweep reverses it).  Default 7,3,4,5,2,6,1,0.  The
 * reference's colouring example (OptimizeProblem.cpp:42-97) fixes the colour
 * classes; this picks their order. */
int hpcgx_set_colour_order(hpcgx_ctx *ctx, const int *order);
/* Uploads one level in the reference's row-pointer form (SparseMatrix.hpp:47-51:
 * nonzerosInRow, mtxIndL, matrixValues) of an nx*ny*nz grid, levels from fine
 * (0) to coarse.  f2c: the level's f2cOperator (MGData.hpp:31), length = rows
 * of level+1, or NULL on the coarsest level.  Colours, reorders and packs the
 * rows into 27-slot sliced ELL on the device. */
int hpcgx_level_upload(hpcgx_ctx *ctx, int level, int nx, int ny, int nz, const char *nonzerosInRow,
                       const int *const *mtxIndL, const double *const *matrixValues, const int *f2c);
/* Same, from packed arrays: row i's entries at cols[i*27 + j], vals[i*27 + j]
 * for j < nnz_row[i].
 * Both uploads accept any rows, as the reference requires of every
 * optimisation (GenerateProblem.cpp:47-49): rows that are the 27-point
 * stencil of the stated grid (ascending columns) get the parity-class layout;
 * other rows (any pattern, <= 27 entries, a diagonal) get a general layout
 * coloured greedily like the reference's example (OptimizeProblem.cpp:43-79),
 * summed in the row's own entry order.  f2c then has (nx/2)(ny/2)(nz/2)
 * entries, the reference's coarse grid (GenerateCoarseProblem.cpp:50-53). */
int hpcgx_level_upload_packed(hpcgx_ctx *ctx, int level, int nx, int ny, int nz,
                              const uint8_t *nnz_row, const int *cols, const double *vals,
                              const int *f2c);
/* Generates the whole hierarchy on the device from the closed forms of
 * GenerateProblem_ref.cpp:131-183 and GenerateCoarseProblem.cpp:71-82
 * (replaces the host generation for large grids).  kappa: NULL for the HPCG
 * matrix, else the per-cell diffusivity of the fine grid in natural order
 * (host pointer, nx*ny*nz) for the variable-coefficient operator
 * a_ij = -(k_i+k_j)/2, a_ii = sum_j (k_i+k_j)/2 + (26-n_i) k_i; coarse
 * levels take kappa by injection. */
/* General rows without a grid: n rows packed as above, f2c the injection
 * onto the n_coarse rows of level+1 (one-to-one) or NULL.  Always the general
 * layout.  Replaces OptimizeProblem.cpp:37 for matrices the caller builds. */
int hpcgx_level_upload_rows(hpcgx_ctx *ctx, int level, int64_t n, const uint8_t *nnz_row, const int *cols,
                            const double *vals, const int *f2c, int64_t n_coarse);
int hpcgx_problem_generate(hpcgx_ctx *ctx, int nx, int ny, int nz, int levels, const double *kappa);
/* out[0..7]: rows, padded rows, nonzeros, nx, ny, nz, stored slots, levels */
int hpcgx_level_info(const hpcgx_ctx *ctx, int level, int64_t *out);
/* The colouring a level is swept in (OptimizeProblem.cpp:42-97): number of
 * colours, whether the level has the general layout, and (if colour is not
 * NULL) the colour of every row in natural order; colours are swept
 * 0..ncolours-1 forward and back down to 0. */
int hpcgx_level_colouring(const hpcgx_ctx *ctx, int level, int *ncolours, int *general, int *colour);
/* The colouring the general layout gives rows packed as for
 * hpcgx_level_upload_rows (OptimizeProblem.cpp:42-97 restated for general
 * patterns): first fit in a breadth-first order over the symmetrised
 * pattern, the order starting from a far row of each component; with an
 * injection f2c[n_coarse] (may be NULL), colours are then sorted by how many
 * injected rows they hold, so that colour is swept last forward.  colour[n];
 * order[n] (may be NULL): the breadth-first sequence.  Host only; needs no
 * device or context. */
int hpcgx_colour_rows(int64_t n, const uint8_t *nnz_row, const int *cols, const int *f2c, int64_t n_coarse,
                      int *colour, int *order, int *ncolours);
/* Replaces the diagonal of a level (ReplaceMatrixDiagonal, SparseMatrix.hpp:139-145);
 * diag is host, natural order.  Lower-precision copies are rebuilt. */
int hpcgx_level_set_diagonal(hpcgx_ctx *ctx, int level, const double *diag);
/* Reads the diagonal back (CopyMatrixDiagonal, SparseMatrix.hpp:126-132). */
int hpcgx_level_get_diagonal(hpcgx_ctx *ctx, int level, double *diag);
/* Reads a level's device matrix back in the reference's natural row order,
 * packed like hpcgx_level_upload_packed (unused slots: col -1, value 0); the
 * inverse of the upload, for inspection and tests (SparseMatrix.hpp:47-51). */
int hpcgx_level_export_packed(hpcgx_ctx *ctx, int level, uint8_t *nnz_row, int *cols, double *vals);
/* Column indices the kernels read (OptimizeProblem.cpp:37 lets the matrix be
 * stored in any form).  HPCGX_COLS_COMPACT (the default): per 64-row slice and
 * slot one 32-bit base; a slice whose stored entries of each slot sit at base
 * + lane reads no per-entry index, other slices read an int8 delta per entry,
 * and slices whose deltas do not fit in int8 keep the explicit 32-bit
 * indices.  Zero-valued entries may point anywhere in the level.  The form is
 * chosen from the uploaded columns alone.  A level on which at least half
 * the slices keep explicit indices reads the explicit indices throughout
 * (faster there); HPCGX_COLS_COMPACT_ALWAYS reads the compact form on those
 * levels too.  HPCGX_COLS_EXPLICIT: the 32-bit index of every entry (the
 * reference's mtxIndL).  Results are the same bit for bit; only the bytes
 * read differ. */
int hpcgx_set_column_format(hpcgx_ctx *ctx, int format);
/* Slices of a level per column form: counts[0] int8 deltas, counts[1]
 * explicit indices, counts[2] no per-entry index (zeros before the compact
 * copy exists); counts[3] 1 if the level's kernels read the compact copy
 * (the compact format is on and fewer than half the slices need explicit
 * indices), else 0.  counts has 4 entries. */
int hpcgx_level_column_forms(const hpcgx_ctx *ctx, int level, int64_t *counts);
/* Builds the fp32 or scaled-fp16 copies of every level for the mixed solvers. */
int hpcgx_prepare_precision(hpcgx_ctx *ctx, int precision);
/* Device bytes held by the context (OptimizeProblemMemoryUse, OptimizeProblem.cpp:103-107). */
double hpcgx_memory_bytes(const hpcgx_ctx *ctx);

/* ---- vectors (Vector.hpp:27-113) --------------------------------------- */

/* Device vector of one level (padded length), zero-filled. */
int hpcgx_vec_alloc(hpcgx_ctx *ctx, int level, double **dvec);
int hpcgx_vec_free(hpcgx_ctx *ctx, double *dvec);
/* host (natural order, n entries) -> device (colour order) */
int hpcgx_vec_upload(hpcgx_ctx *ctx, int level, double *dvec, const double *host, int64_t n);
/* device (colour order) -> host (natural order, n entries) */
int hpcgx_vec_download(hpcgx_ctx *ctx, int level, const double *dvec, double *host, int64_t n);
/* Device-to-device order conversion: dnat (natural order, n entries, device)
 * -> dvec (colour order, padded), and back.  For callers that keep their own
 * device buffers (e.g. torch tensors). */
int hpcgx_vec_from_natural(hpcgx_ctx *ctx, int level, double *dvec, const double *dnat);
int hpcgx_vec_to_natural(hpcgx_ctx *ctx, int level, const double *dvec, double *dnat);
/* ZeroVector (Vector.hpp:55-60) and friends: every true entry := value. */
int hpcgx_vec_fill(hpcgx_ctx *ctx, int level, double *dvec, double value);

/* ---- kernels ------------------------------------------------------------ */

/* y = A x on one level (ComputeSPMV, ComputeSPMV.cpp:40-45; arithmetic of
 * ComputeSPMV_ref.cpp:61-70, summed in the same order: bit-exact). */
int hpcgx_spmv(hpcgx_ctx *ctx, int level, const double *x, double *y);
/* One symmetric Gauss-Seidel step, 8 colours forward then backward
 * (ComputeSYMGS, ComputeSYMGS.cpp:50-55; row arithmetic of
 * ComputeSYMGS_ref.cpp:67-100). */
int hpcgx_symgs(hpcgx_ctx *ctx, int level, const double *r, double *x);
/* rc = (r - A x) at the injected rows of `level`, written in level+1 order
 * (ComputeSPMV_ref + ComputeRestriction_ref.cpp:51, fused: only the
 * injected rows' residual is formed). */
int hpcgx_restrict(hpcgx_ctx *ctx, int level, const double *r, const double *x, double *rc);
/* x[f2c] += xc (ComputeProlongation_ref.cpp:49); xc in level+1 order. */
int hpcgx_prolong(hpcgx_ctx *ctx, int level, const double *xc, double *x);
/* V-cycle from `level` down (ComputeMG, ComputeMG.cpp:33-38; structure of
 * ComputeMG_ref.cpp:39-63). */
int hpcgx_mg(hpcgx_ctx *ctx, int level, const double *r, double *x);
/* The V-cycle in reduced precision (no reference counterpart): x = MG(r) with
 * r rounded to fp32, the fp32 (HPCGX_PREC_FP32) or diagonally scaled fp16
 * (HPCGX_PREC_FP16) copies of every level, fp32 arithmetic, x widened to fp64.
 * Level-0 device vectors; needs hpcgx_prepare_precision. */
int hpcgx_mg_mixed(hpcgx_ctx *ctx, int precision, const double *r, double *x);
/* result = x . y over a whole level-0 vector, n = its row count (ComputeDotProduct,
 * ComputeDotProduct.cpp:41-47).  Fixed-order reduction: the same inputs give
 * the same bits on every run. */
int hpcgx_dot(hpcgx_ctx *ctx, int64_t n, const double *x, const double *y, double *result);
/* w = alpha x + beta y over a whole level-0 vector, n = its row count
 * (ComputeWAXPBY, ComputeWAXPBY.cpp:42-48; bit-exact).  The vector calls
 * refuse any other n (HPCGX_EINVAL): device vectors are in colour order, so a
 * shorter length would not select the reference's first n rows. */
int hpcgx_waxpby(hpcgx_ctx *ctx, int64_t n, double alpha, const double *x, double beta,
                 const double *y, double *w);
/* w = alpha x + beta y and result = w . w in one pass (no reference counterpart;
 * the pair WAXPBY+DOT of CG.cpp:122-123). */
int hpcgx_waxpby_dot(hpcgx_ctx *ctx, int64_t n, double alpha, const double *x, double beta,
                     const double *y, double *w, double *result);
/* hpcgx_dot that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and bits as hpcgx_dot, but the partial sums live in a
 * static 2.1 MB device array that each device holds until the process ends.
 * The problem must already be set up (HPCGX_ESTATE otherwise), and level 0
 * may have at most 2^28 padded rows (HPCGX_EINVAL beyond).  Every call on one
 * device shares that array: calls that overlap on one device (other threads,
 * other contexts' streams) race and must be serialised by the caller. */
int hpcgx_dot_noalloc(hpcgx_ctx *ctx, int64_t n, const double *x, const double *y, double *result);

/* ---- solver (CG, CG.cpp:59-143) ---------------------------------------- */

/* Preconditioned CG on level 0.  b, x: device vectors (colour order); x holds
 * the initial guess and receives the solution.  HPCGX_CG_FP64 follows the
 * CG.cpp recurrence exactly (stops when normr/normr0 <= tol or after
 * max_iter).  The mixed variants run a flexible (Polak-Ribiere) PCG with an
 * fp32 / scaled-fp16 MG and fp64 outer vectors, stop when the fp64 true
 * residual ||b - A x|| / normr0 <= tol, and need hpcgx_prepare_precision.
 * times: NULL or the reference's times[] (CG.cpp:133-141 slots 0..5,
 * accumulated).  hist: NULL or max_iter+1 doubles, normr after each
 * iteration (hist[0] = normr0).  With tol = 0 the FP64 variant stops only
 * on a norm that is zero or NaN; it runs that test on the device and waits
 * for the host once per 64 iterations (results are the same). */
int hpcgx_cg(hpcgx_ctx *ctx, int variant, const double *b, double *x, int max_iter, double tol,
             int do_preconditioning, int *niters, double *normr, double *normr0, double *times,
             double *hist);

/* ---- variable-coefficient adjoint (no reference counterpart) ----------- */

/* g = dL/dkappa for L(x(kappa)) given the adjoint lam = A^-1 dL/dx:
 * g_k = -1/2 sum_{j in N(k)} (lam_k - lam_j)(x_k - x_j) - (26 - n_k) lam_k x_k.
 * lam, x: level-0 device vectors; g: device vector (colour order). */
int hpcgx_kappa_grad(hpcgx_ctx *ctx, const double *lam, const double *x, double *g);

#ifdef __cplusplus
}
#endif
#endif
