/*
 * hpcg_oracle.h -- CPU restatement of the HPCG 3.1 reference arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.  The product
 * (hpcg_amd/libhpcgx.so) does not link, load or call it.
 *
 * Every function names the reference file:line whose arithmetic it restates
 * (paths are relative to the reference's src/ directory).  Matrices are held
 * as the reference holds them: per row, up to 27 entries in ascending column
 * order (GenerateProblem_ref.cpp:152-172), stored compactly in [n*27] arrays.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   generation, spmv, symgs (natural), mg (natural), restriction,
 *   prolongation, dot, waxpby, cg        -> bit-exact against the reference
 *                                           compiled from its own sources
 *                                           (oracle/_ref) and the committed
 *                                           golden fixtures, plus A*1 == b
 *   coloured symgs / mg                  -> reduces to a textbook dense
 *                                           Gauss-Seidel on a permuted matrix
 *                                           (numpy brute force on tiny grids),
 *                                           colour independence, MG symmetry
 *   kappa operator                       -> kappa == 1 reproduces the HPCG
 *                                           matrix bit-for-bit; symmetry;
 *                                           row sums == (26-n_i)*kappa_i
 *   kappa gradient                       -> central finite differences of an
 *                                           exact dense solve (numpy) on tiny
 *                                           grids
 *   mixed-precision MG / flexible CG     -> fp32 MG == fp64 MG to fp32
 *                                           rounding; scaled fp16 MG == fp64
 *                                           MG to fp16 rounding on
 *                                           hierarchies whose transfer factor
 *                                           sigma = sqrt(d_f/d_c) is far from
 *                                           1 (each level scaled by its own
 *                                           random diagonal) and on the kappa
 *                                           operator, a check that planted
 *                                           sigma mistakes are shown to fail;
 *                                           converged solutions checked by an
 *                                           fp64 true residual
 *   general levels (orc_create_rows,     -> natural-order spmv / symgs / mg /
 *   orc_link)                               cg bit-exact against the
 *                                           reference build with the same
 *                                           rows written into its own
 *                                           matrices (ref_replace_rows)
 *   orc_colour_general                   -> a first fit in its breadth-first
 *                                           order (restated in numpy) with
 *                                           independent colours; the order
 *                                           is a breadth-first sequence of
 *                                           every component
 *   coloured symgs on general levels     -> dense textbook Gauss-Seidel in
 *                                           the colour order
 *                                           (tests/test_oracle_general.py)
 */
#ifndef HPCG_ORACLE_H
#define HPCG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_level orc_level;

/* Build an L-level hierarchy (fine grid nx*ny*nz, each level halves every
 * dimension).  kappa == NULL gives the HPCG matrix; otherwise kappa holds the
 * per-cell diffusivity of the finest grid in natural order and every level
 * gets the variable-coefficient operator, coarse kappa taken by injection. */
orc_level *orc_create(int nx, int ny, int nz, int levels, const double *kappa);
void orc_destroy(orc_level *L);
orc_level *orc_coarse(orc_level *L);
int orc_nrow(const orc_level *L);
int64_t orc_nnz(const orc_level *L);
void orc_dims(const orc_level *L, int *nxyz);
const unsigned char *orc_nnz_row(const orc_level *L);
const int *orc_cols(const orc_level *L);      /* [n*27], row i at i*27 */
const double *orc_vals(const orc_level *L);   /* [n*27] */
const int *orc_f2c(const orc_level *L);       /* [n_coarse] or NULL */
const double *orc_kappa(const orc_level *L);  /* [n] or NULL */

void orc_rhs(const orc_level *L, double *b, double *x, double *xexact);
void orc_get_diagonal(const orc_level *L, double *d);
void orc_set_diagonal(orc_level *L, const double *d);

/* Colouring.  parity class p = (x&1) | (y&1)<<1 | (z&1)<<2 ; order[k] is the
 * parity class swept k-th in the forward sweep (backward sweep reverses it).
 * Applied to every level of the hierarchy.  Default: 7,3,4,5,2,6,1,0. */
void orc_set_colour_order(orc_level *L, const int order[8]);
void orc_colour_greedy(const orc_level *L, int *colour, int *ncolours);

/* General (unstructured) levels.  orc_create_rows builds one level from
 * arbitrary rows (<= 27 entries each, row order kept, one diagonal) and
 * colours it with orc_colour_general; orc_link attaches a coarse level with
 * an injection f2c (and reorders the fine level's colours so the colour of
 * the injected rows comes last).  orc_set_colouring replaces the colouring of one level
 * (colours swept in ascending order); -1 if two rows of one colour are
 * coupled. */
orc_level *orc_create_rows(int n, const unsigned char *nnz_row, const int *cols, const double *vals);
int orc_link(orc_level *fine, orc_level *coarse, const int *f2c);
void orc_colour_general(const orc_level *L, int *colour, int *ncolours);
void orc_colour_general_order(const orc_level *L, int *colour, int *ncolours, int *order);
/* colours sorted by their count of injected rows (orc_link applies it) */
void orc_order_colours(int n, int ncolours, int *colour, const int *f2c, int nc);
void orc_level_colours(const orc_level *L, int *colour);  /* the level's colouring now */
int orc_set_colouring(orc_level *L, const int *colour, int ncolours);
int orc_ncolours(const orc_level *L);
const int *orc_colour_rows(const orc_level *L);   /* rows grouped by colour */
const int *orc_colour_starts(const orc_level *L); /* [ncolours+1] */

/* Kernels (natural row order, full-length vectors).  mode selects the
 * Gauss-Seidel ordering: 0 natural (the reference), 1 coloured, 2 coloured
 * with the redundant backward step of the last colour elided (the device's). */
void orc_spmv(const orc_level *L, const double *x, double *y);
void orc_symgs(const orc_level *L, const double *r, double *x, int mode);
void orc_restrict(const orc_level *L, const double *r, const double *Axf, double *rc);
void orc_prolong(const orc_level *L, const double *xc, double *x);
void orc_mg(orc_level *L, const double *r, double *x, int mode);
double orc_dot(int64_t n, const double *x, const double *y);
void orc_waxpby(int64_t n, double alpha, const double *x, double beta, const double *y, double *w);
int orc_cg(orc_level *L, const double *b, double *x, int max_iter, double tol, int precond,
           int mode, int *niters, double *normr, double *normr0, double *hist);

/* Mixed precision (no reference counterpart); Gauss-Seidel in mode 2. */
void orc_mg_f32(orc_level *L, const float *r, float *x);
void orc_mg_f16(orc_level *L, const float *r, float *x);
int orc_cg_mixed(orc_level *L, const double *b, double *x, int max_iter, double tol, int mg_prec,
                 int *niters, double *normr, double *normr0, double *hist);

/* Variable-coefficient adjoint: g_k = -1/2 sum_{j in N(k)} (lam_k-lam_j)(x_k-x_j)
 *                                     - (26-n_k) lam_k x_k                      */
void orc_kappa_grad(const orc_level *L, const double *lam, const double *x, double *g);

/* Plants a known mistake in the scaled fp16 V-cycle so that a test can show
 * that its check catches it: 0 none, 1 sigma inverted (sqrt(d_c/d_f)),
 * 2 sigma dropped (1).  Applies to fp16 copies built afterwards.  Tests only. */
void orc_plant_fault(int which);

/* fp16 helpers (round to nearest even) shared with tests */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);

#ifdef __cplusplus
}
#endif
#endif
