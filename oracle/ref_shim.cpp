/*
 * ref_shim.cpp -- extern "C" access to the HPCG 3.1 reference's own code.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
 * with the unmodified reference sources (read in place from the reference
 * checkout) into oracle/_ref/libhpcg_ref.so.  Tests use it to pin the C
 * restatement in hpcg_oracle.c and to write the golden fixtures; bench.py's
 * cpu_baseline leg times the reference CG through it.  The product never
 * loads it.
 *
 * Everything here only drives reference functions: the problem comes from
 * GenerateGeometry / GenerateProblem / SetupHalo / GenerateCoarseProblem
 * exactly as main.cpp:121-144 builds it, and the kernels are the _ref
 * functions themselves.
 */
#include <cstdint>
#include <cstring>
#include <vector>

#include "CG_ref.hpp"
#include "CGData.hpp"
#include "ComputeDotProduct_ref.hpp"
#include "ComputeMG_ref.hpp"
#include "ComputeSPMV_ref.hpp"
#include "ComputeSYMGS_ref.hpp"
#include "ComputeWAXPBY_ref.hpp"
#include "GenerateCoarseProblem.hpp"
#include "GenerateGeometry.hpp"
#include "GenerateProblem.hpp"
#include "Geometry.hpp"
#include "SetupHalo.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"
#include "mytimer.hpp"

namespace {
struct RefProblem {
  SparseMatrix A;
  Vector b, x, xexact;
  CGData data;
  int levels;
};

SparseMatrix *level_of(RefProblem *p, int level) {
  SparseMatrix *m = &p->A;
  for (int l = 0; l < level && m; ++l) m = m->Ac;
  return m;
}

Vector wrap(double *v, int n) {
  Vector w;
  w.localLength = n;
  w.values = v;
  w.optimizationData = 0;
  return w;
}
}  // namespace

extern "C" {

/* main.cpp:121-144 for one rank: geometry, fine problem, halo, coarse levels. */
void *ref_create(int nx, int ny, int nz, int levels) {
  RefProblem *p = new RefProblem;
  Geometry *geom = new Geometry;
  GenerateGeometry(1, 0, 1, 0, 0, 0, nx, ny, nz, 0, 0, 0, geom);
  InitializeSparseMatrix(p->A, geom);
  GenerateProblem(p->A, &p->b, &p->x, &p->xexact);
  SetupHalo(p->A);
  SparseMatrix *cur = &p->A;
  for (int l = 1; l < levels; ++l) {
    GenerateCoarseProblem(*cur);
    cur = cur->Ac;
  }
  p->levels = levels;
  InitializeSparseCGData(p->A, p->data);
  return p;
}

void ref_destroy(void *h) {
  RefProblem *p = static_cast<RefProblem *>(h);
  DeleteMatrix(p->A);
  DeleteCGData(p->data);
  DeleteVector(p->b);
  DeleteVector(p->x);
  DeleteVector(p->xexact);
  delete p;
}

/* out: nrow, ncol, nnz(low 31 bits are enough for the fixture grids), nx, ny, nz */
int ref_level_info(void *h, int level, int64_t *out) {
  SparseMatrix *m = level_of(static_cast<RefProblem *>(h), level);
  if (!m) return 1;
  out[0] = m->localNumberOfRows;
  out[1] = m->localNumberOfColumns;
  out[2] = m->localNumberOfNonzeros;
  out[3] = m->geom->nx;
  out[4] = m->geom->ny;
  out[5] = m->geom->nz;
  return 0;
}

/* Packs one level's row-pointer arrays into [n*27] arrays (unused slots:
 * col -1, value 0), the diagonal slot per row, and f2c (length nrow of the
 * next level, or nothing on the coarsest). */
int ref_level_export(void *h, int level, unsigned char *nnz_row, int *cols, double *vals,
                     int *diag_slot, int *f2c) {
  SparseMatrix *m = level_of(static_cast<RefProblem *>(h), level);
  if (!m) return 1;
  for (int i = 0; i < m->localNumberOfRows; ++i) {
    const int k = m->nonzerosInRow[i];
    nnz_row[i] = (unsigned char)k;
    for (int j = 0; j < 27; ++j) {
      cols[(int64_t)i * 27 + j] = j < k ? m->mtxIndL[i][j] : -1;
      vals[(int64_t)i * 27 + j] = j < k ? m->matrixValues[i][j] : 0.0;
    }
    diag_slot[i] = (int)(m->matrixDiagonal[i] - m->matrixValues[i]);
  }
  if (m->mgData && f2c) {
    const int nc = m->Ac->localNumberOfRows;
    std::memcpy(f2c, m->mgData->f2cOperator, sizeof(int) * nc);
  }
  return 0;
}

/* Replaces the rows of one level in place with arbitrary rows of the same
 * count (<= 27 entries each, one of them the diagonal), writing into the
 * reference's own per-row arrays (GenerateProblem_ref.cpp:104-110) the
 * fields its kernels read: nonzerosInRow, mtxIndL, matrixValues,
 * matrixDiagonal (the diagonal entry, as :160 sets it).  f2c (may be
 * NULL) replaces the level's f2cOperator for its nrow(level+1) coarse rows.
 * Returns nonzero if a row cannot be stored. */
int ref_replace_rows(void *h, int level, const unsigned char *nnz_row, const int *cols, const double *vals,
                     const int *f2c) {
  SparseMatrix *m = level_of(static_cast<RefProblem *>(h), level);
  if (!m) return 1;
  const int n = m->localNumberOfRows;
  int64_t nnz = 0;
  for (int i = 0; i < n; ++i) {
    const int k = nnz_row[i];
    if (k < 1 || k > 27) return 2; /* numberOfNonzerosPerRow, GenerateProblem_ref.cpp:67 */
    int ds = -1;
    for (int j = 0; j < k; ++j) {
      const int c = cols[(int64_t)i * 27 + j];
      if (c < 0 || c >= m->localNumberOfColumns) return 3;
      m->mtxIndL[i][j] = c;
      m->matrixValues[i][j] = vals[(int64_t)i * 27 + j];
      if (c == i && ds < 0) ds = j;
    }
    if (ds < 0) return 4;
    m->nonzerosInRow[i] = (char)k;
    m->matrixDiagonal[i] = &m->matrixValues[i][ds];
    nnz += k;
  }
  m->localNumberOfNonzeros = nnz;
  if (f2c && m->mgData && m->Ac)
    std::memcpy(m->mgData->f2cOperator, f2c, sizeof(int) * m->Ac->localNumberOfRows);
  return 0;
}

void ref_rhs(void *h, double *b, double *x, double *xexact) {
  RefProblem *p = static_cast<RefProblem *>(h);
  const int n = p->A.localNumberOfRows;
  std::memcpy(b, p->b.values, sizeof(double) * n);
  std::memcpy(x, p->x.values, sizeof(double) * n);
  std::memcpy(xexact, p->xexact.values, sizeof(double) * n);
}

/* Writes through the reference's own ReplaceMatrixDiagonal (SparseMatrix.hpp:139-145). */
void ref_set_diagonal(void *h, int level, const double *d) {
  SparseMatrix *m = level_of(static_cast<RefProblem *>(h), level);
  Vector dv = wrap(const_cast<double *>(d), m->localNumberOfRows);
  ReplaceMatrixDiagonal(*m, dv);
}

int ref_spmv(void *h, int level, double *x, double *y) {
  SparseMatrix *m = level_of(static_cast<RefProblem *>(h), level);
  Vector xv = wrap(x, m->localNumberOfColumns), yv = wrap(y, m->localNumberOfRows);
  return ComputeSPMV_ref(*m, xv, yv);
}

int ref_symgs(void *h, int level, double *r, double *x) {
  SparseMatrix *m = level_of(static_cast<RefProblem *>(h), level);
  Vector rv = wrap(r, m->localNumberOfRows), xv = wrap(x, m->localNumberOfColumns);
  return ComputeSYMGS_ref(*m, rv, xv);
}

int ref_mg(void *h, int level, double *r, double *x) {
  SparseMatrix *m = level_of(static_cast<RefProblem *>(h), level);
  Vector rv = wrap(r, m->localNumberOfRows), xv = wrap(x, m->localNumberOfColumns);
  return ComputeMG_ref(*m, rv, xv);
}

double ref_dot(int n, double *x, double *y) {
  Vector xv = wrap(x, n), yv = wrap(y, n);
  double r = 0.0, t = 0.0;
  ComputeDotProduct_ref(n, xv, yv, r, t);
  return r;
}

int ref_waxpby(int n, double alpha, double *x, double beta, double *y, double *w) {
  Vector xv = wrap(x, n), yv = wrap(y, n), wv = wrap(w, n);
  return ComputeWAXPBY_ref(n, alpha, xv, beta, yv, wv);
}

/* CG_ref on b (NULL: the problem's own b) and the caller's x (in/out).
 * times: the reference's 10-slot timing vector (may be NULL). */
int ref_cg(void *h, const double *b, double *x, int max_iter, double tol, int precond, int *niters,
           double *normr, double *normr0, double *times) {
  RefProblem *p = static_cast<RefProblem *>(h);
  const int n = p->A.localNumberOfRows;
  std::memcpy(p->x.values, x, sizeof(double) * n);
  std::vector<double> t(10, 0.0);
  std::vector<double> bcopy;
  Vector bv = p->b;
  if (b) {
    bcopy.assign(b, b + n);
    bv = wrap(&bcopy[0], n);
  }
  const int ierr = CG_ref(p->A, p->data, bv, p->x, max_iter, tol, *niters, *normr, *normr0,
                          &t[0], precond != 0);
  std::memcpy(x, p->x.values, sizeof(double) * n);
  if (times) std::memcpy(times, &t[0], sizeof(double) * 10);
  return ierr;
}

double ref_timer(void) { return mytimer(); }
}
