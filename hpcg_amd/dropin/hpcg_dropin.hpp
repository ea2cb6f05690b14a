// hpcg_dropin.hpp -- the drop-in data structures of the HPCG hooks.
//
// Same types, member names, member order and inline helpers as the
// reference's Geometry.hpp, Vector.hpp, MGData.hpp, SparseMatrix.hpp and
// CGData.hpp (one-rank build, HPCG_NO_MPI), so the reference's driver and
// validation code compile unchanged against them; the include guards are the
// reference's, so when this header is included first (e.g. `-include
// hpcg_dropin.hpp`) the reference's own copies of those headers are skipped.
//
// What differs is device bookkeeping behind the reference's
// `optimizationData` slots (SparseMatrix.hpp:64, Vector.hpp:34): the helpers
// that change host data (ZeroVector, ScaleVectorValue, CopyVector,
// FillRandomVector, ReplaceMatrixDiagonal) tell the device mirror, so the
// hooks upload only what changed; DeleteMatrix/DeleteVector free the device
// copies.  Code that writes Vector::values directly after OptimizeProblem
// must call HPCG_VectorHostModified(v) (the reference's driver and tests
// never do).
#ifndef HPCG_DROPIN_HPP
#define HPCG_DROPIN_HPP

#define GEOMETRY_HPP
#define VECTOR_HPP
#define MGDATA_HPP
#define SPARSEMATRIX_HPP
#define CGDATA_HPP

#include <cassert>
#include <cstdlib>
#include <unordered_map>
#include <vector>

// ---------------------------------------------------------------- Geometry.hpp

typedef int local_int_t;
#ifdef HPCG_NO_LONG_LONG
typedef int global_int_t;
#else
typedef long long global_int_t;
#endif

struct Geometry_STRUCT {
  int size;
  int rank;
  int numThreads;
  local_int_t nx;
  local_int_t ny;
  local_int_t nz;
  int npx;
  int npy;
  int npz;
  int pz;
  int npartz;
  int *partz_ids;
  local_int_t *partz_nz;
  int ipx;
  int ipy;
  int ipz;
  global_int_t gnx;
  global_int_t gny;
  global_int_t gnz;
  global_int_t gix0;
  global_int_t giy0;
  global_int_t giz0;
};
typedef struct Geometry_STRUCT Geometry;

// Rank owning a global row (Geometry.hpp:86-119): z partitions of varying nz.
inline int ComputeRankOfMatrixRow(const Geometry &geom, global_int_t index) {
  const global_int_t plane = geom.gny * geom.gnx;
  global_int_t iz = index / plane;
  const global_int_t iy = (index - iz * plane) / geom.gnx;
  const global_int_t ix = index % geom.gnx;
  int ipz = 0, prev = 0;
  for (int i = 0; i < geom.npartz; ++i) {
    const int part_nz = geom.partz_nz[i];
    const int span = geom.partz_ids[i] - prev;
    if (iz <= (global_int_t)part_nz * span) {
      ipz += (int)(iz / part_nz);
      break;
    }
    ipz += span;
    iz -= (global_int_t)part_nz * span;
    prev = span;
  }
  const int ipy = (int)(iy / geom.ny), ipx = (int)(ix / geom.nx);
  return ipx + ipy * geom.npx + ipz * geom.npy * geom.npx;
}

inline void DeleteGeometry(Geometry &geom) {
  delete[] geom.partz_nz;
  delete[] geom.partz_ids;
}

// ---------------------------------------------------------------- Vector.hpp

struct Vector_STRUCT {
  local_int_t localLength;
  double *values;
  void *optimizationData;  // device mirror (struct hpcgx_dropin_vector)
};
typedef struct Vector_STRUCT Vector;

extern "C" {
// Device-mirror notifications (hpcg_amd/dropin/dropin_state.cpp).
void HPCG_VectorHostModified(Vector *v);
void HPCG_VectorHostZeroed(Vector *v);
void HPCG_VectorRelease(Vector *v);
}

inline void InitializeVector(Vector &v, local_int_t localLength) {
  v.localLength = localLength;
  v.values = new double[localLength];
  v.optimizationData = 0;
}

inline void ZeroVector(Vector &v) {
  double *vv = v.values;
  for (local_int_t i = 0; i < v.localLength; ++i) vv[i] = 0.0;
  if (v.optimizationData) HPCG_VectorHostZeroed(&v);
}

inline void ScaleVectorValue(Vector &v, local_int_t index, double value) {
  assert(index >= 0 && index < v.localLength);
  v.values[index] *= value;
  if (v.optimizationData) HPCG_VectorHostModified(&v);
}

// Values in [1, 2) from the C library's rand(), as Vector.hpp:81-86.
inline void FillRandomVector(Vector &v) {
  double *vv = v.values;
  for (local_int_t i = 0; i < v.localLength; ++i) vv[i] = rand() / (double)(RAND_MAX) + 1.0;
  if (v.optimizationData) HPCG_VectorHostModified(&v);
}

inline void CopyVector(const Vector &v, Vector &w) {
  assert(w.localLength >= v.localLength);
  const double *vv = v.values;
  double *wv = w.values;
  for (local_int_t i = 0; i < v.localLength; ++i) wv[i] = vv[i];
  if (w.optimizationData) HPCG_VectorHostModified(&w);
}

inline void DeleteVector(Vector &v) {
  if (v.optimizationData) HPCG_VectorRelease(&v);
  delete[] v.values;
  v.localLength = 0;
}

// ---------------------------------------------------------------- MGData.hpp

struct MGData_STRUCT {
  int numberOfPresmootherSteps;
  int numberOfPostsmootherSteps;
  local_int_t *f2cOperator;
  Vector *rc;
  Vector *xc;
  Vector *Axf;
  void *optimizationData;
};
typedef struct MGData_STRUCT MGData;

inline void InitializeMGData(local_int_t *f2cOperator, Vector *rc, Vector *xc, Vector *Axf, MGData &data) {
  data.numberOfPresmootherSteps = 1;
  data.numberOfPostsmootherSteps = 1;
  data.f2cOperator = f2cOperator;
  data.rc = rc;
  data.xc = xc;
  data.Axf = Axf;
  data.optimizationData = 0;
}

inline void DeleteMGData(MGData &data) {
  delete[] data.f2cOperator;
  DeleteVector(*data.Axf);
  DeleteVector(*data.rc);
  DeleteVector(*data.xc);
  delete data.Axf;
  delete data.rc;
  delete data.xc;
}

// ---------------------------------------------------------------- SparseMatrix.hpp

typedef std::unordered_map<global_int_t, local_int_t> GlobalToLocalMap;

struct SparseMatrix_STRUCT {
  char *title;
  Geometry *geom;
  global_int_t totalNumberOfRows;
  global_int_t totalNumberOfNonzeros;
  local_int_t localNumberOfRows;
  local_int_t localNumberOfColumns;
  local_int_t localNumberOfNonzeros;
  char *nonzerosInRow;
  global_int_t **mtxIndG;
  local_int_t **mtxIndL;
  double **matrixValues;
  double **matrixDiagonal;
  GlobalToLocalMap globalToLocalMap;
  std::vector<global_int_t> localToGlobalMap;
  mutable bool isDotProductOptimized;
  mutable bool isSpmvOptimized;
  mutable bool isMgOptimized;
  mutable bool isWaxpbyOptimized;
  mutable struct SparseMatrix_STRUCT *Ac;
  mutable MGData *mgData;
  void *optimizationData;  // device state (struct hpcgx_dropin_matrix)
};
typedef struct SparseMatrix_STRUCT SparseMatrix;

extern "C" {
void HPCG_MatrixDiagonalReplaced(SparseMatrix *A, const double *diag);
void HPCG_MatrixRelease(SparseMatrix *A);
// Frees the matrix arrays if the drop-in's GenerateProblem allocated them
// (one block per array) and returns 1; returns 0 for any other matrix.
int HPCG_MatrixStorageRelease(SparseMatrix *A);
}

inline void InitializeSparseMatrix(SparseMatrix &A, Geometry *geom) {
  A.title = 0;
  A.geom = geom;
  A.totalNumberOfRows = 0;
  A.totalNumberOfNonzeros = 0;
  A.localNumberOfRows = 0;
  A.localNumberOfColumns = 0;
  A.localNumberOfNonzeros = 0;
  A.nonzerosInRow = 0;
  A.mtxIndG = 0;
  A.mtxIndL = 0;
  A.matrixValues = 0;
  A.matrixDiagonal = 0;
  A.isDotProductOptimized = true;
  A.isSpmvOptimized = true;
  A.isMgOptimized = true;
  A.isWaxpbyOptimized = true;
  A.mgData = 0;
  A.Ac = 0;
  A.optimizationData = 0;
}

inline void CopyMatrixDiagonal(SparseMatrix &A, Vector &diagonal) {
  assert(A.localNumberOfRows == diagonal.localLength);
  for (local_int_t i = 0; i < A.localNumberOfRows; ++i) diagonal.values[i] = *(A.matrixDiagonal[i]);
  if (diagonal.optimizationData) HPCG_VectorHostModified(&diagonal);
}

inline void ReplaceMatrixDiagonal(SparseMatrix &A, Vector &diagonal) {
  assert(A.localNumberOfRows == diagonal.localLength);
  for (local_int_t i = 0; i < A.localNumberOfRows; ++i) *(A.matrixDiagonal[i]) = diagonal.values[i];
  if (A.optimizationData) HPCG_MatrixDiagonalReplaced(&A, diagonal.values);
}

// Frees the matrix as SparseMatrix.hpp:151-183 does.  Its arrays are either
// the drop-in GenerateProblem's blocks (freed by HPCG_MatrixStorageRelease),
// one allocation per array (HPCG_CONTIGUOUS_ARRAYS, GenerateProblem_ref.cpp:116-124)
// or one allocation per row (GenerateProblem_ref.cpp:105-112).
inline void DeleteMatrix(SparseMatrix &A) {
  if (A.optimizationData) HPCG_MatrixRelease(&A);
  if (!HPCG_MatrixStorageRelease(&A) && A.matrixValues) {
#ifdef HPCG_CONTIGUOUS_ARRAYS
    if (A.localNumberOfRows > 0) {
      delete[] A.matrixValues[0];
      delete[] A.mtxIndG[0];
      delete[] A.mtxIndL[0];
    }
#else
    for (local_int_t i = 0; i < A.localNumberOfRows; ++i) {
      delete[] A.matrixValues[i];
      delete[] A.mtxIndG[i];
      delete[] A.mtxIndL[i];
    }
#endif
  }
  if (A.title) delete[] A.title;
  if (A.nonzerosInRow) delete[] A.nonzerosInRow;
  if (A.mtxIndG) delete[] A.mtxIndG;
  if (A.mtxIndL) delete[] A.mtxIndL;
  if (A.matrixValues) delete[] A.matrixValues;
  if (A.matrixDiagonal) delete[] A.matrixDiagonal;
  if (A.geom != 0) {
    DeleteGeometry(*A.geom);
    delete A.geom;
    A.geom = 0;
  }
  if (A.Ac != 0) {
    DeleteMatrix(*A.Ac);
    delete A.Ac;
    A.Ac = 0;
  }
  if (A.mgData != 0) {
    DeleteMGData(*A.mgData);
    delete A.mgData;
    A.mgData = 0;
  }
}

// ---------------------------------------------------------------- CGData.hpp

struct CGData_STRUCT {
  Vector r;
  Vector z;
  Vector p;
  Vector Ap;
};
typedef struct CGData_STRUCT CGData;

inline void InitializeSparseCGData(SparseMatrix &A, CGData &data) {
  InitializeVector(data.r, A.localNumberOfRows);
  InitializeVector(data.z, A.localNumberOfColumns);
  InitializeVector(data.p, A.localNumberOfColumns);
  InitializeVector(data.Ap, A.localNumberOfRows);
}

inline void DeleteCGData(CGData &data) {
  DeleteVector(data.r);
  DeleteVector(data.z);
  DeleteVector(data.p);
  DeleteVector(data.Ap);
}

// ---------------------------------------------------------------- hooks
// The reference's optimisation hooks with their exact signatures.

int CG(const SparseMatrix &A, CGData &data, const Vector &b, Vector &x, const int max_iter,
       const double tolerance, int &niters, double &normr, double &normr0, double *times,
       bool doPreconditioning);                                                   // CG.hpp:22-24
int ComputeSPMV(const SparseMatrix &A, Vector &x, Vector &y);                     // ComputeSPMV.hpp:20
int ComputeSYMGS(const SparseMatrix &A, const Vector &r, Vector &x);              // ComputeSYMGS.hpp:20
int ComputeMG(const SparseMatrix &A, const Vector &r, Vector &x);                 // ComputeMG.hpp:20
int ComputeDotProduct(const local_int_t n, const Vector &x, const Vector &y, double &result,
                      double &time_allreduce, bool &isOptimized);                 // ComputeDotProduct.hpp:18-19
int ComputeWAXPBY(const local_int_t n, const double alpha, const Vector &x, const double beta,
                  const Vector &y, Vector &w, bool &isOptimized);                 // ComputeWAXPBY.hpp:18-19
int OptimizeProblem(SparseMatrix &A, CGData &data, Vector &b, Vector &x, Vector &xexact);  // OptimizeProblem.hpp:22
double OptimizeProblemMemoryUse(const SparseMatrix &A);                          // OptimizeProblem.hpp:28
void ExchangeHalo(const SparseMatrix &A, Vector &x);                              // ExchangeHalo.hpp:19
void GenerateProblem(SparseMatrix &A, Vector *b, Vector *x, Vector *xexact);      // GenerateProblem.hpp:19
void SetupHalo(SparseMatrix &A);                                                  // SetupHalo.hpp:19

#endif  // HPCG_DROPIN_HPP
