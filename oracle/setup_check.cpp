// setup_check -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// Holds the drop-in's problem-setup hooks (hpcg_amd/dropin/setup_hooks.cpp:
// GenerateProblem, SetupHalo) to the reference's own GenerateProblem_ref and
// SetupHalo_ref, compiled from the reference's sources by oracle/Makefile:
//   1. on every level of the hierarchy (the fine grid and each coarse grid
//      GenerateCoarseProblem.cpp:50-53 makes), both build the level and every
//      array and field must be equal: nonzerosInRow, mtxIndG, mtxIndL,
//      matrixValues, matrixDiagonal (as offsets into the row), the local-to-
//      global map, the counts and b, x, xexact; the global-to-local map too
//      where the build reads it (MPI or HPCG_DETAILED_DEBUG; else the drop-in
//      leaves it empty);
//   2. the reference's driver sequence (main.cpp:131-158) runs on the
//      drop-in hooks: GenerateProblem, SetupHalo, GenerateCoarseProblem three
//      times, then CheckProblem on every level (its asserts are live: this
//      file is built without NDEBUG);
//   3. DeleteMatrix frees both storage forms (run under AddressSanitizer by
//      the Makefile).
// Usage: setup_check nx ny nz  -> prints "setup_check ok ..." or aborts.
#include <cassert>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "hpcgx.h"
#include "CheckProblem.hpp"
#include "GenerateCoarseProblem.hpp"
#include "GenerateGeometry.hpp"
#include "GenerateProblem.hpp"
#include "GenerateProblem_ref.hpp"
#include "SetupHalo.hpp"
#include "SetupHalo_ref.hpp"

// The device-mirror notifications the drop-in header calls; no device here.
extern "C" {
void HPCG_VectorHostModified(Vector *) {}
void HPCG_VectorHostZeroed(Vector *) {}
void HPCG_VectorRelease(Vector *) {}
void HPCG_MatrixDiagonalReplaced(SparseMatrix *, const double *) {}
void HPCG_MatrixRelease(SparseMatrix *) {}
int hpcgx_device_init(int) { return 0; }
int hpcgx_create(int, hpcgx_ctx **out) {
  *out = nullptr;
  return -3;  // no device here: the drop-in then opens none
}
}

static int fails = 0;
#define CHECK(cond, ...)                        \
  do {                                          \
    if (!(cond)) {                              \
      std::fprintf(stderr, "MISMATCH: " __VA_ARGS__); \
      std::fprintf(stderr, "\n");               \
      if (++fails > 20) std::abort();           \
    }                                           \
  } while (0)

static Geometry *make_geometry(local_int_t nx, local_int_t ny, local_int_t nz) {
  Geometry *g = new Geometry;
  GenerateGeometry(1, 0, 1, 0, 0, 0, nx, ny, nz, 1, 1, 1, g);
  return g;
}

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void compare_level(local_int_t nx, local_int_t ny, local_int_t nz, bool vectors) {
  SparseMatrix R, D;
  Vector rb, rx, rxe, db, dx, dxe;
  InitializeSparseMatrix(R, make_geometry(nx, ny, nz));
  InitializeSparseMatrix(D, make_geometry(nx, ny, nz));
  double t0 = now();
  GenerateProblem_ref(R, vectors ? &rb : 0, vectors ? &rx : 0, vectors ? &rxe : 0);
  SetupHalo_ref(R);
  double t1 = now();
  GenerateProblem(D, vectors ? &db : 0, vectors ? &dx : 0, vectors ? &dxe : 0);
  SetupHalo(D);
  double t2 = now();
  const local_int_t n = R.localNumberOfRows;
  CHECK(D.localNumberOfRows == n, "rows");
  CHECK(D.localNumberOfColumns == R.localNumberOfColumns, "columns");
  CHECK(D.localNumberOfNonzeros == R.localNumberOfNonzeros, "local nonzeros");
  CHECK(D.totalNumberOfRows == R.totalNumberOfRows, "total rows");
  CHECK(D.totalNumberOfNonzeros == R.totalNumberOfNonzeros, "total nonzeros");
  CHECK(D.title == R.title, "title");
  CHECK(D.localToGlobalMap == R.localToGlobalMap, "localToGlobalMap");
#if !defined(HPCG_NO_MPI) || defined(HPCG_DETAILED_DEBUG)
  CHECK(D.globalToLocalMap == R.globalToLocalMap, "globalToLocalMap");
#else
  // one rank, no detailed debug: nothing reads the map, and the drop-in
  // leaves it empty (hpcg_amd/dropin/setup_hooks.cpp)
  CHECK(D.globalToLocalMap.empty(), "globalToLocalMap left empty");
#endif
  for (local_int_t i = 0; i < n && fails == 0; ++i) {
    const int k = R.nonzerosInRow[i];
    CHECK(D.nonzerosInRow[i] == k, "nonzerosInRow[%d]", i);
    CHECK(D.matrixDiagonal[i] - D.matrixValues[i] == R.matrixDiagonal[i] - R.matrixValues[i], "diagonal[%d]", i);
    for (int j = 0; j < k; ++j) {
      CHECK(D.matrixValues[i][j] == R.matrixValues[i][j], "value[%d][%d]", i, j);
      CHECK(D.mtxIndG[i][j] == R.mtxIndG[i][j], "mtxIndG[%d][%d]", i, j);
      CHECK(D.mtxIndL[i][j] == R.mtxIndL[i][j], "mtxIndL[%d][%d]", i, j);
    }
  }
  if (vectors) {
    CHECK(db.localLength == rb.localLength && dx.localLength == rx.localLength && dxe.localLength == rxe.localLength,
          "vector lengths");
    CHECK(std::memcmp(db.values, rb.values, n * sizeof(double)) == 0, "b");
    CHECK(std::memcmp(dx.values, rx.values, n * sizeof(double)) == 0, "x");
    CHECK(std::memcmp(dxe.values, rxe.values, n * sizeof(double)) == 0, "xexact");
    DeleteVector(rb); DeleteVector(rx); DeleteVector(rxe);
    DeleteVector(db); DeleteVector(dx); DeleteVector(dxe);
  }
  std::printf("level %dx%dx%d: reference %.3f s, drop-in %.3f s\n", nx, ny, nz, t1 - t0, t2 - t1);
  DeleteMatrix(R);  // one allocation per row
  DeleteMatrix(D);  // the drop-in's blocks
}

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: setup_check nx ny nz\n");
    return 2;
  }
  const local_int_t nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]);
  // 1. level by level against the reference
  for (int l = 0; l < 4; ++l) compare_level(nx >> l, ny >> l, nz >> l, l == 0);
  // 2. the driver's sequence on the drop-in hooks, checked by CheckProblem
  SparseMatrix A;
  InitializeSparseMatrix(A, make_geometry(nx, ny, nz));
  Vector b, x, xexact;
  double t0 = now();
  GenerateProblem(A, &b, &x, &xexact);
  SetupHalo(A);
  SparseMatrix *cur = &A;
  for (int level = 1; level < 4; ++level) {
    GenerateCoarseProblem(*cur);
    cur = cur->Ac;
  }
  const double setup = now() - t0;
  cur = &A;
  Vector *cb = &b, *cx = &x, *cxe = &xexact;
  for (int level = 0; level < 4; ++level) {
    CheckProblem(*cur, cb, cx, cxe);
    cur = cur->Ac;
    cb = cx = cxe = 0;
  }
  DeleteVector(b);
  DeleteVector(x);
  DeleteVector(xexact);
  DeleteMatrix(A);
  if (fails) {
    std::fprintf(stderr, "setup_check: %d mismatches\n", fails);
    return 1;
  }
  std::printf("setup_check ok %dx%dx%d: driver setup on the drop-in hooks %.3f s\n", nx, ny, nz, setup);
  return 0;
}
