// The reference's optimisation hooks (CG, ComputeSPMV, ComputeSYMGS,
// ComputeMG, ComputeDotProduct, ComputeWAXPBY, OptimizeProblem,
// OptimizeProblemMemoryUse, ExchangeHalo) with their exact signatures,
// implemented over the hpcgx C ABI.  A reference build uses this file in
// place of its own CG.cpp, ComputeSPMV.cpp, ComputeSYMGS.cpp, ComputeMG.cpp,
// ComputeDotProduct.cpp, ComputeWAXPBY.cpp, OptimizeProblem.cpp and
// ExchangeHalo.cpp, and links libhpcgx.so.  Every hook returns 0 on success
// and a nonzero code on failure (the reference's convention); none falls
// back to a CPU path.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "dropin_state.hpp"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int fail(hpcgx_ctx *ctx, const char *hook, int rc) {
  std::fprintf(stderr, "hpcg dropin: %s failed (%d): %s\n", hook, rc,
               ctx ? hpcgx_error_string(ctx) : "no device context (OptimizeProblem not called)");
  return rc ? rc : 1;
}

// Context and level of a matrix prepared by OptimizeProblem.
bool device_of(const SparseMatrix &A, hpcgx_ctx **ctx, int *level) {
  hpcgx_dropin_matrix *dm = dropin_matrix(A);
  if (!dm || !dm->ctx) return false;
  *ctx = dm->ctx;
  *level = dm->level;
  return true;
}

// Level-0 vectors of the dot / WAXPBY hooks, which take no matrix.
bool device_of_length(local_int_t n, hpcgx_ctx **ctx, int *level) {
  *ctx = dropin_context();
  if (!*ctx) return false;
  *level = dropin_level_of_length(*ctx, n);
  return *level >= 0;
}

}  // namespace

// OptimizeProblem.cpp:37: colour, reorder and pack every level on the device.
// With HPCG_DROPIN_TIMING set, the time of each phase goes to stderr.
int OptimizeProblem(SparseMatrix &A, CGData &data, Vector &b, Vector &x, Vector &xexact) {
  (void)data;
  const bool timing = std::getenv("HPCG_DROPIN_TIMING") != nullptr;
  double t = now_s();
  auto phase = [&](const char *what) {
    if (!timing) return;
    const double t1 = now_s();
    std::fprintf(stderr, "hpcg dropin: OptimizeProblem %s %.4f s\n", what, t1 - t);
    t = t1;
  };
  hpcgx_ctx *ctx = dropin_take_started_context();
  phase("device start-up (rest of)");
  int rc = 0;
  if (!ctx && (rc = hpcgx_create(dropin_device(), &ctx))) return fail(nullptr, "OptimizeProblem", rc);
  phase("context");
  int level = 0;
  for (SparseMatrix *m = &A; m; m = m->Ac, ++level) {
    const local_int_t *f2c = m->mgData ? m->mgData->f2cOperator : nullptr;
    rc = hpcgx_level_upload(ctx, level, m->geom->nx, m->geom->ny, m->geom->nz, m->nonzerosInRow,
                            m->mtxIndL, m->matrixValues, f2c);
    if (rc) return fail(ctx, "OptimizeProblem (level upload)", rc);
    m->optimizationData = new hpcgx_dropin_matrix{ctx, level, level == 0};
    phase(level == 0 ? "level 0 upload" : "coarse level upload");
  }
  if ((rc = hpcgx_prepare_precision(ctx, HPCGX_PREC_FP64))) return fail(ctx, "OptimizeProblem", rc);
  phase("solver buffers and column forms");
  dropin_set_context(ctx);
  double *d;
  if ((rc = dropin_sync_in(ctx, 0, b, &d)) || (rc = dropin_sync_in(ctx, 0, x, &d)) ||
      (rc = dropin_sync_in(ctx, 0, xexact, &d)))
    return fail(ctx, "OptimizeProblem (vectors)", rc);
  phase("vectors");
  return 0;
}

// OptimizeProblem.cpp:103-107: bytes held on the device.
double OptimizeProblemMemoryUse(const SparseMatrix &A) {
  hpcgx_dropin_matrix *dm = dropin_matrix(A);
  return dm && dm->ctx ? hpcgx_memory_bytes(dm->ctx) : 0.0;
}

// One rank, one GPU: there is no halo (SetupHalo_ref.cpp:58-65 makes ncol == nrow).
void ExchangeHalo(const SparseMatrix &A, Vector &x) {
  (void)A;
  (void)x;
}

// ComputeSPMV.cpp:40-45: y = A x.
int ComputeSPMV(const SparseMatrix &A, Vector &x, Vector &y) {
  hpcgx_ctx *ctx;
  int level, rc;
  if (!device_of(A, &ctx, &level)) return fail(nullptr, "ComputeSPMV", 0);
  double *dx;
  if ((rc = dropin_sync_in(ctx, level, x, &dx))) return fail(ctx, "ComputeSPMV", rc);
  hpcgx_dropin_vector *my = dropin_mirror(ctx, level, y);  // y is output only
  if (!my) return fail(ctx, "ComputeSPMV", HPCGX_ENOMEM);
  if ((rc = hpcgx_spmv(ctx, level, dx, my->d)) || (rc = dropin_sync_out(ctx, level, y)))
    return fail(ctx, "ComputeSPMV", rc);
  return 0;
}

// ComputeSYMGS.cpp:50-55: one symmetric Gauss-Seidel step.
int ComputeSYMGS(const SparseMatrix &A, const Vector &r, Vector &x) {
  hpcgx_ctx *ctx;
  int level, rc;
  if (!device_of(A, &ctx, &level)) return fail(nullptr, "ComputeSYMGS", 0);
  double *dr, *dx;
  if ((rc = dropin_sync_in(ctx, level, r, &dr)) || (rc = dropin_sync_in(ctx, level, x, &dx)) ||
      (rc = hpcgx_symgs(ctx, level, dr, dx)) || (rc = dropin_sync_out(ctx, level, x)))
    return fail(ctx, "ComputeSYMGS", rc);
  return 0;
}

// ComputeMG.cpp:33-38: the V-cycle.
int ComputeMG(const SparseMatrix &A, const Vector &r, Vector &x) {
  hpcgx_ctx *ctx;
  int level, rc;
  if (!device_of(A, &ctx, &level)) return fail(nullptr, "ComputeMG", 0);
  double *dr, *dx;
  if ((rc = dropin_sync_in(ctx, level, r, &dr))) return fail(ctx, "ComputeMG", rc);
  hpcgx_dropin_vector *mx = dropin_mirror(ctx, level, x);
  if (!mx) return fail(ctx, "ComputeMG", HPCGX_ENOMEM);
  dx = mx->d;  // x is output only (ComputeMG_ref.cpp:42 zeroes it)
  if ((rc = hpcgx_mg(ctx, level, dr, dx)) || (rc = dropin_sync_out(ctx, level, x)))
    return fail(ctx, "ComputeMG", rc);
  return 0;
}

// ComputeDotProduct.cpp:41-47.
int ComputeDotProduct(const local_int_t n, const Vector &x, const Vector &y, double &result,
                      double &time_allreduce, bool &isOptimized) {
  (void)isOptimized;
  hpcgx_ctx *ctx;
  int level, rc;
  if (!device_of_length(n, &ctx, &level) || level != 0) return fail(dropin_context(), "ComputeDotProduct", 0);
  double *dx, *dy;
  if ((rc = dropin_sync_in(ctx, level, x, &dx)) || (rc = dropin_sync_in(ctx, level, y, &dy)) ||
      (rc = hpcgx_dot(ctx, n, dx, dy, &result)))
    return fail(ctx, "ComputeDotProduct", rc);
  time_allreduce += 0.0;
  return 0;
}

// ComputeWAXPBY.cpp:42-48: w = alpha x + beta y (w may alias x or y).
int ComputeWAXPBY(const local_int_t n, const double alpha, const Vector &x, const double beta,
                  const Vector &y, Vector &w, bool &isOptimized) {
  (void)isOptimized;
  hpcgx_ctx *ctx;
  int level, rc;
  if (!device_of_length(n, &ctx, &level) || level != 0) return fail(dropin_context(), "ComputeWAXPBY", 0);
  double *dx, *dy, *dw;
  if ((rc = dropin_sync_in(ctx, level, x, &dx)) || (rc = dropin_sync_in(ctx, level, y, &dy)))
    return fail(ctx, "ComputeWAXPBY", rc);
  hpcgx_dropin_vector *mw = dropin_mirror(ctx, level, w);
  if (!mw) return fail(ctx, "ComputeWAXPBY", HPCGX_ENOMEM);
  dw = mw->d;
  if ((rc = hpcgx_waxpby(ctx, n, alpha, dx, beta, dy, dw)) || (rc = dropin_sync_out(ctx, level, w)))
    return fail(ctx, "ComputeWAXPBY", rc);
  return 0;
}

// CG.cpp:59-143: the whole solve runs on the device; b and x move once per
// call (only when the host copies changed), the timing slots are filled as
// the reference fills them.
int CG(const SparseMatrix &A, CGData &data, const Vector &b, Vector &x, const int max_iter,
       const double tolerance, int &niters, double &normr, double &normr0, double *times,
       bool doPreconditioning) {
  (void)data;
  const double t_begin = now_s();
  hpcgx_ctx *ctx;
  int level, rc;
  if (!device_of(A, &ctx, &level) || level != 0) return fail(nullptr, "CG", 0);
  double *db, *dx;
  if ((rc = dropin_sync_in(ctx, 0, b, &db)) || (rc = dropin_sync_in(ctx, 0, x, &dx)))
    return fail(ctx, "CG", rc);
  double t[10] = {0};
  niters = 0;
  rc = hpcgx_cg(ctx, HPCGX_CG_FP64, db, dx, max_iter, tolerance, doPreconditioning ? 1 : 0, &niters, &normr,
                &normr0, t, nullptr);
  if (rc || (rc = dropin_sync_out(ctx, 0, x))) return fail(ctx, "CG", rc);
  for (int i = 1; i < 6; ++i) times[i] += t[i];
  times[0] += now_s() - t_begin;
  return 0;
}
