// Device mirrors behind the reference's optimizationData slots.
//
// A matrix level carries {context, level}; a vector carries its device copy
// and two version counters: host_ver counts host-side changes (the helpers in
// hpcg_dropin.hpp bump it), dev_ver is the host version the device copy
// holds.  A hook uploads an input only when the two differ, and after writing
// an output it downloads it, so host and device agree when it returns.
#include "dropin_state.hpp"

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <unordered_set>

namespace {
std::mutex g_mu;
hpcgx_ctx *g_ctx = nullptr;                            // context made by OptimizeProblem
std::unordered_set<hpcgx_dropin_vector *> g_mirrors;   // live mirrors of g_ctx
}  // namespace

hpcgx_ctx *dropin_context() { return g_ctx; }

void dropin_set_context(hpcgx_ctx *ctx) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_ctx = ctx;
}

hpcgx_dropin_matrix *dropin_matrix(const SparseMatrix &A) {
  return static_cast<hpcgx_dropin_matrix *>(A.optimizationData);
}

int dropin_level_of_length(hpcgx_ctx *ctx, local_int_t n) {
  int64_t info[8];
  for (int l = 0; hpcgx_level_info(ctx, l, info) == HPCGX_OK; ++l)
    if (info[0] == n) return l;
  return -1;
}

hpcgx_dropin_vector *dropin_mirror(hpcgx_ctx *ctx, int level, const Vector &cv) {
  Vector &v = const_cast<Vector &>(cv);
  hpcgx_dropin_vector *m = static_cast<hpcgx_dropin_vector *>(v.optimizationData);
  if (m && (m->ctx != ctx || m->level != level)) {
    HPCG_VectorRelease(&v);
    m = nullptr;
  }
  if (!m) {
    m = new hpcgx_dropin_vector();
    m->ctx = ctx;
    m->level = level;
    if (hpcgx_vec_alloc(ctx, level, &m->d) != HPCGX_OK) {
      delete m;
      return nullptr;
    }
    m->host_ver = 1;
    m->dev_ver = 0;
    v.optimizationData = m;
    std::lock_guard<std::mutex> lk(g_mu);
    g_mirrors.insert(m);
  }
  return m;
}

int dropin_sync_in(hpcgx_ctx *ctx, int level, const Vector &v, double **dptr) {
  hpcgx_dropin_vector *m = dropin_mirror(ctx, level, v);
  if (!m) return HPCGX_ENOMEM;
  int rc = HPCGX_OK;
  if (m->dev_ver != m->host_ver) {
    if (m->zero_pending)
      rc = hpcgx_vec_fill(ctx, level, m->d, 0.0);
    else
      rc = hpcgx_vec_upload(ctx, level, m->d, v.values, v.localLength);
    if (rc) return rc;
    m->dev_ver = m->host_ver;
    m->zero_pending = false;
  }
  *dptr = m->d;
  return HPCGX_OK;
}

int dropin_sync_out(hpcgx_ctx *ctx, int level, Vector &v) {
  hpcgx_dropin_vector *m = static_cast<hpcgx_dropin_vector *>(v.optimizationData);
  if (!m) return HPCGX_ESTATE;
  const int rc = hpcgx_vec_download(ctx, level, m->d, v.values, v.localLength);
  if (rc) return rc;
  m->host_ver += 1;
  m->dev_ver = m->host_ver;
  m->zero_pending = false;
  return HPCGX_OK;
}

void dropin_release_all(hpcgx_ctx *ctx) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (hpcgx_dropin_vector *m : g_mirrors)
    if (m->ctx == ctx) {  // device memory goes with the context
      m->ctx = nullptr;
      m->d = nullptr;
    }
  if (g_ctx == ctx) g_ctx = nullptr;
}

extern "C" {

void HPCG_VectorHostModified(Vector *v) {
  hpcgx_dropin_vector *m = static_cast<hpcgx_dropin_vector *>(v->optimizationData);
  if (!m) return;
  m->host_ver += 1;
  m->zero_pending = false;
}

void HPCG_VectorHostZeroed(Vector *v) {
  hpcgx_dropin_vector *m = static_cast<hpcgx_dropin_vector *>(v->optimizationData);
  if (!m) return;
  m->host_ver += 1;
  m->zero_pending = true;
}

void HPCG_VectorRelease(Vector *v) {
  hpcgx_dropin_vector *m = static_cast<hpcgx_dropin_vector *>(v->optimizationData);
  if (!m) return;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_mirrors.erase(m);
  }
  if (m->ctx && m->d) hpcgx_vec_free(m->ctx, m->d);
  delete m;
  v->optimizationData = 0;
}

void HPCG_MatrixDiagonalReplaced(SparseMatrix *A, const double *diag) {
  hpcgx_dropin_matrix *dm = dropin_matrix(*A);
  if (!dm || !dm->ctx) return;
  if (hpcgx_level_set_diagonal(dm->ctx, dm->level, diag) != HPCGX_OK) {
    std::fprintf(stderr, "hpcg dropin: diagonal update failed: %s\n", hpcgx_error_string(dm->ctx));
    std::abort();
  }
}

void HPCG_MatrixRelease(SparseMatrix *A) {
  hpcgx_dropin_matrix *dm = dropin_matrix(*A);
  if (!dm) return;
  if (dm->owner && dm->ctx) {
    dropin_release_all(dm->ctx);
    hpcgx_destroy(dm->ctx);
    // coarse levels still point at the context; clear them before they are deleted
    for (SparseMatrix *m = A->Ac; m; m = m->Ac) {
      hpcgx_dropin_matrix *cm = dropin_matrix(*m);
      if (cm) cm->ctx = nullptr;
    }
  }
  delete dm;
  A->optimizationData = 0;
}

}  // extern "C"
