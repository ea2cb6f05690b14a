// Device-mirror state shared by the drop-in hooks (see dropin_state.cpp).
#ifndef HPCG_DROPIN_STATE_HPP
#define HPCG_DROPIN_STATE_HPP

#include <cstdint>

#include "hpcg_dropin.hpp"
#include "hpcgx.h"

struct hpcgx_dropin_matrix {
  hpcgx_ctx *ctx;
  int level;
  bool owner;  // level 0 owns the context
};

struct hpcgx_dropin_vector {
  hpcgx_ctx *ctx = nullptr;
  int level = 0;
  double *d = nullptr;       // device copy, colour row order
  uint64_t host_ver = 0;     // bumped by every host-side change
  uint64_t dev_ver = 0;      // host version the device copy holds
  bool zero_pending = false; // last host change was ZeroVector
};

hpcgx_ctx *dropin_context();
void dropin_set_context(hpcgx_ctx *ctx);
hpcgx_dropin_matrix *dropin_matrix(const SparseMatrix &A);
int dropin_level_of_length(hpcgx_ctx *ctx, local_int_t n);
hpcgx_dropin_vector *dropin_mirror(hpcgx_ctx *ctx, int level, const Vector &v);
// makes the device copy current and returns it
int dropin_sync_in(hpcgx_ctx *ctx, int level, const Vector &v, double **dptr);
// copies the device result to the host; both copies then agree
int dropin_sync_out(hpcgx_ctx *ctx, int level, Vector &v);
void dropin_release_all(hpcgx_ctx *ctx);
// the device the hooks use: $HPCG_DEVICE, default 0
int dropin_device();
// waits for the device start-up that GenerateProblem began and returns the
// context it opened (once; null if it did not start or failed)
hpcgx_ctx *dropin_take_started_context();

#endif
