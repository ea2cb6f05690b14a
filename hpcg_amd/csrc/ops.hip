// Kernel drivers: SpMV, coloured symmetric Gauss-Seidel, fused residual +
// restriction, prolongation, the V-cycle in fp64 / fp32 / scaled fp16, and
// the vector reductions.
#include <initializer_list>
#include <mutex>

#include "kernels.hpp"

namespace hpcgx {

static inline unsigned tiles(int64_t rows) { return (unsigned)((rows + kBlock - 1) / kBlock); }

template <class VT>
static const VT *vals_of(const Level &L);
template <>
const double *vals_of<double>(const Level &L) { return L.vals; }
template <>
const float *vals_of<float>(const Level &L) { return L.vals32; }
template <>
const __half *vals_of<__half>(const Level &L) { return L.vals16; }

// rz (fp64 stencil levels only): the step also leaves the partials of r.x
// over its rows there, tiles(rows) of them.
template <class VT, class XT, class ACC, int ZK, bool FINE, bool SAFE = false>
static void gs_block(Ctx *c, const Level &L, const VT *v, const XT *r, XT *x, int k, double *rz = nullptr) {
  const int64_t r0 = L.lay.start[k], nr = L.lay.start[k + 1] - r0;
  if (nr <= 0) return;
  const ColIdx ci = col_idx(c, L);
  // Coarse fp64 steps of blocks under HPCGX_ONE_WAVE_ROWS (2^20) rows launch
  // as one-wave blocks, which spreads them over more CUs: with explicit
  // columns at two waves per SIMD a 192^3 set measured 2-6 % faster, 512^3
  // about 5 %; with compact columns and no occupancy limit, 1245 against 1219
  // GFLOP/s at 192^3 (512^3 within its run-to-run spread); with paired values,
  // 94.6-95.4 ms per 192^3 set against 95.5-95.6 with the threshold at 2^16
  // (the 96^3 level's steps in 256-thread blocks) and 96.1-96.8 with none.
  // The fine level, the other precisions and the r.z form keep 256-thread
  // blocks.
#ifndef HPCGX_ONE_WAVE_ROWS
#define HPCGX_ONE_WAVE_ROWS (1 << 20)
#endif
  const unsigned bs = (!FINE && sizeof(VT) == 8 && nr < (int64_t)(HPCGX_ONE_WAVE_ROWS)) ? kSlice : kBlock;
  const unsigned g = (unsigned)((nr + bs - 1) / bs);
  hipStream_t st = c->stream;
  if constexpr (ZK < 0 && sizeof(VT) == 8 && sizeof(XT) == 8) {
    if (rz) {
      // 256-thread blocks: the partials are per block of 4 waves, tiles(nr) of them
      if (ci.cbase)
        k_gs<VT, XT, ACC, -1, FINE, true, false, true><<<tiles(nr), kBlock, 0, st>>>(ci, v, r, x, r0, nr, nullptr, rz);
      else
        k_gs<VT, XT, ACC, -1, FINE, false, false, true><<<tiles(nr), kBlock, 0, st>>>(ci, v, r, x, r0, nr, nullptr, rz);
      return;
    }
  }
  if constexpr (ZK <= 0) {  // general levels: full rows, or the first block from x = 0
    if (L.dslot) {
      if (ci.cbase)
        k_gs<VT, XT, ACC, ZK, FINE, true, true><<<g, bs, 0, st>>>(ci, v, r, x, r0, nr, L.dslot);
      else
        k_gs<VT, XT, ACC, ZK, FINE, false, true><<<g, bs, 0, st>>>(ci, v, r, x, r0, nr, L.dslot);
      return;
    }
  }
  if (ci.cbase)
    k_gs<VT, XT, ACC, ZK, FINE, true, false, false, SAFE><<<g, bs, 0, st>>>(ci, v, r, x, r0, nr, nullptr);
  else
    k_gs<VT, XT, ACC, ZK, FINE, false, false, false, SAFE><<<g, bs, 0, st>>>(ci, v, r, x, r0, nr, nullptr);
}

// The zero-start forward sweep of the default order: block k reads only the
// blocks swept before it.
template <class VT, class XT, class ACC, bool FINE, bool SAFE>
static void zero_start_forward(Ctx *c, const Level &L, const VT *v, const XT *r, XT *x) {
  gs_block<VT, XT, ACC, 0, FINE, SAFE>(c, L, v, r, x, 0);
  gs_block<VT, XT, ACC, 1, FINE, SAFE>(c, L, v, r, x, 1);
  gs_block<VT, XT, ACC, 2, FINE, SAFE>(c, L, v, r, x, 2);
  gs_block<VT, XT, ACC, 3, FINE, SAFE>(c, L, v, r, x, 3);
  gs_block<VT, XT, ACC, 4, FINE, SAFE>(c, L, v, r, x, 4);
  gs_block<VT, XT, ACC, 5, FINE, SAFE>(c, L, v, r, x, 5);
  gs_block<VT, XT, ACC, 6, FINE, SAFE>(c, L, v, r, x, 6);
  gs_block<VT, XT, ACC, 7, FINE, SAFE>(c, L, v, r, x, 7);
}

template <class XT>
__global__ __launch_bounds__(kBlock) void k_zero(XT *__restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = (XT)0;
}

static bool default_order(const Level &L) {
  if (L.lay.general) return false;
  for (int k = 0; k < kColours; ++k)
    if (L.lay.cls[k] != kDefaultOrder[k]) return false;
  return true;
}

// One symmetric Gauss-Seidel step over the colour blocks.  Forward sweep
// blocks 0..ncol-1, backward sweep ncol-2..0: the backward step of the last
// block would recompute it from exactly the neighbour values its forward step
// used, so it is elided.  zero_start: x is known to be zero on entry (the
// V-cycle's first sweep); in the default stencil order x is then not read
// before it is written; otherwise (general levels, other orders) it is
// zero-filled first, and block 0, whose neighbours are all still zero, reads
// only its diagonal (x_i = r_i / a_ii, what the full step gives bit for bit).
// rz: the steps that leave their rows final (forward step of the last block,
// then the backward steps) write r.x partials at rz, *nrz of them in all.
// untrusted_x: x may hold anything on a zero start (a caller's vector, or a
// coarse level's vector after a V-cycle on a non-finite input); the solvers
// clear their own level-0 buffer when a solve begins and skip the guard.
template <class VT, class XT, class ACC, bool FINE>
static void sweep(Ctx *c, const Level &L, const XT *r, XT *x, bool zero_start, double *rz = nullptr,
                  int64_t *nrz = nullptr, bool untrusted_x = false) {
  const VT *v = vals_of<VT>(L);
  if (zero_start && !default_order(L)) {
    k_zero<XT><<<tiles(L.lay.npad), kBlock, 0, c->stream>>>(x, L.lay.npad);
    gs_block<VT, XT, ACC, 0, FINE>(c, L, v, r, x, 0);
    for (int k = 1; k < L.lay.ncol; ++k) gs_block<VT, XT, ACC, -1, FINE>(c, L, v, r, x, k);
  } else if (zero_start) {
    if (untrusted_x) zero_start_forward<VT, XT, ACC, FINE, true>(c, L, v, r, x);
    else zero_start_forward<VT, XT, ACC, FINE, false>(c, L, v, r, x);
  } else {
    for (int k = 0; k + 1 < L.lay.ncol; ++k) gs_block<VT, XT, ACC, -1, FINE>(c, L, v, r, x, k);
    const int last = L.lay.ncol - 1;
    gs_block<VT, XT, ACC, -1, FINE>(c, L, v, r, x, last, rz);
    if (rz) rz += tiles(L.lay.start[last + 1] - L.lay.start[last]);
  }
  for (int k = L.lay.ncol - 2; k >= 0; --k) {
    gs_block<VT, XT, ACC, -1, FINE>(c, L, v, r, x, k, rz);
    if (rz) rz += tiles(L.lay.start[k + 1] - L.lay.start[k]);
  }
  if (nrz) {
    *nrz = 0;
    for (int k = 0; k < L.lay.ncol; ++k) *nrz += tiles(L.lay.start[k + 1] - L.lay.start[k]);
  }
}

template <class VT, class XT, class ACC>
static int symgs_launch(Ctx *c, int l, const XT *r, XT *x, bool zero_start, double *rz = nullptr,
                        int64_t *nrz = nullptr, bool untrusted_x = false) {
  const Level &L = c->levels[l];
  if (l == 0) sweep<VT, XT, ACC, true>(c, L, r, x, zero_start, rz, nrz, untrusted_x);
  else sweep<VT, XT, ACC, false>(c, L, r, x, zero_start, rz, nrz, untrusted_x);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

// V-cycle of ComputeMG_ref.cpp:39-63: x = 0; pre-smooth; residual + restrict;
// recurse; prolong; post-smooth.  The coarsest level gets one smoothing step.
// rz / nrz: r.x partials of the result from the post-smoother (see sweep);
// fp64 stencil levels with a coarse level only.  untrusted_x: x at level l
// may hold anything (the caller's vector); the coarse levels always take
// this form.
template <class VT, class XT, class ACC, bool SCALE>
static int vcycle(Ctx *c, int l, const XT *r, XT *x, double *rz = nullptr, int64_t *nrz = nullptr,
                  bool untrusted_x = false) {
  const Level &L = c->levels[l];
  int rc = symgs_launch<VT, XT, ACC>(c, l, r, x, true, nullptr, nullptr, untrusted_x);
  if (rc || l + 1 >= (int)c->levels.size()) return rc;
  Level &C = c->levels[l + 1];
  XT *rc_v, *xc_v;
  if constexpr (sizeof(XT) == 8) {
    rc_v = reinterpret_cast<XT *>(C.r);
    xc_v = reinterpret_cast<XT *>(C.x);
  } else {
    rc_v = reinterpret_cast<XT *>(C.r32);
    xc_v = reinterpret_cast<XT *>(C.x32);
  }
  const int64_t nc = L.ninj, r0 = L.lay.start[L.lay.inj];
  const ColIdx ci = col_idx(c, L);
  if (ci.cbase)
    k_restrict<VT, XT, ACC, SCALE, true><<<tiles(nc), kBlock, 0, c->stream>>>(ci, vals_of<VT>(L), r, x, rc_v,
                                                                             L.inj, L.sigma, r0, nc, L.frow);
  else
    k_restrict<VT, XT, ACC, SCALE, false><<<tiles(nc), kBlock, 0, c->stream>>>(ci, vals_of<VT>(L), r, x, rc_v,
                                                                              L.inj, L.sigma, r0, nc, L.frow);
  HPCGX_CHECK_LAUNCH(c);
  // the coarse level's x may hold anything: a V-cycle on a non-finite input
  // leaves non-finite values in the context's coarse vectors
  if ((rc = vcycle<VT, XT, ACC, SCALE>(c, l + 1, rc_v, xc_v, nullptr, nullptr, true))) return rc;
  k_prolong<XT, ACC, SCALE><<<tiles(nc), kBlock, 0, c->stream>>>(x, xc_v, L.inj, L.sigma, r0, nc, L.frow);
  HPCGX_CHECK_LAUNCH(c);
  return symgs_launch<VT, XT, ACC>(c, l, r, x, false, rz, nrz);
}

// vector kernels

__global__ __launch_bounds__(kBlock) void k_waxpby(int64_t n, double alpha,
                                                   const double *__restrict__ x, double beta,
                                                   const double *__restrict__ y,
                                                   double *__restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  w[i] = waxpby1(alpha, x[i], beta, y[i]);
}

// out = x . y (ComputeDotProduct_ref.cpp:54-64, fixed-order tree)
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const double *__restrict__ x,
                                                const double *__restrict__ y, double *partials) {
  double a[kChunk], b[kChunk];
#pragma unroll
  for (int j = 0; j < kChunk; ++j) {
    const int64_t i = chunk_index(j);
    a[j] = i < n ? x[i] : 0.0;
    b[j] = i < n ? y[i] : 0.0;
  }
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kChunk; ++j) acc += a[j] * b[j];
  const double v[1] = {acc};
  block_partials<1>(v, partials);
}

// w = alpha x + beta y ; out = w . w   (WAXPBY then DOT of CG.cpp:88-89 / 122-123)
__global__ __launch_bounds__(kBlock) void k_waxpby_dot(int64_t n, double alpha,
                                                       const double *__restrict__ x, double beta,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ w, double *partials) {
  double a[kChunk], b[kChunk];
#pragma unroll
  for (int j = 0; j < kChunk; ++j) {
    const int64_t i = chunk_index(j);
    a[j] = i < n ? x[i] : 0.0;
    b[j] = i < n ? y[i] : 0.0;
  }
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kChunk; ++j) {
    const int64_t i = chunk_index(j);
    const double v = waxpby1(alpha, a[j], beta, b[j]);
    if (i < n) {
      w[i] = v;
      acc += v * v;
    }
  }
  const double v[1] = {acc};
  block_partials<1>(v, partials);
}


int mg_fp64(Ctx *c, int level, const double *r, double *x, double *rz, int64_t *nrz, bool untrusted_x) {
  if (rz && (level + 1 >= (int)c->levels.size() || c->levels[level].lay.general))
    return set_err(c, HPCGX_EINVAL, "fused r.z needs a stencil level with a coarse level");
  return vcycle<double, double, double, false>(c, level, r, x, rz, nrz, untrusted_x);
}

int mg_fp32(Ctx *c, int level, const float *r, float *x, bool untrusted_x) {
  return vcycle<float, float, float, false>(c, level, r, x, nullptr, nullptr, untrusted_x);
}

int zero_vec(Ctx *c, double *x, int64_t n) {
  k_zero<double><<<tiles(n), kBlock, 0, c->stream>>>(x, n);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int zero_vec(Ctx *c, float *x, int64_t n) {
  k_zero<float><<<tiles(n), kBlock, 0, c->stream>>>(x, n);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

__global__ __launch_bounds__(kBlock) void k_scale(const float *__restrict__ a, const float *__restrict__ s,
                                                  float *__restrict__ o, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) o[i] = a[i] * s[i];
}

// The scaled fp16 V-cycle proper: xhat from rhat = dinv r (a solver that
// applies the scalings itself).
int mg_fp16_scaled(Ctx *c, const float *rhat, float *xhat) {
  return vcycle<__half, float, float, true>(c, 0, rhat, xhat);
}

// Scaled fp16 V-cycle on level 0: rhat = dinv r, V-cycle with D^-1/2 A D^-1/2,
// x = dinv xhat.
int mg_fp16(Ctx *c, const float *r, float *x, bool untrusted_x) {
  Level &L = c->levels[0];
  const int64_t np = L.lay.npad;
  k_scale<<<tiles(np), kBlock, 0, c->stream>>>(r, L.dinv, L.r16, np);
  HPCGX_CHECK_LAUNCH(c);
  int rc = vcycle<__half, float, float, true>(c, 0, L.r16, x, nullptr, nullptr, untrusted_x);
  if (rc) return rc;
  k_scale<<<tiles(np), kBlock, 0, c->stream>>>(x, L.dinv, x, np);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

__global__ __launch_bounds__(kBlock) void k_to_single(const double *__restrict__ a, float *__restrict__ o,
                                                      int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) o[i] = (float)a[i];
}

__global__ __launch_bounds__(kBlock) void k_to_double(const float *__restrict__ a, double *__restrict__ o,
                                                      int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) o[i] = (double)a[i];
}

int spmv_fp64(Ctx *c, int level, const double *x, double *y) {
  const Level &L = c->levels[level];
  const ColIdx ci = col_idx(c, L);
  if (ci.cbase)
    k_spmv<double, double, double, true><<<tiles(L.lay.npad), kBlock, 0, c->stream>>>(ci, L.vals, x, y, L.lay.npad);
  else
    k_spmv<double, double, double, false><<<tiles(L.lay.npad), kBlock, 0, c->stream>>>(ci, L.vals, x, y, L.lay.npad);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int symgs_fp64(Ctx *c, int level, const double *r, double *x) {
  return symgs_launch<double, double, double>(c, level, r, x, false);
}

// ---------------------------------------------------------------------------
// Static scratch of hpcgx_dot_noalloc, the dot product for callers that may
// not allocate device memory.  It holds the partial sums of the producer
// blocks of a level of up to kNoAllocRows padded rows (one per 1024 rows),
// their folded sums and the finishing kernel's scalars: 2.1 MB, which each
// device keeps until the process ends.  Every call on one device shares it.
// Each call overwrites all it reads (every producer block writes its
// partial, the fold every folded sum, the finishing kernel the result), so
// nothing needs clearing.
constexpr int64_t kNoAllocRows = (int64_t)1 << 28;
constexpr int64_t kNaBlocks = kNoAllocRows / (kChunk * kBlock);
constexpr int64_t kNaFolded = kNaBlocks;                                    // offset of the folded sums
constexpr int64_t kNaScal = kNaFolded + (kNaBlocks / kFold + 1 + 31) / 32 * 32;  // offset of the scalars
constexpr int64_t kNaDoubles = kNaScal + S_COUNT;
static_assert(chunk_blocks(kNoAllocRows) == kNaBlocks, "one partial per producer block");
static __device__ __align__(256) double scratch[kNaDoubles];

// The device address of `scratch` on each device (the array's name is only a
// host-side handle), looked up once per device by noalloc_scratch_init.
constexpr int kNaDevices = 64;
static double *g_scratch[kNaDevices] = {};
static std::mutex g_scratch_mu;

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNaDevices) return HPCGX_OK;  // the entry point refuses such devices
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return HPCGX_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HPCGX_EHIP;
  g_scratch[device] = static_cast<double *>(p);
  return HPCGX_OK;
}

static double *noalloc_scratch(int device) {
  if (device < 0 || device >= kNaDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[device];
}

}  // namespace hpcgx

using namespace hpcgx;

static int check_level(hpcgx_ctx *c, int level, const char *who) {
  if (!c) return HPCGX_EINVAL;
  if (level < 0 || level >= (int)c->levels.size())
    return set_err(c, HPCGX_EINVAL, std::string(who) + ": no such level");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, HPCGX_EHIP, "hipSetDevice failed");
  return ensure_ready(c);
}

// Null pointers and vectors of the wrong size (those made by hpcgx_vec_alloc)
// are refused before any launch.
static int check_vecs(hpcgx_ctx *c, int64_t elems, const char *who, std::initializer_list<const void *> vs) {
  for (const void *v : vs) {
    if (!v) return set_err(c, HPCGX_EINVAL, std::string(who) + ": null vector");
    int rc = check_vec(c, v, elems, who);
    if (rc) return rc;
  }
  return HPCGX_OK;
}

static int64_t npad_of(hpcgx_ctx *c, int level) { return c->levels[level].lay.npad; }

extern "C" {

int hpcgx_spmv(hpcgx_ctx *c, int level, const double *x, double *y) {
  int rc = check_level(c, level, "spmv");
  if (rc) return rc;
  if ((rc = check_vecs(c, npad_of(c, level), "spmv", {x, y}))) return rc;
  if (x == y) return set_err(c, HPCGX_EINVAL, "spmv: x and y must be distinct device vectors");
  return spmv_fp64(c, level, x, y);
}

int hpcgx_symgs(hpcgx_ctx *c, int level, const double *r, double *x) {
  int rc = check_level(c, level, "symgs");
  if (rc) return rc;
  if ((rc = check_vecs(c, npad_of(c, level), "symgs", {r, x}))) return rc;
  return symgs_fp64(c, level, r, x);
}

int hpcgx_restrict(hpcgx_ctx *c, int level, const double *r, const double *x, double *rc_out) {
  int rc = check_level(c, level, "restrict");
  if (rc) return rc;
  if (level + 1 >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "restrict: level has no coarse level");
  if ((rc = check_vecs(c, npad_of(c, level), "restrict", {r, x})) ||
      (rc = check_vecs(c, npad_of(c, level + 1), "restrict", {rc_out})))
    return rc;
  const Level &L = c->levels[level];
  const int64_t nc = L.ninj;
  const ColIdx ci = col_idx(c, L);
  if (ci.cbase)
    k_restrict<double, double, double, false, true><<<tiles(nc), kBlock, 0, c->stream>>>(
        ci, L.vals, r, x, rc_out, L.inj, nullptr, L.lay.start[L.lay.inj], nc, L.frow);
  else
    k_restrict<double, double, double, false, false><<<tiles(nc), kBlock, 0, c->stream>>>(
        ci, L.vals, r, x, rc_out, L.inj, nullptr, L.lay.start[L.lay.inj], nc, L.frow);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int hpcgx_prolong(hpcgx_ctx *c, int level, const double *xc, double *x) {
  int rc = check_level(c, level, "prolong");
  if (rc) return rc;
  if (level + 1 >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "prolong: level has no coarse level");
  if ((rc = check_vecs(c, npad_of(c, level + 1), "prolong", {xc})) ||
      (rc = check_vecs(c, npad_of(c, level), "prolong", {x})))
    return rc;
  const Level &L = c->levels[level];
  const int64_t nc = L.ninj;
  k_prolong<double, double, false><<<tiles(nc), kBlock, 0, c->stream>>>(x, xc, L.inj, nullptr,
                                                                        L.lay.start[L.lay.inj], nc, L.frow);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int hpcgx_mg(hpcgx_ctx *c, int level, const double *r, double *x) {
  int rc = check_level(c, level, "mg");
  if (rc) return rc;
  if ((rc = check_vecs(c, npad_of(c, level), "mg", {r, x}))) return rc;
  if (r == x) return set_err(c, HPCGX_EINVAL, "mg: r and x must be distinct device vectors");
  return mg_fp64(c, level, r, x, nullptr, nullptr, true);
}

int hpcgx_mg_mixed(hpcgx_ctx *c, int precision, const double *r, double *x) {
  int rc = check_level(c, 0, "mg_mixed");
  if (rc) return rc;
  if ((rc = check_vecs(c, npad_of(c, 0), "mg_mixed", {r, x}))) return rc;
  if (r == x) return set_err(c, HPCGX_EINVAL, "mg_mixed: r and x must be distinct device vectors");
  const int need = precision == HPCGX_PREC_FP32 ? 1 : precision == HPCGX_PREC_FP16 ? 2 : 0;
  if (!need) return set_err(c, HPCGX_EINVAL, "mg_mixed: precision must be HPCGX_PREC_FP32 or HPCGX_PREC_FP16");
  if (!(c->prepared & need)) return set_err(c, HPCGX_ESTATE, "mg_mixed: call hpcgx_prepare_precision first");
  const int64_t np = c->levels[0].lay.npad;
  k_to_single<<<tiles(np), kBlock, 0, c->stream>>>(r, c->cg_r32, np);
  HPCGX_CHECK_LAUNCH(c);
  rc = precision == HPCGX_PREC_FP32 ? mg_fp32(c, 0, c->cg_r32, c->cg_z32, true)
                                    : mg_fp16(c, c->cg_r32, c->cg_z32, true);
  if (rc) return rc;
  k_to_double<<<tiles(np), kBlock, 0, c->stream>>>(c->cg_z32, x, np);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

// The vector hooks take the reference's n = nrow (ComputeDotProduct.cpp:41,
// ComputeWAXPBY.cpp:42: the whole level-0 vector).  Device vectors are in
// colour order, where the first k entries are not the first k natural rows,
// so any other length has no reference meaning and is refused; npad (the
// padded length, padding entries zero) is accepted as the same vector.
static int check_n(hpcgx_ctx *c, int64_t n, const char *who) {
  int rc = check_level(c, 0, who);
  if (rc) return rc;
  const Layout &lay = c->levels[0].lay;
  if (n != lay.n && n != lay.npad)
    return set_err(c, HPCGX_EINVAL, std::string(who) + ": n must be the level-0 row count (" +
                                        std::to_string((long long)lay.n) + ")");
  return HPCGX_OK;
}

// n = the level's row count covers the whole padded vector.
static int64_t span(hpcgx_ctx *c, int64_t) { return c->levels[0].lay.npad; }

int hpcgx_dot(hpcgx_ctx *c, int64_t n, const double *x, const double *y, double *result) {
  int rc = check_n(c, n, "dot");
  if (rc) return rc;
  if (!result) return set_err(c, HPCGX_EINVAL, "dot: null argument");
  const int64_t m = span(c, n);
  if ((rc = check_vecs(c, m, "dot", {x, y}))) return rc;
  k_dot<<<chunk_blocks(m), kBlock, 0, c->stream>>>(m, x, y, c->partials);
  HPCGX_CHECK_LAUNCH(c);
  if ((rc = finish(c, 1, chunk_blocks(m), F_DOT))) return rc;
  HPCGX_HIP(c, hipMemcpyAsync(c->hbuf, c->scal + S_DOT, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  *result = c->hbuf[0];
  return HPCGX_OK;
}

int hpcgx_waxpby(hpcgx_ctx *c, int64_t n, double alpha, const double *x, double beta, const double *y,
                 double *w) {
  int rc = check_n(c, n, "waxpby");
  if (rc) return rc;
  const int64_t m = span(c, n);
  if ((rc = check_vecs(c, m, "waxpby", {x, y, w}))) return rc;
  k_waxpby<<<tiles(m), kBlock, 0, c->stream>>>(m, alpha, x, beta, y, w);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int hpcgx_waxpby_dot(hpcgx_ctx *c, int64_t n, double alpha, const double *x, double beta, const double *y,
                     double *w, double *result) {
  int rc = check_n(c, n, "waxpby_dot");
  if (rc) return rc;
  if (!result) return set_err(c, HPCGX_EINVAL, "waxpby_dot: null argument");
  const int64_t m = span(c, n);
  if ((rc = check_vecs(c, m, "waxpby_dot", {x, y, w}))) return rc;
  k_waxpby_dot<<<chunk_blocks(m), kBlock, 0, c->stream>>>(m, alpha, x, beta, y, w, c->partials);
  HPCGX_CHECK_LAUNCH(c);
  if ((rc = finish(c, 1, chunk_blocks(m), F_DOT))) return rc;
  HPCGX_HIP(c, hipMemcpyAsync(c->hbuf, c->scal + S_DOT, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  *result = c->hbuf[0];
  return HPCGX_OK;
}

// hpcgx_dot with its partial sums in the static scratch instead of the
// context's buffers: the same kernels in the same order, so the same bits.
// Allocates nothing; the problem must already be set up (the context ready).
int hpcgx_dot_noalloc(hpcgx_ctx *c, int64_t n, const double *x, const double *y, double *result) {
  if (!c) return HPCGX_EINVAL;
  if (c->levels.empty() || !c->ready)
    return set_err(c, HPCGX_ESTATE, "dot_noalloc: no problem set up (it allocates nothing, so it cannot set one up)");
  if (!result) return set_err(c, HPCGX_EINVAL, "dot_noalloc: null argument");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, HPCGX_EHIP, "hipSetDevice failed");
  const Layout &lay = c->levels[0].lay;
  if (n != lay.n && n != lay.npad)
    return set_err(c, HPCGX_EINVAL, "dot_noalloc: n must be the level-0 row count (" +
                                        std::to_string((long long)lay.n) + ")");
  const int64_t m = lay.npad;
  if (m > kNoAllocRows)
    return set_err(c, HPCGX_EINVAL, "dot_noalloc: the level has " + std::to_string((long long)m) +
                                        " padded rows; its static scratch covers at most 2^28 (use hpcgx_dot)");
  double *s = noalloc_scratch(c->device);
  if (!s) return set_err(c, HPCGX_ESTATE, "dot_noalloc: no static scratch on this device");
  int rc = check_vecs(c, m, "dot_noalloc", {x, y});
  if (rc) return rc;
  int64_t grid = chunk_blocks(m);
  k_dot<<<(unsigned)grid, kBlock, 0, c->stream>>>(m, x, y, s);
  HPCGX_CHECK_LAUNCH(c);
  const double *src = s;
  if (grid > kFoldAbove) {
    const int64_t g2 = (grid + kFold - 1) / kFold;
    k_fold<1><<<(unsigned)g2, kBlock, 0, c->stream>>>(s, grid, s + kNaFolded);
    src = s + kNaFolded;
    grid = g2;
  }
  k_finish<1><<<1, kFinishThreads, 0, c->stream>>>(src, grid, s + kNaScal, F_DOT, nullptr);
  HPCGX_CHECK_LAUNCH(c);
  HPCGX_HIP(c, hipMemcpyAsync(c->hbuf, s + kNaScal + S_DOT, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  *result = c->hbuf[0];
  return HPCGX_OK;
}

}  // extern "C"
