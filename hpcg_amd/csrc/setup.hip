// Context lifecycle, problem upload / on-device generation, colour layout,
// lower-precision copies and vector transfers.  This is the OptimizeProblem
// side of the library (OptimizeProblem.cpp:37-107): colouring, reordering and
// packing happen once, here; the iteration kernels only read the result.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>

#include "kernels.hpp"

namespace hpcgx {

// ---------------------------------------------------------------------------
// helpers

int set_err(Ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int dalloc(Ctx *c, void **p, size_t bytes, bool zero) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_err(c, HPCGX_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  if (zero) {
    e = hipMemsetAsync(*p, 0, bytes, c->stream);
    if (e != hipSuccess) return set_err(c, HPCGX_EHIP, "hipMemsetAsync failed");
  }
  c->bytes += (int64_t)bytes;
  return HPCGX_OK;
}

int check_vec(Ctx *c, const void *v, int64_t elems, const char *who) {
  auto it = c->user_vecs.find(v);
  if (it != c->user_vecs.end() && it->second < elems)
    return set_err(c, HPCGX_EINVAL, std::string(who) + ": vector holds " + std::to_string((long long)it->second) +
                                        " entries, " + std::to_string((long long)elems) + " needed");
  return HPCGX_OK;
}

void dfree(Ctx *c, void *p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  c->bytes -= (int64_t)(bytes ? bytes : 8);
}

// Runs f(lo, hi) over up to 16 contiguous ranges of [0, n) on host threads.
template <class F>
static void parallel_ranges(int64_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 65536) nt = 1;
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto &t : th) t.join();
}

template <class F>
static void parallel_for(int64_t n, F f) {
  parallel_ranges(n, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) f(i);
  });
}

Layout make_layout(int nx, int ny, int nz, const int *order) {
  Layout L;
  L.nx = nx;
  L.ny = ny;
  L.nz = nz;
  L.n = (int64_t)nx * ny * nz;
  L.start[0] = 0;
  for (int k = 0; k < kColours; ++k) {
    L.cls[k] = order[k];
    L.pos[order[k]] = k;
  }
  L.inj = L.pos[0];
  for (int k = 0; k < kColours; ++k) {
    const int c = L.cls[k];
    L.nb[k][0] = (nx + 1 - (c & 1)) / 2;
    L.nb[k][1] = (ny + 1 - ((c >> 1) & 1)) / 2;
    L.nb[k][2] = (nz + 1 - ((c >> 2) & 1)) / 2;
    L.count[k] = (int64_t)L.nb[k][0] * L.nb[k][1] * L.nb[k][2];
    const int64_t padded = (L.count[k] + kSlice - 1) / kSlice * kSlice;
    L.start[k + 1] = L.start[k] + padded;
  }
  L.npad = L.start[kColours];
  // 8 strips of Y when every block has the same Y extent and a plane of a
  // block is large (>= 128 rows of Y: at 192^3 and below the vectors stay in
  // the 256 MB last-level cache anyway)
  const int nby = L.nb[0][1];
  bool strips = ny % 2 == 0 && nby >= 128 && nby % 8 == 0;
  for (int k = 1; k < kColours; ++k) strips = strips && L.nb[k][1] == nby;
#ifndef HPCGX_STRIPS
#define HPCGX_STRIPS 1
#endif
  L.strip = HPCGX_STRIPS && strips ? nby / 8 : 0;
  return L;
}

static LayoutDev to_dev(const Layout &L) {
  LayoutDev d;
  d.nx = L.nx;
  d.ny = L.ny;
  d.nz = L.nz;
  d.strip = L.strip;
  for (int k = 0; k <= kColours; ++k) d.start[k] = L.start[k];
  for (int k = 0; k < kColours; ++k) {
    for (int a = 0; a < 3; ++a) d.nb[k][a] = L.nb[k][a];
    d.cls[k] = L.cls[k];
    d.pos[k] = L.pos[k];
  }
  return d;
}

// Partial-sum scratch for reductions over grids of up to `grid` blocks.
static size_t reduce_elems(int64_t grid) { return (size_t)(2 * grid + 16) + (size_t)(2 * (grid / kFold + 1) + 16); }

int ensure_reduce(Ctx *c, int64_t grid) {
  if (c->reduce_cap >= grid) return HPCGX_OK;
  if (c->partials) {
    HPCGX_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, c->partials, reduce_elems(c->reduce_cap) * sizeof(double));
  }
  c->partials = nullptr;
  c->folded = nullptr;
  c->reduce_cap = 0;
  int rc = dalloc(c, (void **)&c->partials, reduce_elems(grid) * sizeof(double), false);
  if (rc) return rc;
  c->folded = c->partials + 2 * grid + 16;
  c->reduce_cap = grid;
  return HPCGX_OK;
}

// One-block finishing kernel of a reduction over `grid` producer blocks.
int finish(Ctx *c, int nv, int64_t grid, int op, double *slot) {
  const double *src = c->partials;
  if (grid > kFoldAbove) {
    const int64_t g2 = (grid + kFold - 1) / kFold;
    if (nv == 1)
      k_fold<1><<<(unsigned)g2, kBlock, 0, c->stream>>>(c->partials, grid, c->folded);
    else
      k_fold<2><<<(unsigned)g2, kBlock, 0, c->stream>>>(c->partials, grid, c->folded);
    src = c->folded;
    grid = g2;
  }
  if (nv == 1)
    k_finish<1><<<1, kFinishThreads, 0, c->stream>>>(src, grid, c->scal, op, slot);
  else
    k_finish<2><<<1, kFinishThreads, 0, c->stream>>>(src, grid, c->scal, op, slot);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int ensure_stage(Ctx *c, int64_t n) {
  if (c->stage_len >= n) return HPCGX_OK;
  if (c->stage) dfree(c, c->stage, c->stage_len * sizeof(double));
  c->stage = nullptr;
  c->stage_len = 0;
  int rc = dalloc(c, (void **)&c->stage, n * sizeof(double), false);
  if (rc) return rc;
  c->stage_len = n;
  return HPCGX_OK;
}

static void free_level(Ctx *c, Level &L) {
  const size_t m = (size_t)L.lay.npad * kSlots;
  dfree(c, L.cols, m * sizeof(int));
  dfree(c, L.vals, m * sizeof(double));
  dfree(c, L.vals32, m * sizeof(float));
  dfree(c, L.vals16, m * sizeof(__half));
  dfree(c, L.cbase, (size_t)(L.lay.npad / kSlice) * kMeta * sizeof(int));
  dfree(c, L.cdelta, m * sizeof(int8_t));
  dfree(c, L.cform, (size_t)(L.lay.npad / kSlice));
  dfree(c, L.dinv, L.lay.npad * sizeof(float));
  dfree(c, L.sigma, L.ninj * sizeof(float));
  dfree(c, L.perm, L.lay.n * sizeof(int));
  dfree(c, L.iperm, L.lay.npad * sizeof(int));
  dfree(c, L.inj, L.ninj * sizeof(int));
  dfree(c, L.frow, L.ninj * sizeof(int));
  dfree(c, L.dslot, L.lay.npad);
  dfree(c, L.r, L.lay.npad * sizeof(double));
  dfree(c, L.x, L.lay.npad * sizeof(double));
  dfree(c, L.r32, L.lay.npad * sizeof(float));
  dfree(c, L.x32, L.lay.npad * sizeof(float));
  dfree(c, L.r16, L.lay.npad * sizeof(float));
  L = Level();
}

static int alloc_level(Ctx *c, Level &L) {
  const size_t m = (size_t)L.lay.npad * kSlots;
  int rc;
  if ((rc = dalloc(c, (void **)&L.cols, m * sizeof(int), false))) return rc;
  if ((rc = dalloc(c, (void **)&L.vals, m * sizeof(double), false))) return rc;
  if ((rc = dalloc(c, (void **)&L.perm, L.lay.n * sizeof(int), false))) return rc;
  if ((rc = dalloc(c, (void **)&L.iperm, L.lay.npad * sizeof(int), false))) return rc;
  return HPCGX_OK;
}

// ---------------------------------------------------------------------------
// natural order <-> colour order

__global__ __launch_bounds__(kBlock) void k_gather_natural(double *__restrict__ dst,
                                                           const double *__restrict__ src,
                                                           const int *__restrict__ iperm,
                                                           int64_t npad) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= npad) return;
  const int i = iperm[p];
  dst[p] = i >= 0 ? src[i] : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_scatter_natural(double *__restrict__ dst,
                                                            const double *__restrict__ src,
                                                            const int *__restrict__ perm,
                                                            int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  dst[i] = src[perm[i]];
}

__global__ __launch_bounds__(kBlock) void k_fill(double *__restrict__ v, const int *__restrict__ iperm,
                                                 double value, int64_t npad) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= npad) return;
  v[p] = iperm[p] >= 0 ? value : 0.0;
}

// ---------------------------------------------------------------------------
// on-device generation: GenerateProblem_ref.cpp:131-183 in colour order

__global__ __launch_bounds__(kBlock) void k_generate(LayoutDev L, int64_t npad, int *__restrict__ cols,
                                                     double *__restrict__ vals,
                                                     int *__restrict__ iperm, int *__restrict__ perm,
                                                     const double *__restrict__ kappa, bool values) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= npad) return;
  int k = 0;
  while (k < 7 && p >= L.start[k + 1]) ++k;
  const int64_t idx = p - L.start[k];
  const int nbx = L.nb[k][0], nby = L.nb[k][1], nbz = L.nb[k][2];
  const size_t b = ell_base(p);
  if (idx >= (int64_t)nbx * nby * nbz) {  // padding: identity row
    iperm[p] = -1;
    for (int s = 0; s < kSlots; ++s) {
      cols[b + s * kSlice] = (int)p;
      vals[val_pos<2>(p, s)] = s == kDiagSlot ? 1.0 : 0.0;
    }
    return;
  }
  const int c = L.cls[k];
  int X, Y, Z;
  block_coords(L, k, idx, X, Y, Z);
  const int ix = 2 * X + (c & 1), iy = 2 * Y + ((c >> 1) & 1), iz = 2 * Z + ((c >> 2) & 1);
  const int64_t i = ((int64_t)iz * L.ny + iy) * L.nx + ix;
  iperm[p] = (int)i;
  perm[i] = (int)p;
  const double ki = kappa ? kappa[i] : 1.0;
  int nn = 0;
  double dsum = 0.0;
  for (int s = 0; s < kSlots; ++s) {
    const int jx = ix + s % 3 - 1, jy = iy + (s / 3) % 3 - 1, jz = iz + s / 9 - 1;
    const bool in = jx >= 0 && jx < L.nx && jy >= 0 && jy < L.ny && jz >= 0 && jz < L.nz;
    if (!in) {
      cols[b + s * kSlice] = (int)p;
      vals[val_pos<2>(p, s)] = 0.0;
      continue;
    }
    cols[b + s * kSlice] = (int)perm_of(L, jx, jy, jz);
    if (!values) {  // structure only: the uploaded values come next
      vals[val_pos<2>(p, s)] = 0.0;
      continue;
    }
    if (s == kDiagSlot) continue;
    ++nn;
    if (kappa) {
      const double kj = kappa[((int64_t)jz * L.ny + jy) * L.nx + jx];
      vals[val_pos<2>(p, s)] = -((ki + kj) * 0.5);
      dsum += (ki + kj) * 0.5;
    } else {
      vals[val_pos<2>(p, s)] = -1.0;
    }
  }
  if (values) vals[val_pos<2>(p, kDiagSlot)] = kappa ? dsum + (double)(26 - nn) * ki : 26.0;
}

// coarse kappa by injection (GenerateCoarseProblem.cpp:71-82 applied to kappa)
__global__ void k_inject_kappa(const double *__restrict__ kf, double *__restrict__ kc, int nxf, int nyf,
                               int ncx, int ncy, int64_t nc) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nc) return;
  const int a = (int)(t % ncx), b = (int)((t / ncx) % ncy), cc = (int)(t / ((int64_t)ncx * ncy));
  kc[t] = kf[((int64_t)(2 * cc) * nyf + 2 * b) * nxf + 2 * a];
}

// inj for generated levels: the t-th injected fine row is coarse natural row t
__global__ void k_inj_generated(LayoutDev Lf, LayoutDev Lc, int *__restrict__ inj, int64_t nc) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nc) return;
  // the t-th row of the fine class-0 block is the fine point (2a, 2b, 2c):
  // coarse point (a, b, c) (GenerateCoarseProblem.cpp:71-82)
  int a, b, cc;
  block_coords(Lf, Lf.pos[0], t, a, b, cc);
  inj[t] = (int)perm_of(Lc, a, b, cc);
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// upload from the reference's structures

// ---------------------------------------------------------------------------
// device-side packing of uploaded stencil rows (OptimizeProblem's heavy part)
//
// The host only gathers the rows, in natural order, into pinned chunks of
// kUpRows rows ([rows][27] columns and values plus the row lengths); the
// copies run while the next chunk is gathered, and one kernel per chunk
// checks each row against the stencil of the grid and scatters its values
// into the slots that k_generate (structure only) laid out.

constexpr int64_t kUpRows = 1 << 18;
constexpr size_t kUpBytes = (size_t)kUpRows * (1 + kSlots * (sizeof(int) + sizeof(double)));

struct UpChunk {
  uint8_t *nr;
  int *cols;
  double *vals;
};
static UpChunk up_view(uint8_t *base) {
  UpChunk u;
  u.vals = reinterpret_cast<double *>(base);
  u.cols = reinterpret_cast<int *>(base + (size_t)kUpRows * kSlots * sizeof(double));
  u.nr = base + (size_t)kUpRows * kSlots * (sizeof(double) + sizeof(int));
  return u;
}

// Row i0 + t of the chunk: natural row i with coordinates (ix, iy, iz).  An
// entry must be an in-grid neighbour (offset in {-1,0,1}^3), entries must
// ascend in stencil order, and the diagonal must be present; a row that is
// not flags err[lane] (the host then packs the level in the general layout).
__global__ __launch_bounds__(kBlock) void k_pack_rows(LayoutDev L, int64_t n, const uint8_t *__restrict__ nr,
                                                      const int *__restrict__ scols,
                                                      const double *__restrict__ svals, int64_t i0,
                                                      int64_t rows, const int *__restrict__ perm,
                                                      double *__restrict__ vals, int *__restrict__ err) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= rows) return;
  const int64_t i = i0 + t;
  const int ix = (int)(i % L.nx), iy = (int)((i / L.nx) % L.ny), iz = (int)(i / ((int64_t)L.nx * L.ny));
  const int64_t p = perm[i];
  const int cnt = nr[t];
  bool ok = cnt >= 1 && cnt <= kSlots, diag = false;
  int last = -1;
  for (int j = 0; ok && j < cnt; ++j) {
    const int col = scols[t * kSlots + j];
    if (col < 0 || col >= n) { ok = false; break; }
    const int dx = (int)(col % L.nx) - ix, dy = (int)((col / L.nx) % L.ny) - iy,
              dz = (int)(col / ((int64_t)L.nx * L.ny)) - iz;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) { ok = false; break; }
    const int s = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
    if (s <= last) { ok = false; break; }
    last = s;
    diag = diag || s == kDiagSlot;
    vals[val_pos<2>(p, s)] = svals[t * kSlots + j];
  }
  if (!ok || !diag) err[threadIdx.x & 63] = 1;
}

static int ensure_upload_staging(Ctx *c) {
  for (int k = 0; k < 2; ++k) {
    if (!c->up_host[k] && hipHostMalloc((void **)&c->up_host[k], kUpBytes, hipHostMallocDefault) != hipSuccess) {
      c->up_host[k] = nullptr;
      return set_err(c, HPCGX_ENOMEM, "pinned upload staging");
    }
    int rc;
    if (!c->up_dev[k] && (rc = dalloc(c, (void **)&c->up_dev[k], kUpBytes, false))) return rc;
    if (!c->up_ev[k]) HPCGX_HIP(c, hipEventCreateWithFlags(&c->up_ev[k], hipEventDisableTiming));
  }
  if (!c->up_err) {
    int rc = dalloc(c, (void **)&c->up_err, 64 * sizeof(int), false);
    if (rc) return rc;
  }
  return HPCGX_OK;
}

// Whether every row is exactly the HPCG matrix's row of the grid
// (GenerateProblem_ref.cpp:152-180): the in-grid neighbours of the 27-point
// stencil in ascending column order, 26 on the diagonal, -1 off it.  Checked
// on host threads, which read the rows at host memory speed; *nnz gets the
// stored entries.  Stops at the first row that differs.
template <class RowAcc>
static bool rows_are_hpcg(RowAcc rows, const Layout &lay, int64_t *nnz) {
  const int64_t nx = lay.nx, ny = lay.ny, nz = lay.nz, n = lay.n;
  std::atomic<bool> ok(true);
  std::atomic<int64_t> total(0);
  // interior rows: all 27 neighbours in the grid, columns i + off[k]; checked
  // in a branch-free loop the compiler vectorises
  int off[kSlots];
  double want[kSlots];
  for (int k = 0; k < kSlots; ++k) {
    off[k] = (int)(((k / 9 - 1) * ny + (k / 3 % 3 - 1)) * nx + (k % 3 - 1));
    want[k] = k == kDiagSlot ? 26.0 : -1.0;
  }
  parallel_ranges(n, [&](int64_t lo, int64_t hi) {
    int64_t sum = 0;
    for (int64_t i = lo; i < hi && ok.load(std::memory_order_relaxed); ++i) {
      const int64_t ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
      const int cnt = rows.nnz(i);
      const int *rc = rows.cols(i);
      const double *rv = rows.vals(i);
      if (ix > 0 && ix + 1 < nx && iy > 0 && iy + 1 < ny && iz > 0 && iz + 1 < nz) {
        if (cnt != kSlots) {
          ok = false;
          break;
        }
        int bad = 0;
        for (int k = 0; k < kSlots; ++k) bad |= (rc[k] != (int)i + off[k]) | (rv[k] != want[k]);
        if (bad) {
          ok = false;
          break;
        }
        sum += kSlots;
        continue;
      }
      int k = 0;
      bool row_ok = true;
      for (int sz = -1; sz <= 1 && row_ok; ++sz) {
        if (iz + sz < 0 || iz + sz >= nz) continue;
        for (int sy = -1; sy <= 1 && row_ok; ++sy) {
          if (iy + sy < 0 || iy + sy >= ny) continue;
          for (int sx = -1; sx <= 1; ++sx) {
            if (ix + sx < 0 || ix + sx >= nx) continue;
            const int64_t col = i + (sz * ny + sy) * nx + sx;
            if (k >= cnt || rc[k] != col || rv[k] != (col == i ? 26.0 : -1.0)) {
              row_ok = false;
              break;
            }
            ++k;
          }
        }
      }
      if (!row_ok || k != cnt) {
        ok = false;
        break;
      }
      sum += cnt;
    }
    total += sum;
  });
  *nnz = total;
  return ok;
}

// Packs a stencil level on the device.  *stencil is false when some row is
// not the stencil of the grid (the level's device arrays are then released).
// Rows that are exactly the HPCG matrix (rows_are_hpcg) are not sent: the
// device writes the same values from the closed form (k_generate, bit-exact
// against GenerateProblem_ref), which at 192^3 saves moving 2.3 GB over the
// bus; any other stencil rows are gathered, copied and checked on the device.
template <class RowAcc>
static int pack_stencil_device(Ctx *c, Level &L, RowAcc rows, bool *stencil) {
  const Layout &lay = L.lay;
  const int64_t n = lay.n, npad = lay.npad;
  int rc;
  *stencil = false;
  int64_t hpcg_nnz = 0;
  if (rows_are_hpcg(rows, lay, &hpcg_nnz)) {
    if ((rc = alloc_level(c, L))) return rc;
    k_generate<<<nblk(npad), kBlock, 0, c->stream>>>(L.ldev, npad, L.cols, L.vals, L.iperm, L.perm, nullptr, true);
    HPCGX_CHECK_LAUNCH(c);
    L.perm_h.resize(n);
    L.iperm_h.resize(npad);
    HPCGX_HIP(c, hipMemcpyAsync(L.perm_h.data(), L.perm, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HPCGX_HIP(c, hipMemcpyAsync(L.iperm_h.data(), L.iperm, npad * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HPCGX_HIP(c, hipStreamSynchronize(c->stream));
    L.nnz = hpcg_nnz;
    *stencil = true;
    return HPCGX_OK;
  }
  if ((rc = ensure_upload_staging(c)) || (rc = alloc_level(c, L))) return rc;
  k_generate<<<nblk(npad), kBlock, 0, c->stream>>>(L.ldev, npad, L.cols, L.vals, L.iperm, L.perm, nullptr, false);
  HPCGX_CHECK_LAUNCH(c);
  HPCGX_HIP(c, hipMemsetAsync(c->up_err, 0, 64 * sizeof(int), c->stream));
  std::atomic<int64_t> nnz(0);
  const int64_t chunks = (n + kUpRows - 1) / kUpRows;
  for (int64_t q = 0; q < chunks; ++q) {
    const int k = (int)(q & 1);
    const int64_t i0 = q * kUpRows, m = std::min<int64_t>(kUpRows, n - i0);
    if (q >= 2) HPCGX_HIP(c, hipEventSynchronize(c->up_ev[k]));  // host buffer free again
    UpChunk h = up_view(c->up_host[k]), d = up_view(c->up_dev[k]);
    parallel_ranges(m, [&](int64_t lo, int64_t hi) {
      int64_t sum = 0;  // one atomic add per range: per-row adds contend on one line
      for (int64_t t = lo; t < hi; ++t) {
        const int64_t i = i0 + t;
        int cnt = rows.nnz(i);
        if (cnt < 1 || cnt > kSlots) cnt = 0;  // row length 0: rejected on the device
        h.nr[t] = (uint8_t)cnt;
        std::memcpy(h.cols + t * kSlots, rows.cols(i), cnt * sizeof(int));
        std::memcpy(h.vals + t * kSlots, rows.vals(i), cnt * sizeof(double));
        sum += cnt;
      }
      nnz += sum;
    });
    HPCGX_HIP(c, hipMemcpyAsync(d.vals, h.vals, (size_t)m * kSlots * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCGX_HIP(c, hipMemcpyAsync(d.cols, h.cols, (size_t)m * kSlots * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HPCGX_HIP(c, hipMemcpyAsync(d.nr, h.nr, (size_t)m, hipMemcpyHostToDevice, c->stream));
    HPCGX_HIP(c, hipEventRecord(c->up_ev[k], c->stream));
    k_pack_rows<<<nblk(m), kBlock, 0, c->stream>>>(L.ldev, n, d.nr, d.cols, d.vals, i0, m, L.perm, L.vals,
                                                   c->up_err);
    HPCGX_CHECK_LAUNCH(c);
  }
  int err[64];
  HPCGX_HIP(c, hipMemcpyAsync(err, c->up_err, sizeof(err), hipMemcpyDeviceToHost, c->stream));
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  bool bad = false;
  for (int e : err) bad = bad || e != 0;
  if (bad) {
    const size_t ms = (size_t)npad * kSlots;
    dfree(c, L.cols, ms * sizeof(int));
    dfree(c, L.vals, ms * sizeof(double));
    dfree(c, L.perm, n * sizeof(int));
    dfree(c, L.iperm, npad * sizeof(int));
    L.cols = nullptr;
    L.vals = nullptr;
    L.perm = nullptr;
    L.iperm = nullptr;
    return HPCGX_OK;
  }
  L.perm_h.resize(n);
  L.iperm_h.resize(npad);
  HPCGX_HIP(c, hipMemcpy(L.perm_h.data(), L.perm, n * sizeof(int), hipMemcpyDeviceToHost));
  HPCGX_HIP(c, hipMemcpy(L.iperm_h.data(), L.iperm, npad * sizeof(int), hipMemcpyDeviceToHost));
  L.nnz = nnz;
  *stencil = true;
  return HPCGX_OK;
}

// Colouring of a general level (rows already validated): the breadth-first
// order and the first-fit colours in it.  False if more than kMaxColours
// colours would be needed.
template <class RowAcc>
static bool colour_general(RowAcc rows, int64_t n, std::vector<int> &order, std::vector<int> &colour, int &ncol) {
  // A symmetric pattern (each off-diagonal entry (i, c) has its (c, i), no
  // column twice in a row) needs no transposed lists: the rows holding column
  // i are then row i's own columns, already marked when such a list would be
  // walked, so the results are the same without them.
  std::atomic<bool> asym(false);
  parallel_for(n, [&](int64_t i) {
    const int cnt = rows.nnz(i);
    const int *rc = rows.cols(i);
    for (int k = 0; k < cnt && !asym.load(std::memory_order_relaxed); ++k) {
      const int c = rc[k];
      if (c == i) continue;
      for (int q = 0; q < k; ++q)
        if (rc[q] == c) asym = true;
      const int cc = rows.nnz(c);
      const int *cr = rows.cols(c);
      bool back = false;
      for (int q = 0; q < cc; ++q) back = back || cr[q] == i;
      if (!back) asym = true;
    }
  });
  const bool sym = !asym;
  // Otherwise, the transpose: the rows j != i that hold column i, ascending.
  std::vector<int64_t> ts;
  std::vector<int> tj;
  if (!sym) {
    ts.assign(n + 1, 0);
    for (int64_t j = 0; j < n; ++j) {
      const int cnt = rows.nnz(j);
      const int *rc = rows.cols(j);
      for (int k = 0; k < cnt; ++k)
        if (rc[k] != j) ts[rc[k] + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) ts[i + 1] += ts[i];
    tj.resize(std::max<int64_t>(ts[n], 1));
    std::vector<int64_t> fill(ts.begin(), ts.end() - 1);
    for (int64_t j = 0; j < n; ++j) {
      const int cnt = rows.nnz(j);
      const int *rc = rows.cols(j);
      for (int k = 0; k < cnt; ++k)
        if (rc[k] != j) tj[fill[rc[k]]++] = (int)j;
    }
  }
  // The passes below walk rows in a queue whose later rows are known: the
  // row 16 places ahead is fetched early, and the per-row flags of the row 8
  // places ahead's columns (the accesses are random under any numbering).
  auto ahead = [&](const std::vector<int> &q, size_t h, const auto *flag) {
    if (h + 16 < q.size()) __builtin_prefetch(rows.cols(q[h + 16]));
    if (h + 8 < q.size()) {
      const int r = q[h + 8];
      const int cnt = rows.nnz(r);
      const int *rc = rows.cols(r);
      for (int k = 0; k < cnt; ++k) __builtin_prefetch(flag + rc[k]);
    }
  };
  // Breadth-first order: components in order of their smallest row s, each
  // from a pseudo-peripheral row (George and Liu, ACM TOMS 5(3), 1979): a
  // pass from x = s gives its last level and y, the row of that level with
  // the fewest couplings; while a pass from y ends deeper than the pass from
  // x, x moves to y (at most 8 moves); at the end x gives way to y when y has
  // fewer couplings.  On a renumbered 27-point stencil the start is a corner
  // of the grid, and the first fit below needs the minimum, 8 colours, where
  // natural order needs 17.
  // Neighbours of row i: its entries in row order (diagonal skipped), then the
  // rows holding column i in ascending order.
  order.clear();
  order.reserve(n);
  {
    std::vector<char> seen(n, 0);
    std::vector<int> stamp(n, -1);
    std::vector<int> level(n, 0);
    int stamp_id = 0;
    auto visit = [&](int i, auto &&take) {
      const int cnt = rows.nnz(i);
      const int *rc = rows.cols(i);
      for (int k = 0; k < cnt; ++k)
        if (rc[k] != i) take(rc[k]);
      if (!sym)
        for (int64_t t = ts[i]; t < ts[i + 1]; ++t) take(tj[t]);
    };
    // couplings of row i in both directions (a symmetric one counts twice)
    auto couplings = [&](int i) {
      const int cnt = rows.nnz(i);
      const int *rc = rows.cols(i);
      int64_t d = 0;
      for (int k = 0; k < cnt; ++k) d += rc[k] != i;
      return sym ? 2 * d : d + ts[i + 1] - ts[i];
    };
    // a breadth-first pass over s's component into q: the depth of its last
    // level and the row of that level with the fewest couplings (first
    // reached on ties).  The pass from the start row is the order itself.
    auto probe_pass = [&](int s, std::vector<int> &q, int &far, int &depth) {
      const int id = stamp_id++;
      q.clear();
      q.push_back(s);
      stamp[s] = id;
      level[s] = 0;
      for (size_t h = 0; h < q.size(); ++h) {
        const int i = q[h];
        ahead(q, h, stamp.data());
        visit(i, [&](int j) {
          if (stamp[j] != id) { stamp[j] = id; level[j] = level[i] + 1; q.push_back(j); }
        });
      }
      depth = level[q.back()];
      size_t u = q.size() - 1;
      while (u > 0 && level[q[u - 1]] == depth) --u;
      far = q[u];
      int64_t best = couplings(far);
      for (++u; u < q.size(); ++u) {
        const int64_t d = couplings(q[u]);
        if (d < best) { best = d; far = q[u]; }
      }
    };
    std::vector<int> qx, qy;  // the passes from x and from its candidate
    qx.reserve(1024);
    qy.reserve(1024);
    for (int64_t s0 = 0; s0 < n; ++s0) {
      if (seen[s0]) continue;
      int x = (int)s0, fx, ex, fy, ey;
      probe_pass(x, qx, fx, ex);
      for (int it = 0; it < 8; ++it) {
        probe_pass(fx, qy, fy, ey);
        if (ey <= ex) {
          if (couplings(fx) < couplings(x)) { x = fx; qx.swap(qy); }
          break;
        }
        x = fx;
        qx.swap(qy);
        ex = ey;
        fx = fy;
      }
      for (const int r : qx) seen[r] = 1;
      order.insert(order.end(), qx.begin(), qx.end());
    }
  }
  // First fit in that order: row i takes the smallest colour no coloured
  // neighbour holds (for either direction of a coupling).
  colour.assign(n, -1);
  int64_t mark[kMaxColours + 1];
  for (auto &m : mark) m = -1;
  ncol = 0;
  for (size_t h = 0; h < order.size(); ++h) {
    const int i = order[h];
    ahead(order, h, colour.data());
    const int cnt = rows.nnz(i);
    const int *rc = rows.cols(i);
    for (int k = 0; k < cnt; ++k)
      if (rc[k] != i && colour[rc[k]] >= 0) mark[colour[rc[k]]] = i;
    if (!sym)
      for (int64_t t = ts[i]; t < ts[i + 1]; ++t)
        if (colour[tj[t]] >= 0) mark[colour[tj[t]]] = i;
    int pick = 0;
    while (mark[pick] == i) ++pick;
    if (pick >= kMaxColours) return false;
    colour[i] = pick;
    ncol = std::max(ncol, pick + 1);
  }
  return true;
}

// Colour order with an injection: colours sorted by how many injected rows
// (f2c targets) they hold, ascending, ties by colour, so the colour of the
// injected rows is swept last forward and first backward (swept last
// backward, its residual would be zero and the V-cycle would lose its
// coarse correction).
static void order_colours(int64_t n, int ncol, std::vector<int> &colour, const int *f2c, int64_t nc) {
  std::vector<int64_t> cnt(ncol, 0);
  for (int64_t i = 0; i < nc; ++i)
    if (f2c[i] >= 0 && f2c[i] < n) cnt[colour[f2c[i]]]++;
  std::vector<int> rank(ncol, 0);
  for (int k = 0; k < ncol; ++k)
    for (int j = 0; j < ncol; ++j)
      if (cnt[j] < cnt[k] || (cnt[j] == cnt[k] && j < k)) rank[k]++;
  for (auto &cl : colour) cl = rank[cl];
}

// One padded row p of a general level in the ELL arrays; returns its entries.
template <class RowAcc>
static int pack_general_row(Level &L, RowAcc &rows, int64_t p, std::vector<int> &hc, std::vector<double> &hv,
                            std::vector<uint8_t> &hd) {
  const size_t b = (size_t)(p >> 6) * (kSlots * kSlice) + (size_t)(p & 63);
  const int i = L.iperm_h[p];
  for (int s = 0; s < kSlots; ++s) {
    hc[b + s * kSlice] = (int)p;
    hv[val_pos<2>(p, s)] = (i < 0 && s == 0) ? 1.0 : 0.0;
  }
  if (i < 0) return 0;
  const int cnt = rows.nnz(i);
  const int *rc = rows.cols(i);
  const double *rv = rows.vals(i);
  int ds = -1;
  for (int j = 0; j < cnt; ++j) {
    hc[b + j * kSlice] = L.perm_h[rc[j]];
    hv[val_pos<2>(p, j)] = rv[j];
    if (rc[j] == i && ds < 0) ds = j;
  }
  hd[p] = (uint8_t)ds;
  L.rowlen_h[i] = (uint8_t)cnt;
  return cnt;
}

// General layout for arbitrary rows (the reference requires every
// optimisation to work for general sparse matrices, GenerateProblem.cpp:47-49).
// Colouring: first fit, like the reference's greedy example
// (OptimizeProblem.cpp:43-79), but in a breadth-first order and over the
// symmetrised pattern (a_ij != 0 or a_ji != 0), so each colour stays free of
// couplings when the pattern is not symmetric.  Blocks in colour order, rows
// in the breadth-first order inside a block, slot j = the row's j-th entry,
// so sums run in the reference's order (ComputeSPMV_ref.cpp:67-68,
// ComputeSYMGS_ref.cpp:74-77).
template <class RowAcc>
static int pack_general(Ctx *c, Level &L, RowAcc rows, const int *f2c, int64_t n_coarse, std::vector<int> &hc,
                        std::vector<double> &hv, std::vector<uint8_t> &hd) {
  const int64_t n = L.lay.n;
  std::atomic<int64_t> bad_row(-1);
  parallel_for(n, [&](int64_t i) {
    const int cnt = rows.nnz(i);
    const int *rc = rows.cols(i);
    bool ok = cnt >= 1 && cnt <= kSlots, diag = false;
    for (int j = 0; ok && j < cnt; ++j) {
      if (rc[j] < 0 || rc[j] >= n) ok = false;
      else if (rc[j] == i) diag = true;
    }
    if (!ok || !diag) bad_row = i;
  });
  if (bad_row >= 0)
    return set_err(c, HPCGX_EINVAL, "level_upload: row " + std::to_string((long long)bad_row) +
                                         " has no diagonal, a column out of range, or not 1..27 entries");
  std::vector<int> order, colour;
  int ncol = 0;
  if (!colour_general(rows, n, order, colour, ncol))
    return set_err(c, HPCGX_EINVAL, "level_upload: greedy colouring needs more than " +
                                        std::to_string(kMaxColours) + " colours");
  if (f2c) order_colours(n, ncol, colour, f2c, n_coarse);
  Layout &lay = L.lay;
  lay.general = true;
  lay.ncol = ncol;
  lay.strip = 0;
  lay.inj = 0;
  for (int k = 0; k < kMaxColours; ++k) lay.count[k] = 0;
  for (int64_t i = 0; i < n; ++i) lay.count[colour[i]]++;
  lay.start[0] = 0;
  for (int k = 0; k < ncol; ++k) lay.start[k + 1] = lay.start[k] + (lay.count[k] + kSlice - 1) / kSlice * kSlice;
  for (int k = ncol + 1; k <= kMaxColours; ++k) lay.start[k] = lay.start[ncol];
  lay.npad = lay.start[ncol];
  if (lay.npad >= ((int64_t)1 << 31)) return set_err(c, HPCGX_EINVAL, "level_upload: too many rows for 32-bit columns");
  const int64_t npad = lay.npad;
  // rows inside a block in the breadth-first order: neighbouring rows land
  // near each other in every block, so the entries a tile of rows gathers
  // are close together whatever the caller's numbering (the order inside a
  // colour changes no result: rows of one colour are not coupled)
  L.perm_h.assign(n, 0);
  L.iperm_h.assign(npad, -1);
  {
    std::vector<int64_t> fill(lay.start, lay.start + ncol);
    for (const int i : order) {
      const int64_t p = fill[colour[i]]++;
      L.perm_h[i] = (int)p;
      L.iperm_h[p] = i;
    }
  }
  const size_t m = (size_t)npad * kSlots;
  hc.assign(m, 0);
  hv.assign(m, 0.0);
  hd.assign(npad, 0);
  L.rowlen_h.assign(n, 0);
  std::atomic<int64_t> nnz(0);
  parallel_ranges(npad, [&](int64_t lo, int64_t hi) {
    int64_t sum = 0;
    for (int64_t p = lo; p < hi; ++p) sum += pack_general_row(L, rows, p, hc, hv, hd);
    nnz += sum;
  });
  L.nnz = nnz;
  return HPCGX_OK;
}

// Uploads one level.  Rows that form a 27-point stencil of the (nx, ny, nz)
// grid get the parity-class layout; any other rows get the general layout.
// f2c: natural-order injection of the n_coarse rows of level + 1, or null.
template <class RowAcc>
static int upload_level(Ctx *c, int level, int nx, int ny, int nz, RowAcc rows, const int *f2c,
                        int64_t n_coarse, bool try_stencil) {
  if (level < 0 || level >= kMaxLevels || nx < 1 || ny < 1 || nz < 1)
    return set_err(c, HPCGX_EINVAL, "level_upload: bad level or grid");
  if ((int)c->levels.size() < level) return set_err(c, HPCGX_ESTATE, "levels must be uploaded fine to coarse");
  if ((int)c->levels.size() > level) {
    for (int l = (int)c->levels.size() - 1; l >= level; --l) free_level(c, c->levels[l]);
    c->levels.resize(level);
  }
  c->ready = false;
  c->prepared = 0;
  c->levels.emplace_back();
  Level &L = c->levels.back();
  L.lay = make_layout(nx, ny, nz, c->order);
  L.ldev = to_dev(L.lay);
  if (L.lay.npad >= ((int64_t)1 << 31) && try_stencil)
    return set_err(c, HPCGX_EINVAL, "level_upload: too many rows for 32-bit columns");
  int rc;
  bool stencil = false;
  if (try_stencil && (rc = pack_stencil_device(c, L, rows, &stencil))) return rc;
  if (!stencil) {
    std::vector<int> hc;
    std::vector<double> hv;
    std::vector<uint8_t> hd;
    if ((rc = pack_general(c, L, rows, f2c, n_coarse, hc, hv, hd))) return rc;
    const Layout &lay = L.lay;
    const size_t m = (size_t)lay.npad * kSlots;
    if ((rc = alloc_level(c, L))) return rc;
    HPCGX_HIP(c, hipMemcpy(L.cols, hc.data(), m * sizeof(int), hipMemcpyHostToDevice));
    HPCGX_HIP(c, hipMemcpy(L.vals, hv.data(), m * sizeof(double), hipMemcpyHostToDevice));
    HPCGX_HIP(c, hipMemcpy(L.perm, L.perm_h.data(), lay.n * sizeof(int), hipMemcpyHostToDevice));
    HPCGX_HIP(c, hipMemcpy(L.iperm, L.iperm_h.data(), lay.npad * sizeof(int), hipMemcpyHostToDevice));
    if ((rc = dalloc(c, (void **)&L.dslot, lay.npad, false))) return rc;
    HPCGX_HIP(c, hipMemcpy(L.dslot, hd.data(), lay.npad, hipMemcpyHostToDevice));
  }
  const Layout &lay = L.lay;
  if (f2c) {
    if (n_coarse < 1 || n_coarse > lay.n) return set_err(c, HPCGX_EINVAL, "level_upload: bad coarse row count for f2c");
    L.f2c_h.assign(f2c, f2c + n_coarse);
  }
  return HPCGX_OK;
}

// Rows of the next level of an nx*ny*nz level: the grid halves in every
// dimension (GenerateCoarseProblem.cpp:50-53).  Negative: error code.
static int64_t grid_coarse_rows(Ctx *c, int nx, int ny, int nz, const int *f2c) {
  if (!f2c) return 0;
  if (nx < 2 || ny < 2 || nz < 2 || nx % 2 || ny % 2 || nz % 2)
    return set_err(c, HPCGX_EINVAL, "level_upload: f2c given for a grid that cannot be halved");
  return (int64_t)(nx / 2) * (ny / 2) * (nz / 2);
}

struct PackedRows {
  const uint8_t *nr;
  const int *c;
  const double *v;
  int nnz(int64_t i) const { return nr[i]; }
  const int *cols(int64_t i) const { return c + i * kSlots; }
  const double *vals(int64_t i) const { return v + i * kSlots; }
};

struct PointerRows {
  const char *nr;
  const int *const *c;
  const double *const *v;
  int nnz(int64_t i) const { return (int)(unsigned char)nr[i]; }
  const int *cols(int64_t i) const { return c[i]; }
  const double *vals(int64_t i) const { return v[i]; }
};

// Builds what the V-cycle needs once every level is present: the injection
// maps, the coarse work vectors and the solver buffers.
static int build_compact_all(Ctx *c);

int ensure_ready(Ctx *c) {
  if (c->ready) return HPCGX_OK;
  if (c->levels.empty()) return set_err(c, HPCGX_ESTATE, "no problem: upload or generate levels first");
  const int nl = (int)c->levels.size();
  int rc;
  for (int l = 0; l + 1 < nl; ++l) {
    Level &F = c->levels[l], &Cc = c->levels[l + 1];
    const int64_t nc = Cc.lay.n;
    if (F.inj) continue;
    if ((int64_t)F.f2c_h.size() != nc)
      return set_err(c, HPCGX_ESTATE, "level " + std::to_string(l) + " has no f2c operator for the " +
                                          std::to_string((long long)nc) + " rows of level " + std::to_string(l + 1));
    // Stencil levels whose f2c is the reference's injection (coarse point
    // (a, b, c) <- fine point (2a, 2b, 2c), GenerateCoarseProblem.cpp:71-82)
    // on grids that halve exactly take the generated map, checked in one
    // pass over f2c instead of inverting it over the fine level.
    if (!F.lay.general && !Cc.lay.general && F.lay.nx == 2 * Cc.lay.nx && F.lay.ny == 2 * Cc.lay.ny &&
        F.lay.nz == 2 * Cc.lay.nz) {
      const int64_t cx = Cc.lay.nx, cy = Cc.lay.ny, fx = F.lay.nx, fy = F.lay.ny;
      bool closed = true;
      for (int64_t k = 0; closed && k < nc; ++k) {
        const int64_t a = k % cx, b = (k / cx) % cy, cz = k / (cx * cy);
        closed = F.f2c_h[k] == ((2 * cz) * fy + 2 * b) * fx + 2 * a;
      }
      if (closed) {
        F.ninj = nc;
        if ((rc = dalloc(c, (void **)&F.inj, nc * sizeof(int), false))) return rc;
        k_inj_generated<<<nblk(nc), kBlock, 0, c->stream>>>(F.ldev, Cc.ldev, F.inj, nc);
        HPCGX_CHECK_LAUNCH(c);
        continue;
      }
    }
    std::vector<int> inv(F.lay.npad, -1);
    for (int64_t k = 0; k < nc; ++k) {
      const int f = F.f2c_h[k];
      if (f < 0 || f >= F.lay.n) return set_err(c, HPCGX_EINVAL, "f2c out of range");
      const int pf = F.perm_h[f];
      if (inv[pf] >= 0) return set_err(c, HPCGX_EINVAL, "f2c must map coarse rows one-to-one onto fine rows");
      inv[pf] = (int)k;
    }
    // Stencil layouts whose injected rows are exactly the class-0 block
    // transfer over that block; otherwise over the list of injected rows.
    const int ib = F.lay.inj;
    bool block = !F.lay.general && !Cc.lay.general && F.lay.count[ib] == nc;
    for (int64_t t = 0; block && t < nc; ++t) block = inv[F.lay.start[ib] + t] >= 0;
    std::vector<int> inj(nc), frow;
    if (block) {
      for (int64_t t = 0; t < nc; ++t) inj[t] = Cc.perm_h[inv[F.lay.start[ib] + t]];
    } else {
      // ascending fine row, so a transfer launch walks the fine level in order
      frow.reserve(nc);
      for (int64_t p = 0; p < F.lay.npad; ++p)
        if (inv[p] >= 0) frow.push_back((int)p);
      for (int64_t t = 0; t < nc; ++t) inj[t] = Cc.perm_h[inv[frow[t]]];
    }
    F.ninj = nc;
    if ((rc = dalloc(c, (void **)&F.inj, nc * sizeof(int), false))) return rc;
    HPCGX_HIP(c, hipMemcpy(F.inj, inj.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    if (!block) {
      if ((rc = dalloc(c, (void **)&F.frow, nc * sizeof(int), false))) return rc;
      HPCGX_HIP(c, hipMemcpy(F.frow, frow.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  for (int l = 1; l < nl; ++l) {
    Level &L = c->levels[l];
    if (!L.r && (rc = dalloc(c, (void **)&L.r, L.lay.npad * sizeof(double)))) return rc;
    if (!L.x && (rc = dalloc(c, (void **)&L.x, L.lay.npad * sizeof(double)))) return rc;
  }
  const int64_t np = c->levels[0].lay.npad;
  // partials of one pass over level 0 (+ one tile per colour block for the
  // r.z partials of the blocks' steps)
  if ((rc = ensure_reduce(c, std::max<int64_t>((np + kBlock - 1) / kBlock, chunk_blocks(np)) + kMaxColours)))
    return rc;
  if (c->cg_len < np) {
    HPCGX_HIP(c, hipStreamSynchronize(c->stream));
    for (double **b : {&c->cg_r, &c->cg_z, &c->cg_p, &c->cg_Ap}) {
      dfree(c, *b, c->cg_len * sizeof(double));
      *b = nullptr;
    }
    c->cg_len = 0;
    for (double **b : {&c->cg_r, &c->cg_z, &c->cg_p, &c->cg_Ap})
      if ((rc = dalloc(c, (void **)b, np * sizeof(double)))) return rc;
    c->cg_len = np;
  }
  if ((rc = build_compact_all(c))) return rc;
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  c->ready = true;
  return HPCGX_OK;
}

// ---------------------------------------------------------------------------
// lower-precision copies

// fp32 copy of the values, from the fp64 layout into the fp32 one
__global__ void k_to_f32(const double *__restrict__ v, float *__restrict__ o, int64_t npad) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= npad) return;
  for (int s = 0; s < kSlots; ++s) o[val_pos<4>(p, s)] = (float)v[val_pos<2>(p, s)];
}

// Address of the diagonal of row p: slot 13 of the stencil layout, dslot[p]
// of a general level.
__device__ __forceinline__ size_t diag_at(const uint8_t *dslot, int64_t p) {
  return val_pos<2>(p, dslot ? dslot[p] : kDiagSlot);
}

// dinv = 1/sqrt(a_ii) per row (padding: a_ii = 1)
__global__ void k_dinv(const double *__restrict__ vals, const uint8_t *__restrict__ dslot,
                       float *__restrict__ dinv, int64_t npad) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= npad) return;
  dinv[p] = (float)(1.0 / sqrt(vals[diag_at(dslot, p)]));
}

// ahat_ij = a_ij * dinv_i * dinv_j in fp64, rounded to fp32 then fp16; diagonal 1
__global__ void k_to_f16_scaled(const int *__restrict__ cols, const double *__restrict__ vals,
                                const uint8_t *__restrict__ dslot, __half *__restrict__ v16,
                                int64_t npad) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= npad) return;
  const size_t b = ell_base(p);
  const int ds = dslot ? dslot[p] : kDiagSlot;
  const double di = 1.0 / sqrt(vals[val_pos<2>(p, ds)]);
  for (int s = 0; s < kSlots; ++s) {
    double a;
    if (s == ds) {
      a = 1.0;
    } else {
      const int cj = cols[b + s * kSlice];
      const double dj = 1.0 / sqrt(vals[diag_at(dslot, cj)]);
      a = vals[val_pos<2>(p, s)] * di * dj;
    }
    v16[val_pos<4>(p, s)] = __float2half_rn((float)a);
  }
}

// sigma_t = sqrt(d_f / d_c) for the t-th injected row
__global__ void k_sigma(const double *__restrict__ fvals, const uint8_t *__restrict__ fdslot,
                        const double *__restrict__ cvals, const uint8_t *__restrict__ cdslot,
                        const int *__restrict__ inj, const int *__restrict__ frow,
                        float *__restrict__ sigma, int64_t row0, int64_t nc) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nc) return;
  const int64_t f = frow ? (int64_t)frow[t] : row0 + t;
  const double df = fvals[diag_at(fdslot, f)];
  const double dc = cvals[diag_at(cdslot, inj[t])];
  sigma[t] = (float)sqrt(df / dc);
}

// Compact column indices, one wave per 64-row slice.  For each slot the
// columns of the rows with a nonzero value (and of the diagonal, whatever its
// value) are base + lane + d with d in [-128, 127]; a zero-valued slot
// (neighbour outside the grid, unused entry) may name any row of the level,
// as its product is an exact zero: it takes its own row when reachable, else
// the nearest reachable one.  A slice whose columns do not fit is wide and
// keeps its explicit indices.  A slice in which every slot's stored entries
// share one offset (column - lane), and whose zero-valued slots may then
// take base + lane (a row of the level), is uniform: base + lane is its
// column, with no per-entry index.  The form depends only on the columns and
// on which entries are zero, never on the matrix being a stencil.
__global__ __launch_bounds__(kBlock) void k_compact_cols(const int *__restrict__ cols,
                                                         const double *__restrict__ vals,
                                                         const uint8_t *__restrict__ dslot,
                                                         int *__restrict__ cbase, int8_t *__restrict__ cdelta,
                                                         uint8_t *__restrict__ cform, int64_t nslices, int64_t npad) {
  const int lane = threadIdx.x & 63;
  const int64_t slice = (int64_t)blockIdx.x * (kBlock / kSlice) + (threadIdx.x >> 6);
  if (slice >= nslices) return;
  const int64_t row = slice * kSlice + lane;
  const size_t b = (size_t)slice * (kSlots * kSlice) + lane;
  const int ds = dslot ? dslot[row] : kDiagSlot;
  bool wide = false, uniform = true;
  for (int s = 0; s < kSlots; ++s) {
    const int col = cols[b + s * kSlice];
    const bool live = vals[val_pos<2>(row, s)] != 0.0 || s == ds;
    int lo = live ? col - lane : INT_MAX, hi = live ? col - lane : INT_MIN;
    for (int m = 32; m >= 1; m >>= 1) {
      lo = min(lo, __shfl_xor(lo, m, 64));
      hi = max(hi, __shfl_xor(hi, m, 64));
    }
    int64_t base;
    if (lo == INT_MAX) {
      base = slice * kSlice;  // no live lane: every lane names its own row
    } else {
      if ((int64_t)hi - lo > 255) wide = true;
      if (hi != lo) uniform = false;
      base = (int64_t)lo + (((int64_t)hi - lo + 1) >> 1);
    }
    // uniform: every lane reads base + lane, which must be a row of the level
    if (base + lane < 0 || base + lane >= npad) uniform = false;
    int64_t target = live ? (int64_t)col : row;
    const int64_t rlo = max<int64_t>(0, base + lane - 128), rhi = min<int64_t>(npad - 1, base + lane + 127);
    if (rlo > rhi) wide = true;
    if (!live) target = min(max(target, rlo), rhi);
    const int64_t d = target - base - lane;
    if (d < -128 || d > 127) wide = true;
    cdelta[b + s * kSlice] = (int8_t)(wide ? 0 : d);
    if (lane == 0) cbase[slice * kMeta + s] = (int)base;
  }
  // wide and uniform are per lane so far; the slice is wide if any lane is,
  // uniform if every lane is
  const bool any_wide = __any(wide), all_uniform = __all(uniform);
  const int form = any_wide ? kFormWide : all_uniform ? kFormUniform : kFormDelta;
  if (lane == 0) cform[slice] = (uint8_t)form;
  if (lane < kMeta - kSlots) cbase[slice * kMeta + kSlots + lane] = lane == 0 ? form : 0;  // the form, then padding
}

static int build_compact_cols(Ctx *c, Level &L) {
  if (L.cbase) return HPCGX_OK;
  const int64_t ns = L.lay.npad / kSlice;
  int rc;
  if ((rc = dalloc(c, (void **)&L.cbase, (size_t)ns * kMeta * sizeof(int), false))) return rc;
  if ((rc = dalloc(c, (void **)&L.cdelta, (size_t)L.lay.npad * kSlots * sizeof(int8_t), false))) return rc;
  if ((rc = dalloc(c, (void **)&L.cform, (size_t)ns, false))) return rc;
  const int64_t grid = (ns + (kBlock / kSlice) - 1) / (kBlock / kSlice);
  k_compact_cols<<<grid, kBlock, 0, c->stream>>>(L.cols, L.vals, L.dslot, L.cbase, L.cdelta, L.cform, ns,
                                                 L.lay.npad);
  HPCGX_CHECK_LAUNCH(c);
  std::vector<uint8_t> f(ns);
  HPCGX_HIP(c, hipMemcpyAsync(f.data(), L.cform, ns, hipMemcpyDeviceToHost, c->stream));
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  int64_t wide = 0;
  for (const uint8_t v : f) wide += v == kFormWide;
  L.read_compact = 2 * wide < ns;
  return HPCGX_OK;
}

// The compact index copies of every level (the default column format).
static int build_compact_all(Ctx *c) {
  if (c->col_format == kColsExplicit) return HPCGX_OK;
  for (auto &L : c->levels) {
    int rc = build_compact_cols(c, L);
    if (rc) return rc;
  }
  return HPCGX_OK;
}

static int build_precision(Ctx *c, int prec) {
  int rc = ensure_ready(c);
  if (rc) return rc;
  const int nl = (int)c->levels.size();
  for (int l = 0; l < nl; ++l) {
    Level &L = c->levels[l];
    const int64_t m = L.lay.npad * kSlots;
    if (prec == HPCGX_PREC_FP32) {
      if (!L.vals32 && (rc = dalloc(c, (void **)&L.vals32, m * sizeof(float), false))) return rc;
      k_to_f32<<<nblk(L.lay.npad), kBlock, 0, c->stream>>>(L.vals, L.vals32, L.lay.npad);
      HPCGX_CHECK_LAUNCH(c);
    } else {
      if (!L.vals16 && (rc = dalloc(c, (void **)&L.vals16, m * sizeof(__half), false))) return rc;
      if (!L.dinv && (rc = dalloc(c, (void **)&L.dinv, L.lay.npad * sizeof(float), false))) return rc;
      k_to_f16_scaled<<<nblk(L.lay.npad), kBlock, 0, c->stream>>>(L.cols, L.vals, L.dslot, L.vals16,
                                                                  L.lay.npad);
      k_dinv<<<nblk(L.lay.npad), kBlock, 0, c->stream>>>(L.vals, L.dslot, L.dinv, L.lay.npad);
      HPCGX_CHECK_LAUNCH(c);
    }
    if (!L.r32 && l > 0 && (rc = dalloc(c, (void **)&L.r32, L.lay.npad * sizeof(float)))) return rc;
    if (!L.x32 && l > 0 && (rc = dalloc(c, (void **)&L.x32, L.lay.npad * sizeof(float)))) return rc;
  }
  if (prec == HPCGX_PREC_FP16) {
    for (int l = 0; l + 1 < nl; ++l) {
      Level &F = c->levels[l];
      const Level &Cc = c->levels[l + 1];
      const int64_t nc = F.ninj;
      if (!F.sigma && (rc = dalloc(c, (void **)&F.sigma, nc * sizeof(float), false))) return rc;
      k_sigma<<<nblk(nc), kBlock, 0, c->stream>>>(F.vals, F.dslot, Cc.vals, Cc.dslot, F.inj, F.frow, F.sigma,
                                                  F.lay.start[F.lay.inj], nc);
      HPCGX_CHECK_LAUNCH(c);
    }
    Level &L0 = c->levels[0];
    if (!L0.r16 && (rc = dalloc(c, (void **)&L0.r16, L0.lay.npad * sizeof(float)))) return rc;
  }
  const int64_t np = c->levels[0].lay.npad;
  if (c->cg32_len < np) {
    HPCGX_HIP(c, hipStreamSynchronize(c->stream));
    for (float **b : {&c->cg_r32, &c->cg_z32}) {
      dfree(c, *b, c->cg32_len * sizeof(float));
      *b = nullptr;
    }
    c->cg32_len = 0;
    for (float **b : {&c->cg_r32, &c->cg_z32})
      if ((rc = dalloc(c, (void **)b, np * sizeof(float)))) return rc;
    c->cg32_len = np;
  }
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  c->prepared |= (prec == HPCGX_PREC_FP32) ? 1 : 2;
  return HPCGX_OK;
}

// diagonal scatter / gather (natural order staging <-> colour-order ELL)
__global__ void k_set_diag(double *__restrict__ vals, const uint8_t *__restrict__ dslot,
                           const int *__restrict__ perm, const double *__restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) vals[diag_at(dslot, perm[i])] = d[i];
}
__global__ void k_get_diag(const double *__restrict__ vals, const uint8_t *__restrict__ dslot,
                           const int *__restrict__ perm, double *__restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) d[i] = vals[diag_at(dslot, perm[i])];
}

}  // namespace hpcgx

using namespace hpcgx;

// ===========================================================================
// C ABI

extern "C" {

int hpcgx_device_init(int device) {
  if (hipSetDevice(device) != hipSuccess) return HPCGX_EHIP;
  if (hipFree(nullptr) != hipSuccess) return HPCGX_EHIP;  // brings up the device's context
  return HPCGX_OK;
}

int hpcgx_create(int device, hpcgx_ctx **out) {
  if (!out) return HPCGX_EINVAL;
  *out = nullptr;
  hpcgx_ctx *c = new hpcgx_ctx();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete c;
    return HPCGX_EHIP;
  }
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return HPCGX_EHIP;
  }
  c->stream = c->own_stream;
  int rc;
  if ((rc = noalloc_scratch_init(device)) || (rc = ensure_reduce(c, 4096)) ||
      (rc = dalloc(c, (void **)&c->scal, S_COUNT * sizeof(double)))) {
    hpcgx_destroy(c);
    return rc;
  }
  if (hipHostMalloc((void **)&c->hbuf, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    hpcgx_destroy(c);
    return HPCGX_ENOMEM;
  }
  for (auto &ev : c->ev) (void)hipEventCreate(&ev);
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    hpcgx_destroy(c);
    return HPCGX_EHIP;
  }
  *out = c;
  return HPCGX_OK;
}

int hpcgx_destroy(hpcgx_ctx *c) {
  if (!c) return HPCGX_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &L : c->levels) free_level(c, L);
  for (auto &kv : c->user_vecs) (void)hipFree(const_cast<void *>(kv.first));
  (void)hipFree(c->cg_r);
  (void)hipFree(c->cg_z);
  (void)hipFree(c->cg_p);
  (void)hipFree(c->cg_Ap);
  (void)hipFree(c->cg_r32);
  (void)hipFree(c->cg_z32);
  (void)hipFree(c->stage);
  (void)hipFree(c->partials);
  (void)hipFree(c->scal);
  (void)hipFree(c->hist);
  if (c->hbuf) (void)hipHostFree(c->hbuf);
  for (int k = 0; k < 2; ++k) {
    if (c->up_host[k]) (void)hipHostFree(c->up_host[k]);
    (void)hipFree(c->up_dev[k]);
    if (c->up_ev[k]) (void)hipEventDestroy(c->up_ev[k]);
  }
  (void)hipFree(c->up_err);
  for (auto &ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : c->evpool) (void)hipEventDestroy(ev);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return HPCGX_OK;
}

const char *hpcgx_error_string(const hpcgx_ctx *c) { return c ? c->err.c_str() : "null context"; }

int hpcgx_set_stream(hpcgx_ctx *c, void *stream) {
  if (!c) return HPCGX_EINVAL;
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  c->stream = (hipStream_t)stream;
  return HPCGX_OK;
}

int hpcgx_sync(hpcgx_ctx *c) {
  if (!c) return HPCGX_EINVAL;
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  return HPCGX_OK;
}

int hpcgx_level_upload_packed(hpcgx_ctx *c, int level, int nx, int ny, int nz, const uint8_t *nnz_row,
                              const int *cols, const double *vals, const int *f2c) {
  if (!c || !nnz_row || !cols || !vals) return set_err(c, HPCGX_EINVAL, "level_upload_packed: null argument");
  HPCGX_HIP(c, hipSetDevice(c->device));
  int64_t nc = 0;
  if ((nc = grid_coarse_rows(c, nx, ny, nz, f2c)) < 0) return (int)nc;
  return upload_level(c, level, nx, ny, nz, PackedRows{nnz_row, cols, vals}, f2c, nc, true);
}

int hpcgx_level_upload_rows(hpcgx_ctx *c, int level, int64_t n, const uint8_t *nnz_row, const int *cols,
                            const double *vals, const int *f2c, int64_t n_coarse) {
  if (!c || !nnz_row || !cols || !vals) return set_err(c, HPCGX_EINVAL, "level_upload_rows: null argument");
  if (n < 1 || n >= ((int64_t)1 << 31)) return set_err(c, HPCGX_EINVAL, "level_upload_rows: bad row count");
  HPCGX_HIP(c, hipSetDevice(c->device));
  return upload_level(c, level, (int)n, 1, 1, PackedRows{nnz_row, cols, vals}, f2c, f2c ? n_coarse : 0, false);
}

int hpcgx_level_upload(hpcgx_ctx *c, int level, int nx, int ny, int nz, const char *nonzerosInRow,
                       const int *const *mtxIndL, const double *const *matrixValues, const int *f2c) {
  if (!c || !nonzerosInRow || !mtxIndL || !matrixValues) return set_err(c, HPCGX_EINVAL, "level_upload: null argument");
  HPCGX_HIP(c, hipSetDevice(c->device));
  int64_t nc = 0;
  if ((nc = grid_coarse_rows(c, nx, ny, nz, f2c)) < 0) return (int)nc;
  return upload_level(c, level, nx, ny, nz, PointerRows{nonzerosInRow, mtxIndL, matrixValues}, f2c, nc, true);
}

int hpcgx_problem_generate(hpcgx_ctx *c, int nx, int ny, int nz, int levels, const double *kappa) {
  if (!c || nx < 1 || ny < 1 || nz < 1 || levels < 1 || levels > kMaxLevels)
    return set_err(c, HPCGX_EINVAL, "problem_generate: bad grid or level count");
  HPCGX_HIP(c, hipSetDevice(c->device));
  for (int l = 1; l < levels; ++l)
    if (((nx >> (l - 1)) & 1) || ((ny >> (l - 1)) & 1) || ((nz >> (l - 1)) & 1))
      return set_err(c, HPCGX_EINVAL, "problem_generate: grid cannot be halved that many times");
  for (int l = (int)c->levels.size() - 1; l >= 0; --l) free_level(c, c->levels[l]);
  c->levels.clear();
  c->ready = false;
  c->prepared = 0;
  double *kdev = nullptr, *kprev = nullptr;
  int rc;
  int64_t kprev_n = 0;
  for (int l = 0; l < levels; ++l) {
    c->levels.emplace_back();
    Level &L = c->levels.back();
    L.lay = make_layout(nx >> l, ny >> l, nz >> l, c->order);
    L.ldev = to_dev(L.lay);
    L.kappa = kappa != nullptr;
    if (L.lay.npad >= ((int64_t)1 << 31)) return set_err(c, HPCGX_EINVAL, "problem_generate: too many rows for 32-bit columns");
    if ((rc = alloc_level(c, L))) return rc;
    if (kappa) {
      if ((rc = dalloc(c, (void **)&kdev, L.lay.n * sizeof(double), false))) return rc;
      if (l == 0) {
        HPCGX_HIP(c, hipMemcpy(kdev, kappa, L.lay.n * sizeof(double), hipMemcpyHostToDevice));
      } else {
        const Layout &F = c->levels[l - 1].lay;
        k_inject_kappa<<<nblk(L.lay.n), kBlock, 0, c->stream>>>(kprev, kdev, F.nx, F.ny, L.lay.nx,
                                                                L.lay.ny, L.lay.n);
        HPCGX_CHECK_LAUNCH(c);
      }
    }
    k_generate<<<nblk(L.lay.npad), kBlock, 0, c->stream>>>(L.ldev, L.lay.npad, L.cols, L.vals, L.iperm,
                                                          L.perm, kdev, true);
    HPCGX_CHECK_LAUNCH(c);
    int64_t nnz = 1;
    for (const int d : {L.lay.nx, L.lay.ny, L.lay.nz}) nnz *= (int64_t)(3 * d - 2);
    L.nnz = nnz;
    if (kprev) {
      HPCGX_HIP(c, hipStreamSynchronize(c->stream));
      dfree(c, kprev, kprev_n * sizeof(double));
    }
    kprev = kdev;
    kprev_n = L.lay.n;
    kdev = nullptr;
  }
  if (kprev) {
    HPCGX_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, kprev, kprev_n * sizeof(double));
  }
  for (int l = 0; l + 1 < levels; ++l) {
    Level &F = c->levels[l];
    const Level &Cc = c->levels[l + 1];
    if ((rc = dalloc(c, (void **)&F.inj, Cc.lay.n * sizeof(int), false))) return rc;
    F.ninj = Cc.lay.n;
    k_inj_generated<<<nblk(Cc.lay.n), kBlock, 0, c->stream>>>(F.ldev, Cc.ldev, F.inj, Cc.lay.n);
    HPCGX_CHECK_LAUNCH(c);
  }
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  return ensure_ready(c);
}

int hpcgx_set_colour_order(hpcgx_ctx *c, const int *order) {
  if (!c || !order) return HPCGX_EINVAL;
  int seen = 0;
  for (int k = 0; k < kColours; ++k) {
    if (order[k] < 0 || order[k] >= kColours || (seen >> order[k]) & 1)
      return set_err(c, HPCGX_EINVAL, "set_colour_order: not a permutation of the 8 parity classes");
    seen |= 1 << order[k];
  }
  for (int k = 0; k < kColours; ++k) c->order[k] = order[k];
  return HPCGX_OK;
}

int hpcgx_level_info(const hpcgx_ctx *c, int level, int64_t *out) {
  if (!c || !out || level < 0 || level >= (int)c->levels.size()) return HPCGX_EINVAL;
  const Level &L = c->levels[level];
  out[0] = L.lay.n;
  out[1] = L.lay.npad;
  out[2] = L.nnz;
  out[3] = L.lay.nx;
  out[4] = L.lay.ny;
  out[5] = L.lay.nz;
  out[6] = L.lay.npad * kSlots;
  out[7] = (int64_t)c->levels.size();
  return HPCGX_OK;
}

int hpcgx_colour_rows(int64_t n, const uint8_t *nnz_row, const int *cols, const int *f2c, int64_t n_coarse,
                      int *colour, int *order, int *ncolours) {
  if (n < 1 || n >= ((int64_t)1 << 31) || !nnz_row || !cols || !colour || !ncolours) return HPCGX_EINVAL;
  for (int64_t i = 0; i < n; ++i) {
    if (nnz_row[i] < 1 || nnz_row[i] > kSlots) return HPCGX_EINVAL;
    for (int k = 0; k < nnz_row[i]; ++k)
      if (cols[i * kSlots + k] < 0 || cols[i * kSlots + k] >= n) return HPCGX_EINVAL;
  }
  std::vector<int> ord, col;
  int nc = 0;
  if (!colour_general(PackedRows{nnz_row, cols, nullptr}, n, ord, col, nc)) return HPCGX_EINVAL;
  if (f2c) {
    if (n_coarse < 0) return HPCGX_EINVAL;
    order_colours(n, nc, col, f2c, n_coarse);
  }
  std::memcpy(colour, col.data(), n * sizeof(int));
  if (order) std::memcpy(order, ord.data(), n * sizeof(int));
  *ncolours = nc;
  return HPCGX_OK;
}

int hpcgx_level_colouring(const hpcgx_ctx *c, int level, int *ncolours, int *general, int *colour) {
  if (!c || level < 0 || level >= (int)c->levels.size()) return HPCGX_EINVAL;
  const Level &L = c->levels[level];
  if (ncolours) *ncolours = L.lay.ncol;
  if (general) *general = L.lay.general ? 1 : 0;
  if (colour && L.lay.general) {
    for (int k = 0; k < L.lay.ncol; ++k)
      for (int64_t p = L.lay.start[k]; p < L.lay.start[k + 1]; ++p)
        if (L.iperm_h[p] >= 0) colour[L.iperm_h[p]] = k;
  } else if (colour) {
    const int nx = L.lay.nx, ny = L.lay.ny;
    for (int64_t i = 0; i < L.lay.n; ++i)
      colour[i] = L.lay.pos[parity_class((int)(i % nx), (int)((i / nx) % ny), (int)(i / ((int64_t)nx * ny)))];
  }
  return HPCGX_OK;
}

int hpcgx_level_set_diagonal(hpcgx_ctx *c, int level, const double *diag) {
  if (!c || !diag || level < 0 || level >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "set_diagonal: bad argument");
  HPCGX_HIP(c, hipSetDevice(c->device));
  Level &L = c->levels[level];
  int rc = ensure_stage(c, L.lay.n);
  if (rc) return rc;
  HPCGX_HIP(c, hipMemcpyAsync(c->stage, diag, L.lay.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_set_diag<<<nblk(L.lay.n), kBlock, 0, c->stream>>>(L.vals, L.dslot, L.perm, c->stage, L.lay.n);
  HPCGX_CHECK_LAUNCH(c);
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  // the lower-precision copies (and the fp16 scaling) follow the new diagonal
  const int prepared = c->prepared;
  if ((prepared & 1) && (rc = build_precision(c, HPCGX_PREC_FP32))) return rc;
  if ((prepared & 2) && (rc = build_precision(c, HPCGX_PREC_FP16))) return rc;
  return HPCGX_OK;
}

int hpcgx_level_get_diagonal(hpcgx_ctx *c, int level, double *diag) {
  if (!c || !diag || level < 0 || level >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "get_diagonal: bad argument");
  HPCGX_HIP(c, hipSetDevice(c->device));
  Level &L = c->levels[level];
  int rc = ensure_stage(c, L.lay.n);
  if (rc) return rc;
  k_get_diag<<<nblk(L.lay.n), kBlock, 0, c->stream>>>(L.vals, L.dslot, L.perm, c->stage, L.lay.n);
  HPCGX_CHECK_LAUNCH(c);
  HPCGX_HIP(c, hipMemcpyAsync(diag, c->stage, L.lay.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  return HPCGX_OK;
}

int hpcgx_prepare_precision(hpcgx_ctx *c, int precision) {
  if (!c || (precision != HPCGX_PREC_FP32 && precision != HPCGX_PREC_FP16 && precision != HPCGX_PREC_FP64))
    return set_err(c, HPCGX_EINVAL, "prepare_precision: bad precision");
  HPCGX_HIP(c, hipSetDevice(c->device));
  if (precision == HPCGX_PREC_FP64) return ensure_ready(c);
  return build_precision(c, precision);
}

int hpcgx_set_column_format(hpcgx_ctx *c, int format) {
  if (!c) return HPCGX_EINVAL;
  if (format != HPCGX_COLS_EXPLICIT && format != HPCGX_COLS_COMPACT && format != HPCGX_COLS_COMPACT_ALWAYS)
    return set_err(c, HPCGX_EINVAL,
                   "set_column_format: HPCGX_COLS_EXPLICIT, HPCGX_COLS_COMPACT or HPCGX_COLS_COMPACT_ALWAYS");
  HPCGX_HIP(c, hipSetDevice(c->device));
  c->col_format = format == HPCGX_COLS_COMPACT ? kColsCompact
                 : format == HPCGX_COLS_COMPACT_ALWAYS ? kColsCompactAlways : kColsExplicit;
  if (c->ready) {
    int rc = build_compact_all(c);
    if (rc) return rc;
    HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  }
  return HPCGX_OK;
}

int hpcgx_level_column_forms(const hpcgx_ctx *c, int level, int64_t *counts) {
  if (!c || !counts) return HPCGX_EINVAL;
  if (level < 0 || level >= (int)c->levels.size()) return HPCGX_EINVAL;
  const Level &L = c->levels[level];
  counts[0] = counts[1] = counts[2] = 0;
  counts[3] = L.cbase && ((c->col_format == kColsCompact && L.read_compact) || c->col_format == kColsCompactAlways);
  if (!L.cbase) return HPCGX_OK;
  const int64_t ns = L.lay.npad / kSlice;
  std::vector<uint8_t> f(ns);
  if (hipSetDevice(c->device) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(f.data(), L.cform, ns, hipMemcpyDeviceToHost) != hipSuccess)
    return HPCGX_EHIP;
  for (const uint8_t v : f) counts[v < 3 ? v : 1] += 1;
  return HPCGX_OK;
}

double hpcgx_memory_bytes(const hpcgx_ctx *c) { return c ? (double)c->bytes : 0.0; }

int hpcgx_vec_alloc(hpcgx_ctx *c, int level, double **dvec) {
  if (!c || !dvec || level < 0 || level >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "vec_alloc: bad argument");
  HPCGX_HIP(c, hipSetDevice(c->device));
  const int64_t np = c->levels[level].lay.npad;
  int rc = dalloc(c, (void **)dvec, np * sizeof(double));
  if (rc) return rc;
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  c->user_vecs[*dvec] = np;
  return HPCGX_OK;
}

int hpcgx_vec_free(hpcgx_ctx *c, double *dvec) {
  if (!c || !dvec) return HPCGX_EINVAL;
  auto it = c->user_vecs.find(dvec);
  if (it == c->user_vecs.end()) return set_err(c, HPCGX_EINVAL, "vec_free: not a vector of this context");
  const int64_t elems = it->second;
  c->user_vecs.erase(it);
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  dfree(c, dvec, elems * sizeof(double));
  return HPCGX_OK;
}

int hpcgx_vec_upload(hpcgx_ctx *c, int level, double *dvec, const double *host, int64_t n) {
  if (!c || !dvec || !host || level < 0 || level >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "vec_upload: bad argument");
  Level &L = c->levels[level];
  if (n != L.lay.n) return set_err(c, HPCGX_EINVAL, "vec_upload: length must equal the level's rows");
  if (int rc0 = check_vec(c, dvec, L.lay.npad, "vec_upload")) return rc0;
  HPCGX_HIP(c, hipSetDevice(c->device));
  int rc = ensure_stage(c, n);
  if (rc) return rc;
  HPCGX_HIP(c, hipMemcpyAsync(c->stage, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_gather_natural<<<nblk(L.lay.npad), kBlock, 0, c->stream>>>(dvec, c->stage, L.iperm, L.lay.npad);
  HPCGX_CHECK_LAUNCH(c);
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  return HPCGX_OK;
}

int hpcgx_vec_download(hpcgx_ctx *c, int level, const double *dvec, double *host, int64_t n) {
  if (!c || !dvec || !host || level < 0 || level >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "vec_download: bad argument");
  Level &L = c->levels[level];
  if (n != L.lay.n) return set_err(c, HPCGX_EINVAL, "vec_download: length must equal the level's rows");
  if (int rc0 = check_vec(c, dvec, L.lay.npad, "vec_download")) return rc0;
  HPCGX_HIP(c, hipSetDevice(c->device));
  int rc = ensure_stage(c, n);
  if (rc) return rc;
  k_scatter_natural<<<nblk(n), kBlock, 0, c->stream>>>(c->stage, dvec, L.perm, n);
  HPCGX_CHECK_LAUNCH(c);
  HPCGX_HIP(c, hipMemcpyAsync(host, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  return HPCGX_OK;
}

int hpcgx_vec_from_natural(hpcgx_ctx *c, int level, double *dvec, const double *dnat) {
  if (!c || !dvec || !dnat || level < 0 || level >= (int)c->levels.size() || dvec == dnat)
    return set_err(c, HPCGX_EINVAL, "vec_from_natural: bad argument");
  Level &L = c->levels[level];
  if (int rc0 = check_vec(c, dvec, L.lay.npad, "vec_from_natural")) return rc0;
  if (int rc0 = check_vec(c, dnat, L.lay.n, "vec_from_natural")) return rc0;
  HPCGX_HIP(c, hipSetDevice(c->device));
  k_gather_natural<<<nblk(L.lay.npad), kBlock, 0, c->stream>>>(dvec, dnat, L.iperm, L.lay.npad);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int hpcgx_vec_to_natural(hpcgx_ctx *c, int level, const double *dvec, double *dnat) {
  if (!c || !dvec || !dnat || level < 0 || level >= (int)c->levels.size() || dvec == dnat)
    return set_err(c, HPCGX_EINVAL, "vec_to_natural: bad argument");
  Level &L = c->levels[level];
  if (int rc0 = check_vec(c, dvec, L.lay.npad, "vec_to_natural")) return rc0;
  if (int rc0 = check_vec(c, dnat, L.lay.n, "vec_to_natural")) return rc0;
  HPCGX_HIP(c, hipSetDevice(c->device));
  k_scatter_natural<<<nblk(L.lay.n), kBlock, 0, c->stream>>>(dnat, dvec, L.perm, L.lay.n);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

int hpcgx_vec_fill(hpcgx_ctx *c, int level, double *dvec, double value) {
  if (!c || !dvec || level < 0 || level >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "vec_fill: bad argument");
  Level &L = c->levels[level];
  if (int rc0 = check_vec(c, dvec, L.lay.npad, "vec_fill")) return rc0;
  HPCGX_HIP(c, hipSetDevice(c->device));
  k_fill<<<nblk(L.lay.npad), kBlock, 0, c->stream>>>(dvec, L.iperm, value, L.lay.npad);
  HPCGX_CHECK_LAUNCH(c);
  return HPCGX_OK;
}

// Test / inspection: the level's matrix in natural row order, packed
// [n*27] with ascending columns (unused slots: col -1, value 0).
int hpcgx_level_export_packed(hpcgx_ctx *c, int level, uint8_t *nnz_row, int *cols, double *vals) {
  if (!c || level < 0 || level >= (int)c->levels.size()) return set_err(c, HPCGX_EINVAL, "level_export: bad argument");
  HPCGX_HIP(c, hipSetDevice(c->device));
  Level &L = c->levels[level];
  const size_t m = (size_t)L.lay.npad * kSlots;
  std::vector<int> hc(m), ip(L.lay.npad);
  std::vector<double> hv(m);
  HPCGX_HIP(c, hipStreamSynchronize(c->stream));
  HPCGX_HIP(c, hipMemcpy(hc.data(), L.cols, m * sizeof(int), hipMemcpyDeviceToHost));
  HPCGX_HIP(c, hipMemcpy(hv.data(), L.vals, m * sizeof(double), hipMemcpyDeviceToHost));
  HPCGX_HIP(c, hipMemcpy(ip.data(), L.iperm, L.lay.npad * sizeof(int), hipMemcpyDeviceToHost));
  const int nx = L.lay.nx, ny = L.lay.ny, nz = L.lay.nz;
  for (int64_t p = 0; p < L.lay.npad; ++p) {
    const int i = ip[p];
    if (i < 0) continue;
    const size_t b = (size_t)(p >> 6) * (kSlots * kSlice) + (size_t)(p & 63);
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
    int k = 0;
    for (int s = 0; s < kSlots; ++s) {
      if (L.lay.general) {
        if (s >= L.rowlen_h[i]) break;
      } else {
        const int jx = ix + s % 3 - 1, jy = iy + (s / 3) % 3 - 1, jz = iz + s / 9 - 1;
        if (jx < 0 || jx >= nx || jy < 0 || jy >= ny || jz < 0 || jz >= nz) continue;
      }
      const int pc = hc[b + s * kSlice];
      cols[(int64_t)i * kSlots + k] = pc >= 0 && pc < L.lay.npad ? ip[pc] : -2;
      vals[(int64_t)i * kSlots + k] = hv[val_pos<2>(p, s)];
      ++k;
    }
    nnz_row[i] = (uint8_t)k;
    for (; k < kSlots; ++k) {
      cols[(int64_t)i * kSlots + k] = -1;
      vals[(int64_t)i * kSlots + k] = 0.0;
    }
  }
  return HPCGX_OK;
}

}  // extern "C"
