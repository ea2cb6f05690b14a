// Internal state of an hpcgx context: per-level device matrices in 27-slot
// sliced ELL, colour layout, work vectors and reduction scratch.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "hpcgx.h"

namespace hpcgx {

constexpr int kSlots = 27;     // one slot per stencil offset (dx,dy,dz) in {-1,0,1}^3
constexpr int kSlice = 64;     // rows per ELL slice = one wavefront
constexpr int kDiagSlot = 13;  // slot of offset (0,0,0)
constexpr int kBlock = 256;    // threads per block = 4 slices
constexpr int kColours = 8;     // parity classes of the stencil layout
constexpr int kMaxColours = 64; // colours of a general (unstructured) level
constexpr int kMaxLevels = 8;

// Column forms of a 64-row slice in the compact index copy.  A zero-valued
// slot (a neighbour outside the grid, an unused entry) may name any row of
// the level, since its product is an exact zero; the diagonal always counts
// as a stored entry.
enum ColForm : uint8_t {
  kFormDelta = 0,    // column = base[slot] + lane + int8 delta[slot][lane]
  kFormWide = 1,     // some delta does not fit in int8: the explicit 32-bit column
  kFormUniform = 2,  // every stored entry of each slot sits at base[slot] + lane: no per-entry index
};

// Ints per slice in the compact index record: the 27 slot bases, then the
// slice's form (ColForm), then padding to 128 bytes.  A wave reads a slice's
// form and bases with scalar loads from one record, in one round trip.
constexpr int kMeta = 32;

// Value layouts of a slice (27 x 64 entries).  The slots are stored in
// units of 1, 2 or 4 slots; within a unit the values of a row are adjacent,
// so a wave reads a unit of its 64 rows with one load of 2 or 4 values per
// lane: wide loads stream from HBM faster than as many narrow ones.
//   W = 2 (fp64 values): 13 pairs and the diagonal, 16 bytes per lane;
//   W = 4 (the fp32 and fp16 copies): 5 quads, 3 pairs and the diagonal,
//         16 (fp32) or 8 (fp16) bytes per lane for a quad.
// The slots of a unit are offsets of the same parity in x, y and z, so their
// neighbours lie in the same colour block: a step that skips the slots of
// some blocks skips whole units.  Column indices keep one 64-entry run per
// slot.  A row's products are summed in slot order whatever the storage
// order.
struct ValUnit {
  int n;     // slots in the unit
  int s[4];  // the slots
};
__host__ __device__ constexpr int val_units(int W) { return W == 2 ? 14 : 9; }
__host__ __device__ constexpr ValUnit val_unit(int W, int u) {
  constexpr ValUnit pairs[14] = {{2, {0, 2}},   {2, {1, 7}},   {2, {3, 5}},   {2, {4, 22}},  {2, {6, 8}},
                                 {2, {9, 11}},  {2, {10, 16}}, {2, {12, 14}}, {1, {13}},     {2, {15, 17}},
                                 {2, {18, 20}}, {2, {19, 25}}, {2, {21, 23}}, {2, {24, 26}}};
  constexpr ValUnit quads[9] = {{4, {0, 2, 6, 8}},    {4, {18, 20, 24, 26}}, {4, {9, 11, 15, 17}},
                                {4, {3, 5, 21, 23}},  {4, {1, 7, 19, 25}},   {2, {12, 14}},
                                {2, {10, 16}},        {2, {4, 22}},          {1, {13}}};
  return W == 2 ? pairs[u] : quads[u];
}
// Offset of unit u from the start of its slice, in values.
__host__ __device__ constexpr int val_unit_off(int W, int u) {
  int off = 0;
  for (int k = 0; k < u; ++k) off += val_unit(W, k).n * kSlice;
  return off;
}
// Per slot 9 bits: bits 0-4 the unit's offset in runs of 64 values, bits
// 5-6 the slot's place in its unit, bits 7-8 log2 of the unit's size.
// Packed 7 to a word so that val_pos is arithmetic on constants, on the host
// and on the device, for a slot known at compile time or not.
__host__ __device__ constexpr uint64_t val_code_word(int W, int word) {
  uint64_t code = 0;
  for (int u = 0; u < val_units(W); ++u) {
    const ValUnit v = val_unit(W, u);
    const uint64_t run = (uint64_t)(val_unit_off(W, u) / kSlice), lw = v.n == 4 ? 2 : v.n == 2 ? 1 : 0;
    for (int h = 0; h < v.n; ++h) {
      const int s = v.s[h];
      if (s / 7 != word) continue;
      code |= (run | ((uint64_t)h << 5) | (lw << 7)) << ((s % 7) * 9);
    }
  }
  return code;
}
template <int W, int K>
struct ValCodeWord {
  static constexpr uint64_t v = val_code_word(W, K);
};
template <int W>
__host__ __device__ constexpr uint32_t val_code(int s) {
  const int k = s / 7;
  const uint64_t w = k == 0 ? ValCodeWord<W, 0>::v : k == 1 ? ValCodeWord<W, 1>::v : k == 2 ? ValCodeWord<W, 2>::v
                                                                                             : ValCodeWord<W, 3>::v;
  return (uint32_t)((w >> ((s - 7 * k) * 9)) & 0x1ffu);
}
// Index of (row, slot s) in a value array of layout W.
template <int W>
__host__ __device__ constexpr size_t val_pos(int64_t row, int s) {
  const uint32_t c = val_code<W>(s);
  return (size_t)(row >> 6) * (kSlots * kSlice) + (size_t)(c & 31) * kSlice +
         ((size_t)(row & 63) << ((c >> 7) & 3)) + ((c >> 5) & 3);
}
// Layout of the value array of element type VT.
template <class VT>
constexpr int val_width = sizeof(VT) == 8 ? 2 : 4;

// Layout checks: every slot in exactly one unit, each unit within one parity
// class, units aligned to their load size, 27 x 64 values in all, and
// val_pos agreeing with the unit table.
template <int W>
constexpr bool val_layout_ok() {
  int seen[kSlots] = {};
  const auto odd = [](int s) { return ((s % 3) != 1) | ((((s / 3) % 3) != 1) << 1) | (((s / 9) != 1) << 2); };
  for (int u = 0; u < val_units(W); ++u) {
    const ValUnit v = val_unit(W, u);
    if (v.n != 1 && v.n != 2 && v.n != 4) return false;
    if (v.n == 4 && W != 4) return false;
    if (val_unit_off(W, u) % v.n) return false;
    for (int h = 0; h < v.n; ++h) {
      if (v.s[h] < 0 || v.s[h] >= kSlots || ++seen[v.s[h]] != 1) return false;
      if (odd(v.s[h]) != odd(v.s[0])) return false;
    }
    if (v.n == 1 && v.s[0] != kDiagSlot) return false;
  }
  if (val_unit_off(W, val_units(W)) != kSlots * kSlice) return false;
  for (int u = 0; u < val_units(W); ++u) {
    const ValUnit v = val_unit(W, u);
    for (int64_t row = 0; row < 2 * kSlice; row += 63)
      for (int h = 0; h < v.n; ++h) {
        const size_t want = (size_t)(row >> 6) * (kSlots * kSlice) + (size_t)val_unit_off(W, u) +
                            (size_t)(row & 63) * v.n + h;
        if (val_pos<W>(row, v.s[h]) != want) return false;
      }
  }
  return true;
}
static_assert(val_layout_ok<2>() && val_layout_ok<4>(), "value layouts");

// Column index format of a context (hpcgx_set_column_format).
enum ColFormat : int { kColsExplicit = 0, kColsCompact = 1, kColsCompactAlways = 2 };

// Device scalars of the CG recurrences (Ctx::scal).
enum Scalar : int {
  S_RTZ = 0,
  S_OLDRTZ,
  S_PAP,
  S_NORMR2,
  S_ALPHA,
  S_BETA,
  S_DOT,
  S_DOT2,
  S_STOP,  // deferred stop test of a tolerance-0 solve: nonzero once a norm was not > 0
  S_XGO,   // the fp64 solver's last r update ran (S_STOP was clear): its x update is due
  S_COUNT = 16
};

// Sweep order: the parity class ((x&1)|(y&1)<<1|(z&1)<<2) of block k.  The
// injected rows (all coordinates even, class 0) are updated last in the
// forward sweep and first in the backward sweep: swept last in the backward
// sweep, their residual would be zero and the V-cycle would lose its
// coarse-grid correction.  Between the all-odd class 7 and class 0 come the
// complementary pairs (3,4), (5,2), (6,1).  Of the 840 orders with class 0
// last (up to axis permutations) this one converges fastest on the 192^3
// problem: it reaches the reference's 50-iteration residual reduction in 50
// iterations, and relative residual 1e-9 in fewer iterations than the
// reference's natural-order Gauss-Seidel at 32^3..96^3.
constexpr int kDefaultOrder[kColours] = {7, 3, 4, 5, 2, 6, 1, 0};

// Row layout of one level.  Stencil layout: rows are grouped in 8 blocks,
// block k holding the rows of parity class cls[k]; inside a block rows keep
// natural order of (x>>1, y>>1, z>>1), and slot s of a row is the stencil
// offset s.  General layout (rows that are not a 27-point stencil of their
// grid): ncol blocks from a first-fit colouring in breadth-first order, rows
// in that breadth-first order inside a block, and slot j of a row is the
// row's j-th entry.  Each block is padded to a
// multiple of 64 rows; padded rows hold the identity row (diagonal 1 in the
// diagonal slot, other slots 0) and zero vectors.
struct Layout {
  int nx = 0, ny = 0, nz = 0;
  int64_t n = 0, npad = 0;
  int ncol = kColours;                   // colour blocks
  bool general = false;                  // first-fit colouring of arbitrary rows
  int64_t start[kMaxColours + 1] = {};   // padded block starts
  int64_t count[kMaxColours] = {};       // true rows per block
  // stencil layout only
  int nb[kColours][3] = {};              // block extents in x, y, z
  int cls[kColours] = {};                // parity class of block k
  int pos[kColours] = {};                // block of parity class c
  int inj = 0;                           // block of class 0 (the injected rows)
  int strip = 0;                         // rows order of a block (block_index), 0: natural
};

Layout make_layout(int nx, int ny, int nz, const int *order);

__host__ __device__ inline int parity_class(int ix, int iy, int iz) {
  return (ix & 1) | ((iy & 1) << 1) | ((iz & 1) << 2);
}

struct LayoutDev {  // by-value kernel argument
  int nx, ny, nz;
  int64_t start[kColours + 1];
  int nb[kColours][3];
  int cls[kColours];
  int pos[kColours];
  int strip;
};

// Row order inside a block of extents (NX, NY, NZ).  strip == 0: natural
// order of (X, Y, Z).  strip == H: the block is cut into NY / H strips of H
// rows of Y, taken one after another, each in natural order of (X, y, Z).
// The 8 XCDs each sweep a contiguous eighth of a block; with 8 strips each
// XCD walks one strip through all Z, so the neighbour entries of one plane
// are re-read while the next plane is in flight and stay in that XCD's L2
// (in natural order a whole plane of a large grid lies between the two
// reads of an entry).
__host__ __device__ inline int64_t block_index(const LayoutDev &L, int k, int X, int Y, int Z) {
  const int64_t NX = L.nb[k][0], NY = L.nb[k][1], NZ = L.nb[k][2];
  if (!L.strip) return ((int64_t)Z * NY + Y) * NX + X;
  const int64_t H = L.strip;
  return (Y / H) * (H * NX * NZ) + ((int64_t)Z * H + Y % H) * NX + X;
}

__host__ __device__ inline void block_coords(const LayoutDev &L, int k, int64_t idx, int &X, int &Y, int &Z) {
  const int64_t NX = L.nb[k][0], NY = L.nb[k][1];
  X = (int)(idx % NX);
  if (!L.strip) {
    Y = (int)((idx / NX) % NY);
    Z = (int)(idx / (NX * NY));
    return;
  }
  const int64_t H = L.strip, NZ = L.nb[k][2];
  const int64_t s = idx / (H * NX * NZ), rem = idx % (H * NX * NZ);
  Z = (int)(rem / (H * NX));
  Y = (int)(s * H + (rem / NX) % H);
}

__host__ __device__ inline int64_t perm_of(const LayoutDev &L, int ix, int iy, int iz) {
  const int k = L.pos[parity_class(ix, iy, iz)];
  return L.start[k] + block_index(L, k, ix >> 1, iy >> 1, iz >> 1);
}

struct Level {
  Layout lay;
  LayoutDev ldev;
  int64_t nnz = 0;
  bool kappa = false;
  // host copies of the permutation
  std::vector<int> perm_h;   // natural -> colour order [n]
  std::vector<int> iperm_h;  // colour order -> natural, -1 on padding [npad]
  std::vector<int> f2c_h;    // natural f2c of this level (uploaded levels), [n_coarse]
  std::vector<uint8_t> rowlen_h;  // general levels: entries per natural row
  // device matrix in 64-row slices: columns slot-major, (slice*27+slot)*64+lane;
  // values in units of slots (val_pos<2> for vals, val_pos<4> for the copies)
  int *cols = nullptr;
  double *vals = nullptr;
  float *vals32 = nullptr;
  __half *vals16 = nullptr;
  // compact column indices (the default column format of every copy): per
  // slice a record of kMeta ints (cbase: the 32-bit base of each slot, then
  // the slice's form, ColForm), and the form again per slice in cform (for
  // launches whose waves span several slices, and the host's counts); the
  // column of (row, slot) is cbase[slice*kMeta+slot] + lane (uniform slices),
  // that plus the int8 cdelta[ell] (delta slices), or cols[ell] (wide slices,
  // whose deltas do not fit in int8)
  int *cbase = nullptr;
  int8_t *cdelta = nullptr;
  uint8_t *cform = nullptr;
  // the kernels read the compact copy on this level: fewer than half of its
  // slices are wide (on a level of mostly wide slices the explicit indices
  // ran faster: a renumbered 128^3 hierarchy, 779 against 747 GFLOP/s)
  bool read_compact = false;
  float *dinv = nullptr;   // fp16 copy: 1/sqrt(a_ii), colour order
  float *sigma = nullptr;  // fp16 copy: sqrt(d_f/d_c) per injected row
  int *perm = nullptr;
  int *iperm = nullptr;
  // Transfer to level+1, over ninj injected rows t: fine row start[inj] + t
  // (stencil layout: the injected rows are one colour block) or frow[t]
  // (general levels); coarse row inj[t].
  int *inj = nullptr;
  int *frow = nullptr;
  int64_t ninj = 0;
  // general levels: slot of the diagonal in each row (stencil: kDiagSlot)
  uint8_t *dslot = nullptr;
  // work vectors used when this level is the coarse level of a V-cycle
  double *r = nullptr, *x = nullptr;
  float *r32 = nullptr, *x32 = nullptr;
  float *r16 = nullptr;  // scaled rhs of the fp16 V-cycle (level 0 only)
};

struct Ctx {
  int device = 0;
  int order[kColours] = {7, 3, 4, 5, 2, 6, 1, 0};  // sweep order of new problems (kDefaultOrder)
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  std::vector<Level> levels;
  std::string err;
  double *partials = nullptr;  // per-block partial sums, 2 x the largest level-0 grid
  double *folded = nullptr;    // second-level partials of large reductions (same allocation)
  int64_t reduce_cap = 0;      // largest grid the scratch covers
  double *scal = nullptr;      // device scalars [S_COUNT]
  double *hist = nullptr;      // device norms of a deferred-stop batch
  int hist_cap = 0;
  std::vector<hipEvent_t> evpool;  // phase events of a deferred-stop batch (6 per iteration)
  double *hbuf = nullptr;      // pinned host scratch
  // solver buffers (level-0 length)
  double *cg_r = nullptr, *cg_z = nullptr, *cg_p = nullptr, *cg_Ap = nullptr;
  float *cg_r32 = nullptr, *cg_z32 = nullptr;
  int64_t cg_len = 0, cg32_len = 0;  // elements the solver buffers hold
  double *stage = nullptr;     // natural-order staging for upload/download
  int64_t stage_len = 0;
  // double-buffered row staging of level uploads (pinned host, device)
  uint8_t *up_host[2] = {}, *up_dev[2] = {};
  hipEvent_t up_ev[2] = {};
  int *up_err = nullptr;       // [64] per-lane error flags of the device packing
  std::unordered_map<const void *, int64_t> user_vecs;  // hpcgx_vec_alloc'd: elements
  int64_t bytes = 0;
  bool ready = false;
  int prepared = 0;  // bit 1: fp32 copies, bit 2: fp16 copies
  int col_format = kColsCompact;  // column indices the kernels read (ColFormat)
  hipEvent_t ev[8] = {};
};

// error handling ------------------------------------------------------------
int set_err(Ctx *c, int code, const std::string &msg);
#define HPCGX_HIP(ctx, call)                                                          \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::hpcgx::set_err((ctx), HPCGX_EHIP,                                      \
                              std::string(#call " : ") + hipGetErrorString(e_));      \
  } while (0)
#define HPCGX_CHECK_LAUNCH(ctx) HPCGX_HIP(ctx, hipGetLastError())

int dalloc(Ctx *c, void **p, size_t bytes, bool zero = true);
// Rejects a vector made by hpcgx_vec_alloc that holds fewer than `elems`
// doubles (a vector of another level, or of a problem since replaced).
// Device pointers from elsewhere are the caller's to size.
int check_vec(Ctx *c, const void *v, int64_t elems, const char *who);
void dfree(Ctx *c, void *p, size_t bytes);
int ensure_ready(Ctx *c);
int ensure_stage(Ctx *c, int64_t n);

// launch helpers (implemented in ops.hip)
// rz: if given, the partials of r.x (result) are left there, *nrz of them.
// untrusted_x: x may hold anything (its old contents must not leak in).
int mg_fp64(Ctx *c, int level, const double *r, double *x, double *rz = nullptr, int64_t *nrz = nullptr,
            bool untrusted_x = false);
int mg_fp32(Ctx *c, int level, const float *r, float *x, bool untrusted_x = false);
int mg_fp16(Ctx *c, const float *r, float *x, bool untrusted_x = false);
int mg_fp16_scaled(Ctx *c, const float *rhat, float *xhat);
int spmv_fp64(Ctx *c, int level, const double *x, double *y);
int symgs_fp64(Ctx *c, int level, const double *r, double *x);
int zero_vec(Ctx *c, double *x, int64_t n);
int zero_vec(Ctx *c, float *x, int64_t n);
int ensure_reduce(Ctx *c, int64_t grid);
// slot (F_NORMR2 only): record the norm there and run the deferred stop test
int finish(Ctx *c, int nv, int64_t grid, int op, double *slot = nullptr);
constexpr unsigned chunk_blocks(int64_t n) { return (unsigned)((n + 4 * 256 - 1) / (4 * 256)); }
// Looks up the static scratch of hpcgx_dot_noalloc on `device` (current).
int noalloc_scratch_init(int device);

}  // namespace hpcgx

struct hpcgx_ctx : hpcgx::Ctx {};
