// Device kernels of the PCG iteration phase (gfx950, wave64).
//
// Matrix storage: 27-slot sliced ELL.  Slot s of a row is the stencil offset
// (dx,dy,dz) = (s%3-1, s/3%3-1, s/9-1); a neighbour outside the grid is a
// zero value whose explicit column is the row itself.  Within a slice of 64
// rows one wavefront reads the fp64 values of a pair of slots of its 64 rows
// as one contiguous 1024-byte run, 16 bytes per lane (the fp32 and fp16
// copies: four slots per load; val_pos), and each slot's column as a
// 256-byte run of explicit indices, a 64-byte run of int8 deltas, or no
// per-entry index at all (ColForm).  Rows are summed in slot
// order, ascending column order, the order in which the reference sums a row
// (ComputeSPMV_ref.cpp:67-68); the zero slots add exact zeros, so sums match
// the reference bit for bit.  The whole library is compiled with
// -ffp-contract=off: products and sums are rounded separately, as in the
// reference's x86-64 build.
#pragma once

#include <type_traits>
#include <utility>

#include "common.hpp"

namespace hpcgx {

__device__ __forceinline__ size_t ell_base(int64_t row) {
  return (size_t)(row >> 6) * (kSlots * kSlice) + (size_t)(row & 63);
}

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD works on
// one contiguous range of tiles and neighbouring rows share that XCD's L2.
// ROT (the fine level's fp64 Gauss-Seidel steps): XCD x starts its range at
// x/8 of it and wraps.  A step's 8 ranges are slabs of one colour block,
// equally spaced in memory, and swept in lockstep they meet the same HBM
// channels at the same time; rotated, the fine 192^3 step took 57.6 us
// instead of 59.3 (explicit indices, one value slot per load; 4 alternating
// processes each).  Whole-level passes (SpMV)
// keep the plain map: there XCD x sweeps block x, and in lockstep the 8 XCDs
// gather x around the same grid position, which the last-level cache serves
// once (rotated, the 512^3 SpMV took 9.2-9.6 ms instead of 8.1-8.5).  The
// fp32 / fp16 steps keep it too (rotated, the mixed 512^3 solves ran 1.7 %
// slower).
template <bool ROT = false>
__device__ __forceinline__ int64_t tile_id() {
  const int64_t G = gridDim.x, b = blockIdx.x;
  const int64_t per = G >> 3, rem = G & 7, xcd = b & 7, idx = b >> 3;
  int64_t r = idx;
  if constexpr (ROT) {
    const int64_t len = per + (xcd < rem ? 1 : 0);
    r += (xcd * len) >> 3;
    if (r >= len) r -= len;
  }
  return xcd * per + (xcd < rem ? xcd : rem) + r;
}

// Occupancy of the fp64 sparse kernels (SpMV, Gauss-Seidel steps,
// restriction).  With explicit 32-bit column indices: at most two waves per
// SIMD, so a wave holds the column, value and gather loads of all its slots
// in flight instead of the default's batches of about 9 at six waves per
// SIMD (measured with one value slot per load: 3 % faster at 512^3, the CG
// SpMV 9 %, 1.7 % at 192^3; eight waves per SIMD was 10 % slower; with
// paired values the explicit-column 192^3 set ran 1058 GFLOP/s at two, 1047
// at four and 1022 at eight, the general 128^3 layout 776, 779 and 670;
// HPCGX_FP64_WAVES overrides it at build time).  With
// compact columns (the default) a uniform slice needs no index load, all
// loads of a row still issue before the first wait (114 VGPRs for the fine
// step with paired values), and no limit is best: measured in alternating
// processes with one value slot per load, the 192^3 sets ran 1245-1249
// GFLOP/s unlimited, 1248-1249 at four waves per SIMD, 1241 at three and
// 1181 at two; the 512^3 FP64 solve 37.9-38.3 ms per iteration unlimited
// against 40.7-41.5 at two.  HPCGX_CI_FP64_WAVES overrides it at build time.  The
// fp32 / fp16 kernels keep the default, and so do steps that read fewer
// than 8 slots (the backend pads every kernel under the limit to its
// register budget).
#ifndef HPCGX_CI_FP64_WAVES
#define HPCGX_CI_FP64_WAVES 8
#endif
#ifndef HPCGX_FP64_WAVES
#define HPCGX_FP64_WAVES 2
#endif
template <class VT, bool CI = false, int SLOTS = 27>
struct SparseOccupancy {
  static constexpr unsigned max_waves =
      (sizeof(VT) == 8 && SLOTS >= 8) ? (CI ? HPCGX_CI_FP64_WAVES : HPCGX_FP64_WAVES) : 8;
};
#define HPCGX_SPARSE_OCCUPANCY(VT, CI) __attribute__((amdgpu_waves_per_eu(1, SparseOccupancy<VT, CI>::max_waves)))

// Matrix loads.  NT (non-temporal): the fine level's matrix is streamed once
// per pass and is far larger than any cache, so it is kept out of the way of
// the vectors; coarse-level matrices fit in the 256 MB Infinity Cache and
// are loaded normally so that the V-cycle's repeated passes can hit there.
template <bool NT = true>
__device__ __forceinline__ int ld_col(const int *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <class ACC, bool NT = true>
__device__ __forceinline__ ACC ld_val(const double *p) {
  if constexpr (NT) return (ACC)__builtin_nontemporal_load(p);
  else return (ACC)*p;
}
template <class ACC, bool NT = true>
__device__ __forceinline__ ACC ld_val(const float *p) {
  if constexpr (NT) return (ACC)__builtin_nontemporal_load(p);
  else return (ACC)*p;
}
template <class ACC, bool NT = true>
__device__ __forceinline__ ACC ld_val(const __half *p) {
  const unsigned short *q = reinterpret_cast<const unsigned short *>(p);
  unsigned short bits;
  if constexpr (NT) bits = __builtin_nontemporal_load(q);
  else bits = *q;
  return (ACC)__half2float(__ushort_as_half(bits));
}

// The values of a unit of a row (val_pos): 2 or 4 adjacent values read with
// one load per lane -- a pair of fp64 values in 16 bytes, a quad of fp32 in
// 16, a quad of fp16 in 8, a pair of fp32 or fp16 in 8 or 4.
typedef double f64x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
template <class V, bool NT>
__device__ __forceinline__ V ld_vec(const V *q) {
  if constexpr (NT) return __builtin_nontemporal_load(q);
  else return *q;
}
template <class ACC>
__device__ __forceinline__ ACC half_lo(unsigned bits) {
  return (ACC)__half2float(__ushort_as_half((unsigned short)(bits & 0xffffu)));
}
template <class ACC>
__device__ __forceinline__ ACC half_hi(unsigned bits) {
  return (ACC)__half2float(__ushort_as_half((unsigned short)(bits >> 16)));
}
template <class ACC, bool NT>
__device__ __forceinline__ void ld_unit(const double *p, ACC *a, std::integral_constant<int, 2>) {
  const f64x2_t v = ld_vec<f64x2_t, NT>(reinterpret_cast<const f64x2_t *>(p));
  a[0] = (ACC)v.x;
  a[1] = (ACC)v.y;
}
template <class ACC, bool NT>
__device__ __forceinline__ void ld_unit(const float *p, ACC *a, std::integral_constant<int, 2>) {
  const f32x2_t v = ld_vec<f32x2_t, NT>(reinterpret_cast<const f32x2_t *>(p));
  a[0] = (ACC)v.x;
  a[1] = (ACC)v.y;
}
template <class ACC, bool NT>
__device__ __forceinline__ void ld_unit(const float *p, ACC *a, std::integral_constant<int, 4>) {
  const f32x4_t v = ld_vec<f32x4_t, NT>(reinterpret_cast<const f32x4_t *>(p));
  a[0] = (ACC)v.x;
  a[1] = (ACC)v.y;
  a[2] = (ACC)v.z;
  a[3] = (ACC)v.w;
}
template <class ACC, bool NT>
__device__ __forceinline__ void ld_unit(const __half *p, ACC *a, std::integral_constant<int, 2>) {
  const unsigned bits = ld_vec<unsigned, NT>(reinterpret_cast<const unsigned *>(p));
  a[0] = half_lo<ACC>(bits);
  a[1] = half_hi<ACC>(bits);
}
template <class ACC, bool NT>
__device__ __forceinline__ void ld_unit(const __half *p, ACC *a, std::integral_constant<int, 4>) {
  const u32x2_t bits = ld_vec<u32x2_t, NT>(reinterpret_cast<const u32x2_t *>(p));
  a[0] = half_lo<ACC>(bits.x);
  a[1] = half_hi<ACC>(bits.x);
  a[2] = half_lo<ACC>(bits.y);
  a[3] = half_hi<ACC>(bits.y);
}

template <class F, int... U>
__device__ __forceinline__ void for_each_unit(F &&f, std::integer_sequence<int, U...>) {
  (f(std::integral_constant<int, U>{}), ...);
}

// Slots of a unit as a mask.
__host__ __device__ constexpr uint32_t unit_mask(const ValUnit &w) {
  uint32_t m = 0;
  for (int h = 0; h < w.n; ++h) m |= 1u << w.s[h];
  return m;
}

// The values of the slots in MASK of a row, into v[slot]; other entries of v
// are left unset.  A unit is read whole when any of its slots is wanted.
constexpr uint32_t kAllSlots = (1u << kSlots) - 1;
template <class ACC, bool NT, uint32_t MASK, class VT>
__device__ __forceinline__ void row_vals(const VT *__restrict__ vals, int64_t row, ACC (&v)[kSlots]) {
  constexpr int W = val_width<VT>;
  const VT *sb = vals + (size_t)(row >> 6) * (kSlots * kSlice);
  const int lane = (int)(row & 63);
  for_each_unit(
      [&](auto uc) {
        constexpr int u = decltype(uc)::value;
        constexpr ValUnit w = val_unit(W, u);
        constexpr int off = val_unit_off(W, u);
        if constexpr ((MASK & unit_mask(w)) != 0) {
          if constexpr (w.n == 1) {
            v[w.s[0]] = ld_val<ACC, NT>(sb + off + lane);
          } else {
            ACC t[w.n];
            ld_unit<ACC, NT>(sb + off + w.n * lane, t, std::integral_constant<int, w.n>{});
#pragma unroll
            for (int h = 0; h < w.n; ++h) v[w.s[h]] = t[h];
          }
        }
      },
      std::make_integer_sequence<int, val_units(W)>{});
}

// Column index of (row, slot): the explicit 32-bit index, or the compact
// form (ColForm) with a per-slice base.
struct ColIdx {
  const int *cols;
  const int *cbase;     // null: explicit indices only
  const int8_t *cdelta;
  const uint8_t *cform;
};

// The columns a launch reads on a level: the compact form when the context
// uses it and the level reads it (Level::read_compact), else the explicit
// indices.
inline ColIdx col_idx(const Ctx *c, const Level &L) {
  if (L.cbase && ((c->col_format == kColsCompact && L.read_compact) || c->col_format == kColsCompactAlways))
    return ColIdx{L.cols, L.cbase, L.cdelta, L.cform};
  return ColIdx{L.cols, nullptr, nullptr, nullptr};
}

// Slice of a row.  UNI: the 64 lanes of the wave hold the 64 rows of one
// slice (launches over 64-aligned rows), so the slice is wave-uniform and the
// per-slice metadata comes through the scalar cache.
template <bool UNI>
__device__ __forceinline__ int slice_of(int64_t row) {
  if constexpr (UNI) return __builtin_amdgcn_readfirstlane((int)(row >> 6));
  else return (int)(row >> 6);
}

// Column of (row, slot s).  CM 0: the explicit 32-bit index; CM 1: base +
// lane + int8 delta; CM 2: base + lane.  A launch picks one per slice
// (with_cols), so each inner loop is straight-line code; base holds the
// slice's 27 slot bases (with_cols provides them).
template <int CM, bool NT>
__device__ __forceinline__ int col_at(const ColIdx &ci, int64_t row, const int *base, size_t b, int s) {
  if constexpr (CM == 0) {
    return ld_col<NT>(ci.cols + b + s * kSlice);
  } else if constexpr (CM == 1) {
    int8_t d;
    if constexpr (NT) d = __builtin_nontemporal_load(ci.cdelta + b + s * kSlice);
    else d = ci.cdelta[b + s * kSlice];
    return base[s] + (int)(row & 63) + (int)d;
  } else {
    return base[s] + (int)(row & 63);
  }
}

template <int V>
using cmode = std::integral_constant<int, V>;

// Runs f(cmode<CM>{}, base) with the column mode of the row's slice and its
// slot bases.  CI: the launch reads compact indices.  UNI: the slice is the
// same for the whole wave: the bases and the form are read together from
// the slice's record before the choice, scalar loads in one round trip
// (small coarse steps are bound by their chain of dependent loads: reading
// the form first and the bases after it cost one more), and the choice is a
// scalar branch.  Otherwise each lane reads its slice's form from cform and
// its bases from the record as the slots need them.
template <bool CI, bool UNI, class F>
__device__ __forceinline__ auto with_cols(const ColIdx &ci, int slice, F &&f) {
  if constexpr (!CI) {
    return f(cmode<0>{}, (const int *)nullptr);
  } else if constexpr (UNI) {
    const int *rec = ci.cbase + (int64_t)slice * kMeta;
    int base[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) base[s] = rec[s];
    const int form = rec[kSlots];
    if (form == kFormUniform) return f(cmode<2>{}, (const int *)base);
    if (form == kFormDelta) return f(cmode<1>{}, (const int *)base);
    return f(cmode<0>{}, (const int *)base);
  } else {
    const int *rec = ci.cbase + (int64_t)slice * kMeta;
    const int form = ci.cform[slice];
    if (form == kFormUniform) return f(cmode<2>{}, rec);
    if (form == kFormDelta) return f(cmode<1>{}, rec);
    return f(cmode<0>{}, rec);
  }
}

// ---------------------------------------------------------------------------
// deterministic reductions: fixed partition, fixed tree, last block finishes

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Deterministic reductions in two kernels.  Each block of the producing
// kernel reduces its values in a fixed tree (shuffles, then 4 waves in order)
// and stores one partial per value with a plain store; a one-block finishing
// kernel then sums the partials in a fixed order and applies the scalar
// update.  Within a kernel no fences or atomics are needed: the kernel
// boundary makes the partials visible (an agent-scope fence inside a kernel
// writes back or invalidates the whole L2 of an XCD, which costs far more
// than the extra launch).
template <int NV>
__device__ __forceinline__ void block_partials(const double (&v)[NV], double *partials) {
  __shared__ double sh[4][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[w][k] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      partials[(size_t)k * gridDim.x + blockIdx.x] = (sh[0][k] + sh[1][k]) + (sh[2][k] + sh[3][k]);
  }
}

// What the finishing kernel does with the totals (t0, t1).
enum FinishOp : int {
  F_DOT = 0,        // scal[S_DOT] = t0
  F_RZ_FIRST,       // rtz = t0                                   (CG.cpp:110)
  F_RZ,             // oldrtz = rtz; rtz = t0; beta = rtz/oldrtz  (CG.cpp:112-114)
  F_PAP,            // pAp = t0; alpha = rtz/pAp                  (CG.cpp:119-120)
  F_NORMR2,         // normr^2 = t0                               (CG.cpp:123)
                    // with a slot: record t0 there; stop test t0 > 0
  F_FCG,            // oldrtz = rtz; rtz = t0; beta = -alpha t1 / oldrtz
};

constexpr int kFinishThreads = 1024;
constexpr int kFinishLoads = 32;  // partials per finishing thread read in one round

// Reductions over more than kFoldAbove partials first fold each run of kFold
// consecutive partials into one (one block per run), so the one-block
// finishing kernel reads few values: its single CU would otherwise take
// ~100 us over the 524288 partials of a 512^3 pass.
constexpr int64_t kFold = 4096;
constexpr int64_t kFoldAbove = 32768;

template <int NV>
__global__ __launch_bounds__(kBlock) void k_fold(const double *__restrict__ in, int64_t G, double *out) {
  double v[NV];
  const int64_t b0 = (int64_t)blockIdx.x * kFold;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double *p = in + (size_t)k * G;
    double a0 = 0.0, a1 = 0.0;
#pragma unroll 4
    for (int64_t j = threadIdx.x; j < kFold; j += 2 * kBlock) {
      const int64_t i = b0 + j;
      if (i < G) a0 += p[i];
      if (i + kBlock < G) a1 += p[i + kBlock];
    }
    v[k] = a0 + a1;
  }
  block_partials<NV>(v, out);
}

template <int NV>
__global__ __launch_bounds__(kFinishThreads) void k_finish(const double *__restrict__ partials, int64_t G,
                                                           double *scal, int op, double *slot) {
  __shared__ double sh[kFinishThreads / 64][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double *p = partials + (size_t)k * G;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    constexpr int64_t T = kFinishThreads;
    const int64_t tid = threadIdx.x;
    if (G <= kFinishLoads * T) {
      // The additions of the loop below, in the same order into the same
      // accumulators, with all loads issued first: one round trip to memory
      // instead of one per group of four.  Element j of this thread is
      // p[tid + j T]; the loop adds it to a(j % 4) in group j / 4 while the
      // group's last element is below G, then the rest to a0.
      double v[kFinishLoads];
#pragma unroll
      for (int j = 0; j < kFinishLoads; ++j) {
        const int64_t i = tid + j * T;
        v[j] = i < G ? p[i] : 0.0;
      }
      const int64_t ngroups = G > tid + 3 * T ? (G - tid - 3 * T - 1) / (4 * T) + 1 : 0;
      const int64_t nmine = G > tid ? (G - tid - 1) / T + 1 : 0;
#pragma unroll
      for (int j = 0; j < kFinishLoads; ++j) {
        if (j / 4 < ngroups) {
          if (j % 4 == 0) a0 += v[j];
          else if (j % 4 == 1) a1 += v[j];
          else if (j % 4 == 2) a2 += v[j];
          else a3 += v[j];
        } else if (j < nmine) {
          a0 += v[j];
        }
      }
    } else {
      int64_t i = tid;
      for (; i + 3 * T < G; i += 4 * T) {
        a0 += p[i];
        a1 += p[i + T];
        a2 += p[i + 2 * T];
        a3 += p[i + 3 * T];
      }
      for (; i < G; i += T) a0 += p[i];
    }
    const double s = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) sh[w][k] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int j = 0; j < kFinishThreads / 64; ++j) s += sh[j][k];
    t[k] = s;
  }
  switch (op) {
    case F_DOT: scal[S_DOT] = t[0]; break;
    case F_RZ_FIRST: scal[S_RTZ] = t[0]; break;
    case F_RZ: {
      const double old = scal[S_RTZ];
      scal[S_OLDRTZ] = old;
      scal[S_RTZ] = t[0];
      scal[S_BETA] = t[0] / old;
      break;
    }
    case F_PAP:
      scal[S_PAP] = t[0];
      scal[S_ALPHA] = scal[S_RTZ] / t[0];
      break;
    case F_NORMR2:
      // Deferred stop test of a tolerance-0 solve (CG.cpp:100 with tolerance
      // 0): with normr0 finite and positive, normr/normr0 > 0 holds exactly
      // when normr^2 > 0.  The first norm that fails (zero or NaN) raises
      // S_STOP; from then on k_cg_update leaves x and r alone and no norm
      // is recorded.  Without a slot the flag is cleared (solver preamble).
      if (slot) {
        if (scal[S_STOP] != 0.0) break;
        *slot = t[0];
        if (!(t[0] > 0.0)) scal[S_STOP] = 1.0;
      } else {
        scal[S_STOP] = 0.0;
      }
      scal[S_NORMR2] = t[0];
      break;
    case F_FCG: {
      const double old = scal[S_RTZ];
      scal[S_OLDRTZ] = old;
      scal[S_RTZ] = t[0];
      scal[S_BETA] = -scal[S_ALPHA] * t[NV - 1] / old;
      break;
    }
  }
}

// Streaming reductions give each block one chunk of kChunk * kBlock
// consecutive elements, kChunk per thread, each a coalesced wave access.
constexpr int kChunk = 4;
__device__ __forceinline__ int64_t chunk_index(int j) {
  return ((int64_t)blockIdx.x * kChunk + j) * kBlock + threadIdx.x;
}

// ---------------------------------------------------------------------------
// sparse kernels (one thread per row, 4 slices per block)

// y = A x over rows [0, nrows); ComputeSPMV_ref.cpp:61-70
template <class VT, class XT, class ACC, bool CI>
__global__ __launch_bounds__(kBlock) HPCGX_SPARSE_OCCUPANCY(VT, CI) void k_spmv(const ColIdx ci,
                                                 const VT *__restrict__ vals,
                                                 const XT *__restrict__ x, XT *__restrict__ y,
                                                 int64_t nrows) {
  const int64_t row = tile_id() * kBlock + threadIdx.x;
  if (row >= nrows) return;
  const size_t b = ell_base(row);
  const int slice = slice_of<true>(row);
  ACC v[kSlots];
  row_vals<ACC, true, kAllSlots>(vals, row, v);
  y[row] = (XT)with_cols<CI, true>(ci, slice, [&](auto cm, const int *base) {
    constexpr int CM = decltype(cm)::value;
    ACC sum = 0;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int c = col_at<CM, true>(ci, row, base, b, s);
      sum += v[s] * (ACC)x[c];
    }
    return sum;
  });
}

// Slots whose neighbour lies in a block swept before block k in the default
// order (the neighbour at offset d has class c ^ oddness(d)).
__host__ __device__ constexpr uint32_t swept_before_mask(int k) {
  int pos[kColours] = {};
  for (int j = 0; j < kColours; ++j) pos[kDefaultOrder[j]] = j;
  uint32_t m = 0;
  for (int s = 0; s < kSlots; ++s) {
    const int odd = ((s % 3) != 1) | ((((s / 3) % 3) != 1) << 1) | (((s / 9) != 1) << 2);
    if (pos[kDefaultOrder[k] ^ odd] < k) m |= 1u << s;
  }
  return m;
}

// The masks keep or skip the slots of a unit together (each unit's
// neighbours lie in one colour block).
constexpr bool masks_keep_units(int W) {
  for (int k = 0; k < kColours; ++k)
    for (int u = 0; u < val_units(W); ++u) {
      const uint32_t um = unit_mask(val_unit(W, u)), m = swept_before_mask(k) & um;
      if (m != 0 && m != um) return false;
    }
  return true;
}
static_assert(masks_keep_units(2) && masks_keep_units(4), "zero-start masks split a value unit");

// One Gauss-Seidel step on the rows of one colour block; ComputeSYMGS_ref.cpp:72-79:
// sum = r_i - sum_j a_ij x_j (diagonal included) + x_i a_ii ; x_i = sum / a_ii.
// ZK >= 0: the step of block ZK in the first forward sweep from x = 0 -- only
// neighbours in blocks swept before ZK can be nonzero; the other terms are
// exact zeros and are not read, and neither is x_i (x may hold anything on
// entry).  The skipped set is a compile-time constant, so the remaining loads
// issue back to back.
// FINE: the finest level (streamed matrix loads; its own instantiation, so
// profiles show the dominant launches separately from the coarse levels').
// ZK == 0 reads no neighbour at all (x_i = r_i / a_ii), which holds for the
// first block of any order, general levels included.
// GEN: general level, the diagonal's slot is dslot[row] (stencil: slot 13).
// gs_row is the update of one row; it returns r_i * x_i (new) for the fused r.z.
template <class VT, class XT, class ACC, int ZK, bool FINE, bool CI, bool GEN, bool SAFE>
__device__ __forceinline__ double gs_row(const ColIdx &ci, const VT *__restrict__ vals,
                                         const XT *__restrict__ r, XT *__restrict__ x, int64_t row,
                                         const uint8_t *__restrict__ dslot) {
  constexpr uint32_t mask = ZK >= 0 ? swept_before_mask(ZK) : kAllSlots;
  const size_t b = ell_base(row);
  const int slice = slice_of<true>(row);
  ACC vr[kSlots];
  row_vals<ACC, FINE, mask>(vals, row, vr);
  ACC d;
  if constexpr (!GEN && ((mask >> kDiagSlot) & 1u)) d = vr[kDiagSlot];
  else d = ld_val<ACC, FINE>(vals + val_pos<val_width<VT>>(row, GEN ? (int)dslot[row] : kDiagSlot));
  auto row_sum = [&](auto cm, const int *base) {
    constexpr int CM = decltype(cm)::value;
    ACC sum = (ACC)r[row];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      if (!((mask >> s) & 1u)) continue;
      const int c = col_at<CM, FINE>(ci, row, base, b, s);
      const ACC v = vr[s];
      const ACC xc = (ACC)x[c];
      // SAFE (from x = 0 into a vector that may hold anything): a zero entry (a neighbour
      // outside the grid) may point at a row not written yet in this sweep,
      // whose old value is not trusted: it contributes 0 * 0, as the
      // reference's zeroed x gives.  The solvers clear their level-0
      // buffer when a solve begins, so their level-0 sweeps skip the select.
      if constexpr (ZK >= 0 && SAFE) sum -= v * (v != (ACC)0 ? xc : (ACC)0);
      else sum -= v * xc;
    }
    return sum;
  };
  ACC sum = with_cols<CI, true>(ci, slice, row_sum);
  if (ZK < 0) sum += (ACC)x[row] * d;
  const XT xn = (XT)(sum / d);
  x[row] = xn;
  return (double)r[row] * (double)xn;
}

// RZ: this step writes the final value of its rows in the V-cycle (the
// post-smoother's last forward block and its backward steps); the block
// also leaves its partial sum of r_i x_i in rz[blockIdx.x], so the solver's
// r.z needs no pass of its own.
template <class VT, class XT, class ACC, int ZK, bool FINE, bool CI = false, bool GEN = false, bool RZ = false,
          bool SAFE = false>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(1, SparseOccupancy<VT, CI, ZK < 0 ? kSlots : __builtin_popcount(swept_before_mask(ZK))>::max_waves)))
void k_gs(const ColIdx ci, const VT *__restrict__ vals,
                                               const XT *__restrict__ r, XT *__restrict__ x,
                                               int64_t row0, int64_t nrows,
                                               const uint8_t *__restrict__ dslot, double *rz = nullptr) {
  static_assert(!GEN || ZK <= 0, "general levels: full rows, or the first block from x = 0 (diagonal only)");
  static_assert(!RZ || ZK < 0, "r.z is taken from the post-smoother");
  // blocks of 256 threads, or of 64 for the small coarse fp64 steps (one wave
  // per block spreads them over more CUs); the r.z form always takes 256
#ifndef HPCGX_FINE_ROT
#define HPCGX_FINE_ROT 1
#endif
  const int64_t t = tile_id<HPCGX_FINE_ROT && FINE && sizeof(VT) == 8>() * (int64_t)blockDim.x + threadIdx.x;
  if constexpr (RZ) {
    // every thread reaches the block reduction; rows past the block add 0
    double acc = 0.0;
    if (t < nrows) acc = gs_row<VT, XT, ACC, ZK, FINE, CI, GEN, SAFE>(ci, vals, r, x, row0 + t, dslot);
    const double v[1] = {acc};
    block_partials<1>(v, rz);
    return;
  }
  if (t >= nrows) return;
  gs_row<VT, XT, ACC, ZK, FINE, CI, GEN, SAFE>(ci, vals, r, x, row0 + t, dslot);
}

// Residual at the injected rows and restriction, fused (ComputeMG_ref.cpp:49-51 +
// ComputeRestriction_ref.cpp:51): rc[inj[t]] = r_f - (A x)_f for the t-th
// injected row f = row0 + t (frow[t] when frow is given).  (A x)_f is summed
// from 0 like SpMV, then subtracted.  SCALE: fp16 V-cycle, times sigma[t].
template <class VT, class XT, class ACC, bool SCALE, bool CI = false>
__global__ __launch_bounds__(kBlock) HPCGX_SPARSE_OCCUPANCY(VT, CI) void k_restrict(const ColIdx ci, const VT *__restrict__ vals,
                                                     const XT *__restrict__ r,
                                                     const XT *__restrict__ x,
                                                     XT *__restrict__ rc, const int *__restrict__ inj,
                                                     const float *__restrict__ sigma, int64_t row0,
                                                     int64_t count, const int *__restrict__ frow) {
  const int64_t t = tile_id() * kBlock + threadIdx.x;
  if (t >= count) return;
  const int64_t row = frow ? (int64_t)frow[t] : row0 + t;
  const size_t b = ell_base(row);
  ACC v[kSlots];
  row_vals<ACC, true, kAllSlots>(vals, row, v);
  auto row_ax = [&](auto cm, const int *base) {
    constexpr int CM = decltype(cm)::value;
    ACC ax = 0;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int c = col_at<CM, true>(ci, row, base, b, s);
      ax += v[s] * (ACC)x[c];
    }
    return ax;
  };
  ACC ax;
  if (frow) {  // rows from a list: slices differ across the wave
    const int slice = slice_of<false>(row);
    ax = with_cols<CI, false>(ci, slice, row_ax);
  } else {
    const int slice = slice_of<true>(row);
    ax = with_cols<CI, true>(ci, slice, row_ax);
  }
  ACC res = (ACC)r[row] - ax;
  if (SCALE) res = res * (ACC)sigma[t];
  rc[inj[t]] = (XT)res;
}

// x_f += xc at the injected rows (ComputeProlongation_ref.cpp:49).
template <class XT, class ACC, bool SCALE>
__global__ __launch_bounds__(kBlock) void k_prolong(XT *__restrict__ x, const XT *__restrict__ xc,
                                                    const int *__restrict__ inj,
                                                    const float *__restrict__ sigma, int64_t row0,
                                                    int64_t count, const int *__restrict__ frow) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int64_t row = frow ? (int64_t)frow[t] : row0 + t;
  ACC v = (ACC)xc[inj[t]];
  if (SCALE) v = v * (ACC)sigma[t];
  x[row] = (XT)((ACC)x[row] + v);
}

// ---------------------------------------------------------------------------
// element-wise helpers

// w = alpha x + beta y with the branches of ComputeWAXPBY_ref.cpp:52-67
__device__ __forceinline__ double waxpby1(double alpha, double x, double beta, double y) {
  if (alpha == 1.0) return x + beta * y;
  if (beta == 1.0) return alpha * x + y;
  return alpha * x + beta * y;
}

}  // namespace hpcgx
