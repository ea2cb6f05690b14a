// The reference's problem-setup hooks, GenerateProblem (GenerateProblem.cpp:44-52)
// and SetupHalo (SetupHalo.cpp:35-42), for the one-rank drop-in build.
//
// Both produce exactly what GenerateProblem_ref / SetupHalo_ref produce
// (CheckProblem.cpp:51-147 asserts it on every level), in the time the host
// memory system allows:
// * the matrix arrays are one allocation each, the row pointers indexing into
//   them (the HPCG_CONTIGUOUS_ARRAYS form of GenerateProblem_ref.cpp:116-124)
//   instead of three allocations per row, on 2 MiB-aligned memory that the
//   kernel may back with huge pages;
// * host threads fill disjoint ranges of rows (first touch, so each thread
//   faults in what it writes), with no lock per row: each thread counts its
//   own nonzeros;
// * the global-to-local map is built only where something reads it: the
//   reference's MPI halo setup (SetupHalo_ref.cpp:137,150) and its
//   HPCG_DETAILED_DEBUG prints (CheckProblem.cpp:89, GenerateProblem_ref.cpp:147,
//   SetupHalo_ref.cpp:86).  A one-rank build without those never reads it,
//   and its 7M single-threaded inserts at 192^3 (0.3 s; an unordered map
//   cannot take concurrent inserts) were most of the setup time.  Where it is
//   built, one more thread builds it while the others fill the arrays;
// * SetupHalo's one-rank index copy (SetupHalo_ref.cpp:57-65) runs on the
//   same threads.
// * the HIP runtime, the device and the hpcgx context start on one more
//   thread at the first GenerateProblem call (hpcgx_device_init, then
//   hpcgx_create), so that start-up overlaps the host fill instead of opening
//   OptimizeProblem (runtime 38 ms on a warm box, 175 ms on a fresh one,
//   context 31 ms, measured); OptimizeProblem takes the context.
// Setup time is charged to the HPCG rating (times[9], ReportResults.cpp:353),
// which is why it matters here.  DeleteMatrix (hpcg_dropin.hpp) frees this
// storage through HPCG_MatrixStorageRelease.
#include <sys/mman.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "dropin_state.hpp"

namespace {

#if !defined(HPCG_NO_MPI) || defined(HPCG_DETAILED_DEBUG)
constexpr bool kBuildGlobalToLocalMap = true;
#else
constexpr bool kBuildGlobalToLocalMap = false;
#endif

struct Storage {
  void *vals, *indG, *indL;
};

// The device start-up thread and the context it opens; joined by
// OptimizeProblem, or at exit.
struct DeviceInit {
  std::mutex mu;
  std::thread th;
  bool started = false;
  hpcgx_ctx *ctx = nullptr;  // opened by the thread, taken by OptimizeProblem
  ~DeviceInit() {
    if (th.joinable()) th.join();
  }
} g_device_init;
std::mutex g_store_mu;
std::unordered_map<const void *, Storage> g_store;  // keyed by the values block

// Host threads for the setup loops: OMP_NUM_THREADS when set (the thread
// budget the reference's own OpenMP loops would get), else the hardware's.
unsigned setup_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (const char *e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) n = (unsigned)v;
  }
  return std::max(1u, std::min(n, 64u));
}

// Runs f(lo, hi) on contiguous ranges of [0, n), one per thread.
template <class F>
void run_ranges(int64_t n, unsigned nt, F f) {
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    if (lo < hi) th.emplace_back([=, &f] { f(t, lo, hi); });
  }
  for (auto &t : th) t.join();
}

void *big_alloc(size_t bytes) {
  const size_t align = (size_t)2 << 20;
  void *p = nullptr;
  if (posix_memalign(&p, align, std::max(bytes, (size_t)1)) != 0) return nullptr;
  if (bytes >= align) (void)madvise(p, bytes, MADV_HUGEPAGE);  // a hint; ignored where unsupported
  return p;
}

}  // namespace

int dropin_device() {
  const char *dev = std::getenv("HPCG_DEVICE");
  return dev ? std::atoi(dev) : 0;
}

hpcgx_ctx *dropin_take_started_context() {
  std::lock_guard<std::mutex> lk(g_device_init.mu);
  if (g_device_init.th.joinable()) g_device_init.th.join();
  hpcgx_ctx *ctx = g_device_init.ctx;
  g_device_init.ctx = nullptr;
  return ctx;
}

extern "C" int HPCG_MatrixStorageRelease(SparseMatrix *A) {
  if (!A || !A->matrixValues || A->localNumberOfRows <= 0) return 0;
  Storage s;
  {
    std::lock_guard<std::mutex> lk(g_store_mu);
    auto it = g_store.find(A->matrixValues[0]);
    if (it == g_store.end()) return 0;
    s = it->second;
    g_store.erase(it);
  }
  std::free(s.vals);
  std::free(s.indG);
  std::free(s.indL);
  return 1;
}

// GenerateProblem.cpp:44-52 -> GenerateProblem_ref.cpp:131-183: the
// 27-point stencil of this rank's nx*ny*nz block of the global grid, values
// 26 on the diagonal and -1 off it, entries in ascending column order,
// b_i = 26 - (nnz_i - 1), x = 0, xexact = 1.
void GenerateProblem(SparseMatrix &A, Vector *b, Vector *x, Vector *xexact) {
  {
    std::lock_guard<std::mutex> lk(g_device_init.mu);
    if (!g_device_init.started) {
      g_device_init.started = true;
      const int dev = dropin_device();
      g_device_init.th = std::thread([dev] {
        hpcgx_ctx *ctx = nullptr;
        if (hpcgx_device_init(dev) == HPCGX_OK && hpcgx_create(dev, &ctx) == HPCGX_OK) g_device_init.ctx = ctx;
      });
    }
  }
  const global_int_t nx = A.geom->nx, ny = A.geom->ny, nz = A.geom->nz;
  const global_int_t gnx = A.geom->gnx, gny = A.geom->gny, gnz = A.geom->gnz;
  const global_int_t gix0 = A.geom->gix0, giy0 = A.geom->giy0, giz0 = A.geom->giz0;
  const local_int_t n = (local_int_t)(nx * ny * nz);
  assert(n > 0);
  const int per_row = 27;
  const size_t slots = (size_t)n * per_row;

  char *nonzerosInRow = new char[n];
  global_int_t **mtxIndG = new global_int_t *[n];
  local_int_t **mtxIndL = new local_int_t *[n];
  double **matrixValues = new double *[n];
  double **matrixDiagonal = new double *[n];
  double *vals = static_cast<double *>(big_alloc(slots * sizeof(double)));
  global_int_t *indG = static_cast<global_int_t *>(big_alloc(slots * sizeof(global_int_t)));
  local_int_t *indL = static_cast<local_int_t *>(big_alloc(slots * sizeof(local_int_t)));
  if (!vals || !indG || !indL) {
    std::fprintf(stderr, "hpcg dropin: GenerateProblem could not allocate %zu matrix slots\n", slots);
    std::abort();
  }
  {
    std::lock_guard<std::mutex> lk(g_store_mu);
    g_store[vals] = Storage{vals, indG, indL};
  }
  if (b) InitializeVector(*b, n);
  if (x) InitializeVector(*x, n);
  if (xexact) InitializeVector(*xexact, n);
  double *bv = b ? b->values : nullptr, *xv = x ? x->values : nullptr, *xev = xexact ? xexact->values : nullptr;
  A.localToGlobalMap.resize(n);
  global_int_t *l2g = A.localToGlobalMap.data();

  // the global-to-local map on its own thread, in row order, where it is read
  A.globalToLocalMap.clear();
  std::thread map_thread;
  if (kBuildGlobalToLocalMap)
    map_thread = std::thread([&A, n, nx, ny, gnx, gny, gix0, giy0, giz0] {
      A.globalToLocalMap.reserve(n);
      for (local_int_t i = 0; i < n; ++i) {
        const global_int_t ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
        A.globalToLocalMap.emplace((giz0 + iz) * gnx * gny + (giy0 + iy) * gnx + gix0 + ix, i);
      }
    });

  const unsigned nt = setup_threads();
  std::vector<global_int_t> counts(nt, 0);
  run_ranges(n, nt, [&](unsigned t, int64_t lo, int64_t hi) {
    global_int_t cnt = 0;
    for (int64_t i = lo; i < hi; ++i) {
      const global_int_t ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
      const global_int_t gix = gix0 + ix, giy = giy0 + iy, giz = giz0 + iz;
      const global_int_t row = giz * gnx * gny + giy * gnx + gix;
      double *v = vals + (size_t)i * per_row;
      global_int_t *g = indG + (size_t)i * per_row;
      matrixValues[i] = v;
      mtxIndG[i] = g;
      mtxIndL[i] = indL + (size_t)i * per_row;
      l2g[i] = row;
      char k = 0;
      for (int sz = -1; sz <= 1; ++sz) {
        if (giz + sz < 0 || giz + sz >= gnz) continue;
        for (int sy = -1; sy <= 1; ++sy) {
          if (giy + sy < 0 || giy + sy >= gny) continue;
          for (int sx = -1; sx <= 1; ++sx) {
            if (gix + sx < 0 || gix + sx >= gnx) continue;
            const global_int_t col = row + sz * gnx * gny + sy * gnx + sx;
            if (col == row) {
              matrixDiagonal[i] = v + k;
              v[k] = 26.0;
            } else {
              v[k] = -1.0;
            }
            g[k] = col;
            ++k;
          }
        }
      }
      nonzerosInRow[i] = k;
      cnt += k;
      if (bv) bv[i] = 26.0 - (double)(k - 1);
      if (xv) xv[i] = 0.0;
      if (xev) xev[i] = 1.0;
    }
    counts[t] = cnt;
  });
  if (map_thread.joinable()) map_thread.join();
  global_int_t nnz = 0;
  for (global_int_t c : counts) nnz += c;
  assert(nnz > 0);

  A.title = 0;
  A.totalNumberOfRows = gnx * gny * gnz;
  A.totalNumberOfNonzeros = nnz;  // one rank: the local count is the total
  A.localNumberOfRows = n;
  A.localNumberOfColumns = n;
  A.localNumberOfNonzeros = (local_int_t)nnz;
  A.nonzerosInRow = nonzerosInRow;
  A.mtxIndG = mtxIndG;
  A.mtxIndL = mtxIndL;
  A.matrixValues = matrixValues;
  A.matrixDiagonal = matrixDiagonal;
}

// SetupHalo.cpp:35-42 -> SetupHalo_ref.cpp:57-65 (one rank): every column is
// a local row, so the local indices are the global ones.  Works for any rows.
void SetupHalo(SparseMatrix &A) {
  const local_int_t n = A.localNumberOfRows;
  const char *nr = A.nonzerosInRow;
  global_int_t **g = A.mtxIndG;
  local_int_t **l = A.mtxIndL;
  run_ranges(n, setup_threads(), [&](unsigned, int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      const int cnt = nr[i];
      for (int j = 0; j < cnt; ++j) l[i][j] = (local_int_t)g[i][j];
    }
  });
}
