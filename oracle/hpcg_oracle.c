/*
 * hpcg_oracle.c -- CPU restatement of the HPCG 3.1 reference arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see
 * hpcg_oracle.h for who may load it and how each function is pinned.
 *
 * Build with -ffp-contract=off: the reference's x86-64 build performs every
 * multiply and add separately (no FMA), and bit-exact parity depends on it.
 * Citations are file:line in the reference's src/ directory.
 */
#include "hpcg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NSLOT 27

struct orc_level {
  int nx, ny, nz, n;
  int64_t nnz;
  unsigned char *nnz_row; /* entries per row (8/12/18/27) */
  int *cols;              /* [n*27] ascending column per row */
  double *vals;           /* [n*27] */
  int *diag_slot;         /* position of the diagonal inside its row */
  double *kappa;          /* per-cell diffusivity, NULL for the HPCG matrix */
  int *f2c;               /* [coarse->n] fine row injected into coarse row */
  int order[8];           /* parity class swept k-th (forward) */
  int *crow;              /* rows grouped by sweep position */
  int *cstart;            /* [ncol+1] first entry of each colour in crow */
  int ncol;               /* colours: 8 parity classes, or a general colouring */
  int general;            /* built from arbitrary rows (orc_create_rows) */
  /* mixed-precision copies (built on first use) */
  float *vals32;
  uint16_t *vals16; /* diagonally scaled, fp16 bits */
  float *dinv;      /* 1/sqrt(a_ii) */
  float *sigma;     /* [coarse->n] sqrt(d_f[f2c]/d_c) */
  /* V-cycle work vectors */
  double *Axf, *rc, *xc;
  float *Axf32, *rc32, *xc32;
  struct orc_level *coarse;
};

static int parity_class(const orc_level *L, int i) {
  int ix = i % L->nx, iy = (i / L->nx) % L->ny, iz = i / (L->nx * L->ny);
  return (ix & 1) | ((iy & 1) << 1) | ((iz & 1) << 2);
}

/* Rows grouped by sweep position, ascending inside a colour. */
static void build_colour_lists(orc_level *L) {
  free(L->cstart);
  L->cstart = (int *)malloc(9 * sizeof(int));
  L->ncol = 8;
  int pos_of_class[8];
  for (int k = 0; k < 8; ++k) pos_of_class[L->order[k]] = k;
  int count[8] = {0};
  for (int i = 0; i < L->n; ++i) count[pos_of_class[parity_class(L, i)]]++;
  L->cstart[0] = 0;
  for (int k = 0; k < 8; ++k) L->cstart[k + 1] = L->cstart[k] + count[k];
  int fill[8];
  for (int k = 0; k < 8; ++k) fill[k] = L->cstart[k];
  for (int i = 0; i < L->n; ++i) L->crow[fill[pos_of_class[parity_class(L, i)]]++] = i;
}

/* Matrix of one level.  Row layout and values restate
 * GenerateProblem_ref.cpp:131-183 for a single rank (global == local):
 * neighbours visited z-major, then y, then x, each from -1 to +1, skipping
 * points outside the grid; diagonal 26, off-diagonals -1.
 * With kappa: a_ij = -(k_i+k_j)*0.5 and
 *             a_ii = sum_{j in N(i), ascending} (k_i+k_j)*0.5 + (26-n_i)*k_i. */
static void generate_level(orc_level *L) {
  const int nx = L->nx, ny = L->ny, nz = L->nz;
  int64_t nnz = 0;
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const int i = (iz * ny + iy) * nx + ix;
        int *c = L->cols + (int64_t)i * NSLOT;
        double *v = L->vals + (int64_t)i * NSLOT;
        int k = 0;
        for (int sz = -1; sz <= 1; ++sz) {
          if (iz + sz < 0 || iz + sz >= nz) continue;
          for (int sy = -1; sy <= 1; ++sy) {
            if (iy + sy < 0 || iy + sy >= ny) continue;
            for (int sx = -1; sx <= 1; ++sx) {
              if (ix + sx < 0 || ix + sx >= nx) continue;
              const int col = i + sz * nx * ny + sy * nx + sx;
              c[k] = col;
              if (col == i) {
                L->diag_slot[i] = k;
                v[k] = 26.0;
              } else {
                v[k] = L->kappa ? -((L->kappa[i] + L->kappa[col]) * 0.5) : -1.0;
              }
              ++k;
            }
          }
        }
        L->nnz_row[i] = (unsigned char)k;
        if (L->kappa) {
          double s = 0.0;
          for (int j = 0; j < k; ++j)
            if (c[j] != i) s += (L->kappa[i] + L->kappa[c[j]]) * 0.5;
          v[L->diag_slot[i]] = s + (double)(26 - (k - 1)) * L->kappa[i];
        }
        for (int j = k; j < NSLOT; ++j) { c[j] = i; v[j] = 0.0; }
        nnz += k;
      }
  L->nnz = nnz;
}

orc_level *orc_create(int nx, int ny, int nz, int levels, const double *kappa) {
  if (nx < 1 || ny < 1 || nz < 1 || levels < 1) return NULL;
  orc_level *L = (orc_level *)calloc(1, sizeof(orc_level));
  L->nx = nx; L->ny = ny; L->nz = nz;
  L->n = nx * ny * nz;
  const int64_t n = L->n;
  L->nnz_row = (unsigned char *)malloc(n);
  L->cols = (int *)malloc(n * NSLOT * sizeof(int));
  L->vals = (double *)malloc(n * NSLOT * sizeof(double));
  L->diag_slot = (int *)malloc(n * sizeof(int));
  L->crow = (int *)malloc(n * sizeof(int));
  if (kappa) {
    L->kappa = (double *)malloc(n * sizeof(double));
    memcpy(L->kappa, kappa, n * sizeof(double));
  }
  generate_level(L);
  /* default sweep order: the device's (parity classes 7,3,4,5,2,6,1,0) */
  {
    static const int dflt[8] = {7, 3, 4, 5, 2, 6, 1, 0};
    for (int k = 0; k < 8; ++k) L->order[k] = dflt[k];
  }
  build_colour_lists(L);
  L->Axf = (double *)calloc(n, sizeof(double));
  L->Axf32 = (float *)calloc(n, sizeof(float));
  if (levels > 1 && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0) {
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    const int nc = cx * cy * cz;
    /* Injection operator, GenerateCoarseProblem.cpp:71-82. */
    L->f2c = (int *)malloc((size_t)nc * sizeof(int));
    for (int izc = 0; izc < cz; ++izc)
      for (int iyc = 0; iyc < cy; ++iyc)
        for (int ixc = 0; ixc < cx; ++ixc)
          L->f2c[(izc * cy + iyc) * cx + ixc] = ((2 * izc) * ny + 2 * iyc) * nx + 2 * ixc;
    double *kc = NULL;
    if (kappa) {
      kc = (double *)malloc((size_t)nc * sizeof(double));
      for (int i = 0; i < nc; ++i) kc[i] = kappa[L->f2c[i]];
    }
    L->coarse = orc_create(cx, cy, cz, levels - 1, kc);
    free(kc);
    L->rc = (double *)calloc(nc, sizeof(double));
    L->xc = (double *)calloc(nc, sizeof(double));
    L->rc32 = (float *)calloc(nc, sizeof(float));
    L->xc32 = (float *)calloc(nc, sizeof(float));
  }
  return L;
}

void orc_destroy(orc_level *L) {
  if (!L) return;
  orc_destroy(L->coarse);
  free(L->nnz_row); free(L->cols); free(L->vals); free(L->diag_slot); free(L->crow); free(L->cstart);
  free(L->kappa); free(L->f2c); free(L->Axf); free(L->rc); free(L->xc);
  free(L->Axf32); free(L->rc32); free(L->xc32);
  free(L->vals32); free(L->vals16); free(L->dinv); free(L->sigma);
  free(L);
}

orc_level *orc_coarse(orc_level *L) { return L->coarse; }
int orc_nrow(const orc_level *L) { return L->n; }
int64_t orc_nnz(const orc_level *L) { return L->nnz; }
void orc_dims(const orc_level *L, int *d) { d[0] = L->nx; d[1] = L->ny; d[2] = L->nz; }
const unsigned char *orc_nnz_row(const orc_level *L) { return L->nnz_row; }
const int *orc_cols(const orc_level *L) { return L->cols; }
const double *orc_vals(const orc_level *L) { return L->vals; }
const int *orc_f2c(const orc_level *L) { return L->f2c; }
const double *orc_kappa(const orc_level *L) { return L->kappa; }

/* b, x, xexact: GenerateProblem_ref.cpp:178-180. */
void orc_rhs(const orc_level *L, double *b, double *x, double *xexact) {
  for (int i = 0; i < L->n; ++i) {
    if (b) b[i] = 26.0 - (double)(L->nnz_row[i] - 1);
    if (x) x[i] = 0.0;
    if (xexact) xexact[i] = 1.0;
  }
}

/* CopyMatrixDiagonal / ReplaceMatrixDiagonal, SparseMatrix.hpp:126-145. */
void orc_get_diagonal(const orc_level *L, double *d) {
  for (int i = 0; i < L->n; ++i) d[i] = L->vals[(int64_t)i * NSLOT + L->diag_slot[i]];
}
void orc_set_diagonal(orc_level *L, const double *d) {
  for (int i = 0; i < L->n; ++i) L->vals[(int64_t)i * NSLOT + L->diag_slot[i]] = d[i];
  free(L->vals32); L->vals32 = NULL;
  free(L->vals16); L->vals16 = NULL;
  free(L->sigma); L->sigma = NULL;
}

void orc_set_colour_order(orc_level *L, const int order[8]) {
  for (; L; L = L->coarse) {
    if (L->general) continue;
    memcpy(L->order, order, 8 * sizeof(int));
    build_colour_lists(L);
  }
}

/* First-fit greedy colouring in row order, OptimizeProblem.cpp:43-79. */
void orc_colour_greedy(const orc_level *L, int *colour, int *ncolours) {
  const int n = L->n;
  for (int i = 0; i < n; ++i) colour[i] = -1;
  int total = 1;
  colour[0] = 0;
  for (int i = 1; i < n; ++i) {
    int used[NSLOT + 1];
    memset(used, 0, sizeof(used));
    const int *c = L->cols + (int64_t)i * NSLOT;
    for (int j = 0; j < L->nnz_row[i]; ++j)
      if (c[j] < i) used[colour[c[j]]] = 1;
    int pick = -1;
    for (int k = 0; k < total; ++k)
      if (!used[k]) { pick = k; break; }
    if (pick < 0) pick = total++;
    colour[i] = pick;
  }
  *ncolours = total;
}

/* Couplings of row i in both directions: its off-diagonal entries plus the
 * rows that hold column i (a symmetric coupling counts twice). */
static int couplings(const orc_level *L, const int *tcnt, int i) {
  int d = tcnt[i + 1] - tcnt[i];
  for (int k = 0; k < L->nnz_row[i]; ++k)
    if (L->cols[(int64_t)i * NSLOT + k] != i) ++d;
  return d;
}

/* Breadth-first pass from s over its component (rows marked with a fresh
 * stamp): the depth of its last level, and the row of that level with the
 * fewest couplings (the first such row reached). */
static void probe_bfs(const orc_level *L, const int *tcnt, const int *tj, int s, int *q, int *stamp,
                      int *level, int *stamp_id, int *far, int *depth) {
  const int id = (*stamp_id)++;
  int h = 0, t = 0;
  q[t++] = s;
  stamp[s] = id;
  level[s] = 0;
  while (h < t) {
    const int i = q[h++];
    for (int k = 0; k < L->nnz_row[i]; ++k) {
      const int c = L->cols[(int64_t)i * NSLOT + k];
      if (c != i && stamp[c] != id) { stamp[c] = id; level[c] = level[i] + 1; q[t++] = c; }
    }
    for (int r = tcnt[i]; r < tcnt[i + 1]; ++r)
      if (stamp[tj[r]] != id) { stamp[tj[r]] = id; level[tj[r]] = level[i] + 1; q[t++] = tj[r]; }
  }
  *depth = level[q[t - 1]];
  int u = t - 1;
  while (u > 0 && level[q[u - 1]] == *depth) --u;
  int best = q[u], bd = couplings(L, tcnt, q[u]);
  for (++u; u < t; ++u) {
    const int d = couplings(L, tcnt, q[u]);
    if (d < bd) { bd = d; best = q[u]; }
  }
  *far = best;
}

/* Colouring of a general level: first fit in breadth-first order over the
 * symmetrised pattern.  The neighbours of row i are its own entries (row
 * order, the diagonal skipped) followed by the rows j that hold column i
 * (ascending j).  Components are taken in order of their smallest row s.
 * The start row is pseudo-peripheral (George and Liu, ACM TOMS 5(3), 1979):
 * a breadth-first pass from x = s gives its last level, and y, the row of
 * that level with the fewest couplings; while a pass from y ends deeper than
 * the pass from x, x moves to y (at most 8 moves); at the end x gives way to
 * y when y has fewer couplings.  On a renumbered 27-point stencil this start
 * is a corner of the grid and the first fit needs 8 colours, the minimum,
 * where natural order needs 17.  Row i takes the smallest colour not held by
 * an already coloured neighbour.  order[] (may be NULL) receives the
 * sequence. */
void orc_colour_general_order(const orc_level *L, int *colour, int *ncolours, int *order) {
  const int n = L->n;
  int *tcnt = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < L->nnz_row[j]; ++k) {
      const int c = L->cols[(int64_t)j * NSLOT + k];
      if (c != j) tcnt[c + 1]++;
    }
  for (int i = 0; i < n; ++i) tcnt[i + 1] += tcnt[i];
  int *tj = (int *)malloc((size_t)(tcnt[n] > 0 ? tcnt[n] : 1) * sizeof(int));
  int *fill = (int *)malloc((size_t)n * sizeof(int));
  memcpy(fill, tcnt, (size_t)n * sizeof(int));
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < L->nnz_row[j]; ++k) {
      const int c = L->cols[(int64_t)j * NSLOT + k];
      if (c != j) tj[fill[c]++] = j;
    }
  int *seq = (int *)malloc((size_t)n * sizeof(int));
  int *probe = (int *)malloc((size_t)n * sizeof(int));
  int *stamp = (int *)malloc((size_t)n * sizeof(int));
  int *level = (int *)malloc((size_t)n * sizeof(int));
  char *seen = (char *)calloc((size_t)n, 1);
  int stamp_id = 0;
  for (int i = 0; i < n; ++i) stamp[i] = -1;
  int len = 0;
  for (int s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    /* probe passes (s0's component is wholly unseen) */
    int x = s0, fx, ex, fy, ey;
    probe_bfs(L, tcnt, tj, x, probe, stamp, level, &stamp_id, &fx, &ex);
    for (int it = 0; it < 8; ++it) {
      probe_bfs(L, tcnt, tj, fx, probe, stamp, level, &stamp_id, &fy, &ey);
      if (ey <= ex) {
        if (couplings(L, tcnt, fx) < couplings(L, tcnt, x)) x = fx;
        break;
      }
      x = fx;
      ex = ey;
      fx = fy;
    }
    const int start = x;
    /* the pass that gives the order */
    int head = len;
    seq[len++] = start;
    seen[start] = 1;
    while (head < len) {
      const int i = seq[head++];
      for (int k = 0; k < L->nnz_row[i]; ++k) {
        const int c = L->cols[(int64_t)i * NSLOT + k];
        if (c != i && !seen[c]) { seen[c] = 1; seq[len++] = c; }
      }
      for (int q = tcnt[i]; q < tcnt[i + 1]; ++q)
        if (!seen[tj[q]]) { seen[tj[q]] = 1; seq[len++] = tj[q]; }
    }
  }
  int *mark = (int *)malloc(((size_t)n + 1) * sizeof(int));
  for (int i = 0; i <= n; ++i) mark[i] = -1;
  for (int i = 0; i < n; ++i) colour[i] = -1;
  int total = 0;
  for (int q = 0; q < n; ++q) {
    const int i = seq[q];
    for (int k = 0; k < L->nnz_row[i]; ++k) {
      const int c = L->cols[(int64_t)i * NSLOT + k];
      if (c != i && colour[c] >= 0) mark[colour[c]] = i;
    }
    for (int r = tcnt[i]; r < tcnt[i + 1]; ++r)
      if (colour[tj[r]] >= 0) mark[colour[tj[r]]] = i;
    int pick = 0;
    while (mark[pick] == i) ++pick;
    colour[i] = pick;
    if (pick + 1 > total) total = pick + 1;
  }
  *ncolours = total;
  if (order) memcpy(order, seq, (size_t)n * sizeof(int));
  free(tcnt); free(tj); free(fill); free(seq); free(probe); free(stamp); free(level); free(seen); free(mark);
}

void orc_colour_general(const orc_level *L, int *colour, int *ncolours) {
  orc_colour_general_order(L, colour, ncolours, NULL);
}

/* Rows grouped by colour (colours swept in ascending order, rows ascending
 * inside a colour).  Returns -1 unless every colour is independent: no row
 * has an off-diagonal entry in a row of its own colour. */
int orc_set_colouring(orc_level *L, const int *colour, int ncolours) {
  const int n = L->n;
  if (ncolours < 1) return -1;
  for (int i = 0; i < n; ++i) {
    if (colour[i] < 0 || colour[i] >= ncolours) return -1;
    for (int k = 0; k < L->nnz_row[i]; ++k) {
      const int c = L->cols[(int64_t)i * NSLOT + k];
      if (c != i && colour[c] == colour[i]) return -1;
    }
  }
  free(L->cstart);
  L->cstart = (int *)calloc((size_t)ncolours + 1, sizeof(int));
  L->ncol = ncolours;
  for (int i = 0; i < n; ++i) L->cstart[colour[i] + 1]++;
  for (int k = 0; k < ncolours; ++k) L->cstart[k + 1] += L->cstart[k];
  int *fill = (int *)malloc((size_t)ncolours * sizeof(int));
  memcpy(fill, L->cstart, (size_t)ncolours * sizeof(int));
  for (int i = 0; i < n; ++i) L->crow[fill[colour[i]]++] = i;
  free(fill);
  return 0;
}

int orc_ncolours(const orc_level *L) { return L->ncol; }
const int *orc_colour_rows(const orc_level *L) { return L->crow; }
const int *orc_colour_starts(const orc_level *L) { return L->cstart; }

/* A level from arbitrary rows held as the reference holds them
 * (SparseMatrix.hpp:44-52): row i has nnz_row[i] <= 27 entries cols/vals at
 * i*27, in the row's own order, one of them the diagonal (its first
 * occurrence is matrixDiagonal, GenerateProblem_ref.cpp:163).  Coloured with
 * orc_colour_general.  NULL if a row is empty, too long, has a column out of
 * range or no diagonal. */
orc_level *orc_create_rows(int n, const unsigned char *nnz_row, const int *cols, const double *vals) {
  if (n < 1) return NULL;
  for (int i = 0; i < n; ++i) {
    if (nnz_row[i] < 1 || nnz_row[i] > NSLOT) return NULL;
    int diag = 0;
    for (int k = 0; k < nnz_row[i]; ++k) {
      const int c = cols[(int64_t)i * NSLOT + k];
      if (c < 0 || c >= n) return NULL;
      if (c == i) diag = 1;
    }
    if (!diag) return NULL;
  }
  orc_level *L = (orc_level *)calloc(1, sizeof(orc_level));
  L->nx = n; L->ny = 1; L->nz = 1;
  L->n = n;
  L->general = 1;
  L->nnz_row = (unsigned char *)malloc((size_t)n);
  L->cols = (int *)malloc((size_t)n * NSLOT * sizeof(int));
  L->vals = (double *)malloc((size_t)n * NSLOT * sizeof(double));
  L->diag_slot = (int *)malloc((size_t)n * sizeof(int));
  L->crow = (int *)malloc((size_t)n * sizeof(int));
  int64_t nnz = 0;
  for (int i = 0; i < n; ++i) {
    L->nnz_row[i] = nnz_row[i];
    nnz += nnz_row[i];
    L->diag_slot[i] = -1;
    for (int k = 0; k < NSLOT; ++k) {
      const int64_t e = (int64_t)i * NSLOT + k;
      if (k < nnz_row[i]) {
        L->cols[e] = cols[e];
        L->vals[e] = vals[e];
        if (cols[e] == i && L->diag_slot[i] < 0) L->diag_slot[i] = k;
      } else {
        L->cols[e] = i;
        L->vals[e] = 0.0;
      }
    }
  }
  L->nnz = nnz;
  for (int k = 0; k < 8; ++k) L->order[k] = k;
  int *colour = (int *)malloc((size_t)n * sizeof(int));
  int nc = 0;
  orc_colour_general(L, colour, &nc);
  orc_set_colouring(L, colour, nc);
  free(colour);
  L->Axf = (double *)calloc((size_t)n, sizeof(double));
  L->Axf32 = (float *)calloc((size_t)n, sizeof(float));
  return L;
}

/* Colour order of a general level with an injection: colours sorted by how
 * many injected rows (f2c targets) they hold, ascending and stable, so the
 * colour of the injected rows is swept last forward and first backward (if
 * it were swept last backward its residual would be zero and the V-cycle
 * would lose its coarse correction).  colour[] is relabelled in place. */
void orc_order_colours(int n, int ncolours, int *colour, const int *f2c, int nc) {
  int *cnt = (int *)calloc((size_t)ncolours, sizeof(int));
  int *rank = (int *)malloc((size_t)ncolours * sizeof(int));
  for (int i = 0; i < nc; ++i) cnt[colour[f2c[i]]]++;
  /* rank of colour k: colours with fewer injected rows first, ties by k */
  for (int k = 0; k < ncolours; ++k) {
    rank[k] = 0;
    for (int j = 0; j < ncolours; ++j)
      if (cnt[j] < cnt[k] || (cnt[j] == cnt[k] && j < k)) rank[k]++;
  }
  for (int i = 0; i < n; ++i) colour[i] = rank[colour[i]];
  free(cnt); free(rank);
}

void orc_level_colours(const orc_level *L, int *colour) {
  for (int k = 0; k < L->ncol; ++k)
    for (int t = L->cstart[k]; t < L->cstart[k + 1]; ++t) colour[L->crow[t]] = k;
}

/* Makes coarse the next level of fine with the injection f2c[n_coarse]
 * (MGData.hpp:56-63: the coarse operator and f2cOperator).  fine takes
 * ownership of coarse.  -1 if f2c is out of range or not one-to-one (the
 * prolongation x[f2c[i]] += xc[i] scatters). */
int orc_link(orc_level *fine, orc_level *coarse, const int *f2c) {
  if (fine->coarse || !coarse) return -1;
  const int nc = coarse->n;
  char *seen = (char *)calloc((size_t)fine->n, 1);
  for (int i = 0; i < nc; ++i) {
    if (f2c[i] < 0 || f2c[i] >= fine->n || seen[f2c[i]]) { free(seen); return -1; }
    seen[f2c[i]] = 1;
  }
  free(seen);
  fine->f2c = (int *)malloc((size_t)nc * sizeof(int));
  memcpy(fine->f2c, f2c, (size_t)nc * sizeof(int));
  fine->coarse = coarse;
  fine->rc = (double *)calloc((size_t)nc, sizeof(double));
  fine->xc = (double *)calloc((size_t)nc, sizeof(double));
  fine->rc32 = (float *)calloc((size_t)nc, sizeof(float));
  fine->xc32 = (float *)calloc((size_t)nc, sizeof(float));
  if (fine->general) {  /* the injected rows' colour goes last */
    int *colour = (int *)malloc((size_t)fine->n * sizeof(int));
    orc_level_colours(fine, colour);
    orc_order_colours(fine->n, fine->ncol, colour, f2c, nc);
    orc_set_colouring(fine, colour, fine->ncol);
    free(colour);
  }
  return 0;
}

/* y = A x, ComputeSPMV_ref.cpp:61-70: per row, left-to-right over the
 * row's entries, starting from 0.0. */
void orc_spmv(const orc_level *L, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < L->n; ++i) {
    const int *c = L->cols + (int64_t)i * NSLOT;
    const double *v = L->vals + (int64_t)i * NSLOT;
    double sum = 0.0;
    for (int j = 0; j < L->nnz_row[i]; ++j) sum += v[j] * x[c[j]];
    y[i] = sum;
  }
}

/* One Gauss-Seidel row update, ComputeSYMGS_ref.cpp:72-79: subtract every
 * entry including the diagonal, add the diagonal term back, divide. */
static inline void gs_row(const orc_level *L, int i, const double *r, double *x) {
  const int *c = L->cols + (int64_t)i * NSLOT;
  const double *v = L->vals + (int64_t)i * NSLOT;
  const double d = v[L->diag_slot[i]];
  double sum = r[i];
  for (int j = 0; j < L->nnz_row[i]; ++j) sum -= v[j] * x[c[j]];
  sum += x[i] * d;
  x[i] = sum / d;
}

/* Symmetric Gauss-Seidel.
 * mode 0: ComputeSYMGS_ref.cpp:67-100 (rows 0..n-1 then n-1..0).
 * mode 1: the same row update, forward over the colours in sweep order, then
 *         backward in reverse order; rows of one colour share no entries, so
 *         their order inside a colour does not change x.
 * mode 2: mode 1 without the backward step of the last forward colour.  That
 *         step recomputes those rows from the neighbour values the forward
 *         step just used, so it differs from mode 1 only by the rounding of
 *         the "- a_ii x_i + x_i a_ii" pair (a few ulp).  The device kernels
 *         implement mode 2. */
void orc_symgs(const orc_level *L, const double *r, double *x, int mode) {
  if (!mode) {
    for (int i = 0; i < L->n; ++i) gs_row(L, i, r, x);
    for (int i = L->n - 1; i >= 0; --i) gs_row(L, i, r, x);
    return;
  }
  for (int k = 0; k < L->ncol; ++k) {
#pragma omp parallel for schedule(static)
    for (int t = L->cstart[k]; t < L->cstart[k + 1]; ++t) gs_row(L, L->crow[t], r, x);
  }
  for (int k = mode == 2 ? L->ncol - 2 : L->ncol - 1; k >= 0; --k) {
#pragma omp parallel for schedule(static)
    for (int t = L->cstart[k]; t < L->cstart[k + 1]; ++t) gs_row(L, L->crow[t], r, x);
  }
}

/* rc = r[f2c] - Axf[f2c], ComputeRestriction_ref.cpp:51. */
void orc_restrict(const orc_level *L, const double *r, const double *Axf, double *rc) {
  for (int i = 0; i < L->coarse->n; ++i) rc[i] = r[L->f2c[i]] - Axf[L->f2c[i]];
}

/* x[f2c] += xc, ComputeProlongation_ref.cpp:49. */
void orc_prolong(const orc_level *L, const double *xc, double *x) {
  for (int i = 0; i < L->coarse->n; ++i) x[L->f2c[i]] += xc[i];
}

/* V-cycle, ComputeMG_ref.cpp:39-63 with one pre- and one post-smoothing
 * sweep (MGData.hpp:51-52) and one sweep on the coarsest level. */
void orc_mg(orc_level *L, const double *r, double *x, int mode) {
  memset(x, 0, (size_t)L->n * sizeof(double));
  if (L->coarse) {
    orc_symgs(L, r, x, mode);
    orc_spmv(L, x, L->Axf);
    orc_restrict(L, r, L->Axf, L->rc);
    orc_mg(L->coarse, L->rc, L->xc, mode);
    orc_prolong(L, L->xc, x);
    orc_symgs(L, r, x, mode);
  } else {
    orc_symgs(L, r, x, mode);
  }
}

/* Serial dot product, ComputeDotProduct_ref.cpp:54-64 (x == y reads x twice). */
double orc_dot(int64_t n, const double *x, const double *y) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

/* w = alpha x + beta y with the reference's three branches,
 * ComputeWAXPBY_ref.cpp:52-67. */
void orc_waxpby(int64_t n, double alpha, const double *x, double beta, const double *y, double *w) {
  if (alpha == 1.0) {
    for (int64_t i = 0; i < n; ++i) w[i] = x[i] + beta * y[i];
  } else if (beta == 1.0) {
    for (int64_t i = 0; i < n; ++i) w[i] = alpha * x[i] + y[i];
  } else {
    for (int64_t i = 0; i < n; ++i) w[i] = alpha * x[i] + beta * y[i];
  }
}

/* Preconditioned CG, the hook version CG.cpp:86-130 (k == 1 forms p with
 * WAXPBY(1,z,0,z), :109).  hist[0] = normr0, hist[k] = normr after k. */
int orc_cg(orc_level *L, const double *b, double *x, int max_iter, double tol, int precond,
           int mode, int *niters, double *normr_out, double *normr0_out, double *hist) {
  const int64_t n = L->n;
  double *r = (double *)malloc(n * sizeof(double));
  double *z = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * sizeof(double));
  double *Ap = (double *)malloc(n * sizeof(double));
  double rtz = 0.0, oldrtz, alpha, beta, pAp;
  memcpy(p, x, n * sizeof(double));
  orc_spmv(L, p, Ap);
  orc_waxpby(n, 1.0, b, -1.0, Ap, r);
  double normr = sqrt(orc_dot(n, r, r));
  const double normr0 = normr;
  if (hist) hist[0] = normr;
  *niters = 0;
  for (int k = 1; k <= max_iter && normr / normr0 > tol; ++k) {
    if (precond) orc_mg(L, r, z, mode);
    else memcpy(z, r, n * sizeof(double));
    if (k == 1) {
      orc_waxpby(n, 1.0, z, 0.0, z, p);
      rtz = orc_dot(n, r, z);
    } else {
      oldrtz = rtz;
      rtz = orc_dot(n, r, z);
      beta = rtz / oldrtz;
      orc_waxpby(n, 1.0, z, beta, p, p);
    }
    orc_spmv(L, p, Ap);
    pAp = orc_dot(n, p, Ap);
    alpha = rtz / pAp;
    orc_waxpby(n, 1.0, x, alpha, p, x);
    orc_waxpby(n, 1.0, r, -alpha, Ap, r);
    normr = sqrt(orc_dot(n, r, r));
    if (hist) hist[k] = normr;
    *niters = k;
  }
  *normr_out = normr;
  *normr0_out = normr0;
  free(r); free(z); free(p); free(Ap);
  return 0;
}

/* ---------------- mixed precision (no reference counterpart) ------------- */

uint16_t orc_f32_to_f16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const int32_t exp = (int32_t)((u >> 23) & 0xff) - 127 + 15;
  uint32_t man = u & 0x7fffffu;
  if (((u >> 23) & 0xff) == 0xff) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0));
  if (exp >= 31) return (uint16_t)(sign | 0x7c00u);
  if (exp <= 0) {
    if (exp < -10) return (uint16_t)sign;
    man |= 0x800000u;
    const int shift = 14 - exp;
    uint32_t h = man >> shift;
    const uint32_t rem = man & ((1u << shift) - 1), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)exp << 10) | (man >> 13);
  const uint32_t rem = man & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  return (uint16_t)(sign | h);
}

float orc_f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1f, man = h & 0x3ffu, u;
  if (exp == 0) {
    if (man == 0) {
      u = sign;
    } else {
      int e = -1;
      do { ++e; man <<= 1; } while (!(man & 0x400u));
      u = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3ffu) << 13);
    }
  } else if (exp == 31) {
    u = sign | 0x7f800000u | (man << 13);
  } else {
    u = sign | ((exp - 15 + 127) << 23) | (man << 13);
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static void ensure_f32(orc_level *L) {
  for (; L; L = L->coarse) {
    if (L->vals32) continue;
    const int64_t m = (int64_t)L->n * NSLOT;
    L->vals32 = (float *)malloc(m * sizeof(float));
    for (int64_t k = 0; k < m; ++k) L->vals32[k] = (float)L->vals[k];
  }
}

static int planted_fault = 0;
void orc_plant_fault(int which) { planted_fault = which; }

/* Symmetric diagonal scaling: ahat_ij = a_ij * dinv_i * dinv_j (in double,
 * then rounded to fp16), dinv = 1/sqrt(a_ii); the diagonal becomes 1. */
static void ensure_f16(orc_level *L) {
  for (orc_level *l = L; l; l = l->coarse) {
    if (l->vals16) continue;
    const int n = l->n;
    free(l->dinv);
    l->dinv = (float *)malloc((size_t)n * sizeof(float));
    double *dd = (double *)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; ++i) dd[i] = 1.0 / sqrt(l->vals[(int64_t)i * NSLOT + l->diag_slot[i]]);
    l->vals16 = (uint16_t *)malloc((size_t)n * NSLOT * sizeof(uint16_t));
    for (int i = 0; i < n; ++i) {
      l->dinv[i] = (float)dd[i];
      for (int j = 0; j < NSLOT; ++j) {
        const int64_t k = (int64_t)i * NSLOT + j;
        const double a = (j == l->diag_slot[i]) ? 1.0 : l->vals[k] * dd[i] * dd[l->cols[k]];
        l->vals16[k] = orc_f32_to_f16((float)a);
      }
    }
    free(dd);
  }
  for (orc_level *l = L; l && l->coarse; l = l->coarse) {
    if (l->sigma) continue;
    const int nc = l->coarse->n;
    l->sigma = (float *)malloc((size_t)nc * sizeof(float));
    for (int i = 0; i < nc; ++i) {
      const int f = l->f2c[i];
      const double df = l->vals[(int64_t)f * NSLOT + l->diag_slot[f]];
      const double dc = l->coarse->vals[(int64_t)i * NSLOT + l->coarse->diag_slot[i]];
      const double sg = planted_fault == 1 ? sqrt(dc / df) : planted_fault == 2 ? 1.0 : sqrt(df / dc);
      l->sigma[i] = (float)sg;
    }
  }
}

/* fp32 row update with the reference's operation order; vals is either the
 * fp32 copy or (half != 0) the scaled fp16 copy. */
static inline void gs_row32(const orc_level *L, int i, const float *r, float *x, int half) {
  const int *c = L->cols + (int64_t)i * NSLOT;
  const int64_t base = (int64_t)i * NSLOT;
  float d, sum = r[i];
  if (half) {
    d = orc_f16_to_f32(L->vals16[base + L->diag_slot[i]]);
    for (int j = 0; j < L->nnz_row[i]; ++j) sum -= orc_f16_to_f32(L->vals16[base + j]) * x[c[j]];
  } else {
    d = L->vals32[base + L->diag_slot[i]];
    for (int j = 0; j < L->nnz_row[i]; ++j) sum -= L->vals32[base + j] * x[c[j]];
  }
  sum += x[i] * d;
  x[i] = sum / d;
}

static void symgs32(const orc_level *L, const float *r, float *x, int half) {
  for (int k = 0; k < L->ncol; ++k) {
#pragma omp parallel for schedule(static)
    for (int t = L->cstart[k]; t < L->cstart[k + 1]; ++t) gs_row32(L, L->crow[t], r, x, half);
  }
  for (int k = L->ncol - 2; k >= 0; --k) {
#pragma omp parallel for schedule(static)
    for (int t = L->cstart[k]; t < L->cstart[k + 1]; ++t) gs_row32(L, L->crow[t], r, x, half);
  }
}

static void spmv32(const orc_level *L, const float *x, float *y, int half) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < L->n; ++i) {
    const int *c = L->cols + (int64_t)i * NSLOT;
    const int64_t base = (int64_t)i * NSLOT;
    float sum = 0.0f;
    for (int j = 0; j < L->nnz_row[i]; ++j) {
      const float a = half ? orc_f16_to_f32(L->vals16[base + j]) : L->vals32[base + j];
      sum += a * x[c[j]];
    }
    y[i] = sum;
  }
}

/* fp32 V-cycle: same structure as orc_mg in single precision.  In the
 * scaled (fp16) variant the level operator is D^-1/2 A D^-1/2 and the
 * transfers carry the factor sigma = sqrt(d_f[f2c]/d_c). */
static void mg32(orc_level *L, const float *r, float *x, int half) {
  memset(x, 0, (size_t)L->n * sizeof(float));
  symgs32(L, r, x, half);
  if (!L->coarse) return;
  spmv32(L, x, L->Axf32, half);
  const int nc = L->coarse->n;
  for (int i = 0; i < nc; ++i) {
    const float res = r[L->f2c[i]] - L->Axf32[L->f2c[i]];
    L->rc32[i] = half ? res * L->sigma[i] : res;
  }
  mg32(L->coarse, L->rc32, L->xc32, half);
  for (int i = 0; i < nc; ++i)
    x[L->f2c[i]] += half ? L->xc32[i] * L->sigma[i] : L->xc32[i];
  symgs32(L, r, x, half);
}

void orc_mg_f32(orc_level *L, const float *r, float *x) {
  ensure_f32(L);
  mg32(L, r, x, 0);
}

void orc_mg_f16(orc_level *L, const float *r, float *x) {
  ensure_f16(L);
  const int n = L->n;
  float *rh = (float *)malloc((size_t)n * sizeof(float));
  for (int i = 0; i < n; ++i) rh[i] = r[i] * L->dinv[i];
  mg32(L, rh, x, 1);
  for (int i = 0; i < n; ++i) x[i] = x[i] * L->dinv[i];
  free(rh);
}

/* Flexible PCG with an fp32 (mg_prec 32) or scaled-fp16 (mg_prec 16) MG
 * preconditioner and fp64 outer vectors.  beta is Polak-Ribiere,
 * beta_k = z_k.(r_k - r_{k-1}) / (z_{k-1}.r_{k-1}) = -alpha_{k-1} z_k.Ap_{k-1} / oldrtz.
 * On reaching tol the true residual b - A x is formed; if it is above tol
 * the recurrence restarts from it (residual replacement). */
int orc_cg_mixed(orc_level *L, const double *b, double *x, int max_iter, double tol, int mg_prec,
                 int *niters, double *normr_out, double *normr0_out, double *hist) {
  const int64_t n = L->n;
  double *r = (double *)malloc(n * sizeof(double));
  double *z = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * sizeof(double));
  double *Ap = (double *)malloc(n * sizeof(double));
  float *r32 = (float *)malloc(n * sizeof(float));
  float *z32 = (float *)malloc(n * sizeof(float));
  orc_spmv(L, x, Ap);
  orc_waxpby(n, 1.0, b, -1.0, Ap, r);
  double normr = sqrt(orc_dot(n, r, r));
  const double normr0 = normr;
  if (hist) hist[0] = normr;
  double rtz = 0.0, oldrtz = 0.0, alpha = 0.0;
  int restart = 1, k = 0;
  *niters = 0;
  while (k < max_iter) {
    if (!(normr / normr0 > tol)) {
      orc_spmv(L, x, Ap);
      orc_waxpby(n, 1.0, b, -1.0, Ap, r);
      normr = sqrt(orc_dot(n, r, r));
      if (!(normr / normr0 > tol)) break;
      restart = 1;
    }
    ++k;
    for (int64_t i = 0; i < n; ++i) r32[i] = (float)r[i];
    if (mg_prec == 16) orc_mg_f16(L, r32, z32);
    else orc_mg_f32(L, r32, z32);
    for (int64_t i = 0; i < n; ++i) z[i] = (double)z32[i];
    if (restart) {
      rtz = orc_dot(n, r, z);
      memcpy(p, z, n * sizeof(double));
      restart = 0;
    } else {
      oldrtz = rtz;
      rtz = orc_dot(n, r, z);
      const double zAp = orc_dot(n, z, Ap);
      const double beta = -alpha * zAp / oldrtz;
      orc_waxpby(n, 1.0, z, beta, p, p);
    }
    orc_spmv(L, p, Ap);
    const double pAp = orc_dot(n, p, Ap);
    alpha = rtz / pAp;
    orc_waxpby(n, 1.0, x, alpha, p, x);
    orc_waxpby(n, 1.0, r, -alpha, Ap, r);
    normr = sqrt(orc_dot(n, r, r));
    if (hist) hist[k] = normr;
    *niters = k;
  }
  *normr_out = normr;
  *normr0_out = normr0;
  free(r); free(z); free(p); free(Ap); free(r32); free(z32);
  return 0;
}

/* dL/dkappa_k = -lam^T (dA/dkappa_k) x
 *             = -1/2 sum_{j in N(k)} (lam_k-lam_j)(x_k-x_j) - (26-n_k) lam_k x_k,
 * neighbours in ascending column order. */
void orc_kappa_grad(const orc_level *L, const double *lam, const double *x, double *g) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < L->n; ++k) {
    const int *c = L->cols + (int64_t)k * NSLOT;
    double s = 0.0;
    for (int j = 0; j < L->nnz_row[k]; ++j) {
      const int col = c[j];
      if (col == k) continue;
      s += (lam[k] - lam[col]) * (x[k] - x[col]);
    }
    g[k] = -0.5 * s - (double)(26 - (L->nnz_row[k] - 1)) * lam[k] * x[k];
  }
}
