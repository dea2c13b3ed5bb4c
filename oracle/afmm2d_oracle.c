/*
 * afmm2d_oracle.c -- CPU restatement of the AFMM2D reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product never links, loads or calls anything in oracle/.
 *
 * Written from the behaviour of the reference (citations are file:line in the
 * reference repository), not from its code: plain C, FP64, OpenMP.
 *
 * Pins (what checks each part, beyond retyping):
 *   kd sort     -- numpy lexsort(key asc, position desc) brute force on
 *                  tie-heavy inputs; bijection; split invariants.
 *   topology    -- tables extracted from the reference's own assignment
 *                  statements (tests/golden/topology_L4.json).
 *   ACA / ID    -- numpy SVD lower bound on rank, interpolation identity
 *                  L2P[skeleton] = I, low-rank residual O(tol), brute-force
 *                  pivot replay of the stopping rule on tiny blocks.
 *   LU solve    -- numpy.linalg.solve.
 *   matvec      -- FP64 dense direct sum (closed form), relative error
 *                  O(tol); linearity; near-field-only identities.
 * Deliberate deviations from the reference (both here and in the product):
 *   - a zero ACA pivot ends the ACA (the reference divides by zero);
 *   - a box of rank 0 contributes no far field (the reference aborts);
 *   - R2 == 0 gives the kernel's diagonal value (the reference evaluates
 *     log(0) for distinct coincident points).
 * Product-only deviation, switchable here (or_set_exact_skeleton_rows): the
 * skeleton rows of L2P are exactly the identity rows the LU solve returns up
 * to rounding; pinned by the matvec agreeing with the solved rows to
 * rounding (tests/test_oracle_matvec.py).
 */
#include "afmm2d_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Kernels: examples/testAFMM2D.cpp:9-20 (a/R), kernel.hpp:39-54 (1/r, log). */

/* The user kernel (kernel id 4): the same header the engine is built with
 * (see afmm2d_amd/user_kernels/anisotropic_log.h for the contract). */
#define AFMM2D_USER_FN static inline
#ifndef AFMM2D_USER_KERNEL_HEADER
#define AFMM2D_USER_KERNEL_HEADER "../afmm2d_amd/user_kernels/anisotropic_log.h"
#endif
#include AFMM2D_USER_KERNEL_HEADER

double or_kernel_value(int kid, const double* kp, double a, double xi, double yi,
                       double xj, double yj) {
  double dx = xi - xj, dy = yi - yj;
  double r2 = dx * dx + dy * dy;
  switch (kid) {
    case OR_KERNEL_LOG:
      return r2 == 0.0 ? 0.0 : 0.5 * log(r2);
    case OR_KERNEL_INV_R:
      return r2 == 0.0 ? 0.0 : 1.0 / sqrt(r2);
    case OR_KERNEL_GAUSS: {
      double s2 = (kp && kp[0] > 0.0) ? kp[0] : 0.01;
      return exp(-r2 / s2);
    }
    case OR_KERNEL_REG_INV_R: {
      double r = sqrt(r2);
      return r < a ? r / a : a / r;
    }
    case OR_KERNEL_USER:
    case OR_KERNEL_USER_SYM:
    case OR_KERNEL_USER_ANTI: {
      /* the user kernel, or its symmetric / antisymmetric part: the
       * reference's algorithm assumes a symmetric kernel (AFMM2DTree.cpp:
       * 1081-1082, 1151); K = S + A runs as one operator per part */
      static const double zero[4] = {0.0, 0.0, 0.0, 0.0};
      const double* p = kp ? kp : zero;
      const double kij = afmm2d_user_kernel(xi, yi, xj, yj, p, a);
      if (kid == OR_KERNEL_USER) return kij;
      const double kji = afmm2d_user_kernel(xj, yj, xi, yi, p, a);
      return kid == OR_KERNEL_USER_SYM ? 0.5 * (kij + kji) : 0.5 * (kij - kji);
    }
  }
  return 0.0;
}

/* One entry per call through a function pointer: the reference makes one
 * virtual getMatrixEntry call per entry (kernel.cpp:3-47). */
typedef struct {
  int kid;
  double kp[4];
  double a;
  const double* x;
  const double* y;
} or_kernel;

typedef double (*or_entry_fn)(const or_kernel*, int, int);

static double entry_impl(const or_kernel* k, int i, int j) {
  return or_kernel_value(k->kid, k->kp, k->a, k->x[i], k->y[i], k->x[j], k->y[j]);
}

/* ------------------------------------------------------------------------- */
/* Levels: AFMM2D.cpp:7 (integer N/M, truncated log4). */

int or_nlevels(int n, int leaf) {
  if (n <= 0 || leaf <= 0) return 0;
  int q = n / leaf;
  if (q <= 0) return 0;
  return (int)(unsigned)(log((double)q) / log(4.0));
}

/* ------------------------------------------------------------------------- */
/* KD partition: KDTree.cpp:37-142 (merge sort, ties take the right run),
 * 200-243 (box numbering [0,3,1,2]). */

typedef struct { double c[2]; double p; } kd_row;

static void kd_merge(kd_row* a, unsigned ls, unsigned ln, unsigned rs, unsigned rn, int key,
                     kd_row* tmp) {
  unsigned jl = ls, jr = rs, j = 0;
  while (jl < ls + ln && jr < rs + rn) {
    if (a[jl].c[key] < a[jr].c[key]) tmp[j++] = a[jl++];
    else tmp[j++] = a[jr++];
  }
  while (jl < ls + ln) tmp[j++] = a[jl++];
  while (jr < rs + rn) tmp[j++] = a[jr++];
  memcpy(a + ls, tmp, (size_t)(ln + rn) * sizeof(kd_row));
}

static void kd_merge_sort(kd_row* a, unsigned s, unsigned n, int key, kd_row* tmp) {
  if (n <= 1) return;
  unsigned nl = n / 2;
  kd_merge_sort(a, s, nl, key, tmp);
  kd_merge_sort(a, s + nl, n - nl, key, tmp);
  kd_merge(a, s, nl, s + nl, n - nl, key, tmp);
}

static void kd_rec(kd_row* a, unsigned s, unsigned n, unsigned idx, unsigned lev, unsigned L,
                   int* counts, int* ncounts, kd_row* tmp) {
  idx %= 2;
  kd_merge_sort(a, s, n, (int)idx, tmp);
  unsigned nl = n / 2, nr = n - nl;
  idx = (idx + 1) % 2;
  kd_merge_sort(a, s, nl, (int)idx, tmp);
  unsigned lb = nl / 2, lt = nl - lb;
  kd_merge_sort(a, s + nl, nr, (int)idx, tmp);
  unsigned rb = nr / 2, rt = nr - rb;
  idx = (idx + 1) % 2;
  if (lev < L) {
    lev++;
    if (lev == L) {
      counts[(*ncounts)++] = (int)lb;
      counts[(*ncounts)++] = (int)lt;
      counts[(*ncounts)++] = (int)rb;
      counts[(*ncounts)++] = (int)rt;
    }
    kd_rec(a, s, lb, idx, lev, L, counts, ncounts, tmp);
    kd_rec(a, s + lb, lt, idx, lev, L, counts, ncounts, tmp);
    kd_rec(a, s + nl, rb, idx, lev, L, counts, ncounts, tmp);
    kd_rec(a, s + nl + rb, rt, idx, lev, L, counts, ncounts, tmp);
  }
}

void or_kd_sort(int n, const double* xy, int L, double* sorted_xy, int64_t* perm, int* kd_counts,
                int* kd_boxnum) {
  kd_row* a = (kd_row*)malloc((size_t)(n > 0 ? n : 1) * sizeof(kd_row));
  kd_row* tmp = (kd_row*)malloc((size_t)(n > 0 ? n : 1) * sizeof(kd_row));
  for (int i = 0; i < n; i++) {
    a[i].c[0] = xy[2 * i];
    a[i].c[1] = xy[2 * i + 1];
    a[i].p = (double)i;
  }
  int nc = 0;
  if (n > 1) kd_rec(a, 0, (unsigned)n, 0, 0, (unsigned)L, kd_counts, &nc, tmp);
  for (int i = 0; i < n; i++) {
    sorted_xy[2 * i] = a[i].c[0];
    sorted_xy[2 * i + 1] = a[i].c[1];
    perm[i] = (int64_t)a[i].p;
  }
  /* Leaf box number of KD-order leaf t: base-4 digits of t, most significant
   * first, mapped through {0,3,1,2}. */
  if (kd_boxnum) {
    static const int map[4] = {0, 3, 1, 2};
    long nleaf = 1L << (2 * L);
    for (long t = 0; t < nleaf; t++) {
      int b = 0;
      for (int d = L - 1; d >= 0; d--) b = 4 * b + map[(t >> (2 * d)) & 3];
      kd_boxnum[t] = b;
    }
  }
  free(a);
  free(tmp);
}

/* ------------------------------------------------------------------------- */
/* Topology: AFMM2DTree.cpp:24-62 (ids: child c of k is 4k+c), 65-967 (lists
 * filled from the parent's neighbours' children), 969-989 (child c offsets:
 * 0=SW, 1=SE, 2=NE, 3=NW).  Restated level by level from the parent's
 * neighbour list; each candidate is placed by its offset on the ring of
 * Chebyshev radius d, counter-clockwise from the ring's SW corner. */

static const int child_dx[4] = {0, 1, 1, 0};
static const int child_dy[4] = {0, 0, 1, 1};

static int ring_slot(int d, int dx, int dy) {
  if (dy == -d) return dx + d;
  if (dx == d) return 2 * d + (dy + d);
  if (dy == d) return 4 * d + (d - dx);
  return 6 * d + (d - dy);
}

static void box_pos(int level, int k, int* px, int* py) {
  int x = 0, y = 0;
  for (int d = level - 1; d >= 0; d--) {
    int c = (k >> (2 * d)) & 3;
    x = 2 * x + child_dx[c];
    y = 2 * y + child_dy[c];
  }
  *px = x;
  *py = y;
}

void or_topology(int level, int* nbr, int* inner, int* outer) {
  /* Build every level up to 'level' from the root (root lists are empty). */
  long nb = 1;
  int* pn = (int*)malloc(sizeof(int) * 8);
  for (int s = 0; s < 8; s++) pn[s] = -1;
  for (int j = 0; j < level; j++) {
    long nc = nb * 4;
    int* cn = (int*)malloc(sizeof(int) * 8 * nc);
    int* ci = (int*)malloc(sizeof(int) * 16 * nc);
    int* co = (int*)malloc(sizeof(int) * 24 * nc);
    for (long i = 0; i < 8 * nc; i++) cn[i] = -1;
    for (long i = 0; i < 16 * nc; i++) ci[i] = -1;
    for (long i = 0; i < 24 * nc; i++) co[i] = -1;
    for (long k = 0; k < nb; k++) {
      for (int c = 0; c < 4; c++) {
        long ch = 4 * k + c;
        int cx, cy;
        box_pos(j + 1, (int)ch, &cx, &cy);
        /* candidates: the parent's own children and its neighbours' children */
        for (int s = -1; s < 8; s++) {
          long pk = s < 0 ? k : pn[8 * k + s];
          if (pk < 0) continue;
          for (int cc = 0; cc < 4; cc++) {
            long b = 4 * pk + cc;
            if (b == ch) continue;
            int bx, by;
            box_pos(j + 1, (int)b, &bx, &by);
            int dx = bx - cx, dy = by - cy;
            int d = abs(dx) > abs(dy) ? abs(dx) : abs(dy);
            if (d == 1) cn[8 * ch + ring_slot(1, dx, dy)] = (int)b;
            else if (d == 2) ci[16 * ch + ring_slot(2, dx, dy)] = (int)b;
            else if (d == 3) co[24 * ch + ring_slot(3, dx, dy)] = (int)b;
          }
        }
      }
    }
    free(pn);
    pn = cn;
    if (j + 1 == level) {
      memcpy(nbr, cn, sizeof(int) * 8 * nc);
      memcpy(inner, ci, sizeof(int) * 16 * nc);
      memcpy(outer, co, sizeof(int) * 24 * nc);
    }
    free(ci);
    free(co);
    nb = nc;
  }
  if (level == 0) {
    for (int s = 0; s < 8; s++) nbr[s] = -1;
    for (int s = 0; s < 16; s++) inner[s] = -1;
    for (int s = 0; s < 24; s++) outer[s] = -1;
  }
  free(pn);
}

/* ------------------------------------------------------------------------- */
/* ACA: ACA.cpp:12-25 (first-max argmax over allowed indices), 27-133. */

static int argmax_allowed(const double* v, const unsigned char* used, int n) {
  int idx = -1;
  double mx = 0.0;
  for (int i = 0; i < n; i++) {
    if (used[i]) continue;
    double a = fabs(v[i]);
    if (idx < 0) { idx = i; mx = a; }
    else if (a > mx) { idx = i; mx = a; }
  }
  return idx;
}

typedef double (*aca_fn)(const void* ctx, int i, int j);

/* Summation order of the ACA's dot products (the norm estimate's terms).
 * 0: ascending index (this restatement's order for ACA.cpp's dot products).
 * 1: the engine's order: 64 partial sums, partial l adding the products of
 * indices l, l + 64, l + 128, ... in ascending order, then combined by the
 * butterfly s_l + s_(l xor d) for d = 32, 16, ..., 1 (a wave's shuffle
 * reduction).  Set for the whole library (not thread-safe). */
static int g_aca_wave_sums = 0;

void or_set_aca_wave_sums(int on) { g_aca_wave_sums = on; }

static double aca_dot(const double* a, const double* b, int n) {
  if (!g_aca_wave_sums) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
  }
  double part[64];
  for (int l = 0; l < 64; l++) {
    double s = 0.0;
    for (int i = l; i < n; i += 64) s += a[i] * b[i];
    part[l] = s;
  }
  for (int d = 32; d > 0; d >>= 1) {
    double nxt[64];
    for (int l = 0; l < 64; l++) nxt[l] = part[l] + part[l ^ d];
    memcpy(part, nxt, sizeof(part));
  }
  return part[0];
}

/* Returns rank; U (n1 x mn, column l at U + l*n1) and V (n2 x mn) are scratch. */
static int aca_core(aca_fn K, const void* ctx, int n1, int n2, double tol, int* row_piv,
                    int* col_piv, double* U, double* V, double* row, unsigned char* used_r,
                    unsigned char* used_c) {
  if (n1 <= 0 || n2 <= 0) return 0;
  int mn = n1 < n2 ? n1 : n2;
  memset(used_r, 0, (size_t)n1);
  memset(used_c, 0, (size_t)n2);
  int l, jstar = -1;
  for (l = 0; l < n1; l++) {
    for (int j = 0; j < n2; j++) row[j] = K(ctx, l, j);
    jstar = argmax_allowed(row, used_c, n2);
    if (fabs(row[jstar]) > tol) break;
  }
  if (l == n1) return 0;
  int istar = l;
  double piv = row[jstar];
  for (int j = 0; j < n2; j++) V[j] = row[j] / piv;
  for (int i = 0; i < n1; i++) U[i] = K(ctx, i, jstar);
  row_piv[0] = istar;
  col_piv[0] = jstar;
  used_c[jstar] = 1;
  used_r[istar] = 1;
  int rank = 1;
  double normS = 0.0, prev = DBL_MAX;
  if (rank < n1) istar = argmax_allowed(U, used_r, n1);
  while (fabs(prev - normS) >= tol * prev && rank < mn) {
    for (int j = 0; j < n2; j++) {
      double r = K(ctx, istar, j);
      for (int q = 0; q < rank; q++) r -= U[(size_t)q * n1 + istar] * V[(size_t)q * n2 + j];
      row[j] = r;
    }
    jstar = argmax_allowed(row, used_c, n2);
    piv = row[jstar];
    if (piv == 0.0) break; /* deviation: the reference divides by zero */
    row_piv[rank] = istar;
    col_piv[rank] = jstar;
    double* v = V + (size_t)rank * n2;
    double* u = U + (size_t)rank * n1;
    for (int j = 0; j < n2; j++) v[j] = row[j] / piv;
    for (int i = 0; i < n1; i++) {
      double c = K(ctx, i, jstar);
      for (int q = 0; q < rank; q++) c -= V[(size_t)q * n2 + jstar] * U[(size_t)q * n1 + i];
      u[i] = c;
    }
    rank++;
    used_c[jstar] = 1;
    used_r[istar] = 1;
    if (rank != 2) prev = normS;
    double nu = aca_dot(u, u, n1), nv = aca_dot(v, v, n2);
    double p = sqrt(nu) * sqrt(nv);
    double s = normS * normS + p * p;
    for (int q = 0; q < rank; q++) {
      const double du = aca_dot(U + (size_t)q * n1, u, n1);
      const double dv = aca_dot(V + (size_t)q * n2, v, n2);
      s += 2.0 * fabs(du * dv);
    }
    normS = sqrt(s);
    if (rank < n1) istar = argmax_allowed(u, used_r, n1);
  }
  return rank;
}

typedef struct {
  int kid;
  const double* kp;
  double a;
  const double* rxy;
  const double* cxy;
} pts_ctx;

static double pts_entry(const void* c, int i, int j) {
  const pts_ctx* p = (const pts_ctx*)c;
  return or_kernel_value(p->kid, p->kp, p->a, p->rxy[2 * i], p->rxy[2 * i + 1], p->cxy[2 * j],
                         p->cxy[2 * j + 1]);
}

int or_aca_points(int kid, const double* kp, double a, double tol, int n1, const double* rxy,
                  int n2, const double* cxy, int* row_piv, int* col_piv) {
  if (n1 <= 0 || n2 <= 0) return 0;
  int mn = n1 < n2 ? n1 : n2;
  double* U = (double*)malloc(sizeof(double) * (size_t)n1 * mn);
  double* V = (double*)malloc(sizeof(double) * (size_t)n2 * mn);
  double* row = (double*)malloc(sizeof(double) * (size_t)n2);
  unsigned char* ur = (unsigned char*)malloc((size_t)n1);
  unsigned char* uc = (unsigned char*)malloc((size_t)n2);
  pts_ctx c = {kid, kp, a, rxy, cxy};
  int r = aca_core(pts_entry, &c, n1, n2, tol, row_piv, col_piv, U, V, row, ur, uc);
  free(U); free(V); free(row); free(ur); free(uc);
  return r;
}

typedef struct {
  int n2;
  const double* A;
} mat_ctx;

static double mat_entry(const void* c, int i, int j) {
  const mat_ctx* m = (const mat_ctx*)c;
  return m->A[(size_t)i * m->n2 + j];
}

int or_aca_matrix(int n1, int n2, const double* A, double tol, int* row_piv, int* col_piv) {
  if (n1 <= 0 || n2 <= 0) return 0;
  int mn = n1 < n2 ? n1 : n2;
  double* U = (double*)malloc(sizeof(double) * (size_t)n1 * mn);
  double* V = (double*)malloc(sizeof(double) * (size_t)n2 * mn);
  double* row = (double*)malloc(sizeof(double) * (size_t)n2);
  unsigned char* ur = (unsigned char*)malloc((size_t)n1);
  unsigned char* uc = (unsigned char*)malloc((size_t)n2);
  mat_ctx c = {n2, A};
  int r = aca_core(mat_entry, &c, n1, n2, tol, row_piv, col_piv, U, V, row, ur, uc);
  free(U); free(V); free(row); free(ur); free(uc);
  return r;
}

/* ------------------------------------------------------------------------- */
/* LU with partial pivoting (first max), for L2P = Ac * D^{-1} via LU(D^T)
 * (AFMM2DTree.cpp:1057-1059). */

/* Solves X * D = Ac for X (n1 x r), all row-major. */
static int l2p_solve(int r, const double* D, int n1, const double* Ac, double* X) {
  double* M = (double*)malloc(sizeof(double) * (size_t)r * r);
  int* piv = (int*)malloc(sizeof(int) * (size_t)r);
  double* b = (double*)malloc(sizeof(double) * (size_t)r);
  /* M = D^T */
  for (int i = 0; i < r; i++)
    for (int j = 0; j < r; j++) M[(size_t)i * r + j] = D[(size_t)j * r + i];
  for (int k = 0; k < r; k++) {
    int p = k;
    double mx = fabs(M[(size_t)k * r + k]);
    for (int i = k + 1; i < r; i++) {
      double v = fabs(M[(size_t)i * r + k]);
      if (v > mx) { mx = v; p = i; }
    }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < r; j++) {
        double t = M[(size_t)k * r + j];
        M[(size_t)k * r + j] = M[(size_t)p * r + j];
        M[(size_t)p * r + j] = t;
      }
    double d = M[(size_t)k * r + k];
    if (d != 0.0)
      for (int i = k + 1; i < r; i++) M[(size_t)i * r + k] /= d;
    for (int i = k + 1; i < r; i++) {
      double f = M[(size_t)i * r + k];
      for (int j = k + 1; j < r; j++) M[(size_t)i * r + j] -= f * M[(size_t)k * r + j];
    }
  }
  /* For each row i of Ac: solve D^T x = Ac[i,:]^T. */
  for (int i = 0; i < n1; i++) {
    for (int j = 0; j < r; j++) b[j] = Ac[(size_t)i * r + j];
    for (int k = 0; k < r; k++)
      if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int k = 0; k < r; k++)
      for (int j = 0; j < k; j++) b[k] -= M[(size_t)k * r + j] * b[j];
    for (int k = r - 1; k >= 0; k--) {
      for (int j = k + 1; j < r; j++) b[k] -= M[(size_t)k * r + j] * b[j];
      b[k] /= M[(size_t)k * r + k];
    }
    for (int j = 0; j < r; j++) X[(size_t)i * r + j] = b[j];
  }
  free(M); free(piv); free(b);
  return 0;
}

int or_solve_right(int r, const double* D, int n1, const double* Ac, double* X) {
  return l2p_solve(r, D, n1, Ac, X);
}

/* ------------------------------------------------------------------------- */
/* Tree (AFMM2DTree.hpp:12-56, AFMM2DBox.hpp:11-38). */

typedef struct {
  int nin, rank;
  int* rows;     /* nin sorted-point indices (leaf points or children skeletons) */
  int* skel;     /* rank sorted-point indices, pivot order */
  double* l2p;   /* nin x rank row-major */
  double* m2l[40];
  double* outv;  /* rank */
  double* inv;   /* rank */
  double* pot;   /* leaf: count */
} or_box;

struct or_tree {
  int n, leaf, tol_pow, L;
  double tol;
  or_kernel k;
  or_entry_fn entry;
  double* sx;
  double* sy;
  int64_t* perm;
  int* kd_counts;
  int* kd_boxnum;
  int* leaf_start; /* by box number */
  int* leaf_count;
  int* nbr[32];
  int* inner[32];
  int* outer[32];
  or_box* box[32];
  double* qs;
  double* ps;
  int assembled;
  int par_mode;
  int exact_skel_rows;
};

int or_levels(const or_tree* t) { return t->L; }
double or_kernel_a(const or_tree* t) { return t->k.a; }

or_tree* or_create(int n, int leaf, int tol_pow, const double* xy, int kid, const double* kp) {
  int L = or_nlevels(n, leaf);
  if (L < 2 || L > 15) return NULL;
  or_tree* t = (or_tree*)calloc(1, sizeof(or_tree));
  t->n = n;
  t->leaf = leaf;
  t->tol_pow = tol_pow;
  t->tol = pow(10.0, -1.0 * tol_pow); /* ACA.cpp:7 */
  t->L = L;
  double* sxy = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  t->perm = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  long nleaf = 1L << (2 * L);
  t->kd_counts = (int*)malloc(sizeof(int) * nleaf);
  t->kd_boxnum = (int*)malloc(sizeof(int) * nleaf);
  or_kd_sort(n, xy, L, sxy, t->perm, t->kd_counts, t->kd_boxnum);
  t->sx = (double*)malloc(sizeof(double) * (size_t)n);
  t->sy = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) { t->sx[i] = sxy[2 * i]; t->sy[i] = sxy[2 * i + 1]; }
  free(sxy);
  /* kernel: a = 0.5^L (AFMM2DTree.cpp:13-19) */
  t->k.kid = kid;
  for (int i = 0; i < 4; i++) t->k.kp[i] = kp ? kp[i] : 0.0;
  t->k.a = ldexp(1.0, -L);
  t->k.x = t->sx;
  t->k.y = t->sy;
  t->entry = entry_impl;
  /* leaf ranges (AFMM2DTree.cpp:1099-1118) */
  t->leaf_start = (int*)malloc(sizeof(int) * nleaf);
  t->leaf_count = (int*)malloc(sizeof(int) * nleaf);
  int start = 0;
  for (long q = 0; q < nleaf; q++) {
    int b = t->kd_boxnum[q];
    t->leaf_start[b] = start;
    t->leaf_count[b] = t->kd_counts[q];
    start += t->kd_counts[q];
  }
  for (int j = 0; j <= L; j++) {
    long nb = 1L << (2 * j);
    t->nbr[j] = (int*)malloc(sizeof(int) * 8 * nb);
    t->inner[j] = (int*)malloc(sizeof(int) * 16 * nb);
    t->outer[j] = (int*)malloc(sizeof(int) * 24 * nb);
    or_topology(j, t->nbr[j], t->inner[j], t->outer[j]);
    t->box[j] = (or_box*)calloc((size_t)nb, sizeof(or_box));
  }
  t->qs = (double*)calloc((size_t)n, sizeof(double));
  t->ps = (double*)calloc((size_t)n, sizeof(double));
  return t;
}

void or_set_parallel_mode(or_tree* t, int mode) { t->par_mode = mode; }
void or_set_exact_skeleton_rows(or_tree* t, int on) { t->exact_skel_rows = on; }

static void free_box(or_box* b) {
  free(b->rows); free(b->skel); free(b->l2p);
  for (int s = 0; s < 40; s++) free(b->m2l[s]);
  free(b->outv); free(b->inv); free(b->pot);
  memset(b, 0, sizeof(*b));
}

void or_destroy(or_tree* t) {
  if (!t) return;
  for (int j = 0; j <= t->L; j++) {
    long nb = 1L << (2 * j);
    for (long k = 0; k < nb; k++) free_box(&t->box[j][k]);
    free(t->box[j]); free(t->nbr[j]); free(t->inner[j]); free(t->outer[j]);
  }
  free(t->sx); free(t->sy); free(t->perm); free(t->kd_counts); free(t->kd_boxnum);
  free(t->leaf_start); free(t->leaf_count); free(t->qs); free(t->ps);
  free(t);
}

static int il_box(const or_tree* t, int j, long k, int s) {
  return s < 16 ? t->inner[j][16 * k + s] : t->outer[j][24 * k + (s - 16)];
}

typedef struct {
  const or_tree* t;
  const int* r;
  const int* c;
} idx_ctx;

static double idx_entry(const void* c, int i, int j) {
  const idx_ctx* p = (const idx_ctx*)c;
  return p->t->entry(&p->t->k, p->r[i], p->c[j]);
}

/* AFMM2DTree.cpp:997-1007: a box's point list is its leaf points, or the
 * concatenation of its children's skeletons in child order 0..3. */
static void box_rows(or_tree* t, int j, long k) {
  or_box* b = &t->box[j][k];
  if (j == t->L) {
    b->nin = t->leaf_count[k];
    b->rows = (int*)malloc(sizeof(int) * (size_t)(b->nin > 0 ? b->nin : 1));
    for (int i = 0; i < b->nin; i++) b->rows[i] = t->leaf_start[k] + i;
  } else {
    int n = 0;
    for (int c = 0; c < 4; c++) n += t->box[j + 1][4 * k + c].rank;
    b->nin = n;
    b->rows = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    n = 0;
    for (int c = 0; c < 4; c++) {
      or_box* ch = &t->box[j + 1][4 * k + c];
      memcpy(b->rows + n, ch->skel, sizeof(int) * (size_t)ch->rank);
      n += ch->rank;
    }
  }
}

/* AFMM2DTree.cpp:1009-1061 */
static void nodes_box(or_tree* t, int j, long k) {
  or_box* b = &t->box[j][k];
  int ncols = 0;
  for (int s = 0; s < 40; s++) {
    int kb = il_box(t, j, k, s);
    if (kb >= 0) ncols += t->box[j][kb].nin;
  }
  int* cols = (int*)malloc(sizeof(int) * (size_t)(ncols > 0 ? ncols : 1));
  ncols = 0;
  for (int s = 0; s < 40; s++) {
    int kb = il_box(t, j, k, s);
    if (kb < 0) continue;
    or_box* ib = &t->box[j][kb];
    memcpy(cols + ncols, ib->rows, sizeof(int) * (size_t)ib->nin);
    ncols += ib->nin;
  }
  int n1 = b->nin, n2 = ncols;
  int mn = n1 < n2 ? n1 : n2;
  b->rank = 0;
  if (mn > 0) {
    double* U = (double*)malloc(sizeof(double) * (size_t)n1 * mn);
    double* V = (double*)malloc(sizeof(double) * (size_t)n2 * mn);
    double* row = (double*)malloc(sizeof(double) * (size_t)n2);
    unsigned char* ur = (unsigned char*)malloc((size_t)n1);
    unsigned char* uc = (unsigned char*)malloc((size_t)n2);
    int* rp = (int*)malloc(sizeof(int) * (size_t)mn);
    int* cp = (int*)malloc(sizeof(int) * (size_t)mn);
    idx_ctx c = {t, b->rows, cols};
    int r = aca_core(idx_entry, &c, n1, n2, t->tol, rp, cp, U, V, row, ur, uc);
    b->rank = r;
    if (r > 0) {
      b->skel = (int*)malloc(sizeof(int) * (size_t)r);
      int* sc = (int*)malloc(sizeof(int) * (size_t)r);
      for (int q = 0; q < r; q++) { b->skel[q] = b->rows[rp[q]]; sc[q] = cols[cp[q]]; }
      double* D = (double*)malloc(sizeof(double) * (size_t)r * r);
      double* Ac = (double*)malloc(sizeof(double) * (size_t)n1 * r);
      for (int a = 0; a < r; a++)
        for (int q = 0; q < r; q++) D[(size_t)a * r + q] = t->entry(&t->k, b->skel[a], sc[q]);
      for (int i = 0; i < n1; i++)
        for (int q = 0; q < r; q++) Ac[(size_t)i * r + q] = t->entry(&t->k, b->rows[i], sc[q]);
      b->l2p = (double*)malloc(sizeof(double) * (size_t)n1 * r);
      l2p_solve(r, D, n1, Ac, b->l2p);
      /* Row rp[q] of Ac is row q of D, so row rp[q] of L2P is e_q up to the
       * solve's rounding; the product applies it as the identity. */
      if (t->exact_skel_rows)
        for (int q = 0; q < r; q++)
          for (int c = 0; c < r; c++) b->l2p[(size_t)rp[q] * r + c] = q == c ? 1.0 : 0.0;
      free(D); free(Ac); free(sc);
    }
    free(U); free(V); free(row); free(ur); free(uc); free(rp); free(cp);
  }
  free(cols);
}

int or_assemble(or_tree* t) {
  if (t->assembled) return 0;
  for (int j = t->L; j >= 2; j--) {
    long nb = 1L << (2 * j);
    for (long k = 0; k < nb; k++) box_rows(t, j, k);
    /* The reference runs the boxes of a level serially (1063-1068) with the
     * parallelism inside getRow/getCol; boxes are independent, so they run in
     * parallel here.  Results are identical. */
#pragma omp parallel for schedule(dynamic, 4)
    for (long k = 0; k < nb; k++) nodes_box(t, j, k);
  }
  /* M2L blocks K(skel_k, skel_kIL) (AFMM2DTree.cpp:1070-1097). */
  for (int j = 2; j <= t->L; j++) {
    long nb = 1L << (2 * j);
#pragma omp parallel for schedule(dynamic, 16)
    for (long k = 0; k < nb; k++) {
      or_box* b = &t->box[j][k];
      b->outv = (double*)calloc((size_t)(b->rank > 0 ? b->rank : 1), sizeof(double));
      b->inv = (double*)calloc((size_t)(b->rank > 0 ? b->rank : 1), sizeof(double));
      if (j == t->L) b->pot = (double*)calloc((size_t)(b->nin > 0 ? b->nin : 1), sizeof(double));
      for (int s = 0; s < 40; s++) {
        int kb = il_box(t, j, k, s);
        if (kb < 0) continue;
        or_box* ib = &t->box[j][kb];
        int r1 = b->rank, r2 = ib->rank;
        double* blk = (double*)malloc(sizeof(double) * (size_t)(r1 * r2 > 0 ? r1 * r2 : 1));
        for (int a = 0; a < r1; a++)
          for (int q = 0; q < r2; q++)
            blk[(size_t)a * r2 + q] = t->entry(&t->k, b->skel[a], ib->skel[q]);
        b->m2l[s] = blk;
      }
    }
  }
  t->assembled = 1;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Matvec: AFMM2D.cpp:46-65; AFMM2DTree.cpp:1120-1225. */

static void m2m_box(or_tree* t, int j, long k) {
  or_box* b = &t->box[j][k];
  int r = b->rank, n = b->nin;
  if (r == 0) return;
  const double* src;
  double* tmp = NULL;
  if (j == t->L) {
    src = t->qs + t->leaf_start[k];
  } else {
    tmp = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int o = 0;
    for (int c = 0; c < 4; c++) {
      or_box* ch = &t->box[j + 1][4 * k + c];
      memcpy(tmp + o, ch->outv, sizeof(double) * (size_t)ch->rank);
      o += ch->rank;
    }
    src = tmp;
  }
  for (int c = 0; c < r; c++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += b->l2p[(size_t)i * r + c] * src[i];
    b->outv[c] = s;
  }
  free(tmp);
}

static void m2l_box(or_tree* t, int j, long k) {
  or_box* b = &t->box[j][k];
  int r = b->rank;
  for (int a = 0; a < r; a++) b->inv[a] = 0.0;
  for (int s = 0; s < 40; s++) {
    int kb = il_box(t, j, k, s);
    if (kb < 0) continue;
    or_box* ib = &t->box[j][kb];
    int r2 = ib->rank;
    const double* blk = b->m2l[s];
    for (int a = 0; a < r; a++) {
      double acc = 0.0;
      for (int q = 0; q < r2; q++) acc += blk[(size_t)a * r2 + q] * ib->outv[q];
      b->inv[a] += acc;
    }
  }
}

static void l2l_box(or_tree* t, int j, long k) {
  or_box* b = &t->box[j][k];
  int r = b->rank, n = b->nin;
  if (j == t->L) {
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int c = 0; c < r; c++) s += b->l2p[(size_t)i * r + c] * b->inv[c];
      b->pot[i] = s;
    }
    return;
  }
  int o = 0;
  for (int c4 = 0; c4 < 4; c4++) {
    or_box* ch = &t->box[j + 1][4 * k + c4];
    for (int i = 0; i < ch->rank; i++) {
      double s = 0.0;
      for (int c = 0; c < r; c++) s += b->l2p[(size_t)(o + i) * r + c] * b->inv[c];
      ch->inv[i] += s;
    }
    o += ch->rank;
  }
}

static void near_box(or_tree* t, long k) {
  int L = t->L;
  or_box* b = &t->box[L][k];
  int n = t->leaf_count[k], st = t->leaf_start[k];
  double* R = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * 512);
  for (int s = 0; s < 9; s++) {
    long kb = s < 8 ? t->nbr[L][8 * k + s] : k;
    if (kb < 0) continue;
    int m = t->leaf_count[kb], sb = t->leaf_start[kb];
    if ((size_t)n * m > (size_t)(n > 0 ? n : 1) * 512) {
      free(R);
      R = (double*)malloc(sizeof(double) * (size_t)n * m);
    }
    /* the reference materialises each block per matvec (1204, 1209) */
    for (int i = 0; i < n; i++)
      for (int q = 0; q < m; q++) R[(size_t)i * m + q] = t->entry(&t->k, st + i, sb + q);
    for (int i = 0; i < n; i++) {
      double acc = 0.0;
      for (int q = 0; q < m; q++) acc += R[(size_t)i * m + q] * t->qs[sb + q];
      b->pot[i] += acc;
    }
  }
  free(R);
}

/* The near-field term alone (evaluate_NearField, AFMM2DTree.cpp:1199-1213):
 * for every leaf, the sum over its 8 neighbours and itself of K(leaf,
 * neighbour) q(neighbour), in the order near_box adds them; y in the
 * caller's order.  Needs no assembly. */
void or_near_field(const or_tree* t, const double* q, double* y) {
  int L = t->L;
  long nl = 1L << (2 * L);
  double* qs = (double*)malloc(sizeof(double) * (size_t)(t->n > 0 ? t->n : 1));
  double* ps = (double*)calloc((size_t)(t->n > 0 ? t->n : 1), sizeof(double));
  for (int i = 0; i < t->n; i++) qs[i] = q[t->perm[i]];
#pragma omp parallel for schedule(dynamic, 16)
  for (long k = 0; k < nl; k++) {
    int n = t->leaf_count[k], st = t->leaf_start[k];
    for (int s = 0; s < 9; s++) {
      long kb = s < 8 ? t->nbr[L][8 * k + s] : k;
      if (kb < 0) continue;
      int m = t->leaf_count[kb], sb = t->leaf_start[kb];
      for (int i = 0; i < n; i++) {
        double acc = 0.0;
        for (int c = 0; c < m; c++) acc += t->entry(&t->k, st + i, sb + c) * qs[sb + c];
        ps[st + i] += acc;
      }
    }
  }
  for (int i = 0; i < t->n; i++) y[t->perm[i]] = ps[i];
  free(qs);
  free(ps);
}

void or_matvec(or_tree* t, const double* q, double* y) {
  int L = t->L;
  for (int i = 0; i < t->n; i++) t->qs[i] = q[t->perm[i]];
  for (int j = L; j >= 2; j--) {
    long nb = 1L << (2 * j);
#pragma omp parallel for schedule(dynamic, 16)
    for (long k = 0; k < nb; k++) m2m_box(t, j, k);
  }
  if (t->par_mode == 0) {
    /* the reference parallelises M2L over levels only (1157) */
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 2; j <= L; j++) {
      long nb = 1L << (2 * j);
      for (long k = 0; k < nb; k++) m2l_box(t, j, k);
    }
  } else {
    for (int j = 2; j <= L; j++) {
      long nb = 1L << (2 * j);
#pragma omp parallel for schedule(dynamic, 16)
      for (long k = 0; k < nb; k++) m2l_box(t, j, k);
    }
  }
  for (int j = 2; j <= L; j++) {
    long nb = 1L << (2 * j);
#pragma omp parallel for schedule(dynamic, 16)
    for (long k = 0; k < nb; k++) l2l_box(t, j, k);
  }
  long nl = 1L << (2 * L);
#pragma omp parallel for schedule(dynamic, 16)
  for (long k = 0; k < nl; k++) near_box(t, k);
  for (long k = 0; k < nl; k++) {
    or_box* b = &t->box[L][k];
    memcpy(t->ps + t->leaf_start[k], b->pot, sizeof(double) * (size_t)t->leaf_count[k]);
  }
  for (int i = 0; i < t->n; i++) y[t->perm[i]] = t->ps[i];
}

/* ------------------------------------------------------------------------- */
/* Accessors */

void or_get_sorted_xy(const or_tree* t, double* xy) {
  for (int i = 0; i < t->n; i++) { xy[2 * i] = t->sx[i]; xy[2 * i + 1] = t->sy[i]; }
}
void or_get_perm(const or_tree* t, int64_t* perm) {
  memcpy(perm, t->perm, sizeof(int64_t) * (size_t)t->n);
}
void or_get_kd_counts(const or_tree* t, int* c) {
  memcpy(c, t->kd_counts, sizeof(int) * ((size_t)1 << (2 * t->L)));
}
void or_get_kd_boxnum(const or_tree* t, int* c) {
  memcpy(c, t->kd_boxnum, sizeof(int) * ((size_t)1 << (2 * t->L)));
}
void or_get_leaf_range(const or_tree* t, int box, int* start, int* count) {
  *start = t->leaf_start[box];
  *count = t->leaf_count[box];
}
int or_get_rank(const or_tree* t, int j, int k) { return t->box[j][k].rank; }
int or_get_nin(const or_tree* t, int j, int k) { return t->box[j][k].nin; }
void or_get_skeleton(const or_tree* t, int j, int k, int* idx) {
  memcpy(idx, t->box[j][k].skel, sizeof(int) * (size_t)t->box[j][k].rank);
}
void or_get_l2p(const or_tree* t, int j, int k, double* l2p) {
  const or_box* b = &t->box[j][k];
  memcpy(l2p, b->l2p, sizeof(double) * (size_t)b->nin * b->rank);
}
int or_get_m2l(const or_tree* t, int j, int k, int s, double* blk) {
  int kb = il_box(t, j, k, s);
  if (kb < 0) return 0;
  const or_box* b = &t->box[j][k];
  memcpy(blk, b->m2l[s], sizeof(double) * (size_t)b->rank * t->box[j][kb].rank);
  return 1;
}

void or_direct_rows(const or_tree* t, int nrows, const int* rows, const double* qs, double* out) {
#pragma omp parallel for schedule(static)
  for (int a = 0; a < nrows; a++) {
    int i = rows[a];
    double s = 0.0;
    for (int j = 0; j < t->n; j++) s += t->entry(&t->k, i, j) * qs[j];
    out[a] = s;
  }
}

void or_direct(const or_tree* t, const double* q, double* y) {
  int n = t->n;
  size_t m = (size_t)(n > 0 ? n : 1);
  double* qs = (double*)calloc(m, sizeof(double));
  double* ys = (double*)calloc(m, sizeof(double));
  int* rows = (int*)calloc(m, sizeof(int));
  for (int i = 0; i < n; i++) { qs[i] = q[t->perm[i]]; rows[i] = i; }
  or_direct_rows(t, n, rows, qs, ys);
  for (int i = 0; i < n; i++) y[t->perm[i]] = ys[i];
  free(qs); free(ys); free(rows);
}

double or_compute_error(const or_tree* t, int box) {
  int st = t->leaf_start[box], n = t->leaf_count[box];
  int* rows = (int*)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
  double* tr = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (int i = 0; i < n; i++) rows[i] = st + i;
  or_direct_rows(t, n, rows, t->qs, tr);
  double e = 0.0, d = 0.0;
  for (int i = 0; i < n; i++) {
    double df = tr[i] - t->ps[st + i];
    e += df * df;
    d += tr[i] * tr[i];
  }
  free(rows); free(tr);
  return sqrt(e) / sqrt(d);
}
