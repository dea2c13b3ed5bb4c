/*
 * afmm2d_oracle.h -- CPU restatement of the AFMM2D reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (afmm2d_amd, libafmm2d.so) never
 * links, loads or calls it.
 *
 * Every function follows a reference file:line (paths relative to the
 * reference repository root); see the comments in afmm2d_oracle.c.
 */
#ifndef AFMM2D_ORACLE_H
#define AFMM2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_KERNEL_LOG = 0, OR_KERNEL_INV_R = 1, OR_KERNEL_GAUSS = 2, OR_KERNEL_REG_INV_R = 3,
       OR_KERNEL_USER = 4, OR_KERNEL_USER_SYM = 5, OR_KERNEL_USER_ANTI = 6 };

typedef struct or_tree or_tree;

/* nLevels = (unsigned)(log((double)(N/M))/log(4))  (AFMM2D.cpp:7). */
int or_nlevels(int n, int leaf);

/* Kernel value on coordinates; R2 == 0 (the diagonal) gives the kernel's
 * diagonal value: 0 for log, 1/r and a/R, 1 for the Gaussian. */
double or_kernel_value(int kernel_id, const double* kparams, double a,
                       double xi, double yi, double xj, double yj);

/* KD partition (KDTree.cpp:37-153, 200-243).  xy interleaved (x0,y0,x1,...).
 * Outputs sorted xy (interleaved), perm[i] = original index of sorted point i,
 * leaf counts in KD order (4^L) and leaf box numbers in KD order (4^L). */
void or_kd_sort(int n, const double* xy, int L, double* sorted_xy, int64_t* perm,
                int* kd_counts, int* kd_boxnum);

/* Quad-tree interaction lists of every box of one level (AFMM2DTree.cpp:24-967):
 * nbr[4^level][8], inner[4^level][16], outer[4^level][24], -1 = absent. */
void or_topology(int level, int* nbr, int* inner, int* outer);

/* Partially pivoted ACA on explicit point lists (ACA.cpp:12-133).  Returns the
 * rank; row_piv/col_piv receive the pivots (capacity min(n1,n2)). */
int or_aca_points(int kernel_id, const double* kparams, double a, double tol,
                  int n1, const double* rxy, int n2, const double* cxy,
                  int* row_piv, int* col_piv);

/* The same ACA on an explicit row-major n1 x n2 matrix. */
int or_aca_matrix(int n1, int n2, const double* A, double tol, int* row_piv, int* col_piv);

/* X (n1 x r) = Ac (n1 x r) * D^{-1} (r x r) through LU with partial pivoting
 * of D^T (AFMM2DTree.cpp:1057-1059); all row-major. */
int or_solve_right(int r, const double* D, int n1, const double* Ac, double* X);

/* Operator lifecycle, mirroring AFMM (AFMM2D.hpp:14-33). */
or_tree* or_create(int n, int leaf, int tol_pow, const double* xy, int kernel_id,
                   const double* kparams);
void or_destroy(or_tree* t);
int or_assemble(or_tree* t);
void or_matvec(or_tree* t, const double* q, double* y);
/* The near-field term alone (leaf self + neighbour blocks), caller order;
 * needs no assembly. */
void or_near_field(const or_tree* t, const double* q, double* y);
/* 0: OpenMP placement of the reference (M2L parallel over levels only);
 * 1: parallel over boxes everywhere.  Results are identical. */
void or_set_parallel_mode(or_tree* t, int mode);
/* 1: store the skeleton rows of every L2P as the exact identity rows (as the
 * product applies them) instead of the LU solve's result; set before
 * or_assemble. */
void or_set_exact_skeleton_rows(or_tree* t, int on);

/* The ACA's dot-product summation order for the whole library: 0 ascending
 * (the default), 1 the engine's wave order (64
 * strided partial sums, then a butterfly).  Test infrastructure: explains
 * rank differences that come from rounding of the norm estimate only. */
void or_set_aca_wave_sums(int on);

/* Accessors. */
int or_levels(const or_tree* t);
double or_kernel_a(const or_tree* t);
void or_get_sorted_xy(const or_tree* t, double* xy);
void or_get_perm(const or_tree* t, int64_t* perm);
void or_get_kd_counts(const or_tree* t, int* counts);
void or_get_kd_boxnum(const or_tree* t, int* boxnum);
void or_get_leaf_range(const or_tree* t, int box, int* start, int* count);
int or_get_rank(const or_tree* t, int level, int box);
int or_get_nin(const or_tree* t, int level, int box);
void or_get_skeleton(const or_tree* t, int level, int box, int* idx);
void or_get_l2p(const or_tree* t, int level, int box, double* l2p); /* nin x rank row-major */
/* M2L block of interaction-list slot s (0-15 inner, 16-39 outer); returns 0 if absent. */
int or_get_m2l(const or_tree* t, int level, int box, int slot, double* blk); /* rank x rank' row-major */

/* Dense direct sum y = K q over all points, original order
 * (tutorial.cpp:97-131 semantics). */
void or_direct(const or_tree* t, const double* q, double* y);
/* Rows of a subset of sorted points against all points (sorted order q). */
void or_direct_rows(const or_tree* t, int nrows, const int* rows, const double* q_sorted, double* out);
/* Relative 2-norm error of one leaf's rows after the last matvec
 * (AFMM2DTree.cpp:1227-1236). */
double or_compute_error(const or_tree* t, int box);

#ifdef __cplusplus
}
#endif

#endif
