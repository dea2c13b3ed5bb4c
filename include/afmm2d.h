/*
 * afmm2d.h -- C ABI of the MI355X AFMM2D matvec engine (libafmm2d.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an afmm2d_status;
 * afmm2d_last_error() gives a message for the calling thread's last failure.
 * Handles are not thread-safe: one caller thread per handle.  Host-pointer
 * entry points synchronise before returning; the device entry points are
 * ordered on the stream they are given.  Calls on one handle share its
 * buffers, so each call's device work is also ordered after the previous
 * call's, whatever stream that one used.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * the vaishna77/AFMM2D repository).
 */
#ifndef AFMM2D_H
#define AFMM2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AFMM2D_OK = 0,
  AFMM2D_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, bad kernel id, non-finite input */
  AFMM2D_ERR_LEVELS = 2,           /* nLevels = log4(N/M) < 2: no far field (AFMM2D.cpp:7) */
  AFMM2D_ERR_STATE = 3,            /* e.g. matvec before assemble */
  AFMM2D_ERR_OUT_OF_MEMORY = 4,
  AFMM2D_ERR_HIP = 5,              /* a HIP runtime call failed */
  AFMM2D_ERR_UNSUPPORTED = 6,      /* a size beyond what this build handles */
  AFMM2D_ERR_NOT_FOUND = 7         /* missing device or symbol */
} afmm2d_status;

/* Built-in kernels.  The reference takes any user-defined
 * userkernel::getMatrixEntry (kernel.hpp:58, defined in the user's file,
 * e.g. examples/testAFMM2D.cpp:9-20); device code needs the kernel compiled
 * in, so the engine ships this registry.  R2 = |x_i - x_j|^2. */
typedef enum {
  AFMM2D_KERNEL_LOG = 0,       /* 0.5*log(R2), 0 when R2 == 0 */
  AFMM2D_KERNEL_INV_R = 1,     /* 1/sqrt(R2), 0 when R2 == 0 */
  AFMM2D_KERNEL_GAUSSIAN = 2,  /* exp(-R2/kparams[0]), kparams[0] defaults to 0.01 */
  AFMM2D_KERNEL_REG_INV_R = 3, /* R < a ? R/a : a/R, a = 0.5^nLevels (testAFMM2D.cpp:9-20) */
  AFMM2D_KERNEL_USER = 4,      /* the user kernel compiled into this library: a header defining
                                  afmm2d_user_kernel(xi, yi, xj, yj, kparams, a), chosen at build
                                  time (the reference's userkernel::getMatrixEntry, kernel.hpp:58);
                                  default afmm2d_amd/user_kernels/anisotropic_log.h.  May be
                                  non-symmetric: the operator runs as the sum of the two below
                                  (declare a symmetric kernel with AFMM2D_KERNEL_USER_SYMMETRIC) */
  AFMM2D_KERNEL_USER_SYM = 5,  /* (K(x,y) + K(y,x)) / 2 of the user kernel */
  AFMM2D_KERNEL_USER_ANTI = 6, /* (K(x,y) - K(y,x)) / 2 of the user kernel */
  AFMM2D_KERNEL_USER_SYMMETRIC = 7 /* the user kernel, declared symmetric by the caller: one
                                      operator (M2L blocks applied both ways, the near field's
                                      K(j,i) = K(i,j)); afmm2d_check_symmetry samples the claim */
} afmm2d_kernel_id;

/* Precision of the stored factors and of the arithmetic of a matvec
 * (afmm2d_matvec_ex, afmm2d_matvec_device_ex).  The reference is FP64 only
 * (definitions.hpp:5-7).  FP64 applies the reference's own factors.  FP32 and
 * BF16 apply an orthonormal basis of the same operator, built from the FP64
 * factors at first use (or by afmm2d_prepare_precision): E_k = Q_k R_k per
 * box, M2L blocks R_k M2L R_b^T, rounded to the precision, products
 * accumulated in FP32, the near field evaluated in FP32, y formed in FP64.
 * Expect a few 1e-7 (FP32) and 1e-3 .. 1e-2 (BF16) relative error on top of
 * the FP64 operator's. */
typedef enum {
  AFMM2D_PRECISION_FP64 = 0,
  AFMM2D_PRECISION_FP32 = 1,
  AFMM2D_PRECISION_BF16 = 2
} afmm2d_precision;

/* Mode flag for afmm2d_matvec_ex / _device_ex: run the blocked passes (the
 * several-RHS kernels on the matrix cores) also for one FP64 right-hand side,
 * which otherwise runs the single-RHS passes. */
#define AFMM2D_MATVEC_BLOCKED 0x100

typedef enum {
  AFMM2D_LOC_INTERLEAVED = 0,  /* x0,y0,x1,y1,... (N x 2 row-major) */
  AFMM2D_LOC_COLUMN_MAJOR = 1  /* all x then all y (Eigen::MatrixXd N x 2) */
} afmm2d_loc_layout;

typedef struct afmm2d_handle_s* afmm2d_handle;

typedef struct {
  int n;                     /* points */
  int levels;                /* nLevels */
  int tol_pow;
  int kernel_id;
  int max_rank;              /* over all boxes */
  int64_t skeleton_total;    /* sum of ranks over levels 2..L */
  int64_t il_pairs;          /* ordered interaction-list pairs, levels 2..L */
  int64_t near_pairs;        /* ordered leaf neighbour pairs incl. self */
  int64_t m2l_bytes;         /* stored M2L coefficients */
  int64_t m2l_columns;       /* M2L columns over all target boxes (one index each) */
  int64_t l2p_bytes;         /* stored L2P coefficients (one layout) */
  int64_t matvec_bytes;      /* algorithmic HBM bytes of one matvec */
  double create_seconds;     /* KD partition + topology */
  double assemble_seconds;   /* ACA + ID + M2L build */
  int64_t m2l_leaf_bytes;    /* stored M2L coefficients of the leaf level's pairs */
  int64_t m2l_leaf_columns;  /* their partial columns (rank_k + rank_b per pair) */
  int64_t skeleton_leaf;     /* sum of the leaf level's ranks */
  int64_t near_waves;        /* near-field waves per SIMD used beside the M2L */
  int64_t blocked_m2l_coeffs;   /* blocked FP64 passes: stored M2L coefficients (each
                                   interaction block once, padded to 4 x 4 tiles) */
  int64_t blocked_basis_coeffs; /* blocked passes: basis coefficients, sum nin_k rank_k */
  int64_t rank_boxes;           /* boxes with an ACA block (min(rows, columns) > 0) */
  int64_t rank_fallbacks;       /* of them, rerun at full capacity (learned capacities) */
  int64_t aca_scratch_peak;     /* largest ACA scratch of one launch (bytes) */
  int64_t aca_launches;         /* ACA launches of the assembly */
  double aca_seconds;           /* ACA phase of the assembly (incl. capacity prediction) */
  int64_t m2l_factored_pairs;   /* interaction pairs stored as ACA factors U V^T */
  int64_t m2l_factored_bytes;   /* their factor bytes (included in m2l_bytes) */
  int64_t m2l_factored_dense_bytes; /* the bytes their dense blocks would take */
  int64_t blocked_m2l_coeffs_bf16;  /* blocked FP32 and BF16 passes: M2L row-block
                                       coefficients, sum rank_k K_k (each block stored for
                                       both boxes) */
  int64_t m2l_storage;              /* afmm2d_m2l_storage of the dense M2L blocks */
} afmm2d_stats;

/* Message of the calling thread's last failure ("" if none). */
const char* afmm2d_last_error(void);
const char* afmm2d_status_string(int status);

/* Replaces AFMM::AFMM (AFMM2D.cpp:3-39): nLevels = log4(N/M) (line 7), KD
 * partition (KDTree.cpp:200-243), tree and interaction lists
 * (AFMM2DTree.cpp:3-62, 65-967), leaf point ranges (1099-1118).  loc is a
 * host array of N points in the given layout.  kparams may be NULL (4
 * doubles otherwise).  device < 0 uses the current HIP device. */
int afmm2d_create(int n, int min_particles_in_leaf, int tol_pow, const double* loc,
                  int loc_layout, int kernel_id, const double* kparams, int device,
                  afmm2d_handle* out);

/* Replaces AFMM::~AFMM (AFMM2D.cpp:76-82). */
int afmm2d_destroy(afmm2d_handle h);

/* Replaces AFMM::assemble (AFMM2D.cpp:41-44): per-box ACA skeletons and L2P
 * (AFMM2DTree.cpp:991-1068), M2L blocks (1070-1097).  The ACA's scratch
 * (up to 16 GiB, or a quarter of the free memory) stays allocated on the
 * device after the assembly for the next one (freeing and reallocating it
 * made assemblies stall in the allocator); afmm2d_release_scratch frees it.
 * Assemblies in one process run one at a time. */
int afmm2d_assemble(afmm2d_handle h);

/* Frees the ACA scratch kept on device (< 0: every device).  No reference
 * counterpart. */
int afmm2d_release_scratch(int device);

/* Replaces AFMM::computeMatVecProduct (AFMM2D.cpp:46-65): y = K b for nrhs
 * right-hand sides, host arrays in the caller's point order, column-major
 * with leading dimension n. */
int afmm2d_matvec(afmm2d_handle h, const double* b, double* y, int nrhs);

/* Same on device-resident vectors (one RHS each call), ordered on 'stream'
 * (a hipStream_t; NULL = the null stream).  Nothing is copied to the host. */
int afmm2d_matvec_device(afmm2d_handle h, const double* d_b, double* d_y, void* stream);

/* AFMM::computeMatVecProduct (AFMM2D.cpp:46-65) for nrhs right-hand sides at
 * once: Y = K B with B and Y column-major (caller point order) with leading
 * dimensions ldb, ldy >= N.  mode = an afmm2d_precision, optionally
 * | AFMM2D_MATVEC_BLOCKED.  FP64 without the flag applies the single-RHS
 * passes to each column; otherwise the blocked passes run, up to 32 columns
 * per pass, each stored factor read once per pass for all its columns.
 * Host arrays; returns after the result is in y. */
int afmm2d_matvec_ex(afmm2d_handle h, const double* b, int64_t ldb, double* y, int64_t ldy,
                     int nrhs, int mode);

/* Same on device-resident B and Y, ordered on 'stream' (a hipStream_t; NULL =
 * the null stream).  One FP64 column without AFMM2D_MATVEC_BLOCKED is
 * afmm2d_matvec_device. */
int afmm2d_matvec_device_ex(afmm2d_handle h, const double* d_b, int64_t ldb, double* d_y,
                            int64_t ldy, int nrhs, int mode, void* stream);

/* Builds the factors of one precision now rather than at the first matvec
 * that uses it (FP32/BF16: the orthonormal basis; any: the blocked layout).
 * No reference counterpart: the reference has one precision. */
int afmm2d_prepare_precision(afmm2d_handle h, int precision);

/* ---- Batches: many problems' matvecs as one ----------------------------
 * The reference has one AFMM object per problem (AFMM2D.hpp:14-33) and
 * applies them one after another; a batch applies the blocked matvec
 * (afmm2d_matvec_device_ex with AFMM2D_MATVEC_BLOCKED) of several handles at
 * once: one launch per pass over the concatenated work of all problems
 * (upward pass aligned at the leaves, downward pass from level 2, the near
 * field one persistent launch per kernel id), optionally captured into one
 * hipGraph.  Each problem's result equals its own blocked matvec bit for
 * bit.  Handles of non-symmetric user kernels are refused
 * (AFMM2D_ERR_UNSUPPORTED).  A batch works in its handles' blocked
 * workspaces: a batch matvec and a call on one of its handles must not
 * overlap (order them on one stream, or synchronize). */
typedef struct afmm2d_batch_s* afmm2d_batch;

typedef struct {
  int launches;    /* kernel launches per batch matvec */
  int depth;       /* M2L chunk ring depth */
  int xcd;         /* M2L XCD-aware task order (0/1) */
  int near_waves;  /* near-field persistent waves per SIMD */
  int rows_m2l;    /* M2L rows per task */
  int captured;    /* a graph is captured and replayed */
  int tuned;       /* parameters from the autotuner (this batch or its cache) */
  int nproblems;
} afmm2d_batch_info;

/* afmm2d_assemble (AFMM::assemble, AFMM2D.cpp:41-44) of every handle. */
int afmm2d_batch_assemble(const afmm2d_handle* handles, int count);

/* A batch over count assembled handles.  d_b[i], d_y[i]: device arrays of
 * problem i (caller order, column-major, leading dimensions ldb[i], ldy[i]
 * >= N_i, nrhs <= 32 columns), fixed for the batch's life.  precision: an
 * afmm2d_precision; an FP64 batch needs every handle on the same M2L
 * storage (afmm2d_set_m2l_storage).  Parameters come from the autotuner's cache for this
 * GPU and shape class when it holds them (AFMM2D_TUNE_CACHE names a file
 * that keeps it across processes), else from the defaults. */
int afmm2d_batch_create(const afmm2d_handle* handles, int count, const double* const* d_b,
                        const int64_t* ldb, double* const* d_y, const int64_t* ldy, int nrhs,
                        int precision, afmm2d_batch* out);

/* Y_i = K_i B_i for every problem, ordered on 'stream' (NULL = null stream):
 * the captured graph when there is one, else eager launches. */
int afmm2d_batch_matvec(afmm2d_batch b, void* stream);

/* Times the candidate launch parameters (M2L rows per task 32/64, ring depth
 * 3/4, XCD order 0/1, near-field waves 1/2/4) on this batch with eager runs
 * on 'stream', keeps the fastest and caches it for the shape class.  A
 * captured graph is captured again.  ms_per_run (may be NULL): the time per
 * batch matvec of the chosen parameters. */
int afmm2d_batch_tune(afmm2d_batch b, void* stream, double* ms_per_run);

/* on = 1: captures the batch matvec into a hipGraph replayed by every later
 * afmm2d_batch_matvec; on = 0: drops it. */
int afmm2d_batch_capture(afmm2d_batch b, int on);

int afmm2d_batch_get_info(afmm2d_batch b, afmm2d_batch_info* info);
int afmm2d_batch_destroy(afmm2d_batch b);

/* ACA-factored M2L (set before afmm2d_assemble; on = 0 by default).  The
 * reference builds every M2L block K(skeleton_k, skeleton_kIL) densely
 * (AFMM2DTree.cpp:1070-1097) and applies it as one GEMV (1156-1178).  With
 * factoring on, assembly also runs the ACA (ACA.cpp:27-133, same tolerance
 * and stopping rule) on each interaction pair's block; a block whose rank-p
 * factorisation U V^T is smaller (p (r_k + r_b) < r_k r_b) is stored as U
 * and V and applied as U (V^T out_b) and V (U^T out_k) by the single-RHS
 * matvec.  Each factored block adds an error of about the ACA tolerance.
 * The blocked passes (AFMM2D_MATVEC_BLOCKED, FP32/BF16) keep dense blocks.
 * Symmetric kernels only (no effect on a non-symmetric user kernel's
 * antisymmetric part).  No reference counterpart. */
int afmm2d_set_m2l_factoring(afmm2d_handle h, int on);

/* Storage of the dense M2L blocks read by the single-RHS matvec (set before
 * afmm2d_assemble; AFMM2D_M2L_FP64 by default).  The reference keeps each
 * block K(skeleton_k, skeleton_kIL) as an FP64 Eigen matrix
 * (AFMM2DTree.cpp:1070-1097).  AFMM2D_M2L_BFP48 keeps each block in 48-bit
 * block floating point: one power-of-two exponent per block and a 48-bit
 * integer mantissa per entry (6 bytes instead of 8), every entry within
 * 2^-46 (1.4e-14) of the block's largest entry; the products are formed
 * and summed in FP64 from the stored values.  The M2L pass, the matvec's
 * largest HBM stream, then reads a quarter fewer bytes.  ACA-factored
 * pairs stay FP64; the blocked passes build their own factors.
 * afmm2d_get_m2l returns the stored values.  No reference counterpart. */
typedef enum { AFMM2D_M2L_FP64 = 0, AFMM2D_M2L_BFP48 = 1 } afmm2d_m2l_storage;
int afmm2d_set_m2l_storage(afmm2d_handle h, int storage);

/* symmetric = 0 when K(i, j) != K(j, i) on one of 4096 sampled pairs of the
 * handle's points (a diagnostic for AFMM2D_KERNEL_USER_SYMMETRIC; 1 for the
 * built-in kernels). */
int afmm2d_check_symmetry(afmm2d_handle h, int* symmetric);

/* How the pair of (level, box, slot) is stored: factor_rank = p for ACA
 * factors (-1: a dense block); box_first = 1 when the stored block (or
 * U V^T) is K(skeleton_box, skeleton_slot), 0 when it is the transpose.
 * AFMM2D_ERR_NOT_FOUND for an empty slot. */
int afmm2d_get_m2l_storage(afmm2d_handle h, int level, int box, int slot, int* factor_rank,
                           int* box_first);

/* ---- Learned ACA capacities -------------------------------------------
 * The reference's ACA (ACA.cpp:27-133, called per box at
 * AFMM2DTree.cpp:1009-1041) holds min(rows, columns) factor columns per box.
 * A rank model -- a 16-64-64-64-1 FP32 MLP with ReLU hidden layers, trained
 * and evaluated by the library's own kernels -- predicts log2(1 + rank) of
 * each box from AFMM2D_RANK_FEATURES inputs (level, point bounding-box
 * half-width and centre, kernel, tolerance, block shape; see
 * afmm2d_get_rank_samples).  With a model set, assembly sizes each box's ACA
 * scratch to its bucketed prediction and groups boxes of equal capacity;
 * the ACA still decides every rank, and a box that reaches its capacity
 * before its stopping test passes runs again at full capacity, so ranks,
 * skeletons and results equal the plain assembly's.  No reference
 * counterpart. */
#define AFMM2D_RANK_FEATURES 16
#define AFMM2D_RANK_PARAMS 9473
typedef struct afmm2d_rank_model_s* afmm2d_rank_model;

/* He-uniform weights from seed, zero biases, on device (< 0: current). */
int afmm2d_rank_model_create(int device, uint64_t seed, afmm2d_rank_model* out);
int afmm2d_rank_model_destroy(afmm2d_rank_model m);
/* AFMM2D_RANK_PARAMS floats: W1 (64x16 row-major), b1, W2 (64x64), b2, W3,
 * b3, W4 (1x64), b4.  Setting them clears the optimizer state. */
int afmm2d_rank_model_get_params(afmm2d_rank_model m, float* params);
int afmm2d_rank_model_set_params(afmm2d_rank_model m, const float* params);
/* out[i] = prediction for features x[i * AFMM2D_RANK_FEATURES ...]. */
int afmm2d_rank_model_predict(afmm2d_rank_model m, const float* x, int64_t n, float* out);
/* Mean squared error of (x, y) and its gradient (AFMM2D_RANK_PARAMS floats). */
int afmm2d_rank_model_gradient(afmm2d_rank_model m, const float* x, const float* y, int n,
                               float* grad, double* loss);
/* Adam (beta 0.9, 0.999, eps 1e-8) over minibatches of a seeded shuffle,
 * epochs passes; loss: the last epoch's mean squared error (may be NULL). */
int afmm2d_rank_model_train(afmm2d_rank_model m, const float* x, const float* y, int64_t n,
                            int epochs, int batch, float lr, uint64_t seed, double* loss);

/* Assemble h with capacities from model m (NULL: the plain ACA), each
 * 2^(prediction + margin) - 1 bucketed.  The model must outlive the
 * assembly; models and handles are not thread-safe. */
int afmm2d_set_rank_model(afmm2d_handle h, afmm2d_rank_model m, float margin);

/* After assembly: one sample per box of levels L..2 (level L first, boxes in
 * quad-tree order): its AFMM2D_RANK_FEATURES features and its ACA rank.
 * count = the number of samples; nothing is written when capacity is
 * smaller (features and ranks may be NULL to query the count). */
int afmm2d_get_rank_samples(afmm2d_handle h, float* features, int* ranks, int64_t capacity,
                            int64_t* count);

/* nLevels (AFMM2D.hpp:20). */
int afmm2d_levels(afmm2d_handle h, int* levels);

/* sorted_Properties (AFMM2D.hpp:24): perm[i] = caller index of KD-sorted point i. */
int afmm2d_get_permutation(afmm2d_handle h, int64_t* perm);

/* sorted_Locations (AFMM2D.hpp:23), interleaved x,y in KD order. */
int afmm2d_get_sorted_locations(afmm2d_handle h, double* xy);

/* NumberOfParticlesInLeaves in KD order (KDTree.cpp:122-132), 4^L entries. */
int afmm2d_get_leaf_counts(afmm2d_handle h, int* counts);

/* Leaf box 'box' owns KD-sorted points [start, start+count)
 * (AFMM2DTree.cpp:1099-1118). */
int afmm2d_get_leaf_range(afmm2d_handle h, int box, int* start, int* count);

/* Interaction lists of every box of 'level' (FMM2DBox::neighborNumbers,
 * innerNumbers, outerNumbers; AFMM2DBox.hpp:17-19): arrays of 8, 16, 24 ints
 * per box, -1 = absent.  Any pointer may be NULL. */
int afmm2d_get_interactions(afmm2d_handle h, int level, int* nbr, int* inner, int* outer);

/* ACA ranks of every box of 'level' (levels 2..L), after assemble. */
int afmm2d_get_ranks(afmm2d_handle h, int level, int* ranks);

/* incoming_checkPoints of a box (AFMM2DBox.hpp:35): skeleton as KD-sorted
 * point indices in pivot order.  *count receives the rank. */
int afmm2d_get_skeleton(afmm2d_handle h, int level, int box, int* idx, int capacity, int* count);

/* L2P of a box (AFMM2DBox.hpp:25), rows x rank, row-major.  Its skeleton
 * rows are exactly the identity rows (the engine stores only the other rows;
 * the reference's LU solve returns them up to rounding). */
int afmm2d_get_l2p(afmm2d_handle h, int level, int box, double* out, int64_t capacity, int* rows,
                   int* cols);

/* M2L[kIL] of a box (AFMM2DBox.hpp:24) for interaction slot 'slot'
 * (0-15 innerNumbers, 16-39 outerNumbers), rank x rank(kIL), row-major.
 * Returns AFMM2D_ERR_NOT_FOUND if the slot is empty. */
int afmm2d_get_m2l(afmm2d_handle h, int level, int box, int slot, double* out, int64_t capacity,
                   int* rows, int* cols);

/* The near-field term alone, y = sum over each leaf's neighbours and itself
 * of K(leaf, box) b(box) (FMM2DTree::evaluate_NearField,
 * AFMM2DTree.cpp:1199-1213), host arrays in caller order.  Needs no
 * assembly; runs the matvec's own near-field kernel. */
int afmm2d_near_field(afmm2d_handle h, const double* b, double* y);

/* Dense direct sum y = K b on the device (the check of tutorial.cpp:97-131),
 * host arrays in caller order. */
int afmm2d_direct(afmm2d_handle h, const double* b, double* y);

/* Rows of the dense direct sum: out[r] = sum_j K(rows[r], j) b[j], rows given
 * in caller order. */
int afmm2d_direct_rows(afmm2d_handle h, int nrows, const int64_t* rows, const double* b,
                       double* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): afmm2d_direct_rows on device arrays without any allocation.
 * d_rows: nrows caller indices (int64); d_b: N charges in caller order;
 * d_out: nrows results; ordered on 'stream'.  The partial sums go to a
 * static device buffer of 2^20 doubles (8 MiB per GPU, held until the
 * process ends) shared by every call on that GPU: calls that overlap on one
 * GPU race.  nrows > 2^20 is AFMM2D_ERR_UNSUPPORTED.  Results are
 * bit-identical to afmm2d_direct_rows. */
int afmm2d_direct_rows_device_noalloc(afmm2d_handle h, int nrows, const int64_t* d_rows,
                                      const double* d_b, double* d_out, void* stream);

/* FMM2DTree::compute_error (AFMM2DTree.cpp:1227-1236): relative 2-norm error
 * of leaf 'box' against the direct sum, for the last matvec's first RHS. */
int afmm2d_compute_error(afmm2d_handle h, int box, double* err);

/* AFMM::evaluateError (AFMM2D.cpp:67-74) with an explicit seed instead of
 * srand(time(NULL)): picks one leaf box, reports it and its error. */
int afmm2d_evaluate_error(afmm2d_handle h, unsigned seed, int* box, double* err);

/* kernel::getMatrixEntry on the KD-sorted points (kernel.hpp:23-26), host side. */
int afmm2d_get_matrix_entry(afmm2d_handle h, int i, int j, double* value);

/* The built-in kernel on explicit coordinates, host side. */
int afmm2d_kernel_value(int kernel_id, const double* kparams, double a, double xi, double yi,
                        double xj, double yj, double* value);

int afmm2d_get_stats(afmm2d_handle h, afmm2d_stats* st);

/* Per-phase device timing with HIP events recorded on the stream each phase
 * runs on.  level 0: off; 1: every phase ("m2m", "m2m_upper", "m2l",
 * "m2l_upper", "l2l", "p2p", "l2p_out"); 2: only "m2l" (the leaf
 * level's M2L launch), one event pair per matvec, so a timed loop is barely
 * perturbed. */
int afmm2d_set_profiling(afmm2d_handle h, int level);
int afmm2d_get_kernel_time(afmm2d_handle h, const char* name, double* total_ms, int64_t* launches);
int afmm2d_reset_kernel_times(afmm2d_handle h);

#ifdef __cplusplus
}
#endif

#endif
