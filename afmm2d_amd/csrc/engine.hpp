// engine.hpp -- the per-problem device state of the AFMM2D engine and the
// launchers each kernel file provides.
//
// HBM layout (levels 2..L, box gid = level_base[j] + k, k the quad-tree id):
//   x, y          KD-sorted point coordinates (N doubles each)
//   perm          KD-sorted index -> caller index (N int32)
//   skx, sky      skeleton points of every box, concatenated in (level, box)
//                 order; box gid owns [soff[gid], soff[gid] + rank[gid])
//   l2p_rm        L2P of every box without its skeleton rows (those are the
//                 identity: row skel_q of L2P is e_q), (nin - rank) x rank
//                 row-major, the other rows in ascending order  (read by M2M)
//   l2p_cm        the same matrices column-major             (read by L2L)
//   rowmap        row of each stored L2P row: box rows t < nin - rank are the
//                 stored rows, rows nin - rank + q the skeleton rows in
//                 skeleton order; a leaf's rows (its points) at
//                 leaf_start, an upper box's rows (its children's skeletons)
//                 at N + the global skeleton offset of its first child
//   m2l           one block per unordered interaction pair (k, b), k < b:
//                 B = K(skeleton_k, skeleton_b), rank_k x rank_b column-major
//                 with leading dimension ld = rank_k rounded up to even (a
//                 zero row when rank_k is odd), so every column starts 16-byte
//                 aligned; K(skeleton_b, skeleton_k) = +-B^T is not stored
//                 (symmetric kernels; -B^T for an antisymmetric part)
//   m2l_part      pair-major M2L partials: pair q holds B out_b (rank_first
//                 values) then B^T out_k (rank_second values), pairs in block
//                 storage order
//   slot_off      kILSlots ints per box gid: where the box's t-th present
//                 interaction slot (slot order) finds its product in
//                 m2l_part, -1 past the present slots
//   near_part     leaf-major near-field partials: leaf A owns 9 slots of
//                 count[A] values at 9 start[A] (slot 0 self, 1 + q
//                 neighbour q), absent neighbours stay zero
//   outv, inv     multipole / local coefficients, indexed like skx
#pragma once

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace afmm {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    if (b <= bytes && p) return;
    release();
    if (b == 0) b = 256;
    AFMM_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

struct LevelInfo {
  int j = 0;
  int nb = 0;                   // boxes at this level (4^j)
  int64_t gbase = 0;            // first gid of the level
  std::vector<int> nin;         // rows of each box's L2P
  std::vector<int> rank;        // ACA rank
  std::vector<int64_t> soff;    // global skeleton offset per box (nb + 1 entries)
  std::vector<int64_t> l2poff;  // element offset into l2p_* (nb + 1)
  int max_nin = 0, max_rank = 0;
  int max_stored = 0;           // largest nin - rank (stored L2P rows)
  // this level's M2L coefficients (rank_k rank_b per pair) and partial
  // columns (rank_k + rank_b per pair)
  int64_t m2l_coeffs = 0, m2l_cols = 0;
  int64_t m2l_fact_coeffs = 0;  // of m2l_coeffs: those stored as ACA factors
};

// Register-tile classes of the M2L pair kernel, cheapest first (VGPRs per
// lane on gfx950: 59, 75, 86, 88, 100, 114, 185): class c holds blocks of up
// to 16 kM2LTileRows[c] rows (leading dimension) and 8 kM2LTileCols[c]
// columns, the whole block in registers.
constexpr int kM2LClasses = 7;
constexpr int kM2LTileRows[kM2LClasses] = {2, 2, 2, 3, 3, 3, 4};
constexpr int kM2LTileCols[kM2LClasses] = {4, 5, 6, 4, 5, 6, 8};
// Streamed classes, after those: class kM2LClasses + s holds blocks of up to
// 16 kM2LStreamRows[s] rows and any number of columns, streamed through
// registers in chunks of columns.  256 rows covers every rank the assembly
// accepts.
constexpr int kM2LStreamClasses = 6;
constexpr int kM2LStreamRows[kM2LStreamClasses] = {4, 6, 8, 10, 12, 16};
constexpr int kM2LAllClasses = kM2LClasses + kM2LStreamClasses;
// Pairs stored as ACA factors U V^T (opt-in), after the dense classes.
constexpr int kM2LFactored = kM2LAllClasses + 1;

inline int m2l_tile_class(int ld, int cols) {
  const int rt = (ld + 15) / 16, ct = (cols + 7) / 8;
  for (int c = 0; c < kM2LClasses; c++)
    if (rt <= kM2LTileRows[c] && ct <= kM2LTileCols[c]) return c;
  for (int s = 0; s < kM2LStreamClasses; s++)
    if (rt <= kM2LStreamRows[s]) return kM2LClasses + s;
  return kM2LAllClasses;  // more than 256 rows: rejected at assembly
}

// Everything the M2L pair kernel needs about one pair, in one record.
// One row pair of a block in the 48-bit storage (M2L_STORE_B48): the
// biased mantissas' low 32 bits, then their top 16 bits packed in one word
// (entry 2i in bits 0-15, 2i + 1 in 16-31).
struct B48Pair {
  uint32_t lo0, lo1;
  uint32_t hi;
};

struct PairDesc {
  int rk, rb;      // ranks of the pair's boxes
  int part_k;      // where B out_b goes (the pair's partials)
  int ld;          // leading dimension of the block: rk rounded up to even
  int64_t off;     // element offset of the block
  int64_t sk, sb;  // skeleton offsets of the boxes
  int part_b;      // where B^T out_k goes: part_k + rk
  int sexp;        // 48-bit block storage: entry = mantissa * 2^sexp (else 0)
};

// Largest interaction list: children of the parent's 9-box neighbourhood
// that are not adjacent, 36 - 9 = 27.  Every box owns this many partial slots.
constexpr int kILSlots = 27;

struct KernelTimer {
  double total_ms = 0.0;
  int64_t launches = 0;
};

// ---- Blocked matvec: several right-hand sides per pass, factors in FP64,
// FP32 or BF16, products on the matrix cores (blocked.hip).
enum Precision { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16 = 2 };
constexpr int kPrecisions = 3;
// Right-hand sides per pass: 16 (one MFMA column tile) or 32 (two).
constexpr int kMrMaxCols = 32;

// One product of the blocked passes: C[c0 + i][:] (+)= sum_k A[16 rt0 + i][k]
// B[row(k)][:] for i < M <= 64, k < K.  A is a box's factor matrix of mbox
// rows and kp columns (K rounded up to the precision's column group) in the
// tiled layout of blocked.hip, in the precision's factor type; B and C are
// row-major vectors of 16 NCT columns in the precision's vector type (FP64
// for FP64, FP32 otherwise); row(k) = bidx ? bidx[b0 + k] : b0 + k.  beta = 1
// adds into C.  Pointers are absolute, so tasks of several problems can share
// a launch.
struct MrTask {
  const void* A;
  const void* B;
  const int* bidx;
  void* C;
  int64_t b0, c0;
  int M, K, kp, mbox, rt0, beta;
  // 0: B rows b0 + k; 1: through the index list bidx; 3: A and B through the
  // micro-column descriptors (MrMicro) at bidx + b0 (the pair-stored M2L)
  int imode, pad;
};

// One micro column (4 columns) of a box's M2L in the FP64 blocked passes.  Every unordered interaction pair (first, second) stores its block
// K(skeleton_first, skeleton_second) once, ceil4(r_first) x ceil4(r_second)
// in 4 x 4 micro-tiles (row-major inside, micro (a, b) at 16 (a nb + b),
// nb = ceil4(r_second) / 4), zero padded.  The first box reads a micro
// column of it (row i of its task: micro row i / 4, 4 contiguous values);
// the second box reads it transposed (its row i is the block's column i:
// micro column i / 4 of a micro row, 4 values 4 apart).
struct MrMicro {
  int64_t a;  // element offset of micro row 0 of this micro column
  int brow;   // the partner's skeleton row of the micro column's first column
  int info;   // bits 0-15: elements between micro rows of the task; 16-24: the
              // last micro row; 25-27: valid columns (1-4, 0 = padding);
              // 28: forward (1) or transposed (0); 29: negated (the
              // transposed block of an antisymmetric kernel)
};

// Near field of one group of up to 64 target points (KD-sorted indices ta ..
// ta + M - 1) against up to 9 source ranges (the leaf's neighbours and
// itself), written to rows out_row .. of the problem's near-field output.
// Sources [0, npair) are pair sources (FP64 plans): the task evaluates
// each such leaf pair once, for its own rows and, transposed, for the
// source leaf's, whose product it stores to the problem's partials at pt
// (the slot of this task's rows in the source leaf's list); the rest are
// applied to the task's rows only.
struct MrNearTask {
  int ta, M, out_row, nsrc;
  int ss[9], sc[9];
  int prob, npair;
  int64_t pt[4];
};

// The near-field partials of one leaf chunk added to its potentials, slot
// after slot: out[(row0 + r) RP + c] += sum_s part[base + s stride +
// (lr0 + r) prs + c] for r < nrows and c < pw (slots keep their places
// whatever the row stride).
struct MrNearRed {
  double* out;
  const double* part;
  int64_t row0, base, stride;  // base: the leaf's slot 0, its first row
  int nrows, nslots;
  int pw;   // columns used in each row (even)
  int prs;  // row stride of the partials (pw or the plan's row width)
  int lr0;  // the chunk's first row within the leaf
  int pad;
};

// Per-problem data of the near field (coordinates, charges, output).
struct MrNearProb {
  const double* x;
  const double* y;
  const void* q;  // charges, rows of 16 NCT vector elements, KD order
  double* out;    // near-field potentials, rows of 16 NCT doubles
  double* part;   // pair sources' transposed products (MrNearTask::pt)
  KParams kp;
  int pw;         // partial columns written (the right-hand sides, rounded up to even)
  int prs;        // the partials' row stride: pw, or the plan's row width
};

// Gather (caller order -> KD order) or final sum and scatter of up to
// kMrIoPoints caller points of one problem.
struct MrIoTask {
  const double* b;        // gather: caller-order charges, column-major, ldb
  double* y;              // scatter: caller-order output, column-major, ldy
  const int* inv_perm;    // caller index -> KD-sorted index
  void* qs;               // KD-order charges (vector type)
  const void* pot;        // KD-order far-field potentials (vector type)
  const double* nearv;    // KD-order near-field potentials (FP64)
  int64_t ldb, ldy;
  int i0, cnt, nrhs, pad;
};
constexpr int kMrIoPoints = 256;

// The factors of one precision, each box's matrix in the tiled layout of
// blocked.hip within a slot of rows x (columns rounded up to 8) elements.
// V: the box's basis, nin x rank; Vt: its transpose, rank x nin; M: the
// box's M2L row block, rank x (sum of its interaction-list boxes' ranks), the
// column blocks in interaction-slot order.
// FP64 uses the reference's interpolative basis; FP32 and BF16 an
// orthonormal basis of the same operator (Q R = L2P, M2L' = R_k M2L R_b^T).
struct MrFactors {
  bool built = false;
  DevBuf V, Vt, M;
};

// A task list on the device, ordered by row tiles (16-row blocks of output,
// 1..4): tasks [cls[c], cls[c + 1]) have c + 1 row tiles, one launch each.
struct MrTaskList {
  DevBuf buf;
  int n = 0;
  int cls[5] = {0, 0, 0, 0, 0};
  int kmax[4] = {0, 0, 0, 0};  // largest K per class
  int mode = 0;                // MrTask::imode of its tasks (one per list)
};

// Task lists and vectors of the blocked matvec for one (precision, column
// tiles) pair.
struct MrPlan {
  int prec = -1, nct = 0;
  DevBuf qs, outv, inv, pot, nearv;
  DevBuf nearp;                 // near-field pair partials (FP64 plans)
  DevBuf near_red;              // their reduction, MrNearRed per leaf chunk
  std::vector<MrNearRed> h_near_red;  // (empty without pairs)
  std::vector<MrTaskList> up, down;  // per level (index j)
  MrTaskList m2l, leaf, near;
  DevBuf probs;
  DevBuf near_ctr;  // work counters of the near field's persistent grid
  // host copies of the task lists (multi-problem batches concatenate them)
  std::vector<std::vector<MrTask>> h_up, h_down;
  std::vector<MrTask> h_m2l, h_leaf;
  std::vector<MrNearTask> h_near;
};

// Launch parameters of the blocked passes that do not change the task
// lists: the M2L's chunk ring depth (3, 4, 6), XCD mapping and rows per
// task (16, 32, 64), the near field's waves per SIMD.  Defaults from the
// environment (AFMM2D_MR_DEPTH, AFMM2D_MR_XCD, AFMM2D_MR_ROWS_M2L,
// AFMM2D_MR_NEAR_WAVES).
struct MrTune {
  int depth = 3, xcd = 0, near_waves = 1, rows_m2l = 64;
};
MrTune mr_default_tune();

// Several problems' blocked matvecs as one: per pass one launch over the
// concatenated task lists of all problems (the upward pass aligned at the
// leaves, the downward pass from level 2, the near field one persistent
// launch per kernel id), optionally captured into one hipGraph and
// replayed.  Charges and outputs are fixed device arrays given at creation
// (caller order, column-major, nrhs columns).  The batch works in the
// problems' own blocked workspaces: a batch run and a call on one of its
// problems must not overlap.
struct MrBatch;
struct Engine;
MrBatch* mr_batch_create(const std::vector<Engine*>& eng, const double* const* d_b,
                         const int64_t* ldb, double* const* d_y, const int64_t* ldy, int nrhs,
                         int prec);
void mr_batch_destroy(MrBatch* b);
// One batched matvec on stream s: the captured graph if there is one, the
// eager launches otherwise.
void mr_batch_run(MrBatch* b, hipStream_t s);
// Times the candidate launch parameters on this batch (eager runs on s) and
// keeps the fastest; results are cached per GPU architecture and shape
// class for the process (and in the file AFMM2D_TUNE_CACHE names, if set).
// A captured graph is captured again with the new parameters.  Returns the
// time per run of the chosen parameters (ms).
double mr_batch_tune(MrBatch* b, hipStream_t s);
// Captures the batched matvec into a graph (on = true) or drops it.
void mr_batch_graph(MrBatch* b, bool on);
struct MrBatchInfo {
  int launches, depth, xcd, near_waves, rows_m2l, captured, tuned, nproblems;
};
MrBatchInfo mr_batch_info(const MrBatch* b);

// Learned rank predictor: a 16-64-64-64-1 FP32 MLP (ReLU hidden layers)
// predicting log2(1 + ACA rank) of a box; trained and evaluated by the
// library's own kernels (rank_model.hip).
constexpr int kRmIn = 16, kRmHid = 64;
constexpr int kRmParams = kRmHid * kRmIn + kRmHid + 2 * (kRmHid * kRmHid + kRmHid) + kRmHid + 1;
struct RankModel {
  int device = 0;
  hipStream_t s = nullptr;
  DevBuf params, grads, adam_m, adam_v;
  DevBuf act[4], out, yb, dout, dact[2], loss_part, wpart;  // per-batch activations
  int cap = 0;                                               // samples they hold
  int64_t step = 0;                                          // Adam steps taken
  RankModel() = default;
  RankModel(const RankModel&) = delete;
  RankModel& operator=(const RankModel&) = delete;
  ~RankModel();
};
RankModel* rank_model_create(int device, uint64_t seed);
void rank_model_set_params(RankModel& m, const float* p);
void rank_model_get_params(RankModel& m, float* p);
// outputs of n samples (n x kRmIn row-major), device arrays, on stream s
void rank_model_predict_device(RankModel& m, const float* d_x, int64_t n, float* d_out,
                               hipStream_t s);
void rank_model_predict(RankModel& m, const float* x, int64_t n, float* out);
// mean squared error of n host samples and its gradient (kRmParams floats)
double rank_model_gradient(RankModel& m, const float* x, const float* y, int n, float* grad);
// Adam over shuffled minibatches (a partial last batch is dropped); returns
// the last epoch's mean squared error
double rank_model_train(RankModel& m, const float* x, const float* y, int64_t n, int epochs,
                        int batch, float lr, uint64_t seed);

struct Engine {
  // problem
  int n = 0, leaf = 0, tol_pow = 0, L = 0, kid = 0, device = 0;
  // the kernel of the whole operator, for direct sums and single entries: kid
  // itself, or K_USER when this engine carries one part of a non-symmetric
  // user kernel
  int kid_full = 0;
  double tol = 0.0;
  KParams kp{0.01, 0.0};

  // learned ACA capacities (null: every box reserves min(rows, columns))
  RankModel* rank_model = nullptr;
  float rank_margin = 0.5f;  // log2 units added to the prediction
  // assembly record: per box (levels L..2) its features and ACA rank
  std::vector<float> rank_feat;
  std::vector<int> rank_label;
  int64_t rank_boxes = 0, rank_fallbacks = 0, aca_scratch_peak = 0;
  int aca_launches = 0;
  double aca_seconds = 0.0;

  // s_aux: set-up work and the near field; s_tree: the upper tree in the
  // schedules that run it beside the leaf-level M2L (created on first use)
  hipStream_t s_aux = nullptr, s_tree = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_tree = nullptr;

  // KD partition
  DevBuf x, y, perm;
  std::vector<int64_t> h_perm;
  std::vector<double> h_xy;  // sorted, interleaved (for host-side entries)
  std::vector<int> kd_counts, kd_boxnum;
  std::vector<int> leaf_start, leaf_count;  // by leaf box number
  int max_leaf = 0;
  DevBuf d_leaf_start, d_leaf_count, d_leaf_nbr;

  // topology: per level, 8/16/24 ints per box, -1 absent
  std::vector<std::vector<int>> nbr, inner, outer;
  int64_t il_pairs = 0, near_pairs = 0;

  // near field, symmetric: every unordered neighbour pair (A < B) and every
  // self pair (A, A) once; near_pair_part[2p], [2p + 1]: where pair p writes
  // A's sums and B's sums (slots of the leaf-major near_part)
  int64_t near_upairs = 0, near_part_total = 0;
  DevBuf near_pair, near_pair_part, near_part;
  DevBuf near_queue;  // chunk counters of the near-field work queue

  // assembly
  bool assembled = false;
  std::vector<LevelInfo> lv;  // indexed by level (0..L), levels < 2 unused
  int64_t skel_total = 0, l2p_total = 0, m2l_total = 0, m2l_cols = 0;
  int max_rank = 0;
  DevBuf skx, sky, skidx;
  DevBuf rank_g, nin_g, soff_g, l2poff_g;
  DevBuf l2p_rm, l2p_cm, rowmap;
  // M2L, symmetric pair storage
  int64_t npairs = 0;
  int pair_max_ld = 0, pair_max_rb = 0;   // largest block shape
  std::vector<int> h_pair_k, h_pair_b;    // box gids: first (block rows), second
  std::vector<int64_t> h_pair_off;        // element offset of each block
  // pairs of group g (0: levels 2..L-1, 1: level L) in tile class c:
  // [m2l_cls_begin[g][c], m2l_cls_end[g][c])
  int64_t m2l_cls_begin[2][kM2LFactored + 1] = {}, m2l_cls_end[2][kM2LFactored + 1] = {};
  // ACA-factored M2L pairs (m2l_factor set before assembly): a pair whose
  // block has an ACA factorisation U V^T of rank p with p (r1 + r2) < r1 r2
  // is stored in m2lf as U (r1 x p) then V (r2 x p), column-major; its
  // PairDesc holds p in ld and the offset into m2lf in off
  bool m2l_factor = false;
  DevBuf m2lf;
  std::vector<int> h_pair_fact;  // per pair: p, or -1 for a dense block
  int64_t m2lf_pairs = 0, m2lf_total = 0, m2lf_dense = 0;  // pairs, factor / dense coefficients
  // Dense block storage (m2l_store set before assembly):
  //   M2L_STORE_F64: m2l, one double per entry;
  //   M2L_STORE_B48: m2l_b48, 48-bit block floating point (b48_split): each
  //   entry a 48-bit integer mantissa, the block's exponent in its
  //   PairDesc; the entries 2i, 2i + 1 of a block (one row pair of a
  //   column) form a 12-byte record (B48Pair) at record index
  //   (off + 2i) / 2.  6 bytes per entry, error <= 2^-46 of the block's
  //   largest entry.  Factored pairs stay FP64 in m2lf.
  int m2l_store = 0;
  DevBuf m2l_b48;
  std::vector<int> h_red_start;           // per box gid (nbox + 1): first entry of h_red_pair
  std::vector<int> h_red_pair;            // per present slot, slot order: pair index
  DevBuf pair_k, pair_b, pair_off, m2l;
  DevBuf m2l_part;                        // pair-major: rank_first + rank_second per pair
  DevBuf slot_off;                        // kILSlots per box gid: partial offset, -1 absent
  DevBuf pair_desc;                       // PairDesc per pair (48 bytes)
  int num_cus = 256;
  int max_lds = 65536;  // LDS bytes a workgroup may allocate
  // schedule options (environment, read at create):
  //   AFMM2D_NEAR_OVERLAP=0   near field on the main stream after M2L
  //   AFMM2D_P2P_WAVES=n      near-field waves per SIMD (persistent grid;
  //                           default: timed at assembly, tune_near_waves)
  //   AFMM2D_NEAR_TUNE=0      the estimate near_waves_per_simd, untimed
  //   AFMM2D_TREE_MODE=m      which upper-tree passes run beside the
  //                           leaf-level M2L (matvec_device's schedules)
  //   AFMM2D_TREE_FUSED=0     upper-tree M2M and L2L one launch per level
  //                           (default: fused over 3-level subtrees)
  bool near_overlap = true;
  int tree_mode = 0;
  bool tree_fused = true;
  int p2p_waves_per_simd = 4;   // set at assembly from the problem's sizes
  bool p2p_waves_fixed = false;  // AFMM2D_P2P_WAVES given
  bool near_tune = true;         // AFMM2D_NEAR_TUNE=0: the estimate, untimed
  // the near field's FP64 tables: kNearLogTab double2 entries of the log
  // table, then kNearExpTab doubles of the exp table
  DevBuf near_tab;
  double create_seconds = 0.0, assemble_seconds = 0.0;

  // matvec state
  DevBuf outv, inv, qs, io_b, io_y;
  bool have_matvec = false;
  std::vector<double> last_qs, last_ps;  // host copies for compute_error

  // blocked matvec: layout shared by the precisions (element offsets per box
  // gid), the interaction-list column index of every M2L row block, the
  // factors per precision and the plans
  bool mr_layout = false;
  std::vector<int64_t> mr_voff, mr_vtoff, mr_moff, mr_coloff;  // per gid
  std::vector<int> mr_vld, mr_vtld, mr_mld, mr_K;              // per gid
  int64_t mr_v_total = 0, mr_vt_total = 0, mr_m_total = 0;     // elements incl. padding
  int64_t mr_m_coeffs = 0, mr_v_coeffs = 0;                    // sum rank K, sum nin rank
  DevBuf mr_colidx, inv_perm_d;
  // pair-stored M2L (FP64): per gid its micro descriptors; per pair
  // its block (element offset, skeleton offsets, ranks)
  bool mr_boxmajor_layout = false;  // mr_moff, mr_coloff, mr_K, mr_colidx (FP32, BF16) built
  std::vector<int64_t> mr_uoff;  // per gid (+1): first descriptor
  struct PairBlk {
    int64_t poff, s1, s2, g1, g2;  // block offset, skeleton offsets, gids
    int r1, r2;
  };
  std::vector<PairBlk> mr_pairs;
  int64_t mr_pm_total = 0;  // pair-layout elements (zero padding included)
  DevBuf mr_micro;
  // the same blocks for the FP64 blocked passes when m2l_store is
  // M2L_STORE_B48: every pair's block and its transpose in 48-bit block
  // floating point (96-byte micro tiles, one exponent per pair), read
  // forward through their own descriptors (tile offset | exponent << 48)
  bool mr_b48 = false;
  DevBuf mr_m48, mr_micro48;
  MrFactors mrf[kPrecisions];
  std::vector<std::unique_ptr<MrPlan>> mr_plans;
  // orthonormal basis (FP32 / BF16): R of every box, rank x rank row-major
  std::vector<int64_t> mr_roff;
  DevBuf mr_R;
  bool mr_ortho = false;

  // profiling
  int profiling = 0;  // 0 off, 1 every phase, 2 the leaf-level M2L only
  std::map<std::string, KernelTimer> timers;
  std::vector<std::array<hipEvent_t, 2>> ev_pool;
  size_t ev_used = 0;
  struct Pending { std::string name; hipEvent_t a, b; };
  std::vector<Pending> pending;

  ~Engine();
  int64_t gid(int j, int k) const { return lv[j].gbase + k; }
  // Event pair around a launch on stream s when profiling is on.
  void prof_begin(const char* name, hipStream_t s, hipEvent_t* a, hipEvent_t* b);
  void prof_end(hipStream_t s, hipEvent_t b);
  void prof_collect();
};

// kd_sort.hip
void kd_partition(Engine& e, const double* h_x, const double* h_y);

// assemble.hip
void assemble_tree(Engine& e);

// api.cpp
int near_waves_per_simd(const Engine& e);
int tune_near_waves(Engine& e, int estimate);

// matvec.hip
void matvec_device(Engine& e, const double* d_b, double* d_y, hipStream_t s);
void near_field_device(Engine& e, const double* d_b, double* d_y, hipStream_t s);
// y += t (n values)
void add_device(double* d_y, const double* d_t, int n, hipStream_t s);
void direct_rows_device(Engine& e, int nrows, const int* d_rows, const double* d_qs, double* d_out,
                        hipStream_t s);
// Direct-sum rows of caller indices d_rows against caller-order charges d_b
// (device arrays), the partial sums in a static device array: no allocation.
constexpr int64_t kNoAllocScratchDoubles = (int64_t)1 << 20;  // 8 MiB
void direct_rows_noalloc(Engine& e, int nrows, const int64_t* d_rows, const double* d_b,
                         double* d_out, hipStream_t s);
void gather_sorted(Engine& e, const double* d_b, double* d_qs, hipStream_t s);

// blocked.hip
// Builds the factors of precision prec (and the FP64 layout they share).
// Frees the ACA scratch kept between assemblies on device (< 0: all).
void release_scratch(int device);
void mr_prepare(Engine& e, int prec);
// y[:, c] = K b[:, c] for c < nrhs (nrhs <= kMrMaxCols), device arrays in
// caller order, column-major with leading dimensions ldb and ldy, on stream s.
void mr_matvec(Engine& e, const double* d_b, int64_t ldb, double* d_y, int64_t ldy, int nrhs,
               int prec, hipStream_t s);
// y[:, c] += t[:, c]
void add_cols_device(double* d_y, int64_t ldy, const double* d_t, int64_t ldt, int n, int nrhs,
                     hipStream_t s);

}  // namespace afmm
