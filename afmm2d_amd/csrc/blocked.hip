// blocked.hip -- the blocked matvec: up to 32 right-hand sides per pass, the
// factors in FP64, FP32 or BF16, every product on the matrix cores (MFMA).
//
// The reference applies its operator to one vector at a time
// (AFMM2D.cpp:46-65: gather, M2M, M2L, L2L, near field, scatter;
// AFMM2DTree.cpp:1129-1225), each pass a set of small GEMVs.  With R vectors
// every pass becomes a set of small GEMMs with R columns, and one read of a
// factor serves all R of them:
//
//   M2M (levels L..2)  out_k = V_k^T src_k            src = leaf charges or the
//                                                     children's multipoles
//   M2L (all levels)   in_k  = M_k [out_b1; ..; out_bm]   M_k = [M2L(k, b1) .. M2L(k, bm)],
//                                                     the box's interaction list in
//                                                     slot order (inner 0..15, outer
//                                                     0..23), as the reference's ACA
//                                                     columns (AFMM2DTree.cpp:997-1034)
//   L2L (levels 2..L-1) in_children += V_k in_k
//   leaf               pot_k = V_k in_k
//   near field         near_k = sum over the leaf's neighbours N0..N7 and
//                      itself of K(leaf, source) q_source (evaluated here)
//   output             y = pot + near, scattered to the caller's order
//
// All five are one kernel each over a list of tasks (MrTask / MrNearTask)
// with absolute pointers, so one launch can carry the boxes of many problems
// (the batched entry points use that).  A task is one wave: a block of up to
// 64 output rows x 16 NCT columns held in MFMA accumulators, the contraction
// in chunks of 16 (FP64, FP32: four v_mfma_*_16x16x4) or 32 (BF16: one
// v_mfma_f32_16x16x32_bf16 per half of the split vector), the next chunk's
// loads issued before the current one is used.
//
// MFMA lane layouts (checked on gfx950): lane l = i + 16 g with i = l mod 16,
// g = l / 16.  A[i][k] and B[k][j] sit in lane i + 16 g / j + 16 g for the
// k of group g; here lane group g holds k = k0 + 4 g + s in MFMA s (so each
// lane reads four consecutive k of an A row), or k0 + 8 g + t for BF16.  The
// 16 x 16 result D[r][j] sits in lane j + 16 g: FP64 r = 4 v + g, FP32 and
// BF16 r = 4 g + v, for register v.
//
// Precisions.  FP64 stores the reference's interpolative basis (L2P with its
// skeleton rows the exact identity) and K(skeleton_k, skeleton_b) blocks.
// Rounding those factors to FP32 amplifies: the multipoles of an
// interpolative basis are large and cancel inside the smooth M2L blocks
// (measured on the CPU: 4e-6 relative error with FP32 factors at N = 16384,
// tolerance 1e-12).  FP32 and BF16 therefore store an orthonormal basis of
// the same operator: bottom-up, E_k = L2P_k (leaves) or diag(R_children)
// L2P_k, E_k = Q_k R_k (Gram-Schmidt twice, FP64), V_k = Q_k and
// M2L'(k, b) = R_k M2L(k, b) R_b^T (FP64, then rounded).  The near field
// of FP32 and BF16 evaluates the kernel in FP32 (FP64 coordinate differences
// rounded to FP32); BF16 rounds the kernel values and splits the FP32
// vectors into two BF16 halves (hi + lo), so its error is that of the BF16
// factors.  FP32 and BF16 accumulate in FP32 on the matrix cores, add the
// near field's per-source partials with compensated (two-sum) summation,
// and form the output y = pot + near in FP64.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <cstdlib>
#include <cstring>

#include "engine.hpp"
#include "near_math.hpp"

namespace afmm {

namespace {

typedef double mr_d4 __attribute__((ext_vector_type(4)));
typedef float mr_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mr_bf8 __attribute__((ext_vector_type(8)));
typedef double mr_d2 __attribute__((ext_vector_type(2)));
typedef uint32_t mr_u4 __attribute__((ext_vector_type(4)));
typedef uint32_t mr_u2 __attribute__((ext_vector_type(2)));

template <int P>
struct MrT;
template <>
struct MrT<PREC_F64> {
  using A = double;
  using V = double;
  using Acc = mr_d4;
  static constexpr int KC = 16;
};
template <>
struct MrT<PREC_F32> {
  using A = float;
  using V = float;
  using Acc = mr_f4;
  static constexpr int KC = 16;
};
template <>
struct MrT<PREC_BF16> {
  using A = uint16_t;  // bf16 bits
  using V = float;
  using Acc = mr_f4;
  static constexpr int KC = 32;
};

// Tiled layout of a factor matrix of mbox rows and kp columns (kp a multiple
// of the column group KG: 4 for FP64 and FP32, 8 for BF16): chunks of
// KC = 4 KG columns one after another; within a chunk its 16-row tiles in
// row order; within a tile, column group g by column group, the KG elements
// of each row (FP64: first the elements 0-1 of every (g, row), then 2-3).  A
// load instruction of the products, lane i + 16 g taking 16 bytes of row i
// of group g, reads one contiguous KiB (less for a tile of fewer rows), and a
// box's matrix streams in order chunk by chunk.
template <int P>
__host__ __device__ inline int64_t tile_off(int row, int k, int mbox, int kp) {
  constexpr int KG = P == PREC_BF16 ? 8 : 4, KC = 4 * KG;
  const int c0 = k - k % KC;
  const int kc = kp - c0 < KC ? kp - c0 : KC;
  const int ng = kc / KG;
  const int kk = k - c0, g = kk / KG, e = kk - g * KG;
  const int rt = row >> 4, i = row & 15;
  const int rows = mbox - 16 * rt < 16 ? mbox - 16 * rt : 16;
  const int64_t base = (int64_t)c0 * mbox + (int64_t)kc * 16 * rt;
  if (P == PREC_F64) return base + ((int64_t)((e >> 1) * ng + g) * rows + i) * 2 + (e & 1);
  return base + ((int64_t)g * rows + i) * KG + e;
}
inline int col_group(int prec) { return prec == PREC_BF16 ? 8 : 4; }
// FP64 stores every interaction block once (pair-stored M2L, read forward by
// one box and transposed by the other); FP32 and BF16 store each box's row
// block (every block twice), whose 16-byte loads both directions share.
// Either way each direction reads its block from memory, so the traffic is
// the same; FP64 halves its M2L memory at equal speed, FP32's transposed
// 4-byte reads measured 2-8% slower than its row blocks.
inline bool pair_m2l(int prec) { return prec == PREC_F64; }
constexpr int kMrCtrStride = 32;  // ints between work counters (128 bytes)

// Pointers taken from task records are generic; loads through the global
// address space avoid flat instructions.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr_w(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

inline size_t factor_bytes(int prec) { return prec == PREC_F64 ? 8 : prec == PREC_F32 ? 4 : 2; }
inline size_t vector_bytes(int prec) { return prec == PREC_F64 ? 8 : 4; }

// Output row and column of accumulator register v of the 16 x 16 tile.
template <int P>
__device__ __forceinline__ void acc_pos(int i, int g, int v, int& row, int& col) {
  if constexpr (P == PREC_F64) {
    row = 4 * v + g;
    col = i;
  } else {
    row = 4 * g + v;
    col = i;
  }
}

__device__ __forceinline__ mr_bf8 bf16_pack(const float (&x)[8]) {
  mr_bf8 r;
#pragma unroll
  for (int t = 0; t < 8; t++) r[t] = (__bf16)x[t];
  return r;
}

// x = hi + lo with both halves in BF16: the product with a BF16 factor
// carries the vector to about 2^-17 relative.
__device__ __forceinline__ void bf16_split(const float (&x)[8], mr_bf8& hi, mr_bf8& lo) {
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const __bf16 h = (__bf16)x[t];
    hi[t] = h;
    lo[t] = (__bf16)(x[t] - (float)h);
  }
}

// Workgroup index mapped so that each XCD (workgroups go to the 8 XCDs
// round robin) takes one contiguous range of tasks: neighbouring boxes share
// interaction-list vectors, which then meet in the same L2.
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  const int x = b & 7, idx = b >> 3, q = nblk >> 3, rem = nblk & 7;
  return x * q + min(x, rem) + idx;
}

// ---------------------------------------------------------------------------
// The batched product C (+)= A B over a task list, one task per wave.
// A chunk's operands as loaded.  Nothing touches the loaded registers until
// the chunk is multiplied (masks, signs and BF16 unpacking are applied in
// gemm_mma): any instruction on them at load time makes the wave wait for
// the load there, and the chunks meant to be in flight would not be.
template <int P, int NCT>
struct GemmFrag {
  using T = MrT<P>;
  // FP64/FP32: a[rt][s] is A[row][k0 + 4 g + s]
  typename T::A a[P == PREC_BF16 ? 1 : 4][4];
  mr_u4 au[P == PREC_BF16 ? 4 : 1];                     // BF16: 8 k of row tile rt
  float bf[P == PREC_BF16 ? 8 : 1][NCT];               // BF16: the vector before the split
  typename T::V b[P == PREC_BF16 ? 1 : 4][NCT];         // FP64/FP32: B[k0 + 4 g + s][j]
  uint32_t bmask;  // bit s (t for BF16): B row valid, else it counts as zero
  uint32_t neg;    // FP64/FP32: sign bit applied to a (transposed antisymmetric blocks)
};

template <int P, int NCT, int NRT, int IM>
__device__ __forceinline__ void gemm_load(GemmFrag<P, NCT>& f, const MrTask& d,
                                          int k0, int i, int g, const int* sidx) {
  using T = MrT<P>;
  using AT = typename T::A;
  using VT = typename T::V;
  constexpr int RP = 16 * NCT;
  constexpr int KG = P == PREC_BF16 ? 8 : 4;  // k per lane group
  constexpr int KC = 4 * KG;
  const auto A = gptr(static_cast<const AT*>(d.A));
  const auto B = gptr(static_cast<const VT*>(d.B));
  const auto bidx = gptr(d.bidx);
  // the chunk's tiles (tile_off): a lane group past the chunk's last column
  // group reads the last one again (its B entries are zero: k >= kp >= K)
  f.neg = 0;
  f.bmask = 0;
  const int kc = min(KC, d.kp - k0);
  const int ng = kc / KG;
  const int gg = min(g, ng - 1);
  const int64_t cbase = (int64_t)k0 * d.mbox;
#pragma unroll
  for (int rt = 0; rt < NRT; rt++) {
    const int rtg = d.rt0 + rt;
    const int rows = min(16, d.mbox - 16 * rtg);
    const int ii = min(i, rows - 1);
    const int64_t tb = cbase + (int64_t)kc * 16 * rtg;
    if constexpr (P == PREC_F64) {
      const mr_d2 u = *((const __attribute__((address_space(1))) mr_d2*)(A + tb + ((int64_t)gg * rows + ii) * 2));
      const mr_d2 v = *((const __attribute__((address_space(1))) mr_d2*)(A + tb + ((int64_t)(ng + gg) * rows + ii) * 2));
      f.a[rt][0] = u.x;
      f.a[rt][1] = u.y;
      f.a[rt][2] = v.x;
      f.a[rt][3] = v.y;
    } else if constexpr (P == PREC_F32) {
      const mr_f4 u = *((const __attribute__((address_space(1))) mr_f4*)(A + tb + ((int64_t)gg * rows + ii) * 4));
      f.a[rt][0] = u.x;
      f.a[rt][1] = u.y;
      f.a[rt][2] = u.z;
      f.a[rt][3] = u.w;
    } else {
      f.au[rt] = *((const __attribute__((address_space(1))) mr_u4*)(A + tb + ((int64_t)gg * rows + ii) * 8));
    }
  }
#pragma unroll
  for (int s = 0; s < KG; s++) {
    const int k = k0 + KG * g + s;
    const bool ok = k < d.K;
    const int kc2 = ok ? k : d.K - 1;
    int64_t row;
    if constexpr (IM == 0) row = d.b0 + kc2;
    else if constexpr (IM == 1) row = sidx[kc2];
    else row = bidx[d.b0 + kc2];
    f.bmask |= (uint32_t)ok << s;
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      if constexpr (P == PREC_BF16) f.bf[s][ct] = B[row * RP + 16 * ct + i];
      else f.b[s][ct] = B[row * RP + 16 * ct + i];
    }
  }
}

// The pair-stored M2L (FP64, FP32): lane group g of chunk k0 reads micro
// column k0 / 4 + g, a forward or a transposed slice of one pair's block,
// and the 4 partner rows that micro column multiplies (rows past its valid
// columns read as zero).  Every lane issues 4 scalar loads per row tile, so
// the two orientations cost the same and a wave never diverges on them.
// IM 3: the descriptors staged in LDS (mic = the wave's copy, padded to
// whole chunks); IM 4: read from global memory (mic = the task's list).
// One descriptor from global memory (one 16-byte load).
__device__ __forceinline__ MrMicro micro_at(const MrMicro* p, int idx) {
  const mr_u4 v = *gptr(reinterpret_cast<const mr_u4*>(p) + idx);
  MrMicro m;
  m.a = (int64_t)(((uint64_t)v.y << 32) | (uint64_t)v.x);
  m.brow = (int)v.z;
  m.info = (int)v.w;
  return m;
}

typedef __attribute__((address_space(3))) mr_u4 lds_u4;

__device__ __forceinline__ MrMicro micro_unpack(const mr_u4 v) {
  MrMicro m;
  m.a = (int64_t)(((uint64_t)v.y << 32) | (uint64_t)v.x);
  m.brow = (int)v.z;
  m.info = (int)v.w;
  return m;
}

// IM 5 and 6: the FP64 blocks in 48-bit block floating point, both
// orientations stored (every read forward): descriptors as IM 3 and 4, the
// micro column's first tile in the low 48 bits of a and the pair's exponent
// in its top 16; a lane reads its micro row as the four entries' low words
// (one 16-byte load) and top 16 bits (one 8-byte load) and turns each into
// the exact double
// (2^52 + u) 2^s - (2^52 + 2^47) 2^s = m 2^s with one FMA.
template <int P, int NCT, int NRT, int IM>
__device__ __forceinline__ void gemm_load_micro(GemmFrag<P, NCT>& f, const MrTask& d, int k0,
                                                int i, int g, const MrMicro* mic) {
  static_assert(P != PREC_BF16, "the pair-stored M2L is FP64 or FP32");
  static_assert(IM < 5 || P == PREC_F64, "48-bit pair storage is FP64");
  using T = MrT<P>;
  using AT = typename T::A;
  using VT = typename T::V;
  constexpr int RP = 16 * NCT;
  const auto A = gptr(static_cast<const AT*>(d.A));
  const auto B = gptr(static_cast<const VT*>(d.B));
  const int mi = k0 / 4 + g;
  MrMicro u;
  if constexpr (IM == 3 || IM == 5) {
    u = micro_unpack(((const lds_u4*)mic)[mi]);  // one 16-byte LDS read
  } else {
    const int nm = d.K / 4;
    u = micro_at(mic, min(mi, nm - 1));
    if (mi >= nm) u.info &= ~(7 << 25);
  }
  const int astr = u.info & 0xFFFF, amax = (u.info >> 16) & 0x1FF, nval = (u.info >> 25) & 7;
  if constexpr (IM >= 5) {
    const uint64_t ua = (uint64_t)u.a;
    const int64_t t0 = (int64_t)(ua & 0xFFFFFFFFFFFFull);
    const double sc = ldexp(1.0, (int)(int16_t)(ua >> 48)), nbias = -kB48Bias * sc;
    const char* base = static_cast<const char*>(d.A);
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) {
      const int ai = min(4 * (d.rt0 + rt) + (i >> 2), amax);
      const char* tp = base + (t0 + (int64_t)ai * astr) * 96;
      const mr_u4 lo = *gptr(reinterpret_cast<const mr_u4*>(tp + (i & 3) * 16));
      const mr_u2 hw = *((const __attribute__((address_space(1))) mr_u2*)(tp + 64 + (i & 3) * 8));
      const uint32_t h0 = hw.x, h1 = hw.y;
      f.a[rt][0] = fma(__hiloint2double((int)(kB48ExpWord | (h0 & 0xFFFFu)), (int)lo.x), sc, nbias);
      f.a[rt][1] = fma(__hiloint2double((int)(kB48ExpWord | (h0 >> 16)), (int)lo.y), sc, nbias);
      f.a[rt][2] = fma(__hiloint2double((int)(kB48ExpWord | (h1 & 0xFFFFu)), (int)lo.z), sc, nbias);
      f.a[rt][3] = fma(__hiloint2double((int)(kB48ExpWord | (h1 >> 16)), (int)lo.w), sc, nbias);
    }
    f.neg = 0;  // the transposed copies carry their sign
    f.bmask = (1u << nval) - 1u;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int64_t row = u.brow + (s < nval ? s : 0);
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) f.b[s][ct] = B[row * RP + 16 * ct + i];
    }
    return;
  }
  const bool fwd = (u.info >> 28) & 1;
  // 4 loads per row tile in either orientation (forward: a micro row's 4
  // contiguous values; transposed: a micro column's, 4 apart): the same
  // number of loads on every path keeps the compiler's wait counts exact
  const int step = fwd ? 1 : 4;
  const int e0 = fwd ? (i & 3) * 4 : (i & 3);
#pragma unroll
  for (int rt = 0; rt < NRT; rt++) {
    const int ai = min(4 * (d.rt0 + rt) + (i >> 2), amax);
    const int64_t base = u.a + (int64_t)ai * astr + e0;
#pragma unroll
    for (int s = 0; s < 4; s++) f.a[rt][s] = A[base + s * step];
  }
  // sign -1 for an antisymmetric kernel's transposed block (exact)
  f.neg = ((uint32_t)u.info >> 29 & 1u) << 31;
  f.bmask = (1u << nval) - 1u;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int64_t row = u.brow + (s < nval ? s : 0);
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) f.b[s][ct] = B[row * RP + 16 * ct + i];
  }
}

template <int P, int NCT, int NRT>
__device__ __forceinline__ void gemm_mma(const GemmFrag<P, NCT>& f,
                                         typename MrT<P>::Acc (&acc)[4][NCT]) {
  if constexpr (P == PREC_BF16) {
    mr_bf8 bh[NCT], bl[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      float x[8];
#pragma unroll
      for (int t = 0; t < 8; t++) x[t] = (f.bmask >> t) & 1 ? f.bf[t][ct] : 0.0f;
      bf16_split(x, bh[ct], bl[ct]);
    }
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) {
      const mr_bf8 a = __builtin_bit_cast(mr_bf8, f.au[rt]);  // element 2t: low half of word t
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) {
        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh[ct], acc[rt][ct], 0, 0, 0);
        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl[ct], acc[rt][ct], 0, 0, 0);
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int rt = 0; rt < NRT; rt++) {
        typename MrT<P>::A a = f.a[rt][s];
        if constexpr (P == PREC_F64) {
          uint64_t bits = __builtin_bit_cast(uint64_t, a) ^ ((uint64_t)f.neg << 32);
          a = __builtin_bit_cast(double, bits);
        } else {
          a = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, a) ^ f.neg);
        }
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
          const typename MrT<P>::V bv = (f.bmask >> s) & 1 ? f.b[s][ct] : (typename MrT<P>::V)0;
          if constexpr (P == PREC_F64)
            acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[rt][ct], 0, 0, 0);
          else
            acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv, acc[rt][ct], 0, 0, 0);
        }
      }
  }
}

// IM: how B rows are found: 0 contiguous (b0 + k), 1 through the task's index
// list staged in LDS first (dynamic LDS: kmax ints per wave), 2 through the
// index list in global memory.  Staging the indices takes the index load off
// the path of every chunk: the vector loads of the chunks in flight then
// depend on nothing.  Chunks go through a ring of three register sets, two
// chunks' loads in flight while one is multiplied.
template <int P, int NCT, int NRT, int IM, int D>
__global__ __launch_bounds__(256) void k_mr_gemm(const MrTask* __restrict__ tasks, int ntasks,
                                                 int xcd_map, int kmax) {
  using T = MrT<P>;
  using VT = typename T::V;
  using Acc = typename T::Acc;
  constexpr int RP = 16 * NCT;
  constexpr int KC = T::KC;
  extern __shared__ __align__(16) int lds_idx[];
  const int l = threadIdx.x & 63, i = l & 15, g = l >> 4;
  const int blk = xcd_map ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int w = threadIdx.x >> 6;
  const int t = __builtin_amdgcn_readfirstlane(blk * 4 + w);
  if (t >= ntasks) return;
  const MrTask d = tasks[t];
  const int M = d.M, K = d.K;
  int* sidx = lds_idx + w * kmax;
  // IM 3: kmax = micro descriptors per wave
  MrMicro* smic = reinterpret_cast<MrMicro*>(lds_idx) + w * kmax;
  const MrMicro* gmic = reinterpret_cast<const MrMicro*>(d.bidx) + d.b0;
  if constexpr (IM == 1) {
    const auto bidx = gptr(d.bidx);
    for (int k = l; k < K; k += 64) sidx[k] = bidx[d.b0 + k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if constexpr (IM == 3 || IM == 5) {
    // the task's descriptors, padded with a copy of the last one marked
    // empty up to whole chunks (4 per chunk)
    const int nm = K / 4, nmp = (K + KC - 1) / KC * 4;
    lds_u4* sm = (lds_u4*)smic;
    for (int u = l; u < nmp; u += 64) {
      mr_u4 x = *gptr(reinterpret_cast<const mr_u4*>(gmic) + min(u, nm - 1));
      if (u >= nm) x.w &= ~(7u << 25);
      sm[u] = x;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  Acc acc[4][NCT];
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) acc[rt][ct] = Acc{};
  const int nch = (K + KC - 1) / KC;
  if (nch > 0) {
    // ring of D register sets: chunk c + D - 1 is loaded while chunk c is
    // multiplied (static indices: the ring position is a compile-time
    // constant in the unrolled loop)
    GemmFrag<P, NCT> f[D];
    // chunks past the last are loaded again as the last one (cache hits,
    // never multiplied): every ring slot issues its loads unconditionally,
    // so the compiler's wait before a chunk's products counts only that
    // chunk's loads, not the ones just issued for later chunks
    auto load = [&](GemmFrag<P, NCT>& fr, int ch) {
      const int k0 = min(ch, nch - 1) * KC;
      if constexpr (IM == 3 || IM == 5) gemm_load_micro<P, NCT, NRT, IM>(fr, d, k0, i, g, smic);
      else if constexpr (IM == 4 || IM == 6) gemm_load_micro<P, NCT, NRT, IM>(fr, d, k0, i, g, gmic);
      else gemm_load<P, NCT, NRT, IM>(fr, d, k0, i, g, sidx);
    };
#pragma unroll
    for (int u = 0; u < D - 1; u++) load(f[u], u);
    for (int c = 0; c < nch; c += D) {
#pragma unroll
      for (int u = 0; u < D; u++) {
        load(f[(u + D - 1) % D], c + u + D - 1);
        if (c + u < nch) gemm_mma<P, NCT, NRT>(f[u], acc);
      }
    }
  }
  const auto C = gptr_w(static_cast<VT*>(d.C));
#pragma unroll
  for (int rt = 0; rt < NRT; rt++) {
#pragma unroll
    for (int ct = 0; ct < NCT; ct++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row, col;
        acc_pos<P>(i, g, v, row, col);
        row += 16 * rt;
        if (row < M) {
          const auto p = C + (d.c0 + row) * RP + 16 * ct + col;
          const VT x = (VT)acc[rt][ct][v];
          *p = d.beta ? *p + x : x;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Near field: one wave per target group, the kernel evaluated on the fly in
// the MFMA A layout (lane i + 16 g: target row i against sources
// k0 + 4 g + s, or k0 + 8 g + t for BF16), the sources' charges as B.
__device__ __forceinline__ void two_sum(float& s, float& c, float x) {
  const float t = s + x;
  const float bp = t - s;
  c += (s - (t - bp)) + (x - bp);
  s = t;
}

// Leading dimension (doubles) of a wave's FP64 kernel tile in LDS: 64
// target rows by 16 sources, one pad column.
constexpr int kNearTileLd = 17;

template <int P, int NCT, int NRT, int KID>
__device__ __forceinline__ void mr_near_task(const MrNearTask& d,
                                             const MrNearProb* __restrict__ probs,
                                             const double2* tab, double* ktile) {
  using T = MrT<P>;
  using VT = typename T::V;
  using Acc = typename T::Acc;
  constexpr int RP = 16 * NCT;
  constexpr int KC = T::KC;
  const int l = threadIdx.x & 63, i = l & 15, g = l >> 4;
  const MrNearProb& pr = probs[d.prob];
  const KParams kp = pr.kp;
  const auto X = gptr(pr.x);
  const auto Y = gptr(pr.y);
  const auto Q = gptr(static_cast<const VT*>(pr.q));
  const int M = d.M, nsrc = d.nsrc;
  double xt[4], yt[4];
#pragma unroll
  for (int rt = 0; rt < 4; rt++) {
    const int r = d.ta + min(16 * rt + i, M - 1);
    xt[rt] = X[r];
    yt[rt] = Y[r];
  }
  const double ninv_p0 = -1.0 / kp.p0, inv_a = 1.0 / kp.a;
  // FP32 constants: log 0.5 ln 2; Gaussian -log2(e)/p0; a/R 1/a and a
  const float c0 = KID == K_LOG ? 0.34657359f
                   : KID == K_GAUSS ? (float)(-1.4426950408889634 / kp.p0)
                                    : (float)inv_a;
  const float c1 = (float)kp.a;
  constexpr double scale = near_scale<KID>;  // of the FP64 sums
  // FP64 pair sources: the task's own charges as the B operand of the
  // transposed products (lane i + 16 g: target 16 rt + 4 tg + g, column
  // 16 ct + i; zero past the task's rows)
  const int npair = P == PREC_F64 ? __builtin_amdgcn_readfirstlane(d.npair) : 0;
  VT qa[P == PREC_F64 ? NRT : 1][4][NCT];
  if constexpr (P == PREC_F64) {
#pragma unroll
    for (int rt = 0; rt < NRT; rt++)
#pragma unroll
      for (int tg = 0; tg < 4; tg++) {
        const int r = 16 * rt + 4 * tg + g;
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
          qa[rt][tg][ct] = npair > 0 ? Q[(int64_t)(d.ta + min(r, M - 1)) * RP + 16 * ct + i] : 0.0;
          if (r >= M) qa[rt][tg][ct] = 0.0;
        }
      }
  }
  Acc acc[4][NCT];  // FP64: the sum; FP32/BF16: one source's partial
  Acc sh[4][NCT], sl[4][NCT];
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      acc[rt][ct] = Acc{};
      sh[rt][ct] = Acc{};
      sl[rt][ct] = Acc{};
    }
  // The sources as a sequence of steps (source src, chunk k0): the loads of
  // the next step are issued before the current one is evaluated, into the
  // other of two buffers (a step past the last reloads the last one and is
  // not evaluated).  Loaded registers are not touched until their step runs:
  // padding sources are masked at the product (bit s of okm).
  constexpr int KS = P == PREC_BF16 ? 8 : 4;  // sources per lane group and step
  struct Buf {
    double sx[KS], sy[KS];
    float qf[P == PREC_BF16 ? KS : 1][NCT];
    VT q[P == PREC_BF16 ? 1 : KS][NCT];
    uint32_t okm;
    int src, k0, cnt;
  };
  auto load = [&](Buf& b, int src, int k0) {
    const int st = __builtin_amdgcn_readfirstlane(d.ss[src]);
    const int cnt = __builtin_amdgcn_readfirstlane(d.sc[src]);
    b.src = src;
    b.k0 = k0;
    b.cnt = cnt;
    b.okm = 0;
#pragma unroll
    for (int u = 0; u < KS; u++) {
      const int k = k0 + KS * g + u;
      const bool ok = k < cnt;
      const int kc = st + (ok ? k : cnt - 1);
      b.okm |= (uint32_t)ok << u;
      b.sx[u] = X[kc];
      b.sy[u] = Y[kc];
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) {
        if constexpr (P == PREC_BF16) b.qf[u][ct] = Q[(int64_t)kc * RP + 16 * ct + i];
        else b.q[u][ct] = Q[(int64_t)kc * RP + 16 * ct + i];
      }
    }
  };
  auto compute = [&](const Buf& b) {
    if constexpr (P == PREC_BF16) {
      mr_bf8 bh[NCT], bl[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) {
        float x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = (b.okm >> u) & 1 ? b.qf[u][ct] : 0.0f;
        bf16_split(x, bh[ct], bl[ct]);
      }
#pragma unroll
      for (int rt = 0; rt < NRT; rt++) {
        float kv[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
          kv[u] = near_kernel_f32<KID>(xt[rt], yt[rt], b.sx[u], b.sy[u], kp, c0, c1);
        const mr_bf8 a = bf16_pack(kv);
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
          acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh[ct], acc[rt][ct], 0, 0, 0);
          acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl[ct], acc[rt][ct], 0, 0, 0);
        }
      }
    } else {
      VT bq[4][NCT];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) bq[u][ct] = (b.okm >> u) & 1 ? b.q[u][ct] : (VT)0;
      const bool pair = P == PREC_F64 && b.src < npair;
#pragma unroll
      for (int rt = 0; rt < NRT; rt++) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if constexpr (P == PREC_F64) {
            double kv;
            if constexpr (KID >= K_USER) {
              kv = kernel_xy<KID>(xt[rt], yt[rt], b.sx[u], b.sy[u], kp);
            } else {
              const double dx = xt[rt] - b.sx[u], dy = yt[rt] - b.sy[u];
              kv = near_kernel<KID>(fma(dx, dx, dy * dy), kp, ninv_p0, inv_a, tab);
            }
            // the tile for the transposed products: row = target, column =
            // source of the step
            if (pair) ktile[(16 * rt + i) * kNearTileLd + 4 * g + u] = kv;
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
              acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv, bq[u][ct], acc[rt][ct], 0, 0, 0);
          } else {
            const float kv = near_kernel_f32<KID>(xt[rt], yt[rt], b.sx[u], b.sy[u], kp, c0, c1);
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
              acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(kv, bq[u][ct], acc[rt][ct], 0, 0, 0);
          }
        }
      }
      if constexpr (P == PREC_F64) {
        if (pair) {
          // the source leaf's rows of this step: D' (16 sources x 16 columns)
          // = K^T (sources x the task's targets, 4 per product, read
          // transposed from the tile) x the task's charges, with the
          // kernel's transpose sign (-1 for an antisymmetric part); stored to
          // the pair's partial slot, rows past the source leaf's count
          // dropped.  A wave's LDS operations run in order, so the tile reads
          // see this step's writes and the next step's writes follow these
          // reads; the wave barriers only keep the compiler from moving them
          // (no wait on the next step's loads in flight).  Two accumulator
          // chains (even and odd target groups) keep the products apart.
          __builtin_amdgcn_wave_barrier();
          Acc accT[2][NCT];
#pragma unroll
          for (int h = 0; h < 2; h++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) accT[h][ct] = Acc{};
#pragma unroll
          for (int rt = 0; rt < NRT; rt++)
#pragma unroll
            for (int tg = 0; tg < 4; tg++) {
              const double a = ktile[(16 * rt + 4 * tg + g) * kNearTileLd + i];
#pragma unroll
              for (int ct = 0; ct < NCT; ct++)
                accT[tg & 1][ct] =
                    __builtin_amdgcn_mfma_f64_16x16x4f64(a, qa[rt][tg][ct], accT[tg & 1][ct], 0, 0, 0);
            }
          __builtin_amdgcn_wave_barrier();
          const auto part = gptr_w(pr.part + d.pt[b.src]);
#pragma unroll
          for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
              int row, col;
              acc_pos<P>(i, g, v, row, col);
              const int k = b.k0 + row;
              // K(source, target) = kTransposeSign K(target, source)^T
              if (k < b.cnt && 16 * ct + col < pr.pw)
                part[(int64_t)k * pr.prs + 16 * ct + col] =
                    (kTransposeSign<KID> * scale) * (accT[0][ct][v] + accT[1][ct][v]);
            }
        }
      }
    }
  };
  // end of a source (FP32/BF16): its partial sum added to the compensated
  // total, the partial restarted
  auto close_source = [&]() {
    if constexpr (P != PREC_F64) {
#pragma unroll
      for (int rt = 0; rt < NRT; rt++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            float s_ = sh[rt][ct][v], c_ = sl[rt][ct][v];
            two_sum(s_, c_, acc[rt][ct][v]);
            sh[rt][ct][v] = s_;
            sl[rt][ct][v] = c_;
            acc[rt][ct][v] = 0.0f;
          }
    }
  };
  if (nsrc > 0) {
    Buf b0, b1;
    int src = 0, k0 = 0;  // the step to evaluate next
    load(b0, 0, 0);
    for (;;) {
      // the step after (src, k0), or (src, k0) again past the end
      int ns = src, nk = k0 + KC;
      if (nk >= __builtin_amdgcn_readfirstlane(d.sc[src])) {
        ns = src + 1;
        nk = 0;
      }
      const bool more = ns < nsrc;
      load(b1, more ? ns : src, more ? nk : k0);
      compute(b0);
      if (ns != src || !more) close_source();
      if (!more) break;
      src = ns;
      k0 = nk;
      ns = src;
      nk = k0 + KC;
      if (nk >= __builtin_amdgcn_readfirstlane(d.sc[src])) {
        ns = src + 1;
        nk = 0;
      }
      const bool more2 = ns < nsrc;
      load(b0, more2 ? ns : src, more2 ? nk : k0);
      compute(b1);
      if (ns != src || !more2) close_source();
      if (!more2) break;
      src = ns;
      k0 = nk;
    }
  }
  const auto out = gptr_w(pr.out);
#pragma unroll
  for (int rt = 0; rt < NRT; rt++) {
#pragma unroll
    for (int ct = 0; ct < NCT; ct++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row, col;
        acc_pos<P>(i, g, v, row, col);
        row += 16 * rt;
        if (row < M) {
          double val;
          if constexpr (P == PREC_F64) val = scale * acc[rt][ct][v];
          else val = (double)sh[rt][ct][v] + (double)sl[rt][ct][v];
          out[(int64_t)(d.out_row + row) * RP + 16 * ct + col] = val;
        }
      }
  }
}

// The near field's launch: a persistent grid of a few waves per SIMD (the
// memory-bound tree passes run beside it) taking tasks from a counter, one
// atomic per wave and task.
template <int P, int NCT, int NRT, int KID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(P == PREC_F64 && NCT == 1 ? 2 : 1)))
void k_mr_p2p(const MrNearTask* __restrict__ tasks, int ntasks,
                                                const MrNearProb* __restrict__ probs,
                                                const double2* __restrict__ ntab,
                                                int* __restrict__ ctr) {
  __shared__ double2 tab[P == PREC_F64 ? kNearTabLen<KID> : 1];
  // FP64: each wave's kernel tile of its pair steps
  __shared__ double ktile[P == PREC_F64 ? 4 * 16 * NRT * kNearTileLd : 1];
  if constexpr (P == PREC_F64) {
    load_near_tab<KID>(tab, ntab, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  const int l = threadIdx.x & 63;
  double* wtile = ktile + (P == PREC_F64 ? (threadIdx.x >> 6) * 16 * NRT * kNearTileLd : 0);
  for (;;) {
    int t = 0;
    if (l == 0) t = atomicAdd(ctr, 1);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= ntasks) break;
    mr_near_task<P, NCT, NRT, KID>(tasks[t], probs, tab, wtile);
  }
}

// Near field with the sources staged in LDS: a workgroup takes one task
// (up to 64 target rows of a leaf, one 16-row tile per wave) and walks its
// source leaves in stages of kNearStage points.  The whole workgroup loads a
// stage's coordinates and charges (whole rows, coalesced) into registers
// while the previous stage is evaluated, then into the other of two LDS
// buffers; every wave evaluates its 16 targets against the stage (the
// kernel on the fly into the MFMA A operand, the charges from LDS as B), so
// the sources are read from memory once per task instead of once per wave.
// Sources past a leaf's count read as the leaf's last point with charge 0.
// Tasks are taken largest first (the list is ordered by row tiles).
constexpr int kNearStage = 64;

template <int P, int NCT, int KID>
__global__ __launch_bounds__(256) void k_mr_p2p_lds(const MrNearTask* __restrict__ tasks,
                                                    int ntasks,
                                                    const MrNearProb* __restrict__ probs,
                                                    const double2* __restrict__ ntab,
                                                    int* __restrict__ ctr) {
  using T = MrT<P>;
  using VT = typename T::V;
  using Acc = typename T::Acc;
  constexpr int RP = 16 * NCT;
  constexpr int KS = P == PREC_BF16 ? 8 : 4;  // sources per lane group and product
  constexpr int KC = 4 * KS;                   // sources per product
  constexpr int SB = kNearStage;
  constexpr int NLD = (SB * RP + 255) / 256;  // charges per thread and stage
  __shared__ double2 tab[P == PREC_F64 ? kNearTabLen<KID> : 1];
  __shared__ double2 sxy[2][SB];
  __shared__ VT sq[2][SB * RP];
  __shared__ int s_task;
  if constexpr (P == PREC_F64) load_near_tab<KID>(tab, ntab, threadIdx.x, 256);
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, i = l & 15, g = l >> 4;
  for (;;) {
    __syncthreads();  // the previous task's LDS reads are done
    if (tid == 0) s_task = atomicAdd(ctr, 1);
    __syncthreads();
    const int tq = s_task;
    if (tq >= ntasks) break;  // the same for every thread
    const MrNearTask& d = tasks[ntasks - 1 - tq];
    const MrNearProb& pr = probs[d.prob];
    const KParams kp = pr.kp;
    const auto X = gptr(pr.x);
    const auto Y = gptr(pr.y);
    const auto Q = gptr(static_cast<const VT*>(pr.q));
    const int M = d.M, nsrc = d.nsrc;
    const bool act = 16 * w < M;  // the wave has target rows
    const int tr = min(16 * w + i, M - 1);
    const double xt = X[d.ta + tr], yt = Y[d.ta + tr];
    const double ninv_p0 = -1.0 / kp.p0, inv_a = 1.0 / kp.a;
    const float c0 = KID == K_LOG ? 0.34657359f
                     : KID == K_GAUSS ? (float)(-1.4426950408889634 / kp.p0)
                                      : (float)inv_a;
    const float c1 = (float)kp.a;
    constexpr double scale = near_scale<KID>;  // of the FP64 sums
    Acc acc[NCT], sh[NCT], sl[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      acc[ct] = Acc{};
      sh[ct] = Acc{};
      sl[ct] = Acc{};
    }
    double2 pxy = make_double2(0.0, 0.0);
    VT pq[NLD];
    auto fetch = [&](int src, int k0) {
      const int st = __builtin_amdgcn_readfirstlane(d.ss[src]);
      const int nv = min(SB, __builtin_amdgcn_readfirstlane(d.sc[src]) - k0);
      if (tid < SB) {
        const int k = st + k0 + min(tid, nv - 1);
        pxy = make_double2(X[k], Y[k]);
      }
#pragma unroll
      for (int u = 0; u < NLD; u++) {
        const int e = tid + 256 * u, row = e / RP, col = e - row * RP;
        pq[u] = (VT)0;
        if (e < SB * RP && row < nv) pq[u] = Q[(int64_t)(st + k0 + row) * RP + col];
      }
    };
    auto commit = [&](int b) {
      if (tid < SB) sxy[b][tid] = pxy;
#pragma unroll
      for (int u = 0; u < NLD; u++) {
        const int e = tid + 256 * u;
        if (e < SB * RP) sq[b][e] = pq[u];
      }
    };
    auto compute = [&](int b, int nv) {
      if (!act) return;
      const int nch = (nv + KC - 1) / KC;
      for (int ch = 0; ch < nch; ch++) {
        const int kb = ch * KC + KS * g;  // this lane group's first source of the product
        if constexpr (P == PREC_BF16) {
          float kv[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const double2 sp = sxy[b][kb + u];
            kv[u] = near_kernel_f32<KID>(xt, yt, sp.x, sp.y, kp, c0, c1);
          }
          const mr_bf8 a = bf16_pack(kv);
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            float x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = sq[b][(kb + u) * RP + 16 * ct + i];
            mr_bf8 bh, bl;
            bf16_split(x, bh, bl);
            acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc[ct], 0, 0, 0);
            acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc[ct], 0, 0, 0);
          }
        } else {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const double2 sp = sxy[b][kb + u];
            typename T::A kv;
            if constexpr (P == PREC_F64) {
              if constexpr (KID >= K_USER) {
                kv = kernel_xy<KID>(xt, yt, sp.x, sp.y, kp);
              } else {
                const double dx = xt - sp.x, dy = yt - sp.y;
                kv = near_kernel<KID>(fma(dx, dx, dy * dy), kp, ninv_p0, inv_a, tab);
              }
            } else {
              kv = near_kernel_f32<KID>(xt, yt, sp.x, sp.y, kp, c0, c1);
            }
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
              const VT bq = sq[b][(kb + u) * RP + 16 * ct + i];
              if constexpr (P == PREC_F64)
                acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv, bq, acc[ct], 0, 0, 0);
              else
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(kv, bq, acc[ct], 0, 0, 0);
            }
          }
        }
      }
    };
    // end of a source leaf (FP32/BF16): its partial sum added to the
    // compensated total, the partial restarted
    auto close_source = [&]() {
      if constexpr (P != PREC_F64) {
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            float s_ = sh[ct][v], c_ = sl[ct][v];
            two_sum(s_, c_, acc[ct][v]);
            sh[ct][v] = s_;
            sl[ct][v] = c_;
            acc[ct][v] = 0.0f;
          }
      }
    };
    int src = 0, k0 = 0, b = 0;
    fetch(0, 0);
    commit(0);
    __syncthreads();
    for (;;) {
      const int cnt = __builtin_amdgcn_readfirstlane(d.sc[src]);
      const int nv = min(SB, cnt - k0);
      int ns = src, nk = k0 + SB;
      if (nk >= cnt) {
        ns = src + 1;
        nk = 0;
      }
      const bool more = ns < nsrc;
      if (more) fetch(ns, nk);  // in flight while this stage is evaluated
      compute(b, nv);
      if (!more || ns != src) close_source();
      if (more) commit(b ^ 1);
      __syncthreads();
      if (!more) break;
      src = ns;
      k0 = nk;
      b ^= 1;
    }
    if (act) {
      const auto out = gptr_w(pr.out);
#pragma unroll
      for (int ct = 0; ct < NCT; ct++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          int row, col;
          acc_pos<P>(i, g, v, row, col);
          const int r = 16 * w + row;
          if (r < M) {
            double val;
            if constexpr (P == PREC_F64) val = scale * acc[ct][v];
            else val = (double)sh[ct][v] + (double)sl[ct][v];
            out[(int64_t)(d.out_row + r) * RP + 16 * ct + col] = val;
          }
        }
    }
  }
}

// The pair partials of the near field added to the potentials, one
// workgroup per leaf chunk, slots in order (deterministic).  The first pw
// columns of each row (pw even, rows and slots whole multiples of 16
// doubles): 16-byte accesses, four slots' loads in flight before their sums.
__global__ __launch_bounds__(256) void k_mr_near_red(const MrNearRed* __restrict__ tasks,
                                                     int rp) {
  const MrNearRed d = tasks[blockIdx.x];
  const int hw = d.pw / 2, hr = rp / 2, hs = d.prs / 2;
  const int tot = d.nrows * hw;
  double2* out = reinterpret_cast<double2*>(d.out + d.row0 * rp);
  const double2* part = reinterpret_cast<const double2*>(d.part + d.base);
  const int64_t st = d.stride / 2;
  const int ns = d.nslots;
  for (int e = threadIdx.x; e < tot; e += 256) {
    const int r = e / hw, c = e - r * hw;
    const int o = r * hr + c, po = (d.lr0 + r) * hs + c;
    double2 s = out[o];
    int sl = 0;
    for (; sl + 4 <= ns; sl += 4) {
      const double2 a0 = part[po + (sl + 0) * st], a1 = part[po + (sl + 1) * st];
      const double2 a2 = part[po + (sl + 2) * st], a3 = part[po + (sl + 3) * st];
      s.x += a0.x;
      s.y += a0.y;
      s.x += a1.x;
      s.y += a1.y;
      s.x += a2.x;
      s.y += a2.y;
      s.x += a3.x;
      s.y += a3.y;
    }
    for (; sl < ns; sl++) {
      const double2 a = part[po + sl * st];
      s.x += a.x;
      s.y += a.y;
    }
    out[o] = s;
  }
}

__global__ void k_mr_zero(int* __restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0;
}

// ---------------------------------------------------------------------------
// Gather and output go through LDS, half a workgroup's kMrIoPoints caller
// points at a time: b and y are read and written along their columns (a
// wave's access is 64 consecutive caller points of one column), qs, pot and
// near along their RP-column rows (one point's row whole), so neither side
// touches a cache line for a few bytes of it.  Workgroup blk of a problem
// takes its caller points [blk kMrIoPoints, (blk + 1) kMrIoPoints) of d.cnt.
constexpr int kMrIoHalf = kMrIoPoints / 2;
static_assert(kMrIoHalf == 128, "the column phase maps 256 threads to 128 points x 2 columns");

// Gather: qs[inv_perm[i]][c] = b[i + c ldb] (0 for c >= nrhs).
template <class VT, int NCT>
__device__ __forceinline__ void mr_gather_block(const MrIoTask& d, int blk) {
  constexpr int RP = 16 * NCT, PPI = 256 / RP, LD = RP + 1;
  __shared__ VT tile[kMrIoHalf * LD];
  __shared__ int sperm[kMrIoHalf];
  const int t = threadIdx.x, c = t % RP, pl = t / RP;
  const int p = t % kMrIoHalf, c0 = t / kMrIoHalf;
  VT* qs = static_cast<VT*>(d.qs);
  for (int h = 0; h < kMrIoPoints; h += kMrIoHalf) {
    const int i0 = blk * kMrIoPoints + h;
    const int cnt = min(kMrIoHalf, d.cnt - i0);
    if (cnt <= 0) break;
    // columns in: point p of columns c0, c0 + 2, ...
    double v[RP / 2];
#pragma unroll
    for (int u = 0; u < RP / 2; u++) {
      const int cc = c0 + 2 * u;
      v[u] = p < cnt && cc < d.nrhs ? d.b[(int64_t)(i0 + p) + (int64_t)cc * d.ldb] : 0.0;
    }
    if (t < cnt) sperm[t] = d.inv_perm[i0 + t];
#pragma unroll
    for (int u = 0; u < RP / 2; u++) tile[p * LD + c0 + 2 * u] = (VT)v[u];
    __syncthreads();
    // rows out
    for (int q = pl; q < cnt; q += PPI) qs[(int64_t)sperm[q] * RP + c] = tile[q * LD + c];
    __syncthreads();
  }
}

// Output: y[i + c ldy] = pot[inv_perm[i]][c] + near[inv_perm[i]][c] in FP64.
template <class VT, int NCT>
__device__ __forceinline__ void mr_out_block(const MrIoTask& d, int blk) {
  constexpr int RP = 16 * NCT, PPI = 256 / RP, LD = RP + 1, NQ = kMrIoHalf / PPI;
  __shared__ double tile[kMrIoHalf * LD];
  __shared__ int sperm[kMrIoHalf];
  const int t = threadIdx.x, c = t % RP, pl = t / RP;
  const int p = t % kMrIoHalf, c0 = t / kMrIoHalf;
  const VT* pot = static_cast<const VT*>(d.pot);
  for (int h = 0; h < kMrIoPoints; h += kMrIoHalf) {
    const int i0 = blk * kMrIoPoints + h;
    const int cnt = min(kMrIoHalf, d.cnt - i0);
    if (cnt <= 0) break;
    if (t < cnt) sperm[t] = d.inv_perm[i0 + t];
    __syncthreads();
    // rows in: every load of the thread issued before the first sum
    VT a[NQ];
    double b[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const int q = min(pl + k * PPI, cnt - 1);
      const int64_t si = sperm[q];
      a[k] = pot[si * RP + c];
      b[k] = d.nearv[si * RP + c];
    }
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const int q = pl + k * PPI;
      if (q < cnt) tile[q * LD + c] = (double)a[k] + b[k];
    }
    __syncthreads();
    // columns out: point p of columns c0, c0 + 2, ...
#pragma unroll
    for (int u = 0; u < RP / 2; u++) {
      const int cc = c0 + 2 * u;
      if (p < cnt && cc < d.nrhs) d.y[(int64_t)(i0 + p) + (int64_t)cc * d.ldy] = tile[p * LD + cc];
    }
    __syncthreads();
  }
}

template <class VT, int NCT>
__global__ __launch_bounds__(256) void k_mr_gather(MrIoTask d) {
  mr_gather_block<VT, NCT>(d, blockIdx.x);
}

template <class VT, int NCT>
__global__ __launch_bounds__(256) void k_mr_out(MrIoTask d) {
  mr_out_block<VT, NCT>(d, blockIdx.x);
}

// Several problems in one launch: problem t owns workgroups
// [tasks[t].i0, tasks[t + 1].i0); workgroup w finds its problem by binary
// search (the last t with tasks[t].i0 <= w).
template <class VT, int NCT, bool GATHER>
__global__ __launch_bounds__(256) void k_mr_io_multi(const MrIoTask* __restrict__ tasks,
                                                     int ntasks) {
  const int w = blockIdx.x;
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].i0 <= w) lo = mid;
    else hi = mid - 1;
  }
  const MrIoTask d = tasks[lo];
  if constexpr (GATHER) mr_gather_block<VT, NCT>(d, w - d.i0);
  else mr_out_block<VT, NCT>(d, w - d.i0);
}

__global__ void k_add_cols(double* __restrict__ y, int64_t ldy, const double* __restrict__ t,
                           int64_t ldt, int n, int nrhs) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * nrhs) return;
  const int64_t c = e / n, i = e - c * n;
  y[i + c * ldy] += t[i + c * ldt];
}

// ---------------------------------------------------------------------------
// Factor construction.
struct MrBoxV {
  int64_t l2poff, rmbase, voff, vtoff;
  int n, r, vld, vtld;
};

// FP64 basis, row-major: V = the box's L2P with its skeleton rows the exact
// identity (row rowmap[nc + q] is e_q, stored row t is row rowmap[t]).
__global__ __launch_bounds__(256) void k_mr_build_v64(const MrBoxV* __restrict__ boxes,
                                                      const double* __restrict__ l2p_rm,
                                                      const int* __restrict__ rowmap,
                                                      double* __restrict__ V) {
  __shared__ int inv_row[4096];  // rows: leaf points (<= 512) or 4 children's ranks (<= 4 x 1024)
  const MrBoxV d = boxes[blockIdx.x];
  const int n = d.n, r = d.r, nc = n - r;
  for (int t = threadIdx.x; t < n; t += blockDim.x) inv_row[rowmap[d.rmbase + t]] = t;
  __syncthreads();
  for (int e = threadIdx.x; e < n * d.vld; e += blockDim.x) {
    const int row = e / d.vld, c = e - row * d.vld;
    double v = 0.0;
    if (c < r) {
      const int t = inv_row[row];
      v = t < nc ? l2p_rm[d.l2poff + (int64_t)t * r + c] : (t - nc == c ? 1.0 : 0.0);
    }
    V[d.voff + e] = v;
  }
}

// A row-major FP64 matrix (M x K, leading dimension ld; trans: the matrix is
// the transpose of the one stored) into the tiled layout of precision P.
struct MrBoxRT {
  int64_t src, dst;
  int M, K, ld, trans;
};
template <int P>
__global__ __launch_bounds__(256) void k_mr_retile(const MrBoxRT* __restrict__ boxes,
                                                   const double* __restrict__ src,
                                                   typename MrT<P>::A* __restrict__ dst) {
  using AT = typename MrT<P>::A;
  const MrBoxRT d = boxes[blockIdx.x];
  const int kg = P == PREC_BF16 ? 8 : 4;
  const int kp = (d.K + kg - 1) / kg * kg;
  const int64_t tot = (int64_t)d.M * kp;
  for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
    const int row = (int)(e / kp), k = (int)(e - (int64_t)row * kp);
    double v = 0.0;
    if (k < d.K) v = d.trans ? src[d.src + (int64_t)k * d.ld + row] : src[d.src + (int64_t)row * d.ld + k];
    AT x;
    if constexpr (P == PREC_F64) x = v;
    else if constexpr (P == PREC_F32) x = (float)v;
    else x = __builtin_bit_cast(uint16_t, (__bf16)(float)v);
    dst[d.dst + tile_off<P>(row, k, d.M, kp)] = x;
  }
}

struct MrBoxM {
  int64_t soff, moff, coloff;
  int r, K, mld, pad;
};

// FP64 M2L row block: M[i][c] = K(skeleton_k[i], point colidx[c]), the same
// formula as the single-RHS pair blocks, written in the FP64 tiled layout
// (columns up to K rounded to 4, the padding zero).
template <int KID>
__global__ __launch_bounds__(256) void k_mr_build_m64(const MrBoxM* __restrict__ boxes,
                                                      const double* __restrict__ skx,
                                                      const double* __restrict__ sky,
                                                      const int* __restrict__ colidx,
                                                      double* __restrict__ Mout, KParams kp) {
  const MrBoxM d = boxes[blockIdx.x];
  const int kp4 = (d.K + 3) / 4 * 4;
  const int64_t tot = (int64_t)d.r * kp4;
  for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
    const int i = (int)(e / kp4), c = (int)(e - (int64_t)i * kp4);
    double v = 0.0;
    if (c < d.K) {
      const int q = colidx[d.coloff + c];
      v = kernel_xy<KID>(skx[d.soff + i], sky[d.soff + i], skx[q], sky[q], kp);
    }
    Mout[d.moff + tile_off<PREC_F64>(i, c, d.r, kp4)] = v;
  }
}

struct MrBoxQR {
  int64_t voff, roff;
  int64_t croff[4];  // children's R offsets
  int n, r, vld, nchild;
  int crank[4];      // children's ranks (their row blocks, in order)
};

// Orthonormal basis of one box: E = V (leaves) or diag(R_children) V, then
// E = Q R by classical Gram-Schmidt with one re-orthogonalisation (FP64).
// Q overwrites Qout (n x r, leading dimension vld), R goes to Rout (r x r).
// Dynamic LDS: vs[nmax], hs[4][rmax], h[rmax] doubles.
__global__ __launch_bounds__(256) void k_mr_qr(const MrBoxQR* __restrict__ boxes,
                                               const double* __restrict__ V,
                                               double* __restrict__ Qout,
                                               double* __restrict__ Rbuf, int nmax, int rmax) {
  extern __shared__ double qr_lds[];
  double* vs = qr_lds;
  double* hs0 = qr_lds + nmax;
  double* h = hs0 + 4 * rmax;
  __shared__ double red[4];
  const MrBoxQR d = boxes[blockIdx.x];
  const int n = d.n, r = d.r, ld = d.vld;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const double* Vb = V + d.voff;
  double* Q = Qout + d.voff;
  double* R = Rbuf + d.roff;
  for (int e = tid; e < r * r; e += 256) R[e] = 0.0;
  // row offsets of the children's blocks
  int cro[5];
  cro[0] = 0;
  for (int u = 0; u < 4; u++) cro[u + 1] = cro[u] + (u < d.nchild ? d.crank[u] : 0);
  __syncthreads();
  for (int c = 0; c < r; c++) {
    // column c of E
    for (int row = tid; row < n; row += 256) {
      double v;
      if (d.nchild == 0) {
        v = Vb[(int64_t)row * ld + c];
      } else {
        int u = 0;
        while (u < 3 && row >= cro[u + 1]) u++;
        const int lr = row - cro[u], rc = d.crank[u];
        const double* Rc = Rbuf + d.croff[u];
        v = 0.0;
        for (int p = lr; p < rc; p++) v = fma(Rc[(int64_t)lr * rc + p], Vb[(int64_t)(cro[u] + p) * ld + c], v);
      }
      vs[row] = v;
    }
    __syncthreads();
    for (int pass = 0; pass < 2 && c > 0; pass++) {
      // h = Q[:, :c]^T v: wave w sums rows w, w + 4, ..; lane p columns p, p + 64, ..
      for (int p = lane; p < c; p += 64) {
        double s = 0.0;
        for (int row = w; row < n; row += 4) s = fma(Q[(int64_t)row * ld + p], vs[row], s);
        hs0[w * rmax + p] = s;
      }
      __syncthreads();
      for (int p = tid; p < c; p += 256) {
        const double hp = (hs0[p] + hs0[rmax + p]) + (hs0[2 * rmax + p] + hs0[3 * rmax + p]);
        h[p] = hp;
        R[(int64_t)p * r + c] += hp;
      }
      __syncthreads();
      for (int row = tid; row < n; row += 256) {
        double v = vs[row];
        for (int p = 0; p < c; p++) v = fma(-Q[(int64_t)row * ld + p], h[p], v);
        vs[row] = v;
      }
      __syncthreads();
    }
    double s = 0.0;
    for (int row = tid; row < n; row += 256) s = fma(vs[row], vs[row], s);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) red[w] = s;
    __syncthreads();
    const double nrm = sqrt((red[0] + red[1]) + (red[2] + red[3]));
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    for (int row = tid; row < n; row += 256) Q[(int64_t)row * ld + c] = vs[row] * inv;
    if (tid == 0) R[(int64_t)c * r + c] = nrm;
    __syncthreads();
  }
  // padding columns of Q stay zero
  for (int e = tid; e < n * (ld - r); e += 256) {
    const int row = e / (ld - r), c = r + e % (ld - r);
    Q[(int64_t)row * ld + c] = 0.0;
  }
}

constexpr int kMaxILBlocks = 27;
struct MrBoxMP {
  int64_t moff, roff;
  int r, K, mld, nblk;
  int boff[kMaxILBlocks], brank[kMaxILBlocks];
  int64_t broff[kMaxILBlocks];
};

// M2L' = R_k M R_b^T block by block (FP64), written in the tiled layout of
// precision P (r x K matrix, columns up to kp zero).
template <int P>
__global__ __launch_bounds__(256) void k_mr_mprime(const MrBoxMP* __restrict__ boxes,
                                                   const double* __restrict__ M64,
                                                   const double* __restrict__ Rbuf,
                                                   typename MrT<P>::A* __restrict__ Mout) {
  using AT = typename MrT<P>::A;
  __shared__ double T[4096];
  const MrBoxMP& d = boxes[blockIdx.x];
  const int r = d.r;
  const int kg = P == PREC_BF16 ? 8 : 4;
  const int kp = (d.K + kg - 1) / kg * kg;
  const double* Rk = Rbuf + d.roff;
  const double* Mk = M64 + d.moff;  // FP64 tiled layout (columns to K rounded to 4)
  const int kp4 = (d.K + 3) / 4 * 4;
  AT* Ok = Mout + d.moff;
  auto put = [&](int i, int c, double v) {
    AT x;
    if constexpr (P == PREC_F64) x = v;
    else if constexpr (P == PREC_F32) x = (float)v;
    else x = __builtin_bit_cast(uint16_t, (__bf16)(float)v);
    Ok[tile_off<P>(i, c, r, kp)] = x;
  };
  // padding columns
  for (int e = threadIdx.x; e < r * (kp - d.K); e += blockDim.x) {
    const int i = e / (kp - d.K), c = d.K + e % (kp - d.K);
    put(i, c, 0.0);
  }
  for (int b = 0; b < d.nblk; b++) {
    const int rb = d.brank[b], off = d.boff[b];
    const double* Rb = Rbuf + d.broff[b];
    const int RI = max(1, min(r, 4096 / max(rb, 1)));  // rows of T per chunk
    for (int i0 = 0; i0 < r; i0 += RI) {
      const int ni = min(RI, r - i0);
      __syncthreads();
      // T[ii][w] = sum_{p >= i} R_k[i][p] M[p][off + w]
      for (int e = threadIdx.x; e < ni * rb; e += blockDim.x) {
        const int ii = e / rb, wv = e - ii * rb, ix = i0 + ii;
        double s = 0.0;
        for (int p = ix; p < r; p++)
          s = fma(Rk[(int64_t)ix * r + p], Mk[tile_off<PREC_F64>(p, off + wv, r, kp4)], s);
        T[ii * rb + wv] = s;
      }
      __syncthreads();
      // out[i][off + u] = sum_{w >= u} T[i][w] R_b[u][w]
      for (int e = threadIdx.x; e < ni * rb; e += blockDim.x) {
        const int ii = e / rb, u = e - ii * rb;
        double s = 0.0;
        for (int wv = u; wv < rb; wv++) s = fma(T[ii * rb + wv], Rb[(int64_t)u * rb + wv], s);
        put(i0 + ii, off + u, s);
      }
    }
  }
}

// Pair-stored M2L blocks (FP64, FP32): element (row, col) of a block with nb
// micro columns (MrMicro).
__host__ __device__ inline int64_t micro_off(int row, int col, int nb) {
  return ((int64_t)(row >> 2) * nb + (col >> 2)) * 16 + (row & 3) * 4 + (col & 3);
}

struct MrPairB {
  int64_t poff, s1, s2, r1off, r2off;
  int r1, r2;
};

// FP64 block K(skeleton_first[row], skeleton_second[col]), the single-RHS
// blocks' formula, zero in the padding to whole micro-tiles.
template <int KID>
__global__ __launch_bounds__(256) void k_mr_build_pairs(const MrPairB* __restrict__ pairs,
                                                        const double* __restrict__ skx,
                                                        const double* __restrict__ sky,
                                                        double* __restrict__ out, KParams kp) {
  const MrPairB d = pairs[blockIdx.x];
  const int R1 = (d.r1 + 3) & ~3, R2 = (d.r2 + 3) & ~3, nb = R2 / 4;
  const int64_t tot = (int64_t)R1 * R2;
  for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
    const int m = (int)(e >> 4), w = (int)(e & 15);
    const int a = m / nb, bb = m - a * nb;
    const int row = 4 * a + (w >> 2), col = 4 * bb + (w & 3);
    double v = 0.0;
    if (row < d.r1 && col < d.r2)
      v = kernel_xy<KID>(skx[d.s1 + row], sky[d.s1 + row], skx[d.s2 + col], sky[d.s2 + col], kp);
    out[d.poff + e] = v;
  }
}

// 48-bit block floating point (b48_split) of the pair-stored FP64 blocks for
// the blocked passes: per pair its block F (the first box's rows) and F's
// transpose times the kernel's transpose sign (the second box's rows), each
// as micro tiles in the FP64 layout's order at tile offsets 2 poff / 16 and
// after F; a tile holds its 16 entries' low 32 bits, then their top 16 bits
// (96 bytes).  One exponent per pair, from the block's largest magnitude.
template <int KID>
__global__ __launch_bounds__(256) void k_mr_b48_pairs(const MrPairB* __restrict__ pairs,
                                                      const double* __restrict__ M64,
                                                      char* __restrict__ out,
                                                      int* __restrict__ sexp) {
  __shared__ double wmax[4];
  const MrPairB d = pairs[blockIdx.x];
  const int R1 = (d.r1 + 3) & ~3, R2 = (d.r2 + 3) & ~3, nb = R2 / 4, nbt = R1 / 4;
  const int64_t tot = (int64_t)R1 * R2;
  const double* F = M64 + d.poff;
  double m = 0.0;
  for (int64_t e = threadIdx.x; e < tot; e += 256) m = fmax(m, fabs(F[e]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
  const int sx = b48_exponent(m);
  char* Fo = out + (d.poff / 16) * 2 * 96;
  char* To = Fo + (tot / 16) * 96;
  for (int64_t e = threadIdx.x; e < tot; e += 256) {
    const int t = (int)(e >> 4), w = (int)(e & 15);
    const int a = t / nb, bb = t - a * nb;
    uint32_t lo;
    uint16_t hi;
    b48_split(F[e], sx, lo, hi);
    reinterpret_cast<uint32_t*>(Fo + (int64_t)t * 96)[w] = lo;
    reinterpret_cast<uint16_t*>(Fo + (int64_t)t * 96 + 64)[w] = hi;
    // entry (4 a + w / 4, 4 bb + w % 4) of F is entry (4 bb + w % 4, 4 a + w / 4) of T
    const int tt = bb * nbt + a, wt = (w & 3) * 4 + (w >> 2);
    b48_split(kTransposeSign<KID> * F[e], sx, lo, hi);
    reinterpret_cast<uint32_t*>(To + (int64_t)tt * 96)[wt] = lo;
    reinterpret_cast<uint16_t*>(To + (int64_t)tt * 96 + 64)[wt] = hi;
  }
  if (threadIdx.x == 0) sexp[blockIdx.x] = sx;
}

// The orthonormal basis' block R_first P R_second^T (FP64 arithmetic) of a
// pair, written in precision P at the block's place in the same layout.
template <int P>
__global__ __launch_bounds__(256) void k_mr_pair_prime(const MrPairB* __restrict__ pairs,
                                                       const double* __restrict__ P64,
                                                       const double* __restrict__ Rbuf,
                                                       typename MrT<P>::A* __restrict__ out) {
  using AT = typename MrT<P>::A;
  __shared__ double T[4096];
  const MrPairB d = pairs[blockIdx.x];
  const int r1 = d.r1, r2 = d.r2;
  const int R1 = (r1 + 3) & ~3, R2 = (r2 + 3) & ~3, nb = R2 / 4;
  const double* R1m = Rbuf + d.r1off;
  const double* R2m = Rbuf + d.r2off;
  const double* Pm = P64 + d.poff;
  AT* O = out + d.poff;
  auto put = [&](int64_t e, double v) {
    if constexpr (P == PREC_F64) O[e] = v;
    else if constexpr (P == PREC_F32) O[e] = (float)v;
    else O[e] = __builtin_bit_cast(uint16_t, (__bf16)(float)v);
  };
  // the padding
  for (int64_t e = threadIdx.x; e < (int64_t)R1 * R2; e += blockDim.x) {
    const int m = (int)(e >> 4), w = (int)(e & 15);
    const int a = m / nb, bb = m - a * nb;
    if (4 * a + (w >> 2) >= r1 || 4 * bb + (w & 3) >= r2) put(e, 0.0);
  }
  const int RI = max(1, min(r1, 4096 / max(r2, 1)));  // rows of T per chunk
  for (int i0 = 0; i0 < r1; i0 += RI) {
    const int ni = min(RI, r1 - i0);
    __syncthreads();
    // T[ii][w] = sum_{p >= i} R_first[i][p] P[p][w]
    for (int e = threadIdx.x; e < ni * r2; e += blockDim.x) {
      const int ii = e / r2, wv = e - ii * r2, ix = i0 + ii;
      double s = 0.0;
      for (int p = ix; p < r1; p++) s = fma(R1m[(int64_t)ix * r1 + p], Pm[micro_off(p, wv, nb)], s);
      T[ii * r2 + wv] = s;
    }
    __syncthreads();
    // out[i][u] = sum_{w >= u} T[i][w] R_second[u][w]
    for (int e = threadIdx.x; e < ni * r2; e += blockDim.x) {
      const int ii = e / r2, u = e - ii * r2;
      double s = 0.0;
      for (int wv = u; wv < r2; wv++) s = fma(T[ii * r2 + wv], R2m[(int64_t)u * r2 + wv], s);
      put(micro_off(i0 + ii, u, nb), s);
    }
  }
}

template <class T>
void upload_vec(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
  if (!h.empty()) {
    AFMM_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    AFMM_HIP(hipStreamSynchronize(s));
  }
}

inline int round8(int v) { return (v + 7) & ~7; }

int64_t level_rbase(const Engine& e, int j, int k) {
  // where box (j, k)'s rowmap starts: a leaf's at leaf_start, an upper box's
  // at N + the skeleton offset of its first child
  return j == e.L ? e.leaf_start[k] : e.n + e.lv[j + 1].soff[4 * (size_t)k];
}

int slot_box(const Engine& e, int j, int k, int q) {
  return q < 16 ? e.inner[j][16 * (size_t)k + q] : e.outer[j][24 * (size_t)k + q - 16];
}

// Layout shared by all precisions: the bases' offsets and leading
// dimensions, and the pair-stored M2L of FP64 (blocks, per-box micro
// descriptors).
void mr_layout(Engine& e) {
  if (e.mr_layout) return;
  const int L = e.L;
  const int64_t nbox = e.lv[L].gbase + e.lv[L].nb;
  e.mr_voff.assign(nbox + 1, 0);
  e.mr_vtoff.assign(nbox + 1, 0);
  e.mr_vld.assign(nbox, 8);
  e.mr_vtld.assign(nbox, 8);
  int64_t vo = 0, vto = 0;
  e.mr_v_coeffs = 0;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k], n = li.nin[k];
      e.mr_vld[g] = round8(std::max(r, 1));
      e.mr_vtld[g] = round8(std::max(n, 1));
      e.mr_voff[g] = vo;
      e.mr_vtoff[g] = vto;
      vo += (int64_t)n * e.mr_vld[g];
      vto += (int64_t)r * e.mr_vtld[g];
      e.mr_v_coeffs += (int64_t)n * r;
    }
  }
  e.mr_voff[nbox] = vo;
  e.mr_vtoff[nbox] = vto;
  e.mr_v_total = vo;
  e.mr_vt_total = vto;
  // pairs: (k, b) with k < b in the interaction list, first = k; blocks in
  // level, box, slot order
  e.mr_pairs.clear();
  int64_t po = 0;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int r1 = li.rank[k];
      if (r1 <= 0) continue;
      for (int q = 0; q < 40; q++) {
        const int b = slot_box(e, j, k, q);
        if (b <= k) continue;
        const int r2 = li.rank[b];
        if (r2 <= 0) continue;
        Engine::PairBlk pb;
        pb.poff = po;
        pb.s1 = li.soff[k];
        pb.s2 = li.soff[b];
        pb.g1 = li.gbase + k;
        pb.g2 = li.gbase + b;
        pb.r1 = r1;
        pb.r2 = r2;
        e.mr_pairs.push_back(pb);
        po += (int64_t)((r1 + 3) & ~3) * ((r2 + 3) & ~3);
      }
    }
  }
  e.mr_pm_total = po;
  // the box-major layout (FP32, BF16) holds every block twice, unpadded
  e.mr_m_coeffs = 0;
  for (const Engine::PairBlk& pb : e.mr_pairs) e.mr_m_coeffs += 2 * (int64_t)pb.r1 * pb.r2;
  // per box: its interaction slots in order, a micro column per 4 partner rows
  std::vector<MrMicro> mic;
  e.mr_uoff.assign(nbox + 1, 0);
  size_t pi = 0;  // pairs were made in the same (level, box, slot) order
  std::vector<std::vector<int64_t>> first_off(L + 1);
  for (int j = 2; j <= L; j++) first_off[j].assign((size_t)e.lv[j].nb * 40, -1);
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      if (li.rank[k] <= 0) continue;
      for (int q = 0; q < 40; q++) {
        const int b = slot_box(e, j, k, q);
        if (b <= k || li.rank[b] <= 0) continue;
        first_off[j][(size_t)k * 40 + q] = e.mr_pairs[pi++].poff;
      }
    }
  }
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      e.mr_uoff[g] = (int64_t)mic.size();
      const int r = li.rank[k];
      if (r <= 0) continue;
      const int Rk = (r + 3) & ~3;
      for (int q = 0; q < 40; q++) {
        const int b = slot_box(e, j, k, q);
        if (b < 0) continue;
        const int rb = li.rank[b];
        if (rb <= 0) continue;
        const int Rb = (rb + 3) & ~3;
        if (b > k) {
          // k first: micro columns of its block, forward
          const int64_t poff = first_off[j][(size_t)k * 40 + q];
          for (int bb = 0; bb < Rb / 4; bb++) {
            MrMicro m;
            m.a = poff + 16 * (int64_t)bb;
            m.brow = (int)(li.soff[b] + 4 * bb);
            const int nval = std::min(4, rb - 4 * bb);
            m.info = (Rb / 4 * 16) | ((Rk / 4 - 1) << 16) | (nval << 25) | (1 << 28);
            mic.push_back(m);
          }
        } else {
          // k second: b's block read transposed, a micro column per micro
          // row of the block; find the slot of k in b's list
          int64_t poff = -1;
          for (int q2 = 0; q2 < 40 && poff < 0; q2++)
            if (slot_box(e, j, b, q2) == k) poff = first_off[j][(size_t)b * 40 + q2];
          if (poff < 0) throw Error(6, "interaction lists are not symmetric");
          const int nbk = Rk / 4;  // micro columns of b's block (k's rows)
          for (int aa = 0; aa < Rb / 4; aa++) {
            MrMicro m;
            m.a = poff + (int64_t)aa * nbk * 16;
            m.brow = (int)(li.soff[b] + 4 * aa);
            const int nval = std::min(4, rb - 4 * aa);
            // an antisymmetric kernel's transposed block enters with sign -1
            m.info = 16 | ((nbk - 1) << 16) | (nval << 25) | (e.kid == K_USER_ANTI ? 1 << 29 : 0);
            mic.push_back(m);
          }
        }
      }
    }
  }
  e.mr_uoff[nbox] = (int64_t)mic.size();
  for (const Engine::PairBlk& pb : e.mr_pairs)
    if ((pb.r1 + 3) / 4 > 512 || (pb.r2 + 3) / 4 > 512 || ((pb.r2 + 3) & ~3) * 4 > 0xFFFF)
      throw Error(6, "rank above the pair-stored M2L's range");
  if (e.skel_total > INT32_MAX) throw Error(6, "skeleton count exceeds 2^31");
  upload_vec(e.mr_micro, mic, e.s_aux);
  e.mr_layout = true;
}

// The box-major M2L row blocks of FP32 and BF16: offsets, leading dimensions,
// the column index of every row block (each block stored for both boxes).
void mr_layout_boxmajor(Engine& e) {
  if (e.mr_boxmajor_layout) return;
  mr_layout(e);
  const int L = e.L;
  const int64_t nbox = e.lv[L].gbase + e.lv[L].nb;
  e.mr_moff.assign(nbox + 1, 0);
  e.mr_coloff.assign(nbox + 1, 0);
  e.mr_mld.assign(nbox, 8);
  e.mr_K.assign(nbox, 0);
  std::vector<int> colidx;
  int64_t mo = 0;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k];
      int K = 0;
      e.mr_coloff[g] = (int64_t)colidx.size();
      for (int q = 0; q < 40; q++) {
        const int b = slot_box(e, j, k, q);
        if (b < 0) continue;
        const int rb = li.rank[b];
        for (int u = 0; u < rb; u++) colidx.push_back((int)(li.soff[b] + u));
        K += rb;
      }
      e.mr_K[g] = K;
      e.mr_mld[g] = round8(std::max(K, 1));
      e.mr_moff[g] = mo;
      mo += (int64_t)r * e.mr_mld[g];
    }
  }
  e.mr_moff[nbox] = mo;
  e.mr_coloff[nbox] = (int64_t)colidx.size();
  e.mr_m_total = mo;
  if ((int64_t)colidx.size() > INT32_MAX) throw Error(6, "M2L column index exceeds 2^31 entries");
  upload_vec(e.mr_colidx, colidx, e.s_aux);
  e.mr_boxmajor_layout = true;
}

// The row-major FP64 basis V (nin x rank, leading dimension vld) of every
// box, the source of the tiled bases.
void build_vrm64(Engine& e, DevBuf& Vrm) {
  mr_layout(e);
  const int L = e.L;
  hipStream_t s = e.s_aux;
  Vrm.alloc(sizeof(double) * std::max<int64_t>(e.mr_v_total, 8));
  AFMM_HIP(hipMemsetAsync(Vrm.p, 0, Vrm.bytes, s));
  std::vector<MrBoxV> bv;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k], n = li.nin[k];
      if (r > 0 && n > 0)
        bv.push_back(MrBoxV{li.l2poff[k], level_rbase(e, j, k), e.mr_voff[g], e.mr_vtoff[g], n, r,
                            e.mr_vld[g], e.mr_vtld[g]});
    }
  }
  DevBuf dbv;
  upload_vec(dbv, bv, s);
  if (!bv.empty()) {
    k_mr_build_v64<<<(unsigned)bv.size(), 256, 0, s>>>(dbv.as<MrBoxV>(), e.l2p_rm.as<double>(),
                                                       e.rowmap.as<int>(), Vrm.as<double>());
    AFMM_LAUNCH_CHECK();
  }
  AFMM_HIP(hipStreamSynchronize(s));
}

// The tiled V and Vt of precision prec from a row-major FP64 basis (Vrm's
// layout), into f.
void retile_basis(Engine& e, int prec, const DevBuf& Vsrc, MrFactors& f) {
  const int L = e.L;
  hipStream_t s = e.s_aux;
  std::vector<MrBoxRT> bv, bt;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k], n = li.nin[k];
      if (r <= 0 || n <= 0) continue;
      bv.push_back(MrBoxRT{e.mr_voff[g], e.mr_voff[g], n, r, e.mr_vld[g], 0});
      bt.push_back(MrBoxRT{e.mr_voff[g], e.mr_vtoff[g], r, n, e.mr_vld[g], 1});
    }
  }
  if (bv.empty()) return;
  DevBuf dbv, dbt;
  upload_vec(dbv, bv, s);
  upload_vec(dbt, bt, s);
  const unsigned nb = (unsigned)bv.size();
#define AFMM_MR_RETILE(P)                                                                       \
  k_mr_retile<P><<<nb, 256, 0, s>>>(dbv.as<MrBoxRT>(), Vsrc.as<double>(),                      \
                                    f.V.as<typename MrT<P>::A>());                              \
  AFMM_LAUNCH_CHECK();                                                                          \
  k_mr_retile<P><<<nb, 256, 0, s>>>(dbt.as<MrBoxRT>(), Vsrc.as<double>(),                      \
                                    f.Vt.as<typename MrT<P>::A>());                             \
  AFMM_LAUNCH_CHECK();
  switch (prec) {
    case PREC_F64: AFMM_MR_RETILE(PREC_F64) break;
    case PREC_F32: AFMM_MR_RETILE(PREC_F32) break;
    default: AFMM_MR_RETILE(PREC_BF16) break;
  }
#undef AFMM_MR_RETILE
  AFMM_HIP(hipStreamSynchronize(s));
}

void alloc_factors(Engine& e, int prec, MrFactors& f) {
  const size_t fb = factor_bytes(prec);
  hipStream_t s = e.s_aux;
  f.V.alloc(fb * std::max<int64_t>(e.mr_v_total, 8));
  f.Vt.alloc(fb * std::max<int64_t>(e.mr_vt_total, 8));
  f.M.alloc(fb * std::max<int64_t>(pair_m2l(prec) ? e.mr_pm_total : e.mr_m_total, 8));
  // boxes of rank 0 keep zero factors (their tasks read none)
  AFMM_HIP(hipMemsetAsync(f.V.p, 0, f.V.bytes, s));
  AFMM_HIP(hipMemsetAsync(f.Vt.p, 0, f.Vt.bytes, s));
  AFMM_HIP(hipMemsetAsync(f.M.p, 0, f.M.bytes, s));
}

std::vector<MrPairB> pair_records(const Engine& e) {
  std::vector<MrPairB> v;
  v.reserve(e.mr_pairs.size());
  for (const Engine::PairBlk& pb : e.mr_pairs) {
    MrPairB d{};
    d.poff = pb.poff;
    d.s1 = pb.s1;
    d.s2 = pb.s2;
    d.r1 = pb.r1;
    d.r2 = pb.r2;
    v.push_back(d);
  }
  return v;
}

void build_f64(Engine& e) {
  MrFactors& f = e.mrf[PREC_F64];
  if (f.built) return;
  {
    DevBuf Vrm;
    build_vrm64(e, Vrm);
    alloc_factors(e, PREC_F64, f);
    retile_basis(e, PREC_F64, Vrm, f);
  }
  hipStream_t s = e.s_aux;
  std::vector<MrPairB> pr = pair_records(e);
  if (!pr.empty()) {
    DevBuf dpr;
    upload_vec(dpr, pr, s);
    AFMM_DISPATCH_KERNEL(e.kid, KID,
                         k_mr_build_pairs<KID><<<(unsigned)pr.size(), 256, 0, s>>>(
                             dpr.as<MrPairB>(), e.skx.as<double>(), e.sky.as<double>(),
                             f.M.as<double>(), e.kp));
    AFMM_LAUNCH_CHECK();
  }
  AFMM_HIP(hipStreamSynchronize(s));
  f.built = true;
}

// The FP64 blocks in 48-bit block floating point for the blocked passes
// (Engine::mr_b48): both orientations of every pair (k_mr_b48_pairs) and a
// descriptor per micro column with the same box order, slot order and
// micro columns as mr_micro, every one forward.
void build_b48(Engine& e) {
  if (e.mr_b48) return;
  build_f64(e);
  hipStream_t s = e.s_aux;
  std::vector<MrPairB> pr = pair_records(e);
  e.mr_m48.alloc(96 * (size_t)std::max<int64_t>(2 * (e.mr_pm_total / 16), 1));
  std::vector<int> sexp(pr.size(), 0);
  if (!pr.empty()) {
    DevBuf dpr, dsx;
    upload_vec(dpr, pr, s);
    dsx.alloc(sizeof(int) * pr.size());
    AFMM_DISPATCH_KERNEL(e.kid, KID,
                         k_mr_b48_pairs<KID><<<(unsigned)pr.size(), 256, 0, s>>>(
                             dpr.as<MrPairB>(), e.mrf[PREC_F64].M.as<double>(),
                             e.mr_m48.as<char>(), dsx.as<int>()));
    AFMM_LAUNCH_CHECK();
    AFMM_HIP(hipMemcpyAsync(sexp.data(), dsx.p, sizeof(int) * pr.size(), hipMemcpyDeviceToHost, s));
    AFMM_HIP(hipStreamSynchronize(s));
  }
  std::unordered_map<int64_t, int> pair_of;  // (first gid, second gid) -> pair
  for (size_t u = 0; u < e.mr_pairs.size(); u++)
    pair_of[((int64_t)e.mr_pairs[u].g1 << 32) | e.mr_pairs[u].g2] = (int)u;
  const int L = e.L;
  std::vector<MrMicro> mic;
  auto tagged = [&](int64_t tile, int pair) {
    return (int64_t)((uint64_t)tile | ((uint64_t)(uint16_t)(int16_t)sexp[pair] << 48));
  };
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int r = li.rank[k];
      if (r <= 0) continue;
      const int Rk = (r + 3) & ~3;
      for (int q = 0; q < 40; q++) {
        const int b = slot_box(e, j, k, q);
        if (b < 0) continue;
        const int rb = li.rank[b];
        if (rb <= 0) continue;
        const int Rb = (rb + 3) & ~3;
        // the pair's first box is the lower one; its FP64 block offset
        const int lo = std::min(k, b), hi = std::max(k, b);
        const int pi = pair_of.at(((li.gbase + lo) << 32) | (li.gbase + hi));
        const int64_t f_off = e.mr_pairs[pi].poff;
        const int64_t t0 = 2 * (f_off / 16);
        const int Rlo = lo == k ? Rk : Rb, Rhi = lo == k ? Rb : Rk;
        if (k == lo) {
          // F: k's rows, b's columns; micro column bb is tile t0 + bb
          for (int bb = 0; bb < Rb / 4; bb++) {
            MrMicro m;
            m.a = tagged(t0 + bb, pi);
            m.brow = (int)(li.soff[b] + 4 * bb);
            const int nval = std::min(4, rb - 4 * bb);
            m.info = (Rb / 4) | ((Rk / 4 - 1) << 16) | (nval << 25) | (1 << 28);
            mic.push_back(m);
          }
        } else {
          // T: k's rows, b's columns (b first); after F's Rlo Rhi / 16 tiles
          const int64_t tt0 = t0 + (int64_t)(Rlo / 4) * (Rhi / 4);
          for (int aa = 0; aa < Rb / 4; aa++) {
            MrMicro m;
            m.a = tagged(tt0 + aa, pi);
            m.brow = (int)(li.soff[b] + 4 * aa);
            const int nval = std::min(4, rb - 4 * aa);
            m.info = (Rb / 4) | ((Rk / 4 - 1) << 16) | (nval << 25) | (1 << 28);
            mic.push_back(m);
          }
        }
      }
    }
  }
  upload_vec(e.mr_micro48, mic, s);
  e.mr_b48 = true;
}

// FP64 box-major M2L row blocks (the FP32 / BF16 transforms' source) into M64.
void build_boxmajor64(Engine& e, DevBuf& M64) {
  mr_layout_boxmajor(e);
  const int L = e.L;
  hipStream_t s = e.s_aux;
  M64.alloc(sizeof(double) * std::max<int64_t>(e.mr_m_total, 8));
  AFMM_HIP(hipMemsetAsync(M64.p, 0, M64.bytes, s));
  std::vector<MrBoxM> bm;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      if (li.rank[k] > 0)
        bm.push_back(MrBoxM{li.soff[k], e.mr_moff[g], e.mr_coloff[g], li.rank[k], e.mr_K[g],
                            e.mr_mld[g], 0});
    }
  }
  if (!bm.empty()) {
    DevBuf dbm;
    upload_vec(dbm, bm, s);
    AFMM_DISPATCH_KERNEL(e.kid, KID,
                         k_mr_build_m64<KID><<<(unsigned)bm.size(), 256, 0, s>>>(
                             dbm.as<MrBoxM>(), e.skx.as<double>(), e.sky.as<double>(),
                             e.mr_colidx.as<int>(), M64.as<double>(), e.kp));
    AFMM_LAUNCH_CHECK();
  }
  AFMM_HIP(hipStreamSynchronize(s));
}

// The orthonormal basis (R of every box) and the reduced-precision factors.
void build_reduced(Engine& e, int prec) {
  MrFactors& f = e.mrf[prec];
  if (f.built) return;
  build_f64(e);
  if (!pair_m2l(prec)) mr_layout_boxmajor(e);
  DevBuf Vrm;
  build_vrm64(e, Vrm);
  const int L = e.L;
  hipStream_t s = e.s_aux;
  const int64_t nbox = e.lv[L].gbase + e.lv[L].nb;
  DevBuf Q64;
  Q64.alloc(sizeof(double) * std::max<int64_t>(e.mr_v_total, 8));
  AFMM_HIP(hipMemsetAsync(Q64.p, 0, Q64.bytes, s));
  if (!e.mr_ortho) {
    e.mr_roff.assign(nbox + 1, 0);
    int64_t ro = 0;
    for (int j = 2; j <= L; j++)
      for (int k = 0; k < e.lv[j].nb; k++) {
        e.mr_roff[e.lv[j].gbase + k] = ro;
        ro += (int64_t)e.lv[j].rank[k] * e.lv[j].rank[k];
      }
    e.mr_roff[nbox] = ro;
    e.mr_R.alloc(sizeof(double) * std::max<int64_t>(ro, 1));
  }
  // bottom-up: a box's E needs its children's R
  for (int j = L; j >= 2; j--) {
    const LevelInfo& li = e.lv[j];
    std::vector<MrBoxQR> bq;
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k], n = li.nin[k];
      if (r <= 0 || n <= 0) continue;
      MrBoxQR d{};
      d.voff = e.mr_voff[g];
      d.roff = e.mr_roff[g];
      d.n = n;
      d.r = r;
      d.vld = e.mr_vld[g];
      d.nchild = j == L ? 0 : 4;
      for (int u = 0; u < 4 && j < L; u++) {
        const int64_t gc = e.lv[j + 1].gbase + 4 * (int64_t)k + u;
        d.croff[u] = e.mr_roff[gc];
        d.crank[u] = e.lv[j + 1].rank[4 * k + u];
      }
      bq.push_back(d);
    }
    if (bq.empty()) continue;
    int nmax = 1, rmax = 1;
    for (const MrBoxQR& d : bq) {
      nmax = std::max(nmax, d.n);
      rmax = std::max(rmax, d.r);
    }
    const size_t shm = sizeof(double) * ((size_t)nmax + 5 * (size_t)rmax);
    if (shm > (size_t)e.max_lds)
      throw Error(6, "a box too large for the orthonormal basis (" + std::to_string(nmax) +
                         " rows, rank " + std::to_string(rmax) + ")");
    DevBuf dbq;
    upload_vec(dbq, bq, s);
    k_mr_qr<<<(unsigned)bq.size(), 256, shm, s>>>(dbq.as<MrBoxQR>(), Vrm.as<double>(),
                                                  Q64.as<double>(), e.mr_R.as<double>(), nmax,
                                                  rmax);
    AFMM_LAUNCH_CHECK();
    AFMM_HIP(hipStreamSynchronize(s));
  }
  e.mr_ortho = true;
  alloc_factors(e, prec, f);
  retile_basis(e, prec, Q64, f);
  if (pair_m2l(prec)) {
    std::vector<MrPairB> pr = pair_records(e);
    for (size_t u = 0; u < pr.size(); u++) {
      pr[u].r1off = e.mr_roff[e.mr_pairs[u].g1];
      pr[u].r2off = e.mr_roff[e.mr_pairs[u].g2];
    }
    if (!pr.empty()) {
      DevBuf dpr;
      upload_vec(dpr, pr, s);
      if (prec == PREC_F32)
        k_mr_pair_prime<PREC_F32><<<(unsigned)pr.size(), 256, 0, s>>>(
            dpr.as<MrPairB>(), e.mrf[PREC_F64].M.as<double>(), e.mr_R.as<double>(),
            f.M.as<float>());
      else
        throw Error(6, "pair-stored M2L in this precision");
      AFMM_LAUNCH_CHECK();
    }
    AFMM_HIP(hipStreamSynchronize(s));
    f.built = true;
    return;
  }
  DevBuf M64;
  build_boxmajor64(e, M64);
  std::vector<MrBoxMP> bm;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k];
      if (r <= 0) continue;
      MrBoxMP d{};
      d.moff = e.mr_moff[g];
      d.roff = e.mr_roff[g];
      d.r = r;
      d.K = e.mr_K[g];
      d.mld = e.mr_mld[g];
      int off = 0;
      for (int q = 0; q < 40; q++) {
        const int b = slot_box(e, j, k, q);
        if (b < 0) continue;
        const int rb = li.rank[b];
        if (rb > 0) {
          if (d.nblk >= kMaxILBlocks) throw Error(6, "interaction list longer than 27");
          d.boff[d.nblk] = off;
          d.brank[d.nblk] = rb;
          d.broff[d.nblk] = e.mr_roff[li.gbase + b];
          d.nblk++;
        }
        off += rb;
      }
      bm.push_back(d);
    }
  }
  if (!bm.empty()) {
    DevBuf dbm;
    upload_vec(dbm, bm, s);
    const unsigned nb = (unsigned)bm.size();
    if (prec == PREC_F32)
      k_mr_mprime<PREC_F32><<<nb, 256, 0, s>>>(dbm.as<MrBoxMP>(), M64.as<double>(),
                                              e.mr_R.as<double>(), f.M.as<float>());
    else
      k_mr_mprime<PREC_BF16><<<nb, 256, 0, s>>>(dbm.as<MrBoxMP>(), M64.as<double>(),
                                               e.mr_R.as<double>(), f.M.as<uint16_t>());
    AFMM_LAUNCH_CHECK();
  }
  AFMM_HIP(hipStreamSynchronize(s));
  f.built = true;
}

// ---------------------------------------------------------------------------
// Plans: the task lists of one (precision, column tiles) pair.

// Rows per task of the M2L and the other passes (AFMM2D_MR_ROWS_M2L,
// AFMM2D_MR_ROWS: 16, 32, 48 or 64), the XCD mapping of the M2L launch
// (AFMM2D_MR_XCD), the near field's waves per SIMD (AFMM2D_MR_NEAR_WAVES)
// and target rows per task (AFMM2D_MR_NEAR_ROWS: 16, 32, 48 or 64), whether
// FP64 near fields evaluate each leaf pair once (AFMM2D_MR_NEAR_PAIRS, on by
// default; 0: from both sides) and the chunk ring depth of the M2L
// (AFMM2D_MR_DEPTH: 3, 4, 6), read once.
struct MrKnobs {
  int rows_m2l = 64, rows = 16, xcd = 0, near_waves = 1, depth = 3, near_rows = 64;
  bool near_pairs = true;
  // fork the near field after the upward pass: 1 always, 0 never, -1
  // (default) in a plan's own matvec when its near field runs the LDS-staged
  // kernel on 16 columns (N = 2^20, FP32: 1.35 against 1.38 ms per pass; 32
  // columns and batches keep the early fork; the results do not depend on
  // the fork point)
  int near_after_up = -1;
  // k_mr_p2p_lds instead of k_mr_p2p: 1 always, 0 never, -1 (default) for
  // FP64 with 32 columns (N = 2^20: 3.49 against 3.90 ms per pass) and FP32
  // (N = 2^19, 16 columns: 0.75 against 0.83 ms; 2^20: 1.39 against 1.40 and,
  // 32 columns, 1.93 against 1.96; N = 2^22 mixture, 32 columns: 9.34 against
  // 9.62).  FP64 with 16 columns keeps the leaf pairs evaluated once (2^20:
  // 2.36 against 2.48 ms; the staged form wins only below 2^17), BF16 the
  // per-wave form (2^20: 1.07 against 1.09 ms)
  int near_lds = -1;
  MrKnobs() {
    auto rd = [](const char* n, int d) {
      const char* v = std::getenv(n);
      return v ? std::atoi(v) : d;
    };
    rows_m2l = std::max(16, std::min(64, rd("AFMM2D_MR_ROWS_M2L", 64))) / 16 * 16;
    rows = std::max(16, std::min(64, rd("AFMM2D_MR_ROWS", 16))) / 16 * 16;
    xcd = rd("AFMM2D_MR_XCD", 0);
    // one near-field wave per SIMD leaves the upward pass room beside it
    // (N = 2^20, 16 RHS: FP64 2.49 against 2.69 ms per pass with 2 waves)
    near_waves = std::max(1, std::min(16, rd("AFMM2D_MR_NEAR_WAVES", 1)));
    depth = rd("AFMM2D_MR_DEPTH", 3);
    near_rows = std::max(16, std::min(64, rd("AFMM2D_MR_NEAR_ROWS", 64))) / 16 * 16;
    near_pairs = rd("AFMM2D_MR_NEAR_PAIRS", 1) != 0;
    near_after_up = rd("AFMM2D_MR_NEAR_AFTER_UP", -1);
    near_lds = rd("AFMM2D_MR_NEAR_LDS", -1);
  }
};
const MrKnobs& knobs() {
  static MrKnobs k;
  return k;
}

// Whether the near field of a plan runs k_mr_p2p_lds (see near_lds).
bool near_uses_lds(int prec, int nct) {
  const int v = knobs().near_lds;
  return v < 0 ? (prec == PREC_F64 && nct == 2) || prec == PREC_F32 : v != 0;
}

}  // namespace

MrTune mr_default_tune() {
  MrTune t;
  t.depth = knobs().depth;
  t.xcd = knobs().xcd;
  t.near_waves = knobs().near_waves;
  t.rows_m2l = knobs().rows_m2l;
  return t;
}

namespace {

void add_row_groups(std::vector<MrTask>& out, const void* A, int kg, const void* B,
                    const int* bidx, int64_t b0, void* C, int64_t c0, int M, int K, int beta,
                    int rows = 0) {
  if (rows <= 0) rows = knobs().rows;
  // the same number of rows in every group of the box (none smaller than
  // needed): ceil(M / ceil(M / rows)) rounded up to 16
  const int ng = (M + rows - 1) / rows;
  const int gr = ng > 0 ? ((M + ng - 1) / ng + 15) / 16 * 16 : rows;
  for (int r0 = 0; r0 < M; r0 += gr) {
    MrTask t;
    t.A = A;
    t.B = B;
    t.bidx = bidx;
    t.C = C;
    t.b0 = b0;
    t.c0 = c0 + r0;
    t.M = std::min(gr, M - r0);
    t.K = K;
    t.kp = (K + kg - 1) / kg * kg;
    t.mbox = M;
    t.rt0 = r0 / 16;
    t.beta = beta;
    t.imode = bidx ? 1 : 0;
    t.pad = 0;
    out.push_back(t);
  }
}

inline int task_k(const MrTask& t) { return t.K; }
inline int task_k(const MrNearTask&) { return 0; }
inline int task_mode(const MrTask& t) { return t.imode; }
inline int task_mode(const MrNearTask&) { return 0; }

// Orders a task list by row tiles and uploads it.
template <class T>
void upload_tasks(MrTaskList& tl, std::vector<T>& tasks, hipStream_t s) {
  std::stable_sort(tasks.begin(), tasks.end(),
                   [](const T& a, const T& b) { return (a.M + 15) / 16 < (b.M + 15) / 16; });
  tl.n = (int)tasks.size();
  for (int c = 0; c < 4; c++) tl.kmax[c] = 0;
  tl.mode = 0;
  for (const T& t : tasks) {
    tl.kmax[(t.M + 15) / 16 - 1] = std::max(tl.kmax[(t.M + 15) / 16 - 1], task_k(t));
    if (task_mode(t) != 0) {
      if (tl.mode != 0 && tl.mode != task_mode(t)) throw Error(6, "mixed task modes in one list");
      tl.mode = task_mode(t);
    }
  }
  for (int c = 0; c < 5; c++) tl.cls[c] = tl.n;
  for (int t = tl.n - 1; t >= 0; t--) tl.cls[(tasks[t].M + 15) / 16 - 1] = t;
  for (int c = 3; c >= 0; c--) tl.cls[c] = std::min(tl.cls[c], tl.cls[c + 1]);
  upload_vec(tl.buf, tasks, s);
}

// M2L tasks of a plan, rows rows per task: in_k = M_k [out_b ...] for every
// box of rank > 0, levels 2..L in order.
void m2l_tasks(const Engine& e, const MrPlan& p, int rows, std::vector<MrTask>& out) {
  const MrFactors& f = e.mrf[p.prec];
  const size_t fb = factor_bytes(p.prec);
  const int kg = col_group(p.prec);
  if (pair_m2l(p.prec)) {
    // pair-stored blocks through the box's micro descriptors: K = 4 per
    // descriptor (48-bit storage: both orientations, mode 5)
    const bool b48 = e.mr_b48 && e.m2l_store == M2L_STORE_B48;
    const MrMicro* mic = (b48 ? e.mr_micro48 : e.mr_micro).as<MrMicro>();
    const void* Ap = b48 ? e.mr_m48.p : f.M.p;
    for (int j = 2; j <= e.L; j++) {
      const LevelInfo& li = e.lv[j];
      for (int k = 0; k < li.nb; k++) {
        const int64_t g = li.gbase + k;
        const int r = li.rank[k];
        if (r <= 0) continue;
        const int64_t u0 = e.mr_uoff[g], nu = e.mr_uoff[g + 1] - u0;
        const size_t first = out.size();
        add_row_groups(out, Ap, kg, p.outv.p, reinterpret_cast<const int*>(mic), u0, p.inv.p,
                       li.soff[k], r, (int)(4 * nu), 0, rows);
        for (size_t t = first; t < out.size(); t++) out[t].imode = b48 ? 5 : 3;
      }
    }
    return;
  }
  for (int j = 2; j <= e.L; j++) {
    const LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k];
      if (r <= 0) continue;
      add_row_groups(out, static_cast<const char*>(f.M.p) + e.mr_moff[g] * fb, kg, p.outv.p,
                     e.mr_colidx.as<int>(), e.mr_coloff[g], p.inv.p, li.soff[k], r, e.mr_K[g], 0,
                     rows);
    }
  }
}

MrPlan& get_plan(Engine& e, int prec, int nct) {
  for (auto& p : e.mr_plans)
    if (p->prec == prec && p->nct == nct) return *p;
  if (prec == PREC_F64) {
    if (e.m2l_store == M2L_STORE_B48) build_b48(e);
    else build_f64(e);
  } else {
    build_reduced(e, prec);
  }
  auto pl = std::make_unique<MrPlan>();
  MrPlan& p = *pl;
  p.prec = prec;
  p.nct = nct;
  const int L = e.L;
  const int RP = 16 * nct;
  const size_t vb = vector_bytes(prec), fb = factor_bytes(prec);
  const int kg = col_group(prec);
  const MrFactors& f = e.mrf[prec];
  hipStream_t s = e.s_aux;
  const int64_t nsk = std::max<int64_t>(e.skel_total, 1);
  p.qs.alloc(vb * RP * (size_t)e.n);
  p.outv.alloc(vb * RP * (size_t)nsk);
  p.inv.alloc(vb * RP * (size_t)nsk);
  p.pot.alloc(vb * RP * (size_t)e.n);
  p.nearv.alloc(sizeof(double) * RP * (size_t)e.n);
  AFMM_HIP(hipMemsetAsync(p.outv.p, 0, p.outv.bytes, s));
  AFMM_HIP(hipMemsetAsync(p.inv.p, 0, p.inv.bytes, s));
  auto fptr = [&](const DevBuf& b, int64_t off) {
    return static_cast<const void*>(static_cast<const char*>(b.p) + off * fb);
  };
  auto vrow = [&](const DevBuf& b) { return b.p; };
  p.up.resize(L + 1);
  p.down.resize(L + 1);
  p.h_up.assign(L + 1, {});
  p.h_down.assign(L + 1, {});
  std::vector<MrTask> m2l, leaf;
  for (int j = 2; j <= L; j++) {
    const LevelInfo& li = e.lv[j];
    std::vector<MrTask> up, down;
    for (int k = 0; k < li.nb; k++) {
      const int64_t g = li.gbase + k;
      const int r = li.rank[k], n = li.nin[k];
      if (r > 0) {
        // M2M: out_k = V^T src
        const void* Bsrc = j == L ? vrow(p.qs) : vrow(p.outv);
        const int64_t b0 = j == L ? e.leaf_start[k] : e.lv[j + 1].soff[4 * (size_t)k];
        if (n > 0)
          add_row_groups(up, fptr(f.Vt, e.mr_vtoff[g]), kg, Bsrc, nullptr, b0, p.outv.p,
                         li.soff[k], r, n, 0);
        else
          add_row_groups(up, fptr(f.Vt, e.mr_vtoff[g]), kg, Bsrc, nullptr, 0, p.outv.p,
                         li.soff[k], r, 0, 0);
      }
      if (j < L) {
        // L2L: children's in += V in_k
        if (r > 0 && n > 0)
          add_row_groups(down, fptr(f.V, e.mr_voff[g]), kg, vrow(p.inv), nullptr, li.soff[k],
                         p.inv.p, e.lv[j + 1].soff[4 * (size_t)k], n, r, 1);
      } else if (n > 0) {
        // leaf: pot = V in_k (zero for a rank-0 leaf)
        add_row_groups(leaf, fptr(f.V, e.mr_voff[g]), kg, vrow(p.inv), nullptr,
                       r > 0 ? li.soff[k] : 0, p.pot.p, e.leaf_start[k], n, r, 0);
      }
    }
    upload_tasks(p.up[j], up, s);
    upload_tasks(p.down[j], down, s);
    p.h_up[j] = std::move(up);
    p.h_down[j] = std::move(down);
  }
  m2l_tasks(e, p, mr_default_tune().rows_m2l, m2l);
  upload_tasks(p.m2l, m2l, s);
  upload_tasks(p.leaf, leaf, s);
  p.h_m2l = std::move(m2l);
  p.h_leaf = std::move(leaf);
  // near field.  FP64 plans: each leaf pair once, by the task of the leaf
  // whose E, NE, N or NW neighbour the other is (neighbour slots 3..6; the
  // opposite slot is (q + 4) mod 8), which also stores the transposed
  // product into a partial slot of the other leaf; every row block of a
  // leaf's W, SW, S and SE neighbours owns one slot of its rows.  Other
  // precisions: every leaf's points against its neighbours N0..N7 and
  // itself.
  std::vector<MrNearTask> near;
  const int nleaf = 1 << (2 * L);
  const int nr = knobs().near_rows;
  const bool pairs = p.prec == PREC_F64 && knobs().near_pairs && !near_uses_lds(p.prec, nct);
  auto nbr_of = [&](int k, int q) { return e.nbr[L][8 * (size_t)k + q]; };
  auto blocks_of = [&](int k) { return (e.leaf_count[k] + nr - 1) / nr; };
  std::unordered_map<int64_t, int64_t> slot_at;  // (leaf, upper slot, row block) -> offset
  auto slot_key = [](int k, int q, int rb) { return ((int64_t)k * 8 + q) * 4096 + rb; };
  std::vector<MrNearRed> red;
  int64_t ptot = 0;
  if (pairs) {
    for (int b = 0; b < nleaf; b++) {
      const int cb = e.leaf_count[b];
      if (cb == 0) continue;
      const int64_t base = ptot;
      int ns = 0;
      for (int q : {7, 0, 1, 2}) {
        const int a = nbr_of(b, q);
        if (a < 0 || e.leaf_count[a] == 0) continue;
        for (int rb = 0; rb < blocks_of(a); rb++) {
          slot_at[slot_key(a, (q + 4) % 8, rb)] = ptot;
          ptot += (int64_t)cb * RP;
          ns++;
        }
      }
      for (int r0 = 0; ns > 0 && r0 < cb; r0 += 64)
        red.push_back(MrNearRed{p.nearv.as<double>(), nullptr, e.leaf_start[b] + r0, base,
                                (int64_t)cb * RP, std::min(64, cb - r0), ns, RP, RP, r0, 0});
    }
  }
  p.nearp.alloc(sizeof(double) * (size_t)std::max<int64_t>(ptot, 1));
  for (MrNearRed& r : red) r.part = p.nearp.as<double>();
  for (int k = 0; k < nleaf; k++) {
    const int st = e.leaf_start[k], n = e.leaf_count[k];
    for (int r0 = 0; r0 < n; r0 += nr) {
      MrNearTask t{};
      t.ta = st + r0;
      t.M = std::min(nr, n - r0);
      t.out_row = st + r0;
      if (pairs) {
        for (int q : {3, 4, 5, 6}) {
          const int b = nbr_of(k, q);
          if (b < 0 || e.leaf_count[b] == 0) continue;
          t.ss[t.nsrc] = e.leaf_start[b];
          t.sc[t.nsrc] = e.leaf_count[b];
          t.pt[t.nsrc] = slot_at.at(slot_key(k, q, r0 / nr));
          t.nsrc++;
        }
        t.npair = t.nsrc;
        t.ss[t.nsrc] = st;
        t.sc[t.nsrc] = n;
        t.nsrc++;
      } else {
        for (int q = 0; q <= 8; q++) {
          const int b = q < 8 ? nbr_of(k, q) : k;
          if (b < 0 || e.leaf_count[b] == 0) continue;
          t.ss[t.nsrc] = e.leaf_start[b];
          t.sc[t.nsrc] = e.leaf_count[b];
          t.nsrc++;
        }
      }
      near.push_back(t);
    }
  }
  upload_vec(p.near_red, red, s);
  p.h_near_red = std::move(red);
  upload_tasks(p.near, near, s);
  p.h_near = std::move(near);
  p.near_ctr.alloc(sizeof(int) * 4 * kMrCtrStride);
  MrNearProb pr{e.x.as<double>(), e.y.as<double>(), p.qs.p, p.nearv.as<double>(),
                p.nearp.as<double>(), e.kp, RP, RP};
  upload_vec(p.probs, std::vector<MrNearProb>{pr}, s);
  AFMM_HIP(hipStreamSynchronize(s));
  e.mr_plans.push_back(std::move(pl));
  return *e.mr_plans.back();
}

template <int P, int NCT, int IM, int D>
void launch_gemm_d(const MrTask* t, int n, int c, int xcd, int kmax, size_t shm, hipStream_t s) {
  const unsigned grid = ceil_div(n, 4);
  switch (c) {
    case 0: k_mr_gemm<P, NCT, 1, IM, D><<<grid, 256, shm, s>>>(t, n, xcd, kmax); break;
    case 1: k_mr_gemm<P, NCT, 2, IM, D><<<grid, 256, shm, s>>>(t, n, xcd, kmax); break;
    case 2: k_mr_gemm<P, NCT, 3, IM, D><<<grid, 256, shm, s>>>(t, n, xcd, kmax); break;
    default: k_mr_gemm<P, NCT, 4, IM, D><<<grid, 256, shm, s>>>(t, n, xcd, kmax); break;
  }
  AFMM_LAUNCH_CHECK();
}

// Indices staged in LDS up to this many bytes per workgroup (4 waves).
constexpr size_t kMrIdxLds = 48 * 1024;

// depth: register sets of the chunk ring (3, 4 or 6)
template <int P, int NCT>
void launch_gemm(const MrTaskList& tl, int xcd, int depth, hipStream_t s) {
  for (int c = 0; c < 4; c++) {
    const int n = tl.cls[c + 1] - tl.cls[c];
    if (n <= 0) continue;
    const MrTask* t = tl.buf.as<MrTask>() + tl.cls[c];
    if (tl.mode == 0) {
      launch_gemm_d<P, NCT, 0, 3>(t, n, c, xcd, 0, 0, s);
    } else if (tl.mode == 3) {
      if constexpr (P == PREC_BF16) {
        throw Error(6, "pair-stored M2L tasks in BF16");
      } else {
        // descriptors per wave: the largest task's, padded to whole chunks
        const int kmm = std::max(4, (tl.kmax[c] + 15) / 16 * 4);
        const size_t shm = sizeof(MrMicro) * 4 * (size_t)kmm;
        if (shm > kMrIdxLds) launch_gemm_d<P, NCT, 4, 3>(t, n, c, xcd, 0, 0, s);
        else if (depth >= 6) launch_gemm_d<P, NCT, 3, 6>(t, n, c, xcd, kmm, shm, s);
        else if (depth == 4) launch_gemm_d<P, NCT, 3, 4>(t, n, c, xcd, kmm, shm, s);
        else launch_gemm_d<P, NCT, 3, 3>(t, n, c, xcd, kmm, shm, s);
      }
    } else if (tl.mode == 5) {
      if constexpr (P != PREC_F64) {
        throw Error(6, "48-bit pair-stored M2L tasks outside FP64");
      } else {
        const int kmm = std::max(4, (tl.kmax[c] + 15) / 16 * 4);
        const size_t shm = sizeof(MrMicro) * 4 * (size_t)kmm;
        if (shm > kMrIdxLds) launch_gemm_d<P, NCT, 6, 3>(t, n, c, xcd, 0, 0, s);
        else if (depth >= 6) launch_gemm_d<P, NCT, 5, 6>(t, n, c, xcd, kmm, shm, s);
        else if (depth == 4) launch_gemm_d<P, NCT, 5, 4>(t, n, c, xcd, kmm, shm, s);
        else launch_gemm_d<P, NCT, 5, 3>(t, n, c, xcd, kmm, shm, s);
      }
    } else {
      const int kmax = std::max(1, tl.kmax[c]);
      const size_t shm = sizeof(int) * 4 * (size_t)kmax;
      if (shm > kMrIdxLds) launch_gemm_d<P, NCT, 2, 3>(t, n, c, xcd, 0, 0, s);
      else if (depth >= 6) launch_gemm_d<P, NCT, 1, 6>(t, n, c, xcd, kmax, shm, s);
      else if (depth == 4) launch_gemm_d<P, NCT, 1, 4>(t, n, c, xcd, kmax, shm, s);
      else launch_gemm_d<P, NCT, 1, 3>(t, n, c, xcd, kmax, shm, s);
    }
  }
}

void launch_gemm_p(int prec, int nct, const MrTaskList& tl, int xcd, int depth, hipStream_t s) {
  if (nct == 1) {
    switch (prec) {
      case PREC_F64: launch_gemm<PREC_F64, 1>(tl, xcd, depth, s); break;
      case PREC_F32: launch_gemm<PREC_F32, 1>(tl, xcd, depth, s); break;
      default: launch_gemm<PREC_BF16, 1>(tl, xcd, depth, s); break;
    }
  } else {
    switch (prec) {
      case PREC_F64: launch_gemm<PREC_F64, 2>(tl, xcd, depth, s); break;
      case PREC_F32: launch_gemm<PREC_F32, 2>(tl, xcd, depth, s); break;
      default: launch_gemm<PREC_BF16, 2>(tl, xcd, depth, s); break;
    }
  }
}

// The near field of a task list (one kernel id): one persistent launch per
// row-tile class, ctr holding 4 counters kMrCtrStride apart.
struct NearLaunch {
  const MrTaskList* tl;
  const MrNearProb* probs;
  const double2* ntab;  // the engine's near-field tables
  int* ctr;
  int kid, num_cus, near_waves;
};

template <int P, int NCT>
void launch_near_t(const NearLaunch& nl, hipStream_t s) {
  const MrTaskList& tl = *nl.tl;
  int* ctr = nl.ctr;
  k_mr_zero<<<1, 4 * kMrCtrStride, 0, s>>>(ctr, 4 * kMrCtrStride);
  AFMM_LAUNCH_CHECK();
  if (near_uses_lds(P, NCT)) {
    // one persistent launch over every task: near_waves (at least 2)
    // workgroups of 4 waves per CU
    if (tl.n <= 0) return;
    const MrNearTask* t = tl.buf.as<MrNearTask>();
    const unsigned grid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(tl.n, (int64_t)nl.num_cus * std::max(2, nl.near_waves)));
    AFMM_DISPATCH_KERNEL(nl.kid, KID,
                         k_mr_p2p_lds<P, NCT, KID><<<grid, 256, 0, s>>>(
                             t, tl.n, nl.probs, nl.ntab, ctr));
    AFMM_LAUNCH_CHECK();
    return;
  }
  for (int c = 0; c < 4; c++) {
    const int n = tl.cls[c + 1] - tl.cls[c];
    if (n <= 0) continue;
    const MrNearTask* t = tl.buf.as<MrNearTask>() + tl.cls[c];
    // waves per SIMD of the persistent grid (4 waves per workgroup)
    const int64_t slots = (int64_t)nl.num_cus * nl.near_waves;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 4), slots));
    int* cc = ctr + c * kMrCtrStride;
#define AFMM_MR_NEAR(NRT)                                                                        \
  AFMM_DISPATCH_KERNEL(nl.kid, KID,                                                              \
                       k_mr_p2p<P, NCT, NRT, KID><<<grid, 256, 0, s>>>(t, n, nl.probs, nl.ntab,   \
                                                                       cc))
    switch (c) {
      case 0: AFMM_MR_NEAR(1); break;
      case 1: AFMM_MR_NEAR(2); break;
      case 2: AFMM_MR_NEAR(3); break;
      default: AFMM_MR_NEAR(4); break;
    }
#undef AFMM_MR_NEAR
    AFMM_LAUNCH_CHECK();
  }
}

void launch_near_p(int prec, int nct, const NearLaunch& nl, hipStream_t s) {
  if (nct == 1) {
    switch (prec) {
      case PREC_F64: launch_near_t<PREC_F64, 1>(nl, s); break;
      case PREC_F32: launch_near_t<PREC_F32, 1>(nl, s); break;
      default: launch_near_t<PREC_BF16, 1>(nl, s); break;
    }
  } else {
    switch (prec) {
      case PREC_F64: launch_near_t<PREC_F64, 2>(nl, s); break;
      case PREC_F32: launch_near_t<PREC_F32, 2>(nl, s); break;
      default: launch_near_t<PREC_BF16, 2>(nl, s); break;
    }
  }
}

// The pair partials' reduction after the near field (none without pairs).
void launch_near_red(const DevBuf& red, size_t n, int nct, hipStream_t s) {
  if (n == 0) return;
  k_mr_near_red<<<(unsigned)n, 256, 0, s>>>(red.as<MrNearRed>(), 16 * nct);
  AFMM_LAUNCH_CHECK();
}

void launch_near_plan(Engine& e, const MrPlan& p, int near_waves, hipStream_t s) {
  NearLaunch nl{&p.near, p.probs.as<MrNearProb>(), e.near_tab.as<double2>(),
                p.near_ctr.as<int>(), e.kid, e.num_cus, near_waves};
  launch_near_p(p.prec, p.nct, nl, s);
  launch_near_red(p.near_red, p.h_near_red.size(), p.nct, s);
}

template <class VT, int NCT>
void launch_io(bool gather, const MrIoTask& d, hipStream_t s) {
  if (d.cnt <= 0) return;
  const unsigned grid = ceil_div(d.cnt, kMrIoPoints);
  if (gather) k_mr_gather<VT, NCT><<<grid, 256, 0, s>>>(d);
  else k_mr_out<VT, NCT><<<grid, 256, 0, s>>>(d);
  AFMM_LAUNCH_CHECK();
}

void launch_io_p(int prec, int nct, bool gather, const MrIoTask& d, hipStream_t s) {
  if (prec == PREC_F64) {
    if (nct == 1) launch_io<double, 1>(gather, d, s);
    else launch_io<double, 2>(gather, d, s);
  } else {
    if (nct == 1) launch_io<float, 1>(gather, d, s);
    else launch_io<float, 2>(gather, d, s);
  }
}

template <class VT, int NCT>
void launch_io_multi_t(bool gather, const MrIoTask* t, int n, int blocks, hipStream_t s) {
  if (gather) k_mr_io_multi<VT, NCT, true><<<blocks, 256, 0, s>>>(t, n);
  else k_mr_io_multi<VT, NCT, false><<<blocks, 256, 0, s>>>(t, n);
  AFMM_LAUNCH_CHECK();
}

void launch_io_multi(int prec, int nct, bool gather, const MrIoTask* t, int n, int blocks,
                     hipStream_t s) {
  if (blocks <= 0) return;
  if (prec == PREC_F64) {
    if (nct == 1) launch_io_multi_t<double, 1>(gather, t, n, blocks, s);
    else launch_io_multi_t<double, 2>(gather, t, n, blocks, s);
  } else {
    if (nct == 1) launch_io_multi_t<float, 1>(gather, t, n, blocks, s);
    else launch_io_multi_t<float, 2>(gather, t, n, blocks, s);
  }
}

// Launches one task list makes (one per non-empty row-tile class).
int list_launches(const MrTaskList& tl) {
  int n = 0;
  for (int c = 0; c < 4; c++) n += tl.cls[c + 1] > tl.cls[c];
  return n;
}

}  // namespace

void mr_prepare(Engine& e, int prec) {
  if (prec == PREC_F64) {
    if (e.m2l_store == M2L_STORE_B48) build_b48(e);
    else build_f64(e);
  } else {
    build_reduced(e, prec);
  }
}

void mr_matvec(Engine& e, const double* d_b, int64_t ldb, double* d_y, int64_t ldy, int nrhs,
               int prec, hipStream_t s) {
  if (nrhs < 1 || nrhs > kMrMaxCols) throw Error(1, "blocked matvec takes 1..32 right-hand sides");
  const int nct = nrhs <= 16 ? 1 : 2;
  MrPlan& p = get_plan(e, prec, nct);
  const int L = e.L;
  const MrTune tn = mr_default_tune();
  MrIoTask io{};
  io.b = d_b;
  io.y = d_y;
  io.inv_perm = e.inv_perm_d.as<int>();
  io.qs = p.qs.p;
  io.pot = p.pot.p;
  io.nearv = p.nearv.as<double>();
  io.ldb = ldb;
  io.ldy = ldy;
  io.i0 = 0;
  io.cnt = e.n;
  io.nrhs = nrhs;
  hipEvent_t a, b;
  e.prof_begin("mr_gather", s, &a, &b);
  launch_io_p(prec, nct, true, io, s);
  e.prof_end(s, b);
  // the near field needs only the gathered charges: it runs beside the tree
  // (AFMM2D_NEAR_OVERLAP=0: after it, on the caller's stream)
  auto fork_near = [&]() {
    AFMM_HIP(hipEventRecord(e.ev_fork, s));
    AFMM_HIP(hipStreamWaitEvent(e.s_aux, e.ev_fork, 0));
    e.prof_begin("mr_p2p", e.s_aux, &a, &b);
    launch_near_plan(e, p, tn.near_waves, e.s_aux);
    e.prof_end(e.s_aux, b);
    AFMM_HIP(hipEventRecord(e.ev_join, e.s_aux));
  };
  const int au = knobs().near_after_up;
  const bool late = au > 0 || (au < 0 && nct == 1 && near_uses_lds(prec, nct));
  if (e.near_overlap && !late) fork_near();
  e.prof_begin("mr_m2m", s, &a, &b);
  for (int j = L; j >= 2; j--) launch_gemm_p(prec, nct, p.up[j], 0, 3, s);
  e.prof_end(s, b);
  if (e.near_overlap && late) fork_near();
  e.prof_begin("mr_m2l", s, &a, &b);
  launch_gemm_p(prec, nct, p.m2l, tn.xcd, tn.depth, s);
  e.prof_end(s, b);
  e.prof_begin("mr_l2l", s, &a, &b);
  for (int j = 2; j < L; j++) launch_gemm_p(prec, nct, p.down[j], 0, 3, s);
  launch_gemm_p(prec, nct, p.leaf, 0, 3, s);
  e.prof_end(s, b);
  if (e.near_overlap) {
    AFMM_HIP(hipStreamWaitEvent(s, e.ev_join, 0));
  } else {
    e.prof_begin("mr_p2p", s, &a, &b);
    launch_near_plan(e, p, tn.near_waves, s);
    e.prof_end(s, b);
  }
  e.prof_begin("mr_out", s, &a, &b);
  launch_io_p(prec, nct, false, io, s);
  e.prof_end(s, b);
}

void add_cols_device(double* d_y, int64_t ldy, const double* d_t, int64_t ldt, int n, int nrhs,
                     hipStream_t s) {
  const int64_t tot = (int64_t)n * nrhs;
  if (tot <= 0) return;
  k_add_cols<<<ceil_div(tot, 256), 256, 0, s>>>(d_y, ldy, d_t, ldt, n, nrhs);
  AFMM_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Batches of problems.

struct MrBatch {
  std::vector<Engine*> eng;
  std::vector<MrPlan*> plans;
  int prec = 0, nct = 1, nrhs = 1, device = 0;
  std::vector<MrTaskList> up, down;  // by step: up aligned at the leaves, down from level 2
  MrTaskList m2l;
  int m2l_rows = 0;
  struct NearGroup {
    int kid = 0;
    const Engine* first = nullptr;  // owner of the near-field tables
    MrTaskList tl;
    DevBuf probs;
  };
  std::vector<NearGroup> near;
  DevBuf ctr;                   // 4 counters per near group
  DevBuf near_red;              // every problem's pair-partial reduction
  size_t n_near_red = 0;
  DevBuf io;                    // MrIoTask per problem, i0 = first workgroup
  int nio = 0, io_blocks = 0;
  hipStream_t side = nullptr;   // near field, beside the tree passes
  hipStream_t cap = nullptr;    // graph capture
  hipEvent_t fork = nullptr, join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  MrTune tune;
  bool tuned = false;
  std::string key;
  ~MrBatch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {

// Tuned parameters per "architecture|shape class", for the process and the
// file AFMM2D_TUNE_CACHE names (lines "key depth xcd near_waves rows_m2l").
std::mutex g_tune_mu;
std::map<std::string, MrTune> g_tune_cache;
bool g_tune_loaded = false;

const char* tune_file() {
  const char* f = std::getenv("AFMM2D_TUNE_CACHE");
  return f && *f ? f : nullptr;
}

void load_tune_cache() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* f = tune_file();
  if (!f) return;
  std::ifstream in(f);
  std::string key;
  MrTune t;
  while (in >> key >> t.depth >> t.xcd >> t.near_waves >> t.rows_m2l) g_tune_cache[key] = t;
}

void store_tune(const std::string& key, const MrTune& t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = t;
  if (const char* f = tune_file()) {
    std::ofstream out(f, std::ios::app);
    out << key << ' ' << t.depth << ' ' << t.xcd << ' ' << t.near_waves << ' ' << t.rows_m2l
        << '\n';
  }
}

// Shape class: precision, column tiles, kernel ids present, log2 of the
// total points and of the number of problems.
std::string batch_key(const MrBatch& b) {
  hipDeviceProp_t prop;
  AFMM_HIP(hipGetDeviceProperties(&prop, b.device));
  std::string arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch.resize(colon);
  int64_t pts = 0;
  unsigned kids = 0;
  for (const Engine* e : b.eng) {
    pts += e->n;
    kids |= 1u << e->kid;
  }
  const int lp = (int)std::lround(std::log2((double)std::max<int64_t>(pts, 1)));
  const int lb = (int)std::lround(std::log2((double)b.eng.size()));
  std::ostringstream k;
  k << arch << "|p" << b.prec << "|c" << b.nct << "|k" << kids << "|n" << lp << "|b" << lb;
  return k.str();
}

// Rebuilds the batch's M2L list with rows rows per task.
void batch_m2l(MrBatch& b, int rows, hipStream_t s) {
  if (rows == b.m2l_rows) return;
  std::vector<MrTask> all;
  for (size_t i = 0; i < b.eng.size(); i++) m2l_tasks(*b.eng[i], *b.plans[i], rows, all);
  AFMM_HIP(hipStreamSynchronize(s));
  upload_tasks(b.m2l, all, s);
  AFMM_HIP(hipStreamSynchronize(s));
  b.m2l_rows = rows;
}

void batch_enqueue(MrBatch& b, hipStream_t s) {
  const MrIoTask* io = b.io.as<MrIoTask>();
  launch_io_multi(b.prec, b.nct, true, io, b.nio, b.io_blocks, s);
  auto fork_near = [&]() {
    AFMM_HIP(hipEventRecord(b.fork, s));
    AFMM_HIP(hipStreamWaitEvent(b.side, b.fork, 0));
    for (size_t g = 0; g < b.near.size(); g++) {
      const MrBatch::NearGroup& ng = b.near[g];
      NearLaunch nl{&ng.tl, ng.probs.as<MrNearProb>(), ng.first->near_tab.as<double2>(),
                    b.ctr.as<int>() + 4 * kMrCtrStride * g, ng.kid, ng.first->num_cus,
                    b.tune.near_waves};
      launch_near_p(b.prec, b.nct, nl, b.side);
    }
    launch_near_red(b.near_red, b.n_near_red, b.nct, b.side);
    AFMM_HIP(hipEventRecord(b.join, b.side));
  };
  const bool late = knobs().near_after_up > 0;
  if (!late) fork_near();
  for (const MrTaskList& tl : b.up) launch_gemm_p(b.prec, b.nct, tl, 0, 3, s);
  if (late) fork_near();
  launch_gemm_p(b.prec, b.nct, b.m2l, b.tune.xcd, b.tune.depth, s);
  for (const MrTaskList& tl : b.down) launch_gemm_p(b.prec, b.nct, tl, 0, 3, s);
  AFMM_HIP(hipStreamWaitEvent(s, b.join, 0));
  launch_io_multi(b.prec, b.nct, false, io, b.nio, b.io_blocks, s);
}

void drop_graph(MrBatch& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
}

}  // namespace

MrBatch* mr_batch_create(const std::vector<Engine*>& eng, const double* const* d_b,
                         const int64_t* ldb, double* const* d_y, const int64_t* ldy, int nrhs,
                         int prec) {
  if (eng.empty()) throw Error(1, "batch without problems");
  if (nrhs < 1 || nrhs > kMrMaxCols) throw Error(1, "batched matvec takes 1..32 right-hand sides");
  auto bp = std::make_unique<MrBatch>();
  MrBatch& b = *bp;
  b.eng = eng;
  b.prec = prec;
  b.nrhs = nrhs;
  b.nct = nrhs <= 16 ? 1 : 2;
  b.device = eng[0]->device;
  int maxL = 0;
  for (size_t i = 0; i < eng.size(); i++) {
    const Engine& e = *eng[i];
    if (e.device != b.device) throw Error(1, "batch problems on different devices");
    if (ldb[i] < e.n || ldy[i] < e.n) throw Error(1, "leading dimension below N");
    // FP64 batches read every problem's blocks through one M2L task list
    if (prec == PREC_F64 && e.m2l_store != eng[0]->m2l_store)
      throw Error(1, "FP64 batch of problems with different M2L storage");
    maxL = std::max(maxL, e.L);
  }
  AFMM_HIP(hipSetDevice(b.device));
  for (Engine* e : eng) b.plans.push_back(&get_plan(*e, prec, b.nct));
  AFMM_HIP(hipStreamCreateWithFlags(&b.side, hipStreamNonBlocking));
  AFMM_HIP(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking));
  AFMM_HIP(hipEventCreateWithFlags(&b.fork, hipEventDisableTiming));
  AFMM_HIP(hipEventCreateWithFlags(&b.join, hipEventDisableTiming));
  hipStream_t s = b.side;
  // upward step u: level L - u of every problem with L - u >= 2; downward
  // step d: level 2 + d (< L) and, at d = L - 2, the leaf pass
  const int nsteps = maxL - 1;
  std::vector<std::vector<MrTask>> up(nsteps), down(nsteps);
  for (size_t i = 0; i < eng.size(); i++) {
    const Engine& e = *eng[i];
    const MrPlan& p = *b.plans[i];
    for (int j = e.L; j >= 2; j--) {
      auto& v = up[e.L - j];
      v.insert(v.end(), p.h_up[j].begin(), p.h_up[j].end());
    }
    for (int j = 2; j < e.L; j++) {
      auto& v = down[j - 2];
      v.insert(v.end(), p.h_down[j].begin(), p.h_down[j].end());
    }
    auto& v = down[e.L - 2];
    v.insert(v.end(), p.h_leaf.begin(), p.h_leaf.end());
  }
  b.up.resize(nsteps);
  b.down.resize(nsteps);
  for (int u = 0; u < nsteps; u++) {
    upload_tasks(b.up[u], up[u], s);
    upload_tasks(b.down[u], down[u], s);
  }
  // near field: one group per kernel id, problems indexed within the group
  for (int kid = 0; kid < 32; kid++) {
    std::vector<MrNearTask> tasks;
    std::vector<MrNearProb> probs;
    const Engine* first = nullptr;
    for (size_t i = 0; i < eng.size(); i++) {
      const Engine& e = *eng[i];
      if (e.kid != kid) continue;
      if (!first) first = &e;
      const MrPlan& p = *b.plans[i];
      const int pi = (int)probs.size();
      // the batch's columns: the partials' unused columns are never written
      probs.push_back(MrNearProb{e.x.as<double>(), e.y.as<double>(), p.qs.p, p.nearv.as<double>(),
                                 p.nearp.as<double>(), e.kp, (nrhs + 1) & ~1, (nrhs + 1) & ~1});
      for (MrNearTask t : p.h_near) {
        t.prob = pi;
        tasks.push_back(t);
      }
    }
    if (!first) continue;
    b.near.emplace_back();
    MrBatch::NearGroup& ng = b.near.back();
    ng.kid = kid;
    ng.first = first;
    upload_tasks(ng.tl, tasks, s);
    upload_vec(ng.probs, probs, s);
  }
  b.ctr.alloc(sizeof(int) * 4 * kMrCtrStride * std::max<size_t>(b.near.size(), 1));
  {
    std::vector<MrNearRed> red;
    for (const MrPlan* p : b.plans) red.insert(red.end(), p->h_near_red.begin(), p->h_near_red.end());
    // the batch's columns only, the partials' rows packed at that width
    for (MrNearRed& r : red) r.pw = r.prs = (nrhs + 1) & ~1;
    b.n_near_red = red.size();
    upload_vec(b.near_red, red, s);
  }
  // gather / output: problem i owns workgroups [i0, i0 + ceil(n / kMrIoPoints))
  std::vector<MrIoTask> io;
  int blocks = 0;
  for (size_t i = 0; i < eng.size(); i++) {
    const Engine& e = *eng[i];
    const MrPlan& p = *b.plans[i];
    MrIoTask t{};
    t.b = d_b[i];
    t.y = d_y[i];
    t.inv_perm = e.inv_perm_d.as<int>();
    t.qs = p.qs.p;
    t.pot = p.pot.p;
    t.nearv = p.nearv.as<double>();
    t.ldb = ldb[i];
    t.ldy = ldy[i];
    t.i0 = blocks;
    t.cnt = e.n;
    t.nrhs = nrhs;
    io.push_back(t);
    const int64_t nb = blocks + (int64_t)ceil_div(e.n, kMrIoPoints);
    if (nb > INT32_MAX) throw Error(1, "batch too large");
    blocks = (int)nb;
  }
  b.nio = (int)io.size();
  b.io_blocks = blocks;
  upload_vec(b.io, io, s);
  // launch parameters: the cached tuning of this shape class, else defaults
  b.key = batch_key(b);
  b.tune = mr_default_tune();
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tune_cache();
    auto it = g_tune_cache.find(b.key);
    if (it != g_tune_cache.end()) {
      b.tune = it->second;
      b.tuned = true;
    }
  }
  batch_m2l(b, b.tune.rows_m2l, s);
  AFMM_HIP(hipStreamSynchronize(s));
  return bp.release();
}

void mr_batch_destroy(MrBatch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->side);
  (void)hipStreamSynchronize(b->cap);
  delete b;
}

void mr_batch_run(MrBatch* b, hipStream_t s) {
  AFMM_HIP(hipSetDevice(b->device));
  if (b->exec) {
    AFMM_HIP(hipGraphLaunch(b->exec, s));
  } else {
    batch_enqueue(*b, s);
  }
}

void mr_batch_graph(MrBatch* b, bool on) {
  AFMM_HIP(hipSetDevice(b->device));
  AFMM_HIP(hipStreamSynchronize(b->cap));
  drop_graph(*b);
  if (!on) return;
  AFMM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  try {
    batch_enqueue(*b, b->cap);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->cap, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  AFMM_HIP(hipStreamEndCapture(b->cap, &b->graph));
  AFMM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

double mr_batch_tune(MrBatch* b, hipStream_t s) {
  AFMM_HIP(hipSetDevice(b->device));
  const bool had_graph = b->exec != nullptr;
  drop_graph(*b);
  std::vector<MrTune> cand;
  for (int rows : {32, 64})
    for (int depth : {3, 4})
      for (int xcd : {0, 1})
        for (int waves : {1, 2, 4}) {
          MrTune t;
          t.rows_m2l = rows;
          t.depth = depth;
          t.xcd = xcd;
          t.near_waves = waves;
          cand.push_back(t);
        }
  hipEvent_t e0, e1;
  AFMM_HIP(hipEventCreate(&e0));
  AFMM_HIP(hipEventCreate(&e1));
  constexpr int kReps = 3;
  double best_ms = 1e30;
  MrTune best = b->tune;
  try {
    for (const MrTune& t : cand) {
      b->tune = t;
      batch_m2l(*b, t.rows_m2l, s);
      batch_enqueue(*b, s);  // warm-up
      AFMM_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < kReps; r++) batch_enqueue(*b, s);
      AFMM_HIP(hipEventRecord(e1, s));
      AFMM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      AFMM_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms / kReps < best_ms) {
        best_ms = ms / kReps;
        best = t;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tune = best;
  batch_m2l(*b, best.rows_m2l, s);
  b->tuned = true;
  store_tune(b->key, best);
  if (had_graph) mr_batch_graph(b, true);
  return best_ms;
}

MrBatchInfo mr_batch_info(const MrBatch* b) {
  MrBatchInfo in{};
  int n = b->io_blocks > 0 ? 2 : 0;
  for (const MrTaskList& tl : b->up) n += list_launches(tl);
  for (const MrTaskList& tl : b->down) n += list_launches(tl);
  n += list_launches(b->m2l);
  for (const MrBatch::NearGroup& g : b->near) n += 1 + list_launches(g.tl);
  if (b->n_near_red > 0) n++;  // the pair partials' reduction
  in.launches = n;
  in.depth = b->tune.depth;
  in.xcd = b->tune.xcd;
  in.near_waves = b->tune.near_waves;
  in.rows_m2l = b->tune.rows_m2l;
  in.captured = b->exec != nullptr;
  in.tuned = b->tuned;
  in.nproblems = (int)b->eng.size();
  return in;
}

}  // namespace afmm
