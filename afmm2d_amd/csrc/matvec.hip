// matvec.hip -- the matvec passes of AFMM2D on the GPU (FP64).
//
// Reference flow (AFMM2D.cpp:46-65, AFMM2DTree.cpp:1120-1225): gather q into
// KD order, M2M upward (levels L..2, out_k = L2P_k^T src), M2L (in_k =
// sum over the interaction list of M2L[k][kIL] out_kIL), L2L downward
// (children.in += L2P_k in_k; leaves: potential = L2P_k in_k), near field
// (potential += K(leaf, neighbour) q_neighbour for 8 neighbours and self),
// scatter back to caller order.
//
// Stream plan (matvec_device, default schedule): the M2M of every level, the
// M2L, the L2L chain and the leaf pass run in order on the caller's stream;
// after the upward pass the near field forks onto the handle's internal
// stream and runs beside the M2L and the L2L chain, and the leaf pass (leaf
// L2P + M2L sums + near field + scatter, fused) joins it.  Other schedules
// are options (see matvec_device).
//
// M2L reads each stored block once: a pair kernel holds B = K(skel_k,
// skel_b) in registers (the whole block up to 64 x 64, larger ones streamed
// in column chunks; 16-byte loads, issued before use) and writes B out_b and
// +-B^T out_k (the sign of an antisymmetric kernel part) as partials; the
// downward pass adds each box's partials in its interaction-slot order, so
// results do not depend on scheduling.
#include <cstdlib>
#include <mutex>

#include "engine.hpp"
#include "near_math.hpp"

namespace afmm {

namespace {

// Wave priority of the memory-bound kernels that run beside the near field
// (M2L, M2M, L2L): they issue their loads ahead of the near field's FP64
// arithmetic, which fills the remaining issue slots.
#ifndef AFMM_MEM_PRIO
#define AFMM_MEM_PRIO 3
#endif

__global__ void k_gather(const double* __restrict__ b, const int* __restrict__ perm,
                         double* __restrict__ qs, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) qs[i] = b[perm[i]];
}

// Batched small GEMV, the core of the upward and downward passes:
// y[cb + c] = sum_i M[i][cb + c] x[i] over one column chunk (rc <= 64
// columns) of a box's matrix M (nr x nc, row-major, leading dimension nc),
// x staged in LDS.  Lane l = h rc + c holds column cb + c and the rows
// i = h + G t (G = 64 / rc row groups, so each step t reads G rc contiguous
// elements); the waves of the workgroup take the steps round robin, B steps
// per batch.  Lanes with h >= G use h = nr and contribute nothing.
struct GemvLane {
  int rc, G, h, c;
};

__device__ __forceinline__ GemvLane gemv_lane(int ncols, int cb, int nr) {
  GemvLane q;
  q.rc = min(64, ncols - cb);
  q.G = 64 / q.rc;
  const int l = threadIdx.x & 63;
  const int h = l / q.rc;
  q.c = l - h * q.rc;
  q.h = h < q.G ? h : nr;
  return q;
}

// Issues the loads of one batch (steps t0, t0 + stride, ...): addresses are
// clamped into the matrix, so no load is conditional.
template <int B>
__device__ __forceinline__ void gemv_load(double (&pv)[B], const double* __restrict__ M, int nr,
                                          int nc, int cb, const GemvLane& q, int t0, int stride) {
  if (nr == 0) return;
#pragma unroll
  for (int b = 0; b < B; b++) {
    const int i = q.h + q.G * (t0 + b * stride);
    pv[b] = M[(int64_t)(i < nr ? i : nr - 1) * nc + cb + q.c];
  }
}

template <int B>
__device__ __forceinline__ double gemv_fma(const double (&pv)[B], double acc, int nr,
                                           const GemvLane& q, int t0, int stride,
                                           const double* xs) {
#pragma unroll
  for (int b = 0; b < B; b++) {
    const int i = q.h + q.G * (t0 + b * stride);
    const double xv = xs[i < nr ? i : 0];
    acc = i < nr ? fma(pv[b], xv, acc) : acc;
  }
  return acc;
}

// The whole column chunk for this lane over the steps t_first, t_first +
// stride, ..., given the first batch already loaded into pv (issued before x
// was staged, so its latency overlaps the staging).
template <int B>
__device__ __forceinline__ double gemv_chunk(double (&pv)[B], const double* __restrict__ M, int nr,
                                             int nc, int cb, const GemvLane& q, const double* xs,
                                             int t_first, int stride) {
  const int T = (nr + q.G - 1) / q.G;
  double acc = 0.0;
  if (t_first < T) acc = gemv_fma<B>(pv, acc, nr, q, t_first, stride, xs);
  for (int t0 = t_first + stride * B; t0 < T; t0 += stride * B) {
    gemv_load<B>(pv, M, nr, nc, cb, q, t0, stride);
    acc = gemv_fma<B>(pv, acc, nr, q, t0, stride, xs);
  }
  return acc;
}

// Sum of the row groups and waves for column threadIdx.x < rc, in a fixed
// order; every thread of the workgroup must call it.
template <int WPB>
__device__ __forceinline__ double reduce_groups(double acc, double* red, const GemvLane& q) {
  if (WPB == 1 && q.G == 1) return acc;
  red[threadIdx.x] = acc;
  __syncthreads();
  double s = 0.0;
  if ((int)threadIdx.x < q.rc) {
    for (int w = 0; w < WPB; w++)
      for (int h = 0; h < q.G; h++) s += red[w * 64 + h * q.rc + threadIdx.x];
  }
  __syncthreads();
  return s;
}

// A box's local coefficients into xs: (has_parent ? the parent's L2L
// contribution : 0) + its M2L partials in interaction-slot order.  so[t] is
// where the box's t-th present slot finds its product in part (-1: none, a
// zero term); the offsets are the same for the whole workgroup, and every
// load is issued at once.
template <int WPB>
__device__ __forceinline__ void local_coeffs(double* xs, int r, const double* __restrict__ par,
                                             int has_parent, const double* __restrict__ part,
                                             const int* __restrict__ so) {
  int o[kILSlots];
#pragma unroll
  for (int t = 0; t < kILSlots; t++) o[t] = so[t];
  for (int c = threadIdx.x; c < r; c += 64 * WPB) {
    double v[kILSlots];
#pragma unroll
    for (int t = 0; t < kILSlots; t++) v[t] = o[t] >= 0 ? part[o[t] + c] : 0.0;
    double acc = has_parent ? par[c] : 0.0;
#pragma unroll
    for (int t = 0; t < kILSlots; t++) acc += v[t];
    xs[c] = acc;
  }
}

// M2M for one level: out_k = L2P_k^T src_k with the stored rows of L2P
// (row-major, (nin - rank) x rank) and the identity on its skeleton rows:
// out_k[c] = sum_t P[t][c] src[rowmap[t]] + src[rowmap[nc + c]], nc = nin -
// rank.  src is staged in LDS in rowmap order.  Leaf level (b != nullptr):
// src_k is the leaf's charges, gathered here from the caller's order
// (b[perm[.]]) and also written to qs in KD order for the near field; upper
// levels: the children's concatenated multipoles (always contiguous).  One
// workgroup of WPB waves per box.
template <int WPB, int B>
__global__ __launch_bounds__(64 * WPB) void k_up(const double* __restrict__ l2p_rm,
                                                 const int64_t* __restrict__ l2poff,
                                                 const int* __restrict__ rank,
                                                 const int* __restrict__ nin,
                                                 const int64_t* __restrict__ soff,
                                                 const int* __restrict__ rowmap,
                                                 const double* __restrict__ src,
                                                 const int64_t* __restrict__ src_off_child,
                                                 const double* __restrict__ b,
                                                 const int* __restrict__ perm,
                                                 const int* __restrict__ leaf_start,
                                                 double* __restrict__ qs,
                                                 double* __restrict__ outv, int64_t gbase,
                                                 int nb) {
  extern __shared__ double xs[];
  __shared__ double red[64 * WPB];
  if (AFMM_MEM_PRIO) __builtin_amdgcn_s_setprio(AFMM_MEM_PRIO);
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  const int n = nin[g];
  const int nc = n - r;
  const double* P = l2p_rm + l2poff[g];
  double* out = outv + soff[g];
  const int w = threadIdx.x >> 6;
  double pv[B];
  GemvLane q;
  if (r > 0 && nc > 0) {
    q = gemv_lane(r, 0, nc);
    gemv_load<B>(pv, P, nc, r, 0, q, w, WPB);
  }
  if (b) {
    const int st = leaf_start[k];
    const int* rm = rowmap + st;
    for (int t = threadIdx.x; t < n; t += 64 * WPB) {
      const int i = rm[t];
      const double v = b[perm[st + i]];
      xs[t] = v;
      qs[st + i] = v;
    }
  } else {
    const int64_t so = src_off_child[4 * k];
    const double* x = src + so;
    const int* rm = rowmap + so;
    for (int t = threadIdx.x; t < n; t += 64 * WPB) xs[t] = x[rm[t]];
  }
  if (r == 0) return;
  __syncthreads();
  if (nc == 0) {
    for (int c = threadIdx.x; c < r; c += 64 * WPB) out[c] = xs[c];
    return;
  }
  for (int cb = 0; cb < r; cb += 64) {
    if (cb > 0) {
      q = gemv_lane(r, cb, nc);
      gemv_load<B>(pv, P, nc, r, cb, q, w, WPB);
    }
    const double acc = gemv_chunk<B>(pv, P, nc, r, cb, q, xs, w, WPB);
    const double s = reduce_groups<WPB>(acc, red, q);
    if ((int)threadIdx.x < q.rc) out[cb + threadIdx.x] = s + xs[nc + cb + threadIdx.x];
  }
}

// Sum over the 32 lanes of each half-wave of 2^S values per lane, as a
// butterfly: the stage with lane mask m (16, 8, ...) pairs lane li with
// li ^ m; each keeps the half of its values selected by its bit m and adds
// the partner's copy of that half.  After S halving stages (masks 16 ..
// 2^(5-S)) the remaining masks are plain pairwise sums, and lane li holds the
// 32-lane sum of value li >> (5 - S).
template <int S>
__device__ __forceinline__ double half_wave_transpose_sum(double (&p)[1 << S]) {
  const int li = threadIdx.x & 31;
#pragma unroll
  for (int st = 0; st < 5; st++) {
    const int m = 16 >> st;
    if (st < S) {
      constexpr int NV = 1 << S;
      const int h = NV >> (st + 1);
      const bool upper = (li & m) != 0;
#pragma unroll
      for (int j = 0; j < NV / 2; j++) {
        if (j < h) {
          const double keep = upper ? p[j + h] : p[j];
          const double give = upper ? p[j] : p[j + h];
          p[j] = keep + __shfl_xor(give, m, 32);
        }
      }
    } else {
      p[0] += __shfl_xor(p[0], m, 32);
    }
  }
  return p[0];
}

// M2M for one level on the column-major stored L2P rows (transposed GEMV):
// out[c] = sum_t P[c nc + t] x[rowmap[t]] + x[rowmap[nc + c]], nc = nin -
// rank.  Half-wave `half` of wave w owns column 2 j + half of the column
// pairs j = w + WPB u (u < 2^S); its 32 lanes own the rows 64 kc + li and
// 64 kc + 32 + li of each 64-row chunk kc < NCH, so every load instruction
// reads 256 contiguous bytes per half-wave with all lanes active (the
// row-major kernel leaves lanes idle when the rank does not divide 64).  All
// loads are issued before the first use; each lane's per-column partial sums
// are then combined over the 32 lanes.  Upper levels only: x is the
// children's concatenated multipoles.
template <int S, int NCH, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_up_t(const double* __restrict__ l2p_cm,
                                                   const int64_t* __restrict__ l2poff,
                                                   const int* __restrict__ rank,
                                                   const int* __restrict__ nin,
                                                   const int64_t* __restrict__ soff,
                                                   const int* __restrict__ rowmap,
                                                   const double* __restrict__ src,
                                                   const int64_t* __restrict__ src_off_child,
                                                   double* __restrict__ outv, int64_t gbase,
                                                   int nb) {
  constexpr int NV = 1 << S;
  if (AFMM_MEM_PRIO) __builtin_amdgcn_s_setprio(AFMM_MEM_PRIO);
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  const int nc = nin[g] - r;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, half = l >> 5, li = l & 31;
  const double* P = l2p_cm + l2poff[g];
  // matrix loads first (they need only the box's scalars)
  double a[NV][NCH][2];
  if (r > 0 && nc > 0) {
#pragma unroll
    for (int u = 0; u < NV; u++) {
      const int c = 2 * (w + WPB * u) + half;
      const int64_t cb = (int64_t)(c < r ? c : r - 1) * nc;
#pragma unroll
      for (int kc = 0; kc < NCH; kc++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const int i = 64 * kc + 32 * e + li;
          a[u][kc][e] = P[cb + (i < nc ? i : nc - 1)];
        }
    }
  } else {
#pragma unroll
    for (int u = 0; u < NV; u++)
#pragma unroll
      for (int kc = 0; kc < NCH; kc++) a[u][kc][0] = a[u][kc][1] = 0.0;
  }
  if (r == 0) return;
  double xv[NCH][2];
  const int64_t so = src_off_child[4 * k];
  const double* x = src + so;
  const int* rm = rowmap + so;
#pragma unroll
  for (int kc = 0; kc < NCH; kc++)
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int i = 64 * kc + 32 * e + li;
      xv[kc][e] = i < nc ? x[rm[i]] : 0.0;
    }
  // the skeleton-row term of the column this lane writes (if it writes one)
  const bool writer = (li & ((32 >> S) - 1)) == 0;
  const int cw = 2 * (w + WPB * (li >> (5 - S))) + half;
  const double xsk = writer && cw < r ? x[rm[nc + cw]] : 0.0;
  double p[NV];
#pragma unroll
  for (int u = 0; u < NV; u++) {
    double s = 0.0;
#pragma unroll
    for (int kc = 0; kc < NCH; kc++)
#pragma unroll
      for (int e = 0; e < 2; e++) s = fma(a[u][kc][e], xv[kc][e], s);
    p[u] = 2 * (w + WPB * u) + half < r ? s : 0.0;
  }
  const double sum = half_wave_transpose_sum<S>(p);
  if (writer && cw < r) outv[soff[g] + cw] = sum + xsk;
}

// Wave-level sums through DPP (data-parallel lane moves) on the two 32-bit
// halves of a double.  Lane layout of the M2L tiles: lane = 8 ci + rp.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
// wave_rol:1: lane i receives lane i + 1's value, lane 63 lane 0's
constexpr int kDppWaveRol1 = 0x134;
// sum over the 8 consecutive lanes 8g..8g+7 (quad_perm xor 1, xor 2, then
// row_half_mirror, which pairs each quad with the other half of its 8 lanes)
__device__ __forceinline__ double sum_lanes8(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  return v;
}
// sum over lanes l, l^8, l^16, ..., l^56 (row_ror 8 swaps the halves of each
// 16-lane row; then two cross-row exchanges)
__device__ __forceinline__ double sum_lanes_stride8(double v) {
  v += dpp_move<0x128>(v);
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// The stored M2L blocks as seen by the pair kernels: element i of a block
// pair-indexed (entries 2 i and 2 i + 1 of the block, one row pair of a
// column), and the factor the pair's products take at the output.
// FP64: one 16-byte load, factor 1.  48-bit block floating point
// (M2L_STORE_B48): one 12-byte load of the row pair's record, its biased
// mantissas turned exactly into integer-valued doubles (the bits of
// 2^52 + u, minus 2^52 + 2^47: one bit operation and one subtraction); the block's
// power of two 2^sexp is the factor, so the sums over the block are those
// of the FP64 products of the stored values.  Block offsets and leading
// dimensions are even, so a row pair is one record.
struct M2LBlocks {
  const double* f64;
  const B48Pair* b48;
};
template <int ST>
struct BlockView;
template <>
struct BlockView<M2L_STORE_F64> {
  const double2* p;
  __device__ __forceinline__ BlockView(const M2LBlocks& b, const PairDesc& d)
      : p(reinterpret_cast<const double2*>(b.f64 + d.off)) {}
  __device__ __forceinline__ double2 operator[](int i) const { return p[i]; }
  __device__ __forceinline__ double scale(const PairDesc&) const { return 1.0; }
};
template <>
struct BlockView<M2L_STORE_B48> {
  const B48Pair* p;
  __device__ __forceinline__ BlockView(const M2LBlocks& b, const PairDesc& d)
      : p(b.b48 + (d.off >> 1)) {}
  __device__ __forceinline__ double2 operator[](int i) const {
    const B48Pair a = p[i];
    return make_double2(
        __hiloint2double((int)(kB48ExpWord | (a.hi & 0xFFFFu)), (int)a.lo0) - kB48Bias,
        __hiloint2double((int)(kB48ExpWord | (a.hi >> 16)), (int)a.lo1) - kB48Bias);
  }
  __device__ __forceinline__ double scale(const PairDesc& d) const { return ldexp(1.0, d.sexp); }
};

// M2L pair kernel: one wave per unordered pair, no LDS.  The block (ld x rb,
// column-major, ld even) is read in tiles of 16 rows x 8 columns, one
// 16-byte load per lane: lane 8 ci + rp holds rows 2 rp, 2 rp + 1 of column
// ci of the tile.  All of a pair's loads are issued before any use (tile
// counts are compile-time: ld <= 16 MAXRT, rb <= 8 MAXCT).  B out_b: each
// lane accumulates its two rows over its columns, then the 8 lanes sharing
// rp are summed; B^T out_k: per column tile, the 8 lanes of a column are
// summed.  Writes B out_b and B^T out_k to the pair's partials (part_k,
// part_b).
template <int MAXRT, int MAXCT, int ST>
__device__ __forceinline__ void m2l_pair(const M2LBlocks& m2l, const PairDesc& d,
                                         const double* __restrict__ outv,
                                         double* __restrict__ part, double tsign) {
  const int lane = threadIdx.x, rp = lane & 7, ci = lane >> 3;
  const int rk = d.rk, rb = d.rb, ld = d.ld;
  double* fw = part + d.part_k;
  double* tw = part + d.part_b;
  const double* xk = outv + d.sk;
  const double* xb = outv + d.sb;
  const BlockView<ST> B2(m2l, d);
  const double sc = B2.scale(d), tsc = tsign * sc;
  // staging registers as double arrays (arrays of double2 stay in scratch)
  double ax[MAXCT][MAXRT], ay[MAXCT][MAXRT], xbv[MAXCT], k0[MAXRT], k1[MAXRT];
#pragma unroll
  for (int u = 0; u < MAXCT; u++) {
    const int c = 8 * u + ci;
    const bool vc = c < rb;
    xbv[u] = xb[vc ? c : 0];
#pragma unroll
    for (int t = 0; t < MAXRT; t++) {
      const int r = 16 * t + 2 * rp;
      const bool v = vc && r < ld;
      const double2 a = B2[v ? ((c * ld + r) >> 1) : 0];  // clamped: always issued
      ax[u][t] = v ? a.x : 0.0;
      ay[u][t] = v ? a.y : 0.0;
    }
    if (!vc) xbv[u] = 0.0;
  }
#pragma unroll
  for (int t = 0; t < MAXRT; t++) {
    const int r = 16 * t + 2 * rp;
    k0[t] = xk[r < rk ? r : 0];
    k1[t] = xk[r + 1 < rk ? r + 1 : 0];
    if (!(r < rk)) k0[t] = 0.0;
    if (!(r + 1 < rk)) k1[t] = 0.0;
  }
#pragma unroll
  for (int u = 0; u < MAXCT; u++) {
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < MAXRT; t++) s += ax[u][t] * k0[t] + ay[u][t] * k1[t];
    s = sum_lanes8(s);
    const int c = 8 * u + ci;
    if (rp == 0 && c < rb) tw[c] = tsc * s;
  }
#pragma unroll
  for (int t = 0; t < MAXRT; t++) {
    double f0 = 0.0, f1 = 0.0;
#pragma unroll
    for (int u = 0; u < MAXCT; u++) {
      f0 += ax[u][t] * xbv[u];
      f1 += ay[u][t] * xbv[u];
    }
    f0 = sum_lanes_stride8(f0);
    f1 = sum_lanes_stride8(f1);
    const int r = 16 * t + 2 * rp;
    if (ci == 0) {
      if (r < rk) fw[r] = sc * f0;
      if (r + 1 < rk) fw[r + 1] = sc * f1;
    }
  }
}

// M2L of pairs stored as ACA factors B = U V^T (U: rk x p, V: rb x p,
// column-major at d.off), one wave per pair, four pairs per workgroup.
// Stage 1: t1 = V^T out_b and t2 = U^T out_k, four lanes per column (each
// summing every fourth row, then a two-step shuffle sum), 16 columns per
// pass.  Stage 2: B out_b = U t1 and B^T out_k = V t2, one lane per row,
// written where the dense pair kernels write them.  p <= 256.
constexpr int kFactLanes = 4;
__global__ __launch_bounds__(256) void k_m2l_fact(const double* __restrict__ F,
                                                  const PairDesc* __restrict__ pd,
                                                  const double* __restrict__ outv,
                                                  double* __restrict__ part, int npairs,
                                                  double tsign) {
  __shared__ double tsh[4][2][256];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + w;
  if (q >= npairs) return;  // whole waves: no barrier below spans waves
  const PairDesc d = pd[q];
  const int rk = d.rk, rb = d.rb, p = d.ld;
  const double* U = F + d.off;
  const double* V = U + (int64_t)rk * p;
  const double* xk = outv + d.sk;
  const double* xb = outv + d.sb;
  double* t1 = tsh[w][0];
  double* t2 = tsh[w][1];
  const int gi = l % kFactLanes, col = l / kFactLanes;
  for (int c0 = 0; c0 < p; c0 += 64 / kFactLanes) {
    const int c = c0 + col;
    double s1 = 0.0, s2 = 0.0;
    if (c < p) {
      const double* Vc = V + (int64_t)c * rb;
      const double* Uc = U + (int64_t)c * rk;
      for (int i = gi; i < rb; i += kFactLanes) s1 = fma(Vc[i], xb[i], s1);
      for (int i = gi; i < rk; i += kFactLanes) s2 = fma(Uc[i], xk[i], s2);
    }
#pragma unroll
    for (int o = kFactLanes / 2; o > 0; o >>= 1) {
      s1 += __shfl_xor(s1, o, 64);
      s2 += __shfl_xor(s2, o, 64);
    }
    if (gi == 0 && c < p) {
      t1[c] = s1;
      t2[c] = s2;
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int i = l; i < rk; i += 64) {
    double acc = 0.0;
    for (int c = 0; c < p; c++) acc = fma(U[i + (int64_t)c * rk], t1[c], acc);
    part[d.part_k + i] = acc;
  }
  for (int i = l; i < rb; i += 64) {
    double acc = 0.0;
    for (int c = 0; c < p; c++) acc = fma(V[i + (int64_t)c * rb], t2[c], acc);
    part[d.part_b + i] = tsign * acc;
  }
}

template <int MAXRT, int MAXCT, int ST>
__global__ __launch_bounds__(64) void k_m2l_reg(const double* __restrict__ m2l,
                                                const B48Pair* __restrict__ m2l_b48,
                                                const PairDesc* __restrict__ pd,
                                                const double* __restrict__ outv,
                                                double* __restrict__ part, int npairs,
                                                double tsign) {
  if (AFMM_MEM_PRIO) __builtin_amdgcn_s_setprio(AFMM_MEM_PRIO);
  const int p = blockIdx.x;
  if (p >= npairs) return;
  m2l_pair<MAXRT, MAXCT, ST>(M2LBlocks{m2l, m2l_b48}, pd[p], outv, part, tsign);
}

// M2L pair kernel for blocks taller or wider than the register tiles: one
// wave per pair, the same lane layout (lane 8 ci + rp holds rows 2 rp,
// 2 rp + 1 of each 16-row tile of column ci of an 8-column tile), all RT row
// tiles of a chunk of CT column tiles per step.  The block streams through
// two register buffers: the next chunk's loads are issued before the current
// chunk is used.  Per chunk, B^T out_k of its columns is summed over the 8
// lanes of each column and written; B out_b accumulates per lane (two rows
// per row tile) and is summed over the column lanes at the end.
template <int RT, int CT, class View>
__device__ __forceinline__ void m2l_load_chunk(const View& B2, int ld, int rb,
                                               int c0, int ci, const int (&rr)[RT],
                                               double (&ax)[CT][RT], double (&ay)[CT][RT],
                                               double (&xv)[CT], const double* __restrict__ xb) {
#pragma unroll
  for (int u = 0; u < CT; u++) {
    const int c = c0 + 8 * u + ci;
    const int cc = min(c, rb - 1);
    xv[u] = xb[cc];
    if (c >= rb) xv[u] = 0.0;
    const int col = (cc * ld) >> 1;
#pragma unroll
    for (int t = 0; t < RT; t++) {
      const double2 a = B2[col + rr[t]];
      ax[u][t] = a.x;
      ay[u][t] = a.y;
    }
  }
}

template <int RT, int CT>
__device__ __forceinline__ void m2l_use_chunk(const double (&ax)[CT][RT], const double (&ay)[CT][RT],
                                              const double (&xv)[CT], const double (&k0)[RT],
                                              const double (&k1)[RT], double (&f0)[RT],
                                              double (&f1)[RT], int c0, int ci, int rp, int rb,
                                              double* __restrict__ tw, double tsign) {
#pragma unroll
  for (int u = 0; u < CT; u++) {
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < RT; t++) s += ax[u][t] * k0[t] + ay[u][t] * k1[t];
    s = sum_lanes8(s);
    const int c = c0 + 8 * u + ci;
    if (rp == 0 && c < rb) tw[c] = tsign * s;
#pragma unroll
    for (int t = 0; t < RT; t++) {
      f0[t] = fma(ax[u][t], xv[u], f0[t]);
      f1[t] = fma(ay[u][t], xv[u], f1[t]);
    }
  }
}

template <int RT, int CT, int ST>
__global__ __launch_bounds__(64) void k_m2l_stream(const double* __restrict__ m2l,
                                                   const B48Pair* __restrict__ m2l_b48,
                                                   const PairDesc* __restrict__ pd,
                                                   const double* __restrict__ outv,
                                                   double* __restrict__ part, int npairs,
                                                   double tsign) {
  if (AFMM_MEM_PRIO) __builtin_amdgcn_s_setprio(AFMM_MEM_PRIO);
  const int p = blockIdx.x;
  if (p >= npairs) return;
  const PairDesc d = pd[p];
  const int lane = threadIdx.x, rp = lane & 7, ci = lane >> 3;
  const int rk = d.rk, rb = d.rb, ld = d.ld;
  double* fw = part + d.part_k;
  double* tw = part + d.part_b;
  const double* xk = outv + d.sk;
  const double* xb = outv + d.sb;
  const BlockView<ST> B2(M2LBlocks{m2l, m2l_b48}, d);
  const double sc = B2.scale(d), tsc = tsign * sc;
  if (ld == 0 || rb == 0) {
    // an empty side: both products are zero
    for (int i = lane; i < rk; i += 64) fw[i] = 0.0;
    for (int i = lane; i < rb; i += 64) tw[i] = 0.0;
    return;
  }
  // Addresses are clamped into the block instead of masking the loads: a
  // row past ld meets k0 = k1 = 0 in B^T out_k and its B out_b sum is never
  // stored; a column past rb meets xv = 0 in B out_b and its B^T out_k sum
  // is never stored.  rr[t]: the double2 index of rows 16 t + 2 rp, + 1
  // within a column, clamped to the last row pair.
  constexpr int W = 8 * CT;  // columns per chunk
  int rr[RT];
#pragma unroll
  for (int t = 0; t < RT; t++) rr[t] = min(16 * t + 2 * rp, ld - 2) >> 1;
  double ax[CT][RT], ay[CT][RT], xv[CT], bx[CT][RT], by[CT][RT], bv[CT];
  m2l_load_chunk<RT, CT>(B2, ld, rb, 0, ci, rr, ax, ay, xv, xb);
  double k0[RT], k1[RT], f0[RT], f1[RT];
#pragma unroll
  for (int t = 0; t < RT; t++) {
    const int r = 16 * t + 2 * rp;
    k0[t] = xk[min(r, rk - 1)];
    k1[t] = xk[min(r + 1, rk - 1)];
    if (!(r < rk)) k0[t] = 0.0;
    if (!(r + 1 < rk)) k1[t] = 0.0;
    f0[t] = 0.0;
    f1[t] = 0.0;
  }
  for (int c0 = 0; c0 < rb; c0 += 2 * W) {
    if (c0 + W < rb) m2l_load_chunk<RT, CT>(B2, ld, rb, c0 + W, ci, rr, bx, by, bv, xb);
    m2l_use_chunk<RT, CT>(ax, ay, xv, k0, k1, f0, f1, c0, ci, rp, rb, tw, tsc);
    if (c0 + W >= rb) break;
    if (c0 + 2 * W < rb) m2l_load_chunk<RT, CT>(B2, ld, rb, c0 + 2 * W, ci, rr, ax, ay, xv, xb);
    m2l_use_chunk<RT, CT>(bx, by, bv, k0, k1, f0, f1, c0 + W, ci, rp, rb, tw, tsc);
  }
#pragma unroll
  for (int t = 0; t < RT; t++) {
    const double g0 = sum_lanes_stride8(f0[t]);
    const double g1 = sum_lanes_stride8(f1[t]);
    const int r = 16 * t + 2 * rp;
    if (ci == 0) {
      if (r < rk) fw[r] = sc * g0;
      if (r + 1 < rk) fw[r + 1] = sc * g1;
    }
  }
}

// L2L for a non-leaf level, with the box's M2L sum fused in: in_k =
// parent contribution + M2L partials (local_coeffs), then the children's
// parent contribution = L2P_k in_k: stored row t (column-major stored rows,
// i.e. row-major rank x nc) goes to child coefficient rowmap[t], skeleton row
// q (the identity) to rowmap[nc + q].  The children's coefficients are
// contiguous.
template <int WPB, int B>
__global__ __launch_bounds__(64 * WPB) void k_down(const double* __restrict__ l2p_cm,
                                                   const int64_t* __restrict__ l2poff,
                                                   const int* __restrict__ rank,
                                                   const int* __restrict__ nin,
                                                   const int64_t* __restrict__ soff,
                                                   const int64_t* __restrict__ child_soff,
                                                   const int* __restrict__ rowmap,
                                                   const double* __restrict__ part,
                                                   const int* __restrict__ slot_off,
                                                   double* __restrict__ inv, int has_parent,
                                                   int64_t gbase, int nb) {
  extern __shared__ double xs[];
  __shared__ double red[64 * WPB];
  if (AFMM_MEM_PRIO) __builtin_amdgcn_s_setprio(AFMM_MEM_PRIO);
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  const int n = nin[g];
  if (n == 0) return;
  const int nc = n - r;
  const double* P = l2p_cm + l2poff[g];
  const int64_t co = child_soff[4 * k];
  double* dst = inv + co;
  const int* rm = rowmap + co;
  const int w = threadIdx.x >> 6;
  double pv[B];
  const int nch = (nc + 63) >> 6;
  if (nch <= 1 || nch > WPB) {
    GemvLane q;
    if (nc > 0) {
      q = gemv_lane(nc, 0, r);
      gemv_load<B>(pv, P, r, nc, 0, q, w, WPB);
    }
    local_coeffs<WPB>(xs, r, inv + soff[g], has_parent, part, slot_off + (int64_t)kILSlots * g);
    __syncthreads();
    for (int c = threadIdx.x; c < r; c += 64 * WPB) dst[rm[nc + c]] = xs[c];
    for (int cb = 0; cb < nc; cb += 64) {
      if (cb > 0) {
        q = gemv_lane(nc, cb, r);
        gemv_load<B>(pv, P, r, nc, cb, q, w, WPB);
      }
      const double acc = gemv_chunk<B>(pv, P, r, nc, cb, q, xs, w, WPB);
      const double s = reduce_groups<WPB>(acc, red, q);
      if ((int)threadIdx.x < q.rc) dst[rm[cb + threadIdx.x]] = s;
    }
    return;
  }
  // the column chunks in parallel: wave w takes chunk w % nch and every
  // nsub-th step of it (the waves w' = w mod nch); one reduction at the end
  const int ch = w % nch, sub = w / nch, nsub = (WPB - 1 - ch) / nch + 1;
  const GemvLane q = gemv_lane(nc, 64 * ch, r);
  gemv_load<B>(pv, P, r, nc, 64 * ch, q, sub, nsub);
  local_coeffs<WPB>(xs, r, inv + soff[g], has_parent, part, slot_off + (int64_t)kILSlots * g);
  __syncthreads();
  for (int c = threadIdx.x; c < r; c += 64 * WPB) dst[rm[nc + c]] = xs[c];
  red[threadIdx.x] = gemv_chunk<B>(pv, P, r, nc, 64 * ch, q, xs, sub, nsub);
  __syncthreads();
  const int i = threadIdx.x;
  if (i < nc) {
    const int ci = i >> 6, c = i & 63;
    const int rc = min(64, nc - 64 * ci), G = 64 / rc, ns = (WPB - 1 - ci) / nch + 1;
    double acc = 0.0;
    for (int u = 0; u < ns; u++)
      for (int h = 0; h < G; h++) acc += red[(ci + nch * u) * 64 + h * rc + c];
    dst[rm[i]] = acc;
  }
}

// Upper-tree passes fused over subtrees: one workgroup of 16 waves per box of
// the chunk's top level jt covers that box's descendants on D <= 3 levels
// (jt .. jt + D - 1; at depth d there are 4^d of them, quad-tree ids
// [4^d k, 4^d (k + 1))).  Each level's boxes split the 16 waves evenly
// (16 / 4^d waves per box), and the levels are separated by workgroup
// barriers only, so a chunk costs one launch instead of D.  Every wave runs
// the same sequence of barriers whatever its box's shape.
struct TreeArgs {
  const double* l2p;      // stored L2P rows: row-major (up) / column-major (down)
  const int64_t* l2poff;
  const int* rank;
  const int* nin;
  const int64_t* soff;
  const int* rowmap_up;   // rowmap of the upper levels (indexed like skx)
  double* vec;            // outv (up) / inv (down)
  const double* part;     // M2L partials (down)
  const int* slot_off;    // partial offsets per box (down)
  int64_t gb[4];          // gbase of levels jt .. jt + 3
  int D;                  // levels in the chunk
  int top_has_parent;     // down: the top level has an L2L contribution
  int xs_stride;          // doubles of LDS per box
};

constexpr int kTreeWaves = 16;

// sum_{k in [k0, k1)} P[k stride + ic] xs[k]: loads in batches of 16
// through two register sets, the next batch issued before the current one is
// used; addresses past k1 are clamped to k1 - 1 and their terms dropped.
// The first batch is already issued into pa.
__device__ __forceinline__ void tree_load(double (&pv)[16], const double* __restrict__ P,
                                          int stride, int ic, int k, int k1) {
#pragma unroll
  for (int q = 0; q < 16; q++) pv[q] = P[(unsigned)(min(k + q, k1 - 1) * stride + ic)];
}
__device__ __forceinline__ double tree_use(const double (&pv)[16], double acc, int k, int k1,
                                           const double* xs) {
#pragma unroll
  for (int q = 0; q < 16; q++) acc = k + q < k1 ? fma(pv[q], xs[k + q], acc) : acc;
  return acc;
}
__device__ __forceinline__ double tree_dot_pre(double (&pa)[16], const double* __restrict__ P,
                                               int stride, int ic, int k0, int k1,
                                               const double* xs) {
  double acc = 0.0;
  if (k0 >= k1) return acc;
  double pb[16];
  for (int k = k0;; k += 32) {
    if (k + 16 < k1) tree_load(pb, P, stride, ic, k + 16, k1);
    acc = tree_use(pa, acc, k, k1, xs);
    if (k + 16 >= k1) break;
    if (k + 32 < k1) tree_load(pa, P, stride, ic, k + 32, k1);
    acc = tree_use(pb, acc, k + 16, k1, xs);
    if (k + 32 >= k1) break;
  }
  return acc;
}
// M2M over a chunk, bottom level first: out_k[c] = sum_t P[t r + c]
// x[rowmap[t]] + x[rowmap[nc + c]], x the children's multipoles (staged in
// LDS in rowmap order).  A box's waves split its column chunks (64 columns)
// and, when they outnumber them, its rows; the row parts are summed in order.
__global__ __launch_bounds__(64 * kTreeWaves) void k_up_tree(TreeArgs a) {
  extern __shared__ double lds[];
  __shared__ double red[kTreeWaves][64];
  if (AFMM_MEM_PRIO) __builtin_amdgcn_s_setprio(AFMM_MEM_PRIO);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  for (int d = a.D - 1; d >= 0; d--) {
    const int m = 1 << (2 * d), WB = kTreeWaves / m;
    const int ub = w / WB, u = w - ub * WB;
    const int64_t k = (int64_t)blockIdx.x * m + ub;
    const int64_t g = a.gb[d] + k;
    const int r = a.rank[g], n = a.nin[g], nc = n - r;
    const int64_t co = a.soff[a.gb[d + 1] + 4 * k];
    double* xs = lds + (int64_t)ub * a.xs_stride;
    {
      const double* x = a.vec + co;
      const int* rm = a.rowmap_up + co;
      for (int t = u * 64 + l; t < n; t += WB * 64) xs[t] = x[rm[t]];
    }
    __syncthreads();
    const double* P = a.l2p + a.l2poff[g];
    double* out = a.vec + a.soff[g];
    const int ncc = (r + 63) >> 6;
    const int np = (ncc > 0 && ncc < WB) ? WB / ncc : 1;
    // np == 1: column chunks u, u + WB, ... over all rows; np > 1: the one
    // part (chunk u mod ncc, rows of part u / ncc) into the reduction rows
    int cc0 = u, t0 = 0, t1 = nc;
    if (np > 1) {
      const int p = u / ncc;
      cc0 = u - p * ncc;
      t0 = p < np ? p * nc / np : 0;
      t1 = p < np ? (p + 1) * nc / np : 0;
    }
    for (int cc = cc0; cc < ncc; cc += WB) {
      const int ic = min(cc * 64 + l, r - 1);
      double pa[16];
      if (t0 < t1) tree_load(pa, P, r, ic, t0, t1);
      const double acc = tree_dot_pre(pa, P, r, ic, t0, t1, xs);
      if (np > 1) {
        red[w][l] = acc;
        break;
      }
      const int c = cc * 64 + l;
      if (c < r) out[c] = acc + xs[nc + c];
    }
    __syncthreads();
    if (np > 1 && u < ncc) {
      const int c = u * 64 + l;
      double acc = 0.0;
      for (int p = 0; p < np; p++) acc += red[ub * WB + p * ncc + u][l];
      if (c < r) out[c] = acc + xs[nc + c];
    }
    __syncthreads();
  }
}

// L2L over a chunk, top level first: in_k = parent contribution + M2L
// partials (slot order), then the children's contributions: stored row t
// (column-major stored rows, P[c nc + t]) to child coefficient rowmap[t],
// skeleton row q to rowmap[nc + q].  A box's waves split its 64-row output
// chunks and, when they outnumber them, the sum over c.
__global__ __launch_bounds__(64 * kTreeWaves) void k_down_tree(TreeArgs a) {
  extern __shared__ double lds[];
  __shared__ double red[kTreeWaves][64];
  if (AFMM_MEM_PRIO) __builtin_amdgcn_s_setprio(AFMM_MEM_PRIO);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  for (int d = 0; d < a.D; d++) {
    const int m = 1 << (2 * d), WB = kTreeWaves / m;
    const int ub = w / WB, u = w - ub * WB;
    const int64_t k = (int64_t)blockIdx.x * m + ub;
    const int64_t g = a.gb[d] + k;
    const int r = a.rank[g], n = a.nin[g], nc = n - r;
    double* xs = lds + (int64_t)ub * a.xs_stride;
    {
      const bool hp = d > 0 || a.top_has_parent;
      const double* par = a.vec + a.soff[g];
      const int* so = a.slot_off + (int64_t)kILSlots * g;
      int o[kILSlots];
#pragma unroll
      for (int t = 0; t < kILSlots; t++) o[t] = so[t];
      for (int c = u * 64 + l; c < r; c += WB * 64) {
        double v[kILSlots];
#pragma unroll
        for (int t = 0; t < kILSlots; t++) v[t] = o[t] >= 0 ? a.part[o[t] + c] : 0.0;
        double acc = hp ? par[c] : 0.0;
#pragma unroll
        for (int t = 0; t < kILSlots; t++) acc += v[t];
        xs[c] = acc;
      }
    }
    __syncthreads();
    const double* P = a.l2p + a.l2poff[g];
    const int64_t co = a.soff[a.gb[d + 1] + 4 * k];
    double* dst = a.vec + co;
    const int* rm = a.rowmap_up + co;
    const int ntc = (nc + 63) >> 6;
    const int np = (ntc > 0 && ntc < WB) ? WB / ntc : 1;
    // np == 1: row chunks u, u + WB, ... over all c; np > 1: the one part
    // (chunk u mod ntc, the c range of part u / ntc) into the reduction rows
    int tc0 = u, c0 = 0, c1 = r;
    if (np > 1) {
      const int p = u / ntc;
      tc0 = u - p * ntc;
      c0 = p < np ? p * r / np : 0;
      c1 = p < np ? (p + 1) * r / np : 0;
    }
    for (int tc = tc0; tc < ntc; tc += WB) {
      const int it = min(tc * 64 + l, nc - 1);
      double pa[16];
      if (c0 < c1) tree_load(pa, P, nc, it, c0, c1);
      const double acc = tree_dot_pre(pa, P, nc, it, c0, c1, xs);
      if (np > 1) {
        red[w][l] = acc;
        break;
      }
      const int t = tc * 64 + l;
      if (t < nc) dst[rm[t]] = acc;
    }
    for (int c = u * 64 + l; c < r; c += WB * 64) dst[rm[nc + c]] = xs[c];
    __syncthreads();
    if (np > 1 && u < ntc) {
      const int t = u * 64 + l;
      double acc = 0.0;
      for (int p = 0; p < np; p++) acc += red[ub * WB + p * ntc + u][l];
      if (t < nc) dst[rm[t]] = acc;
    }
    __syncthreads();
  }
}

constexpr int kNearSlices = 8;          // work-queue slices of the near field
constexpr int kNearCounterStride = 32;  // ints between the slice counters (128 B)
constexpr int kNearChunk = 4;           // pairs per counter increment

// Near field, symmetric: one unordered leaf pair (A, B), A <= B, per
// workgroup of nw waves at a time.  B is taken in chunks of 64 points
// staged in LDS (x, y, charge: 1.5 KiB); A's chunks stay in registers
// (compile-time count NCH, so the arrays are statically indexed), wave w
// taking the chunks ca = w, w + nw, ...  For a B chunk, lane t of an A chunk
// meets source (t + s) mod 64 at step s, so the 64 lanes touch 64 distinct
// sources; each kernel value is evaluated once and used for both
// directions: the A-side sum in a register, the B-side sum in LDS (one FP64
// LDS add per step) in a row of its own per A chunk, so every row has one
// writer and the rows are added in chunk order: the sums do not depend on
// how the chunks were spread over waves.  A B chunk's sums are complete
// after its A chunks and written out then; a self pair (A, A) visits the
// chunk pairs ca <= cb, the diagonal ones by steps 1..32 (step 32 on lanes
// 0..31 only) and folds chunk cb's B-side sums into that chunk's A-side
// sums.
template <int KID, int NCH>
__global__ __launch_bounds__(256) void k_p2p(const double* __restrict__ x,
                                             const double* __restrict__ y,
                                             const double* __restrict__ qs,
                                             const int* __restrict__ leaf_start,
                                             const int* __restrict__ leaf_count,
                                             const int* __restrict__ pairs,
                                             const int* __restrict__ pair_part,
                                             double* __restrict__ part, int npairs, KParams kp,
                                             const double2* __restrict__ ntab,
                                             int* __restrict__ queue) {
  __shared__ double2 tab[kNearTabLen<KID>];
  // one address per step for each array, 8-byte lane stride: no bank
  // conflicts.  The staged chunk is stored twice (entries t and t + 64), so
  // source t + s of lane t needs no wrap: each lane's addresses are its base
  // plus an immediate offset.  The B-side sums of an A chunk collect in a
  // row of bc: in entry t for source t when they travel in registers, in
  // entries j and j + 64 (LDS adds) on the diagonal chunk of a self pair;
  // the rows are added when the chunk is done.
  __shared__ double bx[128], by[128], bq[128], bc[NCH][128];
  __shared__ int grab;
  const int t = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  load_near_tab<KID>(tab, ntab, threadIdx.x, blockDim.x);
  // padding: zero charge, far away for the built-in kernels (K stays finite
  // there); the user kernel is only known to be finite at real points, so
  // its padding repeats the leaf's first point
  const double ninv_p0 = KID == K_GAUSS ? -1.0 / kp.p0 : 0.0;
  const double inv_a = KID == K_REG_INV_R ? 1.0 / kp.a : 0.0;
  const double scale = near_scale<KID>;
  // work queue: the pairs are cut into kNearSlices slices, each with its own
  // chunk counter (one cache line apart); a workgroup drains the slice of
  // its own index first and then steals from the others, so a workgroup that
  // becomes resident late (its slots held by the kernels beside it) only
  // finds less work
  int slice = blockIdx.x % kNearSlices, tried = 0, p = 0, p_end = 0;
  for (;;) {
    if (p >= p_end) {
      __syncthreads();
      if (threadIdx.x == 0) grab = atomicAdd(&queue[slice * kNearCounterStride], 1);
      __syncthreads();
      const int c = grab;
      const int s0 = (int)((int64_t)npairs * slice / kNearSlices);
      const int s1 = (int)((int64_t)npairs * (slice + 1) / kNearSlices);
      p = s0 + c * kNearChunk;
      if (p >= s1) {
        if (++tried == kNearSlices) break;
        slice = (slice + 1) % kNearSlices;
        p_end = 0;
        continue;
      }
      p_end = min(p + kNearChunk, s1);
    }
    const int A = pairs[2 * p], B = pairs[2 * p + 1];
    const bool self = A == B;
    const int sa = leaf_start[A], na = leaf_count[A];
    const int sb = leaf_start[B], nb = leaf_count[B];
    const int nca = (na + 63) >> 6, ncb = (nb + 63) >> 6;
    double* out = part + pair_part[2 * p];       // A's sums
    double* outb = part + pair_part[2 * p + 1];  // B's sums
    const double pax = KID >= K_USER ? x[sa] : 1.0e30, pay = KID >= K_USER ? y[sa] : 1.0e30;
    const double pbx = KID >= K_USER ? x[sb] : 1.0e30, pby = KID >= K_USER ? y[sb] : 1.0e30;
    double xa[NCH], ya[NCH], qa[NCH], row[NCH];
#pragma unroll
    for (int ca = 0; ca < NCH; ca++) {
      const int i = 64 * ca + t;
      const bool v = ca < nca && ca % nw == w && i < na;
      xa[ca] = v ? x[sa + i] : pax;
      ya[ca] = v ? y[sa + i] : pay;
      qa[ca] = v ? qs[sa + i] : 0.0;
      row[ca] = 0.0;
    }
    for (int cb = 0; cb < ncb; cb++) {
      __syncthreads();  // the previous chunk's sums have been read
      if (w == 0) {
        const int j = 64 * cb + t;
        const bool v = j < nb;
        const double vx = v ? x[sb + j] : pbx, vy = v ? y[sb + j] : pby;
        const double vq = v ? qs[sb + j] : 0.0;
        bx[t] = vx;
        bx[t + 64] = vx;
        by[t] = vy;
        by[t + 64] = vy;
        bq[t] = vq;
        bq[t + 64] = vq;
      }
#pragma unroll
      for (int ca = 0; ca < NCH; ca++)
        if (ca % nw == w) {
          bc[ca][t] = 0.0;
          bc[ca][t + 64] = 0.0;
        }
      __syncthreads();
      // all A chunks of a full leaf four at a time against one staged
      // source: four independent evaluation chains per step and a quarter of
      // the LDS reads per evaluation
      const bool quad = NCH % 4 == 0 && nw == 1 && !self && nca == NCH;
      if (quad) {
#pragma unroll
        for (int cq = 0; cq < NCH; cq += 4) {
          double xq[4], yq[4], qq[4], aq[4], bs[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            xq[u] = xa[cq + u];
            yq[u] = ya[cq + u];
            qq[u] = qa[cq + u];
            aq[u] = row[cq + u];
            bs[u] = 0.0;
          }
#pragma unroll 4
          for (int s = 0; s < 64; s++) {
            const int j = t + s;
            const double sx = bx[j], sy = by[j], sq = bq[j];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const double v = pair_value<KID>(xq[u], yq[u], sx, sy, kp, ninv_p0, inv_a, tab);
              aq[u] = fma(v, sq, aq[u]);
              bs[u] = dpp_move<kDppWaveRol1>(fma(kTransposeSign<KID> * v, qq[u], bs[u]));
            }
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            bc[cq + u][t] = bs[u];
            row[cq + u] = aq[u];
          }
        }
      }
#pragma unroll
      for (int cp = 0; cp < NCH; cp += 2) {
        if (quad) break;
        if (cp + 1 < NCH && nw == 1 && !self && cp + 1 < nca) {
          // two A chunks per step against one staged source: half the LDS
          // reads per evaluation; each row gets the same adds in the same
          // order as one chunk at a time
          const double x0 = xa[cp], y0 = ya[cp], q0 = qa[cp];
          const double x1 = xa[cp + 1], y1 = ya[cp + 1], q1 = qa[cp + 1];
          double acc0 = row[cp], acc1 = row[cp + 1];
          // B-side sums in registers that travel with the sources: at step
          // s lane t holds source t + s's sum, adds its term and passes it
          // to lane t - 1 (wave_rol:1), so after 64 steps lane t holds
          // source t's
          double bs0 = 0.0, bs1 = 0.0;
#pragma unroll 8
          for (int s = 0; s < 64; s++) {
            const int j = t + s;
            const double sx = bx[j], sy = by[j], sq = bq[j];
            const double v0 = pair_value<KID>(x0, y0, sx, sy, kp, ninv_p0, inv_a, tab);
            const double v1 = pair_value<KID>(x1, y1, sx, sy, kp, ninv_p0, inv_a, tab);
            acc0 = fma(v0, sq, acc0);
            acc1 = fma(v1, sq, acc1);
            bs0 = dpp_move<kDppWaveRol1>(fma(kTransposeSign<KID> * v0, q0, bs0));
            bs1 = dpp_move<kDppWaveRol1>(fma(kTransposeSign<KID> * v1, q1, bs1));
          }
          bc[cp][t] = bs0;  // entries t + 64 of both rows stay zero
          bc[cp + 1][t] = bs1;
          row[cp] = acc0;
          row[cp + 1] = acc1;
          continue;
        }
#pragma unroll
        for (int ca = cp; ca < cp + 2 && ca < NCH; ca++) {
          if (ca >= nca || ca % nw != w || (self && ca > cb)) continue;
          const bool diag = self && ca == cb;
          const double xi = xa[ca], yi = ya[ca], qi = qa[ca];
          double acc = row[ca];
          int s = 0;
          if (diag) {
            // step 0: the point with itself (R2 = 0)
            acc += (KID >= K_USER ? kernel_xy<KID>(xi, yi, xi, yi, kp)
                                  : near_kernel<KID>(0.0, kp, ninv_p0, inv_a, tab)) *
                   qi;
            s = 1;
          }
          if (!diag) {
            // the B-side sums travel with the sources (as above)
            double bs = 0.0;
#pragma unroll 2
            for (; s < 64; s++) {
              const int j = t + s;
              const double v = pair_value<KID>(xi, yi, bx[j], by[j], kp, ninv_p0, inv_a, tab);
              acc = fma(v, bq[j], acc);
              bs = dpp_move<kDppWaveRol1>(fma(kTransposeSign<KID> * v, qi, bs));
            }
            bc[ca][t] = bs;
          }
          const int s_end = diag ? 32 : 0;
#pragma unroll 2
          for (; s < s_end; s++) {
            const int j = t + s;
            const double v = pair_value<KID>(xi, yi, bx[j], by[j], kp, ninv_p0, inv_a, tab);
            acc = fma(v, bq[j], acc);
            atomicAdd(&bc[ca][j], kTransposeSign<KID> * v * qi);
          }
          if (diag && t < 32) {
            const int j = t + 32;
            const double v = pair_value<KID>(xi, yi, bx[j], by[j], kp, ninv_p0, inv_a, tab);
            acc = fma(v, bq[j], acc);
            atomicAdd(&bc[ca][j], kTransposeSign<KID> * v * qi);
          }
          row[ca] = acc;
        }
      }
      __syncthreads();
      // chunk cb's B-side sums are complete: the rows of A's chunks, in order
      double colsum = 0.0;
#pragma unroll
      for (int ca = 0; ca < NCH; ca++)
        if (ca < nca) colsum += bc[ca][t] + bc[ca][t + 64];
      if (self) {
#pragma unroll
        for (int ca = 0; ca < NCH; ca++)
          if (ca == cb && ca % nw == w) row[ca] += colsum;
      } else if (w == 0) {
        const int j = 64 * cb + t;
        if (j < nb) outb[j] = scale * colsum;
      }
    }
#pragma unroll
    for (int ca = 0; ca < NCH; ca++) {
      const int i = 64 * ca + t;
      if (ca < nca && ca % nw == w && i < na) out[i] = scale * row[ca];
    }
    p++;
  }
}

// Near-field term alone: the 9 partial slots of every leaf point summed and
// scattered to the caller's order.  One workgroup per leaf box.
__global__ __launch_bounds__(256) void k_near_out(const double* __restrict__ near_part,
                                                  const int* __restrict__ leaf_start,
                                                  const int* __restrict__ leaf_count,
                                                  const int* __restrict__ perm,
                                                  double* __restrict__ yout, int nleaves) {
  const int k = blockIdx.x;
  if (k >= nleaves) return;
  const int st = leaf_start[k], n = leaf_count[k];
  const double* nblk = near_part + 9 * (int64_t)st;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < 9; t++) s += nblk[t * n + i];
    yout[perm[st + i]] = s;
  }
}

__global__ void k_add(double* __restrict__ y, const double* __restrict__ t, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += t[i];
}

__global__ void k_zero_ints(int* __restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0;
}

// Leaf level, fused: in_k (parent contribution + M2L partials), potential
// = L2P_k in_k + the leaf's 9 near-field partial slots (self, then
// neighbours 0..7; absent ones are zero), scattered to the caller's order.
// Stored L2P row t (column-major stored rows) is point rowmap[t]; skeleton
// point rowmap[nc + q] takes in_k[q] (the identity row).
template <int WPB, int B>
__global__ __launch_bounds__(64 * WPB) void k_leaf(const double* __restrict__ l2p_cm,
                                                   const int64_t* __restrict__ l2poff,
                                                   const int* __restrict__ rank,
                                                   const int64_t* __restrict__ soff,
                                                   const int* __restrict__ leaf_start,
                                                   const int* __restrict__ leaf_count,
                                                   const int* __restrict__ rowmap,
                                                   const double* __restrict__ part,
                                                   const int* __restrict__ slot_off,
                                                   const double* __restrict__ inv, int has_parent,
                                                   const double* __restrict__ near_part,
                                                   const int* __restrict__ perm,
                                                   double* __restrict__ yout, int64_t gbase,
                                                   int nb) {
  extern __shared__ double xs[];
  __shared__ double red[64 * WPB];
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  const int st = leaf_start[k], n = leaf_count[k];
  if (n == 0) return;
  const int nc = n - r;
  const double* P = l2p_cm + l2poff[g];
  const double* nblk = near_part + 9 * (int64_t)st;  // 9 slots of n values
  const int* rm = rowmap + st;
  const int w = threadIdx.x >> 6;
  double pv[B];
  GemvLane q;
  if (nc > 0) {
    q = gemv_lane(nc, 0, r);
    gemv_load<B>(pv, P, r, nc, 0, q, w, WPB);
  }
  // the near-field values and destinations of this thread's first row (stored
  // rows first, then the skeleton rows), loaded early
  double nv[9];
  int dst_i = 0;
  {
    const int i = rm[min((int)threadIdx.x, n - 1)];
#pragma unroll
    for (int t = 0; t < 9; t++) nv[t] = nblk[t * n + i];
    dst_i = perm[st + i];
  }
  local_coeffs<WPB>(xs, r, inv + soff[g], has_parent, part, slot_off + (int64_t)kILSlots * g);
  __syncthreads();
  int row = threadIdx.x;  // the row whose nv/dst_i this thread holds
  for (int cb = 0; cb < nc; cb += 64) {
    if (cb > 0) {
      q = gemv_lane(nc, cb, r);
      gemv_load<B>(pv, P, r, nc, cb, q, w, WPB);
      row = cb + threadIdx.x;
      const int i = rm[min(row, n - 1)];
#pragma unroll
      for (int t = 0; t < 9; t++) nv[t] = nblk[t * n + i];
      dst_i = perm[st + i];
    }
    const double acc = gemv_chunk<B>(pv, P, r, nc, cb, q, xs, w, WPB);
    const double s = reduce_groups<WPB>(acc, red, q);
    if ((int)threadIdx.x < q.rc) {
      double nf = 0.0;
#pragma unroll
      for (int t = 0; t < 9; t++) nf += nv[t];
      yout[dst_i] = s + nf;
    }
  }
  // skeleton rows: thread u takes the rows v >= nc with v = u mod 64 WPB;
  // the values it already holds are reused when their row is v
  int v0 = threadIdx.x;
  while (v0 < nc) v0 += 64 * WPB;
  for (int v = v0; v < n; v += 64 * WPB) {
    if (v != row) {
      const int i = rm[v];
#pragma unroll
      for (int t = 0; t < 9; t++) nv[t] = nblk[t * n + i];
      dst_i = perm[st + i];
    }
    double nf = 0.0;
#pragma unroll
    for (int t = 0; t < 9; t++) nf += nv[t];
    yout[dst_i] = xs[v - nc] + nf;
  }
}

// Dense direct sum over all points for a set of rows, split over column
// chunks into partial sums: part[chunk * nrows + row].  CALLER = false: rows
// are KD-sorted indices (rows32) and q the KD-order charges; CALLER = true:
// rows are caller indices (rows64, mapped through inv_perm) and q the
// caller-order charges, read as q[perm[j]], the same values in the same order.
template <int KID, bool CALLER>
__global__ __launch_bounds__(256) void k_direct(const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ q,
                                                const int* __restrict__ rows32,
                                                const int64_t* __restrict__ rows64,
                                                const int* __restrict__ perm,
                                                const int* __restrict__ inv_perm, int nrows, int n,
                                                int chunk, double* __restrict__ part, KParams kp) {
  __shared__ double sx[256], sy[256], sq[256];
  const int rr = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * chunk;
  const int c1 = min(n, c0 + chunk);
  int row = 0;
  if (rr < nrows) row = CALLER ? inv_perm[rows64[rr]] : rows32[rr];
  const double xi = x[row], yi = y[row];
  double acc = 0.0;
  for (int cb = c0; cb < c1; cb += 256) {
    const int cn = min(256, c1 - cb);
    __syncthreads();
    if ((int)threadIdx.x < cn) {
      const int j = cb + threadIdx.x;
      sx[threadIdx.x] = x[j];
      sy[threadIdx.x] = y[j];
      sq[threadIdx.x] = CALLER ? q[perm[j]] : q[j];
    }
    __syncthreads();
    for (int t = 0; t < cn; t++) acc += kernel_xy<KID>(xi, yi, sx[t], sy[t], kp) * sq[t];
  }
  if (rr < nrows) part[(int64_t)blockIdx.y * nrows + rr] = acc;
}

__global__ void k_sum_parts(const double* __restrict__ part, int nrows, int nchunk,
                            double* __restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += part[(int64_t)c * nrows + r];
  out[r] = s;
}

}  // namespace

namespace {

#ifndef AFMM_GEMV_BATCH
#define AFMM_GEMV_BATCH 16
#endif
constexpr int kGemvBatch = AFMM_GEMV_BATCH;  // row steps per batch of loads in the GEMV
// batch sizes of the two largest levels' kernels (leaf-level M2M, level
// L-1 L2L, leaf pass), where one wave serves one box
// M2M by the transposed kernel where the level's shape allows (0: always
// the row-major kernel)
#ifndef AFMM_UP_T
#define AFMM_UP_T 1
#endif
#ifndef AFMM_UP_T_MIN_WPB
#define AFMM_UP_T_MIN_WPB 2
#endif
// waves per box at level L-1 (0: the waves_per_box rule)
#ifndef AFMM_WPB_UP_LM1
#define AFMM_WPB_UP_LM1 0
#endif
#ifndef AFMM_WPB_DOWN_LM1
#define AFMM_WPB_DOWN_LM1 0
#endif
#ifndef AFMM_B_UP_LEAF
#define AFMM_B_UP_LEAF AFMM_GEMV_BATCH
#endif
#ifndef AFMM_B_DOWN_LAST
#define AFMM_B_DOWN_LAST AFMM_GEMV_BATCH
#endif
#ifndef AFMM_B_LEAF
#define AFMM_B_LEAF AFMM_GEMV_BATCH
#endif

// Waves per box: the kernels are bound by memory latency (a few dependent
// round trips per box), so what counts is the number of boxes resident at
// once.  Large levels get one wave per box; small levels spread each box
// over up to kMaxWavesPerBox waves, for ~kTargetWaves waves per level.
#ifndef AFMM_TARGET_WAVES
#define AFMM_TARGET_WAVES 4096
#endif
#ifndef AFMM_MAX_WAVES_PER_BOX
#define AFMM_MAX_WAVES_PER_BOX 8
#endif
constexpr int64_t kTargetWaves = AFMM_TARGET_WAVES;
constexpr int kMaxWavesPerBox = AFMM_MAX_WAVES_PER_BOX;

int waves_per_box(int nb) {
  int w = 1;
  while (w < kMaxWavesPerBox && (int64_t)nb * w < kTargetWaves) w *= 2;
  return w;
}

}  // namespace

#define AFMM_DISPATCH_WPB(w, WPB, ...)    \
  switch (w) {                            \
    case 1: {                             \
      constexpr int WPB = 1;              \
      __VA_ARGS__;                        \
    } break;                              \
    case 2: {                             \
      constexpr int WPB = 2;              \
      __VA_ARGS__;                        \
    } break;                              \
    case 4: {                             \
      constexpr int WPB = 4;              \
      __VA_ARGS__;                        \
    } break;                              \
    case 8: {                             \
      constexpr int WPB = 8;              \
      __VA_ARGS__;                        \
    } break;                              \
    default: {                            \
      constexpr int WPB = 16;             \
      __VA_ARGS__;                        \
    } break;                              \
  }

namespace {

// M2M of level j (out_k = L2P_k^T src); the leaf level gathers the charges
// d_b (caller's order) itself.
// Shape of the transposed M2M kernel for a level (upper levels only): NCH
// row chunks of 64,
// WPB waves per box, 2^S column pairs per wave; at most 32 matrix loads per
// lane in flight (2^S NCH <= 16).  False if the level needs the row-major
// kernel (ranks above 64 or more than 256 stored rows).
bool up_t_shape(const LevelInfo& li, int& S, int& NCH, int& WPB) {
  if (!AFMM_UP_T || li.max_rank < 1 || li.max_rank > 64 || li.max_stored > 256) return false;
  NCH = ceil_div(std::max(1, li.max_stored), 64);
  const int max_s = NCH == 1 ? 4 : NCH == 2 ? 3 : 2;
  const int V = (li.max_rank + 1) / 2;
  for (int w : {2, 4, 8}) {
    if (w < AFMM_UP_T_MIN_WPB) continue;
    const int vw = ceil_div(V, w);
    int sh = 1;
    while ((1 << sh) < vw) sh++;
    if (sh <= max_s) {
      S = sh;
      WPB = w;
      return true;
    }
  }
  return false;
}

template <int S, int NCH>
void launch_up_t_wpb(Engine& e, int j, int WPB, hipStream_t s) {
  const LevelInfo& li = e.lv[j];
  const int64_t* soff = e.soff_g.as<int64_t>();
#define AFMM_UP_T_ARGS                                                                          \
  e.l2p_cm.as<double>(), e.l2poff_g.as<int64_t>(), e.rank_g.as<int>(), e.nin_g.as<int>(), soff, \
      e.rowmap.as<int>() + e.n, e.outv.as<double>(), soff + e.lv[j + 1].gbase,                 \
      e.outv.as<double>(), li.gbase, li.nb
  switch (WPB) {
    case 2: k_up_t<S, NCH, 2><<<li.nb, 128, 0, s>>>(AFMM_UP_T_ARGS); break;
    case 4: k_up_t<S, NCH, 4><<<li.nb, 256, 0, s>>>(AFMM_UP_T_ARGS); break;
    default: k_up_t<S, NCH, 8><<<li.nb, 512, 0, s>>>(AFMM_UP_T_ARGS); break;
  }
#undef AFMM_UP_T_ARGS
  AFMM_LAUNCH_CHECK();
}

void launch_up_t(Engine& e, int j, int S, int NCH, int WPB, hipStream_t s) {
  switch (NCH * 8 + S) {
    case 8 + 1: launch_up_t_wpb<1, 1>(e, j, WPB, s); break;
    case 8 + 2: launch_up_t_wpb<2, 1>(e, j, WPB, s); break;
    case 8 + 3: launch_up_t_wpb<3, 1>(e, j, WPB, s); break;
    case 8 + 4: launch_up_t_wpb<4, 1>(e, j, WPB, s); break;
    case 16 + 1: launch_up_t_wpb<1, 2>(e, j, WPB, s); break;
    case 16 + 2: launch_up_t_wpb<2, 2>(e, j, WPB, s); break;
    case 16 + 3: launch_up_t_wpb<3, 2>(e, j, WPB, s); break;
    case 24 + 1: launch_up_t_wpb<1, 3>(e, j, WPB, s); break;
    case 24 + 2: launch_up_t_wpb<2, 3>(e, j, WPB, s); break;
    case 32 + 1: launch_up_t_wpb<1, 4>(e, j, WPB, s); break;
    default: launch_up_t_wpb<2, 4>(e, j, WPB, s); break;
  }
}

void launch_up(Engine& e, int j, const double* d_b, hipStream_t s) {
  const LevelInfo& li = e.lv[j];
  // upper levels: the transposed kernel (measured faster there); the leaf
  // level gathers its charges from the caller's order (random reads) and
  // runs as fast on the row-major kernel
  if (j < e.L) {
    int S = 0, NCH = 0, WPB = 0;
    if (up_t_shape(li, S, NCH, WPB)) {
      launch_up_t(e, j, S, NCH, WPB, s);
      return;
    }
  }
  const int wpb = (j == e.L - 1 && AFMM_WPB_UP_LM1 > 0) ? AFMM_WPB_UP_LM1 : waves_per_box(li.nb);
  const size_t shm = sizeof(double) * std::max(1, li.max_nin);
  const bool leaf = j == e.L;
  const int64_t* soff = e.soff_g.as<int64_t>();
  AFMM_DISPATCH_WPB(wpb, WPB,
                    if (leaf) {
                      k_up<WPB, AFMM_B_UP_LEAF><<<li.nb, 64 * WPB, shm, s>>>(
                          e.l2p_rm.as<double>(), e.l2poff_g.as<int64_t>(), e.rank_g.as<int>(),
                          e.nin_g.as<int>(), soff, e.rowmap.as<int>(), nullptr, nullptr, d_b,
                          e.perm.as<int>(),
                          e.d_leaf_start.as<int>(), e.qs.as<double>(), e.outv.as<double>(),
                          li.gbase, li.nb);
                    } else {
                      k_up<WPB, kGemvBatch><<<li.nb, 64 * WPB, shm, s>>>(
                          e.l2p_rm.as<double>(), e.l2poff_g.as<int64_t>(), e.rank_g.as<int>(),
                          e.nin_g.as<int>(), soff, e.rowmap.as<int>() + e.n, e.outv.as<double>(),
                          soff + e.lv[j + 1].gbase, nullptr, nullptr, nullptr, nullptr,
                          e.outv.as<double>(), li.gbase, li.nb);
                    });
  AFMM_LAUNCH_CHECK();
}

// M2L pair blocks of group g (0: levels 2..L-1, 1: level L): one launch
// per register-tile class present (pairs are stored grouped by class).
// column tiles (8 columns each) per streamed chunk of the streamed classes;
// narrow chunks keep the registers low, so more M2L waves fit beside the
// near field's (N = 1e6, leaf 100, a/R, tol 1e-14: 2.46 ms against 2.62 ms
// with 4/3/2 tiles; equal at 1e-10 and 1e-12)
#ifndef AFMM_SCT4
#define AFMM_SCT4 2
#endif
#ifndef AFMM_SCT6
#define AFMM_SCT6 2
#endif
#ifndef AFMM_SCT8
#define AFMM_SCT8 1
#endif
#ifndef AFMM_SCT10
#define AFMM_SCT10 1
#endif
static_assert(kM2LAllClasses == 13 && kM2LStreamRows[0] == 4 && kM2LStreamRows[1] == 6 &&
                  kM2LStreamRows[2] == 8 && kM2LStreamRows[3] == 10 &&
                  kM2LStreamRows[4] == 12 && kM2LStreamRows[5] == 16,
              "launch_m2l's cases follow the class tables");
template <int ST>
void launch_m2l_classes(Engine& e, int g, hipStream_t s) {
  const PairDesc* pd0 = e.pair_desc.as<PairDesc>();
  double* part = e.m2l_part.as<double>();
  const double* outv = e.outv.as<double>();
  const double* B = e.m2l.as<double>();
  const B48Pair* Bh = e.m2l_b48.as<B48Pair>();
  for (int c = 0; c <= kM2LAllClasses; c++) {
    const int64_t p0 = e.m2l_cls_begin[g][c], p1 = e.m2l_cls_end[g][c];
    if (p1 <= p0) continue;
    const PairDesc* pd = pd0 + p0;
    const int np = (int)(p1 - p0);
    // K(second, first) = ts K(first, second)^T: -1 for an antisymmetric kernel
    const double ts = e.kid == K_USER_ANTI ? -1.0 : 1.0;
#define AFMM_M2L_ARGS <<<np, 64, 0, s>>>(B, Bh, pd, outv, part, np, ts)
    switch (c) {
      case 0: k_m2l_reg<2, 4, ST> AFMM_M2L_ARGS; break;
      case 1: k_m2l_reg<2, 5, ST> AFMM_M2L_ARGS; break;
      case 2: k_m2l_reg<2, 6, ST> AFMM_M2L_ARGS; break;
      case 3: k_m2l_reg<3, 4, ST> AFMM_M2L_ARGS; break;
      case 4: k_m2l_reg<3, 5, ST> AFMM_M2L_ARGS; break;
      case 5: k_m2l_reg<3, 6, ST> AFMM_M2L_ARGS; break;
      case 6: k_m2l_reg<4, 8, ST> AFMM_M2L_ARGS; break;
      case 7: k_m2l_stream<4, AFMM_SCT4, ST> AFMM_M2L_ARGS; break;
      case 8: k_m2l_stream<6, AFMM_SCT6, ST> AFMM_M2L_ARGS; break;
      case 9: k_m2l_stream<8, AFMM_SCT8, ST> AFMM_M2L_ARGS; break;
      case 10: k_m2l_stream<10, AFMM_SCT10, ST> AFMM_M2L_ARGS; break;
      case 11: k_m2l_stream<12, 1, ST> AFMM_M2L_ARGS; break;
      case 12: k_m2l_stream<16, 1, ST> AFMM_M2L_ARGS; break;
      default: throw Error(6, "M2L block taller than 256 rows");
    }
#undef AFMM_M2L_ARGS
    AFMM_LAUNCH_CHECK();
  }
}

void launch_m2l(Engine& e, int g, hipStream_t s) {
  const PairDesc* pd0 = e.pair_desc.as<PairDesc>();
  double* part = e.m2l_part.as<double>();
  const double* outv = e.outv.as<double>();
  if (e.m2l_store == M2L_STORE_B48) launch_m2l_classes<M2L_STORE_B48>(e, g, s);
  else launch_m2l_classes<M2L_STORE_F64>(e, g, s);
  const int64_t f0 = e.m2l_cls_begin[g][kM2LFactored], f1 = e.m2l_cls_end[g][kM2LFactored];
  if (f1 > f0) {
    const double ts = e.kid == K_USER_ANTI ? -1.0 : 1.0;
    k_m2l_fact<<<(unsigned)ceil_div(f1 - f0, 4), 256, 0, s>>>(e.m2lf.as<double>(), pd0 + f0,
                                                              outv, part, (int)(f1 - f0), ts);
    AFMM_LAUNCH_CHECK();
  }
}

// L2L of level j (2 <= j < L), with the level's M2L sums fused in.
void launch_down(Engine& e, int j, hipStream_t s) {
  const LevelInfo& li = e.lv[j];
  const int wpb =
      (j == e.L - 1 && AFMM_WPB_DOWN_LM1 > 0) ? AFMM_WPB_DOWN_LM1 : waves_per_box(li.nb);
  const size_t shm = sizeof(double) * std::max(1, li.max_rank);
  const int64_t* soff = e.soff_g.as<int64_t>();
  AFMM_DISPATCH_WPB(wpb, WPB,
                    if (j == e.L - 1) {
                      k_down<WPB, AFMM_B_DOWN_LAST><<<li.nb, 64 * WPB, shm, s>>>(
                          e.l2p_cm.as<double>(), e.l2poff_g.as<int64_t>(), e.rank_g.as<int>(),
                          e.nin_g.as<int>(), soff, soff + e.lv[j + 1].gbase,
                          e.rowmap.as<int>() + e.n, e.m2l_part.as<double>(), e.slot_off.as<int>(),
                          e.inv.as<double>(), j > 2 ? 1 : 0,
                          li.gbase, li.nb);
                    } else {
                      k_down<WPB, kGemvBatch><<<li.nb, 64 * WPB, shm, s>>>(
                          e.l2p_cm.as<double>(), e.l2poff_g.as<int64_t>(), e.rank_g.as<int>(),
                          e.nin_g.as<int>(), soff, soff + e.lv[j + 1].gbase,
                          e.rowmap.as<int>() + e.n, e.m2l_part.as<double>(), e.slot_off.as<int>(),
                          e.inv.as<double>(), j > 2 ? 1 : 0,
                          li.gbase, li.nb);
                    });
  AFMM_LAUNCH_CHECK();
}

// Near field: persistent grid, a few waves per SIMD, leaving wave slots to
// the memory-bound kernels that run beside it.
void launch_near(Engine& e, hipStream_t s) {
  const int nq = kNearSlices * kNearCounterStride;
  e.near_queue.alloc(sizeof(int) * nq);
  k_zero_ints<<<1, nq, 0, s>>>(e.near_queue.as<int>(), nq);
  AFMM_LAUNCH_CHECK();
  // Wave slots the near field may take beside the memory-bound kernels
  // (p2p_waves_per_simd per SIMD); 0: one workgroup per pair.  With fewer
  // pairs than slots, several waves share a pair (up to 4, and no more than
  // the chunks of the largest leaf).
  const int nch = (e.max_leaf + 63) / 64;
  const int64_t slots = (int64_t)4 * e.num_cus * e.p2p_waves_per_simd;
  int nw = 1;
  if (e.p2p_waves_per_simd > 0 && e.near_upairs > 0 && e.near_upairs < slots)
    nw = (int)std::min<int64_t>(std::min(4, nch), (slots + e.near_upairs - 1) / e.near_upairs);
  nw = std::max(1, nw);
  const int64_t cap = e.p2p_waves_per_simd > 0 ? slots / nw : e.near_upairs;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(e.near_upairs, cap));
#define AFMM_NEAR_LAUNCH(KID, NCH)                                                          \
  k_p2p<KID, NCH><<<grid, 64 * nw, 0, s>>>(                                                \
      e.x.as<double>(), e.y.as<double>(), e.qs.as<double>(), e.d_leaf_start.as<int>(),     \
      e.d_leaf_count.as<int>(), e.near_pair.as<int>(), e.near_pair_part.as<int>(),         \
      e.near_part.as<double>(), (int)e.near_upairs, e.kp, e.near_tab.as<double2>(),       \
      e.near_queue.as<int>())
  // A's chunks in registers: arrays for the largest leaf (up to 512 points)
  AFMM_DISPATCH_KERNEL(e.kid, KID,
                       if (nch <= 1) AFMM_NEAR_LAUNCH(KID, 1);
                       else if (nch <= 2) AFMM_NEAR_LAUNCH(KID, 2);
                       else if (nch <= 4) AFMM_NEAR_LAUNCH(KID, 4);
                       else AFMM_NEAR_LAUNCH(KID, 8));
#undef AFMM_NEAR_LAUNCH
  AFMM_LAUNCH_CHECK();
}

// Leaf level: local coefficients, leaf L2P, near-field sums, scatter.
void launch_leaf(Engine& e, double* d_y, hipStream_t s) {
  const int L = e.L;
  const LevelInfo& li = e.lv[L];
  const int wpb = waves_per_box(li.nb);
  const size_t shm = sizeof(double) * std::max(1, li.max_rank);
  AFMM_DISPATCH_WPB(wpb, WPB,
                    k_leaf<WPB, AFMM_B_LEAF><<<li.nb, 64 * WPB, shm, s>>>(
                        e.l2p_cm.as<double>(), e.l2poff_g.as<int64_t>(), e.rank_g.as<int>(),
                        e.soff_g.as<int64_t>(), e.d_leaf_start.as<int>(),
                        e.d_leaf_count.as<int>(), e.rowmap.as<int>(), e.m2l_part.as<double>(),
                        e.slot_off.as<int>(), e.inv.as<double>(), L > 2 ? 1 : 0, e.near_part.as<double>(), e.perm.as<int>(), d_y,
                        li.gbase, li.nb));
  AFMM_LAUNCH_CHECK();
}

// Chunks of the fused upper-tree passes: levels [jt, jb], 3 at a time from
// level L - 1 up, down to level 4 (4^jt >= 256 workgroups: with fewer, one
// launch per level measured faster).  Levels above the chunks, or all levels
// when the pass runs level by level (AFMM2D_TREE_FUSED=0 or a level whose
// rows do not fit the workgroup's LDS), use the per-level kernels.
// (the static LDS of the tree kernels: the reduction rows)
constexpr size_t kTreeStaticLds = 8 * 1024;
constexpr int kTreeMinTop = 4;
std::vector<std::pair<int, int>> tree_chunks(const Engine& e, bool up) {
  std::vector<std::pair<int, int>> ch;
  if (!e.tree_fused || e.L < 3) return ch;
  for (int j = 2; j < e.L; j++) {
    const int need = up ? e.lv[j].max_nin : e.lv[j].max_rank;
    if ((size_t)kTreeWaves * need * sizeof(double) + kTreeStaticLds > (size_t)e.max_lds) return ch;
  }
  for (int jb = e.L - 1; jb >= kTreeMinTop;) {
    const int jt = std::max(kTreeMinTop, jb - 2);
    ch.push_back({jt, jb});
    jb = jt - 1;
  }
  return ch;
}

void launch_tree(Engine& e, int jt, int jb, bool up, hipStream_t s) {
  TreeArgs a{};
  a.l2p = up ? e.l2p_rm.as<double>() : e.l2p_cm.as<double>();
  a.l2poff = e.l2poff_g.as<int64_t>();
  a.rank = e.rank_g.as<int>();
  a.nin = e.nin_g.as<int>();
  a.soff = e.soff_g.as<int64_t>();
  a.rowmap_up = e.rowmap.as<int>() + e.n;
  a.vec = up ? e.outv.as<double>() : e.inv.as<double>();
  a.part = e.m2l_part.as<double>();
  a.slot_off = e.slot_off.as<int>();
  a.D = jb - jt + 1;
  for (int d = 0; d <= a.D; d++) a.gb[d] = e.lv[jt + d].gbase;
  a.top_has_parent = jt > 2 ? 1 : 0;
  int stride = 1;
  for (int j = jt; j <= jb; j++) stride = std::max(stride, up ? e.lv[j].max_nin : e.lv[j].max_rank);
  a.xs_stride = stride;
  const size_t shm = sizeof(double) * (size_t)kTreeWaves * stride;
  const unsigned grid = 1u << (2 * jt);
  if (up) k_up_tree<<<grid, 64 * kTreeWaves, shm, s>>>(a);
  else k_down_tree<<<grid, 64 * kTreeWaves, shm, s>>>(a);
  AFMM_LAUNCH_CHECK();
}

void upper_up(Engine& e, hipStream_t s) {
  hipEvent_t a, b;
  e.prof_begin("m2m_upper", s, &a, &b);
  const auto ch = tree_chunks(e, true);
  int j = e.L - 1;
  for (const auto& c : ch) {
    launch_tree(e, c.first, c.second, true, s);
    j = c.first - 1;
  }
  for (; j >= 2; j--) launch_up(e, j, nullptr, s);
  e.prof_end(s, b);
}

void upper_m2l(Engine& e, hipStream_t s) {
  if (e.L <= 2) return;
  hipEvent_t a, b;
  e.prof_begin("m2l_upper", s, &a, &b);
  launch_m2l(e, 0, s);
  e.prof_end(s, b);
}

void down_chain(Engine& e, hipStream_t s) {
  hipEvent_t a, b;
  e.prof_begin("l2l", s, &a, &b);
  const auto ch = tree_chunks(e, false);
  const int jf = ch.empty() ? e.L : ch.back().first;  // first fused level
  for (int j = 2; j < jf; j++) launch_down(e, j, s);
  for (size_t i = ch.size(); i-- > 0;) launch_tree(e, ch[i].first, ch[i].second, false, s);
  e.prof_end(s, b);
}

}  // namespace

void gather_sorted(Engine& e, const double* d_b, double* d_qs, hipStream_t s) {
  k_gather<<<ceil_div(e.n, 256), 256, 0, s>>>(d_b, e.perm.as<int>(), d_qs, e.n);
  AFMM_LAUNCH_CHECK();
}

// Schedules (Engine::tree_mode; the near field forks onto its own stream
// unless near_overlap is off, in which case everything runs on s in order):
//   0  s: M2M (all levels; the leaf level gathers the charges into KD
//      order), M2L (upper levels, then leaf level),
//      L2L, leaf pass; the near field forks after the M2M.
//   1  after the leaf-level M2M, the upper tree (M2M, M2L and L2L of levels
//      < L) runs on a third stream beside the leaf-level M2L.
//   2  as 0, but after the upper levels' M2L the L2L chain runs on a third
//      stream beside the leaf-level M2L.
//   3  as 0, but the near field forks after the upper levels' M2L.
void matvec_device(Engine& e, const double* d_b, double* d_y, hipStream_t s) {
  const int L = e.L;
  e.qs.alloc(sizeof(double) * e.n);
  hipEvent_t a, b;
  const bool near_side = e.near_overlap;
  const int mode = e.near_overlap ? e.tree_mode : 0;
  if ((mode == 1 || mode == 2) && !e.s_tree)
    AFMM_HIP(hipStreamCreateWithFlags(&e.s_tree, hipStreamNonBlocking));

  e.prof_begin("m2m", s, &a, &b);
  launch_up(e, L, d_b, s);
  e.prof_end(s, b);
  if (mode == 1) {
    AFMM_HIP(hipEventRecord(e.ev_fork, s));
    AFMM_HIP(hipStreamWaitEvent(e.s_tree, e.ev_fork, 0));
    upper_up(e, e.s_tree);
    upper_m2l(e, e.s_tree);
    down_chain(e, e.s_tree);
    AFMM_HIP(hipEventRecord(e.ev_tree, e.s_tree));
  } else {
    upper_up(e, s);
  }

  auto fork_near = [&]() {
    AFMM_HIP(hipEventRecord(e.ev_fork, s));
    AFMM_HIP(hipStreamWaitEvent(e.s_aux, e.ev_fork, 0));
    e.prof_begin("p2p", e.s_aux, &a, &b);
    launch_near(e, e.s_aux);
    e.prof_end(e.s_aux, b);
    AFMM_HIP(hipEventRecord(e.ev_join, e.s_aux));
  };
  if (near_side && mode != 3) fork_near();
  if (mode != 1) upper_m2l(e, s);
  if (near_side && mode == 3) fork_near();
  if (mode == 2) {
    AFMM_HIP(hipEventRecord(e.ev_fork, s));
    AFMM_HIP(hipStreamWaitEvent(e.s_tree, e.ev_fork, 0));
    down_chain(e, e.s_tree);
    AFMM_HIP(hipEventRecord(e.ev_tree, e.s_tree));
  }

  e.prof_begin("m2l", s, &a, &b);
  launch_m2l(e, 1, s);
  e.prof_end(s, b);

  const bool tree_side = mode == 1 || mode == 2;
  if (!tree_side) down_chain(e, s);
  if (!near_side) {
    e.prof_begin("p2p", s, &a, &b);
    launch_near(e, s);
    e.prof_end(s, b);
  }
  if (near_side) AFMM_HIP(hipStreamWaitEvent(s, e.ev_join, 0));
  if (tree_side) AFMM_HIP(hipStreamWaitEvent(s, e.ev_tree, 0));
  e.prof_begin("l2p_out", s, &a, &b);
  launch_leaf(e, d_y, s);
  e.prof_end(s, b);
}

void add_device(double* d_y, const double* d_t, int n, hipStream_t s) {
  k_add<<<ceil_div(n, 256), 256, 0, s>>>(d_y, d_t, n);
  AFMM_LAUNCH_CHECK();
}

void near_field_device(Engine& e, const double* d_b, double* d_y, hipStream_t s) {
  e.qs.alloc(sizeof(double) * e.n);
  gather_sorted(e, d_b, e.qs.as<double>(), s);
  hipEvent_t a = nullptr, b = nullptr;
  e.prof_begin("p2p", s, &a, &b);
  launch_near(e, s);
  e.prof_end(s, b);
  const int nleaves = 1 << (2 * e.L);
  k_near_out<<<nleaves, 256, 0, s>>>(e.near_part.as<double>(), e.d_leaf_start.as<int>(),
                                      e.d_leaf_count.as<int>(), e.perm.as<int>(), d_y, nleaves);
  AFMM_LAUNCH_CHECK();
}

namespace {

// Column chunks of the direct sum for nrows rows: enough workgroups to fill
// the device (nrows * nchunk <= 2^20 for every nrows <= 2^20).
int direct_chunks(int n, int nrows, int& chunk) {
  const int rb = (int)ceil_div(nrows, 256);
  int nchunk = std::max(1, std::min(4096 / rb, (int)ceil_div(n, 4096)));
  chunk = (int)ceil_div(n, nchunk);
  return (int)ceil_div(n, chunk);
}

}  // namespace

void direct_rows_device(Engine& e, int nrows, const int* d_rows, const double* d_qs,
                        double* d_out, hipStream_t s) {
  if (nrows <= 0) return;
  int chunk = 0;
  const int nchunk = direct_chunks(e.n, nrows, chunk);
  DevBuf part;
  part.alloc(sizeof(double) * (size_t)nrows * nchunk);
  dim3 grid(ceil_div(nrows, 256), nchunk);
  AFMM_DISPATCH_KERNEL(e.kid_full, KID,
                       k_direct<KID, false><<<grid, 256, 0, s>>>(
                           e.x.as<double>(), e.y.as<double>(), d_qs, d_rows, nullptr, nullptr,
                           nullptr, nrows, e.n, chunk, part.as<double>(), e.kp));
  AFMM_LAUNCH_CHECK();
  k_sum_parts<<<ceil_div(nrows, 256), 256, 0, s>>>(part.as<double>(), nrows, nchunk, d_out);
  AFMM_LAUNCH_CHECK();
  AFMM_HIP(hipStreamSynchronize(s));
}

// ---- Direct-sum rows without allocation.
//
// The partial sums live in a file-scope device array (each GPU holds its own
// copy until the process ends).  Its device address comes from
// hipGetSymbolAddress, once per device with that device current; the array's
// name is never used as an address in host code.  Every call on a GPU shares
// that GPU's array: calls that overlap on one GPU (two streams, two threads)
// race.  Nothing in it needs to start at zero: each call writes every partial
// it reads.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};  // device -> address of its scratch

double* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(6, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return g_scratch_ptr[device];
  int cur = -1;
  AFMM_HIP(hipGetDevice(&cur));
  if (cur != device) AFMM_HIP(hipSetDevice(device));
  double* p = nullptr;
  const hipError_t st = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (cur != device) (void)hipSetDevice(cur);
  if (st != hipSuccess || !p)
    throw Error(5, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st));
  g_scratch_ptr[device] = p;
  return p;
}
}  // namespace

void direct_rows_noalloc(Engine& e, int nrows, const int64_t* d_rows, const double* d_b,
                         double* d_out, hipStream_t s) {
  if (nrows <= 0) return;
  int chunk = 0;
  const int nchunk = direct_chunks(e.n, nrows, chunk);
  if ((int64_t)nrows * nchunk > kNoAllocScratchDoubles)
    throw Error(6, "direct rows: partial sums exceed the static scratch");
  double* part = scratch_ptr(e.device);
  dim3 grid(ceil_div(nrows, 256), nchunk);
  AFMM_DISPATCH_KERNEL(e.kid_full, KID,
                       k_direct<KID, true><<<grid, 256, 0, s>>>(
                           e.x.as<double>(), e.y.as<double>(), d_b, nullptr, d_rows,
                           e.perm.as<int>(), e.inv_perm_d.as<int>(), nrows, e.n, chunk, part,
                           e.kp));
  AFMM_LAUNCH_CHECK();
  k_sum_parts<<<ceil_div(nrows, 256), 256, 0, s>>>(part, nrows, nchunk, d_out);
  AFMM_LAUNCH_CHECK();
}

}  // namespace afmm
