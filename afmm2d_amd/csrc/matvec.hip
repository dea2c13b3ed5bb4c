// matvec.hip -- the matvec passes of AFMM2D on the GPU (FP64).
//
// Reference flow (AFMM2D.cpp:46-65, AFMM2DTree.cpp:1120-1225): gather q into
// KD order, M2M upward (levels L..2, out_k = L2P_k^T src), M2L (in_k =
// sum over the interaction list of M2L[k][kIL] out_kIL), L2L downward
// (children.in += L2P_k in_k; leaves: potential = L2P_k in_k), near field
// (potential += K(leaf, neighbour) q_neighbour for 8 neighbours and self),
// scatter back to caller order.
//
// Stream plan: the near field depends only on the gathered q.  It is forked
// onto a second stream after the upward pass and runs concurrently with the
// HBM-bound M2L and L2L; the last kernel joins both (leaf L2P + near field +
// scatter, fused).
//
// M2L reads each symmetric block once: the pair kernel stages
// B = K(skel_k, skel_b) in LDS with 16-byte loads and writes B out_b and
// B^T out_k as partials; the reduction adds each box's partials in its
// interaction-slot order, so results do not depend on scheduling.
#include "engine.hpp"

namespace afmm {

namespace {

__global__ void k_gather(const double* __restrict__ b, const int* __restrict__ perm,
                         double* __restrict__ qs, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) qs[i] = b[perm[i]];
}

// M2M: one workgroup of 4 waves per box; out[c] = sum_i L2P[i][c] src[i]
// (row-major L2P: lanes over c read each row coalesced, the 4 waves split
// the rows and are summed through LDS).  src is the leaf's q segment (leaf
// level) or the children's concatenated multipoles (always contiguous).
__global__ __launch_bounds__(256) void k_m2m(const double* __restrict__ l2p_rm,
                                             const int64_t* __restrict__ l2poff,
                                             const int* __restrict__ rank,
                                             const int* __restrict__ nin,
                                             const int64_t* __restrict__ soff,
                                             const double* __restrict__ src,
                                             const int64_t* __restrict__ src_off_child,
                                             const int* __restrict__ src_off_leaf,
                                             double* __restrict__ outv, int64_t gbase, int nb) {
  __shared__ double red[4][64];
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  if (r == 0) return;
  const int n = nin[g];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double* P = l2p_rm + l2poff[g];
  const double* x = src + (src_off_leaf ? (int64_t)src_off_leaf[k] : src_off_child[4 * k]);
  double* out = outv + soff[g];
  for (int cb = 0; cb < r; cb += 64) {
    const int c = cb + lane;
    double acc = 0.0;
    if (c < r) {
#pragma unroll 4
      for (int i = w; i < n; i += 4) acc += P[(int64_t)i * r + c] * x[i];
    }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0 && c < r) out[c] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    __syncthreads();
  }
}

// M2L pair kernel: one workgroup of 4 waves per pair.  LDS holds the block
// (tile doubles), out_k and out_b (rcap each) and the per-wave partial sums
// (2 x 4 x rcap).  The 4 waves split the columns for B out_b and the rows
// for B^T out_k; the partials are then summed in wave order.
__global__ __launch_bounds__(256) void k_m2l_pairs(const double* __restrict__ m2l,
                                                   const int* __restrict__ pk,
                                                   const int* __restrict__ pb,
                                                   const int64_t* __restrict__ poff,
                                                   const int* __restrict__ ppart,
                                                   const int* __restrict__ rank,
                                                   const int64_t* __restrict__ soff,
                                                   const double* __restrict__ outv,
                                                   double* __restrict__ part, int tile, int rcap) {
  extern __shared__ double lds[];
  double* T = lds;
  double* xk = lds + tile;
  double* xb = xk + rcap;
  double* rf = xb + rcap;      // [4][rcap]
  double* rt = rf + 4 * rcap;  // [4][rcap]
  const int p = blockIdx.x;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int k = pk[p], b = pb[p];
  const int rk = rank[k], rb = rank[b];
  double* fw = part + ppart[p];
  double* tw = fw + rk;
  const double* B = m2l + poff[p];
  for (int i = tid; i < rk; i += 256) xk[i] = outv[soff[k] + i];
  for (int c = tid; c < rb; c += 256) xb[c] = outv[soff[b] + c];
  for (int u = tid; u < 4 * rcap; u += 256) rf[u] = 0.0;
  const int tot = rk * rb;
  // columns [c0, c1) of the block are resident in T, at T[i + (c - c0) rk]
  const int chunk = tot <= tile ? rb : max(1, tile / rk);
  for (int c0 = 0; c0 < rb; c0 += chunk) {
    const int c1 = min(rb, c0 + chunk);
    const int ne = (c1 - c0) * rk;
    __syncthreads();
    if (chunk == rb) {
      const double2* B2 = reinterpret_cast<const double2*>(B);
      double2* T2 = reinterpret_cast<double2*>(T);
      const int n2 = ne >> 1;
      int e = tid;
      for (; e + 256 < n2; e += 512) {
        double2 v0 = B2[e], v1 = B2[e + 256];
        T2[e] = v0;
        T2[e + 256] = v1;
      }
      for (; e < n2; e += 256) T2[e] = B2[e];
      if ((ne & 1) && tid == 0) T[ne - 1] = B[ne - 1];
    } else {
      for (int e = tid; e < ne; e += 256) T[e] = B[(int64_t)c0 * rk + e];
    }
    __syncthreads();
    // B out_b: wave w takes every 4th column of the chunk
    for (int i = lane; i < rk; i += 64) {
      double acc = 0.0;
#pragma unroll 4
      for (int c = c0 + w; c < c1; c += 4) acc += T[i + (c - c0) * rk] * xb[c];
      rf[w * rcap + i] += acc;
    }
    // B^T out_k: wave w takes a quarter of the rows
    const int q = (rk + 3) >> 2;
    const int i0 = w * q, i1 = min(rk, i0 + q);
    for (int c = c0 + lane; c < c1; c += 64) {
      const double* Tc = T + (c - c0) * rk;
      double acc = 0.0;
#pragma unroll 4
      for (int i = i0; i < i1; i++) acc += Tc[i] * xk[i];
      rt[w * rcap + c] = acc;
    }
  }
  __syncthreads();
  for (int i = tid; i < rk; i += 256)
    fw[i] = (rf[i] + rf[rcap + i]) + (rf[2 * rcap + i] + rf[3 * rcap + i]);
  for (int c = tid; c < rb; c += 256)
    tw[c] = (rt[c] + rt[rcap + c]) + (rt[2 * rcap + c] + rt[3 * rcap + c]);
}

// M2L pair kernel, persistent and software-pipelined: each workgroup walks
// pairs p, p + G, p + 2G, ...; while it computes one pair from LDS the
// 16-byte loads of the next two pairs are in flight into two register sets
// (NV double2 per thread each, enough for blocks of up to NV * 512
// coefficients).  Same products and summation order as k_m2l_pairs.
template <int NV>
__global__ __launch_bounds__(256) void k_m2l_pipe(const double* __restrict__ m2l,
                                                  const PairDesc* __restrict__ pd,
                                                  const double* __restrict__ outv,
                                                  double* __restrict__ part, int npairs,
                                                  int tile, int rcap) {
  extern __shared__ double lds[];
  double* T = lds;
  double* xk = lds + tile;
  double* xb = xk + rcap;
  double* rf = xb + rcap;      // [4][rcap]
  double* rt = rf + 4 * rcap;  // [4][rcap]
  double2* T2 = reinterpret_cast<double2*>(T);
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  // two staging register sets, A and B (double arrays: an array of double2
  // is not promoted to registers by the compiler)
  double ax[NV], ay[NV], bx[NV], by[NV];
  double at = 0.0, ak = 0.0, ab = 0.0, bt = 0.0, bk = 0.0, bb = 0.0;
  // Loads are unconditional (indices clamped into the block), so they issue
  // back to back; each set is consumed by a commit later in the loop.
#define AFMM_M2L_LOAD(X, Y, TL, XK, XB, D)                                             \
  {                                                                                    \
    const int tot_ = (D).rk * (D).rb, n2_ = tot_ >> 1;                                 \
    const double2* B2_ = reinterpret_cast<const double2*>(m2l + (n2_ > 0 ? (D).off : 0)); \
    const int last_ = n2_ > 0 ? n2_ - 1 : 0;                                           \
    _Pragma("unroll") for (int v = 0; v < NV; v++) {                                   \
      const double2 t_ = B2_[min(tid + 256 * v, last_)];                               \
      X[v] = t_.x;                                                                     \
      Y[v] = t_.y;                                                                     \
    }                                                                                  \
    TL = m2l[tot_ > 0 ? (D).off + tot_ - 1 : 0];                                       \
    XK = outv[(D).rk > 0 ? (D).sk + min(tid, (D).rk - 1) : 0];                         \
    XB = outv[(D).rb > 0 ? (D).sb + min(tid, (D).rb - 1) : 0];                         \
  }
#define AFMM_M2L_COMMIT(X, Y, TL, XK, XB, D)                                           \
  {                                                                                    \
    const int tot_ = (D).rk * (D).rb, n2_ = tot_ >> 1;                                 \
    _Pragma("unroll") for (int v = 0; v < NV; v++) {                                   \
      const int e_ = tid + 256 * v;                                                    \
      if (e_ < n2_) T2[e_] = make_double2(X[v], Y[v]);                                 \
    }                                                                                  \
    if ((tot_ & 1) && tid == 0) T[tot_ - 1] = TL;                                      \
    if (tid < (D).rk) xk[tid] = XK;                                                    \
    if (tid < (D).rb) xb[tid] = XB;                                                    \
  }
#define AFMM_M2L_COMPUTE(D)                                                            \
  {                                                                                    \
    const int rk = (D).rk, rb = (D).rb;                                                \
    for (int i = lane; i < rk; i += 64) {                                              \
      double acc = 0.0;                                                                \
      _Pragma("unroll 4") for (int c = w; c < rb; c += 4) acc += T[i + c * rk] * xb[c]; \
      rf[w * rcap + i] = acc;                                                          \
    }                                                                                  \
    const int q_ = (rk + 3) >> 2;                                                      \
    const int i0 = w * q_, i1 = min(rk, i0 + q_);                                      \
    for (int c = lane; c < rb; c += 64) {                                              \
      const double* Tc = T + c * rk;                                                   \
      double acc = 0.0;                                                                \
      _Pragma("unroll 4") for (int i = i0; i < i1; i++) acc += Tc[i] * xk[i];          \
      rt[w * rcap + c] = acc;                                                          \
    }                                                                                  \
    __syncthreads();                                                                   \
    double* fw = part + (D).part;                                                      \
    double* tw = fw + rk;                                                              \
    for (int i = tid; i < rk; i += 256)                                                \
      fw[i] = (rf[i] + rf[rcap + i]) + (rf[2 * rcap + i] + rf[3 * rcap + i]);          \
    for (int c = tid; c < rb; c += 256)                                                \
      tw[c] = (rt[c] + rt[rcap + c]) + (rt[2 * rcap + c] + rt[3 * rcap + c]);          \
  }
  const int G = gridDim.x;
  int p = blockIdx.x;
  if (p >= npairs) return;
  PairDesc d0 = pd[p];
  AFMM_M2L_LOAD(ax, ay, at, ak, ab, d0);
  PairDesc d1 = d0;
  if (p + G < npairs) {
    d1 = pd[p + G];
    AFMM_M2L_LOAD(bx, by, bt, bk, bb, d1);
  }
  AFMM_M2L_COMMIT(ax, ay, at, ak, ab, d0);
  __syncthreads();
  for (;;) {
    // pair p is in LDS, p + G in set B; set A is free
    PairDesc d2 = d0;
    if (p + 2 * G < npairs) {
      d2 = pd[p + 2 * G];
      AFMM_M2L_LOAD(ax, ay, at, ak, ab, d2);
    }
    AFMM_M2L_COMPUTE(d0);
    if (p + G >= npairs) break;
    AFMM_M2L_COMMIT(bx, by, bt, bk, bb, d1);
    __syncthreads();
    // pair p + G is in LDS, p + 2G in set A; set B is free
    PairDesc d3 = d1;
    if (p + 3 * G < npairs) {
      d3 = pd[p + 3 * G];
      AFMM_M2L_LOAD(bx, by, bt, bk, bb, d3);
    }
    AFMM_M2L_COMPUTE(d1);
    if (p + 2 * G >= npairs) break;
    AFMM_M2L_COMMIT(ax, ay, at, ak, ab, d2);
    __syncthreads();
    p += 2 * G;
    d0 = d2;
    d1 = d3;
  }
#undef AFMM_M2L_LOAD
#undef AFMM_M2L_COMMIT
#undef AFMM_M2L_COMPUTE
}

// in_k = sum of its interaction slots' partials, in slot order.
__global__ __launch_bounds__(64) void k_m2l_reduce(const double* __restrict__ part,
                                                   const int* __restrict__ red_start,
                                                   const int* __restrict__ red_part,
                                                   const int* __restrict__ rank,
                                                   const int64_t* __restrict__ soff,
                                                   double* __restrict__ inv, int nbox) {
  const int g = blockIdx.x;
  if (g >= nbox) return;
  const int r = rank[g];
  const int t0 = red_start[g], t1 = red_start[g + 1];
  double* out = inv + soff[g];
  for (int i = threadIdx.x; i < r; i += 64) {
    double acc = 0.0;
    for (int t = t0; t < t1; t++) acc += part[red_part[t] + i];
    out[i] = acc;
  }
}

// L2L for a non-leaf level: children.in += L2P_k in_k (column-major L2P,
// threads over rows).  The children's coefficients are contiguous.
__global__ __launch_bounds__(256) void k_l2l(const double* __restrict__ l2p_cm,
                                             const int64_t* __restrict__ l2poff,
                                             const int* __restrict__ rank,
                                             const int* __restrict__ nin,
                                             const int64_t* __restrict__ soff,
                                             const int64_t* __restrict__ child_soff,
                                             double* __restrict__ inv, int64_t gbase, int nb) {
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  if (r == 0) return;
  const int n = nin[g];
  const double* P = l2p_cm + l2poff[g];
  const double* x = inv + soff[g];
  double* dst = inv + child_soff[4 * k];
  for (int i = threadIdx.x; i < n; i += 256) {
    double acc = 0.0;
#pragma unroll 8
    for (int c = 0; c < r; c++) acc += P[i + (int64_t)c * n] * x[c];
    dst[i] += acc;
  }
}

// log(x) for normal x > 0 by Tang's table method (ACM TOMS 16(4), 1990):
// x = 2^e m, m in [1,2); j = the top 7 bits of m's fraction;
// c_j = 1 + (j + 1/2)/128, t = (RN(1/c_j), -log RN(1/c_j));
// log x = e ln2 + t.y + log1p(m t.x - 1), |m t.x - 1| < 2^-8, log1p by its
// Taylor polynomial to degree 6 (truncation < 4e-18).
__device__ __forceinline__ double log_tang(double x, const double2* __restrict__ tab,
                                           double ln2_hi, double ln2_lo) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
  const int ex = (int)(bits >> 52) - 1023;
  const int j = (int)((bits >> 45) & 127);
  const double m = __longlong_as_double((long long)((bits & 0x000FFFFFFFFFFFFFull) |
                                                    0x3FF0000000000000ull));
  const double2 t = tab[j];
  const double r = fma(m, t.x, -1.0);
  double p = fma(r, -1.0 / 6.0, 1.0 / 5.0);
  p = fma(r, p, -0.25);
  p = fma(r, p, 1.0 / 3.0);
  p = fma(r, p, -0.5);
  p = r * p;
  p = fma(r, p, r);  // log1p(r)
  const double e = (double)ex;
  return fma(e, ln2_hi, t.y) + fma(e, ln2_lo, p);
}

template <int KID>
__device__ __forceinline__ double near_kernel(double r2, const KParams& kp, const double2* tab,
                                              double ln2_hi, double ln2_lo) {
  if constexpr (KID == K_LOG) {
    // 2 x the log kernel; the factor 1/2 is applied to the sums
    return r2 == 0.0 ? 0.0 : log_tang(r2, tab, ln2_hi, ln2_lo);
  } else {
    return kernel_r2<KID>(r2, kp);
  }
}

// Near field, symmetric: one wave per unordered leaf pair (A, B), A <= B.
// Lane t holds target point t of an A chunk of 64; at step s it meets
// source point (t + s) mod 64 of a B chunk staged in LDS, so the 64 lanes
// touch 64 distinct sources and the source-side sums (for B's points) are
// accumulated in LDS with one FP64 LDS add per step.  Each kernel value is
// evaluated once and used for both directions.  A self pair (A, A) runs
// steps 0..32 on its diagonal chunk pairs (step 32 on lanes 0..31 only).
template <int KID>
__global__ __launch_bounds__(64) void k_p2p(const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const double* __restrict__ qs,
                                            const int* __restrict__ leaf_start,
                                            const int* __restrict__ leaf_count,
                                            const int* __restrict__ pairs,
                                            const int* __restrict__ pair_part,
                                            double* __restrict__ part, int npairs, int maxpad,
                                            KParams kp, const double2* __restrict__ logtab,
                                            double ln2_hi, double ln2_lo) {
  extern __shared__ double sm[];
  double2* tab = reinterpret_cast<double2*>(sm);  // 128 entries
  double* bx = sm + 256;
  double* by = bx + maxpad;
  double* bq = by + maxpad;
  double* bc = bq + maxpad;  // source-side sums
  const int t = threadIdx.x;
  if constexpr (KID == K_LOG) {
    for (int u = t; u < 128; u += 64) tab[u] = logtab[u];
  }
  // padding: far away, zero charge; K stays finite there
  const double FAR = 1.0e30;
  for (int p = blockIdx.x; p < npairs; p += gridDim.x) {
    const int A = pairs[2 * p], B = pairs[2 * p + 1];
    const bool self = A == B;
    const int sa = leaf_start[A], na = leaf_count[A];
    const int sb = leaf_start[B], nb = leaf_count[B];
    const int pb = (nb + 63) & ~63, pa = (na + 63) & ~63;
    __syncthreads();
    for (int j = t; j < pb; j += 64) {
      const bool v = j < nb;
      bx[j] = v ? x[sb + j] : FAR;
      by[j] = v ? y[sb + j] : FAR;
      bq[j] = v ? qs[sb + j] : 0.0;
      bc[j] = 0.0;
    }
    __syncthreads();
    double* out = part + pair_part[p];
    for (int ca = 0; ca < pa; ca += 64) {
      const int i = ca + t;
      const bool vi = i < na;
      const double xi = vi ? x[sa + i] : FAR;
      const double yi = vi ? y[sa + i] : FAR;
      const double qi = vi ? qs[sa + i] : 0.0;
      double row = 0.0;
      for (int cb = self ? ca : 0; cb < pb; cb += 64) {
        const bool diag = self && cb == ca;
        int s = 0;
        if (diag) {
          // step 0: the point with itself (R2 = 0)
          row += near_kernel<KID>(0.0, kp, tab, ln2_hi, ln2_lo) * qi;
          s = 1;
        }
        const int s_end = diag ? 32 : 64;
#pragma unroll 2
        for (; s < s_end; s++) {
          const int j = cb + ((t + s) & 63);
          const double dx = xi - bx[j], dy = yi - by[j];
          const double v = near_kernel<KID>(fma(dx, dx, dy * dy), kp, tab, ln2_hi, ln2_lo);
          row = fma(v, bq[j], row);
          atomicAdd(&bc[j], v * qi);
        }
        if (diag && t < 32) {
          const int j = cb + t + 32;
          const double dx = xi - bx[j], dy = yi - by[j];
          const double v = near_kernel<KID>(fma(dx, dx, dy * dy), kp, tab, ln2_hi, ln2_lo);
          row = fma(v, bq[j], row);
          atomicAdd(&bc[j], v * qi);
        }
      }
      if (self) {
        // B = A: the source-side sums of this chunk arrive from chunks <= ca,
        // all of which precede it in this loop
        if (vi) out[i] = row;
      } else {
        if (vi) out[i] = row;
      }
    }
    __syncthreads();
    const double scale = KID == K_LOG ? 0.5 : 1.0;
    if (self) {
      for (int i = t; i < na; i += 64) out[i] = scale * (out[i] + bc[i]);
    } else {
      for (int i = t; i < na; i += 64) out[i] = scale * out[i];
      for (int j = t; j < nb; j += 64) out[na + j] = scale * bc[j];
    }
  }
}

// Leaves: y[perm[i]] = (L2P_k in_k)[i] + the near-field partials of the
// leaf's self pair and its neighbour slots 0..7, in that order.
__global__ __launch_bounds__(64) void k_out(const double* __restrict__ l2p_cm,
                                            const int64_t* __restrict__ l2poff,
                                            const int* __restrict__ rank,
                                            const int64_t* __restrict__ soff,
                                            const int* __restrict__ leaf_start,
                                            const int* __restrict__ leaf_count,
                                            const double* __restrict__ inv,
                                            const double* __restrict__ near_part,
                                            const int* __restrict__ slot_part,
                                            const int* __restrict__ perm,
                                            double* __restrict__ yout, int64_t gbase, int nb) {
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  const int st = leaf_start[k], n = leaf_count[k];
  const double* P = l2p_cm + l2poff[g];
  const double* x = inv + soff[g];
  int so[9];
#pragma unroll
  for (int s = 0; s < 9; s++) so[s] = slot_part[9 * k + s];
  for (int i = threadIdx.x; i < n; i += 64) {
    double acc = 0.0;
#pragma unroll 8
    for (int c = 0; c < r; c++) acc += P[i + (int64_t)c * n] * x[c];
    double nf = 0.0;
#pragma unroll
    for (int s = 0; s < 9; s++)
      if (so[s] >= 0) nf += near_part[so[s] + i];
    yout[perm[st + i]] = acc + nf;
  }
}

// Dense direct sum over all points for a set of rows, split over column
// chunks into partial sums: part[chunk * nrows + row].
template <int KID>
__global__ __launch_bounds__(256) void k_direct(const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ qs,
                                                const int* __restrict__ rows, int nrows, int n,
                                                int chunk, double* __restrict__ part, KParams kp) {
  __shared__ double sx[256], sy[256], sq[256];
  const int rr = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * chunk;
  const int c1 = min(n, c0 + chunk);
  const int row = rr < nrows ? rows[rr] : 0;
  const double xi = x[row], yi = y[row];
  double acc = 0.0;
  for (int cb = c0; cb < c1; cb += 256) {
    const int cn = min(256, c1 - cb);
    __syncthreads();
    if ((int)threadIdx.x < cn) {
      sx[threadIdx.x] = x[cb + threadIdx.x];
      sy[threadIdx.x] = y[cb + threadIdx.x];
      sq[threadIdx.x] = qs[cb + threadIdx.x];
    }
    __syncthreads();
    for (int q = 0; q < cn; q++) acc += kernel_xy<KID>(xi, yi, sx[q], sy[q], kp) * sq[q];
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

void gather_sorted(Engine& e, const double* d_b, double* d_qs, hipStream_t s) {
  k_gather<<<ceil_div(e.n, 256), 256, 0, s>>>(d_b, e.perm.as<int>(), d_qs, e.n);
  AFMM_LAUNCH_CHECK();
}

void matvec_device(Engine& e, const double* d_b, double* d_y, hipStream_t s) {
  const int L = e.L;
  const int nleaf = 1 << (2 * L);
  e.qs.alloc(sizeof(double) * e.n);
  hipEvent_t a, b;
  const int64_t* soff = e.soff_g.as<int64_t>();
  const int64_t* l2poff = e.l2poff_g.as<int64_t>();
  const int* rank = e.rank_g.as<int>();
  const int* nin = e.nin_g.as<int>();

  e.prof_begin("gather", s, &a, &b);
  gather_sorted(e, d_b, e.qs.as<double>(), s);
  e.prof_end(s, b);

  // upward pass
  e.prof_begin("m2m", s, &a, &b);
  k_m2m<<<nleaf, 256, 0, s>>>(e.l2p_rm.as<double>(), l2poff, rank, nin, soff, e.qs.as<double>(),
                              nullptr, e.d_leaf_start.as<int>(), e.outv.as<double>(),
                              e.lv[L].gbase, nleaf);
  AFMM_LAUNCH_CHECK();
  for (int j = L - 1; j >= 2; j--) {
    const LevelInfo& li = e.lv[j];
    k_m2m<<<li.nb, 256, 0, s>>>(e.l2p_rm.as<double>(), l2poff, rank, nin, soff,
                                e.outv.as<double>(), soff + e.lv[j + 1].gbase, nullptr,
                                e.outv.as<double>(), li.gbase, li.nb);
    AFMM_LAUNCH_CHECK();
  }
  e.prof_end(s, b);

  // fork the near field onto the auxiliary stream: it overlaps the
  // memory-bound M2L and L2L, not the short latency-bound M2M chain
  hipStream_t sn = e.near_overlap ? e.s_aux : s;
  if (e.near_overlap) {
    AFMM_HIP(hipEventRecord(e.ev_fork, s));
    AFMM_HIP(hipStreamWaitEvent(e.s_aux, e.ev_fork, 0));
  }
  e.prof_begin("p2p", sn, &a, &b);
  {
    const int maxpad = (e.max_leaf + 63) & ~63;
    const size_t shm = sizeof(double) * (256 + 4 * (size_t)maxpad);
    const unsigned grid = (unsigned)std::min<int64_t>(e.near_upairs, 16384);
    AFMM_DISPATCH_KERNEL(e.kid, KID,
                         k_p2p<KID><<<grid, 64, shm, sn>>>(
                             e.x.as<double>(), e.y.as<double>(), e.qs.as<double>(),
                             e.d_leaf_start.as<int>(), e.d_leaf_count.as<int>(),
                             e.near_pair.as<int>(), e.near_pair_part.as<int>(),
                             e.near_part.as<double>(), (int)e.near_upairs, maxpad, e.kp,
                             e.log_tab.as<double2>(), e.ln2_hi, e.ln2_lo));
    AFMM_LAUNCH_CHECK();
  }
  e.prof_end(sn, b);
  if (e.near_overlap) AFMM_HIP(hipEventRecord(e.ev_join, e.s_aux));



  // M2L over every level at once: pair blocks, then per-box reduction
  const int nbox = (int)(e.lv[L].gbase + nleaf);
  const int rcap = std::max(e.max_rank, 1);
  const size_t shm = sizeof(double) * ((size_t)e.pair_tile + 10 * (size_t)rcap);
  e.prof_begin("m2l", s, &a, &b);
  if (e.npairs > 0) {
    // pipelined kernel when every block fits the registers of a workgroup
    // (ranks up to 256 for the outv staging), chunked kernel otherwise
    const int need = (e.pair_tile + 511) / 512;
    const unsigned grid = (unsigned)std::min<int64_t>(e.npairs, (int64_t)e.m2l_blocks_per_cu * e.num_cus);
    const PairDesc* pd = e.pair_desc.as<PairDesc>();
    double* part = e.m2l_part.as<double>();
    const double* outv = e.outv.as<double>();
    const bool ok = e.max_rank <= 256;
    const double* B = e.m2l.as<double>();
    const int np = (int)e.npairs;
#define AFMM_PIPE(NVV) k_m2l_pipe<NVV><<<grid, 256, shm, s>>>(B, pd, outv, part, np, e.pair_tile, rcap)
    if (ok && need <= 2) AFMM_PIPE(2);
    else if (ok && need <= 3) AFMM_PIPE(3);
    else if (ok && need <= 4) AFMM_PIPE(4);
    else if (ok && need <= 5) AFMM_PIPE(5);
    else if (ok && need <= 6) AFMM_PIPE(6);
    else if (ok && need <= 8) AFMM_PIPE(8);
    else if (ok && need <= 12) AFMM_PIPE(12);
    else
      k_m2l_pairs<<<(unsigned)e.npairs, 256, shm, s>>>(
          B, e.pair_k.as<int>(), e.pair_b.as<int>(), e.pair_off.as<int64_t>(),
          e.pair_part.as<int>(), rank, soff, outv, part, e.pair_tile, rcap);
#undef AFMM_PIPE
    AFMM_LAUNCH_CHECK();
  }
  e.prof_end(s, b);
  e.prof_begin("m2l_reduce", s, &a, &b);
  k_m2l_reduce<<<nbox, 64, 0, s>>>(e.m2l_part.as<double>(), e.red_start.as<int>(),
                                   e.red_part.as<int>(), rank, soff, e.inv.as<double>(), nbox);
  AFMM_LAUNCH_CHECK();
  e.prof_end(s, b);

  // downward pass
  e.prof_begin("l2l", s, &a, &b);
  for (int j = 2; j < L; j++) {
    const LevelInfo& li = e.lv[j];
    k_l2l<<<li.nb, 256, 0, s>>>(e.l2p_cm.as<double>(), l2poff, rank, nin, soff,
                                soff + e.lv[j + 1].gbase, e.inv.as<double>(), li.gbase, li.nb);
    AFMM_LAUNCH_CHECK();
  }
  e.prof_end(s, b);

  if (e.near_overlap) AFMM_HIP(hipStreamWaitEvent(s, e.ev_join, 0));
  e.prof_begin("l2p_out", s, &a, &b);
  k_out<<<nleaf, 64, 0, s>>>(e.l2p_cm.as<double>(), l2poff, rank, soff, e.d_leaf_start.as<int>(),
                             e.d_leaf_count.as<int>(), e.inv.as<double>(),
                             e.near_part.as<double>(), e.near_slot_part.as<int>(),
                             e.perm.as<int>(), d_y, e.lv[L].gbase, nleaf);
  AFMM_LAUNCH_CHECK();
  e.prof_end(s, b);
}

void direct_rows_device(Engine& e, int nrows, const int* d_rows, const double* d_qs,
                        double* d_out, hipStream_t s) {
  if (nrows <= 0) return;
  const int rb = (int)ceil_div(nrows, 256);
  // enough workgroups to fill the device
  int nchunk = std::max(1, std::min(4096 / rb, (int)ceil_div(e.n, 4096)));
  const int chunk = (int)ceil_div(e.n, nchunk);
  nchunk = (int)ceil_div(e.n, chunk);
  DevBuf part;
  part.alloc(sizeof(double) * (size_t)nrows * nchunk);
  dim3 grid(rb, nchunk);
  AFMM_DISPATCH_KERNEL(e.kid, KID,
                       k_direct<KID><<<grid, 256, 0, s>>>(e.x.as<double>(), e.y.as<double>(), d_qs,
                                                          d_rows, nrows, e.n, chunk,
                                                          part.as<double>(), e.kp));
  AFMM_LAUNCH_CHECK();
  k_sum_parts<<<ceil_div(nrows, 256), 256, 0, s>>>(part.as<double>(), nrows, nchunk, d_out);
  AFMM_LAUNCH_CHECK();
  AFMM_HIP(hipStreamSynchronize(s));
}

}  // namespace afmm
