// matvec.hip -- the matvec passes of AFMM2D on the GPU (FP64).
//
// Reference flow (AFMM2D.cpp:46-65, AFMM2DTree.cpp:1120-1225): gather q into
// KD order, M2M upward (levels L..2, out_k = L2P_k^T src), M2L (in_k =
// sum over the interaction list of M2L[k][kIL] out_kIL), L2L downward
// (children.in += L2P_k in_k; leaves: potential = L2P_k in_k), near field
// (potential += K(leaf, neighbour) q_neighbour for 8 neighbours and self),
// scatter back to caller order.
//
// Stream plan: the near field depends only on the gathered q, so it runs on
// a second stream concurrently with the HBM-bound M2M/M2L/L2L chain; the
// last kernel joins both (leaf L2P + near field + scatter, fused).
#include "engine.hpp"

namespace afmm {

namespace {

__global__ void k_gather(const double* __restrict__ b, const int* __restrict__ perm,
                         double* __restrict__ qs, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) qs[i] = b[perm[i]];
}

// One wave per box: out[c] = sum_i L2P[i][c] src[i] (row-major L2P, lanes
// over c, coalesced along each row).  src is the leaf's q segment (leaf
// level) or the children's concatenated multipoles (always contiguous).
__global__ __launch_bounds__(64) void k_m2m(const double* __restrict__ l2p_rm,
                                            const int64_t* __restrict__ l2poff,
                                            const int* __restrict__ rank,
                                            const int* __restrict__ nin,
                                            const int64_t* __restrict__ soff,
                                            const double* __restrict__ src,
                                            const int64_t* __restrict__ src_off_child,
                                            const int* __restrict__ src_off_leaf,
                                            double* __restrict__ outv, int64_t gbase, int nb) {
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  if (r == 0) return;
  const int n = nin[g];
  const double* P = l2p_rm + l2poff[g];
  const double* x = src + (src_off_leaf ? (int64_t)src_off_leaf[k] : src_off_child[4 * k]);
  for (int c = threadIdx.x; c < r; c += 64) {
    double acc = 0.0;
#pragma unroll 8
    for (int i = 0; i < n; i++) acc += P[(int64_t)i * r + c] * x[i];
    outv[soff[g] + c] = acc;
  }
}

// M2L: one wave per target box over all levels.  The box's blocks form one
// rank x C column-major matrix; a wave reads G = floor(64/rank) columns per
// load (lane l: row l % rank, column group l / rank) so each load moves
// G*rank contiguous doubles.  Source coefficients are staged in LDS.
constexpr int M2L_CHUNK = 2048;

__global__ __launch_bounds__(64) void k_m2l(const double* __restrict__ m2l,
                                            const int* __restrict__ m2l_src,
                                            const int64_t* __restrict__ m2loff,
                                            const int64_t* __restrict__ m2lcoff,
                                            const int* __restrict__ rank,
                                            const int64_t* __restrict__ soff,
                                            const double* __restrict__ outv,
                                            double* __restrict__ inv, int nbox) {
  __shared__ double xs[M2L_CHUNK];
  __shared__ double red[64];
  const int g = blockIdx.x;
  if (g >= nbox) return;
  const int r = rank[g];
  if (r == 0) return;
  const int lane = threadIdx.x;
  const int64_t c0 = m2lcoff[g];
  const int C = (int)(m2lcoff[g + 1] - c0);
  const double* B = m2l + m2loff[g];
  double* out = inv + soff[g];
  if (r <= 64) {
    const int G = 64 / r;
    const int row = lane % r, grp = lane / r;
    const bool active = grp < G;
    double acc = 0.0;
    for (int cb = 0; cb < C; cb += M2L_CHUNK) {
      const int cn = min(M2L_CHUNK, C - cb);
      __syncthreads();
      for (int c = lane; c < cn; c += 64) xs[c] = outv[m2l_src[c0 + cb + c]];
      __syncthreads();
      if (active) {
        const double* Bc = B + (int64_t)cb * r;
        int c = grp;
#pragma unroll 4
        for (; c + 3 * G < cn; c += 4 * G) {
          double b0 = Bc[(int64_t)c * r + row];
          double b1 = Bc[(int64_t)(c + G) * r + row];
          double b2 = Bc[(int64_t)(c + 2 * G) * r + row];
          double b3 = Bc[(int64_t)(c + 3 * G) * r + row];
          acc += b0 * xs[c];
          acc += b1 * xs[c + G];
          acc += b2 * xs[c + 2 * G];
          acc += b3 * xs[c + 3 * G];
        }
        for (; c < cn; c += G) acc += Bc[(int64_t)c * r + row] * xs[c];
      }
    }
    red[lane] = active ? acc : 0.0;
    __syncthreads();
    if (lane < r) {
      double s = 0.0;
      for (int q = 0; q < G; q++) s += red[q * r + lane];
      out[lane] = s;
    }
  } else {
    for (int rb = 0; rb < r; rb += 64) {
      const int row = rb + lane;
      double acc = 0.0;
      for (int cb = 0; cb < C; cb += M2L_CHUNK) {
        const int cn = min(M2L_CHUNK, C - cb);
        __syncthreads();
        for (int c = lane; c < cn; c += 64) xs[c] = outv[m2l_src[c0 + cb + c]];
        __syncthreads();
        if (row < r) {
          const double* Bc = B + (int64_t)cb * r;
          for (int c = 0; c < cn; c++) acc += Bc[(int64_t)c * r + row] * xs[c];
        }
      }
      if (row < r) out[row] = acc;
    }
  }
}

// L2L for a non-leaf level: children.in += L2P_k in_k (column-major L2P,
// lanes over rows).  The children's coefficients are contiguous.
__global__ __launch_bounds__(64) void k_l2l(const double* __restrict__ l2p_cm,
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
  for (int i = threadIdx.x; i < n; i += 64) {
    double acc = 0.0;
#pragma unroll 8
    for (int c = 0; c < r; c++) acc += P[i + (int64_t)c * n] * x[c];
    dst[i] += acc;
  }
}

// Near field: one workgroup of 64 per target leaf; the 8 neighbours and the
// leaf itself are staged through LDS and the kernel is evaluated on the fly.
constexpr int P2P_STAGE = 512;

template <int KID>
__global__ __launch_bounds__(64) void k_p2p(const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const double* __restrict__ qs,
                                            const int* __restrict__ leaf_start,
                                            const int* __restrict__ leaf_count,
                                            const int* __restrict__ nbr,
                                            double* __restrict__ near, int nleaf, KParams kp) {
  __shared__ double sx[P2P_STAGE], sy[P2P_STAGE], sq[P2P_STAGE];
  const int k = blockIdx.x;
  if (k >= nleaf) return;
  const int st = leaf_start[k], cnt = leaf_count[k];
  const int lane = threadIdx.x;
  for (int tb = 0; tb < cnt; tb += 64) {
    const int i = tb + lane;
    const double xi = i < cnt ? x[st + i] : 0.0;
    const double yi = i < cnt ? y[st + i] : 0.0;
    double acc = 0.0;
    for (int s = 0; s < 9; s++) {
      const int kb = s < 8 ? nbr[8 * k + s] : k;
      if (kb < 0) continue;
      const int sb = leaf_start[kb], m = leaf_count[kb];
      for (int mb = 0; mb < m; mb += P2P_STAGE) {
        const int mm = min(P2P_STAGE, m - mb);
        __syncthreads();
        for (int q = lane; q < mm; q += 64) {
          sx[q] = x[sb + mb + q];
          sy[q] = y[sb + mb + q];
          sq[q] = qs[sb + mb + q];
        }
        __syncthreads();
        for (int q = 0; q < mm; q++) {
          double r2 = (xi - sx[q]) * (xi - sx[q]) + (yi - sy[q]) * (yi - sy[q]);
          acc += kernel_r2<KID>(r2, kp) * sq[q];
        }
      }
    }
    if (i < cnt) near[st + i] = acc;
  }
}

// Leaves: y[perm[i]] = (L2P_k in_k)[i] + near[i].
__global__ __launch_bounds__(64) void k_out(const double* __restrict__ l2p_cm,
                                            const int64_t* __restrict__ l2poff,
                                            const int* __restrict__ rank,
                                            const int64_t* __restrict__ soff,
                                            const int* __restrict__ leaf_start,
                                            const int* __restrict__ leaf_count,
                                            const double* __restrict__ inv,
                                            const double* __restrict__ near,
                                            const int* __restrict__ perm,
                                            double* __restrict__ yout, int64_t gbase, int nb) {
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int64_t g = gbase + k;
  const int r = rank[g];
  const int st = leaf_start[k], n = leaf_count[k];
  const double* P = l2p_cm + l2poff[g];
  const double* x = inv + soff[g];
  for (int i = threadIdx.x; i < n; i += 64) {
    double acc = 0.0;
#pragma unroll 8
    for (int c = 0; c < r; c++) acc += P[i + (int64_t)c * n] * x[c];
    yout[perm[st + i]] = acc + near[st + i];
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
  e.pot_near.alloc(sizeof(double) * e.n);
  hipEvent_t a, b;
  const int64_t* soff = e.soff_g.as<int64_t>();
  const int64_t* l2poff = e.l2poff_g.as<int64_t>();
  const int* rank = e.rank_g.as<int>();
  const int* nin = e.nin_g.as<int>();

  e.prof_begin("gather", s, &a, &b);
  gather_sorted(e, d_b, e.qs.as<double>(), s);
  e.prof_end(s, b);

  // fork the near field onto the auxiliary stream
  AFMM_HIP(hipEventRecord(e.ev_fork, s));
  AFMM_HIP(hipStreamWaitEvent(e.s_aux, e.ev_fork, 0));
  e.prof_begin("p2p", e.s_aux, &a, &b);
  AFMM_DISPATCH_KERNEL(e.kid, KID,
                       k_p2p<KID><<<nleaf, 64, 0, e.s_aux>>>(
                           e.x.as<double>(), e.y.as<double>(), e.qs.as<double>(),
                           e.d_leaf_start.as<int>(), e.d_leaf_count.as<int>(),
                           e.d_leaf_nbr.as<int>(), e.pot_near.as<double>(), nleaf, e.kp));
  AFMM_LAUNCH_CHECK();
  e.prof_end(e.s_aux, b);
  AFMM_HIP(hipEventRecord(e.ev_join, e.s_aux));

  // upward pass
  e.prof_begin("m2m", s, &a, &b);
  k_m2m<<<nleaf, 64, 0, s>>>(e.l2p_rm.as<double>(), l2poff, rank, nin, soff, e.qs.as<double>(),
                             nullptr, e.d_leaf_start.as<int>(), e.outv.as<double>(),
                             e.lv[L].gbase, nleaf);
  AFMM_LAUNCH_CHECK();
  for (int j = L - 1; j >= 2; j--) {
    const LevelInfo& li = e.lv[j];
    k_m2m<<<li.nb, 64, 0, s>>>(e.l2p_rm.as<double>(), l2poff, rank, nin, soff,
                               e.outv.as<double>(), soff + e.lv[j + 1].gbase, nullptr,
                               e.outv.as<double>(), li.gbase, li.nb);
    AFMM_LAUNCH_CHECK();
  }
  e.prof_end(s, b);

  // M2L over every level at once
  const int nbox = (int)(e.lv[L].gbase + nleaf);
  e.prof_begin("m2l", s, &a, &b);
  k_m2l<<<nbox, 64, 0, s>>>(e.m2l.as<double>(), e.m2l_src.as<int>(), e.m2loff_g.as<int64_t>(),
                            e.m2lcoff_g.as<int64_t>(), rank, soff, e.outv.as<double>(),
                            e.inv.as<double>(), nbox);
  AFMM_LAUNCH_CHECK();
  e.prof_end(s, b);

  // downward pass
  e.prof_begin("l2l", s, &a, &b);
  for (int j = 2; j < L; j++) {
    const LevelInfo& li = e.lv[j];
    k_l2l<<<li.nb, 64, 0, s>>>(e.l2p_cm.as<double>(), l2poff, rank, nin, soff,
                               soff + e.lv[j + 1].gbase, e.inv.as<double>(), li.gbase, li.nb);
    AFMM_LAUNCH_CHECK();
  }
  e.prof_end(s, b);

  AFMM_HIP(hipStreamWaitEvent(s, e.ev_join, 0));
  e.prof_begin("l2p_out", s, &a, &b);
  k_out<<<nleaf, 64, 0, s>>>(e.l2p_cm.as<double>(), l2poff, rank, soff, e.d_leaf_start.as<int>(),
                             e.d_leaf_count.as<int>(), e.inv.as<double>(),
                             e.pot_near.as<double>(), e.perm.as<int>(), d_y, e.lv[L].gbase, nleaf);
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
