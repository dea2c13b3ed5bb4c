// kd_sort.hip -- KD partition of the points on the GPU.
//
// The reference (KDTree.cpp:91-142) visits every box of levels 0..L: it
// merge-sorts the box's points by x, splits at floor(n/2), merge-sorts each
// half by y and splits each half at floor(half/2), giving the children
// [LB, LT, RB, RT] = quad-tree children [0, 3, 1, 2] (KDTree.cpp:217-220).
// Its merge takes the right run on ties (KDTree.cpp:47), which orders equal
// keys by descending position in the input of that sort; the order here is
// therefore the total order (key ascending, position descending), which makes
// every sort deterministic and identical to the reference, ties included.
//
// All boxes of one level are sorted by one segmented bitonic network: every
// segment is padded virtually to the same power of two P with +inf keys.  The
// network uses the min-to-lower-index form (first step of each stage compares
// i with its mirror), so padding never moves and compares against it are
// skipped.  Stages up to the LDS tile run inside one workgroup; longer
// distances run one global step per launch.
#include <climits>

#include "engine.hpp"

namespace afmm {

namespace {

constexpr int TILE = 2048;  // elements per LDS tile (12 B each, 24 KiB)

__device__ __forceinline__ bool before(double ka, int ia, double kb, int ib) {
  return ka < kb || (ka == kb && ia > ib);
}

// Index of pair t for a half-cleaner of distance j (log2 lj): insert a zero
// bit at position lj.
__device__ __forceinline__ int pair_lo(int t, int lj) {
  return ((t >> lj) << (lj + 1)) | (t & ((1 << lj) - 1));
}

__global__ void k_init_keys(const double* __restrict__ src, double* __restrict__ key,
                            int* __restrict__ id, int n) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    key[p] = src[p];
    id[p] = p;
  }
}

// Loads one tile (tile elements of one segment's virtual space) into LDS,
// runs either the full sort of the tile (stages 2..tile) or, when
// merge_stage > 0, the half-cleaner steps j = tile/2..1 of that stage.
__global__ __launch_bounds__(TILE / 2) void k_tile(double* __restrict__ key, int* __restrict__ id,
                                                   const int* __restrict__ seg_off, int P, int tile,
                                                   int merge_stage) {
  __shared__ double sk[TILE];
  __shared__ int si[TILE];
  const int tiles_per_seg = P / tile;
  const int seg = blockIdx.x / tiles_per_seg;
  const int tbase = (blockIdx.x % tiles_per_seg) * tile;
  const int base = seg_off[seg];
  const int n = seg_off[seg + 1] - base;
  for (int v = threadIdx.x; v < tile; v += blockDim.x) {
    int g = tbase + v;
    if (g < n) {
      sk[v] = key[base + g];
      si[v] = id[base + g];
    } else {
      sk[v] = __builtin_inf();
      si[v] = INT_MIN;
    }
  }
  __syncthreads();
  const int half = tile / 2;
  auto cx = [&](int i, int l) {
    double ki = sk[i], kl = sk[l];
    int ii = si[i], il = si[l];
    if (before(kl, il, ki, ii)) {
      sk[i] = kl; si[i] = il;
      sk[l] = ki; si[l] = ii;
    }
  };
  if (merge_stage == 0) {
    for (int kk = 2; kk <= tile; kk <<= 1) {
      const int lh = __builtin_ctz(kk >> 1);
      for (int t = threadIdx.x; t < half; t += blockDim.x) {
        int i = pair_lo(t, lh);
        int l = i ^ (kk - 1);
        cx(i, l);
      }
      __syncthreads();
      for (int j = kk >> 2; j > 0; j >>= 1) {
        const int lj = __builtin_ctz(j);
        for (int t = threadIdx.x; t < half; t += blockDim.x) {
          int i = pair_lo(t, lj);
          cx(i, i | j);
        }
        __syncthreads();
      }
    }
  } else {
    for (int j = tile >> 1; j > 0; j >>= 1) {
      const int lj = __builtin_ctz(j);
      for (int t = threadIdx.x; t < half; t += blockDim.x) {
        int i = pair_lo(t, lj);
        cx(i, i | j);
      }
      __syncthreads();
    }
  }
  for (int v = threadIdx.x; v < tile; v += blockDim.x) {
    int g = tbase + v;
    if (g < n) {
      key[base + g] = sk[v];
      id[base + g] = si[v];
    }
  }
}

// One global compare-exchange step over every segment: the mirror step of
// stage kk (mirror = 1) or a half-cleaner of distance j.
__global__ void k_step(double* __restrict__ key, int* __restrict__ id,
                       const int* __restrict__ seg_off, int nseg, int P, int kk, int j, int mirror) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t pairs = (int64_t)nseg * (P / 2);
  if (t >= pairs) return;
  const int seg = (int)(t / (P / 2));
  const int tt = (int)(t % (P / 2));
  const int base = seg_off[seg];
  const int n = seg_off[seg + 1] - base;
  int i, l;
  if (mirror) {
    i = pair_lo(tt, __builtin_ctz(kk >> 1));
    l = i ^ (kk - 1);
  } else {
    i = pair_lo(tt, __builtin_ctz(j));
    l = i | j;
  }
  if (l >= n) return;
  double ki = key[base + i], kl = key[base + l];
  int ii = id[base + i], il = id[base + l];
  if (before(kl, il, ki, ii)) {
    key[base + i] = kl; id[base + i] = il;
    key[base + l] = ki; id[base + l] = ii;
  }
}

__global__ void k_apply(const int* __restrict__ id, const double* __restrict__ sx,
                        const double* __restrict__ sy, const int* __restrict__ sp,
                        double* __restrict__ dx, double* __restrict__ dy, int* __restrict__ dp,
                        int n) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    int s = id[p];
    dx[p] = sx[s];
    dy[p] = sy[s];
    dp[p] = sp[s];
  }
}

__global__ void k_iota(int* __restrict__ a, int n) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) a[p] = p;
}

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Sorts every segment [off[s], off[s+1]) of (key, id) by (key asc, id desc).
void segmented_sort(double* key, int* id, const std::vector<int>& off, DevBuf& d_off,
                    hipStream_t s) {
  const int nseg = (int)off.size() - 1;
  int maxn = 1;
  for (int q = 0; q < nseg; q++) maxn = std::max(maxn, off[q + 1] - off[q]);
  const int P = next_pow2(std::max(maxn, 2));
  d_off.alloc(sizeof(int) * off.size());
  AFMM_HIP(hipMemcpyAsync(d_off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, s));
  AFMM_HIP(hipStreamSynchronize(s));  // off is a host temporary
  const int tile = std::min(P, TILE);
  const unsigned nblk = (unsigned)((int64_t)nseg * (P / tile));
  const int thr = std::max(tile / 2, 64);
  k_tile<<<nblk, thr, 0, s>>>(key, id, d_off.as<int>(), P, tile, 0);
  AFMM_LAUNCH_CHECK();
  const int64_t pairs = (int64_t)nseg * (P / 2);
  const unsigned gblk = ceil_div(pairs, 256);
  for (int kk = 2 * tile; kk <= P; kk <<= 1) {
    k_step<<<gblk, 256, 0, s>>>(key, id, d_off.as<int>(), nseg, P, kk, 0, 1);
    AFMM_LAUNCH_CHECK();
    for (int j = kk >> 2; j >= tile; j >>= 1) {
      k_step<<<gblk, 256, 0, s>>>(key, id, d_off.as<int>(), nseg, P, kk, j, 0);
      AFMM_LAUNCH_CHECK();
    }
    k_tile<<<nblk, thr, 0, s>>>(key, id, d_off.as<int>(), P, tile, kk);
    AFMM_LAUNCH_CHECK();
  }
}

}  // namespace

void kd_partition(Engine& e, const double* h_x, const double* h_y) {
  const int n = e.n, L = e.L;
  hipStream_t s = e.s_aux;
  DevBuf bx[2], by[2], bp[2], key, id, d_off;
  for (int q = 0; q < 2; q++) {
    bx[q].alloc(sizeof(double) * n);
    by[q].alloc(sizeof(double) * n);
    bp[q].alloc(sizeof(int) * n);
  }
  key.alloc(sizeof(double) * n);
  id.alloc(sizeof(int) * n);
  AFMM_HIP(hipMemcpyAsync(bx[0].p, h_x, sizeof(double) * n, hipMemcpyHostToDevice, s));
  AFMM_HIP(hipMemcpyAsync(by[0].p, h_y, sizeof(double) * n, hipMemcpyHostToDevice, s));
  const unsigned nblk = ceil_div(n, 256);
  k_iota<<<nblk, 256, 0, s>>>(bp[0].as<int>(), n);
  AFMM_LAUNCH_CHECK();

  // Segment sizes of every level (KD order): a box of n points has children
  // LB = floor(floor(n/2)/2), LT, RB, RT (KDTree.cpp:96-114).
  std::vector<int> sizes{n};
  int cur = 0;
  for (int d = 0; d <= L && n > 1; d++) {
    // x-sort of the level-d boxes
    std::vector<int> off(sizes.size() + 1, 0);
    for (size_t q = 0; q < sizes.size(); q++) off[q + 1] = off[q] + sizes[q];
    k_init_keys<<<nblk, 256, 0, s>>>(bx[cur].as<double>(), key.as<double>(), id.as<int>(), n);
    AFMM_LAUNCH_CHECK();
    segmented_sort(key.as<double>(), id.as<int>(), off, d_off, s);
    k_apply<<<nblk, 256, 0, s>>>(id.as<int>(), bx[cur].as<double>(), by[cur].as<double>(),
                                 bp[cur].as<int>(), bx[1 - cur].as<double>(),
                                 by[1 - cur].as<double>(), bp[1 - cur].as<int>(), n);
    AFMM_LAUNCH_CHECK();
    cur = 1 - cur;
    // y-sort of each half
    std::vector<int> hoff(2 * sizes.size() + 1, 0);
    std::vector<int> next;
    next.reserve(4 * sizes.size());
    for (size_t q = 0; q < sizes.size(); q++) {
      int m = sizes[q], nl = m / 2, nr = m - nl;
      hoff[2 * q + 1] = hoff[2 * q] + nl;
      hoff[2 * q + 2] = hoff[2 * q + 1] + nr;
      next.push_back(nl / 2);
      next.push_back(nl - nl / 2);
      next.push_back(nr / 2);
      next.push_back(nr - nr / 2);
    }
    k_init_keys<<<nblk, 256, 0, s>>>(by[cur].as<double>(), key.as<double>(), id.as<int>(), n);
    AFMM_LAUNCH_CHECK();
    segmented_sort(key.as<double>(), id.as<int>(), hoff, d_off, s);
    k_apply<<<nblk, 256, 0, s>>>(id.as<int>(), bx[cur].as<double>(), by[cur].as<double>(),
                                 bp[cur].as<int>(), bx[1 - cur].as<double>(),
                                 by[1 - cur].as<double>(), bp[1 - cur].as<int>(), n);
    AFMM_LAUNCH_CHECK();
    cur = 1 - cur;
    if (d < L) sizes.swap(next);
  }
  // sizes now holds the 4^L leaf counts in KD order
  e.kd_counts = sizes;
  e.x.alloc(sizeof(double) * n);
  e.y.alloc(sizeof(double) * n);
  e.perm.alloc(sizeof(int) * n);
  AFMM_HIP(hipMemcpyAsync(e.x.p, bx[cur].p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  AFMM_HIP(hipMemcpyAsync(e.y.p, by[cur].p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  AFMM_HIP(hipMemcpyAsync(e.perm.p, bp[cur].p, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
  std::vector<int> hp(n);
  std::vector<double> hx(n), hy(n);
  AFMM_HIP(hipMemcpyAsync(hp.data(), bp[cur].p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  AFMM_HIP(hipMemcpyAsync(hx.data(), bx[cur].p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  AFMM_HIP(hipMemcpyAsync(hy.data(), by[cur].p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  AFMM_HIP(hipStreamSynchronize(s));
  e.h_perm.assign(hp.begin(), hp.end());
  e.h_xy.resize(2 * (size_t)n);
  for (int i = 0; i < n; i++) {
    e.h_xy[2 * i] = hx[i];
    e.h_xy[2 * i + 1] = hy[i];
  }
}

}  // namespace afmm
