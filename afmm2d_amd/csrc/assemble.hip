// assemble.hip -- operator assembly on the GPU: per-box ACA skeletons, the
// interpolative L2P basis and the M2L blocks.
//
// Levels run L -> 2 (AFMM2DTree.cpp:991-995).  At level j the rows of box k
// are its leaf points (j = L) or the concatenated skeletons of its children
// in child order 0..3; the columns are the rows of its interaction-list boxes,
// inner slots 0..15 then outer slots 0..23 (AFMM2DTree.cpp:997-1034).  Since
// children 4k..4k+3 are consecutive and skeletons are stored box after box,
// every box's rows are one contiguous range of the level's source points.
//
// ACA (ACA.cpp:27-133): one workgroup per box (256 threads, 1024 on levels
// with few boxes), the rank loop is sequential,
// each step's row/column evaluation, residual update, argmax and norm
// reductions are parallel over the workgroup.  Pivot choice and stopping rule
// follow the reference exactly (first maximum, |prev - normS| >= tol*prev,
// normS^2 += (|u||v|)^2 + sum_l 2|U_l.u||V_l.v| including the new pair,
// rank-2 quirk).  Products and differences are not contracted into FMAs so
// residuals round as the CPU oracle's do.
//
// ID (AFMM2DTree.cpp:1042-1060): L2P = Ac D^{-1} with D = K(skeleton rows,
// skeleton columns), through LU with partial pivoting of D^T.  Row skel_q of
// Ac is row q of D, so row skel_q of L2P is e_q (the LU solve returns it up to
// rounding amplified by cond(D)); only the other rows are solved and stored,
// and the skeleton rows are applied as the identity.
//
// M2L (AFMM2DTree.cpp:1070-1097): the reference builds K(skeleton_k,
// skeleton_kIL) for every ordered pair; here each unordered pair is built
// once, B = K(skeleton_first, skeleton_second) column-major, in the
// orientation whose register-tile class is cheaper, and the matvec applies
// B and +-B^T (the kernel is symmetric, or the antisymmetric part of one).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "engine.hpp"

namespace afmm {

namespace {

constexpr int ACA_THREADS = 256;
constexpr int MAXR = 1024;      // largest ACA rank a workgroup can hold
constexpr int MAXOWN = 64;      // owned rows/cols per thread (mask bits)
constexpr int RLDS = 84;        // ID matrices up to 84x84 live in LDS (< 64 KiB)

struct Cand {
  double a;  // |value|
  double v;  // value
  int i;     // index, -1 = none
};

__device__ __forceinline__ Cand cbetter(const Cand& x, const Cand& y) {
  if (y.i < 0) return x;
  if (x.i < 0) return y;
  return (y.a > x.a || (y.a == x.a && y.i < x.i)) ? y : x;
}

__device__ __forceinline__ Cand wave_argmax(Cand c) {
  for (int off = 32; off > 0; off >>= 1) {
    Cand o;
    o.a = __shfl_xor(c.a, off, 64);
    o.v = __shfl_xor(c.v, off, 64);
    o.i = __shfl_xor(c.i, off, 64);
    c = cbetter(c, o);
  }
  return c;
}

__device__ __forceinline__ double wave_sum(double s) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

// Workgroup argmax; every thread gets the result.
__device__ Cand block_argmax(Cand c, Cand* s4) {
  c = wave_argmax(c);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) s4[w] = c;
  __syncthreads();
  Cand r = s4[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) r = cbetter(r, s4[q]);
  __syncthreads();
  return r;
}

struct AcaArgs {
  const double* src_x;
  const double* src_y;
  const int* rstart;   // per box: first row in the level's source points
  const int* n1;       // per box: rows
  const int* n2;       // per box: columns
  const int* il;       // per box: 40 interaction slots (-1 absent)
  const int* boxes;    // boxes of this batch
  const int64_t* scr_off;  // per batch entry: scratch offset (doubles)
  double* scr;
  const int64_t* piv_off;  // per box: pivot output offset
  int* rank;
  int* prow;           // local row index of each pivot
  double* pcx;         // coordinates of each pivot column's point
  double* pcy;
  double tol;
  KParams kp;
  // learned capacities: per batch entry the rank the scratch holds (null:
  // min(rows, columns)); overflow[box] = 1 when a box reached it before its
  // stopping test passed (its result is then discarded and recomputed)
  const int* cap;
  int* overflow;
};

template <int KID, int NT>
__global__ __launch_bounds__(NT) void k_aca(AcaArgs a) {
#pragma clang fp contract(off)
  __shared__ double coefU[MAXR];
  __shared__ double coefV[MAXR];
  __shared__ double dU[MAXR + 1];
  __shared__ double dV[MAXR + 1];
  __shared__ int il_start[40], il_pref[41];
  __shared__ int nslot;
  __shared__ Cand s4[NT / 64];
  __shared__ double s_pt[4];
  __shared__ double s_norm[2];

  const int tid = threadIdx.x;
  const int b = a.boxes[blockIdx.x];
  const int n1 = a.n1[b], n2 = a.n2[b];
  const int mn = min(n1, n2);
  const int64_t po = a.piv_off[b];
  if (mn <= 0) {
    if (tid == 0) a.rank[b] = 0;
    return;
  }
  const int cmax = a.cap ? min(a.cap[blockIdx.x], mn) : mn;
  double* U = a.scr + a.scr_off[blockIdx.x];
  double* V = U + (int64_t)n1 * cmax;
  double* cx = V + (int64_t)n2 * cmax;
  double* cy = cx + n2;
  const int r0 = a.rstart[b];
  if (tid == 0) {
    int p = 0, c = 0;
    for (int s = 0; s < 40; s++) {
      int kb = a.il[40 * b + s];
      if (kb < 0) continue;
      il_start[c] = a.rstart[kb];
      il_pref[c] = p;
      p += a.n1[kb];
      c++;
    }
    il_pref[c] = p;
    nslot = c;
  }
  __syncthreads();
  // column points, each owned by thread (j mod NT)
  for (int j = tid; j < n2; j += NT) {
    int s = 0;
    while (s + 1 < nslot && il_pref[s + 1] <= j) s++;
    int p = il_start[s] + (j - il_pref[s]);
    cx[j] = a.src_x[p];
    cy[j] = a.src_y[p];
  }
  uint64_t cmask = 0, rmask = 0;  // used columns / rows owned by this thread

  // First pivot: first row whose largest entry exceeds tol (ACA.cpp:56-63).
  int istar = -1, jstar = -1;
  double piv = 0.0;
  for (int l = 0; l < n1; l++) {
    const double px = a.src_x[r0 + l], py = a.src_y[r0 + l];
    Cand c{0.0, 0.0, -1};
    for (int j = tid, m = 0; j < n2; j += NT, m++) {
      double v = kernel_xy<KID>(px, py, cx[j], cy[j], a.kp);
      V[j] = v;
      c = cbetter(c, Cand{fabs(v), v, j});
    }
    Cand r = block_argmax(c, s4);
    if (fabs(r.v) > a.tol) {
      istar = l;
      jstar = r.i;
      piv = r.v;
      break;
    }
  }
  if (istar < 0) {
    if (tid == 0) a.rank[b] = 0;
    return;
  }
  if (tid == (jstar & (NT - 1))) {
    s_pt[0] = cx[jstar];
    s_pt[1] = cy[jstar];
    cmask |= 1ull << (jstar / NT);
  }
  if (tid == (istar & (NT - 1))) rmask |= 1ull << (istar / NT);
  for (int j = tid; j < n2; j += NT) V[j] = V[j] / piv;
  __syncthreads();
  double cjx = s_pt[0], cjy = s_pt[1];
  if (tid == 0) {
    a.prow[po] = istar;
    a.pcx[po] = cjx;
    a.pcy[po] = cjy;
  }
  Cand cr{0.0, 0.0, -1};
  for (int i = tid, m = 0; i < n1; i += NT, m++) {
    double u = kernel_xy<KID>(a.src_x[r0 + i], a.src_y[r0 + i], cjx, cjy, a.kp);
    U[i] = u;
    if (!((rmask >> m) & 1)) cr = cbetter(cr, Cand{fabs(u), u, i});
  }
  int rank = 1;
  Cand nr = block_argmax(cr, s4);
  istar = nr.i;
  double normS = 0.0, prev = DBL_MAX;
  bool capped = false;

  while (fabs(prev - normS) >= a.tol * prev && rank < mn && istar >= 0) {
    if (rank >= cmax) {  // the ACA goes on beyond the capacity held
      capped = true;
      break;
    }
    const int t = rank;
    for (int l = tid; l < t; l += NT) coefU[l] = U[(int64_t)l * n1 + istar];
    __syncthreads();
    // Phase A: residual row of pivot row istar.
    const double px = a.src_x[r0 + istar], py = a.src_y[r0 + istar];
    double* Vt = V + (int64_t)t * n2;
    Cand c{0.0, 0.0, -1};
    for (int j = tid, m = 0; j < n2; j += NT, m++) {
      double r = kernel_xy<KID>(px, py, cx[j], cy[j], a.kp);
      for (int l = 0; l < t; l++) r -= coefU[l] * V[(int64_t)l * n2 + j];
      Vt[j] = r;
      if (!((cmask >> m) & 1)) c = cbetter(c, Cand{fabs(r), r, j});
    }
    Cand rc = block_argmax(c, s4);
    jstar = rc.i;
    piv = rc.v;
    if (jstar < 0 || piv == 0.0) break;  // a zero pivot ends the ACA
    for (int l = tid; l < t; l += NT) coefV[l] = V[(int64_t)l * n2 + jstar];
    if (tid == (jstar & (NT - 1))) {
      s_pt[0] = cx[jstar];
      s_pt[1] = cy[jstar];
      cmask |= 1ull << (jstar / NT);
    }
    if (tid == (istar & (NT - 1))) rmask |= 1ull << (istar / NT);
    __syncthreads();
    cjx = s_pt[0];
    cjy = s_pt[1];
    if (tid == 0) {
      a.prow[po + t] = istar;
      a.pcx[po + t] = cjx;
      a.pcy[po + t] = cjy;
    }
    // Phase B: v = row / pivot; residual column u.
    for (int j = tid; j < n2; j += NT) Vt[j] = Vt[j] / piv;
    double* Ut = U + (int64_t)t * n1;
    Cand cu{0.0, 0.0, -1};
    for (int i = tid, m = 0; i < n1; i += NT, m++) {
      double u = kernel_xy<KID>(a.src_x[r0 + i], a.src_y[r0 + i], cjx, cjy, a.kp);
      for (int l = 0; l < t; l++) u -= coefV[l] * U[(int64_t)l * n1 + i];
      Ut[i] = u;
      if (!((rmask >> m) & 1)) cu = cbetter(cu, Cand{fabs(u), u, i});
    }
    rank = t + 1;
    nr = block_argmax(cu, s4);  // includes a barrier: U_t, V_t complete
    // Phase C: U_l.u and V_l.v for l = 0..t.
    {
      const int w = tid >> 6, lane = tid & 63;
      for (int l = w; l <= t; l += NT / 64) {
        double su = 0.0, sv = 0.0;
        const double* Ul = U + (int64_t)l * n1;
        const double* Vl = V + (int64_t)l * n2;
        for (int i = lane; i < n1; i += 64) su += Ul[i] * Ut[i];
        for (int j = lane; j < n2; j += 64) sv += Vl[j] * Vt[j];
        su = wave_sum(su);
        sv = wave_sum(sv);
        if (lane == 0) {
          dU[l] = su;
          dV[l] = sv;
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      double pv = prev;
      if (rank != 2) pv = normS;
      double p = sqrt(dU[t]) * sqrt(dV[t]);
      double s = normS * normS + p * p;
      for (int l = 0; l <= t; l++) s += 2.0 * fabs(dU[l] * dV[l]);
      s_norm[0] = pv;
      s_norm[1] = sqrt(s);
    }
    __syncthreads();
    prev = s_norm[0];
    normS = s_norm[1];
    istar = nr.i;
    __syncthreads();
  }
  if (tid == 0) {
    a.rank[b] = rank;
    if (a.overflow) a.overflow[b] = capped;
  }
}

struct IdArgs {
  const double* src_x;
  const double* src_y;
  const int* rstart;
  const int* n1;
  const int* rank;
  const int64_t* piv_off;
  const int* prow;
  const double* pcx;
  const double* pcy;
  const int64_t* l2p_off;  // per box, local level offset
  double* l2p_rm;
  double* l2p_cm;
  int* rowmap;             // per row of the level's sources (rstart-indexed)
  double* mscr;            // global LU scratch when rank > RLDS
  const int64_t* m_off;
  int use_lds;
  KParams kp;
};

template <int KID>
__global__ __launch_bounds__(256) void k_id(IdArgs a) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  __shared__ int perm[MAXR];
  __shared__ int s_p;
  const int b = blockIdx.x, tid = threadIdx.x;
  const int r = a.rank[b];
  const int n1 = a.n1[b];
  const int r0 = a.rstart[b];
  int* rmap = a.rowmap + r0;
  if (r <= 0) {
    for (int i = tid; i < n1; i += blockDim.x) rmap[i] = i;
    return;
  }
  const int64_t po = a.piv_off[b];
  double* M = a.use_lds ? lds : a.mscr + a.m_off[b];
  // M = D^T: M[i][q] = K(skeleton row q, skeleton column i)
  for (int e = tid; e < r * r; e += blockDim.x) {
    int i = e / r, q = e % r;
    int rq = r0 + a.prow[po + q];
    M[e] = kernel_xy<KID>(a.src_x[rq], a.src_y[rq], a.pcx[po + i], a.pcy[po + i], a.kp);
  }
  __syncthreads();
  for (int k = 0; k < r; k++) {
    if (tid < 64) {
      Cand c{0.0, 0.0, -1};
      for (int i = k + tid; i < r; i += 64) {
        double v = M[i * r + k];
        c = cbetter(c, Cand{fabs(v), v, i});
      }
      c = wave_argmax(c);
      if (tid == 0) s_p = c.i;
    }
    __syncthreads();
    const int p = s_p;
    if (tid == 0) perm[k] = p;
    if (p != k)
      for (int q = tid; q < r; q += blockDim.x) {
        double t0 = M[k * r + q];
        M[k * r + q] = M[p * r + q];
        M[p * r + q] = t0;
      }
    __syncthreads();
    const double d = M[k * r + k];
    if (d != 0.0)
      for (int i = k + 1 + tid; i < r; i += blockDim.x) M[i * r + k] = M[i * r + k] / d;
    __syncthreads();
    const int w = r - k - 1;
    for (int e = tid; e < w * w; e += blockDim.x) {
      int i = k + 1 + e / w, q = k + 1 + e % w;
      M[i * r + q] = M[i * r + q] - M[i * r + k] * M[k * r + q];
    }
    __syncthreads();
  }
  double* rm = a.l2p_rm + a.l2p_off[b];
  double* cm = a.l2p_cm + a.l2p_off[b];
  const int nc = n1 - r;  // stored rows
  for (int i = tid; i < n1; i += blockDim.x) {
    // skeleton row q goes to position nc + q; the others keep their order
    int sq = -1, below = 0;
    for (int u = 0; u < r; u++) {
      const int pr = a.prow[po + u];
      sq = pr == i ? u : sq;
      below += pr < i;
    }
    if (sq >= 0) {
      rmap[nc + sq] = i;
      continue;
    }
    const int t = i - below;
    rmap[t] = i;
    double* x = rm + (int64_t)t * r;
    const double xi = a.src_x[r0 + i], yi = a.src_y[r0 + i];
    for (int q = 0; q < r; q++) x[q] = kernel_xy<KID>(xi, yi, a.pcx[po + q], a.pcy[po + q], a.kp);
    for (int k = 0; k < r; k++) {
      int p = perm[k];
      if (p != k) {
        double t0 = x[k];
        x[k] = x[p];
        x[p] = t0;
      }
    }
    for (int k = 0; k < r; k++) {
      double s = x[k];
      for (int q = 0; q < k; q++) s = s - M[k * r + q] * x[q];
      x[k] = s;
    }
    for (int k = r - 1; k >= 0; k--) {
      double s = x[k];
      for (int q = k + 1; q < r; q++) s = s - M[k * r + q] * x[q];
      x[k] = s / M[k * r + k];
    }
    for (int q = 0; q < r; q++) cm[t + (int64_t)q * nc] = x[q];
  }
}

// Skeleton points of a level: skx[sloc[k] + q] = src point prow[q] of box k.
__global__ void k_skeleton(const double* __restrict__ src_x, const double* __restrict__ src_y,
                           const int* __restrict__ src_g, const int* __restrict__ rstart,
                           const int* __restrict__ rank, const int64_t* __restrict__ piv_off,
                           const int* __restrict__ prow, const int64_t* __restrict__ sloc,
                           double* __restrict__ skx, double* __restrict__ sky,
                           int* __restrict__ skg, int nb) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  const int r = rank[b];
  for (int q = threadIdx.x; q < r; q += blockDim.x) {
    int p = rstart[b] + prow[piv_off[b] + q];
    skx[sloc[b] + q] = src_x[p];
    sky[sloc[b] + q] = src_y[p];
    skg[sloc[b] + q] = src_g ? src_g[p] : p;
  }
}

// M2L block of one unordered interaction pair (k, b), k < b:
// B[i + c * ld] = K(skeleton_k[i], skeleton_b[c]) for i < rank_k, 0 for the
// padding row i = rank_k of an odd rank_k (ld = rank_k rounded up to even).
template <int KID>
__global__ __launch_bounds__(64) void k_m2l_pair_build(
    const double* __restrict__ skx, const double* __restrict__ sky, const int* __restrict__ rank,
    const int64_t* __restrict__ soff, const int* __restrict__ pk, const int* __restrict__ pb,
    const int64_t* __restrict__ poff, double* __restrict__ m2l, KParams kp) {
  const int64_t p = blockIdx.x;
  if (poff[p] < 0) return;  // stored as ACA factors
  const int k = pk[p], b = pb[p];
  const int rk = rank[k], rb = rank[b];
  const int64_t sk = soff[k], sb = soff[b];
  double* out = m2l + poff[p];
  const int ld = rk + (rk & 1);
  const int tot = ld * rb;
  for (int e = threadIdx.x; e < tot; e += 64) {
    const int c = e / ld, i = e - c * ld;
    out[e] = i < rk ? kernel_xy<KID>(skx[sk + i], sky[sk + i], skx[sb + c], sky[sb + c], kp) : 0.0;
  }
}

// 48-bit block floating point of the dense M2L blocks (b48_split): one wave
// per pair finds its block's largest magnitude, then writes the block's row
// pair records and its exponent to the pair's descriptor.
__global__ __launch_bounds__(64) void k_m2l_b48(const double* __restrict__ src,
                                                const int* __restrict__ rank,
                                                const int* __restrict__ pk,
                                                const int* __restrict__ pb,
                                                const int64_t* __restrict__ poff,
                                                B48Pair* __restrict__ dst,
                                                PairDesc* __restrict__ pd) {
  const int64_t p = blockIdx.x;
  const int64_t off = poff[p];
  if (off < 0) return;  // stored as ACA factors
  const int rk = rank[pk[p]], rb = rank[pb[p]];
  const int tot = (rk + (rk & 1)) * rb;  // even: whole row pairs
  double m = 0.0;
  for (int i = threadIdx.x; i < tot; i += 64) m = fmax(m, fabs(src[off + i]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  const int sx = b48_exponent(m);
  for (int i = threadIdx.x; i < tot / 2; i += 64) {
    uint32_t l0, l1;
    uint16_t h0, h1;
    b48_split(src[off + 2 * i], sx, l0, h0);
    b48_split(src[off + 2 * i + 1], sx, l1, h1);
    dst[(off >> 1) + i] = B48Pair{l0, l1, (uint32_t)h0 | ((uint32_t)h1 << 16)};
  }
  if (threadIdx.x == 0) pd[p].sexp = sx;
}

template <class T>
void upload(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
  if (!h.empty()) {
    AFMM_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    AFMM_HIP(hipStreamSynchronize(s));  // h may go out of scope after return
  }
}

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t free_device_bytes() {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)8 << 30;
  return fr;
}

// The ACA scratch (up to 16 GiB) is kept per device between assemblies and
// only grows: freeing and allocating buffers of this size in a loop made
// later assemblies wait seconds in the allocator (on the driver's clearing
// of the freed memory) at random.  afmm2d_release_scratch frees it.
// AFMM2D_SCRATCH_POOL=0: a buffer per assembly, freed at its end.
std::mutex g_pool_mu;
std::map<int, DevBuf> g_pool;
// assemblies share the pooled scratch: one at a time per process
std::mutex g_assembly_mu;

bool pool_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("AFMM2D_SCRATCH_POOL");
    return !(v && v[0] == '0');
  }();
  return on;
}

struct Scratch {
  DevBuf own;  // when the pool is off
  int device;
  // budget: the largest batch the caller makes.  A pooled buffer that must
  // grow past half of it grows to all of it, so that the pool grows at most
  // once more (a batch of a larger problem fills the budget to within a box,
  // and growing by those few MB freed and re-allocated 16 GiB, which stalled)
  double* get(size_t doubles, int64_t& peak, size_t budget) {
    size_t bytes = sizeof(double) * std::max<size_t>(doubles, 1);
    if (!pool_enabled()) {
      own.alloc(bytes);
      return own.as<double>();
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    DevBuf& b = g_pool[device];
    if (bytes > b.bytes && 2 * bytes > budget) bytes = std::max(bytes, budget);
    b.alloc(bytes);  // grows only
    peak = std::max<int64_t>(peak, (int64_t)b.bytes);
    return b.as<double>();
  }
};

struct BBox {
  double x0 = DBL_MAX, y0 = DBL_MAX, x1 = -DBL_MAX, y1 = -DBL_MAX;
  void add(const BBox& o) {
    x0 = std::min(x0, o.x0);
    y0 = std::min(y0, o.y0);
    x1 = std::max(x1, o.x1);
    y1 = std::max(y1, o.y1);
  }
  bool empty() const { return x0 > x1; }
};

// The rank model's inputs for box k of level j (kRmIn floats): level, the
// half-width of the box's point bounding box (log2), its centre relative to
// the domain's, the kernel (one-hot: log, 1/r, Gaussian, a/r, user), log10
// of the tolerance, the ACA block's rows and columns (log2), the share of
// interaction slots present, the kernel's length parameter (log10) and
// whether the box is a leaf.
void rank_features(const Engine& e, int j, int n1, int n2, int nil, const BBox& bb,
                   const BBox& dom, float* f) {
  for (int u = 0; u < kRmIn; u++) f[u] = 0.0f;
  const double dcx = 0.5 * (dom.x0 + dom.x1), dcy = 0.5 * (dom.y0 + dom.y1);
  const double dhw = std::max(1e-300, 0.5 * std::max(dom.x1 - dom.x0, dom.y1 - dom.y0));
  double hw = 0.0, cx = dcx, cy = dcy;
  if (!bb.empty()) {
    hw = 0.5 * std::max(bb.x1 - bb.x0, bb.y1 - bb.y0);
    cx = 0.5 * (bb.x0 + bb.x1);
    cy = 0.5 * (bb.y0 + bb.y1);
  }
  f[0] = (float)(j / 10.0);
  f[1] = (float)(std::log2(std::max(hw, 1e-12)) / 10.0);
  f[2] = (float)((cx - dcx) / dhw);
  f[3] = (float)((cy - dcy) / dhw);
  const int k = e.kid_full >= K_USER ? 4 : e.kid_full;
  f[4 + k] = 1.0f;
  f[9] = (float)(-e.tol_pow / 10.0);
  f[10] = (float)(std::log2(1.0 + n1) / 10.0);
  f[11] = (float)(std::log2(1.0 + n2) / 16.0);
  f[12] = (float)(nil / 40.0);
  if (e.kid_full == K_GAUSS) f[13] = (float)(std::log10(e.kp.p0) / 4.0);
  if (e.kid_full == K_REG_INV_R) f[13] = (float)(std::log10(e.kp.a) / 4.0);
  f[14] = j == e.L ? 1.0f : 0.0f;
}

// ACA capacity from a prediction of log2(1 + rank): 2^(pred + margin) - 1
// rounded up to a bucket (multiples of 8 to 64, of 16 to 128, of 32 beyond),
// at least 4, at most min(rows, columns).
int rank_capacity(float pred, float margin, int mn) {
  const double r = std::exp2(std::min(20.0, (double)pred + margin)) - 1.0;
  int c = std::max(4, (int)std::ceil(r));
  if (c <= 64) c = (c + 7) / 8 * 8;
  else if (c <= 128) c = (c + 15) / 16 * 16;
  else c = (c + 31) / 32 * 32;
  return std::min(c, mn);
}

// Factors of pair entry e of an ACA batch (its U: n1 x p at src, its V:
// n2 x p at src + n1 cap) to dst (U then V, column-major).
struct FactCopy {
  int64_t src, dst;
  int n1, n2, p, cap;
};

__global__ __launch_bounds__(256) void k_copy_factors(const FactCopy* __restrict__ fc,
                                                      const double* __restrict__ scr,
                                                      double* __restrict__ dst) {
  const FactCopy f = fc[blockIdx.x];
  const int64_t nu = (int64_t)f.n1 * f.p, nv = (int64_t)f.n2 * f.p;
  for (int64_t t = threadIdx.x; t < nu; t += blockDim.x) dst[f.dst + t] = scr[f.src + t];
  const double* v = scr + f.src + (int64_t)f.n1 * f.cap;
  for (int64_t t = threadIdx.x; t < nv; t += blockDim.x) dst[f.dst + nu + t] = v[t];
}

// Per-pair ACA of the M2L blocks K(skeleton_first, skeleton_second): the
// box ACA kernel run on the global skeleton points, one entry per pair
// whose only interaction slot is a column entry for the second box.
// p[i] = the factorisation's rank for pairs stored factored (p (r1 + r2) <
// r1 r2 and p <= 256), -1 otherwise; foff[i] = their offsets in e.m2lf.
void pair_aca(Engine& e, const std::vector<int>& first, const std::vector<int>& second,
              const std::vector<int>& rank_g, const std::vector<int64_t>& soff_g,
              std::vector<int>& p, std::vector<int64_t>& foff) {
  hipStream_t s = e.s_aux;
  const int P = (int)first.size();
  p.assign(P, -1);
  foff.assign(P, -1);
  if (P == 0) return;
  if (e.skel_total > INT32_MAX) throw Error(6, "skeleton count exceeds 2^31");
  std::vector<int> rstart(2 * (size_t)P), n1(2 * (size_t)P), n2(2 * (size_t)P, 0),
      il(40 * (size_t)P, -1);
  std::vector<int64_t> pivoff(P + 1, 0);
  for (int i = 0; i < P; i++) {
    rstart[i] = (int)soff_g[first[i]];
    n1[i] = rank_g[first[i]];
    rstart[P + i] = (int)soff_g[second[i]];
    n1[P + i] = rank_g[second[i]];
    n2[i] = rank_g[second[i]];
    il[40 * (size_t)i] = P + i;
    pivoff[i + 1] = pivoff[i] + std::min(n1[i], n2[i]);
  }
  DevBuf d_rstart, d_n1, d_n2, d_il, d_pivoff, d_rank, d_prow, d_pcx, d_pcy, d_boxes, d_scroff,
      d_fc;
  Scratch scr{DevBuf{}, e.device};
  int64_t peak_unused = 0;
  upload(d_rstart, rstart, s);
  upload(d_n1, n1, s);
  upload(d_n2, n2, s);
  upload(d_il, il, s);
  upload(d_pivoff, pivoff, s);
  const int64_t npiv = std::max<int64_t>(pivoff[P], 1);
  d_rank.alloc(sizeof(int) * P);
  d_prow.alloc(sizeof(int) * npiv);
  d_pcx.alloc(sizeof(double) * npiv);
  d_pcy.alloc(sizeof(double) * npiv);
  const size_t budget = std::min<size_t>((size_t)4 << 30,
                                         std::max<size_t>((size_t)1 << 28, free_device_bytes() / 8));
  std::vector<DevBuf> parts;           // factors of each batch
  std::vector<int64_t> part_base;      // their offsets in m2lf
  int64_t ftot = 0;
  int i0 = 0;
  while (i0 < P) {
    std::vector<int> boxes;
    std::vector<int64_t> soffs;
    int64_t tot = 0;
    int i = i0;
    for (; i < P; i++) {
      const int mn = std::min(n1[i], n2[i]);
      int64_t need = (int64_t)(n1[i] + n2[i]) * mn + 2 * (int64_t)n2[i];
      need = (need + 31) & ~(int64_t)31;
      if (!boxes.empty() && (size_t)(tot + need) * sizeof(double) > budget) break;
      boxes.push_back(i);
      soffs.push_back(tot);
      tot += need;
    }
    i0 = i;
    upload(d_boxes, boxes, s);
    upload(d_scroff, soffs, s);
    double* d_scr = scr.get((size_t)tot, peak_unused, budget);
    AcaArgs a{e.skx.as<double>(), e.sky.as<double>(), d_rstart.as<int>(), d_n1.as<int>(),
              d_n2.as<int>(), d_il.as<int>(), d_boxes.as<int>(), d_scroff.as<int64_t>(),
              d_scr, d_pivoff.as<int64_t>(), d_rank.as<int>(), d_prow.as<int>(),
              d_pcx.as<double>(), d_pcy.as<double>(), e.tol, e.kp, nullptr, nullptr};
    AFMM_DISPATCH_KERNEL(e.kid, KID,
                         k_aca<KID, ACA_THREADS><<<(unsigned)boxes.size(), ACA_THREADS, 0, s>>>(a));
    AFMM_LAUNCH_CHECK();
    std::vector<int> rk(P);
    AFMM_HIP(hipMemcpyAsync(rk.data(), d_rank.p, sizeof(int) * P, hipMemcpyDeviceToHost, s));
    AFMM_HIP(hipStreamSynchronize(s));
    std::vector<FactCopy> fc;
    int64_t local = 0;
    for (size_t u = 0; u < boxes.size(); u++) {
      const int q = boxes[u];
      const int r1 = n1[q], r2 = n2[q], pr = rk[q];
      if (pr > 256 || (int64_t)pr * (r1 + r2) >= (int64_t)r1 * r2) continue;
      p[q] = pr;
      foff[q] = ftot + local;
      fc.push_back(FactCopy{soffs[u], local, r1, r2, pr, std::min(r1, r2)});
      local += (int64_t)pr * (r1 + r2);
    }
    if (!fc.empty()) {
      parts.emplace_back();
      parts.back().alloc(sizeof(double) * std::max<int64_t>(local, 1));
      part_base.push_back(ftot);
      upload(d_fc, fc, s);
      k_copy_factors<<<(unsigned)fc.size(), 256, 0, s>>>(d_fc.as<FactCopy>(), d_scr,
                                                         parts.back().as<double>());
      AFMM_LAUNCH_CHECK();
      AFMM_HIP(hipStreamSynchronize(s));  // the next batch reuses the scratch
      ftot += local;
    }
  }
  e.m2lf.alloc(sizeof(double) * std::max<int64_t>(ftot, 1));
  for (size_t u = 0; u < parts.size(); u++) {
    const int64_t n = (u + 1 < part_base.size() ? part_base[u + 1] : ftot) - part_base[u];
    if (n > 0)
      AFMM_HIP(hipMemcpyAsync(e.m2lf.as<double>() + part_base[u], parts[u].p, sizeof(double) * n,
                              hipMemcpyDeviceToDevice, s));
  }
  AFMM_HIP(hipStreamSynchronize(s));
  e.m2lf_total = ftot;
}

}  // namespace

void assemble_tree(Engine& e) {
  std::lock_guard<std::mutex> assembly_lock(g_assembly_mu);
  const int L = e.L;
  hipStream_t s = e.s_aux;
  e.lv.assign(L + 1, LevelInfo{});
  // per-level skeleton points, kept until the global arrays are built
  std::vector<DevBuf> lskx(L + 2), lsky(L + 2), lskg(L + 2), lrm(L + 1), lcm(L + 1),
      lrow(L + 1);
  std::vector<int64_t> nrows(L + 1, 0);  // source rows of each level
  std::vector<std::vector<int64_t>> sloc(L + 2);
  Scratch scr{DevBuf{}, e.device};
  DevBuf d_rstart, d_n1, d_n2, d_il, d_boxes, d_scroff, d_pivoff, d_rank, d_prow, d_pcx,
      d_pcy, d_sloc, d_l2poff, d_moff, d_mscr, d_cap, d_over, d_feat, d_pred;
  // point bounding boxes: leaves from their points, parents from children
  std::vector<BBox> bbox(1 << (2 * L));
  BBox dom;
  for (int k = 0; k < (1 << (2 * L)); k++) {
    for (int i = e.leaf_start[k]; i < e.leaf_start[k] + e.leaf_count[k]; i++) {
      const double px = e.h_xy[2 * (size_t)i], py = e.h_xy[2 * (size_t)i + 1];
      bbox[k].x0 = std::min(bbox[k].x0, px);
      bbox[k].x1 = std::max(bbox[k].x1, px);
      bbox[k].y0 = std::min(bbox[k].y0, py);
      bbox[k].y1 = std::max(bbox[k].y1, py);
    }
    dom.add(bbox[k]);
  }
  e.rank_feat.clear();
  e.rank_label.clear();
  e.rank_boxes = e.rank_fallbacks = e.aca_scratch_peak = 0;
  e.aca_launches = 0;
  e.aca_seconds = 0.0;

  for (int j = L; j >= 2; j--) {
    LevelInfo& li = e.lv[j];
    li.j = j;
    li.nb = 1 << (2 * j);
    const int nb = li.nb;
    std::vector<int> rstart(nb), n1(nb), n2(nb, 0);
    const double* src_x;
    const double* src_y;
    const int* src_g;
    if (j == L) {
      for (int k = 0; k < nb; k++) {
        rstart[k] = e.leaf_start[k];
        n1[k] = e.leaf_count[k];
      }
      src_x = e.x.as<double>();
      src_y = e.y.as<double>();
      src_g = nullptr;
    } else {
      const std::vector<int64_t>& cs = sloc[j + 1];
      for (int k = 0; k < nb; k++) {
        rstart[k] = (int)cs[4 * k];
        n1[k] = (int)(cs[4 * k + 4] - cs[4 * k]);
      }
      src_x = lskx[j + 1].as<double>();
      src_y = lsky[j + 1].as<double>();
      src_g = lskg[j + 1].as<int>();
    }
    std::vector<int> il(40 * (size_t)nb);
    for (int k = 0; k < nb; k++) {
      for (int q = 0; q < 16; q++) il[40 * k + q] = e.inner[j][16 * k + q];
      for (int q = 0; q < 24; q++) il[40 * k + 16 + q] = e.outer[j][24 * k + q];
      int c = 0;
      for (int q = 0; q < 40; q++)
        if (il[40 * k + q] >= 0) c += n1[il[40 * k + q]];
      n2[k] = c;
    }
    // Levels with few boxes (large ones) give each box a workgroup of 1024
    // threads; the ranks and pivots are the same for either size (every
    // value is computed alike, the argmax is order-free and each dot product
    // is one wave's).
    const int aca_nt = nb < 2 * e.num_cus ? 1024 : ACA_THREADS;
    std::vector<int64_t> pivoff(nb + 1, 0);
    int maxn = 0;
    for (int k = 0; k < nb; k++) {
      int mn = std::min(n1[k], n2[k]);
      pivoff[k + 1] = pivoff[k] + std::max(mn, 0);
      maxn = std::max(maxn, std::max(n1[k], n2[k]));
      if (mn > MAXR)
        throw Error(6, "ACA block " + std::to_string(n1[k]) + "x" + std::to_string(n2[k]) +
                           " exceeds the supported rank capacity");
      if ((n1[k] + aca_nt - 1) / aca_nt > MAXOWN || (n2[k] + aca_nt - 1) / aca_nt > MAXOWN)
        throw Error(6, "ACA block too large for one workgroup");
    }
    upload(d_rstart, rstart, s);
    upload(d_n1, n1, s);
    upload(d_n2, n2, s);
    upload(d_il, il, s);
    upload(d_pivoff, pivoff, s);
    const int64_t npiv = std::max<int64_t>(pivoff[nb], 1);
    d_rank.alloc(sizeof(int) * nb);
    d_prow.alloc(sizeof(int) * npiv);
    d_pcx.alloc(sizeof(double) * npiv);
    d_pcy.alloc(sizeof(double) * npiv);

    // bounding boxes of this level (children -> parents above the leaves)
    if (j < L) {
      std::vector<BBox> up(nb);
      for (int k = 0; k < nb; k++)
        for (int c = 0; c < 4; c++) up[k].add(bbox[4 * k + c]);
      bbox.swap(up);
    }
    const size_t f0 = e.rank_feat.size();
    e.rank_feat.resize(f0 + (size_t)nb * kRmIn);
    for (int k = 0; k < nb; k++) {
      int nil = 0;
      for (int q = 0; q < 40; q++) nil += il[40 * k + q] >= 0;
      rank_features(e, j, n1[k], n2[k], nil, bbox[k], dom, &e.rank_feat[f0 + (size_t)k * kRmIn]);
    }
    const double t_aca = now_seconds();
    // capacities: min(rows, columns), or the rank model's bucketed prediction
    std::vector<int> cap(nb);
    for (int k = 0; k < nb; k++) cap[k] = std::max(0, std::min(n1[k], n2[k]));
    std::vector<int> order(nb);
    std::iota(order.begin(), order.end(), 0);
    const bool learned = e.rank_model != nullptr;
    if (learned) {
      d_feat.alloc(sizeof(float) * (size_t)nb * kRmIn);
      d_pred.alloc(sizeof(float) * (size_t)nb);
      AFMM_HIP(hipMemcpyAsync(d_feat.p, &e.rank_feat[f0], sizeof(float) * (size_t)nb * kRmIn,
                              hipMemcpyHostToDevice, s));
      rank_model_predict_device(*e.rank_model, d_feat.as<float>(), nb, d_pred.as<float>(), s);
      std::vector<float> pred(nb);
      AFMM_HIP(hipMemcpyAsync(pred.data(), d_pred.p, sizeof(float) * nb, hipMemcpyDeviceToHost, s));
      AFMM_HIP(hipStreamSynchronize(s));
      for (int k = 0; k < nb; k++)
        if (cap[k] > 0) cap[k] = rank_capacity(pred[k], e.rank_margin, cap[k]);
      // buckets: boxes of equal capacity run side by side
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cap[a] < cap[b]; });
      d_over.alloc(sizeof(int) * nb);
      AFMM_HIP(hipMemsetAsync(d_over.p, 0, sizeof(int) * nb, s));
    }
    // ACA in batches bounded by a scratch budget of at most 16 GiB, one
    // buffer reused by every batch and level: a level's whole scratch at once
    // (24-55 GB at N = 1e6) made assembly take 4-5 s instead of ~1 s in some
    // runs; 4 GiB batches left too few boxes per launch at the leaf level.
    // AFMM2D_ACA_BUDGET_MB (at least 16) sets it instead: a process that
    // leaves the engine little device memory
    size_t budget = std::min<size_t>((size_t)16 << 30,
                                     std::max<size_t>((size_t)1 << 30, free_device_bytes() / 4));
    if (const char* v = std::getenv("AFMM2D_ACA_BUDGET_MB"))
      budget = (size_t)std::max(16L, std::atol(v)) << 20;
    // boxes: the order to run; capped: hold cap[] (else full capacity)
    auto run_aca = [&](const std::vector<int>& boxes_in, bool capped) {
      size_t b0 = 0;
      while (b0 < boxes_in.size()) {
        std::vector<int> boxes, caps;
        std::vector<int64_t> soffs;
        int64_t tot = 0;
        size_t q = b0;
        for (; q < boxes_in.size(); q++) {
          const int k = boxes_in[q];
          const int mn = std::max(0, std::min(n1[k], n2[k]));
          const int c = capped ? cap[k] : mn;
          int64_t need = (int64_t)(n1[k] + n2[k]) * c + 2 * (int64_t)n2[k];
          need = (need + 31) & ~(int64_t)31;
          if (!boxes.empty() && (size_t)(tot + need) * sizeof(double) > budget) break;
          boxes.push_back(k);
          caps.push_back(c);
          soffs.push_back(tot);
          tot += need;
        }
        b0 = q;
        upload(d_boxes, boxes, s);
        upload(d_scroff, soffs, s);
        if (capped) upload(d_cap, caps, s);
        int64_t held = 0;
        double* d_scr = scr.get((size_t)tot, held, budget);
        e.aca_scratch_peak = std::max<int64_t>(e.aca_scratch_peak, 8 * tot);
        AcaArgs a{src_x, src_y, d_rstart.as<int>(), d_n1.as<int>(), d_n2.as<int>(), d_il.as<int>(),
                  d_boxes.as<int>(), d_scroff.as<int64_t>(), d_scr,
                  d_pivoff.as<int64_t>(), d_rank.as<int>(), d_prow.as<int>(), d_pcx.as<double>(),
                  d_pcy.as<double>(), e.tol, e.kp, capped ? d_cap.as<int>() : nullptr,
                  capped ? d_over.as<int>() : nullptr};
        if (aca_nt == 1024) {
          AFMM_DISPATCH_KERNEL(e.kid, KID, k_aca<KID, 1024><<<(unsigned)boxes.size(), 1024, 0, s>>>(a));
        } else {
          AFMM_DISPATCH_KERNEL(e.kid, KID,
                               k_aca<KID, ACA_THREADS><<<(unsigned)boxes.size(), ACA_THREADS, 0, s>>>(a));
        }
        AFMM_LAUNCH_CHECK();
        e.aca_launches++;
      }
    };
    run_aca(order, learned);
    if (learned) {
      // boxes whose ACA went beyond the predicted capacity run again in full
      std::vector<int> over(nb);
      AFMM_HIP(hipMemcpyAsync(over.data(), d_over.p, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
      AFMM_HIP(hipStreamSynchronize(s));
      std::vector<int> redo;
      for (int k = 0; k < nb; k++)
        if (over[k]) redo.push_back(k);
      e.rank_fallbacks += (int64_t)redo.size();
      run_aca(redo, false);
    }
    for (int k = 0; k < nb; k++) e.rank_boxes += std::min(n1[k], n2[k]) > 0;
    li.rank.resize(nb);
    AFMM_HIP(hipMemcpyAsync(li.rank.data(), d_rank.p, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    AFMM_HIP(hipStreamSynchronize(s));
    e.aca_seconds += now_seconds() - t_aca;
    e.rank_label.insert(e.rank_label.end(), li.rank.begin(), li.rank.end());
    li.nin = n1;
    // local skeleton offsets and L2P offsets
    sloc[j].assign(nb + 1, 0);
    std::vector<int64_t> l2poff(nb + 1, 0), moff(nb + 1, 0);
    int maxr = 0;
    for (int k = 0; k < nb; k++) {
      int r = li.rank[k];
      maxr = std::max(maxr, r);
      sloc[j][k + 1] = sloc[j][k] + r;
      l2poff[k + 1] = l2poff[k] + (int64_t)r * (n1[k] - r);
      moff[k + 1] = moff[k] + (int64_t)r * r;
    }
    e.max_rank = std::max(e.max_rank, maxr);
    li.l2poff = l2poff;
    const int64_t nsk = std::max<int64_t>(sloc[j][nb], 1);
    lskx[j].alloc(sizeof(double) * nsk);
    lsky[j].alloc(sizeof(double) * nsk);
    lskg[j].alloc(sizeof(int) * nsk);
    upload(d_sloc, sloc[j], s);
    k_skeleton<<<nb, 64, 0, s>>>(src_x, src_y, src_g, d_rstart.as<int>(), d_rank.as<int>(),
                                 d_pivoff.as<int64_t>(), d_prow.as<int>(), d_sloc.as<int64_t>(),
                                 lskx[j].as<double>(), lsky[j].as<double>(), lskg[j].as<int>(), nb);
    AFMM_LAUNCH_CHECK();
    // ID
    lrm[j].alloc(sizeof(double) * std::max<int64_t>(l2poff[nb], 1));
    lcm[j].alloc(sizeof(double) * std::max<int64_t>(l2poff[nb], 1));
    upload(d_l2poff, l2poff, s);
    // a leaf's points start at leaf_start (KD order, not box order)
    nrows[j] = j == L ? (int64_t)e.n : sloc[j + 1][4 * (size_t)nb];
    lrow[j].alloc(sizeof(int) * std::max<int64_t>(nrows[j], 1));
    const int use_lds = maxr <= RLDS;
    if (!use_lds) {
      upload(d_moff, moff, s);
      d_mscr.alloc(sizeof(double) * std::max<int64_t>(moff[nb], 1));
    }
    if (maxr > MAXR) throw Error(6, "rank exceeds the supported capacity");
    IdArgs ia{src_x, src_y, d_rstart.as<int>(), d_n1.as<int>(), d_rank.as<int>(),
              d_pivoff.as<int64_t>(), d_prow.as<int>(), d_pcx.as<double>(), d_pcy.as<double>(),
              d_l2poff.as<int64_t>(), lrm[j].as<double>(), lcm[j].as<double>(), lrow[j].as<int>(),
              use_lds ? nullptr : d_mscr.as<double>(), use_lds ? nullptr : d_moff.as<int64_t>(),
              use_lds, e.kp};
    const size_t shm = use_lds ? sizeof(double) * (size_t)maxr * maxr : 0;
    AFMM_DISPATCH_KERNEL(e.kid, KID, k_id<KID><<<nb, 256, shm, s>>>(ia));
    AFMM_LAUNCH_CHECK();
    AFMM_HIP(hipStreamSynchronize(s));
  }

  // Global numbering: levels 2..L in order, boxes in quad-tree order.
  int64_t gb = 0, sb = 0, lb = 0;
  for (int j = 2; j <= L; j++) {
    LevelInfo& li = e.lv[j];
    li.gbase = gb;
    li.soff.resize(li.nb + 1);
    for (int k = 0; k <= li.nb; k++) li.soff[k] = sb + sloc[j][k];
    for (int k = 0; k <= li.nb; k++) li.l2poff[k] += lb;
    gb += li.nb;
    sb += sloc[j][li.nb];
    lb = li.l2poff[li.nb];
  }
  const int64_t nbox = gb;
  e.skel_total = sb;
  e.l2p_total = lb;
  e.skx.alloc(sizeof(double) * std::max<int64_t>(sb, 1));
  e.sky.alloc(sizeof(double) * std::max<int64_t>(sb, 1));
  e.skidx.alloc(sizeof(int) * std::max<int64_t>(sb, 1));
  e.l2p_rm.alloc(sizeof(double) * std::max<int64_t>(lb, 1));
  e.l2p_cm.alloc(sizeof(double) * std::max<int64_t>(lb, 1));
  e.rowmap.alloc(sizeof(int) * (size_t)(e.n + std::max<int64_t>(sb, 1)));
  for (int j = 2; j <= L; j++) {
    // the leaves' rows are the points; an upper level's rows are the
    // skeletons of the level below
    const int64_t rbase = j == L ? 0 : e.n + e.lv[j + 1].soff[0];
    if (nrows[j] > 0)
      AFMM_HIP(hipMemcpyAsync(e.rowmap.as<int>() + rbase, lrow[j].p, sizeof(int) * nrows[j],
                              hipMemcpyDeviceToDevice, s));
  }
  for (int j = 2; j <= L; j++) {
    LevelInfo& li = e.lv[j];
    int64_t ns = li.soff[li.nb] - li.soff[0];
    int64_t nl = li.l2poff[li.nb] - li.l2poff[0];
    if (ns > 0) {
      AFMM_HIP(hipMemcpyAsync(e.skx.as<double>() + li.soff[0], lskx[j].p, sizeof(double) * ns,
                              hipMemcpyDeviceToDevice, s));
      AFMM_HIP(hipMemcpyAsync(e.sky.as<double>() + li.soff[0], lsky[j].p, sizeof(double) * ns,
                              hipMemcpyDeviceToDevice, s));
      AFMM_HIP(hipMemcpyAsync(e.skidx.as<int>() + li.soff[0], lskg[j].p, sizeof(int) * ns,
                              hipMemcpyDeviceToDevice, s));
    }
    if (nl > 0) {
      AFMM_HIP(hipMemcpyAsync(e.l2p_rm.as<double>() + li.l2poff[0], lrm[j].p,
                              sizeof(double) * nl, hipMemcpyDeviceToDevice, s));
      AFMM_HIP(hipMemcpyAsync(e.l2p_cm.as<double>() + li.l2poff[0], lcm[j].p,
                              sizeof(double) * nl, hipMemcpyDeviceToDevice, s));
    }
  }
  AFMM_HIP(hipStreamSynchronize(s));
  for (int j = 0; j <= L + 1; j++) {
    lskx[j].release(); lsky[j].release(); lskg[j].release();
  }
  for (int j = 0; j <= L; j++) { lrm[j].release(); lcm[j].release(); lrow[j].release(); }
  scr.own.release();
  d_mscr.release();

  // Per-box arrays over all levels.
  std::vector<int> rank_g(nbox), nin_g(nbox);
  std::vector<int64_t> soff_g(nbox + 1), l2poff_g(nbox + 1);
  for (int j = 2; j <= L; j++) {
    LevelInfo& li = e.lv[j];
    li.max_nin = 0;
    li.max_rank = 0;
    li.max_stored = 0;
    for (int k = 0; k < li.nb; k++) {
      int64_t g = li.gbase + k;
      li.max_nin = std::max(li.max_nin, li.nin[k]);
      li.max_rank = std::max(li.max_rank, li.rank[k]);
      li.max_stored = std::max(li.max_stored, li.nin[k] - li.rank[k]);
      rank_g[g] = li.rank[k];
      nin_g[g] = li.nin[k];
      soff_g[g] = li.soff[k];
      l2poff_g[g] = li.l2poff[k];
    }
  }
  soff_g[nbox] = sb;
  l2poff_g[nbox] = lb;
  upload(e.rank_g, rank_g, s);
  upload(e.nin_g, nin_g, s);
  upload(e.soff_g, soff_g, s);
  upload(e.l2poff_g, l2poff_g, s);

  // M2L layout: every unordered interaction pair once.  Each pair is stored
  // as B = K(skeleton_first, skeleton_second) in the orientation whose
  // register-tile class is cheaper (m2l_tile_class), and within two groups
  // (levels 2..L-1, level L) the pairs are ordered by class, so that one
  // launch per (group, class) covers them.  Partials are pair-major: pair q
  // writes its two products next to each other, pairs in storage order, so
  // the pair kernels' writes form one sequential stream; box g's t-th present
  // interaction slot (slot order) reads its product at slot_off[27 g + t].
  e.h_pair_k.clear();
  e.h_pair_b.clear();
  e.pair_max_ld = 0;
  e.pair_max_rb = 0;
  e.h_pair_off.clear();
  e.h_red_start.assign(nbox + 1, 0);
  e.h_red_pair.clear();
  std::unordered_map<int64_t, int> pair_of;
  int64_t off = 0, part = 0;
  e.il_pairs = 0;
  e.m2l_total = 0;
  auto slot_box = [&](int j, int k, int q) {
    return q < 16 ? e.inner[j][16 * (size_t)k + q] : e.outer[j][24 * (size_t)k + q - 16];
  };
  struct Cand {
    int first, second, cls, group;
  };
  std::vector<Cand> cand;
  for (int j = 2; j <= L; j++) {
    LevelInfo& li = e.lv[j];
    li.m2l_coeffs = li.m2l_cols = li.m2l_fact_coeffs = 0;
    for (int k = 0; k < li.nb; k++) {
      for (int q = 0; q < 40; q++) {
        int b = slot_box(j, k, q);
        if (b < 0 || b <= k) continue;
        const int gk = (int)(li.gbase + k), gbx = (int)(li.gbase + b);
        const int rk = li.rank[k], rb = li.rank[b];
        const int ca = m2l_tile_class(rk + (rk & 1), rb);
        const int cb = m2l_tile_class(rb + (rb & 1), rk);
        if (cb < ca) cand.push_back(Cand{gbx, gk, cb, j == L ? 1 : 0});
        else cand.push_back(Cand{gk, gbx, ca, j == L ? 1 : 0});
        part += rk + rb;
        e.m2l_total += (int64_t)rk * rb;
        li.m2l_coeffs += (int64_t)rk * rb;
        li.m2l_cols += rk + rb;
      }
    }
  }
  // ACA-factored pairs (opt-in; symmetric kernels: the pair's other
  // direction applies V U^T)
  std::vector<int> fact_p(cand.size(), -1);
  std::vector<int64_t> fact_off(cand.size(), -1);
  e.m2lf_pairs = e.m2lf_total = e.m2lf_dense = 0;
  e.m2lf.release();
  if (e.m2l_factor && e.kid != K_USER_ANTI && !cand.empty()) {
    std::vector<int> fi(cand.size()), se(cand.size());
    for (size_t i = 0; i < cand.size(); i++) {
      fi[i] = cand[i].first;
      se[i] = cand[i].second;
    }
    pair_aca(e, fi, se, rank_g, soff_g, fact_p, fact_off);
    for (size_t i = 0; i < cand.size(); i++) {
      if (fact_p[i] < 0) continue;
      const int r1 = rank_g[cand[i].first], r2 = rank_g[cand[i].second];
      const int64_t dense = (int64_t)r1 * r2, stored = (int64_t)fact_p[i] * (r1 + r2);
      cand[i].cls = kM2LFactored;
      e.m2lf_pairs++;
      e.m2lf_dense += dense;
      e.m2l_total += stored - dense;
      if (cand[i].group == 1) {  // the leaf launch's bytes
        e.lv[L].m2l_coeffs += stored - dense;
        e.lv[L].m2l_fact_coeffs += stored;
      }
    }
  }
  std::vector<int> order(cand.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int u, int v) {
    return cand[u].group != cand[v].group ? cand[u].group < cand[v].group
                                          : cand[u].cls < cand[v].cls;
  });
  for (int g = 0; g < 2; g++)
    for (int c = 0; c <= kM2LFactored; c++) e.m2l_cls_begin[g][c] = e.m2l_cls_end[g][c] = 0;
  e.h_pair_fact.clear();
  std::vector<int64_t> build_off;  // per pair: dense block offset, -1 when factored
  for (size_t i = 0; i < order.size(); i++) {
    const Cand& cd = cand[order[i]];
    const int qidx = (int)i;
    if (i == 0 || cand[order[i - 1]].group != cd.group || cand[order[i - 1]].cls != cd.cls)
      e.m2l_cls_begin[cd.group][cd.cls] = qidx;
    e.m2l_cls_end[cd.group][cd.cls] = qidx + 1;
    const int rf = rank_g[cd.first], rs = rank_g[cd.second];
    const int lo = std::min(cd.first, cd.second), hi = std::max(cd.first, cd.second);
    pair_of[((int64_t)lo << 32) | hi] = qidx;
    e.h_pair_k.push_back(cd.first);
    e.h_pair_b.push_back(cd.second);
    const int fp = fact_p[order[i]];
    e.h_pair_fact.push_back(fp);
    if (fp >= 0) {
      e.h_pair_off.push_back(fact_off[order[i]]);
      build_off.push_back(-1);
      continue;
    }
    e.h_pair_off.push_back(off);
    build_off.push_back(off);
    const int ld = rf + (rf & 1);
    off += (int64_t)ld * rs;  // ld even: every block and column 16-byte aligned
    e.pair_max_ld = std::max(e.pair_max_ld, ld);
    e.pair_max_rb = std::max(e.pair_max_rb, rs);
  }
  e.npairs = (int64_t)e.h_pair_k.size();
  if (part > INT32_MAX) throw Error(6, "M2L partial buffer exceeds 2^31 entries");
  // pair-major partials: pair q writes B out_b then B^T out_k at pp[q]
  std::vector<int> part_k(e.npairs, 0), part_b(e.npairs, 0);
  {
    int64_t pp = 0;
    for (int64_t q = 0; q < e.npairs; q++) {
      part_k[q] = (int)pp;
      part_b[q] = (int)(pp + rank_g[e.h_pair_k[q]]);
      pp += rank_g[e.h_pair_k[q]] + rank_g[e.h_pair_b[q]];
    }
  }
  std::vector<int> slot_off((size_t)kILSlots * nbox, -1);
  for (int j = 2; j <= L; j++) {
    LevelInfo& li = e.lv[j];
    for (int k = 0; k < li.nb; k++) {
      const int gk = (int)(li.gbase + k);
      e.h_red_start[gk] = (int)e.h_red_pair.size();
      int t = 0;
      for (int q = 0; q < 40; q++) {
        int b = slot_box(j, k, q);
        if (b < 0) continue;
        if (t >= kILSlots) throw Error(6, "interaction list longer than 27");
        const int gbx = (int)(li.gbase + b);
        const int lo = std::min(gk, gbx), hi = std::max(gk, gbx);
        const int pidx = pair_of.at(((int64_t)lo << 32) | hi);
        e.h_red_pair.push_back(pidx);
        // the box's t-th present slot reads the pair's product for it
        slot_off[(size_t)kILSlots * gk + t] = gk == e.h_pair_k[pidx] ? part_k[pidx] : part_b[pidx];
        t++;
        e.il_pairs++;
      }
    }
  }
  e.h_red_start[nbox] = (int)e.h_red_pair.size();
  e.m2l_cols = part;
  // the single-RHS pair kernels hold up to 256 rows per block; operators
  // with larger ranks run their matvecs through the blocked passes
  upload(e.slot_off, slot_off, s);
  upload(e.pair_k, e.h_pair_k, s);
  upload(e.pair_b, e.h_pair_b, s);
  upload(e.pair_off, build_off, s);
  {
    std::vector<PairDesc> pd(e.npairs);
    for (int64_t q = 0; q < e.npairs; q++) {
      const int gk = e.h_pair_k[q], gbx = e.h_pair_b[q];
      const int ld = e.h_pair_fact[q] >= 0 ? e.h_pair_fact[q] : rank_g[gk] + (rank_g[gk] & 1);
      pd[q] = PairDesc{rank_g[gk], rank_g[gbx], part_k[q], ld, e.h_pair_off[q], soff_g[gk],
                       soff_g[gbx], part_b[q], 0};
    }
    upload(e.pair_desc, pd, s);
  }
  // one spare double2 / double past the end: the pair kernels clamp the
  // addresses of empty blocks (rank 0) to their start, which may be the end
  e.m2l.alloc(sizeof(double) * (off + 2));
  e.m2l_part.alloc(sizeof(double) * (size_t)std::max<int64_t>(part, 2));
  AFMM_HIP(hipMemsetAsync(e.m2l_part.p, 0, e.m2l_part.bytes, s));
  if (e.npairs > 0) {
    AFMM_DISPATCH_KERNEL(e.kid, KID,
                         k_m2l_pair_build<KID><<<(unsigned)e.npairs, 64, 0, s>>>(
                             e.skx.as<double>(), e.sky.as<double>(), e.rank_g.as<int>(),
                             e.soff_g.as<int64_t>(), e.pair_k.as<int>(), e.pair_b.as<int>(),
                             e.pair_off.as<int64_t>(), e.m2l.as<double>(), e.kp));
    AFMM_LAUNCH_CHECK();
  }
  e.m2l_b48.release();
  if (e.m2l_store == M2L_STORE_B48) {
    // the FP64 blocks become row pair records (the spare one zero)
    const int64_t n = (off + 2) / 2;
    e.m2l_b48.alloc(sizeof(B48Pair) * n);
    AFMM_HIP(hipMemsetAsync(e.m2l_b48.p, 0, e.m2l_b48.bytes, s));
    if (e.npairs > 0) {
      k_m2l_b48<<<(unsigned)e.npairs, 64, 0, s>>>(
          e.m2l.as<double>(), e.rank_g.as<int>(), e.pair_k.as<int>(), e.pair_b.as<int>(),
          e.pair_off.as<int64_t>(), e.m2l_b48.as<B48Pair>(), e.pair_desc.as<PairDesc>());
      AFMM_LAUNCH_CHECK();
    }
    AFMM_HIP(hipStreamSynchronize(s));
    e.m2l.release();
  }
  AFMM_HIP(hipStreamSynchronize(s));
  e.outv.alloc(sizeof(double) * (sb + 1));
  e.inv.alloc(sizeof(double) * std::max<int64_t>(sb, 1));
  AFMM_HIP(hipMemsetAsync(e.outv.p, 0, e.outv.bytes, s));
  AFMM_HIP(hipMemsetAsync(e.inv.p, 0, e.inv.bytes, s));
  AFMM_HIP(hipStreamSynchronize(s));
}

void release_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (device < 0) g_pool.clear();
  else g_pool.erase(device);
}

}  // namespace afmm
