// api.cpp -- the C ABI of libafmm2d.so: problem set-up, the quad-tree
// interaction lists, and the host side of assemble/matvec.  The numerical
// work runs in the HIP kernels of kd_sort.hip, assemble.hip and matvec.hip;
// this file only moves data and builds index tables.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>

#include "../../include/afmm2d.h"
#include "engine.hpp"

using namespace afmm;

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return AFMM2D_OK;
  } catch (const Error& e) {
    return fail(e.status, e.what());
  } catch (const std::bad_alloc&) {
    return fail(AFMM2D_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, e.what());
  }
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// nLevels = (unsigned)(log((double)(N/M))/log(4))  (AFMM2D.cpp:7)
int levels_of(int n, int leaf) {
  if (n <= 0 || leaf <= 0) return -1;
  int q = n / leaf;
  if (q <= 0) return 0;
  return (int)(unsigned)(std::log((double)q) / std::log(4.0));
}

// Quad-tree position of box k at level j: base-4 digits, child 0 = SW
// (0,0), 1 = SE (1,0), 2 = NE (1,1), 3 = NW (0,1) (AFMM2DTree.cpp:969-989).
void box_xy(int j, int k, int& x, int& y) {
  x = y = 0;
  for (int d = j - 1; d >= 0; d--) {
    int c = (k >> (2 * d)) & 3;
    x = 2 * x + (c == 1 || c == 2);
    y = 2 * y + (c >= 2);
  }
}

int box_id(int j, int x, int y) {
  int k = 0;
  for (int d = j - 1; d >= 0; d--) {
    int xb = (x >> d) & 1, yb = (y >> d) & 1;
    int c = yb ? (xb ? 2 : 3) : (xb ? 1 : 0);
    k = 4 * k + c;
  }
  return k;
}

// Offset of slot s on the square ring of Chebyshev radius d, counter-
// clockwise from the ring's SW corner (the slot numbering of
// neighborNumbers / innerNumbers / outerNumbers drawn in
// AFMM2DTree.cpp:65-943).
void ring_offset(int d, int s, int& dx, int& dy) {
  if (s <= 2 * d) { dx = -d + s; dy = -d; return; }
  s -= 2 * d;
  if (s <= 2 * d) { dx = d; dy = -d + s; return; }
  s -= 2 * d;
  if (s <= 2 * d) { dx = d - s; dy = d; return; }
  s -= 2 * d;
  dx = -d;
  dy = d - s;
}

// Interaction lists of every box of level j.  Ring 1 holds the neighbours;
// rings 2 and 3 hold the interaction list: the children of the parent's
// neighbours that are not adjacent.  Ring 2 always lies inside the parent's
// neighbourhood; ring 3 does on the two sides away from the box's own
// parent-relative position.
void topology_level(int j, std::vector<int>& nbr, std::vector<int>& inner,
                    std::vector<int>& outer) {
  const int nb = 1 << (2 * j), side = 1 << j;
  nbr.assign(8 * (size_t)nb, -1);
  inner.assign(16 * (size_t)nb, -1);
  outer.assign(24 * (size_t)nb, -1);
  if (j == 0) return;
  for (int k = 0; k < nb; k++) {
    int ix, iy;
    box_xy(j, k, ix, iy);
    for (int d = 1; d <= 3; d++) {
      for (int s = 0; s < 8 * d; s++) {
        int dx, dy;
        ring_offset(d, s, dx, dy);
        int bx = ix + dx, by = iy + dy;
        if (bx < 0 || by < 0 || bx >= side || by >= side) continue;
        if (std::abs((bx >> 1) - (ix >> 1)) > 1 || std::abs((by >> 1) - (iy >> 1)) > 1) continue;
        int b = box_id(j, bx, by);
        if (d == 1) nbr[8 * (size_t)k + s] = b;
        else if (d == 2) inner[16 * (size_t)k + s] = b;
        else outer[24 * (size_t)k + s] = b;
      }
    }
  }
}

}  // namespace

namespace afmm {

Engine::~Engine() {
  if (device >= 0) (void)hipSetDevice(device);
  for (auto& p : ev_pool) {
    (void)hipEventDestroy(p[0]);
    (void)hipEventDestroy(p[1]);
  }
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (ev_tree) (void)hipEventDestroy(ev_tree);
  if (s_aux) (void)hipStreamDestroy(s_aux);
  if (s_tree) (void)hipStreamDestroy(s_tree);
}

void Engine::prof_begin(const char* name, hipStream_t s, hipEvent_t* a, hipEvent_t* b) {
  *a = *b = nullptr;
  if (!profiling) return;
  if (profiling == 2 && std::strcmp(name, "m2l") != 0 && std::strcmp(name, "mr_m2l") != 0) return;
  if (ev_used == ev_pool.size()) {
    std::array<hipEvent_t, 2> p;
    AFMM_HIP(hipEventCreate(&p[0]));
    AFMM_HIP(hipEventCreate(&p[1]));
    ev_pool.push_back(p);
  }
  *a = ev_pool[ev_used][0];
  *b = ev_pool[ev_used][1];
  ev_used++;
  AFMM_HIP(hipEventRecord(*a, s));
  pending.push_back(Pending{name, *a, *b});
}

void Engine::prof_end(hipStream_t s, hipEvent_t b) {
  if (!profiling || !b) return;
  AFMM_HIP(hipEventRecord(b, s));
}

void Engine::prof_collect() {
  for (auto& p : pending) {
    AFMM_HIP(hipEventSynchronize(p.b));
    float ms = 0.f;
    AFMM_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    KernelTimer& t = timers[p.name];
    t.total_ms += ms;
    t.launches += 1;
  }
  pending.clear();
  ev_used = 0;
}

// Near-field waves per SIMD for the overlapped schedule: the near field runs
// beside the HBM-bound M2L and should end about when the M2L does.  With
// R = (near-field time alone) / (M2L time alone), estimated from the
// problem's sizes and the rates measured on MI355X (~9e11 near-field kernel
// evaluations/s, ~5 TB/s for the M2L blocks and partials), the best counts
// measured were 3 at R ~ 0.55 (N = 2^20, tolerance 1e-12) and 6 at R ~ 1.2
// (tolerance 1e-8): round(5 R), clamped to 2..8.  With larger leaves (the
// published shapes: 100-391 points, a/R kernel) the count matters little
// from 4 up (N = 1e6: within 5% for 2..8; N = 1e5: 0.55 ms at 4..8,
// 0.91 ms at 2).
int near_waves_per_simd(const Engine& e) {
  double evals = 0.0;
  const int nleaf = (int)e.leaf_count.size();
  for (int a = 0; a < nleaf; a++) {
    const double na = e.leaf_count[a];
    evals += 0.5 * na * na;  // self pair: each unordered point pair once
    for (int q = 0; q < 8; q++) {
      const int b = e.nbr[e.L][8 * (size_t)a + q];
      if (b > a) evals += na * e.leaf_count[b];
    }
  }
  const double t_near = evals / 9.0e11;
  const double t_m2l = (8.0 * e.m2l_total + 16.0 * e.m2l_cols) / 5.0e12;
  if (t_m2l <= 0.0) return 8;
  return std::max(2, std::min(8, (int)std::lround(5.0 * t_near / t_m2l)));
}

// The near-field wave count measured on the assembled operator: matvecs
// back to back with each candidate (the estimate above, 2, 4 and 8 waves per
// SIMD), keeping the fastest.  After 8 warm-up matvecs the candidates take
// turns over 3 rounds (one warm-up and a timed batch of 4 each), and each
// keeps its best batch, so a clock still settling favours none of them.  The
// estimate cannot see
// every interaction between the two: the near field's persistent workgroups
// keep their registers until its queue drains, and with M2L blocks of rank
// above 64 (~200-VGPR kernels) the M2L beside it may find no room (N = 1e6,
// leaf 100, a/R, tolerance 1e-12: 1.90 ms at 8 waves, 1.66 ms at 2), while a
// near field that dominates needs its waves (N = 1e5: 0.55 ms at 4, 0.91 at
// 2).  Results do not depend on the count.
int tune_near_waves(Engine& e, int estimate) {
  std::vector<int> cand = {estimate, 2, 4, 8};
  std::sort(cand.begin(), cand.end());
  cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  const size_t nb = sizeof(double) * e.n;
  e.io_b.alloc(nb);
  e.io_y.alloc(nb);
  {
    std::vector<double> q(e.n);
    std::mt19937_64 rng(7);
    for (double& v : q) v = (double)(rng() >> 11) * 0x1.0p-53 - 0.5;
    AFMM_HIP(hipMemcpy(e.io_b.p, q.data(), nb, hipMemcpyHostToDevice));
  }
  hipStream_t s = nullptr;
  hipEvent_t a, b;
  AFMM_HIP(hipEventCreate(&a));
  AFMM_HIP(hipEventCreate(&b));
  const int saved_prof = e.profiling;
  e.profiling = 0;
  e.p2p_waves_per_simd = estimate;
  for (int i = 0; i < 8; i++) matvec_device(e, e.io_b.as<double>(), e.io_y.as<double>(), s);
  std::vector<float> t(cand.size(), 1e30f);
  for (int round = 0; round < 3; round++) {
    for (size_t ci = 0; ci < cand.size(); ci++) {
      e.p2p_waves_per_simd = cand[ci];
      matvec_device(e, e.io_b.as<double>(), e.io_y.as<double>(), s);
      // back to back, as a caller's loop runs them
      AFMM_HIP(hipEventRecord(a, s));
      for (int i = 0; i < 4; i++) matvec_device(e, e.io_b.as<double>(), e.io_y.as<double>(), s);
      AFMM_HIP(hipEventRecord(b, s));
      AFMM_HIP(hipEventSynchronize(b));
      float x = 0.f;
      AFMM_HIP(hipEventElapsedTime(&x, a, b));
      t[ci] = std::min(t[ci], x / 4);
    }
  }
  int best = estimate;
  float best_ms = 1e30f;
  for (size_t ci = 0; ci < cand.size(); ci++)
    if (t[ci] < best_ms) {
      best_ms = t[ci];
      best = cand[ci];
    }
  // the estimate when it is within 1% of the fastest
  for (size_t ci = 0; ci < cand.size(); ci++)
    if (cand[ci] == estimate && t[ci] <= best_ms * 1.01f) best = estimate;
  e.profiling = saved_prof;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

}  // namespace afmm

// e: the operator, or the symmetric part of a non-symmetric user kernel;
// e2: that kernel's antisymmetric part (null otherwise).  y = e(b) + e2(b).
struct afmm2d_handle_s {
  std::unique_ptr<Engine> e;
  std::unique_ptr<Engine> e2;
  // set once every engine has assembled: a failure in the second engine
  // leaves the handle unassembled, so no matvec launches on a partial one
  bool assembled = false;
  // recorded at the end of every call's device work; a call on another
  // stream first waits for it (the calls share the handle's buffers)
  hipEvent_t done = nullptr;
  hipStream_t last = nullptr;
  bool have_done = false;
  ~afmm2d_handle_s() {
    if (done) (void)hipEventDestroy(done);
  }
};

namespace {

// Whether the user kernel is symmetric on this point set: K(i, j) == K(j, i)
// bit for bit on 4096 sampled pairs (a fixed seed), about half of them
// within a leaf's neighbourhood in KD order, where most kernels vary fastest.
bool user_kernel_symmetric(const Engine& e) {
  std::mt19937_64 rng(12345);
  for (int t = 0; t < 4096; t++) {
    const int i = (int)(rng() % (uint64_t)e.n);
    int j = (int)(rng() % (uint64_t)e.n);
    if (t & 1) j = std::min(e.n - 1, i + (int)(rng() % 64));
    const double xi = e.h_xy[2 * (size_t)i], yi = e.h_xy[2 * (size_t)i + 1];
    const double xj = e.h_xy[2 * (size_t)j], yj = e.h_xy[2 * (size_t)j + 1];
    const double a = kernel_host(K_USER, e.kp, xi, yi, xj, yj);
    const double b = kernel_host(K_USER, e.kp, xj, yj, xi, yi);
    if (!(a == b)) return false;
  }
  return true;
}

// Orders stream s after the handle's previous call (on another stream).
void begin_on(afmm2d_handle h, hipStream_t s) {
  if (h->have_done && h->last != s) AFMM_HIP(hipStreamWaitEvent(s, h->done, 0));
}
void end_on(afmm2d_handle h, hipStream_t s) {
  AFMM_HIP(hipEventRecord(h->done, s));
  h->last = s;
  h->have_done = true;
}

// Ranks the single-RHS M2L pair kernels hold (their largest tile class).
constexpr int kSingleMaxRank = 256;

bool single_ok(afmm2d_handle h) {
  return h->e->max_rank <= kSingleMaxRank && (!h->e2 || h->e2->max_rank <= kSingleMaxRank);
}

// y = the operator applied to b, device arrays in caller order, on stream s.
// Operators with ranks beyond the single-RHS kernels' run the FP64 blocked
// passes for the one column.
void run_matvec(afmm2d_handle h, const double* d_b, double* d_y, hipStream_t s) {
  if (!single_ok(h)) {
    mr_matvec(*h->e, d_b, h->e->n, d_y, h->e->n, 1, PREC_F64, s);
    if (h->e2) {
      Engine& e2 = *h->e2;
      e2.io_y.alloc(sizeof(double) * e2.n);
      mr_matvec(e2, d_b, e2.n, e2.io_y.as<double>(), e2.n, 1, PREC_F64, s);
      add_device(d_y, e2.io_y.as<const double>(), e2.n, s);
    }
    return;
  }
  matvec_device(*h->e, d_b, d_y, s);
  if (h->e2) {
    Engine& e2 = *h->e2;
    e2.io_y.alloc(sizeof(double) * e2.n);
    matvec_device(e2, d_b, e2.io_y.as<double>(), s);
    add_device(d_y, e2.io_y.as<const double>(), e2.n, s);
  }
}

constexpr int kModePrecisionMask = 0xFF;
constexpr int kModeBlocked = 0x100;  // AFMM2D_MATVEC_BLOCKED

// Y = K B for nrhs columns (device, caller order) in the given mode: one FP64
// column runs the single-RHS passes, anything else the blocked passes in
// chunks of up to 32 columns.
void run_matvec_ex(afmm2d_handle h, const double* d_b, int64_t ldb, double* d_y, int64_t ldy,
                   int nrhs, int mode, hipStream_t s) {
  const int prec = mode & kModePrecisionMask;
  if (prec == PREC_F64 && nrhs == 1 && !(mode & kModeBlocked)) {
    run_matvec(h, d_b, d_y, s);
    return;
  }
  for (int c0 = 0; c0 < nrhs; c0 += kMrMaxCols) {
    const int nc = std::min(kMrMaxCols, nrhs - c0);
    const double* b = d_b + (int64_t)c0 * ldb;
    double* y = d_y + (int64_t)c0 * ldy;
    mr_matvec(*h->e, b, ldb, y, ldy, nc, prec, s);
    if (h->e2) {
      Engine& e2 = *h->e2;
      e2.io_y.alloc(sizeof(double) * (size_t)e2.n * nc);
      mr_matvec(e2, b, ldb, e2.io_y.as<double>(), e2.n, nc, prec, s);
      add_cols_device(y, ldy, e2.io_y.as<const double>(), e2.n, e2.n, nc, s);
    }
  }
}

int check_mode(int mode) {
  const int prec = mode & kModePrecisionMask;
  if (prec < 0 || prec >= kPrecisions || (mode & ~(kModePrecisionMask | kModeBlocked)))
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "unknown precision mode");
  return AFMM2D_OK;
}

}  // namespace

extern "C" {

const char* afmm2d_last_error(void) { return g_last_error.c_str(); }

const char* afmm2d_status_string(int s) {
  switch (s) {
    case AFMM2D_OK: return "ok";
    case AFMM2D_ERR_INVALID_ARGUMENT: return "invalid argument";
    case AFMM2D_ERR_LEVELS: return "fewer than 2 tree levels";
    case AFMM2D_ERR_STATE: return "invalid state";
    case AFMM2D_ERR_OUT_OF_MEMORY: return "out of memory";
    case AFMM2D_ERR_HIP: return "HIP error";
    case AFMM2D_ERR_UNSUPPORTED: return "unsupported size";
    case AFMM2D_ERR_NOT_FOUND: return "not found";
  }
  return "unknown status";
}

int afmm2d_create(int n, int leaf, int tol_pow, const double* loc, int layout, int kernel_id,
                  const double* kparams, int device, afmm2d_handle* out) {
  if (!out) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!loc) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "loc is NULL");
  if (n <= 1 || leaf <= 0) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "need N > 1 and leaf > 0");
  if ((kernel_id < 0 || kernel_id >= kNumKernels) && kernel_id != AFMM2D_KERNEL_USER_SYMMETRIC)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "unknown kernel id");
  // the user kernel declared symmetric runs as one operator, as it is
  const bool declared_symmetric = kernel_id == AFMM2D_KERNEL_USER_SYMMETRIC;
  if (declared_symmetric) kernel_id = K_USER;
  if (layout != AFMM2D_LOC_INTERLEAVED && layout != AFMM2D_LOC_COLUMN_MAJOR)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "unknown location layout");
  if (tol_pow < 0 || tol_pow > 300) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "tol_pow out of range");
  const int L = levels_of(n, leaf);
  if (L < 2)
    return fail(AFMM2D_ERR_LEVELS, "nLevels = log4(N/M) = " + std::to_string(L) +
                                       " < 2: no interaction lists exist");
  if (L > 13) return fail(AFMM2D_ERR_UNSUPPORTED, "more than 13 levels");
  if (n / (1 << (2 * L)) + 1 > 512)
    return fail(AFMM2D_ERR_UNSUPPORTED, "leaves above 512 points are not supported");
  return guarded([&] {
  // one engine per operator: the kernel itself, or for a non-symmetric user
  // kernel its symmetric and its antisymmetric part (same points and tree)
  auto make_engine = [&](int kid) {
    auto e = std::make_unique<Engine>();
    if (device < 0) AFMM_HIP(hipGetDevice(&device));
    AFMM_HIP(hipSetDevice(device));
    e->device = device;
    e->n = n;
    e->leaf = leaf;
    e->tol_pow = tol_pow;
    e->tol = std::pow(10.0, -1.0 * tol_pow);  // ACA.cpp:7
    e->L = L;
    e->kid = kid;
    e->kid_full = kernel_id;
    e->kp.p0 = (kparams && kparams[0] > 0.0) ? kparams[0] : 0.01;
    e->kp.a = std::ldexp(1.0, -L);  // a = 0.5^L (AFMM2DTree.cpp:13-19)
    if (kparams && kernel_id == AFMM2D_KERNEL_REG_INV_R && kparams[1] > 0.0) e->kp.a = kparams[1];
    for (int i = 0; i < 4; i++) e->kp.u[i] = kparams ? kparams[i] : 0.0;
    AFMM_HIP(hipDeviceGetAttribute(&e->num_cus, hipDeviceAttributeMultiprocessorCount, device));
    AFMM_HIP(hipDeviceGetAttribute(&e->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    if (const char* v = std::getenv("AFMM2D_NEAR_OVERLAP")) e->near_overlap = std::atoi(v) != 0;
    if (const char* v = std::getenv("AFMM2D_TREE_MODE"))
      e->tree_mode = std::max(0, std::min(3, std::atoi(v)));
    if (const char* v = std::getenv("AFMM2D_NEAR_TUNE")) e->near_tune = std::atoi(v) != 0;
    if (const char* v = std::getenv("AFMM2D_TREE_FUSED")) e->tree_fused = std::atoi(v) != 0;
    if (const char* v = std::getenv("AFMM2D_P2P_WAVES")) {
      e->p2p_waves_per_simd = std::max(0, std::min(16, std::atoi(v)));
      e->p2p_waves_fixed = true;
    }
    // one internal stream per handle (a process has few hardware queues,
    // and streams that share one run in order): it carries the set-up work
    // and the near field beside the caller's stream
    AFMM_HIP(hipStreamCreateWithFlags(&e->s_aux, hipStreamNonBlocking));
    AFMM_HIP(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    AFMM_HIP(hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming));
    AFMM_HIP(hipEventCreateWithFlags(&e->ev_tree, hipEventDisableTiming));
    double t0 = now_s();
    std::vector<double> hx(n), hy(n);
    for (int i = 0; i < n; i++) {
      hx[i] = layout == AFMM2D_LOC_INTERLEAVED ? loc[2 * (size_t)i] : loc[i];
      hy[i] = layout == AFMM2D_LOC_INTERLEAVED ? loc[2 * (size_t)i + 1] : loc[(size_t)n + i];
      if (!std::isfinite(hx[i]) || !std::isfinite(hy[i]))
        throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "non-finite coordinate at index " + std::to_string(i));
    }
    kd_partition(*e, hx.data(), hy.data());
    // Leaf ranges (AFMM2DTree.cpp:1099-1118) and KD-order box numbers
    // (KDTree.cpp:211-223).
    const int nleaf = 1 << (2 * L);
    e->kd_boxnum.resize(nleaf);
    static const int map[4] = {0, 3, 1, 2};
    for (int t = 0; t < nleaf; t++) {
      int b = 0;
      for (int d = L - 1; d >= 0; d--) b = 4 * b + map[(t >> (2 * d)) & 3];
      e->kd_boxnum[t] = b;
    }
    e->leaf_start.assign(nleaf, 0);
    e->leaf_count.assign(nleaf, 0);
    int st = 0;
    for (int t = 0; t < nleaf; t++) {
      int b = e->kd_boxnum[t];
      e->leaf_start[b] = st;
      e->leaf_count[b] = e->kd_counts[t];
      e->max_leaf = std::max(e->max_leaf, e->kd_counts[t]);
      st += e->kd_counts[t];
    }
    e->nbr.resize(L + 1);
    e->inner.resize(L + 1);
    e->outer.resize(L + 1);
    for (int j = 0; j <= L; j++) topology_level(j, e->nbr[j], e->inner[j], e->outer[j]);
    e->near_pairs = nleaf;
    for (int v : e->nbr[L]) e->near_pairs += v >= 0;
    {
      // unordered near-field pairs: self pairs and neighbour pairs A < B.
      // Partials are leaf-major: leaf A's slot s (0 self, 1 + q neighbour q)
      // holds count[A] values at 9 start[A] + s count[A]; a pair writes A's
      // sums to A's slot for B and B's sums to B's slot for A.
      std::vector<int> pair, ppart;
      auto slot_off = [&](int a, int s) {
        return (int64_t)9 * e->leaf_start[a] + (int64_t)s * e->leaf_count[a];
      };
      if ((int64_t)9 * e->n > INT32_MAX)
        throw Error(AFMM2D_ERR_UNSUPPORTED, "near-field partials exceed 2^31");
      for (int a = 0; a < nleaf; a++) {
        pair.push_back(a);
        pair.push_back(a);
        ppart.push_back((int)slot_off(a, 0));
        ppart.push_back((int)slot_off(a, 0));
        for (int q = 0; q < 8; q++) {
          int b = e->nbr[L][8 * (size_t)a + q];
          if (b <= a) continue;
          int qb = -1;
          for (int u = 0; u < 8; u++)
            if (e->nbr[L][8 * (size_t)b + u] == a) qb = u;
          if (qb < 0) throw Error(AFMM2D_ERR_UNSUPPORTED, "neighbour relation not symmetric");
          pair.push_back(a);
          pair.push_back(b);
          ppart.push_back((int)slot_off(a, 1 + q));
          ppart.push_back((int)slot_off(b, 1 + qb));
        }
      }
      e->near_upairs = (int64_t)pair.size() / 2;
      e->near_part_total = (int64_t)9 * e->n;
      e->near_pair.alloc(sizeof(int) * pair.size());
      e->near_pair_part.alloc(sizeof(int) * ppart.size());
      e->near_part.alloc(sizeof(double) * (size_t)std::max<int64_t>(e->near_part_total, 1));
      AFMM_HIP(hipMemsetAsync(e->near_part.p, 0, e->near_part.bytes, e->s_aux));
      AFMM_HIP(hipMemcpyAsync(e->near_pair.p, pair.data(), sizeof(int) * pair.size(),
                              hipMemcpyHostToDevice, e->s_aux));
      AFMM_HIP(hipMemcpyAsync(e->near_pair_part.p, ppart.data(), sizeof(int) * ppart.size(),
                              hipMemcpyHostToDevice, e->s_aux));
      AFMM_HIP(hipStreamSynchronize(e->s_aux));
    }
    // The near field's tables, in extended precision: log2_near's (RN(1/c_j),
    // -log2 RN(1/c_j) - 1) for c_0 = 1 and c_j = 1 + (j + 1/2)/256, then
    // exp_near's RN(2^(j/64)).
    {
      std::vector<double> tab(2 * kNearLogTab + kNearExpTab);
      for (int j = 0; j < kNearLogTab; j++) {
        const long double c = j == 0 ? 1.0L : 1.0L + ((long double)j + 0.5L) / kNearLogTab;
        const double invc = (double)(1.0L / c);
        tab[2 * j] = invc;
        tab[2 * j + 1] = (double)(-log2l((long double)invc) - 1.0L);
      }
      for (int j = 0; j < kNearExpTab; j++)
        tab[2 * kNearLogTab + j] = (double)exp2l((long double)j / kNearExpTab);
      e->near_tab.alloc(sizeof(double) * tab.size());
      AFMM_HIP(hipMemcpyAsync(e->near_tab.p, tab.data(), sizeof(double) * tab.size(),
                              hipMemcpyHostToDevice, e->s_aux));
      AFMM_HIP(hipStreamSynchronize(e->s_aux));
    }
    hipStream_t s = e->s_aux;
    e->d_leaf_start.alloc(sizeof(int) * nleaf);
    e->d_leaf_count.alloc(sizeof(int) * nleaf);
    e->d_leaf_nbr.alloc(sizeof(int) * 8 * (size_t)nleaf);
    AFMM_HIP(hipMemcpyAsync(e->d_leaf_start.p, e->leaf_start.data(), sizeof(int) * nleaf,
                            hipMemcpyHostToDevice, s));
    AFMM_HIP(hipMemcpyAsync(e->d_leaf_count.p, e->leaf_count.data(), sizeof(int) * nleaf,
                            hipMemcpyHostToDevice, s));
    AFMM_HIP(hipMemcpyAsync(e->d_leaf_nbr.p, e->nbr[L].data(), sizeof(int) * 8 * (size_t)nleaf,
                            hipMemcpyHostToDevice, s));
    AFMM_HIP(hipStreamSynchronize(s));
    {
      // caller index -> KD-sorted index, on the device
      std::vector<int> ip(n);
      for (int i = 0; i < n; i++) ip[e->h_perm[i]] = i;
      e->inv_perm_d.alloc(sizeof(int) * (size_t)n);
      AFMM_HIP(hipMemcpy(e->inv_perm_d.p, ip.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
    }
    e->create_seconds = now_s() - t0;
    return e;
  };
    auto e = make_engine(kernel_id);
    std::unique_ptr<Engine> e2;
    if (kernel_id == K_USER && !declared_symmetric) {
      // not declared symmetric: the operator is its symmetric part plus its
      // antisymmetric part, whatever a sample of entries suggests (a kernel
      // asymmetric on a small region passes any sample).  A sampled check
      // that finds K(i, j) != K(j, i) is kept as a diagnostic for callers
      // that declared symmetry wrongly: see afmm2d_check_symmetry.
      e->kid = K_USER_SYM;
      e2 = make_engine(K_USER_ANTI);
    }
    auto* hh = new afmm2d_handle_s{std::move(e), std::move(e2)};
    if (hipEventCreateWithFlags(&hh->done, hipEventDisableTiming) != hipSuccess) {
      delete hh;
      throw Error(AFMM2D_ERR_HIP, "hipEventCreateWithFlags failed");
    }
    *out = hh;
  });
}

int afmm2d_destroy(afmm2d_handle h) {
  if (!h) return AFMM2D_OK;
  return guarded([&] {
    for (Engine* e : {h->e.get(), h->e2.get()}) {
      if (!e) continue;
      (void)hipSetDevice(e->device);
      (void)hipStreamSynchronize(e->s_aux);
      if (e->s_tree) (void)hipStreamSynchronize(e->s_tree);
    }
    delete h;
  });
}

int afmm2d_assemble(afmm2d_handle h) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "handle is NULL");
  if (h->assembled) return AFMM2D_OK;
  return guarded([&] {
    if (h->have_done) AFMM_HIP(hipEventSynchronize(h->done));
    for (Engine* ep : {h->e.get(), h->e2.get()}) {
      if (!ep) continue;
      Engine& e = *ep;
      AFMM_HIP(hipSetDevice(e.device));
      if (e.assembled) continue;
      double t0 = now_s();
      assemble_tree(e);
      e.assembled = true;
      if (!e.p2p_waves_fixed) {
        e.p2p_waves_per_simd = near_waves_per_simd(e);
        if (e.near_tune && e.max_rank <= kSingleMaxRank)
          e.p2p_waves_per_simd = tune_near_waves(e, e.p2p_waves_per_simd);
      }
      e.assemble_seconds = now_s() - t0;
    }
    h->assembled = true;
  });
}

int afmm2d_release_scratch(int device) {
  return guarded([&] { release_scratch(device); });
}

int afmm2d_matvec_device(afmm2d_handle h, const double* d_b, double* d_y, void* stream) {
  if (!h || !d_b || !d_y) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "matvec before assemble");
  return guarded([&] {
    Engine& e = *h->e;
    AFMM_HIP(hipSetDevice(e.device));
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)0;
    begin_on(h, s);
    run_matvec(h, d_b, d_y, s);
    end_on(h, s);
  });
}

int afmm2d_matvec_device_ex(afmm2d_handle h, const double* d_b, int64_t ldb, double* d_y,
                            int64_t ldy, int nrhs, int mode, void* stream) {
  if (!h || !d_b || !d_y || nrhs <= 0) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  if (ldb < h->e->n || ldy < h->e->n)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "leading dimension below N");
  if (check_mode(mode) != AFMM2D_OK) return AFMM2D_ERR_INVALID_ARGUMENT;
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "matvec before assemble");
  return guarded([&] {
    Engine& e = *h->e;
    AFMM_HIP(hipSetDevice(e.device));
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)0;
    begin_on(h, s);
    run_matvec_ex(h, d_b, ldb, d_y, ldy, nrhs, mode, s);
    end_on(h, s);
  });
}

int afmm2d_matvec_ex(afmm2d_handle h, const double* b, int64_t ldb, double* y, int64_t ldy,
                     int nrhs, int mode) {
  if (!h || !b || !y || nrhs <= 0) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  if (ldb < h->e->n || ldy < h->e->n)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "leading dimension below N");
  if (check_mode(mode) != AFMM2D_OK) return AFMM2D_ERR_INVALID_ARGUMENT;
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "matvec before assemble");
  return guarded([&] {
    Engine& e = *h->e;
    AFMM_HIP(hipSetDevice(e.device));
    const int n = e.n;
    const size_t col = sizeof(double) * n;
    // device columns packed with leading dimension n
    e.io_b.alloc(col * nrhs);
    e.io_y.alloc(col * nrhs);
    hipStream_t s = nullptr;  // the null stream plays the caller's stream
    begin_on(h, s);
    if (ldb == n) {
      AFMM_HIP(hipMemcpyAsync(e.io_b.p, b, col * nrhs, hipMemcpyHostToDevice, s));
    } else {
      for (int r = 0; r < nrhs; r++)
        AFMM_HIP(hipMemcpyAsync(e.io_b.as<double>() + (size_t)r * n, b + (size_t)r * ldb, col,
                                hipMemcpyHostToDevice, s));
    }
    const int prec = mode & kModePrecisionMask;
    if (prec == PREC_F64 && !(mode & kModeBlocked)) {
      // one column at a time through the single-RHS passes (the reference's
      // computeMatVecProduct per vector)
      for (int r = 0; r < nrhs; r++)
        run_matvec(h, e.io_b.as<double>() + (size_t)r * n, e.io_y.as<double>() + (size_t)r * n, s);
    } else {
      run_matvec_ex(h, e.io_b.as<double>(), n, e.io_y.as<double>(), n, nrhs, mode, s);
    }
    if (ldy == n) {
      AFMM_HIP(hipMemcpyAsync(y, e.io_y.p, col * nrhs, hipMemcpyDeviceToHost, s));
    } else {
      for (int r = 0; r < nrhs; r++)
        AFMM_HIP(hipMemcpyAsync(y + (size_t)r * ldy, e.io_y.as<double>() + (size_t)r * n, col,
                                hipMemcpyDeviceToHost, s));
    }
    end_on(h, s);
    AFMM_HIP(hipStreamSynchronize(s));
    if (e.profiling) e.prof_collect();
    e.last_qs.resize(n);
    e.last_ps.resize(n);
    for (int i = 0; i < n; i++) {
      e.last_qs[i] = b[e.h_perm[i]];
      e.last_ps[i] = y[e.h_perm[i]];
    }
    e.have_matvec = true;
  });
}

int afmm2d_matvec(afmm2d_handle h, const double* b, double* y, int nrhs) {
  if (!h || !b || !y || nrhs <= 0) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  return afmm2d_matvec_ex(h, b, h->e->n, y, h->e->n, nrhs, AFMM2D_PRECISION_FP64);
}

int afmm2d_prepare_precision(afmm2d_handle h, int precision) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle");
  if (precision < 0 || precision >= kPrecisions)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "unknown precision");
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "not assembled");
  return guarded([&] {
    for (Engine* ep : {h->e.get(), h->e2.get()}) {
      if (!ep) continue;
      AFMM_HIP(hipSetDevice(ep->device));
      if (h->have_done) AFMM_HIP(hipEventSynchronize(h->done));
      mr_prepare(*ep, precision);
    }
  });
}

struct afmm2d_batch_s {
  MrBatch* b = nullptr;
  ~afmm2d_batch_s() { mr_batch_destroy(b); }
};

int afmm2d_batch_assemble(const afmm2d_handle* handles, int count) {
  if (!handles || count < 0) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  for (int i = 0; i < count; i++) {
    if (!handles[i]) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle in batch");
    const int st = afmm2d_assemble(handles[i]);
    if (st != AFMM2D_OK) return st;
  }
  return AFMM2D_OK;
}

int afmm2d_batch_create(const afmm2d_handle* handles, int count, const double* const* d_b,
                        const int64_t* ldb, double* const* d_y, const int64_t* ldy, int nrhs,
                        int precision, afmm2d_batch* out) {
  if (!out) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL output");
  *out = nullptr;
  if (!handles || count <= 0 || !d_b || !ldb || !d_y || !ldy || nrhs < 1 || nrhs > kMrMaxCols)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  if (precision < 0 || precision >= kPrecisions)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "unknown precision");
  for (int i = 0; i < count; i++) {
    afmm2d_handle h = handles[i];
    if (!h || !d_b[i] || !d_y[i]) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL in batch");
    if (!h->assembled) return fail(AFMM2D_ERR_STATE, "batch of an unassembled handle");
    if (h->e2) return fail(AFMM2D_ERR_UNSUPPORTED, "batch of a non-symmetric user kernel");
    if (ldb[i] < h->e->n || ldy[i] < h->e->n)
      return fail(AFMM2D_ERR_INVALID_ARGUMENT, "leading dimension below N");
  }
  return guarded([&] {
    std::vector<Engine*> eng;
    for (int i = 0; i < count; i++) {
      afmm2d_handle h = handles[i];
      AFMM_HIP(hipSetDevice(h->e->device));
      if (h->have_done) AFMM_HIP(hipEventSynchronize(h->done));
      eng.push_back(h->e.get());
    }
    auto bb = std::make_unique<afmm2d_batch_s>();
    bb->b = mr_batch_create(eng, d_b, ldb, d_y, ldy, nrhs, precision);
    *out = bb.release();
  });
}

int afmm2d_batch_matvec(afmm2d_batch b, void* stream) {
  if (!b) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL batch");
  return guarded([&] { mr_batch_run(b->b, stream ? (hipStream_t)stream : (hipStream_t)0); });
}

int afmm2d_batch_tune(afmm2d_batch b, void* stream, double* ms_per_run) {
  if (!b) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL batch");
  return guarded([&] {
    const double ms = mr_batch_tune(b->b, stream ? (hipStream_t)stream : (hipStream_t)0);
    if (ms_per_run) *ms_per_run = ms;
  });
}

int afmm2d_batch_capture(afmm2d_batch b, int on) {
  if (!b) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL batch");
  return guarded([&] { mr_batch_graph(b->b, on != 0); });
}

int afmm2d_batch_get_info(afmm2d_batch b, afmm2d_batch_info* info) {
  if (!b || !info) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  const MrBatchInfo in = mr_batch_info(b->b);
  info->launches = in.launches;
  info->depth = in.depth;
  info->xcd = in.xcd;
  info->near_waves = in.near_waves;
  info->rows_m2l = in.rows_m2l;
  info->captured = in.captured;
  info->tuned = in.tuned;
  info->nproblems = in.nproblems;
  return AFMM2D_OK;
}

int afmm2d_batch_destroy(afmm2d_batch b) {
  if (!b) return AFMM2D_OK;
  return guarded([&] { delete b; });
}

struct afmm2d_rank_model_s {
  std::unique_ptr<RankModel> m;
};

int afmm2d_rank_model_create(int device, uint64_t seed, afmm2d_rank_model* out) {
  if (!out) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL output");
  *out = nullptr;
  return guarded([&] {
    int dev = device;
    if (dev < 0) AFMM_HIP(hipGetDevice(&dev));
    auto m = std::make_unique<afmm2d_rank_model_s>();
    m->m.reset(rank_model_create(dev, seed));
    *out = m.release();
  });
}

int afmm2d_rank_model_destroy(afmm2d_rank_model m) {
  if (!m) return AFMM2D_OK;
  return guarded([&] { delete m; });
}

int afmm2d_rank_model_get_params(afmm2d_rank_model m, float* params) {
  if (!m || !params) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  return guarded([&] { rank_model_get_params(*m->m, params); });
}

int afmm2d_rank_model_set_params(afmm2d_rank_model m, const float* params) {
  if (!m || !params) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  return guarded([&] { rank_model_set_params(*m->m, params); });
}

int afmm2d_rank_model_predict(afmm2d_rank_model m, const float* x, int64_t n, float* out) {
  if (!m || n < 0 || (n > 0 && (!x || !out))) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  return guarded([&] { rank_model_predict(*m->m, x, n, out); });
}

int afmm2d_rank_model_gradient(afmm2d_rank_model m, const float* x, const float* y, int n,
                               float* grad, double* loss) {
  if (!m || !x || !y || !grad || n <= 0) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  return guarded([&] {
    const double l = rank_model_gradient(*m->m, x, y, n, grad);
    if (loss) *loss = l;
  });
}

int afmm2d_rank_model_train(afmm2d_rank_model m, const float* x, const float* y, int64_t n,
                            int epochs, int batch, float lr, uint64_t seed, double* loss) {
  if (!m || !x || !y || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0f))
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  return guarded([&] {
    const double l = rank_model_train(*m->m, x, y, n, epochs, batch, lr, seed);
    if (loss) *loss = l;
  });
}

int afmm2d_check_symmetry(afmm2d_handle h, int* symmetric) {
  if (!h || !symmetric) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  return guarded([&] {
    *symmetric = h->e->kid_full != K_USER || user_kernel_symmetric(*h->e) ? 1 : 0;
  });
}

int afmm2d_set_m2l_factoring(afmm2d_handle h, int on) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle");
  if (h->assembled) return fail(AFMM2D_ERR_STATE, "M2L factoring set after assembly");
  for (Engine* e : {h->e.get(), h->e2.get()})
    if (e) e->m2l_factor = on != 0;
  return AFMM2D_OK;
}

int afmm2d_set_m2l_storage(afmm2d_handle h, int storage) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle");
  if (storage != AFMM2D_M2L_FP64 && storage != AFMM2D_M2L_BFP48)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "unknown M2L storage");
  if (h->assembled) return fail(AFMM2D_ERR_STATE, "M2L storage set after assembly");
  for (Engine* e : {h->e.get(), h->e2.get()})
    if (e) e->m2l_store = storage == AFMM2D_M2L_BFP48 ? M2L_STORE_B48 : M2L_STORE_F64;
  return AFMM2D_OK;
}

int afmm2d_set_rank_model(afmm2d_handle h, afmm2d_rank_model m, float margin) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle");
  if (!std::isfinite(margin)) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "margin not finite");
  if (h->assembled) return fail(AFMM2D_ERR_STATE, "rank model set after assembly");
  if (m && m->m->device != h->e->device)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "rank model on another device");
  for (Engine* e : {h->e.get(), h->e2.get()}) {
    if (!e) continue;
    e->rank_model = m ? m->m.get() : nullptr;
    e->rank_margin = margin;
  }
  return AFMM2D_OK;
}

int afmm2d_get_rank_samples(afmm2d_handle h, float* features, int* ranks, int64_t capacity,
                            int64_t* count) {
  if (!h || !count) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "not assembled");
  const Engine& e = *h->e;
  const int64_t n = (int64_t)e.rank_label.size();
  *count = n;
  if (capacity < n) return AFMM2D_OK;
  if (features) std::memcpy(features, e.rank_feat.data(), sizeof(float) * kRmIn * (size_t)n);
  if (ranks) std::memcpy(ranks, e.rank_label.data(), sizeof(int) * (size_t)n);
  return AFMM2D_OK;
}

int afmm2d_levels(afmm2d_handle h, int* levels) {
  if (!h || !levels) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  *levels = h->e->L;
  return AFMM2D_OK;
}

int afmm2d_get_permutation(afmm2d_handle h, int64_t* perm) {
  if (!h || !perm) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  std::memcpy(perm, h->e->h_perm.data(), sizeof(int64_t) * h->e->n);
  return AFMM2D_OK;
}

int afmm2d_get_sorted_locations(afmm2d_handle h, double* xy) {
  if (!h || !xy) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  std::memcpy(xy, h->e->h_xy.data(), sizeof(double) * 2 * (size_t)h->e->n);
  return AFMM2D_OK;
}

int afmm2d_get_leaf_counts(afmm2d_handle h, int* counts) {
  if (!h || !counts) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  std::memcpy(counts, h->e->kd_counts.data(), sizeof(int) * h->e->kd_counts.size());
  return AFMM2D_OK;
}

int afmm2d_get_leaf_range(afmm2d_handle h, int box, int* start, int* count) {
  if (!h || !start || !count) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  if (box < 0 || box >= (int)h->e->leaf_start.size())
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "box out of range");
  *start = h->e->leaf_start[box];
  *count = h->e->leaf_count[box];
  return AFMM2D_OK;
}

int afmm2d_get_interactions(afmm2d_handle h, int level, int* nbr, int* inner, int* outer) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle");
  Engine& e = *h->e;
  if (level < 0 || level > e.L) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "level out of range");
  if (nbr) std::memcpy(nbr, e.nbr[level].data(), sizeof(int) * e.nbr[level].size());
  if (inner) std::memcpy(inner, e.inner[level].data(), sizeof(int) * e.inner[level].size());
  if (outer) std::memcpy(outer, e.outer[level].data(), sizeof(int) * e.outer[level].size());
  return AFMM2D_OK;
}

int afmm2d_get_ranks(afmm2d_handle h, int level, int* ranks) {
  if (!h || !ranks) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "not assembled");
  if (level < 2 || level > e.L) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "level out of range");
  std::memcpy(ranks, e.lv[level].rank.data(), sizeof(int) * e.lv[level].nb);
  return AFMM2D_OK;
}

int afmm2d_get_skeleton(afmm2d_handle h, int level, int box, int* idx, int capacity, int* count) {
  if (!h || !count) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "not assembled");
  if (level < 2 || level > e.L || box < 0 || box >= e.lv[level].nb)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "level/box out of range");
  return guarded([&] {
    const LevelInfo& li = e.lv[level];
    int r = li.rank[box];
    *count = r;
    if (r > capacity || !idx) throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "capacity too small");
    AFMM_HIP(hipMemcpy(idx, e.skidx.as<int>() + li.soff[box], sizeof(int) * r, hipMemcpyDeviceToHost));
  });
}

int afmm2d_get_l2p(afmm2d_handle h, int level, int box, double* out, int64_t capacity, int* rows,
                   int* cols) {
  if (!h || !rows || !cols) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "not assembled");
  if (level < 2 || level > e.L || box < 0 || box >= e.lv[level].nb)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "level/box out of range");
  return guarded([&] {
    const LevelInfo& li = e.lv[level];
    *rows = li.nin[box];
    *cols = li.rank[box];
    const int n = *rows, r = *cols, nc = n - r;
    int64_t m = (int64_t)n * r;
    if (m > capacity || (!out && m > 0)) throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "capacity too small");
    if (m == 0) return;
    // stored rows plus the identity on the skeleton rows, in box row order
    std::vector<double> st((size_t)nc * r);
    std::vector<int> rmap(n);
    const int64_t roff = level == e.L ? e.leaf_start[box] : e.n + e.lv[level + 1].soff[4 * (size_t)box];
    if (nc > 0)
      AFMM_HIP(hipMemcpy(st.data(), e.l2p_rm.as<double>() + li.l2poff[box], sizeof(double) * st.size(),
                         hipMemcpyDeviceToHost));
    AFMM_HIP(hipMemcpy(rmap.data(), e.rowmap.as<int>() + roff, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::fill(out, out + m, 0.0);
    for (int t = 0; t < nc; t++)
      std::copy(st.begin() + (size_t)t * r, st.begin() + (size_t)(t + 1) * r, out + (int64_t)rmap[t] * r);
    for (int c = 0; c < r; c++) out[(int64_t)rmap[nc + c] * r + c] = 1.0;
  });
}

int afmm2d_get_m2l_storage(afmm2d_handle h, int level, int box, int slot, int* factor_rank,
                           int* box_first) {
  if (!h || !factor_rank || !box_first) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "not assembled");
  if (level < 2 || level > e.L || box < 0 || box >= e.lv[level].nb || slot < 0 || slot >= 40)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "level/box/slot out of range");
  const LevelInfo& li = e.lv[level];
  auto slot_box = [&](int s) {
    return s < 16 ? e.inner[level][16 * (size_t)box + s] : e.outer[level][24 * (size_t)box + s - 16];
  };
  if (slot_box(slot) < 0) return fail(AFMM2D_ERR_NOT_FOUND, "empty interaction slot");
  int t = e.h_red_start[li.gbase + box];
  for (int q = 0; q < slot; q++) t += slot_box(q) >= 0;
  const int p = e.h_red_pair[t];
  *factor_rank = e.h_pair_fact.empty() ? -1 : e.h_pair_fact[p];
  *box_first = e.h_pair_k[p] == (int)(li.gbase + box);
  return AFMM2D_OK;
}

int afmm2d_get_m2l(afmm2d_handle h, int level, int box, int slot, double* out, int64_t capacity,
                   int* rows, int* cols) {
  if (!h || !rows || !cols) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  if (!h->assembled) return fail(AFMM2D_ERR_STATE, "not assembled");
  if (level < 2 || level > e.L || box < 0 || box >= e.lv[level].nb || slot < 0 || slot >= 40)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "level/box/slot out of range");
  const LevelInfo& li = e.lv[level];
  auto slot_box = [&](int s) {
    return s < 16 ? e.inner[level][16 * (size_t)box + s] : e.outer[level][24 * (size_t)box + s - 16];
  };
  const int kb = slot_box(slot);
  if (kb < 0) return fail(AFMM2D_ERR_NOT_FOUND, "empty interaction slot");
  return guarded([&] {
    // position of this slot among the box's present slots
    int t = e.h_red_start[li.gbase + box];
    for (int q = 0; q < slot; q++) t += slot_box(q) >= 0;
    const int p = e.h_red_pair[t];
    const int r = li.rank[box], rc = li.rank[kb];
    *rows = r;
    *cols = rc;
    if ((int64_t)r * rc > capacity || (!out && r * rc > 0))
      throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "capacity too small");
    // the pair stores K(skeleton_lo, skeleton_hi) column-major with the
    // first box's rank rounded up to even as leading dimension
    const bool first = e.h_pair_k[p] == (int)(li.gbase + box);
    const int rows_lo = first ? r : rc, cols_lo = first ? rc : r;
    int ld = rows_lo + (rows_lo & 1);
    std::vector<double> blk((size_t)ld * cols_lo);
    const int fp = e.h_pair_fact.empty() ? -1 : e.h_pair_fact[p];
    if (fp >= 0) {
      // stored as U (rows_lo x fp) and V (cols_lo x fp): the block is U V^T
      ld = rows_lo;
      std::vector<double> uv((size_t)fp * (rows_lo + cols_lo));
      if (!uv.empty())
        AFMM_HIP(hipMemcpy(uv.data(), e.m2lf.as<double>() + e.h_pair_off[p],
                           sizeof(double) * uv.size(), hipMemcpyDeviceToHost));
      const double* U = uv.data();
      const double* V = U + (size_t)fp * rows_lo;
      std::fill(blk.begin(), blk.end(), 0.0);
      for (int c = 0; c < cols_lo; c++)
        for (int i = 0; i < rows_lo; i++) {
          double acc = 0.0;
          for (int l = 0; l < fp; l++) acc += U[i + (size_t)l * rows_lo] * V[c + (size_t)l * cols_lo];
          blk[(size_t)c * ld + i] = acc;
        }
    } else if (!blk.empty() && e.m2l_store == M2L_STORE_B48) {
      // blk.size() is even (even leading dimension): whole row pair records
      std::vector<B48Pair> rec(blk.size() / 2);
      PairDesc d;
      AFMM_HIP(hipMemcpy(rec.data(), e.m2l_b48.as<B48Pair>() + e.h_pair_off[p] / 2,
                         sizeof(B48Pair) * rec.size(), hipMemcpyDeviceToHost));
      AFMM_HIP(hipMemcpy(&d, e.pair_desc.as<PairDesc>() + p, sizeof(PairDesc),
                         hipMemcpyDeviceToHost));
      for (size_t i = 0; i < rec.size(); i++) {
        blk[2 * i] = b48_join(rec[i].lo0, (uint16_t)(rec[i].hi & 0xFFFF), d.sexp);
        blk[2 * i + 1] = b48_join(rec[i].lo1, (uint16_t)(rec[i].hi >> 16), d.sexp);
      }
    } else if (!blk.empty()) {
      AFMM_HIP(hipMemcpy(blk.data(), e.m2l.as<double>() + e.h_pair_off[p],
                         sizeof(double) * blk.size(), hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < r; i++)
      for (int c = 0; c < rc; c++)
        out[(size_t)i * rc + c] = first ? blk[(size_t)c * ld + i] : blk[(size_t)i * ld + c];
  });
}

int afmm2d_direct_rows(afmm2d_handle h, int nrows, const int64_t* rows, const double* b,
                       double* out) {
  if (!h || !rows || !b || !out || nrows < 0) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  return guarded([&] {
    Engine& e = *h->e;
    AFMM_HIP(hipSetDevice(e.device));
    // caller index -> KD-sorted index
    std::vector<int> inv(e.n);
    for (int i = 0; i < e.n; i++) inv[e.h_perm[i]] = i;
    std::vector<int> srows(nrows);
    for (int r = 0; r < nrows; r++) {
      if (rows[r] < 0 || rows[r] >= e.n) throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "row out of range");
      srows[r] = inv[rows[r]];
    }
    hipStream_t s = nullptr;  // the null stream plays the caller's stream
    begin_on(h, s);
    DevBuf d_b, d_qs, d_rows, d_out;
    d_b.alloc(sizeof(double) * e.n);
    d_qs.alloc(sizeof(double) * e.n);
    d_rows.alloc(sizeof(int) * std::max(nrows, 1));
    d_out.alloc(sizeof(double) * std::max(nrows, 1));
    AFMM_HIP(hipMemcpyAsync(d_b.p, b, sizeof(double) * e.n, hipMemcpyHostToDevice, s));
    if (nrows > 0)
      AFMM_HIP(hipMemcpyAsync(d_rows.p, srows.data(), sizeof(int) * nrows, hipMemcpyHostToDevice, s));
    gather_sorted(e, d_b.as<double>(), d_qs.as<double>(), s);
    direct_rows_device(e, nrows, d_rows.as<int>(), d_qs.as<double>(), d_out.as<double>(), s);
    if (nrows > 0)
      AFMM_HIP(hipMemcpyAsync(out, d_out.p, sizeof(double) * nrows, hipMemcpyDeviceToHost, s));
    AFMM_HIP(hipStreamSynchronize(s));
  });
}

int afmm2d_direct_rows_device_noalloc(afmm2d_handle h, int nrows, const int64_t* d_rows,
                                      const double* d_b, double* d_out, void* stream) {
  if (!h || nrows < 0 || (nrows > 0 && (!d_rows || !d_b || !d_out)))
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "bad argument");
  if ((int64_t)nrows > kNoAllocScratchDoubles)
    return fail(AFMM2D_ERR_UNSUPPORTED, "more rows than the static scratch holds (2^20)");
  return guarded([&] {
    Engine& e = *h->e;
    AFMM_HIP(hipSetDevice(e.device));
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)0;
    begin_on(h, s);
    direct_rows_noalloc(e, nrows, d_rows, d_b, d_out, s);
    end_on(h, s);
  });
}

int afmm2d_near_field(afmm2d_handle h, const double* b, double* y) {
  if (!h || !b || !y) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  return guarded([&] {
    Engine& e = *h->e;
    AFMM_HIP(hipSetDevice(e.device));
    const size_t nb = sizeof(double) * e.n;
    e.io_b.alloc(nb);
    e.io_y.alloc(nb);
    hipStream_t s = nullptr;
    begin_on(h, s);
    AFMM_HIP(hipMemcpyAsync(e.io_b.p, b, nb, hipMemcpyHostToDevice, s));
    near_field_device(e, e.io_b.as<double>(), e.io_y.as<double>(), s);
    if (h->e2) {
      Engine& e2 = *h->e2;
      e2.io_y.alloc(nb);
      near_field_device(e2, e.io_b.as<double>(), e2.io_y.as<double>(), s);
      add_device(e.io_y.as<double>(), e2.io_y.as<const double>(), e.n, s);
    }
    AFMM_HIP(hipMemcpyAsync(y, e.io_y.p, nb, hipMemcpyDeviceToHost, s));
    end_on(h, s);
    AFMM_HIP(hipStreamSynchronize(s));
    if (e.profiling) e.prof_collect();
  });
}

int afmm2d_direct(afmm2d_handle h, const double* b, double* y) {
  if (!h || !b || !y) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  std::vector<int64_t> rows(h->e->n);
  for (int i = 0; i < h->e->n; i++) rows[i] = i;
  return afmm2d_direct_rows(h, h->e->n, rows.data(), b, y);
}

int afmm2d_compute_error(afmm2d_handle h, int box, double* err) {
  if (!h || !err) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  if (!e.have_matvec) return fail(AFMM2D_ERR_STATE, "no host matvec has run yet");
  if (box < 0 || box >= (int)e.leaf_start.size()) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "box out of range");
  return guarded([&] {
    AFMM_HIP(hipSetDevice(e.device));
    const int st = e.leaf_start[box], n = e.leaf_count[box];
    hipStream_t s = nullptr;  // the null stream plays the caller's stream
    DevBuf d_qs, d_rows, d_out;
    d_qs.alloc(sizeof(double) * e.n);
    d_rows.alloc(sizeof(int) * std::max(n, 1));
    d_out.alloc(sizeof(double) * std::max(n, 1));
    std::vector<int> rows(n);
    for (int i = 0; i < n; i++) rows[i] = st + i;
    AFMM_HIP(hipMemcpyAsync(d_qs.p, e.last_qs.data(), sizeof(double) * e.n, hipMemcpyHostToDevice, s));
    if (n > 0) AFMM_HIP(hipMemcpyAsync(d_rows.p, rows.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    direct_rows_device(e, n, d_rows.as<int>(), d_qs.as<double>(), d_out.as<double>(), s);
    std::vector<double> tr(n);
    if (n > 0) AFMM_HIP(hipMemcpy(tr.data(), d_out.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; i++) {
      double d = tr[i] - e.last_ps[st + i];
      num += d * d;
      den += tr[i] * tr[i];
    }
    *err = std::sqrt(num) / std::sqrt(den);
  });
}

int afmm2d_evaluate_error(afmm2d_handle h, unsigned seed, int* box, double* err) {
  if (!h || !box || !err) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  std::mt19937 rng(seed);
  *box = (int)(rng() % (unsigned)h->e->leaf_start.size());
  return afmm2d_compute_error(h, *box, err);
}

int afmm2d_get_matrix_entry(afmm2d_handle h, int i, int j, double* value) {
  if (!h || !value) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  if (i < 0 || j < 0 || i >= e.n || j >= e.n) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "index out of range");
  *value = kernel_host(e.kid_full, e.kp, e.h_xy[2 * (size_t)i], e.h_xy[2 * (size_t)i + 1],
                       e.h_xy[2 * (size_t)j], e.h_xy[2 * (size_t)j + 1]);
  return AFMM2D_OK;
}

int afmm2d_kernel_value(int kernel_id, const double* kparams, double a, double xi, double yi,
                        double xj, double yj, double* value) {
  if (!value) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  if (kernel_id < 0 || kernel_id >= kNumKernels)
    return fail(AFMM2D_ERR_INVALID_ARGUMENT, "unknown kernel id");
  KParams p{(kparams && kparams[0] > 0.0) ? kparams[0] : 0.01, a, {0.0, 0.0, 0.0, 0.0}};
  for (int i = 0; i < 4; i++) p.u[i] = kparams ? kparams[i] : 0.0;
  *value = kernel_host(kernel_id, p, xi, yi, xj, yj);
  return AFMM2D_OK;
}

int afmm2d_get_stats(afmm2d_handle h, afmm2d_stats* st) {
  if (!h || !st) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  Engine& e = *h->e;
  std::memset(st, 0, sizeof(*st));
  st->n = e.n;
  st->levels = e.L;
  st->tol_pow = e.tol_pow;
  st->kernel_id = e.kid_full;
  st->max_rank = e.max_rank;
  st->skeleton_total = e.skel_total;
  st->il_pairs = e.il_pairs;
  st->near_pairs = e.near_pairs;
  // dense blocks at 8 or 6 bytes per entry, ACA factors at 8
  const int64_t bpe = e.m2l_store == M2L_STORE_B48 ? 6 : 8;
  const int64_t m2l_stored_bytes = bpe * (e.m2l_total - e.m2lf_total) + 8 * e.m2lf_total;
  st->m2l_bytes = m2l_stored_bytes;
  st->m2l_columns = e.m2l_cols;
  st->l2p_bytes = 8 * e.l2p_total;  // stored rows (the skeleton rows are the identity)
  // One matvec reads q (gather) and the coordinates (near field), L2P twice
  // (M2M and L2L), every stored M2L coefficient and its column index once,
  // the multipole/local vectors, and writes y.
  // One matvec reads b and perm (gather), writes and rereads q in KD order,
  // reads the coordinates (near field), writes and rereads the near-field
  // potentials, reads perm again and writes y; it reads the stored L2P rows
  // twice (M2M, L2L) with their row maps, every stored M2L block once,
  // writes and rereads the M2L partials, and touches the multipole/local
  // vectors a few times.
  st->matvec_bytes = e.assembled
                         ? (int64_t)e.n * (8 + 4 + 8 + 8 + 16 + 8 + 8 + 4 + 8 + 2 * 4) +
                               2 * 8 * e.l2p_total + m2l_stored_bytes + 16 * e.m2l_cols +
                               4 * e.il_pairs + (6 * 8 + 2 * 4) * e.skel_total
                         : 0;
  st->create_seconds = e.create_seconds;
  st->assemble_seconds = e.assemble_seconds;
  st->m2l_leaf_bytes = e.assembled ? bpe * (e.lv[e.L].m2l_coeffs - e.lv[e.L].m2l_fact_coeffs) +
                                         8 * e.lv[e.L].m2l_fact_coeffs
                                   : 0;
  st->m2l_storage = e.m2l_store == M2L_STORE_B48 ? AFMM2D_M2L_BFP48 : AFMM2D_M2L_FP64;
  st->m2l_leaf_columns = e.assembled ? e.lv[e.L].m2l_cols : 0;
  st->skeleton_leaf = e.assembled ? e.lv[e.L].soff[e.lv[e.L].nb] - e.lv[e.L].soff[0] : 0;
  st->near_waves = e.p2p_waves_per_simd;
  st->blocked_m2l_coeffs = e.mr_pm_total;
  st->blocked_m2l_coeffs_bf16 = e.mr_m_coeffs;
  st->blocked_basis_coeffs = e.mr_v_coeffs;
  st->rank_boxes = e.rank_boxes;
  st->rank_fallbacks = e.rank_fallbacks;
  st->aca_scratch_peak = e.aca_scratch_peak;
  st->aca_launches = e.aca_launches;
  st->aca_seconds = e.aca_seconds;
  st->m2l_factored_pairs = e.m2lf_pairs;
  st->m2l_factored_bytes = 8 * e.m2lf_total;
  st->m2l_factored_dense_bytes = 8 * e.m2lf_dense;
  return AFMM2D_OK;
}

int afmm2d_set_profiling(afmm2d_handle h, int level) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle");
  if (level < 0 || level > 2) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "profiling level 0..2");
  h->e->profiling = level;
  return AFMM2D_OK;
}

int afmm2d_get_kernel_time(afmm2d_handle h, const char* name, double* total_ms, int64_t* launches) {
  if (!h || !name || !total_ms || !launches) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL argument");
  int st = guarded([&] {
    AFMM_HIP(hipSetDevice(h->e->device));
    h->e->prof_collect();
  });
  if (st != AFMM2D_OK) return st;
  auto it = h->e->timers.find(name);
  if (it == h->e->timers.end()) {
    *total_ms = 0.0;
    *launches = 0;
    return AFMM2D_OK;
  }
  *total_ms = it->second.total_ms;
  *launches = it->second.launches;
  return AFMM2D_OK;
}

int afmm2d_reset_kernel_times(afmm2d_handle h) {
  if (!h) return fail(AFMM2D_ERR_INVALID_ARGUMENT, "NULL handle");
  return guarded([&] {
    AFMM_HIP(hipSetDevice(h->e->device));
    h->e->prof_collect();
    h->e->timers.clear();
  });
}

}  // extern "C"
