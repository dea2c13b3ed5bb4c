// engine.hpp -- the per-problem device state of the AFMM2D engine and the
// launchers each kernel file provides.
//
// HBM layout (levels 2..L, box gid = level_base[j] + k, k the quad-tree id):
//   x, y          KD-sorted point coordinates (N doubles each)
//   perm          KD-sorted index -> caller index (N int32)
//   skx, sky      skeleton points of every box, concatenated in (level, box)
//                 order; box gid owns [soff[gid], soff[gid] + rank[gid])
//   l2p_rm        L2P of every box, rows x rank row-major   (read by M2M)
//   l2p_cm        the same matrices column-major             (read by L2L)
//   m2l           per target box, the blocks of its interaction list
//                 concatenated column-wise into one rank x C column-major
//                 matrix (C = sum of the list's ranks); m2l_src[c] is the
//                 skeleton offset of column c's source coefficient
//   outv, inv     multipole / local coefficients, indexed like skx
#pragma once

#include <array>
#include <map>
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
  std::vector<int64_t> m2loff;  // element offset into m2l (nb + 1)
  std::vector<int64_t> m2lcoff; // column offset into m2l_src (nb + 1)
};

struct KernelTimer {
  double total_ms = 0.0;
  int64_t launches = 0;
};

struct Engine {
  // problem
  int n = 0, leaf = 0, tol_pow = 0, L = 0, kid = 0, device = 0;
  double tol = 0.0;
  KParams kp{0.01, 0.0};

  hipStream_t s_main = nullptr, s_aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

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

  // assembly
  bool assembled = false;
  std::vector<LevelInfo> lv;  // indexed by level (0..L), levels < 2 unused
  int64_t skel_total = 0, l2p_total = 0, m2l_total = 0, m2l_cols = 0;
  int max_rank = 0;
  DevBuf skx, sky, skidx;
  DevBuf rank_g, nin_g, soff_g, l2poff_g, m2loff_g, m2lcoff_g;
  DevBuf l2p_rm, l2p_cm, m2l, m2l_src;
  double create_seconds = 0.0, assemble_seconds = 0.0;

  // matvec state
  DevBuf outv, inv, qs, pot_near, io_b, io_y;
  bool have_matvec = false;
  std::vector<double> last_qs, last_ps;  // host copies for compute_error

  // profiling
  bool profiling = false;
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

// matvec.hip
void matvec_device(Engine& e, const double* d_b, double* d_y, hipStream_t s);
void direct_rows_device(Engine& e, int nrows, const int* d_rows, const double* d_qs, double* d_out,
                        hipStream_t s);
void gather_sorted(Engine& e, const double* d_b, double* d_qs, hipStream_t s);

}  // namespace afmm
