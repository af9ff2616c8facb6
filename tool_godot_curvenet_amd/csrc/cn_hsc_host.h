// cn_hsc_host.h -- host construction of one level of the hierarchical Schur
// preconditioner (FP64, CSR in and out), used by cn_hsc_build.
//
// A level splits the rows of A into an independent set F (greedy, rows by
// (off-diagonal count, index) ascending) and the rest C, forms the Schur
// complement S = A_CC - A_CF D_F^-1 A_FC, drops small fill with diagonal
// compensation, and returns the restriction R = -A_CF D_F^-1 and the
// prolongation P = -D_F^-1 A_FC.  The per-row work runs on host threads;
// every entry is still summed in the one order the single-threaded form
// uses (S entries over ascending f, diagonal compensation over ascending
// columns), so the level is bit-identical whatever the thread count.
#pragma once
#include <cstdint>
#include <memory>
#include <new>
#include <utility>
#include <vector>

#include "curvenet_la.h"

namespace cn {

// an allocator whose resize() leaves new elements uninitialised, so large
// arrays are first written (and their pages first touched) by the parallel
// passes that fill them, not by one zeroing thread
template <typename T>
struct DefaultInit : std::allocator<T> {
  template <typename U>
  struct rebind { using other = DefaultInit<U>; };
  using std::allocator<T>::allocator;
  template <typename U>
  void construct(U *p) noexcept { ::new ((void *)p) U; }
  template <typename U, typename... Args>
  void construct(U *p, Args &&...args) { ::new ((void *)p) U(std::forward<Args>(args)...); }
};
template <typename T>
using HostVec = std::vector<T, DefaultInit<T>>;

struct HostCsr {
  HostVec<int32_t> ptr, col;
  HostVec<double> val;
  int32_t n() const { return (int32_t)ptr.size() - 1; }
  double at(int32_t i, int32_t j) const;  // entry (i, j) (rows sorted by column), else 0
};

struct HostLevel {
  int32_t n = 0;
  HostVec<int32_t> F, C;
  HostVec<double> dinv;
  HostVec<int32_t> r_ptr, r_col, p_ptr, p_col;  // r_col: level-l rows of F vertices
  HostVec<double> r_val, p_val;
};

// One level of A; false when no level should be built (then A is the
// coarse matrix: |F| < min_frac n or F empty).  `threads` <= 0 takes
// host_threads().
bool hsc_build_level(const HostCsr &A, const cn_hsc_opts &o, HostLevel &L, HostCsr &next,
                     int threads = 0);

// Threads for host-side set-up work: $CN_HOST_THREADS, else
// $OMP_NUM_THREADS, else the hardware's count; at most 64.
int host_threads();

}  // namespace cn
