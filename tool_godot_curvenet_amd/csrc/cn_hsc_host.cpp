// cn_hsc_host.cpp -- one level of the hierarchical Schur preconditioner on
// host threads (built with -ffp-contract=off: the level must equal the
// oracle's restatement bit for bit, so no multiply-add may be fused).
//
// After Krishnan, Fattal, Szeliski 2013 (HSC).  The per-vertex step is the
// spec's HSC.eliminateVertex (HierarchicalSparsifyCompensate.lean:96-125):
// eliminating v adds w_vi w_vj / deg(v) to every pair of v's neighbours.
// On an SPD matrix the pivot is the diagonal a_vv and the update is
// s_ij -= a_iv a_vj / a_vv.  A level eliminates a whole independent set F
// at once, which is exact because F's rows do not couple:
// S = A_CC - A_CF D_F^-1 A_FC.
//
// Threading: the greedy independent set is one ordered pass (it is what
// fixes F); everything after it is row-parallel.  A row of S, its drop
// decisions, its compensated diagonal, and its rows of R and P read only A
// and the finished rows of S, and each sum keeps the order of the
// sequential form, so the result does not depend on the thread count.
#include "cn_hsc_host.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <numeric>
#include <thread>

namespace cn {

double HostCsr::at(int32_t i, int32_t j) const {
  auto b = col.begin() + ptr[i], e = col.begin() + ptr[i + 1];
  auto it = std::lower_bound(b, e, j);
  return (it != e && *it == j) ? val[it - col.begin()] : 0.0;
}

int host_threads() {
  static const int n = [] {
    int v = 0;
    if (const char *e = std::getenv("CN_HOST_THREADS")) v = std::atoi(e);
    if (v <= 0)
      if (const char *e = std::getenv("OMP_NUM_THREADS")) v = std::atoi(e);
    if (v <= 0) v = (int)std::thread::hardware_concurrency();
    return std::clamp(v, 1, 64);
  }();
  return n;
}

namespace {

// chunks of at least `grain` rows, at most `threads` of them
int64_t n_chunks(int64_t n, int threads, int64_t grain) {
  return std::max<int64_t>(1, std::min<int64_t>(threads, n / std::max<int64_t>(grain, 1)));
}

// f(begin, end, chunk) over [0, n) in n_chunks(n, threads, grain) contiguous
// chunks, chunk 0 on the calling thread.  An exception in any chunk is
// rethrown here after every thread has joined.
template <typename Fn>
void parallel_rows(int64_t n, int threads, int64_t grain, Fn &&f) {
  const int64_t chunks = n_chunks(n, threads, grain);
  std::vector<std::exception_ptr> err((size_t)chunks);
  auto run = [&](int64_t c) {
    try {
      f(n * c / chunks, n * (c + 1) / chunks, (int)c);
    } catch (...) {
      err[(size_t)c] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  th.reserve((size_t)chunks);
  for (int64_t c = 1; c < chunks; ++c) {
    try {
      th.emplace_back(run, c);
    } catch (...) {
      run(c);  // no thread to be had: do the chunk here
    }
  }
  run(0);
  for (auto &t : th) t.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

// exclusive prefix of per-row counts into ptr (size rows + 1)
void counts_to_ptr(const HostVec<int32_t> &cnt, HostVec<int32_t> &ptr) {
  ptr.resize(cnt.size() + 1);
  int64_t run = 0;
  for (size_t i = 0; i < cnt.size(); ++i) {
    ptr[i] = (int32_t)run;
    run += cnt[i];
  }
  ptr[cnt.size()] = (int32_t)run;
}

}  // namespace

bool hsc_build_level(const HostCsr &A, const cn_hsc_opts &o, HostLevel &L, HostCsr &next,
                     int threads) {
  const int T = threads > 0 ? threads : host_threads();
  const int32_t n = A.n();
  HostVec<int32_t> deg(n);
  parallel_rows(n, T, 65536, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) {
      int32_t d = 0;
      for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) d += A.col[q] != i;
      deg[i] = d;
    }
  });
  // visit order: (off-diagonal count, index) ascending, as a counting sort
  // by count that keeps index order within a count
  HostVec<int32_t> order(n);
  {
    const int32_t dmax = n ? *std::max_element(deg.begin(), deg.end()) : 0;
    std::vector<int64_t> start((size_t)dmax + 2, 0);
    for (int32_t i = 0; i < n; ++i) ++start[deg[i] + 1];
    for (int32_t d = 0; d <= dmax; ++d) start[d + 1] += start[d];
    for (int32_t i = 0; i < n; ++i) order[start[deg[i]]++] = i;
  }
  std::vector<char> isF(n, 0), blocked(n, 0);
  for (int32_t v : order) {
    if (blocked[v]) continue;
    isF[v] = 1;
    blocked[v] = 1;
    for (int32_t q = A.ptr[v]; q < A.ptr[v + 1]; ++q) blocked[A.col[q]] = 1;
  }
  HostVec<int32_t> F, C, cidx(n);
  for (int32_t i = 0; i < n; ++i) {
    cidx[i] = -1;
    if (isF[i]) F.push_back(i);
    else { cidx[i] = (int32_t)C.size(); C.push_back(i); }
  }
  if ((double)F.size() < o.min_frac * n || F.empty()) return false;
  const int32_t nC = (int32_t)C.size(), nF = (int32_t)F.size();

  // S, row by row.  Row c starts as A_CC's row; then for every F neighbour f
  // of c (ascending), for every neighbour c2 of f (ascending):
  // S[c][c2] -= (a_cf * a_fc2) / a_ff.  Each entry sees its terms in
  // ascending f, the order of the f-major loop of the restatement.  Chunks
  // keep their rows (sorted by column) in their own arrays.
  struct Part {
    HostVec<int32_t> col;
    HostVec<double> val;
    HostVec<char> orig;
  };
  const int64_t sgrain = 2048;
  std::vector<Part> part((size_t)n_chunks(nC, T, sgrain));
  HostVec<int32_t> scnt(nC);
  parallel_rows(nC, T, sgrain, [&](int64_t b, int64_t e, int c) {
    Part &P = part[(size_t)c];
    std::vector<int32_t> slot(nC, -1);  // position of a column in the row's list
    std::vector<int32_t> lu, idx;
    std::vector<double> la;
    std::vector<char> lo;
    for (int64_t t = b; t < e; ++t) {
      const int32_t r = C[t];
      lu.clear();
      la.clear();
      lo.clear();
      for (int32_t q = A.ptr[r]; q < A.ptr[r + 1]; ++q) {
        const int32_t j = A.col[q];
        if (isF[j]) continue;
        slot[cidx[j]] = (int32_t)lu.size();
        lu.push_back(cidx[j]);
        la.push_back(A.val[q]);
        lo.push_back(1);
      }
      for (int32_t q = A.ptr[r]; q < A.ptr[r + 1]; ++q) {
        const int32_t f = A.col[q];
        if (!isF[f]) continue;
        const double acf = A.val[q];
        const double d = A.at(f, f);
        for (int32_t w = A.ptr[f]; w < A.ptr[f + 1]; ++w) {
          const int32_t c2 = A.col[w];
          if (c2 == f) continue;
          const int32_t u = cidx[c2];
          int32_t s = slot[u];
          if (s < 0) {
            s = (int32_t)lu.size();
            slot[u] = s;
            lu.push_back(u);
            la.push_back(0.0);
            lo.push_back(0);
          }
          la[s] = la[s] - acf * A.val[w] / d;
        }
      }
      idx.resize(lu.size());
      std::iota(idx.begin(), idx.end(), 0);
      std::sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return lu[x] < lu[y]; });
      for (int32_t m : idx) {
        P.col.push_back(lu[m]);
        P.val.push_back(la[m]);
        P.orig.push_back(lo[m]);
        slot[lu[m]] = -1;
      }
      scnt[t] = (int32_t)lu.size();
    }
  });
  HostVec<int32_t> sptr;
  counts_to_ptr(scnt, sptr);
  HostVec<int32_t> scol(sptr[nC]);
  HostVec<double> sval(sptr[nC]);
  HostVec<char> sorig(sptr[nC]);
  parallel_rows(nC, T, sgrain, [&](int64_t b, int64_t, int c) {
    const Part &P = part[(size_t)c];
    std::copy(P.col.begin(), P.col.end(), scol.begin() + sptr[b]);
    std::copy(P.val.begin(), P.val.end(), sval.begin() + sptr[b]);
    std::copy(P.orig.begin(), P.orig.end(), sorig.begin() + sptr[b]);
  });
  part.clear();
  part.shrink_to_fit();

  // diagonal of S before any compensation
  HostVec<double> sdiag(nC);
  parallel_rows(nC, T, 65536, [&](int64_t b, int64_t e, int) {
    for (int64_t t = b; t < e; ++t) {
      double d = 0.0;
      for (int32_t m = sptr[t]; m < sptr[t + 1]; ++m)
        if (scol[m] == t) d = sval[m];
      sdiag[t] = d;
    }
  });
  // drops: a fill entry (i, j), i < j, is dropped when it is small against
  // sqrt(s_ii s_jj); its weight (s_ij, or |s_ij| with CN_HSC_DIAG, from row
  // i) goes to both diagonals.  Row t sees its dropped pairs in ascending
  // column order -- the pairs (i, t) of rows i < t, then its own (t, j) --
  // which is the order in which a pass over the rows in ascending order
  // would add them.
  HostVec<char> dropped(sptr[nC]);
  HostVec<double> ndiag(nC);
  const bool absw = o.compensate == CN_HSC_DIAG;
  parallel_rows(nC, T, 8192, [&](int64_t b, int64_t e, int) {
    for (int64_t t = b; t < e; ++t) {
      double nd = sdiag[t];
      for (int32_t m = sptr[t]; m < sptr[t + 1]; ++m) {
        dropped[m] = 0;
        const int32_t j = scol[m];
        if (j == t) continue;
        double s;
        if (j > t) {
          if (sorig[m]) continue;
          s = sval[m];
          if (!(std::fabs(s) < o.theta * std::sqrt(sdiag[t] * sdiag[j]))) continue;
        } else {
          const auto rb = scol.begin() + sptr[j], re = scol.begin() + sptr[j + 1];
          const auto it = std::lower_bound(rb, re, (int32_t)t);
          if (it == re || *it != t) continue;
          const int64_t mm = it - scol.begin();
          if (sorig[mm]) continue;
          s = sval[mm];
          if (!(std::fabs(s) < o.theta * std::sqrt(sdiag[j] * sdiag[t]))) continue;
        }
        const double w = absw ? std::fabs(s) : s;
        nd = nd + w;
        dropped[m] = 1;
      }
      ndiag[t] = nd;
    }
  });
  {
    HostVec<int32_t> ncnt(nC);
    parallel_rows(nC, T, 65536, [&](int64_t b, int64_t e, int) {
      for (int64_t t = b; t < e; ++t) {
        int32_t k = 0;
        for (int32_t m = sptr[t]; m < sptr[t + 1]; ++m) k += !dropped[m];
        ncnt[t] = k;
      }
    });
    counts_to_ptr(ncnt, next.ptr);
    next.col.resize(next.ptr[nC]);
    next.val.resize(next.ptr[nC]);
    parallel_rows(nC, T, 65536, [&](int64_t b, int64_t e, int) {
      for (int64_t t = b; t < e; ++t) {
        int32_t o2 = next.ptr[t];
        for (int32_t m = sptr[t]; m < sptr[t + 1]; ++m) {
          if (dropped[m]) continue;
          next.col[o2] = scol[m];
          next.val[o2] = scol[m] == t ? ndiag[t] : sval[m];
          ++o2;
        }
      }
    });
  }

  L.n = n;
  L.F = std::move(F);
  L.C = std::move(C);
  L.dinv.resize(nF);
  HostVec<int32_t> rcnt(nC), pcnt(nF);
  parallel_rows(nC, T, 65536, [&](int64_t b, int64_t e, int) {
    for (int64_t t = b; t < e; ++t) {
      const int32_t r = L.C[t];
      int32_t k = 0;
      for (int32_t q = A.ptr[r]; q < A.ptr[r + 1]; ++q) k += isF[A.col[q]] != 0;
      rcnt[t] = k;
    }
  });
  parallel_rows(nF, T, 65536, [&](int64_t b, int64_t e, int) {
    for (int64_t x = b; x < e; ++x) {
      const int32_t f = L.F[x];
      L.dinv[x] = 1.0 / A.at(f, f);
      int32_t k = 0;
      for (int32_t q = A.ptr[f]; q < A.ptr[f + 1]; ++q) k += A.col[q] != f;
      pcnt[x] = k;
    }
  });
  counts_to_ptr(rcnt, L.r_ptr);
  counts_to_ptr(pcnt, L.p_ptr);
  L.r_col.resize(L.r_ptr[nC]);
  L.r_val.resize(L.r_ptr[nC]);
  L.p_col.resize(L.p_ptr[nF]);
  L.p_val.resize(L.p_ptr[nF]);
  parallel_rows(nC, T, 65536, [&](int64_t b, int64_t e, int) {
    for (int64_t t = b; t < e; ++t) {
      const int32_t r = L.C[t];
      int32_t o2 = L.r_ptr[t];
      for (int32_t q = A.ptr[r]; q < A.ptr[r + 1]; ++q) {
        const int32_t j = A.col[q];
        if (!isF[j]) continue;
        L.r_col[o2] = j;
        L.r_val[o2] = -A.val[q] / A.at(j, j);
        ++o2;
      }
    }
  });
  parallel_rows(nF, T, 65536, [&](int64_t b, int64_t e, int) {
    for (int64_t x = b; x < e; ++x) {
      const int32_t f = L.F[x];
      const double d = A.at(f, f);
      int32_t o2 = L.p_ptr[x];
      for (int32_t q = A.ptr[f]; q < A.ptr[f + 1]; ++q) {
        const int32_t j = A.col[q];
        if (j == f) continue;
        L.p_col[o2] = cidx[j];
        L.p_val[o2] = -A.val[q] / d;
        ++o2;
      }
    }
  });
  return true;
}

}  // namespace cn
