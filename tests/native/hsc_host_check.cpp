// hsc_host_check.cpp -- test wrapper around the library's host level
// builder (tool_godot_curvenet_amd/csrc/cn_hsc_host.cpp): one level of a host
// CSR matrix, for the CPU tests that compare it with the oracle's HSC
// restatement and across thread counts.  Built by the test that loads it.
#include <cstring>
#include <exception>

#include "cn_hsc_host.h"

namespace {
struct Result {
  cn::HostLevel L;
  cn::HostCsr next;
};
}  // namespace

extern "C" {

// One level of (n, ptr, col, val); null when no level is built (or on an
// error: *err is then set to 1).
void *hsc_level_host(int32_t n, const int32_t *ptr, const int32_t *col, const double *val,
                     double theta, int32_t compensate, double min_frac, int32_t threads,
                     int32_t *err) {
  *err = 0;
  try {
    cn::HostCsr A;
    A.ptr.assign(ptr, ptr + n + 1);
    A.col.assign(col, col + ptr[n]);
    A.val.assign(val, val + ptr[n]);
    cn_hsc_opts o;
    o.theta = theta;
    o.coarse_max = 0;
    o.max_levels = 1;
    o.min_frac = min_frac;
    o.compensate = compensate;
    Result *r = new Result();
    if (!cn::hsc_build_level(A, o, r->L, r->next, threads)) {
      delete r;
      return nullptr;
    }
    return r;
  } catch (const std::exception &) {
    *err = 1;
    return nullptr;
  }
}

// sizes and contents: 0 F, 1 C, 2 dinv, 3 r_ptr, 4 r_col, 5 r_val, 6 p_ptr,
// 7 p_col, 8 p_val, 9 next ptr, 10 next col, 11 next val
int64_t hsc_level_size(void *h, int32_t which) {
  const Result &r = *static_cast<Result *>(h);
  switch (which) {
    case 0: return (int64_t)r.L.F.size();
    case 1: return (int64_t)r.L.C.size();
    case 2: return (int64_t)r.L.dinv.size();
    case 3: return (int64_t)r.L.r_ptr.size();
    case 4: return (int64_t)r.L.r_col.size();
    case 5: return (int64_t)r.L.r_val.size();
    case 6: return (int64_t)r.L.p_ptr.size();
    case 7: return (int64_t)r.L.p_col.size();
    case 8: return (int64_t)r.L.p_val.size();
    case 9: return (int64_t)r.next.ptr.size();
    case 10: return (int64_t)r.next.col.size();
    case 11: return (int64_t)r.next.val.size();
  }
  return -1;
}

void hsc_level_copy(void *h, int32_t which, void *dst) {
  const Result &r = *static_cast<Result *>(h);
  auto cp = [&](const auto &v) { std::memcpy(dst, v.data(), v.size() * sizeof(v[0])); };
  switch (which) {
    case 0: cp(r.L.F); break;
    case 1: cp(r.L.C); break;
    case 2: cp(r.L.dinv); break;
    case 3: cp(r.L.r_ptr); break;
    case 4: cp(r.L.r_col); break;
    case 5: cp(r.L.r_val); break;
    case 6: cp(r.L.p_ptr); break;
    case 7: cp(r.L.p_col); break;
    case 8: cp(r.L.p_val); break;
    case 9: cp(r.next.ptr); break;
    case 10: cp(r.next.col); break;
    case 11: cp(r.next.val); break;
  }
}

void hsc_level_free(void *h) { delete static_cast<Result *>(h); }
}
