// cn_prof.h -- per-kernel timing of a solve or an assembly with HIP events (enabled by
// cn_profile_enable): a ProfScope records an event pair on the stream around
// the launches it encloses; cn_profile_query sums them by name.  The events
// are timing-only (read after the solve's final stream synchronization), so
// they are created without the system-scope fence: a fenced event writes
// back and invalidates the caches at every record, which slows the kernels
// after it and so the solve being measured.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

namespace cn {

struct Prof {
  bool on = false;
  std::string only;  // non-empty: only scopes of this name record events
  std::vector<hipEvent_t> pool;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> recs;
  size_t used = 0;
  std::map<std::string, std::pair<int, double>> result;
};
extern thread_local Prof g_prof;
// clear the records (at the start of a timed entry point) / sum the records
// by name into the results cn_profile_query reads (after its last launch)
void prof_reset();
void prof_collect();

struct ProfScope {
  int idx = -1;
  hipStream_t s;
  ProfScope(const char *name, hipStream_t st) : s(st) {
    Prof &P = g_prof;
    if (!P.on || (!P.only.empty() && P.only != name)) return;
    if (P.used + 2 > P.pool.size()) {
      if (P.pool.size() >= 65536) return;
      size_t add = std::max<size_t>(256, P.pool.size());
      for (size_t i = 0; i < add; ++i) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
        P.pool.push_back(e);
      }
    }
    hipEvent_t a = P.pool[P.used++], b = P.pool[P.used++];
    P.recs.push_back({name, a, b});
    idx = (int)P.recs.size() - 1;
    (void)hipEventRecord(a, s);
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(g_prof.recs[idx].b, s);
  }
};


}  // namespace cn
