// Launch-boundary timeline (see trace.hpp) and its C entry points.
#include "trace.hpp"

#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/half_amd.h"
#include "common.hpp"

namespace half {
namespace {

struct Mark {
  const char *name;  // nullptr closes a sequence
  hipEvent_t ev;
};

bool g_enabled = false;
std::vector<Mark> g_marks;
std::vector<hipEvent_t> g_pool;
size_t g_used = 0;

hipEvent_t next_event() {
  if (g_used == g_pool.size()) {
    hipEvent_t e;
    HALF_HIP(hipEventCreate(&e));
    g_pool.push_back(e);
  }
  return g_pool[g_used++];
}

}  // namespace

void trace_mark(const char *name, hipStream_t s) {
  if (!g_enabled) return;
  hipEvent_t e = next_event();
  HALF_HIP(hipEventRecord(e, s));
  g_marks.push_back({name, e});
}

void trace_close(hipStream_t s) { trace_mark(nullptr, s); }

bool trace_enabled() { return g_enabled; }

}  // namespace half

using namespace half;

extern "C" {

int half_amd_trace(int enable, char *error, int capacity) {
  if (error && capacity > 0) error[0] = '\0';
  g_marks.clear();
  g_used = 0;
  g_enabled = enable != 0;
  return 0;
}

int half_amd_trace_read(int max_kernels, char *names, double *total_ms, int64_t *launches, int32_t *nkinds,
                        char *error, int capacity) {
  if (error && capacity > 0) error[0] = '\0';
  if (!nkinds) return 1;
  try {
    HALF_HIP(hipDeviceSynchronize());
    std::map<std::string, std::pair<double, int64_t>> agg;
    std::vector<std::string> order;
    for (size_t i = 0; i + 1 < g_marks.size(); ++i) {
      if (!g_marks[i].name) continue;
      float ms = 0;
      HALF_HIP(hipEventElapsedTime(&ms, g_marks[i].ev, g_marks[i + 1].ev));
      auto it = agg.find(g_marks[i].name);
      if (it == agg.end()) {
        order.push_back(g_marks[i].name);
        agg[g_marks[i].name] = {ms, 1};
      } else {
        it->second.first += ms;
        it->second.second += 1;
      }
    }
    *nkinds = static_cast<int32_t>(order.size());
    for (int i = 0; i < static_cast<int>(order.size()) && i < max_kernels; ++i) {
      if (names) {
        std::strncpy(names + 64 * i, order[i].c_str(), 63);
        names[64 * i + 63] = '\0';
      }
      if (total_ms) total_ms[i] = agg[order[i]].first;
      if (launches) launches[i] = agg[order[i]].second;
    }
    return 0;
  } catch (const Error &e) {
    if (error && capacity > 0) {
      std::strncpy(error, e.what(), capacity - 1);
      error[capacity - 1] = '\0';
    }
    return e.code;
  }
}

}  // extern "C"
