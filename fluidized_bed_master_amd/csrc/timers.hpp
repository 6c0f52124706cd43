// Per-kernel device timers: HIP events recorded on the context's stream
// around selected launches, reduced to (count, total ms) on request.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace sophia {

struct Timers {
  bool on = false;
  hipStream_t stream = nullptr;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<long, double>> acc;  // name -> (launches, ms)
  hipEvent_t open_a = nullptr;
  std::string open_name;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  void begin(const char* name) {
    if (!on) return;
    open_a = get();
    open_name = name;
    (void)hipEventRecord(open_a, stream);
  }
  void end() {
    if (!on || !open_a) return;
    hipEvent_t b = get();
    (void)hipEventRecord(b, stream);
    pending.push_back({open_name, open_a, b});
    open_a = nullptr;
    if (pending.size() > 4096) collect();
  }
  // count a launch without timing it (0 ms)
  void count(const std::string& name) {
    if (on) acc[name].first += 1;
  }
  void collect() {
    if (pending.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (auto& p : pending) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, p.a, p.b);
      auto& s = acc[p.name];
      s.first += 1;
      s.second += ms;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void reset() {
    collect();
    acc.clear();
  }
  ~Timers() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// RAII scope for one timed launch group
struct TimeScope {
  Timers* t;
  TimeScope(Timers* tt, const char* name) : t(tt) {
    if (t) t->begin(name);
  }
  ~TimeScope() {
    if (t) t->end();
  }
};

}  // namespace sophia
