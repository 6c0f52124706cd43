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
  // open scopes, innermost last: a scope nested in another closes first and
  // both are timed
  std::vector<Pending> open;

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
    Pending p{name, get(), nullptr};
    (void)hipEventRecord(p.a, stream);
    open.push_back(p);
  }
  void end() {
    if (!on || open.empty()) return;
    Pending p = open.back();
    open.pop_back();
    p.b = get();
    (void)hipEventRecord(p.b, stream);
    pending.push_back(p);
    if (pending.size() > 4096 && open.empty()) collect();
  }
  // time [a, b) recorded by the caller on another stream under this name
  void add_span(const char* name, hipEvent_t a, hipEvent_t b) {
    if (on) pending.push_back({name, a, b});
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
  // stop timing: events of scopes still open go back to the pool
  void off() {
    for (auto& p : open) pool.push_back(p.a);
    open.clear();
    on = false;
  }
  ~Timers() {
    for (auto& p : open) (void)hipEventDestroy(p.a);
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
