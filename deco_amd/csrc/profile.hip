// Phase timer tree with the reference's labels and report format (PROFILE, core/profiler.hpp:1-178).
//
// start(name) enters the child `name` of the current node (re-entering a name accumulates, as
// core/profiler.hpp:66-81 does), stop() adds the elapsed time and returns to the parent. It is
// off by default; when on, every boundary first waits for the library stream, so device work
// is attributed to the phase that queued it (which serialises the host and the device: a
// profiled solve is slower than an unprofiled one). Host wall clock, steady_clock.
#include "common.hpp"
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace deco {

namespace {
struct Node {
  std::string name;
  double t0 = 0.0, total = 0.0;
  Node* parent = nullptr;
  std::vector<std::unique_ptr<Node>> kids;
};

struct Profiler {
  bool on = false;
  std::unique_ptr<Node> root{new Node{"root"}};
  Node* cur = root.get();
  int depth = 0;

  static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
  void device_wait() {
    if (ctx().stream) (void)hipStreamSynchronize(ctx().stream);
  }
  void start(const char* name) {
    Node* n = nullptr;
    for (auto& k : cur->kids)
      if (k->name == name) { n = k.get(); break; }
    if (!n) {
      cur->kids.emplace_back(new Node{name});
      n = cur->kids.back().get();
      n->parent = cur;
    }
    device_wait();
    n->t0 = now();
    cur = n;
    ++depth;
  }
  void stop() {
    if (!cur->parent) fail(DECO_ERR_STATE, "profile stop without a matching start");
    device_wait();
    cur->total += now() - cur->t0;
    cur = cur->parent;
    --depth;
  }
  void reset() {
    root.reset(new Node{"root"});
    cur = root.get();
    depth = 0;
  }
  // core/profiler.hpp:89-112
  void print(const Node* n, int d, double parent_time, std::string& out) const {
    const double pct = parent_time == 0.0 ? 100.0 : n->total / parent_time * 100.0;
    char line[256];
    std::snprintf(line, sizeof line, "%*s\xe2\x94\x94\xe2\x94\x80\xe2\x94\x80 %-20s%6.2f%% %*.5fs\n", d * 4, "",
                  n->name.c_str(), pct, 25 - 4 * d, n->total);
    out += line;
    for (auto& k : n->kids) print(k.get(), d + 1, n->total, out);
  }
  std::string report() const {
    std::string out = "\n--------------------------------\nProfile of Execution Times (seconds):\n";
    for (auto& k : root->kids) print(k.get(), 0, 0.0, out);
    return out;
  }
  const Node* find(const Node* n, const std::string& name) const {
    if (n->name == name) return n;
    for (auto& k : n->kids)
      if (const Node* f = find(k.get(), name)) return f;
    return nullptr;
  }
};

Profiler& prof() {
  static Profiler p;
  return p;
}
}  // namespace

void prof_start(const char* name) {
  if (prof().on) prof().start(name);
}
void prof_stop() {
  if (prof().on) prof().stop();
}

}  // namespace deco

using namespace deco;

namespace deco {
ByteCount& byte_count() {
  static ByteCount c;
  return c;
}
}  // namespace deco

extern "C" {

int deco_count_bytes(int on) {
  return guard([&] {
    ByteCount& c = byte_count();
    c.on = on != 0;
    c.bytes = 0.0;
    c.launches = 0;
  });
}

int deco_counted_bytes(double* bytes, long long* launches) {
  return guard([&] {
    const ByteCount& c = byte_count();
    if (bytes) *bytes = c.bytes;
    if (launches) *launches = c.launches;
  });
}

int deco_profile_enable(int on) {
  return guard([&] {
    if (!on && prof().depth != 0) fail(DECO_ERR_STATE, "profile phases still open");
    prof().on = on != 0;
  });
}

int deco_profile_reset(void) {
  return guard([&] { prof().reset(); });
}

int deco_profile_start(const char* name) {
  return guard([&] {
    if (!name) fail(DECO_ERR_ARG, "null name");
    prof_start(name);
  });
}

int deco_profile_stop(void) {
  return guard([&] { prof_stop(); });
}

int deco_profile_time(const char* name, double* seconds) {
  return guard([&] {
    if (!name || !seconds) fail(DECO_ERR_ARG, "null argument");
    const auto* n = prof().find(prof().root.get(), name);
    *seconds = n ? n->total : 0.0;
  });
}

int deco_profile_report(char* buf, size_t len, size_t* needed) {
  return guard([&] {
    const std::string r = prof().report();
    if (needed) *needed = r.size() + 1;
    if (buf && len) {
      const size_t k = std::min(len - 1, r.size());
      std::memcpy(buf, r.data(), k);
      buf[k] = 0;
    }
  });
}

}  // extern "C"
