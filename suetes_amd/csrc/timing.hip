// Optional per-kernel timing of the hot kernels, for measurement runs.
//
// When enabled, every launch of a timed kernel is bracketed by a pair of HIP
// events on the launch's own stream, and its algorithmic bytes (each input
// and output array counted once) are recorded on the host.  Launches that a
// device stop/gate flag retires do no work: a one-thread probe launched just
// before them records the flag, and such launches add time but no bytes.
// Disabled (the default), the hooks cost one branch.
#include <mutex>
#include <vector>

#include "common.h"

namespace sisl {

namespace {
struct Rec {
  int slot;          // kind * 2 + (FP32 working type)
  double bytes;      // algorithmic bytes if the launch does work
  hipEvent_t a, b;
  int probe;         // index into the device skip log, -1 if ungated
};
struct Timer {
  std::mutex mu;
  bool on = false;
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  int *d_skip = nullptr;  // device log of retired launches
  int cap = 0, used = 0;
};
Timer &timer() {
  static Timer t;
  return t;
}
hipEvent_t take_event(Timer &t) {
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

__global__ void skip_probe_kernel(const int *stop, const int *gate, int *out) {
  *out = ((stop && *stop) || (gate && !*gate)) ? 1 : 0;
}

bool timer_on() { return timer().on; }

int timer_begin(int kind, bool fp32, double bytes, const int *stop, const int *gate, hipStream_t s) {
  Timer &t = timer();
  if (!t.on) return -1;
  std::lock_guard<std::mutex> g(t.mu);
  Rec r{kind * 2 + (fp32 ? 1 : 0), bytes, take_event(t), take_event(t), -1};
  if (!r.a || !r.b) return -1;
  if ((stop || gate) && t.used < t.cap) {
    r.probe = t.used++;
    hipLaunchKernelGGL(skip_probe_kernel, dim3(1), dim3(1), 0, s, stop, gate, t.d_skip + r.probe);
  }
  hipEventRecord(r.a, s);
  t.recs.push_back(r);
  return (int)t.recs.size() - 1;
}

void timer_end(int id, hipStream_t s) {
  if (id < 0) return;
  Timer &t = timer();
  std::lock_guard<std::mutex> g(t.mu);
  if (id < (int)t.recs.size()) hipEventRecord(t.recs[id].b, s);
}

}  // namespace sisl

extern "C" {

int sisl_timer_enable(int32_t on) {
  sisl::Timer &t = sisl::timer();
  std::lock_guard<std::mutex> g(t.mu);
  if (on) {
    if (!t.d_skip) {
      t.cap = 1 << 20;
      if (hipMalloc(&t.d_skip, sizeof(int) * t.cap) != hipSuccess) {
        t.d_skip = nullptr;
        t.cap = 0;
        return SISL_ERR_LAUNCH;
      }
    }
    for (auto &r : t.recs) {
      t.pool.push_back(r.a);
      t.pool.push_back(r.b);
    }
    t.recs.clear();
    t.used = 0;
  }
  t.on = on != 0;
  return SISL_OK;
}

int sisl_timer_read(int32_t kind, int32_t fp32, double *total_ms, int64_t *launches,
                    int64_t *working_launches, double *bytes) {
  if (kind < 0 || kind >= SISL_T_COUNT || !total_ms || !launches || !working_launches || !bytes)
    return SISL_ERR_ARG;
  sisl::Timer &t = sisl::timer();
  std::lock_guard<std::mutex> g(t.mu);
  std::vector<int> skip(t.used);
  if (t.used && hipMemcpy(skip.data(), t.d_skip, sizeof(int) * t.used, hipMemcpyDeviceToHost) != hipSuccess)
    return SISL_ERR_LAUNCH;
  double ms = 0, by = 0;
  int64_t n = 0, nw = 0;
  const int slot = kind * 2 + (fp32 ? 1 : 0);
  for (auto &r : t.recs) {
    if (r.slot != slot) continue;
    if (hipEventSynchronize(r.b) != hipSuccess) return SISL_ERR_LAUNCH;
    float e = 0;
    if (hipEventElapsedTime(&e, r.a, r.b) != hipSuccess) return SISL_ERR_LAUNCH;
    ms += e;
    ++n;
    if (r.probe < 0 || !skip[r.probe]) {
      ++nw;
      by += r.bytes;
    }
  }
  *total_ms = ms;
  *launches = n;
  *working_launches = nw;
  *bytes = by;
  return SISL_OK;
}

}  // extern "C"
