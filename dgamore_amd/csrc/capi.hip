// Library-wide state of libdgami: last-error text and the device workspace.
#include <mutex>
#include <vector>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"

namespace dgami {

namespace {
thread_local std::string g_last_error;

struct Slot {
  void* ptr = nullptr;
  size_t bytes = 0;
  int device = -1;
};
std::mutex g_ws_mutex;
std::vector<Slot> g_slots;
// While captured graphs are alive (g_holds > 0) an outgrown buffer may still be named by a
// graph node, so it is kept here instead of freed, until the last hold is released.
int g_holds = 0;
std::vector<void*> g_retired;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

namespace {
struct Profile {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  double bytes = 0, flops = 0;
};
std::mutex g_prof_mutex;
Profile g_prof;
}  // namespace

bool profile_on() { return g_prof.on; }

int profile_begin(hipStream_t stream, hipEvent_t* start) {
  DGAMI_HIP_CHECK(hipEventCreate(start));
  DGAMI_HIP_CHECK(hipEventRecord(*start, stream));
  return 0;
}

int profile_end(hipStream_t stream, hipEvent_t start, double bytes, double flops) {
  hipEvent_t stop;
  DGAMI_HIP_CHECK(hipEventCreate(&stop));
  DGAMI_HIP_CHECK(hipEventRecord(stop, stream));
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  g_prof.events.emplace_back(start, stop);
  g_prof.bytes += bytes;
  g_prof.flops += flops;
  return 0;
}

static void profile_clear() {
  for (auto& e : g_prof.events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  g_prof.events.clear();
  g_prof.bytes = g_prof.flops = 0;
}
const char* last_error() { return g_last_error.c_str(); }

void* workspace(int slot, size_t bytes) {
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("workspace: hipGetDevice failed");
    return nullptr;
  }
  if (slot >= static_cast<int>(g_slots.size())) g_slots.resize(slot + 1);
  Slot& s = g_slots[slot];
  if (s.ptr != nullptr && s.bytes >= bytes && s.device == dev) return s.ptr;
  if (s.ptr != nullptr) {
    if (g_holds > 0) {
      g_retired.push_back(s.ptr);  // a live graph may still use it
    } else if (hipDeviceSynchronize() != hipSuccess || hipFree(s.ptr) != hipSuccess) {
      // the previous buffer may still be in use by queued work on any stream
      set_error("workspace: could not release the previous buffer");
      return nullptr;
    }
    s.ptr = nullptr;
    s.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  if (hipMalloc(&s.ptr, want) != hipSuccess) {
    s.ptr = nullptr;
    set_error("workspace: hipMalloc of " + std::to_string(want) + " bytes failed");
    return nullptr;
  }
  s.bytes = want;
  s.device = dev;
  return s.ptr;
}

namespace {
struct ExtraStreams {
  hipStream_t stream[MAX_EXTRA_STREAMS] = {};
  hipEvent_t fork[MAX_EXTRA_STREAMS] = {}, join[MAX_EXTRA_STREAMS] = {};
};
std::mutex g_xs_mutex;
std::vector<ExtraStreams> g_xs;  // per device
}  // namespace

int extra_stream_fork(hipStream_t caller, int k, hipStream_t* out) {
  DGAMI_ARG_CHECK(k >= 1 && k <= MAX_EXTRA_STREAMS, "extra stream index out of range");
  int dev = 0;
  DGAMI_HIP_CHECK(hipGetDevice(&dev));
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_xs_mutex);
    if (dev >= static_cast<int>(g_xs.size())) g_xs.resize(dev + 1);
    ExtraStreams& x = g_xs[dev];
    if (x.stream[k - 1] == nullptr) {
      DGAMI_HIP_CHECK(hipStreamCreateWithFlags(&x.stream[k - 1], hipStreamNonBlocking));
      DGAMI_HIP_CHECK(hipEventCreateWithFlags(&x.fork[k - 1], hipEventDisableTiming));
      DGAMI_HIP_CHECK(hipEventCreateWithFlags(&x.join[k - 1], hipEventDisableTiming));
    }
    s = x.stream[k - 1];
    fork = x.fork[k - 1];
  }
  DGAMI_HIP_CHECK(hipEventRecord(fork, caller));
  DGAMI_HIP_CHECK(hipStreamWaitEvent(s, fork, 0));
  *out = s;
  return 0;
}

int extra_stream_join(hipStream_t caller, int k) {
  DGAMI_ARG_CHECK(k >= 1 && k <= MAX_EXTRA_STREAMS, "extra stream index out of range");
  int dev = 0;
  DGAMI_HIP_CHECK(hipGetDevice(&dev));
  hipStream_t s = nullptr;
  hipEvent_t join = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_xs_mutex);
    s = g_xs[dev].stream[k - 1];
    join = g_xs[dev].join[k - 1];
  }
  DGAMI_HIP_CHECK(hipEventRecord(join, s));
  DGAMI_HIP_CHECK(hipStreamWaitEvent(caller, join, 0));
  return 0;
}

}  // namespace dgami

extern "C" {

const char* dgami_last_error(void) { return dgami::last_error(); }

int dgami_profile_enable(int on) {
  std::lock_guard<std::mutex> lock(dgami::g_prof_mutex);
  dgami::profile_clear();
  dgami::g_prof.on = on != 0;
  return 0;
}

int dgami_profile_read(double* ms, double* bytes, double* flops, int64_t* launches) {
  using namespace dgami;
  DGAMI_ARG_CHECK(ms && bytes && flops && launches, "profile_read: null pointer");
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  double total = 0;
  for (auto& e : g_prof.events) {
    DGAMI_HIP_CHECK(hipEventSynchronize(e.second));
    float t = 0.f;
    DGAMI_HIP_CHECK(hipEventElapsedTime(&t, e.first, e.second));
    total += t;
  }
  *ms = total;
  *bytes = g_prof.bytes;
  *flops = g_prof.flops;
  *launches = static_cast<int64_t>(g_prof.events.size());
  return 0;
}

int dgami_workspace_hold(void) {
  std::lock_guard<std::mutex> lock(dgami::g_ws_mutex);
  ++dgami::g_holds;
  return dgami::g_holds;
}

int dgami_workspace_release(void) {
  using namespace dgami;
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  DGAMI_ARG_CHECK(g_holds > 0, "workspace_release: no hold to release");
  if (--g_holds == 0 && !g_retired.empty()) {
    DGAMI_HIP_CHECK(hipDeviceSynchronize());
    for (void* p : g_retired) DGAMI_HIP_CHECK(hipFree(p));
    g_retired.clear();
  }
  return g_holds;
}

int dgami_version(void) { return 1; }

}  // extern "C"
