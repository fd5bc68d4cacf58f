// Library-wide state of libdgami: last-error text and the device workspace.
#include <mutex>
#include <vector>

#include "dgami.h"
#include "dgami_internal.hpp"

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
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
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
    // the previous buffer may still be in use by queued work on any stream
    if (hipDeviceSynchronize() != hipSuccess || hipFree(s.ptr) != hipSuccess) {
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

}  // namespace dgami

extern "C" {

const char* dgami_last_error(void) { return dgami::last_error(); }

int dgami_version(void) { return 1; }

}  // extern "C"
