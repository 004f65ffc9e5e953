// torque/error.h -- the reference's exception type (src/error.h:8-14) and the
// translation of C-ABI status codes into it.
#pragma once

#include <stdexcept>
#include <string>

#include "torque_hip.h"

struct Error : public std::runtime_error {
  explicit Error(const std::string &error_message)
      : std::runtime_error("Error: " + error_message) {}
  explicit Error(const char *error_message)
      : std::runtime_error("Error: " + std::string(error_message)) {}
};

namespace torque {
namespace gpu {

// Throws Error with the handle's message for a failed call (the reference
// exits the process on device errors, src/gpu/util/lib_helper.cuh:39-77).
inline void check(tq_status s, tq_handle_t h = nullptr) {
  if (s == TQ_SUCCESS) return;
  std::string msg = tq_status_string(s);
  if (h) {
    std::string extra = tq_last_error(h);
    if (!extra.empty() && extra != msg) msg = extra;
  }
  throw Error(msg);
}

// Process-wide handle on device 0 for the calls whose reference signature
// carries no handle (construction, query, modify, hard_transpose).
inline tq_handle_t default_handle() {
  static tq_handle_t h = [] {
    tq_handle_t x = nullptr;
    check(tq_create(&x, 0));
    return x;
  }();
  return h;
}

}  // namespace gpu
}  // namespace torque
