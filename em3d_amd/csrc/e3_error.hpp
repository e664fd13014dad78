// Per-thread last-error message shared by every C-ABI entry point.
#pragma once

#include <string>

namespace e3 {

inline std::string& last_error_store() {
  static thread_local std::string s;
  return s;
}

inline int set_error(const std::string& msg) {
  last_error_store() = msg;
  return 1;
}

}  // namespace e3
