// Per-thread message of the last failing C ABI call, shared by every entry point of the
// library (ibis_last_error).
#pragma once

#include <stdio.h>

namespace ibis {

inline char* err_buf() {
  static thread_local char buf[512] = "";
  return buf;
}

inline int fail(int code, const char* msg) {
  snprintf(err_buf(), 512, "%s", msg);
  return code;
}

template <class A, class... B> inline int fail(int code, const char* fmt, A a, B... b) {
  snprintf(err_buf(), 512, fmt, a, b...);
  return code;
}

}  // namespace ibis
