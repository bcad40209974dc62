// errors.h -- thread-local last-error message behind hlb_last_error().
#pragma once

#include <string>

namespace hlb {

std::string& last_error();

// Records msg as the calling thread's last error and returns code.
inline int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

}  // namespace hlb
