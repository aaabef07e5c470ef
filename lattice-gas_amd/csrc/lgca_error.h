// Error plumbing shared by every translation unit of the library: the
// thread-local message behind lg_last_error and the helper that sets it.
#pragma once
#include <string>

#include "../../include/lgca.h"

namespace lgca_internal {

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace lgca_internal
