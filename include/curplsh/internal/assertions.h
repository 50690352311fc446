// Precondition checks of the C++ host layer. Like the reference
// (include/internal/assertions.h:7-14) a violated precondition prints the
// failed condition and aborts; the C ABI below the classes returns status
// codes instead, and curplsh::checkStatus turns a failure back into this
// abort with the library's message.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "rplsh.h"

#define host_assert(X)                                                                   \
  do {                                                                                   \
    if (!(X)) {                                                                          \
      std::fprintf(stderr, "Assertion '%s' failed at %s:%d\n", #X, __FILE__, __LINE__); \
      std::abort();                                                                      \
    }                                                                                    \
  } while (0)

namespace curplsh {

// Aborts with the C ABI's last error message if status is not RPLSH_OK.
inline void checkStatus(rplsh_status status, const char* what) {
  if (status != RPLSH_OK) {
    std::fprintf(stderr, "curplsh: %s failed (status %d): %s\n", what, (int)status,
                 rplsh_last_error());
    std::abort();
  }
}

}  // namespace curplsh
