// Error reporting shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

// Records the message of the last failure on this thread (read back by
// kmt_last_error) and returns `code`.
int kmt_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIP_TRY(expr)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return kmt_fail(e_ == hipErrorOutOfMemory ? KMT_ERR_MEMORY : KMT_ERR_HIP, "%s failed: %s (%s:%d)", \
                      #expr, hipGetErrorString(e_), __FILE__, __LINE__);                              \
  } while (0)

#define KMT_TRY(expr)              \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != KMT_OK) return rc_; \
  } while (0)
