// Error reporting shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "effsc_hip.h"
#include "shake.hpp"

namespace effsc_hip {
// Records the thread-local last error (effsc_hip_last_error) and returns code.
int set_error(int code, const std::string& msg);
const char* last_error();
}  // namespace effsc_hip

// Host transcript handle (effsc_hip_transcript_new / _new_test).
struct effsc_hip_transcript {
  int kind;  // 0: Fiat-Shamir, 1: test RNG
  effsc_hip::FsTranscript fs;
  uint64_t rng;
};

#define HIPCHK(expr)                                                                                       \
  do {                                                                                                     \
    hipError_t e_ = (expr);                                                                                \
    if (e_ != hipSuccess)                                                                                  \
      return effsc_hip::set_error(EFFSC_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TRY(expr)                          \
  do {                                     \
    int rc_ = (expr);                      \
    if (rc_ != EFFSC_HIP_OK) return rc_;   \
  } while (0)
