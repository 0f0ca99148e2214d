// Shared host/device helpers for the diffren_amd HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace dr {

// Per-thread last error message returned by dr_last_error().
void set_error(const std::string& msg);
const char* get_error();

// Status codes returned by every C-ABI entry point.
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kWorkspaceTooSmall = 2,
  kLaunchFailed = 3,
  kUnsupported = 4,
};

inline int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

}  // namespace dr

#define DR_CHECK_ARG(cond, msg)                                   \
  do {                                                            \
    if (!(cond)) return dr::fail(dr::kInvalidArgument, (msg));    \
  } while (0)

#define DR_CHECK_LAUNCH(what)                                                 \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess)                                                     \
      return dr::fail(dr::kLaunchFailed,                                      \
                      std::string(what) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define DR_CHECK_HIP(call)                                                    \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return dr::fail(dr::kLaunchFailed,                                      \
                      std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

// Wave (64 lanes on CDNA) helpers.
namespace dr {
__device__ inline int lane_id() { return __lane_id(); }
__device__ inline uint64_t lanemask_lt() {
  return (uint64_t(1) << lane_id()) - 1;
}
}  // namespace dr
