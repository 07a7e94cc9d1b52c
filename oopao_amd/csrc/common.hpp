// Shared host-side plumbing for the C-ABI: error reporting and twiddle tables.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <utility>

namespace oopao {

// Error codes returned by every C-ABI entry point (0 = success).
enum : int {
  OK = 0,
  ERR_ARG = 1,       // invalid argument (shape, size, pointer)
  ERR_UNSUPPORTED = 2,  // geometry outside the compiled set
  ERR_HIP = 3,       // HIP runtime error
  ERR_WORKSPACE = 4  // workspace too small
};

void set_error(const char* fmt, ...);
const char* get_error();

#define OOPAO_CHECK_ARG(cond, ...)        \
  do {                                    \
    if (!(cond)) {                        \
      ::oopao::set_error(__VA_ARGS__);    \
      return ::oopao::ERR_ARG;            \
    }                                     \
  } while (0)

#define OOPAO_HIP(call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      ::oopao::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                       \
      return ::oopao::ERR_HIP;                                                            \
    }                                                                                     \
  } while (0)

#define OOPAO_LAUNCH_CHECK()                                                                      \
  do {                                                                                            \
    hipError_t e_ = hipGetLastError();                                                            \
    if (e_ != hipSuccess) {                                                                       \
      ::oopao::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__,     \
                         __LINE__);                                                               \
      return ::oopao::ERR_HIP;                                                                    \
    }                                                                                             \
  } while (0)

// Device table tw[m] = exp(-2 pi i m / N), m = 0..N-1, one per (device, N),
// built once from double-precision values and kept for the process lifetime.
int twiddles(int N, const float2** out);

}  // namespace oopao
