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

// Logical task of workgroup i out of G such that consecutive tasks run on one XCD: workgroups are
// dispatched round-robin over the 8 XCDs (i % 8), each XCD with its own L2, so tasks that share
// cache lines should land on the same one. A bijection of [0, G).
__device__ __forceinline__ int xcd_task(int i, int G) {
  const int per = (G + 7) / 8;
  const int tall = (G % 8) ? (G % 8) : 8;
  const int x = i & 7, slot = i >> 3;
  return x < tall ? x * per + slot : tall * per + (x - tall) * (per - 1) + slot;
}

// init + p[0] + p[stride] + ... + p[(count - 1) stride], added in index order (the same result as the
// plain loop), with the loads issued eight at a time: in the plain loop each load waited for the one before
// (the additions form a chain), count round trips to memory.
template <typename T>
__device__ __forceinline__ T sum_in_order(const T* p, size_t stride, int count, T init = T(0)) {
  T s = init;
  for (int k0 = 0; k0 < count; k0 += 8) {
    T v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (k0 + j < count) ? p[(size_t)(k0 + j) * stride] : T(0);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (k0 + j < count) s += v[j];
  }
  return s;
}

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
// built once from double-precision values and kept for the process lifetime
// (allocated and copied synchronously on first use: oopao_prepare_tables builds
// it ahead, e.g. before a graph capture).
int twiddles(int N, const float2** out);

// Workgroups of `block` threads that the current device keeps resident at once for `kernel`: compute
// units x workgroups per compute unit at the kernel's register / LDS occupancy. Cached per (device,
// kernel); queries device attributes only (nothing is enqueued, so it is safe under stream capture).
int resident_workgroups(const void* kernel, int block, int* out);

}  // namespace oopao
