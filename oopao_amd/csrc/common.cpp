#include "common.hpp"

#include <vector>

namespace oopao {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

const char* get_error() { return g_error.c_str(); }

int twiddles(int N, const float2** out) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, float2*> cache;
  int dev = 0;
  OOPAO_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(dev, N);
  auto it = cache.find(key);
  if (it != cache.end()) {
    *out = it->second;
    return OK;
  }
  std::vector<float2> h(N);
  for (int m = 0; m < N; ++m) {
    // angle reduced exactly on the rational m/N before scaling by 2 pi
    long double frac = (long double)m / (long double)N;
    long double ang = -2.0L * 3.141592653589793238462643383279502884L * frac;
    h[m] = make_float2((float)std::cos((double)ang), (float)std::sin((double)ang));
  }
  float2* d = nullptr;
  OOPAO_HIP(hipMalloc(&d, sizeof(float2) * N));
  OOPAO_HIP(hipMemcpy(d, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  cache[key] = d;
  *out = d;
  return OK;
}

int resident_workgroups(const void* kernel, int block, int* out) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, int> cache;
  int dev = 0;
  OOPAO_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(dev, kernel);
  auto it = cache.find(key);
  if (it != cache.end()) {
    *out = it->second;
    return OK;
  }
  int cus = 0, per_cu = 0;
  OOPAO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  OOPAO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
  const int v = cus * per_cu > 0 ? cus * per_cu : 1;
  cache[key] = v;
  *out = v;
  return OK;
}

}  // namespace oopao

extern "C" const char* oopao_last_error(void) { return oopao::get_error(); }
