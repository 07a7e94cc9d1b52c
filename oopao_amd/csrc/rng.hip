// Counter-based normal draws on CDNA4 for the atmosphere's device RNG mode.
//
// The reference draws every extruded ring's noise and every boiling screen from numpy's RandomState on
// the host (Atmosphere.py:400-417, phaseStats.py:176-192). In the device mode the same draws come from
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011): pair p of draw d on stream s is the Philox block of
// counter (p, d_lo, d_hi, s) under the layer's key, mapped to two 53-bit uniforms and a Box-Muller pair.
// Any draw is a pure function of (key, stream, draw, index): no state on the device, no host traffic.
#include "common.hpp"

namespace oopao {

__device__ __forceinline__ void philox_round(unsigned& c0, unsigned& c1, unsigned& c2, unsigned& c3, unsigned k0,
                                             unsigned k1) {
  const unsigned long long p0 = (unsigned long long)c0 * 0xD2511F53u;
  const unsigned long long p1 = (unsigned long long)c2 * 0xCD9E8D57u;
  const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
  const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
  const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
  c0 = n0;
  c1 = lo1;
  c2 = n2;
  c3 = lo0;
}

__device__ __forceinline__ void philox4x32_10(unsigned& c0, unsigned& c1, unsigned& c2, unsigned& c3, unsigned k0,
                                              unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c0, c1, c2, c3, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// out[j], j < count: normal j of draw `draw` on stream `stream` (pair j / 2 of the draw's counter blocks)
__global__ void __launch_bounds__(256) k_normal_philox(double* __restrict__ out, long long count, unsigned key0,
                                                       unsigned key1, unsigned stream, unsigned long long draw) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (2 * p >= count) return;
  unsigned c0 = (unsigned)p, c1 = (unsigned)draw, c2 = (unsigned)(draw >> 32), c3 = stream;
  philox4x32_10(c0, c1, c2, c3, key0, key1);
  const double u1 = ((double)(c0 >> 5) * 67108864.0 + (double)(c1 >> 6)) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(c2 >> 5) * 67108864.0 + (double)(c3 >> 6)) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(1.0 - u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  out[2 * p] = rad * c;
  if (2 * p + 1 < count) out[2 * p + 1] = rad * s;
}

}  // namespace oopao

using namespace oopao;

extern "C" int oopao_normal_philox(double* out, long long count, unsigned key0, unsigned key1, unsigned stream,
                                   unsigned long long draw, void* hip_stream) {
  OOPAO_CHECK_ARG(out != nullptr && count > 0 && count < (1LL << 33), "invalid arguments count=%lld", count);
  const long long pairs = (count + 1) / 2;
  hipLaunchKernelGGL(k_normal_philox, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(hip_stream), out, count, key0, key1, stream, draw);
  OOPAO_LAUNCH_CHECK();
  return OK;
}
