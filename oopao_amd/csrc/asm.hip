// Angular-spectrum (Fresnel) propagation of the source field through the layers, for the
// scintillation mode of the atmosphere, on CDNA4.
//
// Reference: Atmosphere.fill_scintillation_support / set_scintillation_support / ASM
// (Atmosphere.py:644-706, 938-972): per source the unit field is multiplied by each layer's
// phase exp(i k OPD_layer), from the highest layer down, and propagated to the next layer (or
// the ground) by
//     E' = fftshift(ifft2(ifftshift(fftshift(fft2(ifftshift(E))) H)))
//     H  = exp(-i pi lambda z (fx^2 + fy^2)),  f on the centred grid (-N/2 .. N/2 - 1) / (N pitch)
// (equal input and output pitch). For even N that is the centred 2-D FFT pair of fft2.hip
// around one multiply by H. The readout gives |E|^2 and the phase as OPD: the wrapped angle, or
// a geometric guide phase plus the wrapped residual angle(E exp(-i k OPD_geo)).
#include "common.hpp"

namespace oopao {

namespace {

__global__ void k_field_phase(const float* __restrict__ opd, float k, float2* __restrict__ field, size_t count,
                              int init) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s, c;
  sincosf(k * opd[i], &s, &c);
  const float2 e = init ? make_float2(1.f, 0.f) : field[i];
  field[i] = make_float2(e.x * c - e.y * s, e.x * s + e.y * c);
}

__global__ void k_apply_transfer(float2* __restrict__ X, const float2* __restrict__ H, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float2 a = X[i], h = H[i];
  X[i] = make_float2(a.x * h.x - a.y * h.y, a.x * h.y + a.y * h.x);
}

__global__ void k_field_readout(const float2* __restrict__ field, const float* __restrict__ geo, float k,
                                float* __restrict__ intensity, float* __restrict__ opd, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float2 e = field[i];
  if (intensity) intensity[i] = e.x * e.x + e.y * e.y;
  if (!opd) return;
  if (geo) {
    // guide + wrapped residual: angle(E exp(-i k geo))
    float s, c;
    sincosf(k * geo[i], &s, &c);
    const float re = e.x * c + e.y * s, im = e.y * c - e.x * s;
    opd[i] = geo[i] + atan2f(im, re) / k;
  } else {
    opd[i] = atan2f(e.y, e.x) / k;
  }
}

unsigned nblocks(size_t count) { return (unsigned)((count + 255) / 256); }

}  // namespace

}  // namespace oopao

using namespace oopao;

extern "C" int oopao_fft2(const void* in, void* out, int N, int B, int dir, int centred, float scale, void* ws,
                          size_t ws_bytes, void* stream);

extern "C" {

int oopao_field_phase(const float* opd, float k, void* field, int count, int init, void* stream) {
  OOPAO_CHECK_ARG(opd && field && count > 0, "invalid arguments");
  hipLaunchKernelGGL(k_field_phase, dim3(nblocks((size_t)count)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     opd, k, reinterpret_cast<float2*>(field), (size_t)count, init);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_asm_bytes(int N, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes && N >= 2, "invalid arguments");
  *bytes = 3 * (size_t)N * N * sizeof(float2);
  return OK;
}

int oopao_asm_propagate(void* field, const void* H, int N, void* ws, size_t ws_bytes, void* stream) {
  OOPAO_CHECK_ARG(field && H && ws, "null pointer argument");
  OOPAO_CHECK_ARG(N >= 2 && (N % 2) == 0, "invalid N=%d (even N required)", N);
  const size_t F = (size_t)N * N;
  if (ws_bytes < 3 * F * sizeof(float2)) {
    set_error("workspace too small");
    return ERR_WORKSPACE;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float2* X = reinterpret_cast<float2*>(ws);
  void* fws = X + F;
  int rc = oopao_fft2(field, X, N, 1, -1, 1, 1.0f, fws, 2 * F * sizeof(float2), stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_apply_transfer, dim3(nblocks(F)), dim3(256), 0, st, X, reinterpret_cast<const float2*>(H), F);
  OOPAO_LAUNCH_CHECK();
  return oopao_fft2(X, field, N, 1, 1, 1, 1.0f / ((float)N * (float)N), fws, 2 * F * sizeof(float2), stream);
}

int oopao_field_readout(const void* field, const float* geo, float k, float* intensity, float* opd, int count,
                        void* stream) {
  OOPAO_CHECK_ARG(field && count > 0 && (intensity || opd), "invalid arguments");
  hipLaunchKernelGGL(k_field_readout, dim3(nblocks((size_t)count)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const float2*>(field), geo, k, intensity,
                     opd, (size_t)count);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
