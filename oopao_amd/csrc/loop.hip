// Closed-loop helpers (tutorials/AO_closed_loop_Pyramid_WFS.py:601-638): the
// residual OPD seen by the WFS and the RMS wavefront errors recorded per
// iteration, in one pass over the pupil:
//   opd = (turb + dm) * pupil,
//   wfe_turb = std over the pupil of turb * pupil, wfe_res = std of opd (nm),
// with the moments accumulated in double, per block then over blocks in a
// fixed order (deterministic), written straight into the device traces.
#include "common.hpp"

namespace oopao {

constexpr int kResidBlocks = 64;

__global__ void __launch_bounds__(256) k_resid_partial(const float* __restrict__ turb, const float* __restrict__ dm,
                                                       const float* __restrict__ pupil, float* __restrict__ opd,
                                                       int count, double* __restrict__ partial) {
  __shared__ double red[4][256];
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const float p = pupil[i];
    const float t = turb[i];
    const float r = (t + (dm ? dm[i] : 0.f)) * p;
    opd[i] = r;
    const double tp = (double)t * (double)p;
    s1 += tp;
    s2 += tp * tp;
    s3 += (double)r;
    s4 += (double)r * (double)r;
  }
  red[0][threadIdx.x] = s1;
  red[1][threadIdx.x] = s2;
  red[2][threadIdx.x] = s3;
  red[3][threadIdx.x] = s4;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// one wavefront: lane b loads block b's moments, fixed-order butterfly sum over the lanes
__global__ void k_resid_final(const double* __restrict__ partial, int nblk, double npix, double* __restrict__ wfe_turb,
                              double* __restrict__ wfe_res) {
  const int lane = threadIdx.x;
  double m[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double v = lane < nblk ? partial[lane * 4 + q] : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    m[q] = v;
  }
  if (lane != 0) return;
  const double mt = m[0] / npix, mr = m[2] / npix;
  const double vt = fmax(m[1] / npix - mt * mt, 0.0), vr = fmax(m[3] / npix - mr * mr, 0.0);
  if (wfe_turb) *wfe_turb = sqrt(vt) * 1e9;
  if (wfe_res) *wfe_res = sqrt(vr) * 1e9;
}

}  // namespace oopao

using namespace oopao;

extern "C" {

int oopao_loop_residual_workspace_bytes(size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  *bytes = (size_t)kResidBlocks * 4 * sizeof(double);
  return OK;
}

int oopao_loop_residual(const float* turb, const float* dm, const float* pupil, float* opd, int count, double npix,
                        void* workspace, double* wfe_turb, double* wfe_res, void* stream) {
  OOPAO_CHECK_ARG(turb && pupil && opd && workspace, "null pointer argument");
  OOPAO_CHECK_ARG(count > 0 && npix > 0, "invalid sizes count=%d npix=%g", count, npix);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  double* partial = reinterpret_cast<double*>(workspace);
  hipLaunchKernelGGL(k_resid_partial, dim3(kResidBlocks), dim3(256), 0, st, turb, dm, pupil, opd, count, partial);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(64), 0, st, partial, kResidBlocks, npix, wfe_turb, wfe_res);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
