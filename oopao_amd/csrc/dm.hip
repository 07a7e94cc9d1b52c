// Deformable-mirror influence functions and OPD contractions on CDNA4.
//
// Reference: DeformableMirror.modesComputation (DeformableMirror.py:572-600)
// builds Gaussian influence functions
//   G_k(X, Y) = exp(-(a (X-x0_k)^2 + 2 b (X-x0_k)(Y-y0_k) + c (Y-y0_k)^2))
// on the grid X = linspace(0, 1, n) * n, stored as modes[n^2, nAct]; the coefs
// setter (651-686) computes OPD = modes @ coefs (GEMV, or GEMM for a batch).
//
// Kernels:
//  * k_gauss_modes   : builds modes (float32) directly on the device.
//  * k_gemv_rows     : OPD[p] = sum_k modes[p,k] c[k] for a [P, K] row-major
//                      matrix (HBM-bound; one wavefront per output row-block,
//                      float4 loads, wavefront reduction). Also used for the
//                      reconstructor product R @ s.
//  * k_sep_opd       : for the zero-misregistration Cartesian DM the Gaussian
//                      factorises, G_k(X,Y) = gx(X - x0_k) gy(Y - y0_k), and the
//                      actuators sit on an nA x nA grid, so
//                      OPD = Gy C Gx^T with C the coefficient grid (zeros at
//                      invalid actuators): two small contractions instead of
//                      an n^2 x nAct GEMV.
#include "common.hpp"

namespace oopao {

__global__ void k_gauss_modes(float* __restrict__ modes, int n, int nact, const float* __restrict__ x0,
                              const float* __restrict__ y0, double a, double b, double c, float sign) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)n * n * nact;
  if (idx >= total) return;
  const int k = (int)(idx % nact);
  const size_t p = idx / nact;
  const int iy = (int)(p / n), ix = (int)(p % n);
  // X = linspace(0, 1, n) * n  (DeformableMirror.py:580)
  const double X = (n > 1 ? (double)ix / (double)(n - 1) : 0.0) * n;
  const double Y = (n > 1 ? (double)iy / (double)(n - 1) : 0.0) * n;
  const double dx = X - (double)x0[k], dy = Y - (double)y0[k];
  modes[idx] = sign * (float)exp(-(a * dx * dx + 2.0 * b * dx * dy + c * dy * dy));
}

// y[r] = alpha * sum_k M[r, k] x[k] + beta * y0[r]; M row-major [R, K] float32.
// One wavefront per row; lanes stride over k with float4 loads when aligned.
__global__ void __launch_bounds__(256) k_gemv_rows(const float* __restrict__ M, const float* __restrict__ x,
                                                   float* __restrict__ y, const float* __restrict__ y0,
                                                   int R, int K, float alpha, float beta) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + wave;
  if (r >= R) return;
  const float* row = M + (size_t)r * K;
  float acc = 0.f;
  if ((K & 3) == 0) {
    const float4* r4 = reinterpret_cast<const float4*>(row);
    const float4* x4 = reinterpret_cast<const float4*>(x);
    for (int k = lane; k < K / 4; k += 64) {
      const float4 m = r4[k];
      const float4 v = x4[k];
      acc = fmaf(m.x, v.x, acc);
      acc = fmaf(m.y, v.y, acc);
      acc = fmaf(m.z, v.z, acc);
      acc = fmaf(m.w, v.w, acc);
    }
  } else {
    for (int k = lane; k < K; k += 64) acc = fmaf(row[k], x[k], acc);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) y[r] = alpha * acc + (y0 ? beta * y0[r] : 0.f);
}

// Batched separable OPD: for each batch j,
//   T[j][a][x] = sum_b C[j][a][b] gx[x][b]      (nA x n)
//   OPD[j][y][x] = sum_a gy[y][a] T[j][a][x]     (n x n)
// gx, gy [n][nA] float32; C [B][nA][nA] float32 (row a = actuator row, along Y).
__global__ void k_sep_stage1(const float* __restrict__ C, const float* __restrict__ gx,
                             float* __restrict__ Tm, int n, int nA, int B) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)B * nA * n;
  if (idx >= total) return;
  const int x = (int)(idx % n);
  const int a = (int)((idx / n) % nA);
  const int j = (int)(idx / ((size_t)n * nA));
  const float* crow = C + ((size_t)j * nA + a) * nA;
  const float* g = gx + (size_t)x * nA;
  float acc = 0.f;
  for (int b = 0; b < nA; ++b) acc = fmaf(crow[b], g[b], acc);
  Tm[idx] = acc;
}

__global__ void k_sep_stage2(const float* __restrict__ Tm, const float* __restrict__ gy,
                             float* __restrict__ opd, int n, int nA, int B) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)B * n * n;
  if (idx >= total) return;
  const int x = (int)(idx % n);
  const int y = (int)((idx / n) % n);
  const int j = (int)(idx / ((size_t)n * n));
  const float* g = gy + (size_t)y * nA;
  const float* tcol = Tm + (size_t)j * nA * n + x;
  float acc = 0.f;
  for (int a = 0; a < nA; ++a) acc = fmaf(g[a], tcol[(size_t)a * n], acc);
  opd[idx] = acc;
}

}  // namespace oopao

using namespace oopao;

extern "C" {

int oopao_dm_gaussian_modes(float* modes, int n, int nact, const float* x0, const float* y0, double a,
                            double b, double c, float sign, void* stream) {
  OOPAO_CHECK_ARG(modes && x0 && y0, "null pointer argument");
  OOPAO_CHECK_ARG(n > 0 && nact > 0, "invalid sizes n=%d nact=%d", n, nact);
  const size_t total = (size_t)n * n * nact;
  hipLaunchKernelGGL(k_gauss_modes, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), modes, n, nact, x0, y0, a, b, c, sign);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_gemv(const float* M, const float* x, float* y, const float* y0, int R, int K, float alpha,
               float beta, void* stream) {
  OOPAO_CHECK_ARG(M && x && y, "null pointer argument");
  OOPAO_CHECK_ARG(R > 0 && K > 0, "invalid sizes R=%d K=%d", R, K);
  OOPAO_CHECK_ARG(((K & 3) != 0) || ((reinterpret_cast<size_t>(M) & 15) == 0 &&
                                      (reinterpret_cast<size_t>(x) & 15) == 0),
                  "M and x must be 16-byte aligned when K %% 4 == 0");
  hipLaunchKernelGGL(k_gemv_rows, dim3((unsigned)((R + 3) / 4)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), M, x, y, y0, R, K, alpha, beta);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_dm_separable_opd(const float* C, const float* gx, const float* gy, float* tmp, float* opd, int n,
                           int nA, int B, void* stream) {
  OOPAO_CHECK_ARG(C && gx && gy && tmp && opd, "null pointer argument");
  OOPAO_CHECK_ARG(n > 0 && nA > 0 && B > 0, "invalid sizes n=%d nA=%d B=%d", n, nA, B);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  size_t t1 = (size_t)B * nA * n;
  hipLaunchKernelGGL(k_sep_stage1, dim3((unsigned)((t1 + 255) / 256)), dim3(256), 0, st, C, gx, tmp, n, nA, B);
  OOPAO_LAUNCH_CHECK();
  size_t t2 = (size_t)B * n * n;
  hipLaunchKernelGGL(k_sep_stage2, dim3((unsigned)((t2 + 255) / 256)), dim3(256), 0, st, tmp, gy, opd, n, nA, B);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
