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
//                      matrix (HBM-bound; four rows per workgroup, K split over
//                      its four wavefronts, each slice of the vector loaded once
//                      for the four rows, float4 loads, fixed-order reduction).
//                      Also used for the reconstructor product R @ s.
//  * k_sep_opd       : for the zero-misregistration Cartesian DM the Gaussian
//                      factorises, G_k(X,Y) = gx(X - x0_k) gy(Y - y0_k), and the
//                      actuators sit on an nA x nA grid, so
//                      OPD = Gy C Gx^T with C the coefficient grid (zeros at
//                      invalid actuators): scatter of the coefficients onto the
//                      grid, then the two small products with coalesced loads.
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
  // A workgroup takes four rows; its four wavefronts split K into four slices and each
  // wavefront accumulates the four rows over its slice (a 16-byte slice of x is loaded once
  // for the four rows, two slices per iteration in flight); the 4 x 4 partial sums are
  // combined through LDS in a fixed order.
  __shared__ float red[4][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = blockIdx.x * 4;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  const float* rows[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) rows[q] = M + (size_t)min(r0 + q, R - 1) * K;
  if ((K & 3) == 0) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const int K4 = K / 4;
    const int per = (K4 + 3) / 4;
    const int k0 = wave * per, k1 = min(K4, k0 + per);
    int k = k0 + lane;
    for (; k + 64 < k1; k += 128) {
      const float4 v0 = x4[k], v1 = x4[k + 64];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 m0 = reinterpret_cast<const float4*>(rows[q])[k];
        const float4 m1 = reinterpret_cast<const float4*>(rows[q])[k + 64];
        acc[q] = fmaf(m0.x, v0.x, fmaf(m0.y, v0.y, fmaf(m0.z, v0.z, fmaf(m0.w, v0.w, acc[q]))));
        acc[q] = fmaf(m1.x, v1.x, fmaf(m1.y, v1.y, fmaf(m1.z, v1.z, fmaf(m1.w, v1.w, acc[q]))));
      }
    }
    for (; k < k1; k += 64) {
      const float4 v = x4[k];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 m = reinterpret_cast<const float4*>(rows[q])[k];
        acc[q] = fmaf(m.x, v.x, fmaf(m.y, v.y, fmaf(m.z, v.z, fmaf(m.w, v.w, acc[q]))));
      }
    }
  } else {
    const int per = (K + 3) / 4;
    const int k0 = wave * per, k1 = min(K, k0 + per);
    for (int k = k0 + lane; k < k1; k += 64) {
      const float v = x[k];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = fmaf(rows[q][k], v, acc[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    float a = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    if (lane == 0) red[wave][q] = a;
  }
  __syncthreads();
  if (threadIdx.x < 4 && r0 + (int)threadIdx.x < R) {
    const int q = threadIdx.x;
    const float a = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    const int r = r0 + q;
    y[r] = alpha * a + (y0 ? beta * y0[r] : 0.f);
  }
}

// Batched separable OPD, for each batch j:
//   C[j][a][b]   = coefs[grid[a*nA+b]][j] (0 where there is no valid actuator)
//   U[j][y][b]   = sum_a gy[y][a] C[j][a][b]
//   opd[j][y][x] = sum_b U[j][y][b] gxT[b][x]
// coefs [nAct][B], grid [nA*nA], gxT [nA][n], gy [n][nA]; lanes run along the
// contiguous index of each product (b, then x) and eight independent partial
// sums keep loads in flight (the grids are small: few waves per SIMD).
__global__ void k_sep_scatter(const float* __restrict__ coefs, const int* __restrict__ grid, float* __restrict__ C,
                              int nA2, int B) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)B * nA2) return;
  const int cell = (int)(idx % nA2);
  const int j = (int)(idx / nA2);
  const int k = grid[cell];
  C[idx] = k >= 0 ? coefs[(size_t)k * B + j] : 0.f;
}

// out[r][c] = sum_k L[r][k] Rm[k][c] for an [R x K] x [K x Cc] product, one output per thread,
// c fastest (coalesced Rm loads, L broadcast within a row)
__global__ void __launch_bounds__(256) k_small_gemm(const float* __restrict__ L, const float* __restrict__ Rm,
                                                    float* __restrict__ out, int Rr, int K, int Cc,
                                                    size_t strideL, size_t strideR, size_t strideO) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (idx >= (size_t)Rr * Cc) return;
  const int c = (int)(idx % Cc);
  const int r = (int)(idx / Cc);
  const float* l = L + j * strideL + (size_t)r * K;
  const float* rm = Rm + j * strideR + c;
  float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int k = 0;
  for (; k + 8 <= K; k += 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = fmaf(l[k + q], rm[(size_t)(k + q) * Cc], a[q]);
  }
  for (; k < K; ++k) a[0] = fmaf(l[k], rm[(size_t)k * Cc], a[0]);
  out[j * strideO + idx] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
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

int oopao_dm_separable_opd(const float* coefs, const int* grid, const float* gxT, const float* gy, float* tmp,
                           float* opd, int n, int nA, int B, void* stream) {
  OOPAO_CHECK_ARG(coefs && grid && gxT && gy && tmp && opd, "null pointer argument");
  OOPAO_CHECK_ARG(n > 0 && nA > 0 && B > 0 && B <= 65535, "invalid sizes n=%d nA=%d B=%d", n, nA, B);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int nA2 = nA * nA;
  float* C = tmp;                       // [B][nA][nA]
  float* U = tmp + (size_t)B * nA2;     // [B][n][nA]
  const size_t t0 = (size_t)B * nA2;
  hipLaunchKernelGGL(k_sep_scatter, dim3((unsigned)((t0 + 255) / 256)), dim3(256), 0, st, coefs, grid, C, nA2, B);
  OOPAO_LAUNCH_CHECK();
  // U[j] = gy [n x nA] . C[j] [nA x nA]
  hipLaunchKernelGGL(k_small_gemm, dim3((unsigned)(((size_t)n * nA + 255) / 256), (unsigned)B), dim3(256), 0, st, gy,
                     C, U, n, nA, nA, (size_t)0, (size_t)nA2, (size_t)n * nA);
  OOPAO_LAUNCH_CHECK();
  // opd[j] = U[j] [n x nA] . gxT [nA x n]
  hipLaunchKernelGGL(k_small_gemm, dim3((unsigned)(((size_t)n * n + 255) / 256), (unsigned)B), dim3(256), 0, st, U,
                     gxT, opd, n, nA, n, (size_t)n * nA, (size_t)0, (size_t)n * n);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
