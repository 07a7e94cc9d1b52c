// Boiling of a frozen-flow layer on CDNA4 (Atmosphere.apply_boiling, Atmosphere.py:387-408).
//
// Per loop step, before the frozen-flow shift, a layer with a boiling time t_b mixes its extended screen
// with a fresh independent screen of the same support: mapShift = alpha mapShift + beta phi with
// alpha = exp(-dt / t_b), beta = sqrt(1 - alpha^2). phi is ft_sh_phase_screen (phaseStats.py:201-271) on
// the M x M support, M = res + 2, in float64:
//   phi_hi = Re fftshift(fft2(fftshift(cn))), cn = (g_re + i g_im) sqrt(PSD) df  (ft_phase_screen,
//            phaseStats.py:153-198), g the host RandomState(boiling_seed) draws in the reference's order;
//   phi_lo = Re sum_t c_t exp(2 pi i (fx_t x + fy_t y)) - its mean over the support (the 3 levels of
//            2 x 2 subharmonics, whose coefficients the host forms from the same draws).
// The centred 2-D transform is W C W^T with W = fftshift . DFT_M . fftshift (a complex128 M x M matrix
// built once per support size): two complex GEMM passes here (M = 326 at an 80 x 80 subaperture, 320 px
// pupil; any M, including the prime factors the layer sizes produce), then the subharmonics, the mean
// and the mix in one pass each.
#include "common.hpp"

namespace oopao {

struct z128 {
  double x, y;
};
__device__ __forceinline__ z128 zfma(z128 a, z128 b, z128 c) {   // a * b + c
  return {fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}

// C[y][x] = (g[y M + x] + i g[M^2 + y M + x]) * amp[y M + x]
__global__ void k_boil_coef(const double* __restrict__ g, const double* __restrict__ amp, z128* __restrict__ C,
                            int M2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M2) return;
  const double a = amp[i];
  C[i] = {g[i] * a, g[M2 + i] * a};
}

// out = A B (complex, M x M x M); with TRANS_B the right operand is B^T, with REAL_OUT only the real part
// is written (as doubles). 16 x 16 output tile per 256-thread workgroup, K staged through LDS in steps of 16.
template <bool TRANS_B, bool REAL_OUT>
__global__ void __launch_bounds__(256) k_zgemm(const z128* __restrict__ A, const z128* __restrict__ B,
                                               void* __restrict__ out, int M) {
  __shared__ z128 As[16][17];
  __shared__ z128 Bs[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  z128 acc{0.0, 0.0};
  for (int k0 = 0; k0 < M; k0 += 16) {
    const int ka = k0 + tx, kb = k0 + ty;
    As[ty][tx] = (row < M && ka < M) ? A[(size_t)row * M + ka] : z128{0.0, 0.0};
    if (TRANS_B) {
      // Bs[k][c] = B[c][k]: read rows of B (coalesced along k), store transposed
      const int brow = blockIdx.x * 16 + ty, bk = k0 + tx;
      Bs[tx][ty] = (brow < M && bk < M) ? B[(size_t)brow * M + bk] : z128{0.0, 0.0};
    } else {
      Bs[ty][tx] = (kb < M && col < M) ? B[(size_t)kb * M + col] : z128{0.0, 0.0};
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc = zfma(As[ty][k], Bs[k][tx], acc);
    __syncthreads();
  }
  if (row < M && col < M) {
    if (REAL_OUT)
      reinterpret_cast<double*>(out)[(size_t)row * M + col] = acc.x;
    else
      reinterpret_cast<z128*>(out)[(size_t)row * M + col] = acc;
  }
}

// lo[y][x] = Re sum_t c_t exp(2 pi i (fx_t xc[x] + fy_t xc[y])), sub[t] = (Re c_t, Im c_t, fx_t, fy_t);
// part[block] = the block's sum of lo (fixed-order tree)
__global__ void __launch_bounds__(256) k_boil_lo(const double* __restrict__ sub, int nsub,
                                                 const double* __restrict__ xc, int M, double* __restrict__ lo,
                                                 double* __restrict__ part) {
  __shared__ double red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (i < M * M) {
    const double x = xc[i % M], y = xc[i / M];
    for (int t = 0; t < nsub; ++t) {
      const double* s = sub + 4 * t;
      double sn, cs;
      sincospi(2.0 * (s[2] * x + s[3] * y), &sn, &cs);
      v = fma(s[0], cs, fma(-s[1], sn, v));
    }
    lo[i] = v;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// map = alpha map + beta (hi + lo - mean(lo)); every workgroup sums the nblk partials in the same order
__global__ void __launch_bounds__(256) k_boil_mix(double* __restrict__ map, const double* __restrict__ hi,
                                                  const double* __restrict__ lo, const double* __restrict__ part,
                                                  int nblk, int M2, double alpha, double beta) {
  __shared__ double red[256];
  double s = 0.0;
  if (lo != nullptr)
    for (int k = threadIdx.x; k < nblk; k += 256) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  const double mean = red[0] / (double)M2;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M2) return;
  const double phi = lo != nullptr ? hi[i] + (lo[i] - mean) : hi[i];
  map[i] = fma(alpha, map[i], beta * phi);
}

// sub[t] = (g[re_t] a_t, g[im_t] a_t, fx_t, fy_t) from table[t] = (re_t, im_t, a_t, fx_t, fy_t): the subharmonic
// coefficients from draws already on the device (the device RNG mode)
__global__ void k_boil_sub(const double* __restrict__ g, const double* __restrict__ table, int nsub,
                           double* __restrict__ sub) {
  const int t = threadIdx.x;
  if (t >= nsub) return;
  const double* e = table + 5 * t;
  const double a = e[2];
  sub[4 * t] = g[(int)e[0]] * a;
  sub[4 * t + 1] = g[(int)e[1]] * a;
  sub[4 * t + 2] = e[3];
  sub[4 * t + 3] = e[4];
}

static size_t al256b(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace oopao

using namespace oopao;

extern "C" {

int oopao_atm_boil_bytes(int M, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr && M > 0 && M <= 8192, "invalid arguments M=%d", M);
  const size_t M2 = (size_t)M * M;
  *bytes = 2 * al256b(M2 * sizeof(z128)) + al256b((M2 + 255) / 256 * sizeof(double));
  return OK;
}

int oopao_atm_boil(double* map, int M, const double* g, const double* amp, const void* W, const double* sub, int nsub,
                   const double* xc, double alpha, double beta, void* ws, size_t ws_bytes, void* stream) {
  OOPAO_CHECK_ARG(map && g && amp && W && ws, "null pointer argument");
  OOPAO_CHECK_ARG(M > 0 && M <= 8192 && nsub >= 0 && (nsub == 0 || (sub && xc)), "invalid arguments M=%d nsub=%d",
                  M, nsub);
  size_t need = 0;
  oopao_atm_boil_bytes(M, &need);
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int M2 = M * M;
  const int nblk = (M2 + 255) / 256;
  char* w = reinterpret_cast<char*>(ws);
  z128* C = reinterpret_cast<z128*>(w);
  z128* T = reinterpret_cast<z128*>(w + al256b((size_t)M2 * sizeof(z128)));
  double* part = reinterpret_cast<double*>(w + 2 * al256b((size_t)M2 * sizeof(z128)));
  const z128* Wz = reinterpret_cast<const z128*>(W);
  hipLaunchKernelGGL(k_boil_coef, dim3(nblk), dim3(256), 0, st, g, amp, C, M2);
  OOPAO_LAUNCH_CHECK();
  const dim3 tiles((M + 15) / 16, (M + 15) / 16);
  // T = W C; hi = Re(T W^T), written over C's memory (C is no longer needed)
  hipLaunchKernelGGL((k_zgemm<false, false>), tiles, dim3(256), 0, st, Wz, C, (void*)T, M);
  OOPAO_LAUNCH_CHECK();
  double* hi = reinterpret_cast<double*>(C);
  hipLaunchKernelGGL((k_zgemm<true, true>), tiles, dim3(256), 0, st, T, Wz, (void*)hi, M);
  OOPAO_LAUNCH_CHECK();
  double* lo = nullptr;
  if (nsub > 0) {
    lo = reinterpret_cast<double*>(T);      // T is no longer needed either
    hipLaunchKernelGGL(k_boil_lo, dim3(nblk), dim3(256), 0, st, sub, nsub, xc, M, lo, part);
    OOPAO_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_boil_mix, dim3(nblk), dim3(256), 0, st, map, hi, lo, part, nblk, M2, alpha, beta);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_atm_boil_subharmonics(const double* g, long long ng, const double* table, int nsub, double* sub,
                                void* stream) {
  OOPAO_CHECK_ARG(g && table && sub && nsub > 0 && nsub <= 256 && ng >= 54, "invalid arguments nsub=%d ng=%lld",
                  nsub, ng);
  hipLaunchKernelGGL(k_boil_sub, dim3(1), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), g, table, nsub,
                     sub);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
