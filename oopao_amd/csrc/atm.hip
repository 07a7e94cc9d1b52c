// Frozen-flow atmosphere update on CDNA4 (Atmosphere.py:410-531, 589-622).
//
// Each layer keeps its extended screen mapShift [(res+2)^2] (float64) on the
// device. Per loop step the reference:
//  * when the accumulated sub-pixel shift crosses a pixel, shifts mapShift by
//    one pixel (cubic warp with an integer translation = exact copy) and
//    extrudes the new outer ring X = A Z + B xi from the inner ring Z
//    (Assemat et al. 2006), add_row, Atmosphere.py:410-421;
//  * shifts mapShift by the residual sub-pixel amount with Catmull-Rom cubic
//    convolution, constant (0) boundary, result clipped to [min, max] of
//    mapShift (scikit-image warp order 3), keeps [1:-1, 1:-1]
//    (Atmosphere.py:513-516);
//  * crops the pupil footprint and sums the layers scaled by
//    lambda_500 / 2 pi * sqrt(fractionalR0) (fill_OPD_support, 589-622).
// Kernels here implement those pieces; the random draws xi come from the
// layer's host RandomState (same stream as the reference).
#include "common.hpp"

namespace oopao {

// dst[r][c] = src[r - sy][c - sx] (0 outside), M x M maps, |sx|,|sy| <= 1 integers; in the same
// launch z[k] = dst[inner_idx[k]] (read from src through the same shift) and the min / max
// accumulator of the next k_minmax is reset.
__device__ __forceinline__ double shifted(const double* __restrict__ src, int M, int sx, int sy, int r, int c) {
  const int rs = r - sy, cs = c - sx;
  return (rs >= 0 && rs < M && cs >= 0 && cs < M) ? src[rs * M + cs] : 0.0;
}
__global__ void k_shift_gather(const double* __restrict__ src, double* __restrict__ dst, int M, int sx, int sy,
                               const int* __restrict__ inner_idx, int n_inner, double* __restrict__ z,
                               unsigned long long* __restrict__ mm) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < M * M) dst[idx] = shifted(src, M, sx, sy, idx / M, idx % M);
  if (idx < n_inner) {
    const int p = inner_idx[idx];
    z[idx] = shifted(src, M, sx, sy, p / M, p % M);
  }
  if (idx == 0) {
    mm[0] = ~0ull;
    mm[1] = 0ull;
    mm[2] = 0ull;
  }
}
// X[r] = sum_k A[r][k] z[k] + sum_k B[r][k] xi[k]  (float64, one wavefront per row)
// Partial dot product of one row with a vector: each lane reads double2 pairs, four
// independent partial sums keep several 16-byte loads in flight.
__device__ __forceinline__ double row_dot(const double* __restrict__ a, const double* __restrict__ v, int K,
                                          int lane) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const int K2 = K / 2;
  const double2* a2 = reinterpret_cast<const double2*>(a);
  const double2* v2 = reinterpret_cast<const double2*>(v);
  const bool aligned = ((reinterpret_cast<size_t>(a) | reinterpret_cast<size_t>(v)) & 15) == 0;
  int k = lane;
  if (aligned) {
    for (; k + 192 < K2; k += 256) {
      const double2 x0 = a2[k], x1 = a2[k + 64], x2 = a2[k + 128], x3 = a2[k + 192];
      const double2 y0 = v2[k], y1 = v2[k + 64], y2 = v2[k + 128], y3 = v2[k + 192];
      s0 = fma(x0.x, y0.x, fma(x0.y, y0.y, s0));
      s1 = fma(x1.x, y1.x, fma(x1.y, y1.y, s1));
      s2 = fma(x2.x, y2.x, fma(x2.y, y2.y, s2));
      s3 = fma(x3.x, y3.x, fma(x3.y, y3.y, s3));
    }
    for (; k < K2; k += 64) {
      const double2 x0 = a2[k], y0 = v2[k];
      s0 = fma(x0.x, y0.x, fma(x0.y, y0.y, s0));
    }
    if ((K & 1) && lane == 0) s1 = fma(a[K - 1], v[K - 1], s1);
  } else {
    for (k = lane; k < K; k += 64) s0 = fma(a[k], v[k], s0);
  }
  return (s0 + s1) + (s2 + s3);
}

// X[r] = A[r,:] z + B[r,:] xi (the outer-ring extrusion, float64), one wavefront per row; the
// row's value is also scattered into the shifted map at outer_idx[r].
__global__ void __launch_bounds__(256) k_extrude(const double* __restrict__ A, const double* __restrict__ z,
                                                 const double* __restrict__ Bm, const double* __restrict__ xi,
                                                 double* __restrict__ X, const int* __restrict__ outer_idx,
                                                 double* __restrict__ map, int R, int KA, int KB) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + wave;
  if (r >= R) return;
  double acc = row_dot(A + (size_t)r * KA, z, KA, lane) + row_dot(Bm + (size_t)r * KB, xi, KB, lane);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) {
    X[r] = acc;
    map[outer_idx[r]] = acc;
  }
}
// min / max of a map: per-block reductions merged with 64-bit atomic min / max on an
// order-preserving integer image of the doubles (the result does not depend on the order of the
// merges); the last block to finish (counted in mm[2]) converts the result back to doubles.
// mm[0..2] must hold (~0, 0, 0) at launch (k_minmax_init or k_shift_gather).
__device__ __forceinline__ unsigned long long dbl_ord(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dbl_unord(unsigned long long o) {
  const unsigned long long u = (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
  return __longlong_as_double((long long)u);
}
__global__ void k_minmax_init(unsigned long long* __restrict__ mm) {
  if (threadIdx.x == 0) {
    mm[0] = ~0ull;
    mm[1] = 0ull;
    mm[2] = 0ull;
  }
}
__global__ void __launch_bounds__(256) k_minmax(const double* __restrict__ map, int n,
                                                unsigned long long* __restrict__ mm) {
  __shared__ double lo[256], hi[256];
  __shared__ int last;
  double a = INFINITY, b = -INFINITY;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double v = map[i];
    a = fmin(a, v);
    b = fmax(b, v);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMin(&mm[0], dbl_ord(lo[0]));
    atomicMax(&mm[1], dbl_ord(hi[0]));
    __threadfence();
    last = atomicAdd(&mm[2], 1ull) == (unsigned long long)(gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const unsigned long long ma = atomicAdd(&mm[0], 0ull);  // device-coherent reads of the merged values
    const unsigned long long mb = atomicAdd(&mm[1], 0ull);
    double* md = reinterpret_cast<double*>(mm);
    md[0] = dbl_unord(ma);
    md[1] = dbl_unord(mb);
  }
}
static int launch_minmax(const double* map, int n, double* mm, hipStream_t st, bool init) {
  unsigned long long* mb = reinterpret_cast<unsigned long long*>(mm);
  if (init) {
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, st, mb);
    OOPAO_LAUNCH_CHECK();
  }
  const int blocks = (int)(((size_t)n + 1023) / 1024 < 256 ? ((size_t)n + 1023) / 1024 : 256);
  hipLaunchKernelGGL(k_minmax, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, map, n, mb);
  OOPAO_LAUNCH_CHECK();
  return OK;
}
__device__ __forceinline__ double cubic(double x, double f0, double f1, double f2, double f3) {
  return f1 + 0.5 * x * (f2 - f0 + x * (2.0 * f0 - 5.0 * f1 + 4.0 * f2 - f3 + x * (3.0 * (f1 - f2) + f3 - f0)));
}

// out[y][x] (+)= w * clip(cubic(map, r = y + r0 - ty, c = x + c0 - tx)), n x n output,
// map M x M, clip range mm[0..1] with the constant-mode cval (0) preserved when
// 0 lies outside [min, max].
__global__ void k_layer_opd(const double* __restrict__ map, int M, const double* __restrict__ mm, double tx,
                            double ty, int r0, int c0, double w, float* __restrict__ out, int n, int accumulate) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int y = idx / n, x = idx % n;
  const double rr = (double)(y + r0) - ty;
  const double cc = (double)(x + c0) - tx;
  const double fr = floor(rr), fc = floor(cc);
  const int ir = (int)fr, ic = (int)fc;
  const double xr = rr - fr, xc = cc - fc;
  double rows[4];
#pragma unroll
  for (int pr = 0; pr < 4; ++pr) {
    const int r = ir - 1 + pr;
    double f[4];
#pragma unroll
    for (int pc = 0; pc < 4; ++pc) {
      const int c = ic - 1 + pc;
      f[pc] = (r >= 0 && r < M && c >= 0 && c < M) ? map[r * M + c] : 0.0;
    }
    rows[pr] = cubic(xc, f[0], f[1], f[2], f[3]);
  }
  double v = cubic(xr, rows[0], rows[1], rows[2], rows[3]);
  const double lo = mm[0], hi = mm[1];
  const bool preserve = !(lo <= 0.0 && 0.0 <= hi);
  if (!(preserve && v == 0.0)) v = fmin(fmax(v, lo), hi);
  const float contrib = (float)(v * w);
  out[idx] = accumulate ? out[idx] + contrib : contrib;
}

}  // namespace oopao

using namespace oopao;

namespace oopao {

// out[r][c] (+)= weight * bilinear(in; y = a r + by, x = a c + bx) on an H x W float32 image,
// 0 outside (scikit-image warp order 1, mode constant, cval 0: neighbours floor/ceil of the
// coordinate, any neighbour outside the image contributes 0). Used for the off-axis footprint's
// sub-pixel translation and the laser-guide-star cone magnification (tools/
// interpolateGeometricalTransformation.py:152-265).
__global__ void k_bilinear_affine(const float* __restrict__ in, int H, int W, float* __restrict__ out, int h, int w,
                                  double a, double by, double bx, double weight, int accumulate) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= h * w) return;
  const int r = idx / w, c = idx % w;
  const double y = a * r + by, x = a * c + bx;
  const double fy = floor(y), fx = floor(x);
  const double cy = ceil(y), cx = ceil(x);
  const double dy = y - fy, dx = x - fx;
  auto px = [&](double rr, double cc) -> double {
    return (rr >= 0.0 && rr <= H - 1 && cc >= 0.0 && cc <= W - 1) ? (double)in[(size_t)rr * W + (size_t)cc] : 0.0;
  };
  const double top = (1.0 - dx) * px(fy, fx) + dx * px(fy, cx);
  const double bot = (1.0 - dx) * px(cy, fx) + dx * px(cy, cx);
  const double v = weight * ((1.0 - dy) * top + dy * bot);
  out[idx] = accumulate ? (float)((double)out[idx] + v) : (float)v;
}

}  // namespace oopao

extern "C" {

int oopao_atm_add_row(const double* map, double* tmp, int M, int sx, int sy, const int* inner_idx, int n_inner,
                      const int* outer_idx, int n_outer, const double* A, const double* Bm, const double* xi,
                      double* z, double* X, double* mm, void* stream) {
  OOPAO_CHECK_ARG(map && tmp && inner_idx && outer_idx && A && Bm && xi && z && X && mm, "null pointer argument");
  OOPAO_CHECK_ARG(M > 2 && n_inner > 0 && n_outer > 0, "invalid sizes");
  OOPAO_CHECK_ARG(sx >= -1 && sx <= 1 && sy >= -1 && sy <= 1, "integer shift must be -1, 0 or 1");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned long long* mb = reinterpret_cast<unsigned long long*>(mm);
  const int cover = M * M > n_inner ? M * M : n_inner;
  hipLaunchKernelGGL(k_shift_gather, dim3((cover + 255) / 256), dim3(256), 0, st, map, tmp, M, sx, sy, inner_idx,
                     n_inner, z, mb);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_extrude, dim3((n_outer + 3) / 4), dim3(256), 0, st, A, z, Bm, xi, X, outer_idx, tmp, n_outer,
                     n_inner, n_outer);
  OOPAO_LAUNCH_CHECK();
  return launch_minmax(tmp, M * M, mm, st, false);
}

int oopao_atm_minmax(const double* map, int count, double* mm, void* stream) {
  OOPAO_CHECK_ARG(map && mm && count > 0, "invalid arguments");
  return launch_minmax(map, count, mm, reinterpret_cast<hipStream_t>(stream), true);
}

int oopao_atm_layer_opd(const double* map, int M, const double* mm, double tx, double ty, int r0, int c0, double w,
                        float* out, int n, int accumulate, void* stream) {
  OOPAO_CHECK_ARG(map && mm && out, "null pointer argument");
  OOPAO_CHECK_ARG(n > 0 && M > 0, "invalid sizes");
  hipLaunchKernelGGL(k_layer_opd, dim3((n * n + 255) / 256), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     map, M, mm, tx, ty, r0, c0, w, out, n, accumulate);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_bilinear_affine(const float* in, int H, int W, float* out, int h, int w, double a, double by, double bx,
                          double weight, int accumulate, void* stream) {
  OOPAO_CHECK_ARG(in && out, "null pointer argument");
  OOPAO_CHECK_ARG(H > 0 && W > 0 && h > 0 && w > 0 && (long long)h * w < (1LL << 31), "invalid sizes");
  hipLaunchKernelGGL(k_bilinear_affine, dim3((unsigned)((h * w + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), in, H, W, out, h, w, a, by, bx, weight, accumulate);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
