// Gain-sensing camera (optical gains of the modulated PWFS from the modulation-camera
// frame, convolutional model) on CDNA4.
//
// Reference: GainSensingCamera.py:78-199 (Chambouleyron et al. 2021). With
// F(x) = fftshift(fft2(fftshift(x))) / N and F'(x) = fftshift(ifft2(fftshift(x))) * N:
//   calibration:    bp[:, :, m] = F(B_m) F'(B_m)   (B_m = mode m zero-padded to N x N)
//                   IR_c = 2 Im(conj(F(mask)) F(mask frame_c / sum frame_c))
//                   S_c = (F(IR_c) F'(IR_c)).flatten() @ bp.reshape(N^2, nModes)
//   optical gains:  IR = ... of the sky frame, S = (F(IR) F'(IR_c)) @ bp, og = Re(S / S_c)
// For a real mode B, F'(B) = conj(F(B)), so bp = |F(B_m)|^2 is real: it is stored as float32
// [N^2][nModes] (frequency-major, the reference's reshape order) and S is one streaming pass
// over it with the complex weights w = F(IR) F'(IR_c).
//
// Kernels: frame sum (double, two phases), mask x frame / sum, impulse response, complex
// weights, padded basis, |X|^2 stored transposed through an LDS tile, and the sensitivity as
// per-chunk float partial sums over 2048 frequencies (one mode per thread, coalesced rows of
// bp) reduced in double in a fixed order.
#include "common.hpp"

namespace oopao {

namespace {

constexpr int kSumBlocks = 256;
constexpr int kSensChunk = 2048;

__global__ void __launch_bounds__(256) k_sum_partial(const float* __restrict__ f, size_t count,
                                                     double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) s += f[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_final(const double* __restrict__ part, int n, double* __restrict__ sum) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sum[0] = red[0];
}

__global__ void k_mask_frame(const float2* __restrict__ mask, const float* __restrict__ frame,
                             const double* __restrict__ sum, float2* __restrict__ out, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float inv = (float)(1.0 / sum[0]);
  const float v = frame[i] * inv;
  out[i] = make_float2(mask[i].x * v, mask[i].y * v);
}

// ir = 2 Im(conj(a) b), stored as a complex array with zero imaginary part
__global__ void k_gsc_ir(const float2* __restrict__ a, const float2* __restrict__ b, float2* __restrict__ ir,
                         size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  ir[i] = make_float2(2.f * (a[i].x * b[i].y - a[i].y * b[i].x), 0.f);
}

__global__ void k_cmul(const float2* __restrict__ a, const float2* __restrict__ b, float2* __restrict__ w,
                       size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  w[i] = make_float2(a[i].x * b[i].x - a[i].y * b[i].y, a[i].x * b[i].y + a[i].y * b[i].x);
}

// out[m][y + pad][x + pad] = basis[m][y][x] (real part), zero elsewhere
__global__ void k_pad_basis(const float* __restrict__ basis, int n, int N, int mb, float2* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)mb * N * N;
  if (i >= total) return;
  const int m = (int)(i / ((size_t)N * N));
  const int r = (int)((i / N) % N), c = (int)(i % N);
  const int pad = (N - n) / 2;
  const int y = r - pad, x = c - pad;
  const float v = (y >= 0 && y < n && x >= 0 && x < n) ? basis[((size_t)m * n + y) * n + x] : 0.f;
  out[i] = make_float2(v, 0.f);
}

// bp[f * M + m0 + m] = |X[m][f]|^2 for m < mb, through a 32 x 32 LDS tile (coalesced both ways)
__global__ void __launch_bounds__(256) k_bp_store(const float2* __restrict__ X, int mb, size_t F,
                                                  float* __restrict__ bp, int M, int m0) {
  __shared__ float tile[32][33];
  const size_t f0 = (size_t)blockIdx.x * 32;
  const int mm0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int m = mm0 + y;
    const size_t f = f0 + tx;
    float v = 0.f;
    if (m < mb && f < F) {
      const float2 z = X[(size_t)m * F + f];
      v = z.x * z.x + z.y * z.y;
    }
    tile[y][tx] = v;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const size_t f = f0 + y;
    const int m = mm0 + tx;
    if (m < mb && f < F) bp[f * M + m0 + m] = tile[tx][y];
  }
}

// part[c][m] = sum over the chunk's frequencies f of bp[f][m] w[f]
__global__ void __launch_bounds__(256) k_sens_partial(const float* __restrict__ bp, const float2* __restrict__ w,
                                                      size_t F, int M, double2* __restrict__ part) {
  __shared__ float2 ws[kSensChunk];
  const size_t f0 = (size_t)blockIdx.x * kSensChunk;
  const int nf = (int)min((size_t)kSensChunk, F - f0);
  for (int j = threadIdx.x; j < nf; j += 256) ws[j] = w[f0 + j];
  __syncthreads();
  const int m = blockIdx.y * 256 + threadIdx.x;
  if (m >= M) return;
  float re0 = 0.f, im0 = 0.f, re1 = 0.f, im1 = 0.f;
  const float* col = bp + f0 * M + m;
  int j = 0;
  for (; j + 1 < nf; j += 2) {
    const float b0 = col[(size_t)j * M], b1 = col[(size_t)(j + 1) * M];
    re0 = fmaf(b0, ws[j].x, re0);
    im0 = fmaf(b0, ws[j].y, im0);
    re1 = fmaf(b1, ws[j + 1].x, re1);
    im1 = fmaf(b1, ws[j + 1].y, im1);
  }
  if (j < nf) {
    const float b0 = col[(size_t)j * M];
    re0 = fmaf(b0, ws[j].x, re0);
    im0 = fmaf(b0, ws[j].y, im0);
  }
  part[(size_t)blockIdx.x * M + m] = make_double2((double)re0 + (double)re1, (double)im0 + (double)im1);
}

__global__ void __launch_bounds__(256) k_sens_final(const double2* __restrict__ part, int nchunks, int M,
                                                    double* __restrict__ S) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= M) return;
  double re = 0.0, im = 0.0;
  for (int c = 0; c < nchunks; ++c) {
    const double2 v = part[(size_t)c * M + m];
    re += v.x;
    im += v.y;
  }
  S[2 * m] = re;
  S[2 * m + 1] = im;
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

unsigned blocks_for(size_t count) { return (unsigned)((count + 255) / 256); }

}  // namespace

}  // namespace oopao

using namespace oopao;

extern "C" int oopao_fft2(const void* in, void* out, int N, int B, int dir, int centred, float scale, void* ws,
                          size_t ws_bytes, void* stream);

extern "C" {

int oopao_gsc_workspace_bytes(int N, int M, int mb, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes, "null pointer argument");
  OOPAO_CHECK_ARG(N >= 2 && M > 0 && mb > 0, "invalid sizes N=%d M=%d mb=%d", N, M, mb);
  const size_t F = (size_t)N * N;
  const size_t nch = (F + kSensChunk - 1) / kSensChunk;
  // impulse response / sensitivity: 2 F float2 + FFT workspace (2 F float2) + sums + partials
  const size_t a = al256(2 * F * sizeof(float2)) + al256(2 * F * sizeof(float2)) +
                   al256((kSumBlocks + 1) * sizeof(double)) + al256(nch * (size_t)M * sizeof(double2));
  // basis product: 2 mb F float2 + FFT workspace (2 mb F float2)
  const size_t b = al256(2 * (size_t)mb * F * sizeof(float2)) + al256(2 * (size_t)mb * F * sizeof(float2));
  *bytes = a > b ? a : b;
  return OK;
}

int oopao_gsc_impulse_response(const void* fmask, const void* mask, const float* frame, int N, void* ir, void* ws,
                               size_t ws_bytes, void* stream) {
  OOPAO_CHECK_ARG(fmask && mask && frame && ir && ws, "null pointer argument");
  OOPAO_CHECK_ARG(N >= 2 && (N % 2) == 0, "invalid N=%d (even N required)", N);
  const size_t F = (size_t)N * N;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* p = reinterpret_cast<char*>(ws);
  float2* mf = reinterpret_cast<float2*>(p);
  float2* fmf = mf + F;
  p += al256(2 * F * sizeof(float2));
  void* fws = p;
  p += al256(2 * F * sizeof(float2));
  double* sums = reinterpret_cast<double*>(p);
  p += al256((kSumBlocks + 1) * sizeof(double));
  if ((size_t)(p - reinterpret_cast<char*>(ws)) > ws_bytes) {
    set_error("workspace too small");
    return ERR_WORKSPACE;
  }
  hipLaunchKernelGGL(k_sum_partial, dim3(kSumBlocks), dim3(256), 0, st, frame, F, sums + 1);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, sums + 1, kSumBlocks, sums);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mask_frame, dim3(blocks_for(F)), dim3(256), 0, st, reinterpret_cast<const float2*>(mask),
                     frame, sums, mf, F);
  OOPAO_LAUNCH_CHECK();
  int rc = oopao_fft2(mf, fmf, N, 1, -1, 1, 1.0f / N, fws, 2 * F * sizeof(float2), stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_gsc_ir, dim3(blocks_for(F)), dim3(256), 0, st, reinterpret_cast<const float2*>(fmask), fmf,
                     reinterpret_cast<float2*>(ir), F);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_gsc_basis_product(const float* basis, int n, int N, int M, int m0, int mb, float* bp, void* ws,
                            size_t ws_bytes, void* stream) {
  OOPAO_CHECK_ARG(basis && bp && ws, "null pointer argument");
  OOPAO_CHECK_ARG(n > 0 && N >= n && ((N - n) % 2) == 0 && (N % 2) == 0, "invalid sizes n=%d N=%d", n, N);
  OOPAO_CHECK_ARG(mb > 0 && mb <= 65535 && m0 >= 0 && m0 + mb <= M, "invalid mode range m0=%d mb=%d M=%d", m0, mb,
                  M);
  const size_t F = (size_t)N * N;
  const size_t need = al256(2 * (size_t)mb * F * sizeof(float2)) + al256(2 * (size_t)mb * F * sizeof(float2));
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float2* padded = reinterpret_cast<float2*>(ws);
  float2* X = padded + (size_t)mb * F;
  void* fws = reinterpret_cast<char*>(ws) + al256(2 * (size_t)mb * F * sizeof(float2));
  hipLaunchKernelGGL(k_pad_basis, dim3(blocks_for((size_t)mb * F)), dim3(256), 0, st,
                     basis + (size_t)m0 * n * n, n, N, mb, padded);
  OOPAO_LAUNCH_CHECK();
  int rc = oopao_fft2(padded, X, N, mb, -1, 1, 1.0f / N, fws, 2 * (size_t)mb * F * sizeof(float2), stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bp_store, dim3((unsigned)((F + 31) / 32), (unsigned)((mb + 31) / 32)), dim3(256), 0, st, X,
                     mb, F, bp, M, m0);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_gsc_sensitivity(const void* ir1, const void* v2, const float* bp, int N, int M, double* S, void* ws,
                          size_t ws_bytes, void* stream) {
  OOPAO_CHECK_ARG(ir1 && v2 && bp && S && ws, "null pointer argument");
  OOPAO_CHECK_ARG(N >= 2 && (N % 2) == 0 && M > 0 && M <= 65535 * 256, "invalid sizes N=%d M=%d", N, M);
  const size_t F = (size_t)N * N;
  const size_t nch = (F + kSensChunk - 1) / kSensChunk;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* p = reinterpret_cast<char*>(ws);
  float2* fir = reinterpret_cast<float2*>(p);
  float2* w = fir + F;
  p += al256(2 * F * sizeof(float2));
  void* fws = p;
  p += al256(2 * F * sizeof(float2));
  p += al256((kSumBlocks + 1) * sizeof(double));
  double2* part = reinterpret_cast<double2*>(p);
  p += al256(nch * (size_t)M * sizeof(double2));
  if ((size_t)(p - reinterpret_cast<char*>(ws)) > ws_bytes) {
    set_error("workspace too small");
    return ERR_WORKSPACE;
  }
  int rc = oopao_fft2(ir1, fir, N, 1, -1, 1, 1.0f / N, fws, 2 * F * sizeof(float2), stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cmul, dim3(blocks_for(F)), dim3(256), 0, st, fir, reinterpret_cast<const float2*>(v2), w, F);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sens_partial, dim3((unsigned)nch, (unsigned)((M + 255) / 256)), dim3(256), 0, st, bp, w, F, M,
                     part);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sens_final, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, part, (int)nch, M, S);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
