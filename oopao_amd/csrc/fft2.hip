// Batched 2-D complex FFT of N x N arrays, optionally centred, on CDNA4.
//
//     out = scale * S( FFT2_dir( S(in) ) ),   S = fftshift (= ifftshift for even N) or identity
//
// i.e. the fftshift(fft2(fftshift(x))) / N convention the reference's gain-sensing camera uses
// (GainSensingCamera.py:147-191). Four launches: a row pass (one wavefront per row, the
// runtime-length LDS Stockham transform of fft_generic.hpp, the input shift folded into the
// load), an LDS-tiled transpose, a second row pass, and a transpose that applies the output
// shift and the scale. Every global access is coalesced; any N with a plan (factors 2..13 and
// other primes, N <= 2048).
#include "common.hpp"
#include "fft_generic.hpp"

namespace oopao {

bool gen_plan_public(int N, GenPlan* p);

namespace {

constexpr int kTile = 32;

// rows: out[row][k] = sum_j in[src_row][(j + s) % N] W^{jk}, src_row = (row % N + s) % N within its array
template <int DIR>
__global__ void __launch_bounds__(64) k_fft2_rows(const float2* __restrict__ in, float2* __restrict__ out,
                                                  const float2* __restrict__ tw, GenPlan p, int shift) {
  extern __shared__ float2 lds2[];
  const int N = p.N;
  float2* a = lds2;
  float2* b = lds2 + N;
  const int lane = threadIdx.x;
  const size_t row = blockIdx.x;
  const size_t arr = row / N;
  const int r = (int)(row % N);
  const int rs = shift ? (r + N / 2) % N : r;
  const float2* src = in + (arr * N + rs) * N;
  for (int j = lane; j < N; j += 64) a[j] = src[shift ? (j + N / 2) % N : j];
  wave_lds_sync();
  float2* res = gen_fft<DIR>(a, b, tw, p, lane);
  float2* dst = out + row * N;
  for (int j = lane; j < N; j += 64) dst[j] = res[j];
}

// out[arr][(c + s) % N][(r + s) % N] = scale * in[arr][r][c]  (tiled through LDS)
__global__ void __launch_bounds__(kTile * 8) k_fft2_transpose(const float2* __restrict__ in,
                                                               float2* __restrict__ out, int N, int shift,
                                                               float scale) {
  __shared__ float2 tile[kTile][kTile + 1];
  const int arr = blockIdx.z;
  const int c0 = blockIdx.x * kTile, r0 = blockIdx.y * kTile;
  const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x / kTile;  // 32 x 8
  const float2* src = in + (size_t)arr * N * N;
  float2* dst = out + (size_t)arr * N * N;
  for (int y = ty; y < kTile; y += 8) {
    const int r = r0 + y, c = c0 + tx;
    if (r < N && c < N) tile[y][tx] = src[(size_t)r * N + c];
  }
  __syncthreads();
  const int s = shift ? N / 2 : 0;
  for (int y = ty; y < kTile; y += 8) {
    const int c = c0 + y, r = r0 + tx;  // output row = input column
    if (r < N && c < N) {
      const float2 v = tile[tx][y];
      dst[(size_t)((c + s) % N) * N + (r + s) % N] = make_float2(scale * v.x, scale * v.y);
    }
  }
}

}  // namespace

}  // namespace oopao

using namespace oopao;

extern "C" {

int oopao_fft2_supported(int N) {
  GenPlan p;
  return (N >= 2 && N <= 2048 && gen_plan_public(N, &p)) ? 1 : 0;
}

int oopao_fft2_bytes(int N, int B, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes, "null pointer argument");
  OOPAO_CHECK_ARG(N >= 2 && B > 0, "invalid sizes N=%d B=%d", N, B);
  *bytes = 2 * (size_t)B * N * N * sizeof(float2);
  return OK;
}

int oopao_fft2(const void* in, void* out, int N, int B, int dir, int centred, float scale, void* ws,
               size_t ws_bytes, void* stream) {
  OOPAO_CHECK_ARG(in && out && ws, "null pointer argument");
  OOPAO_CHECK_ARG(dir == 1 || dir == -1, "dir must be -1 (forward, exp(-2 pi i jk/N)) or +1 (inverse)");
  OOPAO_CHECK_ARG(!centred || (N % 2) == 0, "a centred transform needs an even N (got %d)", N);
  OOPAO_CHECK_ARG(B > 0 && B <= 65535, "invalid batch B=%d", B);
  GenPlan p;
  if (N < 2 || N > 2048 || !gen_plan_public(N, &p)) {
    set_error("no FFT plan for N=%d", N);
    return ERR_UNSUPPORTED;
  }
  const size_t need = 2 * (size_t)B * N * N * sizeof(float2);
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float2* A = reinterpret_cast<float2*>(ws);
  float2* T = A + (size_t)B * N * N;
  const size_t rows = (size_t)B * N;
  const size_t lds = 2 * (size_t)N * sizeof(float2);
  const dim3 tgrid((N + kTile - 1) / kTile, (N + kTile - 1) / kTile, B);
  if (dir < 0)
    hipLaunchKernelGGL(k_fft2_rows<-1>, dim3((unsigned)rows), dim3(64), lds, st,
                       reinterpret_cast<const float2*>(in), A, tw, p, centred);
  else
    hipLaunchKernelGGL(k_fft2_rows<1>, dim3((unsigned)rows), dim3(64), lds, st,
                       reinterpret_cast<const float2*>(in), A, tw, p, centred);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fft2_transpose, tgrid, dim3(kTile * 8), 0, st, A, T, N, 0, 1.0f);
  OOPAO_LAUNCH_CHECK();
  if (dir < 0)
    hipLaunchKernelGGL(k_fft2_rows<-1>, dim3((unsigned)rows), dim3(64), lds, st, T, A, tw, p, 0);
  else
    hipLaunchKernelGGL(k_fft2_rows<1>, dim3((unsigned)rows), dim3(64), lds, st, T, A, tw, p, 0);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fft2_transpose, tgrid, dim3(kTile * 8), 0, st, A, reinterpret_cast<float2*>(out), N,
                     centred, scale);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
