// Adjoint (vector-Jacobian product) of the modulated PWFS frame, fast path.
//
// Forward (oopao_pwfs_frames): per point i, E_i = M u_i with u_i the pupil
// field times the point's tip/tilt phasors and M = K_y (x) K_x, K = IFFT_N .
// diag(m') . FFT_N applied to the zero-padded pupil (the offset and the 1/N
// are folded into m'); frame = bin4x4(sum_i |E_i|^2).
// Given G = dL/dframe, g(Y, X) = G(Y/4, X/4) is dL/dI_i for every point, and
//   dL/dphi(y, x) = sum_i 2 Im( w_i(y, x) conj(u_i(y, x)) ),  w_i = M^H (g . E_i),
// with M^H = K_y^H (x) K_x^H, K^H = crop_n . IFFT_N . diag(conj m') . FFT_N
// (the same transforms, conjugated mask, outputs kept on the pupil support).
// No forward state is stored: the backward recomputes each point's field
// (rows pass), then
//   k_bwd_cols : per (point, column X): E column = K_y(column), times g,
//                then K_y^H, keeping the n pupil rows -> Z_i [n][N] (tiled)
//   k_bwd_rows : per (point chunk, pupil row y): K_x^H of Z_i's row, keeping
//                the n pupil columns, 2 Im(w conj(u)) summed over the chunk
//   k_adj_sum  : fixed-order sum over chunks, times d(phi)/d(opd).
// With column bands (a cotangent that is zero outside some detector columns, as for any loss of the
// slopesMaps signal, which reads only the quadrant columns) k_bwd_cols transforms the band columns
// only and k_bwd_rows reads the other columns of Z as zero: those E columns meet a zero g and add
// nothing to w.
#include "common.hpp"
#include "fft_core.hpp"
#include "pwfs_fast.hpp"

extern "C" int oopao_pwfs_rows(const float* opd, float scale, const float* amp, const void* tx, int P,
                               const int* pstart, const void* ty, const void* mx, int n, int N, int T, int B,
                               int pts_per_task, int lanes, const int* bands, void* inter,
                               size_t inter_bytes, void* stream);

namespace oopao {

// Plans whose backward kernels copy the twiddle prefix into LDS (as the forward passes do): those of the
// forward columns / rows passes, less the ones where the copy would lower the occupancy of the backward
// kernel (columns N = 1536, 3 -> 2 waves per SIMD; N = 1920, more spills; rows N = 512, 5 -> 4). At
// N = 320 the columns kernel drops from 4 to 3 waves per SIMD with the copy and still runs faster
// (config 4, 64 screens: adjoint 1.82 -> 1.61 ms, 1.76 ms with the copy in the rows kernel only).
// Adjoint of 8 screens at N = 256 / 512: 170 -> 130 us / 726 -> 613 us; 4 at N = 1024: 2.61 -> 1.87 ms.
__host__ __device__ constexpr bool tw_lds_bwd_cols(int N, int TPT) {
  return tw_lds_cols(N, TPT) && N != 1536 && N != 1920;
}
__host__ __device__ constexpr bool tw_lds_bwd_rows(int N) { return tw_lds_rows(N, false) && N != 512; }

// Z layout: [B*T][N/4][n][4] complex64 (the forward intermediate's tiling).
template <int N, int TPT>
__global__ void __launch_bounds__(64 * kWaves)
k_bwd_cols(const float2* __restrict__ inter, const float2* __restrict__ my, const float* __restrict__ gframe,
           const float2* __restrict__ tw, float2* __restrict__ Z, int B, int T, Bands bd) {
  constexpr int n = N / 4;
  constexpr int E = N / TPT;
  constexpr int G = 64 / TPT;
  constexpr int NG = N / 4;
  constexpr bool kWaveCols = (TPT == 64);
  constexpr bool kTwLds = tw_lds_bwd_cols(N, TPT);
  constexpr int kTw = kTwLds ? sep_tw_len<N, TPT>() : 1;
  static_assert(!kTwLds || kTw <= N, "twiddle prefix within the table");
  __shared__ float2 lds[kWaves * G][lds_len(N)];
  __shared__ float2 twl[kTw];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane / TPT;
  const int t = lane % TPT;
  const long task = kWaveCols ? (long)blockIdx.x : (long)blockIdx.x * kWaves + wave;
  const int ng = (bd.a1 - bd.a0) + (bd.b1 - bd.b0);   // computed detector columns
  const int g = band_col((int)(task % ng), bd);
  const long bi = task / ng;  // screen * T + point
  // the twiddle prefix into LDS, with every wavefront of the workgroup present (the last workgroup's
  // wavefronts past the tasks leave below)
  if constexpr (kTwLds) {
    copy_twiddles<kTw, 64 * kWaves>(twl, tw);
    __syncthreads();
  }
  const float2* twp = kTwLds ? static_cast<const float2*>(twl) : tw;
  if (bi >= (long)B * T) return;  // uniform over the wavefront
  const int b = (int)(bi / T);
  float2* xl = lds[wave * G + h];
  const float2* tile = inter + ((size_t)bi * NG + g) * (size_t)n * 4;
  float2* zt = Z + ((size_t)bi * NG + g) * (size_t)n * 4;
  const float* gf = gframe + (size_t)b * NG * NG;
  // With kBuf every per-element access goes through a buffer instruction: a lane offset in one VGPR and
  // a compile-time uniform offset per element (TPT is a multiple of 4, so row (t + TPT m) / 4 is
  // t / 4 + m TPT / 4), instead of a 64-bit address per element.
  constexpr bool kBuf = buf_bwd_cols(N);
  const BufferRsrc col = buffer_rsrc(tile, (unsigned)n * 4 * sizeof(float2));
  const BufferRsrc zr = buffer_rsrc(zt, (unsigned)n * 4 * sizeof(float2));
  const BufferRsrc msk = buffer_rsrc(my, (unsigned)N * sizeof(float2));
  const BufferRsrc gr = buffer_rsrc(gf, (unsigned)(NG * NG) * sizeof(float));
  const unsigned msk_lane = (unsigned)t * sizeof(float2);
  const unsigned g_lane = (unsigned)(((t >> 2) * NG + g) * (int)sizeof(float));
#pragma unroll 1
  for (int c0 = 0; c0 < (kWaveCols ? 1 : 4); c0 += G) {
    const int c = kWaveCols ? wave : c0 + h;
    const unsigned col_lane = (unsigned)(((t >> 2) * 4 + c) * (int)sizeof(float2));
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = kBuf ? buffer_load_c64(col, col_lane, m * TPT * (int)sizeof(float2))
                  : tile[((t + TPT * m) >> 2) * 4 + c];
    fft_fwd_pruned<N, TPT, 4, true>(v, xl, twp, t);
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = cmul(v[m], kBuf ? buffer_load_c64(msk, msk_lane, m * TPT * (int)sizeof(float2))
                             : launder(my)[t + TPT * m]);
    fft_full<N, TPT, +1, true>(v, xl, twp, t);
    // E column at rows Y = t + TPT m; dL/dI there is the detector gradient of pixel (Y/4, X/4)
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = cscale(v[m], kBuf ? buffer_load_f32(gr, g_lane, m * (TPT / 4) * NG * (int)sizeof(float))
                               : gf[((t + TPT * m) >> 2) * NG + g]);
    fft_full<N, TPT, -1, true>(v, xl, twp, t);
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = cmulc(v[m], kBuf ? buffer_load_c64(msk, msk_lane, m * TPT * (int)sizeof(float2))
                              : launder(my)[t + TPT * m]);
    fft_full<N, TPT, +1, true>(v, xl, twp, t);
    // keep the pupil rows y < n (elements m < E/4)
#pragma unroll
    for (int m = 0; m < E / 4; ++m) {
      if constexpr (kBuf)
        buffer_store_c64(v[m], zr, (unsigned)((t * 4 + c) * (int)sizeof(float2)), m * TPT * 4 * (int)sizeof(float2));
      else
        zt[(t + TPT * m) * 4 + c] = v[m];
    }
  }
}

// gpart[b][chunk][y][x] = 2 scale sum_{i in chunk} Im(w_i(y,x) conj(u_i(y,x)))
template <int N, int TPT>
__global__ void __launch_bounds__(64 * kWaves)
k_bwd_rows(const float2* __restrict__ Z, const float2* __restrict__ mx, const float2* __restrict__ field,
           const float2* __restrict__ txs, const float2* __restrict__ tys, const float2* __restrict__ tw,
           float* __restrict__ gpart, int B, int T, int nchunk, float scale2, Bands bd) {
  constexpr int n = N / 4;
  constexpr int E = N / TPT;
  constexpr int G = 64 / TPT;
  constexpr int NG = N / 4;
  constexpr int RPB = kWaves * G;
  constexpr int nrb = (n + RPB - 1) / RPB;
  constexpr bool kTwLds = tw_lds_bwd_rows(N);
  constexpr int kTw = kTwLds ? sep_tw_len<N, TPT>() : 1;
  static_assert(!kTwLds || kTw <= N, "twiddle prefix within the table");
  __shared__ float2 lds[kWaves * G][lds_len(N)];
  __shared__ float2 twl[kTw];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane / TPT;
  const int t = lane % TPT;
  int bid = blockIdx.x;
  const int rb = bid % nrb;
  bid /= nrb;
  const int chunk = bid % nchunk;
  const int b = bid / nchunk;
  if (b >= B) return;  // uniform over the workgroup
  if constexpr (kTwLds) {
    copy_twiddles<kTw, 64 * kWaves>(twl, tw);
    __syncthreads();
  }
  const float2* twp = kTwLds ? static_cast<const float2*>(twl) : tw;
  const int y_raw = rb * RPB + wave * G + h;
  const bool active = y_raw < n;
  const int y = active ? y_raw : n - 1;
  const int i0 = (int)(((long)chunk * T) / nchunk);
  const int i1 = (int)(((long)(chunk + 1) * T) / nchunk);
  float2* xl = lds[wave * G + h];
  const float2* fr = field + ((size_t)b * n + y) * n;
  float acc[E / 4];
#pragma unroll
  for (int m = 0; m < E / 4; ++m) acc[m] = 0.f;
  for (int i = i0; i < i1; ++i) {
    const float2* zt = Z + ((size_t)b * T + i) * (size_t)NG * n * 4;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      // columns outside the bands were not computed: their w contribution is zero
      const int X = t + TPT * m;
      const int cm = (TPT / 4) * m, tq = t >> 2;   // X >> 2 = t / 4 + cm, compared on scalar-shifted limits
      const bool in = (tq >= bd.a0 - cm && tq < bd.a1 - cm) || (tq >= bd.b0 - cm && tq < bd.b1 - cm);
      v[m] = in ? zt[((size_t)(X >> 2) * n + y) * 4 + (X & 3)] : make_float2(0.f, 0.f);
    }
    fft_full<N, TPT, -1>(v, xl, twp, t);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmulc(v[m], launder(mx)[t + TPT * m]);
    fft_full<N, TPT, +1>(v, xl, twp, t);
    const float2 tyv = tys[(size_t)i * n + y];
    const float2* txi = launder(txs) + (size_t)i * n;
#pragma unroll
    for (int m = 0; m < E / 4; ++m) {
      const int x = t + TPT * m;
      const float2 u = cmul(cmul(fr[x], txi[x]), tyv);
      acc[m] += v[m].y * u.x - v[m].x * u.y;  // Im(w conj(u))
    }
  }
  if (!active) return;
  float* dst = gpart + (((size_t)b * nchunk + chunk) * n + y) * n;
#pragma unroll
  for (int m = 0; m < E / 4; ++m) dst[t + TPT * m] = acc[m] * scale2;
}

// grad[b][p] = sum_k gpart[b][k][p] (fixed order)
__global__ void k_adj_sum(const float* __restrict__ gpart, float* __restrict__ grad, int B, int nchunk, int count) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)B * count) return;
  const int p = (int)(idx % count);
  const size_t b = idx / count;
  grad[idx] = sum_in_order(gpart + (b * nchunk) * count + p, (size_t)count, nchunk);
}

static size_t al256a(size_t x) { return (x + 255) & ~(size_t)255; }

// rows-pass workspace (intermediate [B*T][N/4][n][4] + pupil field [B][n][n]), as in the forward
static size_t adj_inter_bytes(int n, int N, int T, int B) {
  return al256a((size_t)B * T * (size_t)N * n * sizeof(float2) + (size_t)B * n * n * sizeof(float2));
}

template <int N>
static int launch_adjoint(const float2* inter, const float2* field, const float2* mx, const float2* my,
                          const float2* txs, const float2* tys, const float* gframe, int T, int B, int nchunk,
                          float scale2, Bands bd, float2* Z, float* gpart, float* grad, hipStream_t st) {
  constexpr int TPT = fast_tpt(N);
  constexpr int n = N / 4;
  constexpr int G = 64 / TPT;
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  {
    const long tasks = (long)B * T * ((bd.a1 - bd.a0) + (bd.b1 - bd.b0));
    dim3 grid((unsigned)(TPT == 64 ? tasks : (tasks + kWaves - 1) / kWaves));
    hipLaunchKernelGGL((k_bwd_cols<N, TPT>), grid, dim3(64 * kWaves), 0, st, inter, my, gframe, tw, Z, B, T, bd);
    OOPAO_LAUNCH_CHECK();
  }
  {
    constexpr int RPB = kWaves * G;
    const int nrb = (n + RPB - 1) / RPB;
    hipLaunchKernelGGL((k_bwd_rows<N, TPT>), dim3((unsigned)(nrb * nchunk * B)), dim3(64 * kWaves), 0, st, Z, mx,
                       field, txs, tys, tw, gpart, B, T, nchunk, scale2, bd);
    OOPAO_LAUNCH_CHECK();
  }
  {
    const size_t total = (size_t)B * n * n;
    hipLaunchKernelGGL(k_adj_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, gpart, grad, B, nchunk,
                       n * n);
    OOPAO_LAUNCH_CHECK();
  }
  return OK;
}

}  // namespace oopao

using namespace oopao;

extern "C" {

int oopao_pwfs_adjoint_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(n > 0 && N == 4 * n && T > 0 && B > 0 && nchunk > 0 && nchunk <= T,
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d", n, N, T, B, nchunk);
  *bytes = adj_inter_bytes(n, N, T, B) + al256a((size_t)B * T * (size_t)N * n * sizeof(float2)) +
           al256a((size_t)B * nchunk * (size_t)n * n * sizeof(float));
  return OK;
}

int oopao_pwfs_adjoint(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                       const void* ty, const void* txs, const void* mx, const void* my, const float* gframe, int n,
                       int N, int T, int B, int nchunk, int pts_per_task, const int* bands, void* workspace,
                       size_t ws_bytes, float* grad, void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && pstart && ty && txs && mx && my && gframe && workspace && grad,
                  "null pointer argument");
  size_t need = 0;
  int rc = oopao_pwfs_adjoint_bytes(n, N, T, B, nchunk, &need);
  if (rc) return rc;
  Bands bd;
  rc = read_bands(bands, N / 4, &bd);
  if (rc) return rc;
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  char* w = reinterpret_cast<char*>(workspace);
  float2* inter = reinterpret_cast<float2*>(w);
  const float2* field = inter + (size_t)B * T * (size_t)N * n;
  float2* Z = reinterpret_cast<float2*>(w + adj_inter_bytes(n, N, T, B));
  float* gpart = reinterpret_cast<float*>(w + adj_inter_bytes(n, N, T, B) +
                                          al256a((size_t)B * T * (size_t)N * n * sizeof(float2)));
  // recompute the per-point row transforms (and the pupil field) exactly as the forward
  rc = oopao_pwfs_rows(opd, scale, amp, tx, P, pstart, ty, mx, n, N, T, B, pts_per_task, 0, bands, inter,
                       adj_inter_bytes(n, N, T, B), stream);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const float2* mxp = reinterpret_cast<const float2*>(mx);
  const float2* myp = reinterpret_cast<const float2*>(my);
  const float2* txsp = reinterpret_cast<const float2*>(txs);
  const float2* tysp = reinterpret_cast<const float2*>(ty);
  switch (N) {
#define OOPAO_CASE(NN) \
  case NN:             \
    return launch_adjoint<NN>(inter, field, mxp, myp, txsp, tysp, gframe, T, B, nchunk, 2.0f * scale, bd, Z, gpart, \
                              grad, st);
    OOPAO_FAST_SIZES(OOPAO_CASE)
#undef OOPAO_CASE
    default:
      break;
  }
  set_error("unsupported padded length N=%d", N);
  return ERR_UNSUPPORTED;
}

}  // extern "C"
