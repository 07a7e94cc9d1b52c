// General modulated Pyramid WFS on CDNA4 (gfx950): any padded length N (prime
// factors handled by fft_generic.hpp), any N x N mask and any integer detector
// binning. This is the path for sensors outside the separable fast path:
// rooftop / rotated / shifted pyramids, spatial filters, the reference's
// default n_pix_separation = 2 geometries (e.g. N = 264), other pixel scales,
// and the unbinned raw frame.
//
// Per screen b and modulation point i (Pyramid.py:612-646, 712-757):
//   pass 1 (k_gen_rows)    : R_i[y,:]  = FFT_N(support row y of u tx_i ty_i), y < n
//   pass 2 (k_gen_cols)    : G_i[:,kx] = IFFT_N(M'[:,kx] . FFT_N(column kx of R_i))
//   pass 3 (k_gen_rows_inv): |IFFT_N(G_i[Y,:])|^2 summed over a chunk of points
//   then a fixed-order sum over chunks (the raw frame) and the b x b sum-binning.
// The pupil sits at offset c0 = N/2 - n/2 of the support as in the reference,
// so the transforms see exactly the reference's padded field; M' is the mask
// with both inverse normalisations (1/N^2) folded in, stored column-major
// (M'[kx][ky]) so that pass 2 reads it contiguously.
//
// Layouts (complex64 unless noted): R [B*T][n][N], G [B*T][N][N],
// partial sums [B][nchunk][N][N] float32, raw [B][N][N] float32,
// frames [B][N/bin][N/bin] float32.
#include "common.hpp"
#include "fft_generic.hpp"

namespace oopao {

extern __shared__ float2 gen_lds[];

// Pass 1: one wavefront per (screen, point, pupil row).
__global__ void __launch_bounds__(64)
k_gen_rows(const float* __restrict__ opd, float scale, const float* __restrict__ amp,
           const float2* __restrict__ tx, const float2* __restrict__ ty, const float2* __restrict__ tw, GenPlan p,
           int n, int T, int c0, float2* __restrict__ inter1) {
  const int N = p.N;
  const int lane = threadIdx.x;
  const size_t task = blockIdx.x;
  const int y = (int)(task % n);
  const size_t bi = task / n;
  const int i = (int)(bi % T);
  const size_t b = bi / T;
  float2* a = gen_lds;
  float2* s = gen_lds + N;
  const float2 tyv = ty[(size_t)i * n + y];
  for (int X = lane; X < N; X += 64) {
    float2 v = make_float2(0.f, 0.f);
    const int x = X - c0;
    if (x >= 0 && x < n) {
      float sn, cs;
      sincosf(opd[(b * n + y) * n + x] * scale, &sn, &cs);
      const float av = amp[(size_t)y * n + x];
      v = cmul(cmul(make_float2(av * cs, av * sn), tx[(size_t)i * n + x]), tyv);
    }
    a[X] = v;
  }
  wave_lds_sync();
  const float2* r = gen_fft<-1>(a, s, tw, p, lane);
  float2* dst = inter1 + task * N;
  for (int X = lane; X < N; X += 64) dst[X] = r[X];
}

// Pass 2: one wavefront per (screen, point, focal column kx), blockDim / 64 adjacent columns per
// workgroup and consecutive column groups on one XCD: the 8-byte column reads of R and writes of G then
// fill whole cache lines in one L1 / L2 instead of reaching memory as partial lines.
__global__ void __launch_bounds__(256)
k_gen_cols(const float2* __restrict__ inter1, const float2* __restrict__ maskT, const float2* __restrict__ tw,
           GenPlan p, int n, int c0, float2* __restrict__ inter2) {
  const int N = p.N;
  const int wpg = (int)(blockDim.x >> 6);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int NG = (N + wpg - 1) / wpg;
  const size_t task = (size_t)xcd_task((int)blockIdx.x, (int)gridDim.x);
  const int kx = (int)(task % NG) * wpg + wave;
  const size_t bi = task / NG;
  if (kx >= N) return;  // whole wavefront; the waves of a workgroup share no data
  float2* a = gen_lds + (size_t)wave * 2 * N;
  float2* s = a + N;
  const float2* src = inter1 + bi * n * N + kx;
  for (int Y = lane; Y < N; Y += 64) {
    const int y = Y - c0;
    a[Y] = (y >= 0 && y < n) ? src[(size_t)y * N] : make_float2(0.f, 0.f);
  }
  wave_lds_sync();
  float2* r = gen_fft<-1>(a, s, tw, p, lane);
  float2* o = (r == a) ? s : a;
  const float2* mcol = maskT + (size_t)kx * N;
  for (int ky = lane; ky < N; ky += 64) r[ky] = cmul(r[ky], mcol[ky]);
  wave_lds_sync();
  const float2* e = gen_fft<+1>(r, o, tw, p, lane);
  float2* dst = inter2 + bi * N * N + kx;
  for (int ky = lane; ky < N; ky += 64) dst[(size_t)ky * N] = e[ky];
}

// Pass 3: one wavefront per (screen, point chunk, detector-plane row Y).
__global__ void __launch_bounds__(64)
k_gen_rows_inv(const float2* __restrict__ inter2, const float2* __restrict__ tw, GenPlan p, int T, int nchunk,
               float* __restrict__ part) {
  const int N = p.N;
  const int lane = threadIdx.x;
  const size_t task = blockIdx.x;
  const int Y = (int)(task % N);
  const size_t bc = task / N;
  const int c = (int)(bc % nchunk);
  const size_t b = bc / nchunk;
  const int i0 = (int)(((long)c * T) / nchunk);
  const int i1 = (int)(((long)(c + 1) * T) / nchunk);
  float2* a = gen_lds;
  float2* s = gen_lds + N;
  float* acc = reinterpret_cast<float*>(gen_lds + 2 * N);
  for (int X = lane; X < N; X += 64) acc[X] = 0.f;
  for (int i = i0; i < i1; ++i) {
    const float2* src = inter2 + ((b * T + i) * N + Y) * N;
    for (int X = lane; X < N; X += 64) a[X] = src[X];
    wave_lds_sync();
    const float2* e = gen_fft<+1>(a, s, tw, p, lane);
    for (int X = lane; X < N; X += 64) acc[X] += cabs2(e[X]);
    wave_lds_sync();
  }
  float* dst = part + task * N;
  for (int X = lane; X < N; X += 64) dst[X] = acc[X];
}

// ---------------------------------------------------------------------------
// Adjoint (vector-Jacobian product) of the general frame. With U_i the support
// field of point i (u_i = amp exp(i phi) tx_i ty_i, the PSF-centring phasor
// folded into tx, ty) and E_i = F+ (M' . F U_i) the forward above (2-D; F the
// forward DFT, F+ its conjugate transpose, the unnormalised inverse; M' holds
// both 1/N), and g = dL/dI (the detector gradient spread over each bin x bin
// block):
//   W_i = F+ (conj(M') . F (g . E_i)),  dL/dphi = sum_i 2 Im(W_i conj(U_i))
// on the pupil support. The modulation points are recomputed by passes 1-2,
// then
//   k_gadj_mid  : per (point, row Y): E row = IFFT(G row), times g, FFT, back
//                 into G's row
//   k_gadj_cols : per (point, column kx): IFFT(conj(M') . FFT(column)), the
//                 pupil rows kept -> Z [B*T][n][N] (over pass 1's buffer)
//   k_gadj_rows : per (screen, point chunk, pupil row y): IFFT of Z's row,
//                 2 scale Im(w conj(u)) on the pupil columns, summed over the
//                 chunk's points
// and the fixed-order sum over chunks.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
k_gadj_mid(float2* __restrict__ inter2, const float* __restrict__ gframe, const float2* __restrict__ tw, GenPlan p,
           int T, int bin) {
  const int N = p.N;
  const int cam = N / bin;
  const int lane = threadIdx.x;
  const size_t task = blockIdx.x;  // (screen * T + point) * N + Y
  const int Y = (int)(task % N);
  const size_t b = task / N / T;
  float2* a = gen_lds;
  float2* s = gen_lds + N;
  float2* row = inter2 + task * N;
  for (int X = lane; X < N; X += 64) a[X] = row[X];
  wave_lds_sync();
  float2* e = gen_fft<+1>(a, s, tw, p, lane);
  float2* o = (e == a) ? s : a;
  const float* grow = gframe + (b * cam + Y / bin) * (size_t)cam;
  for (int X = lane; X < N; X += 64) e[X] = cscale(e[X], grow[X / bin]);
  wave_lds_sync();
  const float2* r = gen_fft<-1>(e, o, tw, p, lane);
  for (int X = lane; X < N; X += 64) row[X] = r[X];
}

// as k_gen_cols (adjacent columns per workgroup, consecutive groups on one XCD), with the conjugate mask,
// all N rows in and the n pupil rows out
__global__ void __launch_bounds__(256)
k_gadj_cols(const float2* __restrict__ inter2, const float2* __restrict__ maskT, const float2* __restrict__ tw,
            GenPlan p, int n, int c0, float2* __restrict__ Z) {
  const int N = p.N;
  const int wpg = (int)(blockDim.x >> 6);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int NG = (N + wpg - 1) / wpg;
  const size_t task = (size_t)xcd_task((int)blockIdx.x, (int)gridDim.x);
  const int kx = (int)(task % NG) * wpg + wave;
  const size_t bi = task / NG;
  if (kx >= N) return;  // whole wavefront; the waves of a workgroup share no data
  float2* a = gen_lds + (size_t)wave * 2 * N;
  float2* s = a + N;
  const float2* src = inter2 + bi * N * N + kx;
  for (int Y = lane; Y < N; Y += 64) a[Y] = src[(size_t)Y * N];
  wave_lds_sync();
  float2* r = gen_fft<-1>(a, s, tw, p, lane);
  float2* o = (r == a) ? s : a;
  const float2* mcol = maskT + (size_t)kx * N;
  for (int ky = lane; ky < N; ky += 64) r[ky] = cmulc(r[ky], mcol[ky]);
  wave_lds_sync();
  const float2* e = gen_fft<+1>(r, o, tw, p, lane);
  float2* dst = Z + bi * (size_t)n * N + kx;
  for (int y = lane; y < n; y += 64) dst[(size_t)y * N] = e[c0 + y];
}

// gpart[b][chunk][y][x] = 2 scale sum_{i in chunk} Im(w_i(y, x) conj(u_i(y, x)))
__global__ void __launch_bounds__(64)
k_gadj_rows(const float2* __restrict__ Z, const float* __restrict__ opd, float scale, const float* __restrict__ amp,
            const float2* __restrict__ tx, const float2* __restrict__ ty, const float2* __restrict__ tw, GenPlan p,
            int n, int T, int nchunk, int c0, float* __restrict__ gpart) {
  const int N = p.N;
  const int lane = threadIdx.x;
  const size_t task = blockIdx.x;  // (screen * nchunk + chunk) * n + y
  const int y = (int)(task % n);
  const size_t bc = task / n;
  const int c = (int)(bc % nchunk);
  const size_t b = bc / nchunk;
  const int i0 = (int)(((long)c * T) / nchunk);
  const int i1 = (int)(((long)(c + 1) * T) / nchunk);
  float2* a = gen_lds;
  float2* s = gen_lds + N;
  float2* fld = gen_lds + 2 * N;                          // amp exp(i scale opd) on the pupil row
  float* acc = reinterpret_cast<float*>(gen_lds + 2 * N + n);
  const float* orow = opd + (b * n + y) * (size_t)n;
  const float* arow = amp + (size_t)y * n;
  for (int x = lane; x < n; x += 64) {
    float sn, cs;
    sincosf(orow[x] * scale, &sn, &cs);
    fld[x] = make_float2(arow[x] * cs, arow[x] * sn);
    acc[x] = 0.f;
  }
  for (int i = i0; i < i1; ++i) {
    const float2* src = Z + ((b * T + i) * n + y) * (size_t)N;
    for (int X = lane; X < N; X += 64) a[X] = src[X];
    wave_lds_sync();
    const float2* w = gen_fft<+1>(a, s, tw, p, lane);
    const float2 tyv = ty[(size_t)i * n + y];
    const float2* txi = tx + (size_t)i * n;
    for (int x = lane; x < n; x += 64) {
      const float2 u = cmul(cmul(fld[x], txi[x]), tyv);
      const float2 wv = w[c0 + x];
      acc[x] += wv.y * u.x - wv.x * u.y;  // Im(w conj(u))
    }
    wave_lds_sync();
  }
  float* dst = gpart + task * (size_t)n;
  for (int x = lane; x < n; x += 64) dst[x] = 2.f * scale * acc[x];
}

// Focal camera: one wavefront per (screen, point chunk, window column);
// sum over points of |FFT_N(column)|^2 / N^2 on the window rows.
__global__ void __launch_bounds__(64)
k_gen_cols_focal(const float2* __restrict__ inter1, const float2* __restrict__ tw, GenPlan p, int n, int c0,
                 int T, int nchunk, int crop, int x0, float* __restrict__ part) {
  const int N = p.N;
  const int lane = threadIdx.x;
  const size_t task = blockIdx.x;
  const int xc = (int)(task % crop);
  const size_t bc = task / crop;
  const int c = (int)(bc % nchunk);
  const size_t b = bc / nchunk;
  const int kx = x0 + xc;
  const int i0 = (int)(((long)c * T) / nchunk);
  const int i1 = (int)(((long)(c + 1) * T) / nchunk);
  float2* a = gen_lds;
  float2* s = gen_lds + N;
  float* acc = reinterpret_cast<float*>(gen_lds + 2 * N);
  for (int Y = lane; Y < crop; Y += 64) acc[Y] = 0.f;
  for (int i = i0; i < i1; ++i) {
    const float2* src = inter1 + (b * T + i) * (size_t)n * N + kx;
    for (int Y = lane; Y < N; Y += 64) {
      const int y = Y - c0;
      a[Y] = (y >= 0 && y < n) ? src[(size_t)y * N] : make_float2(0.f, 0.f);
    }
    wave_lds_sync();
    const float2* e = gen_fft<-1>(a, s, tw, p, lane);
    for (int yc = lane; yc < crop; yc += 64) acc[yc] += cabs2(e[x0 + yc]);
    wave_lds_sync();
  }
  const float inv = 1.0f / ((float)N * (float)N);
  float* dst = part + bc * (size_t)crop * crop;
  for (int yc = lane; yc < crop; yc += 64) dst[(size_t)yc * crop + xc] = acc[yc] * inv;
}

// out[b][q] = sum_k part[b][k][q] for q < count (fixed order over k)
__global__ void k_gen_sum(const float* __restrict__ part, float* __restrict__ out, int B, int nchunk, size_t count) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)B * count) return;
  const size_t q = idx % count;
  const size_t b = idx / count;
  out[idx] = sum_in_order(part + (b * nchunk) * count + q, count, nchunk);
}

// frames[b][R][C] = sum of the bin x bin block of raw (row-major order)
__global__ void k_gen_bin(const float* __restrict__ raw, float* __restrict__ frames, int B, int N, int bin) {
  const int cam = N / bin;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)B * cam * cam) return;
  const int C = (int)(idx % cam);
  const int R = (int)((idx / cam) % cam);
  const size_t b = idx / ((size_t)cam * cam);
  const float* src = raw + (b * N + (size_t)R * bin) * N + (size_t)C * bin;
  float s = 0.f;
  for (int dy = 0; dy < bin; ++dy) {
    float r = 0.f;
    for (int dx = 0; dx < bin; ++dx) r += src[(size_t)dy * N + dx];
    s += r;
  }
  frames[idx] = s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
constexpr int kGenMaxN = 2048;  // two N-point LDS buffers + an N-float accumulator stay under 64 KB

static bool gen_plan(int N, GenPlan* p) {
  if (N < 2 || N > kGenMaxN) return false;
  p->N = N;
  p->nst = 0;
  int m = N;
  auto push = [&](int r) {
    if (p->nst >= kGenMaxStages) return false;
    p->R[p->nst++] = r;
    m /= r;
    return true;
  };
  while (m % 8 == 0)
    if (!push(8)) return false;
  if (m % 4 == 0 && !push(4)) return false;
  if (m % 2 == 0 && !push(2)) return false;
  for (int f = 3; m > 1; f += 2)
    while (m % f == 0)
      if (!push(f)) return false;
  return true;
}

// the same plan for other translation units (the 2-D FFT entry point)
bool gen_plan_public(int N, GenPlan* p) { return gen_plan(N, p); }

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

static int gen_nchunk_check(int T, int nchunk) { return nchunk > 0 && nchunk <= T; }

static size_t gen_lds_bytes(int N, int extra_floats) {
  return (size_t)2 * N * sizeof(float2) + (size_t)extra_floats * sizeof(float);
}

}  // namespace oopao

using namespace oopao;

extern "C" {

int oopao_pwfs_generic_supported(int n, int N, int bin) {
  GenPlan p;
  if (n < 1 || n > N || bin < 1 || N % bin != 0) return 0;
  return gen_plan(N, &p) ? 1 : 0;
}

int oopao_pwfs_generic_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(n > 0 && n <= N && T > 0 && B > 0 && gen_nchunk_check(T, nchunk),
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d", n, N, T, B, nchunk);
  *bytes = al256((size_t)B * T * n * (size_t)N * sizeof(float2)) +
           al256((size_t)B * T * (size_t)N * N * sizeof(float2)) +
           al256((size_t)B * nchunk * (size_t)N * N * sizeof(float)) + al256((size_t)B * (size_t)N * N * sizeof(float));
  return OK;
}

int oopao_pwfs_generic_frames(const float* opd, float scale, const float* amp, const void* tx, const void* ty,
                              const void* maskT, int n, int N, int T, int B, int bin, int nchunk, void* workspace,
                              size_t ws_bytes, float* frames, float* raw, void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && ty && maskT && workspace && (frames || raw), "null pointer argument");
  OOPAO_CHECK_ARG(n > 0 && n <= N && T > 0 && B > 0 && gen_nchunk_check(T, nchunk),
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d", n, N, T, B, nchunk);
  OOPAO_CHECK_ARG(bin >= 1 && N % bin == 0, "binning %d does not divide N=%d", bin, N);
  GenPlan p;
  if (!gen_plan(N, &p)) {
    set_error("padded length N=%d outside 2..%d", N, kGenMaxN);
    return ERR_UNSUPPORTED;
  }
  size_t need = 0;
  oopao_pwfs_generic_bytes(n, N, T, B, nchunk, &need);
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* w = reinterpret_cast<char*>(workspace);
  float2* inter1 = reinterpret_cast<float2*>(w);
  w += al256((size_t)B * T * n * (size_t)N * sizeof(float2));
  float2* inter2 = reinterpret_cast<float2*>(w);
  w += al256((size_t)B * T * (size_t)N * N * sizeof(float2));
  float* part = reinterpret_cast<float*>(w);
  w += al256((size_t)B * nchunk * (size_t)N * N * sizeof(float));
  float* rawbuf = raw ? raw : reinterpret_cast<float*>(w);
  const int c0 = N / 2 - n / 2;
  const size_t l2 = gen_lds_bytes(N, 0);
  const size_t l3 = gen_lds_bytes(N, N);
  hipLaunchKernelGGL(k_gen_rows, dim3((unsigned)((size_t)B * T * n)), dim3(64), l2, st, opd, scale, amp,
                     reinterpret_cast<const float2*>(tx), reinterpret_cast<const float2*>(ty), tw, p, n, T, c0,
                     inter1);
  OOPAO_LAUNCH_CHECK();
  // columns per workgroup: 4 while their LDS (2 N complex each) stays within 64 KB, else 2
  const int wpg = (4 * l2 <= 65536) ? 4 : 2;
  const size_t ncol_groups = (size_t)B * T * ((N + wpg - 1) / wpg);
  if (ncol_groups > 0x7FFFFFFF) {
    set_error("too many column tasks (B*T*N/%d = %zu)", wpg, ncol_groups);
    return ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(k_gen_cols, dim3((unsigned)ncol_groups), dim3(64 * wpg), wpg * l2, st, inter1,
                     reinterpret_cast<const float2*>(maskT), tw, p, n, c0, inter2);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gen_rows_inv, dim3((unsigned)((size_t)B * nchunk * N)), dim3(64), l3, st, inter2, tw, p, T,
                     nchunk, part);
  OOPAO_LAUNCH_CHECK();
  const size_t npix = (size_t)B * N * N;
  hipLaunchKernelGGL(k_gen_sum, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, part, rawbuf, B, nchunk,
                     (size_t)N * N);
  OOPAO_LAUNCH_CHECK();
  if (frames) {
    const int cam = N / bin;
    const size_t nout = (size_t)B * cam * cam;
    hipLaunchKernelGGL(k_gen_bin, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, rawbuf, frames, B, N, bin);
    OOPAO_LAUNCH_CHECK();
  }
  return OK;
}

int oopao_pwfs_generic_focal_bytes(int n, int N, int T, int B, int nchunk, int crop, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(n > 0 && n <= N && T > 0 && B > 0 && gen_nchunk_check(T, nchunk) && crop > 0 && crop <= N,
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d crop=%d", n, N, T, B, nchunk, crop);
  *bytes = al256((size_t)B * T * n * (size_t)N * sizeof(float2)) +
           al256((size_t)B * nchunk * (size_t)crop * crop * sizeof(float));
  return OK;
}

int oopao_pwfs_generic_focal(const float* opd, float scale, const float* amp, const void* tx, const void* ty, int n,
                             int N, int T, int B, int nchunk, int crop, void* workspace, size_t ws_bytes, float* out,
                             void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && ty && workspace && out, "null pointer argument");
  size_t need = 0;
  int rc = oopao_pwfs_generic_focal_bytes(n, N, T, B, nchunk, crop, &need);
  if (rc) return rc;
  GenPlan p;
  if (!gen_plan(N, &p)) {
    set_error("padded length N=%d outside 2..%d", N, kGenMaxN);
    return ERR_UNSUPPORTED;
  }
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  const float2* tw = nullptr;
  rc = twiddles(N, &tw);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float2* inter1 = reinterpret_cast<float2*>(workspace);
  float* part = reinterpret_cast<float*>(reinterpret_cast<char*>(workspace) +
                                         al256((size_t)B * T * n * (size_t)N * sizeof(float2)));
  const int c0 = N / 2 - n / 2;
  const int x0 = N / 2 - crop / 2;
  hipLaunchKernelGGL(k_gen_rows, dim3((unsigned)((size_t)B * T * n)), dim3(64), gen_lds_bytes(N, 0), st, opd, scale,
                     amp, reinterpret_cast<const float2*>(tx), reinterpret_cast<const float2*>(ty), tw, p, n, T, c0,
                     inter1);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gen_cols_focal, dim3((unsigned)((size_t)B * nchunk * crop)), dim3(64), gen_lds_bytes(N, crop),
                     st, inter1, tw, p, n, c0, T, nchunk, crop, x0, part);
  OOPAO_LAUNCH_CHECK();
  const size_t cnt = (size_t)crop * crop;
  hipLaunchKernelGGL(k_gen_sum, dim3((unsigned)((B * cnt + 255) / 256)), dim3(256), 0, st, part, out, B, nchunk, cnt);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_pwfs_generic_adjoint_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(n > 0 && n <= N && T > 0 && B > 0 && gen_nchunk_check(T, nchunk),
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d", n, N, T, B, nchunk);
  *bytes = al256((size_t)B * T * n * (size_t)N * sizeof(float2)) +
           al256((size_t)B * T * (size_t)N * N * sizeof(float2)) +
           al256((size_t)B * nchunk * (size_t)n * n * sizeof(float));
  return OK;
}

int oopao_pwfs_generic_adjoint(const float* opd, float scale, const float* amp, const void* tx, const void* ty,
                               const void* maskT, const float* gframe, int n, int N, int T, int B, int bin,
                               int nchunk, void* workspace, size_t ws_bytes, float* grad, void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && ty && maskT && gframe && workspace && grad, "null pointer argument");
  size_t need = 0;
  int rc = oopao_pwfs_generic_adjoint_bytes(n, N, T, B, nchunk, &need);
  if (rc) return rc;
  OOPAO_CHECK_ARG(bin >= 1 && N % bin == 0, "binning %d does not divide N=%d", bin, N);
  GenPlan p;
  if (!gen_plan(N, &p)) {
    set_error("padded length N=%d outside 2..%d", N, kGenMaxN);
    return ERR_UNSUPPORTED;
  }
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  const size_t rows_tasks = (size_t)B * T * N;
  const int wpg = (4 * gen_lds_bytes(N, 0) <= 65536) ? 4 : 2;
  const size_t col_groups = (size_t)B * T * ((N + wpg - 1) / wpg);
  if (rows_tasks > 0x7FFFFFFF || col_groups > 0x7FFFFFFF) {
    set_error("too many transform tasks (B*T*N = %zu)", rows_tasks);
    return ERR_UNSUPPORTED;
  }
  const float2* tw = nullptr;
  rc = twiddles(N, &tw);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* w = reinterpret_cast<char*>(workspace);
  float2* inter1 = reinterpret_cast<float2*>(w);
  w += al256((size_t)B * T * n * (size_t)N * sizeof(float2));
  float2* inter2 = reinterpret_cast<float2*>(w);
  w += al256((size_t)B * T * (size_t)N * N * sizeof(float2));
  float* gpart = reinterpret_cast<float*>(w);
  const int c0 = N / 2 - n / 2;
  const float2* txp = reinterpret_cast<const float2*>(tx);
  const float2* typ = reinterpret_cast<const float2*>(ty);
  const float2* mT = reinterpret_cast<const float2*>(maskT);
  const size_t l2 = gen_lds_bytes(N, 0);
  // the forward's first two passes recompute every point's column-transformed field
  hipLaunchKernelGGL(k_gen_rows, dim3((unsigned)((size_t)B * T * n)), dim3(64), l2, st, opd, scale, amp, txp, typ,
                     tw, p, n, T, c0, inter1);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gen_cols, dim3((unsigned)col_groups), dim3(64 * wpg), wpg * l2, st, inter1, mT, tw, p, n, c0,
                     inter2);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gadj_mid, dim3((unsigned)rows_tasks), dim3(64), l2, st, inter2, gframe, tw, p, T, bin);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gadj_cols, dim3((unsigned)col_groups), dim3(64 * wpg), wpg * l2, st, inter2, mT, tw, p, n, c0,
                     inter1);
  OOPAO_LAUNCH_CHECK();
  const size_t l3 = l2 + (size_t)n * (sizeof(float2) + sizeof(float));
  hipLaunchKernelGGL(k_gadj_rows, dim3((unsigned)((size_t)B * nchunk * n)), dim3(64), l3, st, inter1, opd, scale,
                     amp, txp, typ, tw, p, n, T, nchunk, c0, gpart);
  OOPAO_LAUNCH_CHECK();
  const size_t cnt = (size_t)n * n;
  hipLaunchKernelGGL(k_gen_sum, dim3((unsigned)((B * cnt + 255) / 256)), dim3(256), 0, st, gpart, grad, B, nchunk,
                     cnt);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
