// Modulated Pyramid WFS forward model on CDNA4 (gfx950).
//
// Per modulation point i and phase screen b the reference computes
//   E_i   = amp * exp(i (phi_b + TT_i))   on the n x n pupil, zero-padded to N x N
//   F_i   = fft2(E_i * phasor) ; G_i = F_i * mask ; I_i = |ifft2(G_i)|^2
//   frame = bin(sum_i I_i)                               (Pyramid.py:612-646, 712-757)
// For a mask whose phase is additively separable, m[k,l] = a[k] + b[l] (the
// default pyramid: m = -c(|x_l| + |y_k|), Pyramid.py:453-480), the 2-D operator
// factorises: ifft2(fft2(S) * my mx^T) = Ky S Kx^T with Kx = IFFT . diag(mx) . FFT
// along X and Ky likewise along Y. The tip/tilt modulation TT_i = px_i Tip +
// py_i Tilt and the PSF-centring phasor are separable too. So the frame is
// computed in two passes with only an n x N intermediate per point:
//   rows pass : C_i[y,:] = ty_i[y] IFFT_N( mx' . FFT_N(pad(u[y,:] tx_p)) )
//   cols pass : E_i[:,X] = IFFT_N( my' . FFT_N(pad(C_i[:,X])) ), |E_i|^2 summed
//               over a chunk of points and over 4 x 4 detector pixels
// The tilt phasor ty_i[y] is a scalar per pupil row, so it commutes with the
// row transforms: the rows pass transforms each distinct tip table tx_p once
// and writes it, scaled by ty_i[y], for every point i of its group (points
// pstart[p] .. pstart[p+1]-1). On the default circular path points i and
// T - i share their tip, which halves the rows-pass transforms.
// followed by a tiny deterministic reduction over point chunks. The zero
// padding offset c0 and the 1/N of each inverse transform are folded into the
// mask vectors mx', my' on the host; tx, ty carry exp(i px lin[x]) and the
// phasor.
//
// Layout of the n x N intermediate (per point, per screen): tiled as
// [N/4][n][4] complex64, so that one detector column (4 pupil-plane columns)
// is one contiguous n*32-byte slab for the columns pass.
#include <algorithm>

#include "common.hpp"
#include "fft_core.hpp"
#include "pwfs_fast.hpp"

namespace oopao {

// Pupil-plane field u[b][y][x] = amp[y][x] exp(i scale opd[b][y][x]), once per
// screen (the reference re-forms it per modulation point, Pyramid.py:619-620).
__global__ void k_field(const float* __restrict__ opd, float scale, const float* __restrict__ amp,
                        float2* __restrict__ u, int n, int B) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)B * n * n;
  if (idx >= total) return;
  const size_t p = idx % ((size_t)n * n);
  float s, c;
  sincosf(opd[idx] * scale, &s, &c);
  const float a = amp[p];
  u[idx] = make_float2(a * c, a * s);
}

// ---------------------------------------------------------------------------
// Rows pass. One lane group (TPT lanes) per pupil row y; each transformed row
// is written once per point of the table's group, times that point's tilt
// phasor. A workgroup covers RPB adjacent rows (so that its stores fill whole
// 128-byte lines of the tiled intermediate). The items (screen b, tip table p,
// row block), row block fastest, are dealt out with a stride of the grid size:
// workgroup w takes items w, w + nwg, w + 2 nwg, ... The host launches the
// resident number of workgroups times the rounds that keep each workgroup
// within pts_per_task items, so the pass runs in rounds of equal work
// (config 2 slopes-only: 2960 items over 768 workgroups, 3 or 4 each, where
// fixed tasks of 6 tables left a second, nearly empty round), and at any
// moment the running workgroups write the slabs of a few consecutive tables,
// every row of them (a contiguous range of items per workgroup instead spread
// the stores over all points at once and was slower).
// With BANDS only the pupil-plane columns of the detector columns in `bd` are
// written (a separate instantiation: the whole-frame kernel carries no column
// tests). FWD_ONLY stops after the forward transform (focal-plane field rows,
// no mask). As in the columns pass, the transforms read their twiddles from
// an LDS copy of the table prefix (tw_lds_rows()).
// ---------------------------------------------------------------------------
template <int N, int TPT, bool FWD_ONLY = false, bool BANDS = false>
__global__ void __launch_bounds__(64 * kWaves)
k_rows_sep(const float* __restrict__ opd, float scale, const float* __restrict__ amp, float2* __restrict__ field,
           const float2* __restrict__ tx, const int* __restrict__ pstart, const float2* __restrict__ ty,
           const float2* __restrict__ mx, const float2* __restrict__ tw, float2* __restrict__ ws, int B, int P,
           int T, long items, long nwg, Bands bd) {
  constexpr int n = N / 4;
  constexpr int E = N / TPT;
  constexpr int G = 64 / TPT;
  constexpr int RPB = kWaves * G;  // rows per block
  constexpr bool kStage = (n % TPT) == 0;   // the input rows staged through LDS (below)
  constexpr bool kFuse = rows_form_field(N, TPT, FWD_ONLY);   // the field formed here (below)
  static_assert(!kFuse || kStage, "the field is formed on the staged path only");
  __shared__ float2 lds[kWaves * G][lds_len(N)];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane / TPT;
  const int t = lane % TPT;
  constexpr int nrb = (n + RPB - 1) / RPB;
  float2* xl = lds[wave * G + h];
  constexpr bool kTwLds = tw_lds_rows(N, FWD_ONLY);
  constexpr int kTw = kTwLds ? sep_tw_len<N, TPT>() : 1;
  static_assert(!kTwLds || kTw <= N, "twiddle prefix within the table");
  __shared__ float2 twl[kTw];
  if constexpr (kTwLds) {
    copy_twiddles<kTw, 64 * kWaves>(twl, tw);
    __syncthreads();
  }
  const float2* twp = kTwLds ? static_cast<const float2*>(twl) : tw;

  for (long it = blockIdx.x; it < items; it += nwg) {
    // item = (b * P + i) * nrb + rb (uniform over the workgroup)
    const long bi = it / nrb;
    const int rb = (int)(it - bi * nrb);
    const int b = (int)(bi / P);
    const int i = (int)(bi - (long)b * P);
    const int y_raw = rb * RPB + wave * G + h;
    const bool active = y_raw < n;
    const int y = active ? y_raw : n - 1;
    const float2* txi = launder(tx) + (size_t)i * n;
    // All of an item's global loads are issued before the first of them is used (the scheduling
    // barriers keep the products below the loads): left to itself the compiler issued one load pair at
    // a time, each followed by a wait for it (and, the counter being shared, for every store in
    // flight), 40 dependent round trips per item.
    float2 v[E];
    if constexpr (kStage) {
      // The row's n field values and the table's n tip values are formed / loaded contiguously (n / TPT
      // each per lane: every load instruction reads distinct elements), multiplied there, and spread to
      // the replicated layout v[m] = x[(t + TPT m) / 4] of the pruned transform through the lane group's
      // exchange buffer, which is free until the transform's first exchange. In the replicated layout
      // each group of four lanes had loaded the same element: four times the load instructions and
      // products. With kFuse the field values amp exp(i scale opd) are formed here, per table, with
      // k_field's expression (so the same values) instead of in a launch of their own, and the items of
      // table 0 store them for the consumers of the workspace's field (the adjoint); otherwise the row's
      // field is re-read per table (L2 hits after the first).
      constexpr int Q = E / 4;
      float2 a[Q], c[Q];
      if constexpr (kFuse) {
        const float* op = launder(opd) + ((size_t)b * n + y) * n;
        const float* ap = launder(amp) + (size_t)y * n;
        float o[Q], am[Q];
#pragma unroll
        for (int j = 0; j < Q; ++j) o[j] = op[t + TPT * j];
#pragma unroll
        for (int j = 0; j < Q; ++j) am[j] = ap[t + TPT * j];
#pragma unroll
        for (int j = 0; j < Q; ++j) c[j] = txi[t + TPT * j];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < Q; ++j) {
          float sn, cs;
          sincosf(o[j] * scale, &sn, &cs);
          a[j] = make_float2(am[j] * cs, am[j] * sn);
        }
        if (i == 0 && active) {
          float2* fo = field + ((size_t)b * n + y) * n;
#pragma unroll
          for (int j = 0; j < Q; ++j) fo[t + TPT * j] = a[j];
        }
      } else {
        const float2* fp = launder(field) + ((size_t)b * n + y) * n;
#pragma unroll
        for (int j = 0; j < Q; ++j) a[j] = fp[t + TPT * j];
#pragma unroll
        for (int j = 0; j < Q; ++j) c[j] = txi[t + TPT * j];
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int j = 0; j < Q; ++j) xl[t + TPT * j] = cmul(a[j], c[j]);
      wave_lds_sync();
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = xl[(t + TPT * m) >> 2];
      wave_lds_sync();
    } else {
      // the field from k_field's launch; the row's values are re-read per table (L2 hits after the first)
      const float2* fp = launder(field) + ((size_t)b * n + y) * n;
      float2 a[E], c[E];
#pragma unroll
      for (int m = 0; m < E; ++m) a[m] = fp[(t + TPT * m) >> 2];
#pragma unroll
      for (int m = 0; m < E; ++m) c[m] = txi[(t + TPT * m) >> 2];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(a[m], c[m]);
    }
    fft_fwd_pruned<N, TPT, 4, true>(v, xl, twp, t);
    if constexpr (!FWD_ONLY) {
      const float2* mxp = launder(mx);
      float2 c[E];
#pragma unroll
      for (int m = 0; m < E; ++m) c[m] = mxp[t + TPT * m];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(v[m], c[m]);
      fft_full<N, TPT, +1, true>(v, xl, twp, t);
    }
    if (active) {
      const int k0 = min(max(pstart[i], 0), T);
      const int k1 = min(max(pstart[i + 1], k0), T);
      // element (X, y) of a point's [N/4][n][4] slab, X = t + TPT m: byte offset lane_off + m * kStride
      // (TPT is a multiple of 4, so X & 3 = t & 3 and X >> 2 = t / 4 + m TPT / 4). Buffer stores take
      // the lane part in one VGPR and the per-m part as a uniform offset, instead of a 64-bit address
      // per element.
      const int tq = t >> 2;
      const unsigned lane_off = (unsigned)(((tq * n + y) * 4 + (t & 3)) * (int)sizeof(float2));
      constexpr int kStride = (TPT / 4) * n * 4 * (int)sizeof(float2);
      constexpr unsigned kSlab = (unsigned)N * n * sizeof(float2);
      for (int k = k0; k < k1; ++k) {
        const float2 tyv = ty[(size_t)k * n + y];
        // wait for the tilt value here, once: with the band tests each store sits in its own branch, and
        // a first use of tyv inside those branches put a vmcnt(0) wait (which on CDNA also waits for every
        // store in flight) at the head of each one, serialising the stores
        asm volatile("" ::"v"(tyv.x), "v"(tyv.y));
        const BufferRsrc dst = buffer_rsrc(ws + ((size_t)b * T + k) * (size_t)(N / 4) * n * 4, kSlab);
        // band limits through an opaque zero: the column tests stay here instead of being hoisted out
        // of both loops as E lane masks (which cost a wave per SIMD of occupancy)
        int z = 0;
        asm volatile("" : "+s"(z));
        const Bands bk{bd.a0 + z, bd.a1 + z, bd.b0 + z, bd.b1 + z};
#pragma unroll
        for (int m = 0; m < E; ++m) {
          // detector column X >> 2 = t / 4 + m TPT / 4: the lane part t / 4 (one register) is compared with
          // the band limits shifted by the uniform part (scalar arithmetic), rather than keeping E column
          // indices in registers
          const int cm = (TPT / 4) * m;
          if (!BANDS || (tq >= bk.a0 - cm && tq < bk.a1 - cm) || (tq >= bk.b0 - cm && tq < bk.b1 - cm))
            buffer_store_c64(cmul(v[m], tyv), dst, lane_off, m * kStride);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Columns pass. Per (screen b, point chunk, detector column g) the 4 pupil-
// plane columns 4g..4g+3 are transformed for every point of the chunk, |E|^2
// accumulated in registers, then binned 4 rows x 4 columns.
//  * 64-lane transforms (one per wavefront): one workgroup per (b, chunk, g),
//    wavefront w takes column 4g+w. The four wavefronts walk the same points
//    in step, so each 32-byte row segment of the slab is fetched once into
//    L1/L2 and shared; the column sums are combined through LDS.
//  * narrower transforms: one wavefront per (b, chunk, g), its 64/TPT lane
//    groups taking the columns.
// Only the detector columns in `bd` are computed (one task per such column).
// For the plans tw_lds_cols() names, the workgroup first copies the twiddle
// entries the plans read into LDS: every stage's twiddles are then LDS reads
// (lgkm queue, ~100-cycle latency) instead of global loads queued behind the
// column loads (the table loads cost ~9 % of the kernel at N = 1280). With
// 64-lane transforms the column sums reuse the exchange buffers' LDS, so the
// workgroup needs 48 KB at N = 1280 and three still fit per CU; at N = 320
// (16 lanes) the 2.5 KB prefix takes the workgroup to 44 KB, three per CU as
// before.
// Output: part[b][chunk][g][r] (detector column-major partial sums).
// ---------------------------------------------------------------------------
template <int N, int TPT>
__global__ void __launch_bounds__(64 * kWaves, fast_cols_min_waves(N))
k_cols_sep(const float2* __restrict__ ws, const float2* __restrict__ my,
           const float2* __restrict__ tw, float* __restrict__ part, int B, int T, int nchunk, Bands bd) {
  constexpr int n = N / 4;
  constexpr int E = N / TPT;
  constexpr int G = 64 / TPT;
  constexpr int NG = N / 4;
  constexpr bool kWaveCols = (TPT == 64);
  constexpr bool kTwLds = tw_lds_cols(N, TPT);
  constexpr int kTw = kTwLds ? sep_tw_len<N, TPT>() : 1;
  static_assert(!kTwLds || kTw <= N, "twiddle prefix within the table");
  static_assert(kWaves * NG <= 2 * kWaves * G * lds_len(N), "column sums fit in the exchange buffers");
  __shared__ float2 lds[kWaves * G][lds_len(N) + cols_group_pad(N, TPT)];
  __shared__ float2 twl[kTw];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane / TPT;
  const int t = lane % TPT;
  const int ng = (bd.a1 - bd.a0) + (bd.b1 - bd.b0);
  int task = kWaveCols ? (int)blockIdx.x : (int)blockIdx.x * kWaves + wave;
  const int g = band_col(task % ng, bd);
  task /= ng;
  const int chunk = task % nchunk;
  const int b = task / nchunk;
  // the twiddle copy and its barrier come before the early exit, with every wavefront of the workgroup
  // present (with narrow transforms the last workgroup's wavefronts past the tasks leave below)
  if constexpr (kTwLds) {
    copy_twiddles<kTw, 64 * kWaves>(twl, tw);
    __syncthreads();
  }
  if (b >= B) return;  // uniform over the wavefront (and over the workgroup when kWaveCols)
  const int i0 = (int)(((long)chunk * T) / nchunk);
  const int i1 = (int)(((long)(chunk + 1) * T) / nchunk);
  float2* xl = lds[wave * G + h];
  const float2* twp = kTwLds ? static_cast<const float2*>(twl) : tw;

  float acc[E];
#pragma unroll
  for (int m = 0; m < E; ++m) acc[m] = 0.f;
  // With kBuf the column values and mask values come through buffer loads: the lane part of each offset
  // in one VGPR, the per-element part (a compile-time multiple of m) as a uniform offset, instead of a
  // 64-bit address computed per element. Element ((t + TPT m) / 4) * 4 + c of a tile is
  // ((t / 4) * 4 + c) + m TPT (TPT is a multiple of 4), below n * 4 = N / 4 * 4.
  constexpr bool kBuf = buf_cols_sep(N);
  const BufferRsrc msk = buffer_rsrc(my, (unsigned)N * sizeof(float2));
  const unsigned msk_lane = (unsigned)t * sizeof(float2);
  for (int i = i0; i < i1; ++i) {
    const float2* tile = ws + (((size_t)b * T + i) * NG + g) * (size_t)n * 4;
    const BufferRsrc col = buffer_rsrc(tile, (unsigned)n * 4 * sizeof(float2));
#pragma unroll 1
    for (int c0 = 0; c0 < (kWaveCols ? 1 : 4); c0 += G) {
      const int c = kWaveCols ? wave : c0 + h;
      const unsigned col_lane = (unsigned)(((t >> 2) * 4 + c) * (int)sizeof(float2));
      float2 v[E];
#pragma unroll
      for (int m = 0; m < E; ++m)
        v[m] = kBuf ? buffer_load_c64(col, col_lane, m * TPT * (int)sizeof(float2))
                    : tile[((t + TPT * m) >> 2) * 4 + c];
      // XOR-swizzled LDS exchange (conflict-free small-span stages), as in the
      // rows pass.
      fft_fwd_pruned<N, TPT, 4, true>(v, xl, twp, t);
#pragma unroll
      for (int m = 0; m < E; ++m)
        v[m] = cmul(v[m], kBuf ? buffer_load_c64(msk, msk_lane, m * TPT * (int)sizeof(float2))
                               : launder(my)[t + TPT * m]);
      fft_full<N, TPT, +1, true>(v, xl, twp, t);
#pragma unroll
      for (int m = 0; m < E; ++m) acc[m] += cabs2(v[m]);
    }
  }
  // rows t..t+3 (a lane quad) form one detector row
  float* dst = part + (((size_t)b * nchunk + chunk) * NG + g) * (size_t)NG;
  if constexpr (kWaveCols) {
    // the column sums go to the exchange buffers once every wavefront's last exchange is done
    float* red = reinterpret_cast<float*>(&lds[0][0]);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float s = acc[m];
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      if ((t & 3) == 0) red[wave * NG + ((t + 64 * m) >> 2)] = s;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < NG; r += 64 * kWaves)
      dst[r] = (red[r] + red[NG + r]) + (red[2 * NG + r] + red[3 * NG + r]);
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float s = acc[m];
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
#pragma unroll
      for (int off = TPT; off < 64; off <<= 1) s += __shfl_xor(s, off);
      if (h == 0 && (t & 3) == 0) dst[(t + TPT * m) >> 2] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Focal-plane camera (Pyramid.py:638-642, 1183-1190): sum over points of
// |fft2(E_i phasor)|^2 / N^2 on the crop x crop window centred at N/2. The
// rows pass in forward-only mode leaves the row-transformed fields in the
// tiled intermediate; here each lane group transforms one column inside the
// window for a chunk of points and keeps the window's rows. (|.|^2 removes the
// zero-padding offset phase, so no mask or shift is applied.)
// part[b][chunk][crop][crop]
// ---------------------------------------------------------------------------
template <int N, int TPT>
__global__ void __launch_bounds__(64 * kWaves)
k_cols_focal(const float2* __restrict__ ws, const float2* __restrict__ tw, float* __restrict__ part, int B, int T,
             int nchunk, int crop) {
  constexpr int n = N / 4;
  constexpr int E = N / TPT;
  constexpr int G = 64 / TPT;
  constexpr int NG = N / 4;
  // (the pruned forward reads a part of the same prefix); the copy lowers this kernel's occupancy at
  // N = 320, 512, 1280 and 1536
  constexpr bool kTwLds = tw_lds_cols(N, TPT) && N != 320 && N != 512 && N != 1280 && N != 1536;
  constexpr int kTw = kTwLds ? sep_tw_len<N, TPT>() : 1;
  static_assert(!kTwLds || kTw <= N, "twiddle prefix within the table");
  __shared__ float2 lds[kWaves * G][lds_len(N)];
  __shared__ float2 twl[kTw];
  if constexpr (kTwLds) {
    copy_twiddles<kTw, 64 * kWaves>(twl, tw);
    __syncthreads();
  }
  const float2* twp = kTwLds ? static_cast<const float2*>(twl) : tw;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane / TPT;
  const int t = lane % TPT;
  const long total = (long)B * nchunk * crop;
  long task = ((long)blockIdx.x * kWaves + wave) * G + h;
  const bool active = task < total;
  if (!active) task = total - 1;
  const int xc = (int)(task % crop);
  const int chunk = (int)((task / crop) % nchunk);
  const int b = (int)(task / ((long)crop * nchunk));
  const int x0 = N / 2 - crop / 2;
  const int X = x0 + xc;
  const int i0 = (int)(((long)chunk * T) / nchunk);
  const int i1 = (int)(((long)(chunk + 1) * T) / nchunk);
  float2* xl = lds[wave * G + h];
  float acc[E];
#pragma unroll
  for (int m = 0; m < E; ++m) acc[m] = 0.f;
  // buffer loads of the column values as in k_cols_sep
  constexpr bool kBuf = buf_cols_focal(N);
  const unsigned col_lane = (unsigned)(((t >> 2) * 4 + (X & 3)) * (int)sizeof(float2));
  for (int i = i0; i < i1; ++i) {
    const float2* tile = ws + (((size_t)b * T + i) * NG + (X >> 2)) * (size_t)n * 4;
    const BufferRsrc col = buffer_rsrc(tile, (unsigned)n * 4 * sizeof(float2));
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = kBuf ? buffer_load_c64(col, col_lane, m * TPT * (int)sizeof(float2))
                  : tile[((t + TPT * m) >> 2) * 4 + (X & 3)];
    fft_fwd_pruned<N, TPT, 4>(v, xl, twp, t);
#pragma unroll
    for (int m = 0; m < E; ++m) acc[m] += cabs2(v[m]);
  }
  if (!active) return;
  const float inv = 1.0f / ((float)N * (float)N);
  float* dst = part + ((size_t)b * nchunk + chunk) * (size_t)crop * crop;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int Y = t + TPT * m - x0;
    if (Y >= 0 && Y < crop) dst[(size_t)Y * crop + xc] = acc[m] * inv;
  }
}

// ---------------------------------------------------------------------------
// Unbinned frame (the reference's raw_data, Pyramid.py:751-757) on the separable path. One lane group
// per (screen b, point chunk, pupil-plane column X): the column's transform pair for every point of the
// chunk, |E|^2 summed in registers, written unbinned: part[b][chunk][X][Y].
// ---------------------------------------------------------------------------
template <int N, int TPT>
__global__ void __launch_bounds__(64 * kWaves)
k_cols_raw(const float2* __restrict__ ws, const float2* __restrict__ my, const float2* __restrict__ tw,
           float* __restrict__ part, int B, int T, int nchunk) {
  constexpr int n = N / 4;
  constexpr int E = N / TPT;
  constexpr int G = 64 / TPT;
  constexpr int NG = N / 4;
  // the columns pass's transforms and prefix; the copy lowers this kernel's occupancy at N = 320, 1280
  // and 2048
  constexpr bool kTwLds = tw_lds_cols(N, TPT) && N != 320 && N != 1280 && N != 2048;
  constexpr int kTw = kTwLds ? sep_tw_len<N, TPT>() : 1;
  static_assert(!kTwLds || kTw <= N, "twiddle prefix within the table");
  __shared__ float2 lds[kWaves * G][lds_len(N)];
  __shared__ float2 twl[kTw];
  if constexpr (kTwLds) {
    copy_twiddles<kTw, 64 * kWaves>(twl, tw);
    __syncthreads();
  }
  const float2* twp = kTwLds ? static_cast<const float2*>(twl) : tw;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane / TPT;
  const int t = lane % TPT;
  const long total = (long)B * nchunk * N;
  long task = ((long)blockIdx.x * kWaves + wave) * G + h;
  const bool active = task < total;
  if (!active) task = total - 1;      // the lane group repeats a valid column and writes nothing
  const int X = (int)(task % N);
  const int chunk = (int)((task / N) % nchunk);
  const int b = (int)(task / ((long)N * nchunk));
  const int i0 = (int)(((long)chunk * T) / nchunk);
  const int i1 = (int)(((long)(chunk + 1) * T) / nchunk);
  float2* xl = lds[wave * G + h];
  float acc[E];
#pragma unroll
  for (int m = 0; m < E; ++m) acc[m] = 0.f;
  // buffer loads of the column and mask values as in k_cols_sep
  constexpr bool kBuf = buf_cols_raw(N);
  const BufferRsrc msk = buffer_rsrc(my, (unsigned)N * sizeof(float2));
  const unsigned msk_lane = (unsigned)t * sizeof(float2);
  const unsigned col_lane = (unsigned)(((t >> 2) * 4 + (X & 3)) * (int)sizeof(float2));
  for (int i = i0; i < i1; ++i) {
    const float2* tile = ws + (((size_t)b * T + i) * NG + (X >> 2)) * (size_t)n * 4;
    const BufferRsrc col = buffer_rsrc(tile, (unsigned)n * 4 * sizeof(float2));
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = kBuf ? buffer_load_c64(col, col_lane, m * TPT * (int)sizeof(float2))
                  : tile[((t + TPT * m) >> 2) * 4 + (X & 3)];
    fft_fwd_pruned<N, TPT, 4, true>(v, xl, twp, t);
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = cmul(v[m], kBuf ? buffer_load_c64(msk, msk_lane, m * TPT * (int)sizeof(float2))
                             : launder(my)[t + TPT * m]);
    fft_full<N, TPT, +1, true>(v, xl, twp, t);
#pragma unroll
    for (int m = 0; m < E; ++m) acc[m] += cabs2(v[m]);
  }
  if (!active) return;
  float* dst = part + (((size_t)b * nchunk + chunk) * N + X) * (size_t)N;
#pragma unroll
  for (int m = 0; m < E; ++m) dst[t + TPT * m] = acc[m];
}

// raw[b][Y][X] = sum_k part[b][k][X][Y] (fixed order over chunks), through a 32 x 32 LDS tile so that
// both the column reads and the row writes are coalesced
__global__ void __launch_bounds__(256) k_raw_reduce(const float* __restrict__ part, float* __restrict__ raw, int N,
                                                    int nchunk) {
  __shared__ float tile[32][33];
  const int b = blockIdx.z;
  const int X0 = blockIdx.y * 32, Y0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int X = X0 + r, Y = Y0 + tx;
    float s = 0.f;
    if (X < N && Y < N)
      s = sum_in_order(part + (((size_t)b * nchunk) * N + X) * (size_t)N + Y, (size_t)N * N, nchunk);
    tile[r][tx] = s;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int Y = Y0 + r, X = X0 + tx;
    if (X < N && Y < N) raw[((size_t)b * N + Y) * N + X] = tile[tx][r];
  }
}

// Detector integration by area: detector pixel R covers raw pixels [R s, (R + 1) s), s = N / cam, and
// takes each raw pixel's flux in proportion to the overlap (a sum over whole pixels when cam divides N,
// the reference's set_binning, tools/tools.py:464-483). frames[b][R][C] = sum w(R, Y) w(C, X) raw[b][Y][X].
__global__ void __launch_bounds__(256) k_bin_area(const float* __restrict__ raw, int N, int cam, double s,
                                                  float* __restrict__ frames, int B) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)B * cam * cam) return;
  const int C = (int)(idx % cam);
  const int R = (int)((idx / cam) % cam);
  const int b = (int)(idx / ((long)cam * cam));
  const double r0 = R * s, r1 = (R + 1) * s, c0 = C * s, c1 = (C + 1) * s;
  const int y0 = (int)floor(r0), y1 = min(N, (int)ceil(r1));
  const int x0 = (int)floor(c0), x1 = min(N, (int)ceil(c1));
  const float* f = raw + (size_t)b * N * N;
  double acc = 0.0;
  for (int Y = y0; Y < y1; ++Y) {
    const double wy = fmin((double)(Y + 1), r1) - fmax((double)Y, r0);
    double row = 0.0;
    for (int X = x0; X < x1; ++X) row += (fmin((double)(X + 1), c1) - fmax((double)X, c0)) * (double)f[(size_t)Y * N + X];
    acc += wy * row;
  }
  frames[idx] = (float)acc;
}

// out[b][p] = sum_k part[b][k][p], p < count  (fixed summation order)
__global__ void k_sum_chunks(const float* __restrict__ part, float* __restrict__ out, int B, int nchunk, int count) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)B * count) return;
  const int p = (int)(idx % count);
  const int b = (int)(idx / count);
  out[idx] = sum_in_order(part + ((size_t)b * nchunk) * count + p, (size_t)count, nchunk);
}

// frame[b][r][g] = sum_k part[b][k][g][r]  (fixed summation order), through a 32 x 32 LDS tile so that the
// reads (along r) and the writes (along g) are both coalesced; columns outside the bands `bd` are set to zero
__global__ void __launch_bounds__(1024) k_reduce_chunks(const float* __restrict__ part, float* __restrict__ frame,
                                                        int B, int nchunk, int cam, Bands bd) {
  __shared__ float tile[32][33];
  const int nt = (cam + 31) / 32;
  int bid = blockIdx.x;
  const int r0 = (bid % nt) * 32;
  bid /= nt;
  const int g0 = (bid % nt) * 32;
  const int b = bid / nt;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // one element per thread
  {
    const int g = g0 + ty, r = r0 + tx;
    float s = 0.f;
    if (g < cam && r < cam && in_bands(g, bd))
      s = sum_in_order(part + (((size_t)b * nchunk) * cam + g) * cam + r, (size_t)cam * cam, nchunk);
    tile[ty][tx] = s;
  }
  __syncthreads();
  const int r = r0 + ty, g = g0 + tx;
  if (g < cam && r < cam) frame[((size_t)b * cam + r) * cam + g] = tile[tx][ty];
}

// ---------------------------------------------------------------------------
// slopesMaps post-processing (Pyramid.py:838-854) in two launches: per-block
// sums of the valid I4Q (double), then every slope normalised by their mean.
// quad[0..3] = row origins, quad[4..7] = column origins of quadrants 1..4.
// ---------------------------------------------------------------------------
// The four quadrant pixels of subaperture p = (py, px), loaded together with its valid flag rather than
// after it (the flag was a round trip to memory before the pixel loads could issue). The quadrants lie
// inside the frame, as the reference cuts them (grabQuadrant); the indices are clamped to the frame all
// the same, so that no subaperture reads outside it.
struct Quad4 {
  float I1, I2, I3, I4;
};
__device__ __forceinline__ Quad4 quad_pixels(const float* f, int cam, const int* quad, int w, int p) {
  const int py = p / w, px = p % w;
  const unsigned last = (unsigned)(cam * cam - 1);
  auto at = [&](int r0, int c0) { return f[min((unsigned)((r0 + py) * cam + c0 + px), last)]; };
  return {at(quad[0], quad[4]), at(quad[1], quad[5]), at(quad[2], quad[6]), at(quad[3], quad[7])};
}

__global__ void __launch_bounds__(256)
k_slopes_partial(const float* __restrict__ frames, int cam, const unsigned char* __restrict__ valid, int w,
                 const int* __restrict__ quad, int nblk, double* __restrict__ part) {
  // part[b][blk] = sum over this block's valid pixels of I1 + I2 + I3 + I4 (fixed-order tree)
  const int b = blockIdx.y;
  const float* f = frames + (size_t)b * cam * cam;
  __shared__ double red[256];
  const int p = blockIdx.x * 256 + threadIdx.x;
  double acc = 0.0;
  if (p < w * w) {
    const bool v = valid[p] != 0;
    const Quad4 I = quad_pixels(f, cam, quad, w, p);
    const double s4 = ((double)I.I1 + (double)I.I2) + ((double)I.I3 + (double)I.I4);
    acc = v ? s4 : 0.0;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)b * nblk + blockIdx.x] = red[0];
}

// Unnormalised slope at flat index q of the [2w][w] map (q < w*w: Sx, else Sy), Pyramid.py:846-853.
__device__ __forceinline__ float raw_slope(const float* f, int cam, const unsigned char* valid, int w,
                                           const int* quad, int q) {
  const int ww = w * w;
  const int p = q < ww ? q : q - ww;
  const bool v = valid[p] != 0;
  const Quad4 I = quad_pixels(f, cam, quad, w, p);
  const float s = q < ww ? (I.I1 - I.I2) + (I.I4 - I.I3) : (I.I1 - I.I4) + (I.I2 - I.I3);
  return v ? s : 0.f;
}

__global__ void __launch_bounds__(256)
k_slopes_finish(const float* __restrict__ frames, int cam, const unsigned char* __restrict__ valid, int w,
                const int* __restrict__ quad, const float* __restrict__ ref2d, float units,
                const int* __restrict__ vidx, int nsig, int nvalid, int nblk, const double* __restrict__ part,
                float* __restrict__ sig2d, float* __restrict__ sig) {
  const int b = blockIdx.y;
  const float* f = frames + (size_t)b * cam * cam;
  const int n2 = 2 * w * w;
  const int i = blockIdx.x * 256 + threadIdx.x;
  // this thread's unnormalised slope and reference come first: their loads are in flight while the
  // block partials are gathered below
  int q = -1;
  if (i < n2)
    q = i;
  else if (i - n2 < nsig)
    q = vidx[i - n2];
  float s = 0.f, ref = 0.f;
  if (q >= 0) {
    s = raw_slope(f, cam, valid, w, quad, q);
    ref = ref2d[q];
  }
  // the normalisation: the block partials loaded by all threads at once into LDS and added there in index
  // order by one thread (before, that thread loaded them one after another from memory)
  __shared__ double pbuf[256];
  __shared__ float inv_s;
  double t = 0.0;
  for (int k0 = 0; k0 < nblk; k0 += 256) {
    if (k0 + (int)threadIdx.x < nblk) pbuf[threadIdx.x] = part[(size_t)b * nblk + k0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) t = sum_in_order(static_cast<const double*>(pbuf), 1, min(256, nblk - k0), t);
    __syncthreads();
  }
  if (threadIdx.x == 0) inv_s = 1.0f / (float)(t / (double)nvalid);
  __syncthreads();
  const float inv = inv_s;
  if (i < n2)
    sig2d[(size_t)b * n2 + i] = (s * inv - ref) * units;
  else if (i - n2 < nsig)
    sig[(size_t)b * nsig + (i - n2)] = (s * inv - ref) * units;
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// intermediate [B*T][N/4][n][4] followed by the pupil field [B][n][n]
static size_t ws_inter_bytes(int n, int N, int T, int B) {
  return align256((size_t)B * T * (size_t)N * n * sizeof(float2) + (size_t)B * n * n * sizeof(float2));
}
static size_t ws_part_bytes(int N, int B, int nchunk) {
  return align256((size_t)B * nchunk * (size_t)(N / 4) * (N / 4) * sizeof(float));
}

template <int N, int TPT, bool FWD_ONLY>
static int launch_rows_plan(const float* opd, float scale, const float* amp, const float2* tx, int P,
                            const int* pstart, const float2* ty, const float2* mx, int T, int B, int pts_per_task,
                            Bands bd, float2* inter, hipStream_t st) {
  constexpr int n = N / 4;
  constexpr int G = 64 / TPT;
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  constexpr int RPB = kWaves * G;
  const int nrb = (n + RPB - 1) / RPB;
  float2* field = inter + (size_t)B * T * (size_t)N * n;
  const bool all = bd.a0 == 0 && bd.a1 == n && bd.b0 == n && bd.b1 == n;
  const void* kern = (FWD_ONLY || all) ? reinterpret_cast<const void*>(&k_rows_sep<N, TPT, FWD_ONLY, false>)
                                       : reinterpret_cast<const void*>(&k_rows_sep<N, TPT, false, true>);
  // workgroups: the resident count, times the rounds that keep every range within pts_per_task tables
  int slots = 0;
  rc = resident_workgroups(kern, 64 * kWaves, &slots);
  if (rc) return rc;
  const long items = (long)B * nrb * P;
  const long per_round = (long)slots * pts_per_task;
  const long nwg = std::min(items, (long)slots * ((items + per_round - 1) / per_round));
  if (!rows_form_field(N, TPT, FWD_ONLY)) {
    // the plans that do not form the field themselves read it from this launch
    const size_t total = (size_t)B * n * n;
    hipLaunchKernelGGL(k_field, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, opd, scale, amp, field,
                       n, B);
    OOPAO_LAUNCH_CHECK();
  }
  dim3 grid((unsigned)nwg);
  if (FWD_ONLY || all)
    hipLaunchKernelGGL((k_rows_sep<N, TPT, FWD_ONLY, false>), grid, dim3(64 * kWaves), 0, st, opd, scale, amp,
                       field, tx, pstart, ty, mx, tw, inter, B, P, T, items, nwg, bd);
  else
    hipLaunchKernelGGL((k_rows_sep<N, TPT, false, true>), grid, dim3(64 * kWaves), 0, st, opd, scale, amp, field,
                       tx, pstart, ty, mx, tw, inter, B, P, T, items, nwg, bd);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

// lanes = lanes per transform of a compiled plan (0 = the default one)
template <int N, bool FWD_ONLY = false>
static int launch_rows(const float* opd, float scale, const float* amp, const float2* tx, int P, const int* pstart,
                       const float2* ty, const float2* mx, int T, int B, int pts_per_task, Bands bd, float2* inter,
                       int lanes, hipStream_t st) {
  constexpr int D = fast_tpt_rows(N);
  if (lanes == 0 || lanes == D)
    return launch_rows_plan<N, D, FWD_ONLY>(opd, scale, amp, tx, P, pstart, ty, mx, T, B, pts_per_task, bd, inter,
                                            st);
  if constexpr (!FWD_ONLY && alt_tpt_rows(N) > 0) {
    if (lanes == alt_tpt_rows(N))
      return launch_rows_plan<N, alt_tpt_rows(N), false>(opd, scale, amp, tx, P, pstart, ty, mx, T, B, pts_per_task,
                                                         bd, inter, st);
  }
  set_error("no %d-lane rows plan compiled for N=%d", lanes, N);
  return ERR_UNSUPPORTED;
}

template <int N, int TPT>
static int launch_cols_plan(const float2* inter, const float2* my, int T, int B, int nchunk, Bands bd, float* part,
                            float* frames, hipStream_t st) {
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  {
    const long tasks = (long)B * nchunk * ((bd.a1 - bd.a0) + (bd.b1 - bd.b0));
    dim3 grid((unsigned)(TPT == 64 ? tasks : (tasks + kWaves - 1) / kWaves));
    hipLaunchKernelGGL((k_cols_sep<N, TPT>), grid, dim3(64 * kWaves), 0, st, inter, my, tw, part, B, T, nchunk,
                       bd);
    OOPAO_LAUNCH_CHECK();
  }
  {
    const int cam = N / 4;
    const int nt = (cam + 31) / 32;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)((size_t)B * nt * nt)), dim3(1024), 0, st, part, frames, B,
                       nchunk, cam, bd);
    OOPAO_LAUNCH_CHECK();
  }
  return OK;
}

template <int N>
static int launch_cols(const float2* inter, const float2* my, int T, int B, int nchunk, Bands bd, float* part,
                       float* frames, int lanes, hipStream_t st) {
  constexpr int D = fast_tpt(N);
  if (lanes == 0 || lanes == D) return launch_cols_plan<N, D>(inter, my, T, B, nchunk, bd, part, frames, st);
  if constexpr (alt_tpt_cols(N) > 0) {
    if (lanes == alt_tpt_cols(N))
      return launch_cols_plan<N, alt_tpt_cols(N)>(inter, my, T, B, nchunk, bd, part, frames, st);
  }
  set_error("no %d-lane columns plan compiled for N=%d", lanes, N);
  return ERR_UNSUPPORTED;
}

template <int N>
static int launch_focal(const float2* inter, int T, int B, int nchunk, int crop, float* part, float* out,
                        hipStream_t st) {
  constexpr int TPT = fast_tpt(N);
  constexpr int G = 64 / TPT;
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  const long groups = (long)B * nchunk * crop;
  const long waves = (groups + G - 1) / G;
  hipLaunchKernelGGL((k_cols_focal<N, TPT>), dim3((unsigned)((waves + kWaves - 1) / kWaves)), dim3(64 * kWaves), 0,
                     st, inter, tw, part, B, T, nchunk, crop);
  OOPAO_LAUNCH_CHECK();
  const size_t total = (size_t)B * crop * crop;
  hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, part, out, B, nchunk,
                     crop * crop);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

template <int N>
static int launch_raw(const float2* inter, const float2* my, int T, int B, int nchunk, float* part, float* raw,
                      hipStream_t st) {
  constexpr int TPT = fast_tpt(N);
  constexpr int G = 64 / TPT;
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  const long groups = (long)B * nchunk * N;
  const long waves = (groups + G - 1) / G;
  hipLaunchKernelGGL((k_cols_raw<N, TPT>), dim3((unsigned)((waves + kWaves - 1) / kWaves)), dim3(64 * kWaves), 0,
                     st, inter, my, tw, part, B, T, nchunk);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_raw_reduce, dim3((N + 31) / 32, (N + 31) / 32, B), dim3(256), 0, st, part, raw, N, nchunk);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

static size_t ws_raw_part_bytes(int N, int B, int nchunk) {
  return align256((size_t)B * nchunk * (size_t)N * N * sizeof(float));
}

static bool fast_supported(int n, int N, int bin) {
  return N == 4 * n && bin == 4 && fast_tpt(N) != 0;
}

static int check_slopes_args(const float* frames, int B, int cam, const unsigned char* valid, int w, const int* quad,
                             const float* ref2d, const int* vidx, int nsig, int nvalid, const float* sig2d,
                             const float* sig) {
  OOPAO_CHECK_ARG(frames && valid && quad && ref2d && vidx && sig2d && sig, "null pointer argument");
  OOPAO_CHECK_ARG(B > 0 && cam > 0 && w > 0 && nsig >= 0 && nvalid > 0,
                  "invalid sizes B=%d cam=%d w=%d nsig=%d nvalid=%d", B, cam, w, nsig, nvalid);
  OOPAO_CHECK_ARG(B <= 65535, "too many frames in one call: %d", B);
  return OK;
}

// slopesMaps of B frames with the per-block valid-I4Q sums in part[B][nblk]
static int launch_slopes(const float* frames, int B, int cam, const unsigned char* valid, int w, const int* quad,
                         const float* ref2d, float units, const int* vidx, int nsig, int nvalid, double* part,
                         float* sig2d, float* sig, hipStream_t st) {
  const int nblk = (w * w + 255) / 256;
  hipLaunchKernelGGL(k_slopes_partial, dim3(nblk, B), dim3(256), 0, st, frames, cam, valid, w, quad, nblk, part);
  OOPAO_LAUNCH_CHECK();
  const int nout = 2 * w * w + nsig;
  hipLaunchKernelGGL(k_slopes_finish, dim3((nout + 255) / 256, B), dim3(256), 0, st, frames, cam, valid, w, quad,
                     ref2d, units, vidx, nsig, nvalid, nblk, part, sig2d, sig);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

// Static scratch of oopao_pwfs_slopes_static: the per-block partial sums of the valid I4Q
// (k_slopes_partial writes every entry it covers before k_slopes_finish reads it, so nothing needs
// clearing between calls). 64 Ki doubles = 512 KiB per device, held from the library's first use on a
// device until the process ends. In host code `scratch` only names the variable: its device address,
// one per device, comes from hipGetSymbolAddress.
constexpr int kSlopesScratch = 65536;
static __device__ __align__(256) double scratch[kSlopesScratch];

static int slopes_scratch(double** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* ptrs[kMaxDevices] = {};
  int dev = 0;
  OOPAO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d outside the %d devices the static scratch supports", dev, kMaxDevices);
    return ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (ptrs[dev] == nullptr) {
    void* p = nullptr;
    OOPAO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));   // with device `dev` current
    ptrs[dev] = static_cast<double*>(p);
  }
  *out = ptrs[dev];
  return OK;
}

// Occupancy of every rows-pass kernel of length N (the grid of a rows launch is the resident number of
// workgroups): queried once here so that a first launch under stream capture finds the values cached.
template <int N, int TPT, bool FWD_ONLY>
static int prime_rows_plan() {
  int slots = 0;
  int rc = resident_workgroups(reinterpret_cast<const void*>(&k_rows_sep<N, TPT, FWD_ONLY, false>), 64 * kWaves,
                               &slots);
  if (rc || FWD_ONLY) return rc;
  return resident_workgroups(reinterpret_cast<const void*>(&k_rows_sep<N, TPT, false, true>), 64 * kWaves, &slots);
}
template <int N>
static int prime_rows() {
  int rc = prime_rows_plan<N, fast_tpt_rows(N), false>();
  if (!rc) rc = prime_rows_plan<N, fast_tpt_rows(N), true>();
  if constexpr (alt_tpt_rows(N) > 0) {
    if (!rc) rc = prime_rows_plan<N, alt_tpt_rows(N), false>();
  }
  return rc;
}

}  // namespace oopao

using namespace oopao;

extern "C" {

int oopao_prepare_tables(int N) {
  OOPAO_CHECK_ARG(N > 0 && N <= 4096, "invalid transform length N=%d", N);
  const float2* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  switch (N) {
#define OOPAO_CASE(NN) \
  case NN:             \
    return prime_rows<NN>();
    OOPAO_FAST_SIZES(OOPAO_CASE)
#undef OOPAO_CASE
    default:
      return OK;
  }
}

int oopao_pwfs_workspace_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(n > 0 && N > 0 && T > 0 && B > 0 && nchunk > 0 && nchunk <= T,
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d", n, N, T, B, nchunk);
  *bytes = ws_inter_bytes(n, N, T, B) + ws_part_bytes(N, B, nchunk);
  return OK;
}

int oopao_pwfs_supported(int n, int N, int bin) { return fast_supported(n, N, bin) ? 1 : 0; }

int oopao_pwfs_plans(int N, int* rows_lanes, int* cols_lanes) {
  OOPAO_CHECK_ARG(rows_lanes && cols_lanes, "null pointer argument");
  rows_lanes[0] = rows_lanes[1] = cols_lanes[0] = cols_lanes[1] = 0;
  switch (N) {
#define OOPAO_CASE(NN)                                                                                \
  case NN:                                                                                            \
    rows_lanes[0] = fast_tpt_rows(NN);                                                                \
    rows_lanes[1] = alt_tpt_rows(NN);                                                                 \
    cols_lanes[0] = fast_tpt(NN);                                                                     \
    cols_lanes[1] = alt_tpt_cols(NN);                                                                 \
    return OK;
    OOPAO_FAST_SIZES(OOPAO_CASE)
#undef OOPAO_CASE
    default:
      break;
  }
  set_error("unsupported padded length N=%d", N);
  return ERR_UNSUPPORTED;
}

static int check_geometry(int n, int N, int bin) {
  if (!fast_supported(n, N, bin)) {
    set_error("unsupported PWFS geometry n=%d N=%d bin=%d (need N = 4 n, 4x4 binning, N in the compiled set)",
              n, N, bin);
    return ERR_UNSUPPORTED;
  }
  return OK;
}

int oopao_pwfs_intermediate_bytes(int n, int N, int T, int B, int nchunk, size_t* inter_bytes,
                                  size_t* part_bytes) {
  OOPAO_CHECK_ARG(inter_bytes && part_bytes, "null pointer argument");
  OOPAO_CHECK_ARG(n > 0 && N > 0 && T > 0 && B > 0 && nchunk > 0 && nchunk <= T,
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d", n, N, T, B, nchunk);
  *inter_bytes = ws_inter_bytes(n, N, T, B);
  *part_bytes = ws_part_bytes(N, B, nchunk);
  return OK;
}

int oopao_pwfs_rows(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                    const void* ty, const void* mx, int n, int N, int T, int B, int pts_per_task, int lanes,
                    const int* bands, void* inter, size_t inter_bytes, void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && pstart && ty && mx && inter, "null pointer argument");
  OOPAO_CHECK_ARG(P > 0 && P <= T && B > 0 && pts_per_task > 0, "invalid P=%d T=%d B=%d pts_per_task=%d", P, T, B,
                  pts_per_task);
  int rc = check_geometry(n, N, 4);
  if (rc) return rc;
  Bands bd;
  rc = read_bands(bands, N / 4, &bd);
  if (rc) return rc;
  if (inter_bytes < ws_inter_bytes(n, N, T, B)) {
    set_error("intermediate buffer too small: %zu < %zu bytes", inter_bytes, ws_inter_bytes(n, N, T, B));
    return ERR_WORKSPACE;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const float2* txp = reinterpret_cast<const float2*>(tx);
  const float2* typ = reinterpret_cast<const float2*>(ty);
  const float2* mxp = reinterpret_cast<const float2*>(mx);
  float2* ip = reinterpret_cast<float2*>(inter);
  switch (N) {
#define OOPAO_CASE(NN) \
  case NN:             \
    return launch_rows<NN>(opd, scale, amp, txp, P, pstart, typ, mxp, T, B, pts_per_task, bd, ip, lanes, st);
    OOPAO_FAST_SIZES(OOPAO_CASE)
#undef OOPAO_CASE
    default:
      break;
  }
  set_error("unsupported padded length N=%d", N);
  return ERR_UNSUPPORTED;
}

int oopao_pwfs_cols(const void* inter, size_t inter_bytes, const void* my, int n, int N, int T, int B, int nchunk,
                    int lanes, const int* bands, float* part, size_t part_bytes, float* frames, void* stream) {
  OOPAO_CHECK_ARG(inter && my && part && frames, "null pointer argument");
  OOPAO_CHECK_ARG(T > 0 && B > 0 && nchunk > 0 && nchunk <= T, "invalid T=%d B=%d nchunk=%d", T, B, nchunk);
  int rc = check_geometry(n, N, 4);
  if (rc) return rc;
  Bands bd;
  rc = read_bands(bands, N / 4, &bd);
  if (rc) return rc;
  if (inter_bytes < ws_inter_bytes(n, N, T, B) || part_bytes < ws_part_bytes(N, B, nchunk)) {
    set_error("buffers too small: intermediate %zu < %zu or partial frames %zu < %zu bytes", inter_bytes,
              ws_inter_bytes(n, N, T, B), part_bytes, ws_part_bytes(N, B, nchunk));
    return ERR_WORKSPACE;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const float2* ip = reinterpret_cast<const float2*>(inter);
  const float2* myp = reinterpret_cast<const float2*>(my);
  switch (N) {
#define OOPAO_CASE(NN) \
  case NN:             \
    return launch_cols<NN>(ip, myp, T, B, nchunk, bd, part, frames, lanes, st);
    OOPAO_FAST_SIZES(OOPAO_CASE)
#undef OOPAO_CASE
    default:
      break;
  }
  set_error("unsupported padded length N=%d", N);
  return ERR_UNSUPPORTED;
}

int oopao_pwfs_frames(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                      const void* ty, const void* mx, const void* my, int n, int N, int T, int B, int bin, int nchunk,
                      int pts_per_task, int lanes_rows, int lanes_cols, const int* bands, void* workspace,
                      size_t ws_bytes, float* frames, void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && pstart && ty && mx && my && workspace && frames, "null pointer argument");
  OOPAO_CHECK_ARG(P > 0 && P <= T && B > 0 && nchunk > 0 && nchunk <= T && pts_per_task > 0,
                  "invalid P=%d T=%d B=%d nchunk=%d pts_per_task=%d", P, T, B, nchunk, pts_per_task);
  int rc = check_geometry(n, N, bin);
  if (rc) return rc;
  const size_t ib = ws_inter_bytes(n, N, T, B), pb = ws_part_bytes(N, B, nchunk);
  if (ws_bytes < ib + pb) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, ib + pb);
    return ERR_WORKSPACE;
  }
  char* w = reinterpret_cast<char*>(workspace);
  float* part = reinterpret_cast<float*>(w + ib);
  rc = oopao_pwfs_rows(opd, scale, amp, tx, P, pstart, ty, mx, n, N, T, B, pts_per_task, lanes_rows, bands, w, ib,
                       stream);
  if (rc) return rc;
  return oopao_pwfs_cols(w, ib, my, n, N, T, B, nchunk, lanes_cols, bands, part, pb, frames, stream);
}

int oopao_pwfs_focal_bytes(int n, int N, int T, int B, int nchunk, int crop, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(n > 0 && N > 0 && T > 0 && B > 0 && nchunk > 0 && nchunk <= T && crop > 0 && crop <= N,
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d crop=%d", n, N, T, B, nchunk, crop);
  *bytes = ws_inter_bytes(n, N, T, B) + align256((size_t)B * nchunk * crop * crop * sizeof(float));
  return OK;
}

int oopao_pwfs_focal(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                     const void* ty, int n, int N, int T, int B, int nchunk, int pts_per_task, int crop,
                     void* workspace, size_t ws_bytes, float* out, void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && pstart && ty && workspace && out, "null pointer argument");
  OOPAO_CHECK_ARG(P > 0 && P <= T && B > 0 && nchunk > 0 && nchunk <= T && pts_per_task > 0 && crop > 0 &&
                      crop <= N,
                  "invalid P=%d T=%d B=%d nchunk=%d pts_per_task=%d crop=%d", P, T, B, nchunk, pts_per_task, crop);
  int rc = check_geometry(n, N, 4);
  if (rc) return rc;
  size_t need = 0;
  oopao_pwfs_focal_bytes(n, N, T, B, nchunk, crop, &need);
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const float2* txp = reinterpret_cast<const float2*>(tx);
  const float2* typ = reinterpret_cast<const float2*>(ty);
  float2* inter = reinterpret_cast<float2*>(workspace);
  float* part = reinterpret_cast<float*>(reinterpret_cast<char*>(workspace) + ws_inter_bytes(n, N, T, B));
  switch (N) {
#define OOPAO_CASE(NN)                                                                                     \
  case NN:                                                                                                 \
    rc = launch_rows<NN, true>(opd, scale, amp, txp, P, pstart, typ, txp, T, B, pts_per_task,              \
                               Bands{0, NN / 4, NN / 4, NN / 4}, inter, 0, st);                           \
    if (rc) return rc;                                                                                     \
    return launch_focal<NN>(inter, T, B, nchunk, crop, part, out, st);
    OOPAO_FAST_SIZES(OOPAO_CASE)
#undef OOPAO_CASE
    default:
      break;
  }
  set_error("unsupported padded length N=%d", N);
  return ERR_UNSUPPORTED;
}

int oopao_pwfs_raw_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(n > 0 && N > 0 && T > 0 && B > 0 && nchunk > 0 && nchunk <= T,
                  "invalid sizes n=%d N=%d T=%d B=%d nchunk=%d", n, N, T, B, nchunk);
  *bytes = ws_inter_bytes(n, N, T, B) + ws_raw_part_bytes(N, B, nchunk);
  return OK;
}

int oopao_pwfs_raw(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                   const void* ty, const void* mx, const void* my, int n, int N, int T, int B, int nchunk,
                   int pts_per_task, void* workspace, size_t ws_bytes, float* raw, void* stream) {
  OOPAO_CHECK_ARG(opd && amp && tx && pstart && ty && mx && my && workspace && raw, "null pointer argument");
  OOPAO_CHECK_ARG(P > 0 && P <= T && B > 0 && B <= 65535 && nchunk > 0 && nchunk <= T && pts_per_task > 0,
                  "invalid P=%d T=%d B=%d nchunk=%d pts_per_task=%d", P, T, B, nchunk, pts_per_task);
  int rc = check_geometry(n, N, 4);
  if (rc) return rc;
  const size_t ib = ws_inter_bytes(n, N, T, B), pb = ws_raw_part_bytes(N, B, nchunk);
  if (ws_bytes < ib + pb) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, ib + pb);
    return ERR_WORKSPACE;
  }
  char* w = reinterpret_cast<char*>(workspace);
  rc = oopao_pwfs_rows(opd, scale, amp, tx, P, pstart, ty, mx, n, N, T, B, pts_per_task, 0, nullptr, w, ib, stream);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const float2* ip = reinterpret_cast<const float2*>(w);
  const float2* myp = reinterpret_cast<const float2*>(my);
  float* part = reinterpret_cast<float*>(w + ib);
  switch (N) {
#define OOPAO_CASE(NN) \
  case NN:             \
    return launch_raw<NN>(ip, myp, T, B, nchunk, part, raw, st);
    OOPAO_FAST_SIZES(OOPAO_CASE)
#undef OOPAO_CASE
    default:
      break;
  }
  set_error("unsupported padded length N=%d", N);
  return ERR_UNSUPPORTED;
}

int oopao_bin_area(const float* raw, int B, int N, int cam, float* frames, void* stream) {
  OOPAO_CHECK_ARG(raw && frames, "null pointer argument");
  OOPAO_CHECK_ARG(B > 0 && N > 0 && cam > 0 && cam <= N, "invalid sizes B=%d N=%d cam=%d", B, N, cam);
  const long total = (long)B * cam * cam;
  hipLaunchKernelGGL(k_bin_area, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), raw, N, cam, (double)N / (double)cam, frames, B);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_pwfs_slopes_workspace_bytes(int B, int w, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes != nullptr, "bytes is null");
  OOPAO_CHECK_ARG(B > 0 && w > 0, "invalid sizes B=%d w=%d", B, w);
  *bytes = (size_t)B * ((w * w + 255) / 256) * sizeof(double);
  return OK;
}

int oopao_pwfs_slopes(const float* frames, int B, int cam, const unsigned char* valid, int w,
                      const int* quad, const float* ref2d, float units, const int* vidx, int nsig,
                      int nvalid, void* workspace, size_t ws_bytes, float* sig2d, float* sig, void* stream) {
  OOPAO_CHECK_ARG(workspace != nullptr, "null pointer argument");
  int rc = check_slopes_args(frames, B, cam, valid, w, quad, ref2d, vidx, nsig, nvalid, sig2d, sig);
  if (rc) return rc;
  const int nblk = (w * w + 255) / 256;
  if (ws_bytes < (size_t)B * nblk * sizeof(double)) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, (size_t)B * nblk * sizeof(double));
    return ERR_WORKSPACE;
  }
  return launch_slopes(frames, B, cam, valid, w, quad, ref2d, units, vidx, nsig, nvalid,
                       reinterpret_cast<double*>(workspace), sig2d, sig, reinterpret_cast<hipStream_t>(stream));
}

int oopao_pwfs_slopes_static(const float* frames, int B, int cam, const unsigned char* valid, int w,
                             const int* quad, const float* ref2d, float units, const int* vidx, int nsig,
                             int nvalid, float* sig2d, float* sig, void* stream) {
  int rc = check_slopes_args(frames, B, cam, valid, w, quad, ref2d, vidx, nsig, nvalid, sig2d, sig);
  if (rc) return rc;
  const long need = (long)B * ((w * w + 255) / 256);
  if (need > kSlopesScratch) {
    set_error("oopao_pwfs_slopes_static: %d frames of %d x %d quadrant pixels need %ld partial sums, more than "
              "the %d of the static scratch; use oopao_pwfs_slopes with a workspace", B, w, w, need,
              kSlopesScratch);
    return ERR_WORKSPACE;
  }
  double* part = nullptr;
  rc = slopes_scratch(&part);
  if (rc) return rc;
  return launch_slopes(frames, B, cam, valid, w, quad, ref2d, units, vidx, nsig, nvalid, part, sig2d, sig,
                       reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
