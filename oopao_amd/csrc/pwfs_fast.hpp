// Compile-time configuration of the separable fast PWFS path (N = 4 n, 4 x 4
// detector binning), shared by the forward kernels and their adjoint.
#pragma once
#include "common.hpp"
#include "fft_core.hpp"

namespace oopao {

// Lane count per transform for each supported padded length N (fast path: the
// pupil occupies N/4 samples and the detector bins 4 x 4). Chosen so every
// radix divides N/TPT and the pruned forward + inverse use few stages. N = 1920
// = 2^7 x 15 takes 32 lanes (60 values per lane: radices 4, 4, 4, 2, 15): its
// kernels run at one wave per SIMD, still several times faster than the general
// path at that size.
__host__ __device__ constexpr int fast_tpt(int N) {
  return N == 64 ? 16 : N == 128 ? 16 : N == 256 ? 32 : N == 320 ? 16 : N == 384 ? 16 : N == 512 ? 32
       : N == 640 ? 16 : N == 768 ? 32 : N == 1024 ? 64 : N == 1280 ? 64 : N == 1536 ? 64 : N == 1920 ? 32 : N == 2048 ? 64 : 0;
}

// Lane count of the rows pass. It may differ from the columns pass's (the intermediate layout does not
// depend on it): at N = 640, 32-lane transforms (20 values per lane, 40 KB of LDS per workgroup) run the
// rows pass 2.6x faster than 16-lane ones (40 values per lane, 80 KB), while the columns pass keeps 16.
__host__ __device__ constexpr int fast_tpt_rows(int N) { return N == 640 ? 32 : fast_tpt(N); }

// Minimum waves per SIMD requested for the columns pass: at N = 2048 its 258 VGPRs fit in 256 without
// spills when asked for two waves, which runs it 25 % faster.
__host__ __device__ constexpr int fast_cols_min_waves(int N) { return N == 2048 ? 2 : 1; }

// Second FFT plan compiled for each pass at the sensor sizes a batched calibration mixes (n = 80, 160,
// 256, 320): a different lane count per transform, hence a different radix plan, registers per lane and
// occupancy. N = 320: 32 lanes (10 values, radices 5 / 2 instead of 4 / 5 on 20; rows pass 167 VGPRs,
// three waves per SIMD instead of two); N = 640 columns: 32 lanes (20 values instead of 40; 146 VGPRs,
// three waves instead of two); N = 1024 / 1280 columns: 32 lanes (32 / 40 values, radix 8 first; two
// waves instead of three). The 32-lane rows passes at N = 1024 / 1280 need 256 VGPRs plus 116-206
// AGPRs of spill at one wave per SIMD and are not compiled. 0 where only the default is compiled.
// The calibration autotuner times both plans per geometry.
__host__ __device__ constexpr int alt_tpt_rows(int N) { return N == 320 ? 32 : 0; }
__host__ __device__ constexpr int alt_tpt_cols(int N) {
  return N == 320 ? 32 : N == 640 ? 32 : N == 1024 ? 32 : N == 1280 ? 32 : 0;
}

// Padding (complex values) after each lane group's exchange buffer in the columns pass. At N = 320 the
// four 16-lane transforms of a wavefront otherwise start their buffers on the same LDS bank; four values
// of padding take the pass from 7.9 to 7.3 us per frame (config 1, 64 screens per call), although the
// extra 0.5 KB per workgroup lowers its occupancy from four waves per SIMD to three. Other narrow sizes
// would lose whole workgroups per CU (N = 640 already uses 80 KB per workgroup) and keep 0.
__host__ __device__ constexpr int cols_group_pad(int N, int TPT) { return (N == 320 && TPT == 16) ? 4 : 0; }

// Length of the twiddle-table prefix the passes read: the pruned forward (from span 4) and the full
// transforms (the adjoint's forward ones read the same entries). The kernels copy it into LDS once per
// workgroup (every stage's twiddles are then LDS reads instead of global loads with 64-bit addresses,
// queued behind the data loads), except for the plans where the copy would lower the kernel's occupancy
// (waves per SIMD from the compiler's VGPR and LDS counts): the columns passes at N = 640 (16 and 32
// lanes) and the 32-lane one at N = 1280, the columns pass at N = 768 (3 x 54.4 KB of LDS no longer fit a
// CU once the allocation is rounded up: measured 19 % slower for whole frames), and the forward-only rows
// passes (focal plane) at N = 320, 512, 640 and 1536. Elsewhere the copy keeps or raises the occupancy (N
// = 256 columns 4 -> 5 waves per SIMD, N = 640 rows 2 -> 3, N = 1536 rows 1 -> 2) and lowers the spills at
// N = 1920 / 2048. Frames of 8 screens, slopes-only bands: N = 256 72 -> 56 us, 320 117 -> 86 us, 512 244
// -> 188 us, 1536 3.51 -> 2.58 ms, 2048 8.0 -> 7.1 ms (config 1, 500 screens: 4.89 -> 3.53 ms slopes-only,
// 8.72 -> 5.97 ms whole); N = 1024 and 1280 (64 lanes) had the copy already.
// The first K twiddle-table entries into LDS, by the NT threads of a workgroup, every load issued before
// the first LDS store: the plain strided loop waited for each load before storing it, one round trip to
// memory per iteration at the start of every workgroup.
template <int K, int NT>
__device__ __forceinline__ void copy_twiddles(float2* dst, const float2* __restrict__ src) {
  constexpr int kIt = (K + NT - 1) / NT;
  float2 v[kIt];
#pragma unroll
  for (int j = 0; j < kIt; ++j) {
    const int e = (int)threadIdx.x + j * NT;
    v[j] = e < K ? src[e] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int j = 0; j < kIt; ++j) {
    const int e = (int)threadIdx.x + j * NT;
    if (e < K) dst[e] = v[j];
  }
}

template <int N, int TPT>
__host__ __device__ constexpr int sep_tw_len() {
  return (plan_tw_max(N, N / TPT, 4, N / 4) > plan_tw_max(N, N / TPT, 1, N) ? plan_tw_max(N, N / TPT, 4, N / 4)
                                                                               : plan_tw_max(N, N / TPT, 1, N)) +
         1;
}
__host__ __device__ constexpr bool tw_lds_cols(int N, int TPT) {
  return !(N == 640 || N == 768 || (N == 1280 && TPT == 32));
}
__host__ __device__ constexpr bool tw_lds_rows(int N, bool fwd_only) {
  return !(fwd_only && (N == 320 || N == 512 || N == 640 || N == 1536));
}

// Buffer instructions for the per-element accesses of the column kernels (column values, mask values,
// and in the adjoint the detector gradient and the Z stores): one VGPR of lane offset and compile-time
// uniform offsets instead of a 64-bit address per element (k_cols_sep at N = 1280: 1697 -> 1643 VALU
// instructions per (point, column) pair; the adjoint's columns kernel at N = 1920: 512 VGPRs and 47
// spilled -> 238, none spilled). The compiler schedules the two forms differently, and the gain follows
// the length, so each kernel takes them where they measured faster (frames of 1 and 8 screens; gradients
// of 1-8 screens): k_cols_sep N = 640 2-6 %, 1280 2-3 %, 1536 6-7 %, 1920 3-4 % (1024 unchanged; 320,
// 384, 768 and 2048 slower by 2-10 %); k_cols_raw N = 640 10 %, 1920 4 %; k_cols_focal N = 512 2 %,
// 768 12 %, 1920 2 %; k_bwd_cols N = 512 2 %, 768 1 %, 1280 1 %, 1920 8 %.
__host__ __device__ constexpr bool buf_cols_sep(int N) {
  return N == 640 || N == 1024 || N == 1280 || N == 1536 || N == 1920;
}
__host__ __device__ constexpr bool buf_cols_raw(int N) { return N == 640 || N == 1920; }
__host__ __device__ constexpr bool buf_cols_focal(int N) { return N == 512 || N == 768 || N == 1920; }
__host__ __device__ constexpr bool buf_bwd_cols(int N) { return N == 512 || N == 768 || N == 1280 || N == 1920; }

// Rows-pass plans that form the pupil field amp exp(i scale opd) themselves, per tip table, instead of
// reading it from a k_field launch: one launch (~4.6 us) fewer per call, for the sincos of every field
// value once per tip table (37 tables at config 2). Staged plans only ((N / 4) % TPT == 0), and where it
// measured faster: frames of 1 and 8 screens at N = 128-768 2-14 %, focal frames at N <= 768 2-9 %; config
// 2 (N = 1280) single screen 4958 -> 4999 frames/s (rows pass 58.8 -> 57.2 us) but 8 screens per call
// 0.8 % slower, focal frames unchanged; N = 1024-2048 unchanged to 1 % slower (the table count grows
// with the modulation points while the launch saved stays the same).
__host__ __device__ constexpr bool rows_form_field(int N, int TPT, bool fwd_only) {
  return (N / 4) % TPT == 0 && (N <= 768 || (N == 1280 && !fwd_only));
}

#define OOPAO_FAST_SIZES(X) X(64) X(128) X(256) X(320) X(384) X(512) X(640) X(768) X(1024) X(1280) X(1536) X(1920) X(2048)

constexpr int kWaves = 4;  // wavefronts per workgroup

// Detector columns a frame computes: [a0, a1) and [b0, b1), 0 <= a0 <= a1 <= b0 <= b1 <= cam. The
// slopesMaps signal reads only the quadrant bands of the camera (Pyramid.py:838-853, 946-961: columns
// [c-e-w, c-e) and [c+e, c+e+w)), so a slopes-only frame transforms those detector columns (4 pupil-
// plane columns each) and no others; the full frame is the band [0, cam).
struct Bands {
  int a0, a1, b0, b1;
};
__device__ __forceinline__ bool in_bands(int g, Bands bd) {
  return (g >= bd.a0 && g < bd.a1) || (g >= bd.b0 && g < bd.b1);
}
// the gi-th computed detector column
__device__ __forceinline__ int band_col(int gi, Bands bd) {
  const int na = bd.a1 - bd.a0;
  return gi < na ? bd.a0 + gi : bd.b0 + (gi - na);
}

// Column bands from the C-ABI argument: null = every detector column.
static inline int read_bands(const int* bands, int cam, Bands* out) {
  if (bands == nullptr) {
    *out = Bands{0, cam, cam, cam};
    return OK;
  }
  const Bands b{bands[0], bands[1], bands[2], bands[3]};
  OOPAO_CHECK_ARG(0 <= b.a0 && b.a0 <= b.a1 && b.a1 <= b.b0 && b.b0 <= b.b1 && b.b1 <= cam &&
                      (b.a1 - b.a0) + (b.b1 - b.b0) > 0,
                  "invalid column bands [%d, %d) [%d, %d) for %d detector columns (need 0 <= a0 <= a1 <= b0 <= b1 "
                  "<= cam and at least one column)", b.a0, b.a1, b.b0, b.b1, cam);
  *out = b;
  return OK;
}


}  // namespace oopao
