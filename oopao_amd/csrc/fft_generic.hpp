// Runtime-length Stockham FFT in LDS, one transform per wavefront.
//
// Used by the general PWFS path, where the padded length N is whatever the
// sensor geometry gives (N = (2 nSubap + separation + 2 edge) * n / nSubap,
// e.g. 264 = 8*3*11 for OOPAO's tutorial sensor). The transform lives in two
// LDS buffers of N complex values (ping-pong); each Stockham stage of radix R
// computes N/R butterflies spread over the 64 lanes:
//   butterfly j (k = j mod Ns): v[r] = in[j + r N/R] * W_{Ns R}^{r k},
//   v = DFT_R(v), out[(j - k) R + k + r Ns] = v[r]
// (the self-sorting form, output in natural order after the last stage).
// Radices 8, 4, 2, 3, 5 use the hand-written butterflies, 7, 11, 13 a
// register-resident direct DFT, any other prime factor a direct DFT that reads
// its inputs from LDS. Twiddles come from the shared table
// tw[m] = exp(-2 pi i m / N) (double-precision values rounded once).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.hpp"

namespace oopao {

constexpr int kGenMaxStages = 24;

struct GenPlan {
  int N;
  int nst;
  int R[kGenMaxStages];
};

// W_N^m for direction DIR (conjugate of the table entry for the inverse)
template <int DIR>
__device__ __forceinline__ float2 gen_tw(const float2* tw, int m) {
  float2 w = tw[m];
  return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// Register-resident direct DFT of a small prime R (7, 11, 13): O(R^2).
template <int R, int DIR>
__device__ __forceinline__ void dft_direct(float2* v, const float2* tw, int N) {
  float2 o[R];
  const int step = N / R;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    float2 acc = v[0];
#pragma unroll
    for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], gen_tw<DIR>(tw, ((r * q) % R) * step)));
    o[q] = acc;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = o[q];
}

template <int R, int DIR>
__device__ __forceinline__ void gen_stage(const float2* in, float2* out, const float2* tw, int N, int Ns,
                                          int lane) {
  const int NR = N / R;
  const int step = N / (Ns * R);
  for (int j = lane; j < NR; j += 64) {
    const int k = j % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[j + r * NR];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], gen_tw<DIR>(tw, r * k * step));
    }
    if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
      Dft<R, DIR>::run(v);
    } else {
      dft_direct<R, DIR>(v, tw, N);
    }
    const int o = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[o + r * Ns] = v[r];
  }
}

// Any radix: outputs formed one at a time from LDS (no register array).
template <int DIR>
__device__ __forceinline__ void gen_stage_any(const float2* in, float2* out, const float2* tw, int N, int R,
                                              int Ns, int lane) {
  const int NR = N / R;
  const int step = N / (Ns * R);
  const int rstep = N / R;
  const int total = NR * R;
  for (int e = lane; e < total; e += 64) {
    const int j = e / R;
    const int q = e - j * R;
    const int k = j % Ns;
    float2 acc = make_float2(0.f, 0.f);
    int rq = 0;  // (r * q) mod R
    for (int r = 0; r < R; ++r) {
      float2 x = in[j + r * NR];
      if (r > 0) {
        if (Ns > 1) x = cmul(x, gen_tw<DIR>(tw, r * k * step));
        x = cmul(x, gen_tw<DIR>(tw, rq * rstep));
      }
      acc = cadd(acc, x);
      rq += q;
      if (rq >= R) rq -= R;
    }
    out[(j - k) * R + k + q * Ns] = acc;
  }
}

// Transforms the N values in a (LDS) using b as scratch; returns the buffer
// holding the result. Must be called by all 64 lanes of the wavefront.
template <int DIR>
__device__ float2* gen_fft(float2* a, float2* b, const float2* tw, const GenPlan& p, int lane) {
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.R[s];
    switch (R) {
      case 2: gen_stage<2, DIR>(a, b, tw, p.N, Ns, lane); break;
      case 3: gen_stage<3, DIR>(a, b, tw, p.N, Ns, lane); break;
      case 4: gen_stage<4, DIR>(a, b, tw, p.N, Ns, lane); break;
      case 5: gen_stage<5, DIR>(a, b, tw, p.N, Ns, lane); break;
      case 7: gen_stage<7, DIR>(a, b, tw, p.N, Ns, lane); break;
      case 8: gen_stage<8, DIR>(a, b, tw, p.N, Ns, lane); break;
      case 11: gen_stage<11, DIR>(a, b, tw, p.N, Ns, lane); break;
      case 13: gen_stage<13, DIR>(a, b, tw, p.N, Ns, lane); break;
      default: gen_stage_any<DIR>(a, b, tw, p.N, R, Ns, lane); break;
    }
    wave_lds_sync();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

}  // namespace oopao
