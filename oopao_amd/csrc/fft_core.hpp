// Register-resident Stockham FFT building blocks for CDNA4 (gfx950).
//
// One transform of length N is owned by TPT lanes of a single wavefront
// (TPT <= 64), each lane holding E = N / TPT complex values in registers in the
// "strided" layout: lane t holds element t + m*TPT in v[m], m = 0..E-1.
// A Stockham stage of radix R does E/R butterflies per lane entirely from its
// own registers (butterfly j = t + b*TPT reads elements j + r*N/R, which are
// v[b + r*E/R]), writes the R outputs to their self-sorted positions in the
// wavefront's LDS exchange buffer, and reads back the strided layout. Because a
// transform never spans two wavefronts, the exchange needs only wavefront-level
// ordering (LDS executes one wavefront's instructions in issue order), not a
// workgroup barrier.
//
// Twiddles come from a per-length table tw[m] = exp(-2*pi*i*m/N) computed on
// the host in double precision (accurate to float rounding).
#pragma once
#include <hip/hip_runtime.h>

namespace oopao {

// Wavefront-scope ordering of LDS accesses between the write and read halves
// of an exchange: a compiler code-motion barrier plus wavefront-scope fences.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float cabs2(float2 a) { return fmaf(a.x, a.x, a.y * a.y); }

// a * (i * DIR): DIR = -1 (forward) multiplies by -i, DIR = +1 by +i.
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// In-register DFTs of radix 2, 3, 4, 5, 8 (natural order in and out).
// DIR = -1: X[k] = sum x[j] exp(-2 pi i jk/R); DIR = +1: exp(+...), unscaled.
// ---------------------------------------------------------------------------
template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]);
    float2 t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]);
    float2 t3 = mul_i<DIR>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    const float s3 = 0.86602540378443864676f;  // sin(2 pi / 3)
    float2 t1 = cadd(a[1], a[2]);
    float2 m = make_float2(fmaf(-0.5f, t1.x, a[0].x), fmaf(-0.5f, t1.y, a[0].y));
    float2 d = cscale(mul_i<DIR>(csub(a[1], a[2])), s3);
    a[0] = cadd(a[0], t1);
    a[1] = cadd(m, d);
    a[2] = csub(m, d);
  }
};

template <int DIR>
struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
    const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
    const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
    const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
    float2 b1 = cadd(a[1], a[4]);
    float2 b2 = cadd(a[2], a[3]);
    float2 d1 = csub(a[1], a[4]);
    float2 d2 = csub(a[2], a[3]);
    float2 a0 = a[0];
    float2 m1 = make_float2(fmaf(c2, b2.x, fmaf(c1, b1.x, a0.x)), fmaf(c2, b2.y, fmaf(c1, b1.y, a0.y)));
    float2 m2 = make_float2(fmaf(c1, b2.x, fmaf(c2, b1.x, a0.x)), fmaf(c1, b2.y, fmaf(c2, b1.y, a0.y)));
    float2 n1 = make_float2(fmaf(s2, d2.x, s1 * d1.x), fmaf(s2, d2.y, s1 * d1.y));
    float2 n2 = make_float2(fmaf(-s1, d2.x, s2 * d1.x), fmaf(-s1, d2.y, s2 * d1.y));
    float2 i1 = mul_i<DIR>(n1);
    float2 i2 = mul_i<DIR>(n2);
    a[0] = cadd(a0, cadd(b1, b2));
    a[1] = cadd(m1, i1);
    a[4] = csub(m1, i1);
    a[2] = cadd(m2, i2);
    a[3] = csub(m2, i2);
  }
};

template <int DIR>
struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    const float r2 = 0.70710678118654752440f;
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, DIR>::run(e);
    Dft<4, DIR>::run(o);
    // o[k] *= exp(DIR * i * pi * k / 4)
    float2 o1 = o[1], o3 = o[3];
    if (DIR < 0) {
      o[1] = make_float2((o1.x + o1.y) * r2, (o1.y - o1.x) * r2);
      o[3] = make_float2((o3.y - o3.x) * r2, -(o3.x + o3.y) * r2);
    } else {
      o[1] = make_float2((o1.x - o1.y) * r2, (o1.x + o1.y) * r2);
      o[3] = make_float2(-(o3.x + o3.y) * r2, (o3.x - o3.y) * r2);
    }
    o[2] = mul_i<DIR>(o[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = cadd(e[k], o[k]);
      a[k + 4] = csub(e[k], o[k]);
    }
  }
};

// Prime-factor (Good-Thomas) DFT of R = R1 R2 points with gcd(R1, R2) = 1: input n = (R2 n1 + R1 n2) mod R,
// output k = (R2 u1 k1 + R1 u2 k2) mod R with u1 = R2^-1 mod R1, u2 = R1^-1 mod R2 (so k1 = k mod R1,
// k2 = k mod R2), hence X = DFT_R1 over n1 of DFT_R2 over n2 with no twiddles between the two (every
// index is a compile-time register name). 20 = 4 x 5: maps 5 n1 + 4 n2 and 5 k1 + 16 k2.
__host__ __device__ constexpr int mod_inverse(int a, int m) {
  int x = 1;
  while (x < m && (a * x) % m != 1) ++x;
  return x;
}
template <int R1, int R2, int DIR>
struct Pfa {
  __device__ __forceinline__ static void run(float2* a) {
    constexpr int R = R1 * R2;
    constexpr int u1 = mod_inverse(R2 % R1, R1), u2 = mod_inverse(R1 % R2, R2);
    float2 y[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) y[n1][n2] = a[(R2 * n1 + R1 * n2) % R];
      Dft<R2, DIR>::run(y[n1]);
    }
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      float2 z[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) z[n1] = y[n1][k2];
      Dft<R1, DIR>::run(z);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) a[(R2 * u1 * k1 + R1 * u2 * k2) % R] = z[k1];
    }
  }
};
template <int DIR>
struct Dft<20, DIR> : Pfa<4, 5, DIR> {};
template <int DIR>
struct Dft<15, DIR> : Pfa<3, 5, DIR> {};
template <int DIR>
struct Dft<24, DIR> : Pfa<8, 3, DIR> {};

// ---------------------------------------------------------------------------
// Compile-time radix plans. Every radix must divide E so that each lane owns
// whole butterflies. Larger radices first: fewer LDS exchanges. When a lane holds
// a multiple of 20 values and the length is 20 times a product of radices
// 8 / 4 / 2, those come first and the last stage is one radix-20 prime-factor
// stage held in one lane: as the last stage it needs no exchange, and it has no
// internal twiddles (N = 1280 with 20 values per lane: 4, 4, 4, 20 instead of
// 4, 4, 4, 4, 5; N = 640 with 20 / 40: 4, 4, 2, 20 / 8, 4, 20 instead of
// 4, 4, 4, 5, 2 / 8, 8, 5, 2). Likewise radix-24 (8 x 3) and radix-15 (3 x 5)
// last stages: N = 384 / 768 / 1536 with 24 values per lane run 8, 2, 24 /
// 8, 4, 24 / 8, 8, 24 instead of 8, 8, 3, 2 / 8, 8, 4, 3 / 8, 8, 8, 3 (frames of
// 8 screens 17 % / 0-4 % / 12 % faster; the focal-plane kernels at 768 and 1536
// 4-11 % slower), and N = 1920 with 60 values per lane runs 4, 4, 4, 2, 15
// instead of 4, 4, 4, 5, 3, 2 (frames 3-4 % faster, the columns pass's spills
// gone). A radix-60
// last stage for N = 1920 was measured: 60 values and their twiddles in one lane
// spilled to scratch.
// ---------------------------------------------------------------------------
// (Measured on gfx950 at N = 1280: plans with the odd radices first run 2x slower, because spans 5,
// 20, 80, 320 do not divide the 64-lane count and every twiddle becomes a per-butterfly load.)
__host__ __device__ constexpr int pow2_radix(int x, int E) {
  return (x % 8 == 0 && E % 8 == 0) ? 8 : (x % 4 == 0 && E % 4 == 0) ? 4 : (x % 2 == 0 && E % 2 == 0) ? 2 : 0;
}
__host__ __device__ constexpr bool pow2_plan(int x, int E) {
  return x == 1 ? true : (pow2_radix(x, E) == 0 ? false : pow2_plan(x / pow2_radix(x, E), E));
}
// the prime-factor radix the plan of rem can end with (0: none)
__host__ __device__ constexpr int pfa_last(int rem, int E) {
  return (E % 20 == 0 && rem % 20 == 0 && pow2_plan(rem / 20, E))   ? 20
         : (E % 24 == 0 && rem % 24 == 0 && pow2_plan(rem / 24, E)) ? 24
         : (E % 15 == 0 && rem % 15 == 0 && pow2_plan(rem / 15, E)) ? 15
                                                                     : 0;
}
__host__ __device__ constexpr int pick_radix(int rem, int E) {
  return pfa_last(rem, E) != 0 ? (rem == pfa_last(rem, E) ? rem : pow2_radix(rem / pfa_last(rem, E), E))
       : (rem % 8 == 0 && E % 8 == 0) ? 8
       : (rem % 4 == 0 && E % 4 == 0) ? 4
       : (rem % 5 == 0 && E % 5 == 0) ? 5
       : (rem % 3 == 0 && E % 3 == 0) ? 3
       : (rem % 2 == 0 && E % 2 == 0) ? 2
       : 0;
}

__host__ __device__ constexpr bool plan_ok(int rem, int E) {
  return rem == 1 ? true : (pick_radix(rem, E) == 0 ? false : plan_ok(rem / pick_radix(rem, E), E));
}

__host__ __device__ constexpr int plan_stages(int rem, int E) {
  return rem == 1 ? 0 : 1 + plan_stages(rem / pick_radix(rem, E), E);
}

// Largest twiddle index read by the stages from span NS on (REM = N / NS still to transform): a stage
// of radix R and span NS > 1 reads tw[r k N / (NS R)] for r < R, k < NS.
__host__ __device__ constexpr int plan_tw_max(int N, int E, int NS, int REM) {
  int mx = 0;
  while (REM > 1) {
    const int R = pick_radix(REM, E);
    if (R == 0) return N - 1;
    if (NS > 1 && (R - 1) * (NS - 1) * (N / (NS * R)) > mx) mx = (R - 1) * (NS - 1) * (N / (NS * R));
    NS *= R;
    REM /= R;
  }
  return mx;
}

// Exchange buffers are unpadded: on gfx950 every LDS address then folds into the instruction's
// immediate offset, which saved registers (occupancy 2 -> 4 waves/SIMD for the 1280-point columns
// pass) and ran 10-30 % faster than one float2 of padding every 32.
__host__ __device__ constexpr int lds_len(int N) { return N; }

// XOR swizzle of the exchange-buffer index for 64-lane transforms: the low 4
// bits (the 16 eight-byte bank groups of one 128-byte LDS row) are XORed with
// 5 * (row mod 4). Reads of the strided layout t + 64m stay conflict-free and
// the span-1 and span-4 radix-4 writes (4-way conflicted unswizzled) become
// conflict-free; the XOR term depends only on the lane, so per-butterfly
// offsets (multiples of 64) remain immediate offsets.
__device__ __forceinline__ int lds_swz(int a) {
  const int rw = (a >> 4) & 3;
  return a ^ (rw | (rw << 2));
}

// Opaque copy of a read-only table pointer. Loads through the returned pointer
// cannot be hoisted above this point, which keeps loop-invariant twiddle and
// mask loads next to their use instead of pinning registers across loops.
// (The opaque value is an integer offset, so the pointer keeps its global
// address space and the loads stay global_load rather than flat_load.)
template <typename T>
__device__ __forceinline__ const T* launder(const T* p) {
  int zero = 0;
  asm volatile("" : "+s"(zero)::"memory");
  return p + zero;
}

// Value of lane (lane & ~3) + c for every lane of each quad (DPP quad_perm
// broadcast; c must be uniform across the wavefront).
__device__ __forceinline__ float dpp_quad_bcast(float x, int c) {
  const int v = __float_as_int(x);
  int r;
  switch (c) {
    case 0: r = __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, false); break;
    case 1: r = __builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, false); break;
    case 2: r = __builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, false); break;
    default: r = __builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, false); break;
  }
  return __int_as_float(r);
}
__device__ __forceinline__ float2 quad_bcast(float2 x, int c) {
  return make_float2(dpp_quad_bcast(x.x, c), dpp_quad_bcast(x.y, c));
}

// Raw buffer access (gfx9 family descriptor: word 3 = 0x00020000, untyped dword data). A buffer store
// takes a per-lane byte offset in a VGPR and a wave-uniform offset in an SGPR, so a sequence of stores
// at constant strides from one lane base needs one address register instead of a 64-bit pointer per
// store. Accesses at or beyond `bytes` are dropped by the hardware.
typedef __amdgpu_buffer_rsrc_t BufferRsrc;
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ BufferRsrc buffer_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void buffer_store_c64(float2 v, BufferRsrc r, unsigned lane_off, int uniform_off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)lane_off, uniform_off, 0);
}
__device__ __forceinline__ float2 buffer_load_c64(BufferRsrc r, unsigned lane_off, int uniform_off) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)lane_off, uniform_off, 0));
}
__device__ __forceinline__ float buffer_load_f32(BufferRsrc r, unsigned lane_off, int uniform_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_off, uniform_off, 0));
}

// One Stockham stage (radix R, already-sorted span NS) on the strided layout.
// When NS divides TPT the twiddle index k = j mod NS equals t mod NS for all of
// the lane's butterflies, so R-1 twiddles are loaded once per stage.
// The last stage (NS R = N) needs no exchange: butterfly j < N/R writes its
// outputs to j + r N/R, the positions its inputs came from, and in the strided
// layout those are the same lane's registers v[b + r E/R].
template <int N, int TPT, int R, int NS, int DIR, bool SWZ = false>
__device__ __forceinline__ void stockham_stage(float2* v, float2* lds, const float2* __restrict__ tw, int t) {
  constexpr bool kLast = (NS * R == N);
  constexpr int E = N / TPT;
  constexpr int NB = E / R;
  constexpr bool kUniform = (TPT % NS) == 0;
  constexpr int TS = N / (NS * R);
  static_assert(E % R == 0, "radix must divide the per-lane element count");
  const float2* twp = launder(tw);
  // XOR-swizzled exchange slots for 64-lane transforms (see lds_swz); the
  // lane-dependent part is computed once per stage, butterflies add constants.
  constexpr bool kSwz = SWZ && (TPT == 64) && ((64 % NS == 0) || (NS % 64 == 0));
  int p0[R];
  if constexpr (kSwz) {
    const int base0 = (t / NS) * (NS * R) + (t % NS);
#pragma unroll
    for (int r = 0; r < R; ++r) p0[r] = lds_swz(base0 + r * NS);
  }
  // a stage with one butterfly per lane loads each twiddle where it is used (no array of R - 1 held)
  constexpr bool kPreload = kUniform && NB > 1;
  float2 wu[R];
  if (kPreload && NS > 1) {
    const int k = t % NS;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = twp[r * k * TS];
      if (DIR > 0) w.y = -w.y;
      wu[r] = w;
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TPT;
    const int k = kUniform ? (t % NS) : (j % NS);
    float2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[b + r * NB];
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w;
        if (kPreload) {
          w = wu[r];
        } else {
          w = twp[r * k * TS];
          if (DIR > 0) w.y = -w.y;
        }
        x[r] = cmul(x[r], w);
      }
    }
    Dft<R, DIR>::run(x);
    if constexpr (kLast) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * NB] = x[r];
    } else if constexpr (kSwz) {
      // swizzled slot of butterfly b = slot of butterfly 0 + a multiple of 64
      const int off = ((64 * b) / NS) * (NS * R) + (64 * b) % NS;   // folds to a constant
#pragma unroll
      for (int r = 0; r < R; ++r) lds[p0[r] + off] = x[r];
    } else {
      const int base = (j / NS) * (NS * R) + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[base + r * NS] = x[r];
    }
  }
  if constexpr (kLast) return;
  wave_lds_sync();
  if constexpr (kSwz) {
    const int q = lds_swz(t);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = lds[q + 64 * m];
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = lds[t + m * TPT];
  }
  wave_lds_sync();
}

// Remaining stages from span NS (REM = N / NS still to transform).
template <int N, int TPT, int DIR, int NS, int REM, bool SWZ = false>
struct Stages {
  static constexpr int E = N / TPT;
  static constexpr int R = pick_radix(REM, E);
  static_assert(R != 0, "length not supported by the radix set for this lane count");
  __device__ __forceinline__ static void run(float2* v, float2* lds, const float2* __restrict__ tw, int t) {
    stockham_stage<N, TPT, R, NS, DIR, SWZ>(v, lds, tw, t);
    Stages<N, TPT, DIR, NS * R, REM / R, SWZ>::run(v, lds, tw, t);
  }
};
template <int N, int TPT, int DIR, int NS, bool SWZ>
struct Stages<N, TPT, DIR, NS, 1, SWZ> {
  __device__ __forceinline__ static void run(float2*, float2*, const float2* __restrict__, int) {}
};

// Full transform of the strided-layout registers (natural order in and out).
// SWZ selects the XOR-swizzled exchange (64-lane transforms only).
template <int N, int TPT, int DIR, bool SWZ = false>
__device__ __forceinline__ void fft_full(float2* v, float2* lds, const float2* __restrict__ tw, int t) {
  Stages<N, TPT, DIR, 1, N, SWZ>::run(v, lds, tw, t);
}

// Forward transform of an input that is non-zero only on [0, N/P): the first
// radix-P stage of such an input replicates x[q/P] into position q, so the
// caller loads v[m] = x[(t + m*TPT) / P] and the remaining stages start at
// span P.
template <int N, int TPT, int P, bool SWZ = false>
__device__ __forceinline__ void fft_fwd_pruned(float2* v, float2* lds, const float2* __restrict__ tw, int t) {
  Stages<N, TPT, -1, P, N / P, SWZ>::run(v, lds, tw, t);
}

}  // namespace oopao
