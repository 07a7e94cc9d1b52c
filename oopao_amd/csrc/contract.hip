// Dense influence-function and reconstructor contractions in BF16 on the CDNA4
// matrix cores, FP32 accumulation, optional split-BF16 compensation.
//
// Reference: the DM coefs setter computes OPD = modes @ coefs
// (DeformableMirror.py:651-686; modes [n^2, nAct], coefs [nAct] or [nAct, B]),
// and the closed loop applies the reconstructor as reconstructor @ wfs.signal
// (tutorials/AO_closed_loop_Pyramid_WFS.py:620-624). Both are
//     out[b][r] = alpha * sum_k A[r][k] X[k][b]
// with A the resident matrix (modes, reconstructor) and X the right-hand sides.
//
// Storage: A and X are converted once to BF16 in the MFMA fragment layout of
// v_mfma_f32_32x32x16_bf16. For a matrix Q[i][k] (i = row of A, or column b of X;
// k = contraction index), IT = ceil(I/32), KT = ceil(K/16), the fragment (kt, it)
// holds 64 lanes x 8 values, lane l carrying Q[32 it + l%32][16 kt + 8 (l/32) + j],
// j = 0..7, and fragments are stored kt-major: element
//     ((kt * IT + it) * 64 + l) * 8 + j,
// for kt < KTs = 8 ceil(KT/8) + 8; everything outside Q is zero (the extra k
// fragments let a GEMM run whole groups of pipeline stages and prefetch past the end
// without bounds checks). A fragment is one 1 KB contiguous block: one 16-byte load
// per lane, fully coalesced, and it lands in LDS exactly as the DMA path writes it
// (lane l's 16 bytes at base + 16 l), to be read back conflict-free.
// Split-BF16: Q = hi + lo with hi = bf16(Q), lo = bf16(Q - hi) (round to
// nearest even); the product uses hi*hi + hi*lo + lo*hi. BF16 keeps 8
// significant bits (unit roundoff 2^-8), so the split keeps Q to 2^-16 and the
// dropped lo*lo term is at most about 2^-16 |A||X|.
//
// Kernels (GEMM output tile 256 b x 256 r per workgroup; fragments streamed by
// global_load_lds_dwordx4 into a ring of LDS stages, one k fragment per stage, held in
// distinct LDS arrays so the compiler knows which DMA writes which slot; workgroup
// order remapped so that the workgroups reading one A slab run on one XCD):
//  * k_pack_bf16      : FP32 [I x K] (any strides) -> hi (and lo) fragments.
//  * k_gemm_bf16_pipe : BF16 (hi*hi). 512 threads, 8 wavefronts (4 x 2), each 2 x 4
//                       MFMA 32x32 blocks, ring of 8 stages, the next stage's fragment
//                       reads interleaved with the current stage's MFMAs.
//  * k_gemm_bf16_big  : split BF16 (3 products). 256 threads, 4 wavefronts (2 x 2),
//                       each 4 x 4 blocks (256 accumulators in AGPRs, one wavefront per
//                       SIMD): 16 fragments read per 48 MFMAs; ring of 4 stages.
//  * k_gemv_bf16      : single right-hand side (HBM-bound): one wavefront per
//                       (row fragment, 32-fragment K chunk), 1 KB coalesced loads,
//                       fixed-order partial sums, then k_gemv_finish.
// Accumulation order over k is the same in every kernel, so a product does not depend
// on the tile shape that computed it.
#include "common.hpp"

#include <type_traits>

namespace oopao {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTB = 8;   // b fragments per workgroup tile (256 columns of X)
constexpr int kTR = 8;   // r fragments per workgroup tile (256 rows of A)
constexpr int kTF = kTB + kTR;
constexpr int kGemmThreads = 512;
constexpr int kKC = 32;  // k fragments per wavefront in the GEMV

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

// s_waitcnt immediate (gfx9 encoding) that waits for vmcnt <= n and leaves expcnt/lgkmcnt free
constexpr int vmcnt_imm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) {
  return __builtin_bit_cast(bf16x8, v);
}

}  // namespace

// One thread per 8-value group of a fragment: Q[i][k] = src[i * si + k * sk].
__global__ void __launch_bounds__(256) k_pack_bf16(const float* __restrict__ src, long long si, long long sk, int I,
                                                   int K, int IT, int KT, uint4* __restrict__ hi,
                                                   uint4* __restrict__ lo) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)KT * IT * 64;
  if (g >= total) return;
  const int l = (int)(g & 63);
  const size_t f = g >> 6;
  const int it = (int)(f % IT), kt = (int)(f / IT);
  const int i = it * 32 + (l & 31);
  const int k0 = kt * 16 + (l >> 5) * 8;
  unsigned h[8], lw[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = k0 + j;
    const float v = (i < I && k < K) ? src[(long long)i * si + (long long)k * sk] : 0.f;
    h[j] = bf16_rne(v);
    const float r = v - __uint_as_float(h[j] << 16);
    lw[j] = bf16_rne(r);
  }
  hi[g] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
  if (lo) lo[g] = make_uint4(lw[0] | (lw[1] << 16), lw[2] | (lw[3] << 16), lw[4] | (lw[5] << 16),
                             lw[6] | (lw[7] << 16));
}

// Same product with the LDS ring held in NBUF distinct LDS arrays (one per slot) and the
// main loop unrolled over the ring, so every slot is a compile-time object: the compiler's
// alias tracking then knows which DMA writes which slot and adds no wait for a read of
// another slot. Fragment reads are ordinary LDS loads issued one stage ahead (two register
// sets), so the reads of stage k + 1 overlap the MFMAs of stage k.
template <int SPLIT, int NBUF>
__global__ void __launch_bounds__(kGemmThreads) k_gemm_bf16_pipe(const uint4* __restrict__ ah,
                                                                 const uint4* __restrict__ al,
                                                                 const uint4* __restrict__ xh,
                                                                 const uint4* __restrict__ xl, int IR, int IB, int KT,
                                                                 float* __restrict__ out, long long ldo, int R, int B,
                                                                 float alpha, int nbt, int nrt) {
  constexpr int kSlot = SPLIT * kTF * 64;  // uint4 per slot
  __shared__ uint4 S0[kSlot], S1[kSlot], S2[kSlot], S3[kSlot], S4[kSlot], S5[kSlot], S6[kSlot], S7[kSlot];
  static_assert(NBUF == 4 || NBUF == 8, "ring of 4 or 8 slots");
  static_assert(SPLIT * NBUF * kTF * 1024 <= 131072, "ring fits in 128 KB of LDS");
  constexpr int kLoads = 2 * SPLIT;
  constexpr int kFr = 6 * SPLIT;  // fragments read per wavefront and stage
  static_assert(kTB == kGemmThreads / 64 && kTR == kTB, "one X and one A fragment per wavefront and stage");

  const int G = nbt * nrt;
  const int i = blockIdx.x;
  const int per_xcd = (G + 7) / 8;
  const int tall = (G % 8) ? (G % 8) : 8;
  const int xcd = i & 7, slotid = i >> 3;
  const int t = (xcd < tall) ? xcd * per_xcd + slotid : tall * per_xcd + (xcd - tall) * (per_xcd - 1) + slotid;
  const int rt = t / nbt, bt = t % nbt;

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int wb = w & 3, wr = w >> 2;

  const int itx = min(bt * kTB + w, IB - 1), ita = min(rt * kTR + w, IR - 1);
  const uint4* gx0 = xh + (size_t)itx * 64 + lane;
  const uint4* ga0 = ah + (size_t)ita * 64 + lane;
  const uint4* gx1 = (SPLIT == 2) ? xl + (size_t)itx * 64 + lane : nullptr;
  const uint4* ga1 = (SPLIT == 2) ? al + (size_t)ita * 64 + lane : nullptr;
  const size_t sx = (size_t)IB * 64, sa = (size_t)IR * 64;

  uint4* const slots[8] = {S0, S1, S2, S3, S4, S5, S6, S7};

  // stage layout inside a slot: [term][fragment: 8 of X then 8 of A][lane]
  auto issue = [&](int kt, uint4* dst) {
    __builtin_amdgcn_global_load_lds(gx0 + kt * sx, (__attribute__((address_space(3))) void*)(dst + w * 64), 16, 0,
                                     0);
    __builtin_amdgcn_global_load_lds(ga0 + kt * sa,
                                     (__attribute__((address_space(3))) void*)(dst + (kTB + w) * 64), 16, 0, 0);
    if constexpr (SPLIT == 2) {
      __builtin_amdgcn_global_load_lds(gx1 + kt * sx,
                                       (__attribute__((address_space(3))) void*)(dst + (kTF + w) * 64), 16, 0, 0);
      __builtin_amdgcn_global_load_lds(ga1 + kt * sa,
                                       (__attribute__((address_space(3))) void*)(dst + (kTF + kTB + w) * 64), 16, 0,
                                       0);
    }
  };
  // fragments of one stage: [term][2 of X, 4 of A]
  auto read = [&](const uint4* src, uint4 (&f)[kFr]) {
#pragma unroll
    for (int s2 = 0; s2 < SPLIT; ++s2) {
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) f[s2 * 6 + mb] = src[(s2 * kTF + 2 * wb + mb) * 64 + lane];
#pragma unroll
      for (int nr = 0; nr < 4; ++nr) f[s2 * 6 + 2 + nr] = src[(s2 * kTF + kTB + 4 * wr + nr) * 64 + lane];
    }
  };

  f32x16 acc[2][4];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nr = 0; nr < 4; ++nr)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[mb][nr][e] = 0.f;

  auto mfma = [&](const uint4 (&f)[kFr]) {
    if constexpr (SPLIT == 2) {
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nr = 0; nr < 4; ++nr) {
          acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f[6 + mb]), as_bf16x8(f[2 + nr]),
                                                                acc[mb][nr], 0, 0, 0);
          acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f[mb]), as_bf16x8(f[6 + 2 + nr]),
                                                                acc[mb][nr], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nr = 0; nr < 4; ++nr)
        acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f[mb]), as_bf16x8(f[2 + nr]), acc[mb][nr],
                                                              0, 0, 0);
  };

  // stages run: KT rounded up to whole groups of NBUF (the stored operands hold zero
  // fragments beyond KT, and at least NBUF - 2 more for the last prefetches)
  const int KTr = ((KT + NBUF - 1) / NBUF) * NBUF;
#pragma unroll
  for (int u = 0; u < NBUF - 1; ++u) issue(u, slots[u]);
  __builtin_amdgcn_s_waitcnt(vmcnt_imm((NBUF - 2) * kLoads));
  asm volatile("s_barrier" ::: "memory");
  uint4 F0[kFr], F1[kFr];
  read(slots[0], F0);

  // one fragment read after each of the first kFr MFMAs of a stage, then the rest
  auto interleave = [&]() {
    constexpr int kM = 8 * (SPLIT == 2 ? 3 : 1);
#pragma unroll
    for (int q = 0; q < kFr; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    if constexpr (kM > kFr) __builtin_amdgcn_sched_group_barrier(0x008, kM - kFr, 0);
  };
  // One stage: wait for this wavefront's DMA of stage kt + 1, barrier (every wavefront's
  // stage kt + 1 landed, stage kt - 1 fully consumed), refill the slot of stage kt - 1 with
  // stage kt + NBUF - 1, read stage kt + 1 into the other register set, MFMAs of stage kt.
  auto step = [&](int kt, auto U) {
    constexpr int u = decltype(U)::value;
    __builtin_amdgcn_s_waitcnt(vmcnt_imm((NBUF - 3) * kLoads));
    asm volatile("s_barrier" ::: "memory");
    issue(kt + NBUF - 1, slots[(u + NBUF - 1) % NBUF]);
    if constexpr (u & 1) {
      read(slots[(u + 1) % NBUF], F0);
      mfma(F1);
    } else {
      read(slots[(u + 1) % NBUF], F1);
      mfma(F0);
    }
    interleave();
  };
  for (int kt0 = 0; kt0 < KTr; kt0 += NBUF) {
    step(kt0 + 0, std::integral_constant<int, 0>());
    step(kt0 + 1, std::integral_constant<int, 1>());
    step(kt0 + 2, std::integral_constant<int, 2>());
    step(kt0 + 3, std::integral_constant<int, 3>());
    if constexpr (NBUF == 8) {
      step(kt0 + 4, std::integral_constant<int, 4>());
      step(kt0 + 5, std::integral_constant<int, 5>());
      step(kt0 + 6, std::integral_constant<int, 6>());
      step(kt0 + 7, std::integral_constant<int, 7>());
    }
  }
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // the last prefetches land before the wavefront ends

  const int r_base = rt * (kTR * 32) + (4 * wr) * 32 + (lane & 31);
  const int b_base = bt * (kTB * 32) + (2 * wb) * 32 + 4 * (lane >> 5);
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nr = 0; nr < 4; ++nr) {
      const int r = r_base + nr * 32;
      if (r >= R) continue;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int bb = b_base + mb * 32 + 8 * (e >> 2) + (e & 3);
        if (bb < B) out[(long long)bb * ldo + r] = alpha * acc[mb][nr][e];
      }
    }
}

// 256-thread variant: 4 wavefronts (2 x 2), each owning a 4 x 4 block of 32 x 32 output
// tiles (256 accumulators, one wavefront per SIMD), so each wavefront reads 8 fragments per
// 16 MFMAs; the same 256 x 256 workgroup tile, DMA ring in NBUF distinct LDS arrays, and the
// next stage's fragment reads interleaved with the current stage's MFMAs (two register sets).
constexpr int kBigThreads = 256;
template <int SPLIT, int NBUF>
__global__ void __launch_bounds__(kBigThreads) k_gemm_bf16_big(const uint4* __restrict__ ah,
                                                               const uint4* __restrict__ al,
                                                               const uint4* __restrict__ xh,
                                                               const uint4* __restrict__ xl, int IR, int IB, int KT,
                                                               float* __restrict__ out, long long ldo, int R, int B,
                                                               float alpha, int nbt, int nrt) {
  constexpr int kSlot = SPLIT * kTF * 64;
  __shared__ uint4 S0[kSlot], S1[kSlot], S2[kSlot], S3[kSlot], S4[kSlot], S5[kSlot], S6[kSlot], S7[kSlot];
  static_assert(NBUF == 4 || NBUF == 8, "ring of 4 or 8 slots");
  static_assert(SPLIT * NBUF * kTF * 1024 <= 131072, "ring fits in 128 KB of LDS");
  constexpr int kLoads = 4 * SPLIT;  // DMA instructions per wavefront and stage
  constexpr int kFr = 8 * SPLIT;     // fragments read per wavefront and stage
  static_assert((NBUF - 2) * kLoads <= 63, "vmcnt range");

  const int G = nbt * nrt;
  const int i = blockIdx.x;
  const int per_xcd = (G + 7) / 8;
  const int tall = (G % 8) ? (G % 8) : 8;
  const int xcd = i & 7, slotid = i >> 3;
  const int t = (xcd < tall) ? xcd * per_xcd + slotid : tall * per_xcd + (xcd - tall) * (per_xcd - 1) + slotid;
  const int rt = t / nbt, bt = t % nbt;

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int wb = w & 1, wr = w >> 1;  // b fragments 4wb..4wb+3, r fragments 4wr..4wr+3

  // DMA: wavefront w copies X fragments 2w, 2w+1 and A fragments 2w, 2w+1 of each stage
  const uint4* gx[2][2];
  const uint4* ga[2][2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int itx = min(bt * kTB + 2 * w + j, IB - 1), ita = min(rt * kTR + 2 * w + j, IR - 1);
    gx[0][j] = xh + (size_t)itx * 64 + lane;
    ga[0][j] = ah + (size_t)ita * 64 + lane;
    gx[1][j] = (SPLIT == 2) ? xl + (size_t)itx * 64 + lane : nullptr;
    ga[1][j] = (SPLIT == 2) ? al + (size_t)ita * 64 + lane : nullptr;
  }
  const size_t sx = (size_t)IB * 64, sa = (size_t)IR * 64;
  uint4* const slots[8] = {S0, S1, S2, S3, S4, S5, S6, S7};

  auto issue = [&](int kt, uint4* dst) {
#pragma unroll
    for (int s2 = 0; s2 < SPLIT; ++s2)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        __builtin_amdgcn_global_load_lds(gx[s2][j] + kt * sx,
                                         (__attribute__((address_space(3))) void*)(dst + (s2 * kTF + 2 * w + j) * 64),
                                         16, 0, 0);
        __builtin_amdgcn_global_load_lds(
            ga[s2][j] + kt * sa, (__attribute__((address_space(3))) void*)(dst + (s2 * kTF + kTB + 2 * w + j) * 64), 16,
            0, 0);
      }
  };
  // F[s2 * 8 + m] = X fragment 4wb + m (m < 4), A fragment 4wr + m - 4 (m >= 4)
  auto read = [&](const uint4* src, uint4 (&f)[kFr]) {
#pragma unroll
    for (int s2 = 0; s2 < SPLIT; ++s2)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        f[s2 * 8 + m] = src[(s2 * kTF + 4 * wb + m) * 64 + lane];
        f[s2 * 8 + 4 + m] = src[(s2 * kTF + kTB + 4 * wr + m) * 64 + lane];
      }
  };

  f32x16 acc[4][4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nr = 0; nr < 4; ++nr)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[mb][nr][e] = 0.f;

  auto mfma = [&](const uint4 (&f)[kFr]) {
    if constexpr (SPLIT == 2) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nr = 0; nr < 4; ++nr) {
          acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f[8 + mb]), as_bf16x8(f[4 + nr]),
                                                                acc[mb][nr], 0, 0, 0);
          acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f[mb]), as_bf16x8(f[8 + 4 + nr]),
                                                                acc[mb][nr], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int nr = 0; nr < 4; ++nr)
        acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f[mb]), as_bf16x8(f[4 + nr]), acc[mb][nr], 0,
                                                              0, 0);
  };
  // schedule: one fragment read after each of the first kFr MFMAs of the stage, then the
  // rest of the MFMAs, so the next stage's fragments have landed before its barrier
  auto interleave = [&]() {
    constexpr int kM = 16 * (SPLIT == 2 ? 3 : 1);
#pragma unroll
    for (int q = 0; q < kFr; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, kM - kFr, 0);
  };

  // stages run: KT rounded up to whole groups of NBUF (the stored operands hold zero
  // fragments beyond KT, and at least NBUF - 2 more for the last prefetches)
  const int KTr = ((KT + NBUF - 1) / NBUF) * NBUF;
#pragma unroll
  for (int u = 0; u < NBUF - 1; ++u) issue(u, slots[u]);
  __builtin_amdgcn_s_waitcnt(vmcnt_imm((NBUF - 2) * kLoads));
  asm volatile("s_barrier" ::: "memory");
  uint4 F0[kFr], F1[kFr];
  read(slots[0], F0);

  // One stage: wait for this wavefront's DMA of stage kt + 1 (NBUF - 3 later stages may
  // still be in flight), barrier (every wavefront's stage kt + 1 landed, stage kt - 1 fully
  // consumed), refill the slot of stage kt - 1 with stage kt + NBUF - 1, read stage kt + 1
  // into the other register set, MFMAs of stage kt.
  auto step = [&](int kt, auto U) {
    constexpr int u = decltype(U)::value;
    __builtin_amdgcn_s_waitcnt(vmcnt_imm((NBUF - 3) * kLoads));
    asm volatile("s_barrier" ::: "memory");
    issue(kt + NBUF - 1, slots[(u + NBUF - 1) % NBUF]);
    if constexpr (u & 1) {
      read(slots[(u + 1) % NBUF], F0);
      mfma(F1);
    } else {
      read(slots[(u + 1) % NBUF], F1);
      mfma(F0);
    }
    interleave();
  };
  for (int kt0 = 0; kt0 < KTr; kt0 += NBUF) {
    step(kt0 + 0, std::integral_constant<int, 0>());
    step(kt0 + 1, std::integral_constant<int, 1>());
    step(kt0 + 2, std::integral_constant<int, 2>());
    step(kt0 + 3, std::integral_constant<int, 3>());
    if constexpr (NBUF == 8) {
      step(kt0 + 4, std::integral_constant<int, 4>());
      step(kt0 + 5, std::integral_constant<int, 5>());
      step(kt0 + 6, std::integral_constant<int, 6>());
      step(kt0 + 7, std::integral_constant<int, 7>());
    }
  }
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // the last prefetches land before the wavefront ends

  const int r_base = rt * (kTR * 32) + (4 * wr) * 32 + (lane & 31);
  const int b_base = bt * (kTB * 32) + (4 * wb) * 32 + 4 * (lane >> 5);
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nr = 0; nr < 4; ++nr) {
      const int r = r_base + nr * 32;
      if (r >= R) continue;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int bb = b_base + mb * 32 + 8 * (e >> 2) + (e & 3);
        if (bb < B) out[(long long)bb * ldo + r] = alpha * acc[mb][nr][e];
      }
    }
}

// Partial GEMV over packed hi fragments: part[kc][r] = sum over the chunk's k of A[r][k] x[k].
// Workgroup = 4 wavefronts = 4 consecutive K chunks of one row fragment.
__global__ void __launch_bounds__(256) k_gemv_bf16(const uint4* __restrict__ ah, int IR, int KT,
                                                   const float* __restrict__ x, int K, float* __restrict__ part,
                                                   int Rp) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int it = blockIdx.x;
  const int kc = blockIdx.y * 4 + w;
  const int kt0 = kc * kKC, kt1 = min(KT, kt0 + kKC);
  float a0 = 0.f, a1 = 0.f;
  const int koff = (lane >> 5) * 8;
  // x[k .. k+8) for fragment kt (zero beyond K)
  auto xload = [&](int kt, float (&xv)[8]) {
    const int k = kt * 16 + koff;
    if (k + 8 <= K) {
      const float4 p = *reinterpret_cast<const float4*>(x + k);
      const float4 q = *reinterpret_cast<const float4*>(x + k + 4);
      xv[0] = p.x; xv[1] = p.y; xv[2] = p.z; xv[3] = p.w;
      xv[4] = q.x; xv[5] = q.y; xv[6] = q.z; xv[7] = q.w;
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) xv[j] = (k + j < K) ? x[k + j] : 0.f;
    }
  };
  auto fma8 = [&](const uint4 v, const float (&xv)[8]) {
    const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0 = fmaf(__uint_as_float(u[j] << 16), xv[2 * j], a0);
      a1 = fmaf(__uint_as_float(u[j] & 0xFFFF0000u), xv[2 * j + 1], a1);
    }
  };
  const uint4* base = ah + (size_t)it * 64 + lane;
  const size_t stride = (size_t)IR * 64;
  int kt = kt0;
  // four 1 KB fragment loads in flight per wavefront
  for (; kt + 4 <= kt1; kt += 4) {
    uint4 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = base[(size_t)(kt + q) * stride];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float xv[8];
      xload(kt + q, xv);
      fma8(v[q], xv);
    }
  }
  for (; kt < kt1; ++kt) {
    float xv[8];
    xload(kt, xv);
    fma8(base[(size_t)kt * stride], xv);
  }
  float s = a0 + a1;
  s += __shfl_xor(s, 32);
  if (lane < 32 && kc * kKC < KT) part[(size_t)kc * Rp + it * 32 + lane] = s;
}

__global__ void __launch_bounds__(256) k_gemv_finish(const float* __restrict__ part, int nkc, int Rp, int R,
                                                     float* __restrict__ y, const float* __restrict__ y0,
                                                     float alpha, float beta) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const float s = sum_in_order(part + r, (size_t)Rp, nkc);
  float v = alpha * s;
  if (y0) v = fmaf(beta, y0[r], v);
  y[r] = v;
}

}  // namespace oopao

using namespace oopao;

namespace {

int frag_counts(int I, int K, int* IT, int* KT) {
  if (I <= 0 || K <= 0) return ERR_ARG;
  *IT = (I + 31) / 32;
  *KT = (K + 15) / 16;
  return OK;
}

// k fragments stored per operand: the KT that hold data, rounded up to a multiple of 8, plus
// 8 zero fragments, so a GEMM ring of up to 8 stages can run whole groups of stages and
// prefetch past the end without a bounds check
int stored_kt(int KT) { return ((KT + 7) / 8) * 8 + 8; }

int gemv_chunks(int KT) { return (KT + kKC - 1) / kKC; }

}  // namespace

extern "C" {

int oopao_bf16_packed_bytes(int I, int K, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes, "null pointer argument");
  int IT, KT;
  OOPAO_CHECK_ARG(frag_counts(I, K, &IT, &KT) == OK, "invalid sizes I=%d K=%d", I, K);
  *bytes = (size_t)IT * stored_kt(KT) * 1024;
  return OK;
}

int oopao_bf16_pack(const float* src, long long si, long long sk, int I, int K, void* hi, void* lo, void* stream) {
  OOPAO_CHECK_ARG(src && hi, "null pointer argument");
  int IT, KT;
  OOPAO_CHECK_ARG(frag_counts(I, K, &IT, &KT) == OK, "invalid sizes I=%d K=%d", I, K);
  OOPAO_CHECK_ARG((reinterpret_cast<size_t>(hi) & 15) == 0 && (reinterpret_cast<size_t>(lo) & 15) == 0,
                  "packed buffers must be 16-byte aligned");
  const int KTs = stored_kt(KT);
  const size_t total = (size_t)IT * KTs * 64;
  hipLaunchKernelGGL(k_pack_bf16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), src, si, sk, I, K, IT, KTs,
                     reinterpret_cast<uint4*>(hi), reinterpret_cast<uint4*>(lo));
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_gemm_bf16(const void* a_hi, const void* a_lo, int R, int K, const void* x_hi, const void* x_lo, int B,
                    float* out, long long ldo, float alpha, int terms, void* stream) {
  OOPAO_CHECK_ARG(a_hi && x_hi && out, "null pointer argument");
  OOPAO_CHECK_ARG(terms == 1 || terms == 3, "terms must be 1 (bf16) or 3 (split bf16), got %d", terms);
  OOPAO_CHECK_ARG(terms == 1 || (a_lo && x_lo), "split bf16 needs the lo parts of both operands");
  OOPAO_CHECK_ARG(R > 0 && K > 0 && B > 0 && ldo >= R, "invalid sizes R=%d K=%d B=%d ldo=%lld", R, K, B, ldo);
  int IR, KT, IB, KT2;
  frag_counts(R, K, &IR, &KT);
  frag_counts(B, K, &IB, &KT2);
  const int nrt = (IR + kTR - 1) / kTR, nbt = (IB + kTB - 1) / kTB;
  OOPAO_CHECK_ARG((long long)nrt * nbt < (1LL << 31), "output too large");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)(nrt * nbt));
  const uint4* ah = reinterpret_cast<const uint4*>(a_hi);
  const uint4* al = reinterpret_cast<const uint4*>(a_lo);
  const uint4* xh = reinterpret_cast<const uint4*>(x_hi);
  const uint4* xl = reinterpret_cast<const uint4*>(x_lo);
  if (terms == 1)
    hipLaunchKernelGGL((k_gemm_bf16_pipe<1, 8>), grid, dim3(kGemmThreads), 0, st, ah, al, xh, xl, IR, IB, KT, out,
                       ldo, R, B, alpha, nbt, nrt);
  else
    hipLaunchKernelGGL((k_gemm_bf16_big<2, 4>), grid, dim3(kBigThreads), 0, st, ah, al, xh, xl, IR, IB, KT, out,
                       ldo, R, B, alpha, nbt, nrt);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

int oopao_gemv_bf16_workspace_bytes(int R, int K, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes, "null pointer argument");
  int IR, KT;
  OOPAO_CHECK_ARG(frag_counts(R, K, &IR, &KT) == OK, "invalid sizes R=%d K=%d", R, K);
  *bytes = (size_t)gemv_chunks(KT) * IR * 32 * sizeof(float);
  return OK;
}

int oopao_gemv_bf16(const void* a_hi, int R, int K, const float* x, float* y, const float* y0, float alpha,
                    float beta, void* ws, size_t ws_bytes, void* stream) {
  OOPAO_CHECK_ARG(a_hi && x && y && ws, "null pointer argument");
  int IR, KT;
  OOPAO_CHECK_ARG(frag_counts(R, K, &IR, &KT) == OK, "invalid sizes R=%d K=%d", R, K);
  OOPAO_CHECK_ARG((reinterpret_cast<size_t>(x) & 15) == 0, "x must be 16-byte aligned");
  const int nkc = gemv_chunks(KT);
  const size_t need = (size_t)nkc * IR * 32 * sizeof(float);
  if (ws_bytes < need) {
    set_error("workspace too small: %zu < %zu bytes", ws_bytes, need);
    return ERR_WORKSPACE;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float* part = reinterpret_cast<float*>(ws);
  hipLaunchKernelGGL(k_gemv_bf16, dim3((unsigned)IR, (unsigned)((nkc + 3) / 4)), dim3(256), 0, st,
                     reinterpret_cast<const uint4*>(a_hi), IR, KT, x, K, part, IR * 32);
  OOPAO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gemv_finish, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, st, part, nkc, IR * 32, R, y,
                     y0, alpha, beta);
  OOPAO_LAUNCH_CHECK();
  return OK;
}

}  // extern "C"
