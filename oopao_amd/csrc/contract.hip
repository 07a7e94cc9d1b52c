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
// k = contraction index) padded to 32*IT x 16*KT, the fragment (kt, it) holds
// 64 lanes x 8 values, lane l carrying Q[32 it + l%32][16 kt + 8 (l/32) + j],
// j = 0..7, and fragments are stored kt-major: element
//     ((kt * IT + it) * 64 + l) * 8 + j.
// A fragment is one 1 KB contiguous block, so every global load is one 16-byte
// load per lane, fully coalesced, and an LDS copy of it is read back conflict-free.
// Split-BF16: Q = hi + lo with hi = bf16(Q), lo = bf16(Q - hi) (round to
// nearest even); the product uses hi*hi + hi*lo + lo*hi. BF16 keeps 8
// significant bits (unit roundoff 2^-8), so the split keeps Q to 2^-16 and the
// dropped lo*lo term is at most about 2^-16 |A||X|.
//
// Kernels:
//  * k_pack_bf16   : FP32 [I x K] (any strides) -> hi (and lo) fragments.
//  * k_gemm_bf16   : 512 threads (8 wavefronts, 4 x 2), 256 b x 256 r output tile,
//                    each wavefront 2 x 4 MFMA 32x32 blocks; fragments staged in
//                    LDS, two buffers, the next stage's global loads in flight
//                    during the current stage's MFMAs; workgroup order remapped so
//                    that the workgroups sharing an A slab run on one XCD.
//  * k_gemv_bf16   : single right-hand side (HBM-bound): one wavefront per
//                    (row fragment, 32-fragment K chunk), 1 KB coalesced loads,
//                    fixed-order partial sums, then k_gemv_finish.
#include "common.hpp"

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

// out[b * ldo + r] = alpha * sum_k A[r][k] X[k][b]; A fragments (IR x KT), X fragments (IB x KT).
// SPLIT = 1: hi*hi; SPLIT = 2: hi*hi + hi*lo + lo*hi. KS k fragments per LDS stage.
template <int SPLIT, int KS>
__global__ void __launch_bounds__(kGemmThreads) k_gemm_bf16(const uint4* __restrict__ ah, const uint4* __restrict__ al,
                                                            const uint4* __restrict__ xh, const uint4* __restrict__ xl,
                                                            int IR, int IB, int KT, float* __restrict__ out,
                                                            long long ldo, int R, int B, float alpha, int nbt,
                                                            int nrt) {
  // [buffer][term][k fragment][tile fragment: 8 of X then 8 of A][lane]
  __shared__ uint4 lds[2][SPLIT][KS][kTF][64];
  constexpr int kPer = 2 * SPLIT * KS;
  static_assert(kTB == kGemmThreads / 64 && kTR == kTB, "one X and one A fragment per wavefront and stage slot");

  // XCD-aware order: workgroup i runs on XCD i % 8; give each XCD a contiguous run of
  // tiles (b fastest) so the workgroups that read one A slab share that XCD's L2.
  const int G = nbt * nrt;
  const int i = blockIdx.x;
  const int per_xcd = (G + 7) / 8;
  const int tall = (G % 8) ? (G % 8) : 8;
  const int xcd = i & 7, slot = i >> 3;
  const int t = (xcd < tall) ? xcd * per_xcd + slot : tall * per_xcd + (xcd - tall) * (per_xcd - 1) + slot;
  const int rt = t / nbt, bt = t % nbt;

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int wb = w & 3, wr = w >> 2;  // wavefront owns b fragments 2wb..2wb+1, r fragments 4wr..4wr+3

  const int nst = (KT + KS - 1) / KS;
  // Stage chunk q of this thread (kPer = 2 SPLIT KS): wavefront w copies fragment w of X
  // (q even) and fragment w of A (q odd) for term q / (2 KS) and k fragment (q / 2) % KS.
  const int itx = bt * kTB + w, ita = rt * kTR + w;
  const bool okx = itx < IB, oka = ita < IR;
  const uint4* srcx[SPLIT];
  const uint4* srca[SPLIT];
  srcx[0] = xh + (size_t)itx * 64 + lane;
  srca[0] = ah + (size_t)ita * 64 + lane;
  if constexpr (SPLIT == 2) {
    srcx[1] = xl + (size_t)itx * 64 + lane;
    srca[1] = al + (size_t)ita * 64 + lane;
  }
  uint4 pre[kPer];
#define OOPAO_STAGE_LOAD(st_)                                                               \
  _Pragma("unroll") for (int q = 0; q < kPer; ++q) {                                        \
    const int s_ = q / (2 * KS), kk_ = (q >> 1) % KS;                                       \
    const int kt_ = (st_) * KS + kk_;                                                       \
    const bool isx_ = (q & 1) == 0;                                                         \
    const bool ok_ = kt_ < KT && (isx_ ? okx : oka);                                        \
    const size_t off_ = (size_t)kt_ * 64 * (isx_ ? IB : IR);                                \
    pre[q] = ok_ ? (isx_ ? srcx[s_] : srca[s_])[off_] : make_uint4(0u, 0u, 0u, 0u);         \
  }
#define OOPAO_STAGE_STORE(buf_)                                                             \
  _Pragma("unroll") for (int q = 0; q < kPer; ++q) {                                        \
    const int s_ = q / (2 * KS), kk_ = (q >> 1) % KS;                                       \
    lds[buf_][s_][kk_][(q & 1) ? kTB + w : w][lane] = pre[q];                               \
  }

  f32x16 acc[2][4];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nr = 0; nr < 4; ++nr)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[mb][nr][e] = 0.f;

  OOPAO_STAGE_LOAD(0)
  OOPAO_STAGE_STORE(0)
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    if (st + 1 < nst) { OOPAO_STAGE_LOAD(st + 1) }
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
      bf16x8 xh8[2], ah8[4];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) xh8[mb] = as_bf16x8(lds[buf][0][kk][2 * wb + mb][lane]);
#pragma unroll
      for (int nr = 0; nr < 4; ++nr) ah8[nr] = as_bf16x8(lds[buf][0][kk][kTB + 4 * wr + nr][lane]);
      if constexpr (SPLIT == 2) {
        bf16x8 xl8[2], al8[4];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) xl8[mb] = as_bf16x8(lds[buf][SPLIT - 1][kk][2 * wb + mb][lane]);
#pragma unroll
        for (int nr = 0; nr < 4; ++nr) al8[nr] = as_bf16x8(lds[buf][SPLIT - 1][kk][kTB + 4 * wr + nr][lane]);
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
          for (int nr = 0; nr < 4; ++nr) {
            acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl8[mb], ah8[nr], acc[mb][nr], 0, 0, 0);
            acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh8[mb], al8[nr], acc[mb][nr], 0, 0, 0);
          }
      }
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nr = 0; nr < 4; ++nr)
          acc[mb][nr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh8[mb], ah8[nr], acc[mb][nr], 0, 0, 0);
    }
    if (st + 1 < nst) { OOPAO_STAGE_STORE(buf ^ 1) }
    __syncthreads();
  }

#undef OOPAO_STAGE_LOAD
#undef OOPAO_STAGE_STORE
  // D lane l, element e: row (b) 8 (e/4) + 4 (l/32) + e%4, column (r) l%32
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
        const int b = b_base + mb * 32 + 8 * (e >> 2) + (e & 3);
        if (b < B) out[(long long)b * ldo + r] = alpha * acc[mb][nr][e];
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
  for (int kt = kt0; kt < kt1; ++kt) {
    const uint4 v = ah[((size_t)kt * IR + it) * 64 + lane];
    const int k = kt * 16 + koff;
    float xv[8];
    if (k + 8 <= K) {
      const float4 p = *reinterpret_cast<const float4*>(x + k);
      const float4 q = *reinterpret_cast<const float4*>(x + k + 4);
      xv[0] = p.x; xv[1] = p.y; xv[2] = p.z; xv[3] = p.w;
      xv[4] = q.x; xv[5] = q.y; xv[6] = q.z; xv[7] = q.w;
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) xv[j] = (k + j < K) ? x[k + j] : 0.f;
    }
    const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0 = fmaf(__uint_as_float(u[j] << 16), xv[2 * j], a0);
      a1 = fmaf(__uint_as_float(u[j] & 0xFFFF0000u), xv[2 * j + 1], a1);
    }
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
  float s = 0.f;
  for (int c = 0; c < nkc; ++c) s += part[(size_t)c * Rp + r];
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

int gemv_chunks(int KT) { return (KT + kKC - 1) / kKC; }

}  // namespace

extern "C" {

int oopao_bf16_packed_bytes(int I, int K, size_t* bytes) {
  OOPAO_CHECK_ARG(bytes, "null pointer argument");
  int IT, KT;
  OOPAO_CHECK_ARG(frag_counts(I, K, &IT, &KT) == OK, "invalid sizes I=%d K=%d", I, K);
  *bytes = (size_t)IT * KT * 1024;
  return OK;
}

int oopao_bf16_pack(const float* src, long long si, long long sk, int I, int K, void* hi, void* lo, void* stream) {
  OOPAO_CHECK_ARG(src && hi, "null pointer argument");
  int IT, KT;
  OOPAO_CHECK_ARG(frag_counts(I, K, &IT, &KT) == OK, "invalid sizes I=%d K=%d", I, K);
  OOPAO_CHECK_ARG((reinterpret_cast<size_t>(hi) & 15) == 0 && (reinterpret_cast<size_t>(lo) & 15) == 0,
                  "packed buffers must be 16-byte aligned");
  const size_t total = (size_t)IT * KT * 64;
  hipLaunchKernelGGL(k_pack_bf16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), src, si, sk, I, K, IT, KT,
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
    hipLaunchKernelGGL((k_gemm_bf16<1, 4>), grid, dim3(kGemmThreads), 0, st, ah, al, xh, xl, IR, IB, KT, out, ldo,
                       R, B, alpha, nbt, nrt);
  else
    hipLaunchKernelGGL((k_gemm_bf16<2, 2>), grid, dim3(kGemmThreads), 0, st, ah, al, xh, xl, IR, IB, KT, out, ldo,
                       R, B, alpha, nbt, nrt);
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
