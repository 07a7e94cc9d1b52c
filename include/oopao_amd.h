/*
 * oopao_amd C-ABI: MI355X (gfx950) kernels behind OOPAO's optical-propagation hot path.
 *
 * OOPAO is pure Python; its only "FFI" for this path is the duck-typed relay
 * protocol (Source.py:176-184, Telescope.py:524-546): each optical object has a
 * tag and relay(src). The Python drop-in classes in oopao_amd/ keep that protocol
 * and reach the GPU through the entry points below (ctypes, oopao_amd/_lib.py).
 * Each entry point names the reference code it replaces.
 *
 * Conventions
 *  - All array pointers are device pointers (HIP), caller-allocated; the library
 *    allocates nothing per call (a per-(device, N) twiddle table is built once,
 *    ahead of use with oopao_prepare_tables).
 *  - Complex arrays are interleaved float pairs (complex64).
 *  - Every call is enqueued on `stream` (a hipStream_t, NULL = default stream)
 *    and returns immediately.
 *  - Return value 0 = success; otherwise an error code (1 invalid argument,
 *    2 unsupported geometry, 3 HIP error, 4 workspace too small) and
 *    oopao_last_error() describes it (thread-local).
 */
#ifndef OOPAO_AMD_H
#define OOPAO_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message for the last failed call on this thread. */
const char* oopao_last_error(void);

/* Build the device twiddle table exp(-2 pi i m / N) of transform length N on the
 * current device (once per device and N; kept until the process ends), and for
 * the separable PWFS lengths query the occupancy of the rows-pass kernels (their
 * grid is the resident number of workgroups; the values are cached). The
 * transform entry points do both on their first call for a new N, the table with
 * a synchronous allocation and copy, which a stream under graph capture does not
 * allow: call this first for every N a captured call will use. */
int oopao_prepare_tables(int N);

/* ---------------------------------------------------------------------------
 * Pyramid WFS (replaces Pyramid.pyramid_transform + the modulation loop and
 * detector binning of Pyramid.wfs_measure: Pyramid.py:612-646, 666-757,
 * 1178-1203; tools/tools.py:464-483)
 * ------------------------------------------------------------------------- */

/* 1 if the separable fast path is compiled for this geometry: N = 4 n,
 * 4 x 4 detector binning, N in {64,128,256,320,384,512,640,768,1024,1280,1536,1920,2048}. */
int oopao_pwfs_supported(int n, int N, int bin);

/* FFT plans compiled for padded length N, as lanes per transform (each lane
 * count fixes the radix plan, the values held per lane and the occupancy):
 * rows_lanes[0] / cols_lanes[0] are the default plans of the rows and columns
 * passes, rows_lanes[1] / cols_lanes[1] a second plan (0 when none is compiled;
 * rows: N = 320; columns: N = 320, 640, 1024, 1280). Both arrays hold 2 ints.
 * No counterpart in the reference: the plan is a tuning choice of this
 * library's transforms, picked per geometry by the calibration autotuner. */
int oopao_pwfs_plans(int N, int* rows_lanes, int* cols_lanes);

/* Workspace needed by oopao_pwfs_frames for B screens of T modulation points
 * reduced in `nchunk` point chunks. */
int oopao_pwfs_workspace_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes);

/* Modulated PWFS detector frames for B phase screens (Pyramid.py:612-646, 712-826).
 *   opd    [B][n][n] float32; phase = opd * scale (src.phase = OPD 2 pi / lambda,
 *          Source.py:294-296)
 *   amp    [n][n]    float32 field amplitude sqrt(src.intensity / nTheta) (Pyramid.py:674)
 *   tx     [P][n]    complex64 distinct tip (x) phasors times the PSF-centring phasor
 *   pstart [P+1]     int32: the points using tip table p are pstart[p] .. pstart[p+1]-1
 *                    (0 = pstart[0] <= ... <= pstart[P] = T; values are clamped to that)
 *   ty     [T][n]    complex64 per-point tilt (y) phasors times the PSF-centring
 *                    phasor (Pyramid.py:334-344, 1129-1139), in the same point order:
 *                    point k's field is u(y, x) tx[p][x] ty[k][y]
 *   mx,my  [N]       complex64 separable factors of the pyramid mask, with the
 *                    zero-padding offset shift and the 1/N of each inverse FFT
 *                    folded in (Pyramid.py:388, 453-480)
 *   frames [B][N/bin][N/bin] float32: sum over points of |ifft2(fft2(E) mask)|^2,
 *          sum-binned (tools/tools.py:464-483)
 * pts_per_task: the most tip tables one rows-pass workgroup transforms per
 * pupil row (the pass launches the resident number of workgroups, times the
 * rounds this cap needs, with equal ranges of work); nchunk: point chunks of
 * the columns pass. Both are tuning knobs: the rows pass writes the same
 * intermediate for any value, nchunk changes only the float summation order
 * over points.
 * bands: null for the whole frame, or a host array {a0, a1, b0, b1} of detector
 * columns (0 <= a0 <= a1 <= b0 <= b1 <= N/4): only the columns in [a0, a1) and
 * [b0, b1) are computed (bit for bit as in the whole frame), the others are
 * written as zero. The slopesMaps signal reads only the quadrant columns
 * [c-e-w, c-e) and [c+e, c+e+w) (Pyramid.py:838-853, 946-961), so a
 * slopes-only measurement passes those and does half the work at config 2.
 * lanes_rows / lanes_cols: lanes per transform of the rows / columns pass, one
 * of the plans oopao_pwfs_plans lists, or 0 for the default (ERR_UNSUPPORTED
 * for a plan not compiled at this N); every plan computes the same frame up to
 * float rounding. */
int oopao_pwfs_frames(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                      const void* ty, const void* mx, const void* my, int n, int N, int T, int B, int bin,
                      int nchunk, int pts_per_task, int lanes_rows, int lanes_cols, const int* bands,
                      void* workspace, size_t ws_bytes, float* frames, void* stream);

/* The two passes of oopao_pwfs_frames as separate calls (same arguments), for
 * callers that own the intermediate buffers or time the passes:
 *   rows pass : inter[B*T][N/4][n][4] complex64 (per pupil row y and point k of
 *               tip table p: ty_k[y] IFFT_N(mx . FFT_N(pad(amp exp(i scale opd[y,:]) tx_p))),
 *               one transform per table, written once per point of its group)
 *   cols pass : per detector column, IFFT_N(my . FFT_N(pad(column))) for every
 *               point, |.|^2 summed over the points of each of `nchunk` chunks
 *               and binned 4x4 into part[B][nchunk][cam][cam], then reduced over
 *               chunks into frames[B][cam][cam].
 * inter_bytes / part_bytes: sizes of the caller's buffers, checked against
 * oopao_pwfs_intermediate_bytes (ERR_WORKSPACE when smaller); bands as in
 * oopao_pwfs_frames (the same bands for both passes); lanes as lanes_rows /
 * lanes_cols of oopao_pwfs_frames. */
int oopao_pwfs_intermediate_bytes(int n, int N, int T, int B, int nchunk, size_t* inter_bytes,
                                  size_t* part_bytes);
int oopao_pwfs_rows(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                    const void* ty, const void* mx, int n, int N, int T, int B, int pts_per_task, int lanes,
                    const int* bands, void* inter, size_t inter_bytes, void* stream);
int oopao_pwfs_cols(const void* inter, size_t inter_bytes, const void* my, int n, int N, int T, int B, int nchunk,
                    int lanes, const int* bands, float* part, size_t part_bytes, float* frames, void* stream);

/* Unbinned modulated frames on the separable path (the reference's raw_data,
 * Pyramid.py:751-757): raw[B][N][N] float32 = sum over points of
 * |ifft2(fft2(E) mask)|^2, same inputs as oopao_pwfs_frames; workspace of
 * oopao_pwfs_raw_bytes bytes. */
int oopao_pwfs_raw_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes);
int oopao_pwfs_raw(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                   const void* ty, const void* mx, const void* my, int n, int N, int T, int B, int nchunk,
                   int pts_per_task, void* workspace, size_t ws_bytes, float* raw, void* stream);

/* Detector integration by area (the reference's sum binning, tools/tools.py:464-483,
 * extended to non-integer factors): frames[b][R][C] = sum_{Y,X} w(R,Y) w(C,X) raw[b][Y][X]
 * with w(R, Y) the overlap of raw pixel [Y, Y+1) with detector pixel [R s, (R+1) s),
 * s = N / cam (a plain sum over s x s pixels when cam divides N). No counterpart
 * for a non-integer s in the reference, whose Pyramid refuses such a sampling
 * (Pyramid.py:216-217). */
int oopao_bin_area(const float* raw, int B, int N, int cam, float* frames, void* stream);

/* Modulation focal-plane camera (Pyramid.py:638-642 modulation_camera_em,
 * 1183-1190 intensity crop): for B screens,
 *   out[b][crop][crop] = sum_k |fft2(pad(E_b tx_p(k) ty_k))|^2 / N^2
 * on the crop x crop window starting at N/2 - crop/2 (crop = the focal
 * camera's resolution). Same field inputs as oopao_pwfs_frames; workspace
 * size from oopao_pwfs_focal_bytes. */
int oopao_pwfs_focal_bytes(int n, int N, int T, int B, int nchunk, int crop, size_t* bytes);
int oopao_pwfs_focal(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                     const void* ty, int n, int N, int T, int B, int nchunk, int pts_per_task, int crop,
                     void* workspace, size_t ws_bytes, float* out, void* stream);

/* Adjoint of oopao_pwfs_frames (no counterpart in the reference, which has no
 * gradients; it differentiates the forward of Pyramid.py:612-646, 712-757):
 * for a detector-frame cotangent gframe [B][cam][cam] = dL/dframe, returns
 * grad [B][n][n] = dL/dopd (times nothing else: scale = dphi/dopd is applied).
 * Same field tables as oopao_pwfs_frames plus txs [T][n], the tip phasor of
 * each point in ty's point order (txs[k] = tx[p] for pstart[p] <= k <
 * pstart[p+1]). The modulation points are recomputed, nothing is stored
 * between forward and backward. Workspace from oopao_pwfs_adjoint_bytes.
 * bands: null, or detector-column bands {a0, a1, b0, b1} as in
 * oopao_pwfs_frames outside which gframe is zero (any loss of the slopesMaps
 * signal, which reads only the quadrant columns): only those columns are
 * transformed, and gframe's other columns are ignored. */
int oopao_pwfs_adjoint_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes);
int oopao_pwfs_adjoint(const float* opd, float scale, const float* amp, const void* tx, int P, const int* pstart,
                       const void* ty, const void* txs, const void* mx, const void* my, const float* gframe, int n,
                       int N, int T, int B, int nchunk, int pts_per_task, const int* bands, void* workspace,
                       size_t ws_bytes, float* grad, void* stream);

/* ---------------------------------------------------------------------------
 * General PWFS path: any padded length N <= 2048, any N x N mask (rooftop,
 * rotation, per-quadrant shifts, spatial filter: Pyramid.py:368-480, 1051-1085),
 * any integer detector binning, and the unbinned raw frame (raw_data,
 * Pyramid.py:751-757). Same per-point field tables tx, ty as the fast path
 * (the pupil sits at offset N/2 - n/2 of the padded support, Pyramid.py:623-625).
 *   maskT [N][N] complex64: the mask TRANSPOSED (maskT[kx][ky] = mask[ky][kx])
 *         and divided by N^2 (both inverse normalisations)
 *   frames [B][N/bin][N/bin] (may be null), raw [B][N][N] (may be null; not
 *         both) float32.
 * ------------------------------------------------------------------------- */
int oopao_pwfs_generic_supported(int n, int N, int bin);
int oopao_pwfs_generic_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes);
int oopao_pwfs_generic_frames(const float* opd, float scale, const float* amp, const void* tx, const void* ty,
                              const void* maskT, int n, int N, int T, int B, int bin, int nchunk,
                              void* workspace, size_t ws_bytes, float* frames, float* raw, void* stream);
/* Focal-plane camera on the general path (same output as oopao_pwfs_focal). */
int oopao_pwfs_generic_focal_bytes(int n, int N, int T, int B, int nchunk, int crop, size_t* bytes);
int oopao_pwfs_generic_focal(const float* opd, float scale, const float* amp, const void* tx, const void* ty,
                             int n, int N, int T, int B, int nchunk, int crop, void* workspace, size_t ws_bytes,
                             float* out, void* stream);
/* Adjoint of oopao_pwfs_generic_frames (no counterpart in the reference, which
 * has no gradients): grad [B][n][n] float32 = dL/dopd for the cotangent
 * gframe [B][N/bin][N/bin] = dL/dframes, same field inputs and mask, every
 * modulation point recomputed (nothing is kept from the forward). For the
 * forward E_i = ifft2(fft2(U_i) mask) of Pyramid.pyramid_transform
 * (Pyramid.py:612-646) it computes sum_i 2 scale Im(W_i conj(U_i)) on the pupil,
 * W_i = ifft2(conj(mask) fft2(g E_i)), g = gframe spread over each bin x bin
 * block. Workspace of oopao_pwfs_generic_adjoint_bytes bytes. */
int oopao_pwfs_generic_adjoint_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes);
int oopao_pwfs_generic_adjoint(const float* opd, float scale, const float* amp, const void* tx, const void* ty,
                               const void* maskT, const float* gframe, int n, int N, int T, int B, int bin,
                               int nchunk, void* workspace, size_t ws_bytes, float* grad, void* stream);

/* slopesMaps post-processing (Pyramid.py:835-854, grabQuadrant 946-961):
 *   frames [B][cam][cam] float32, valid [w][w] uint8 (validI4Q),
 *   quad[8] int32 (row origins of quadrants 1..4, then column origins),
 *   ref2d [2w][w] float32 (referenceSignal_2D), units = slopesUnits,
 *   vidx [nsig] int32 flat indices of validSignal in the [2w][w] map,
 *   nvalid = number of valid pixels of validI4Q.
 * Outputs sig2d [B][2w][w] (signal_2D) and sig [B][nsig] (signal). The normalisation
 * (mean of the valid I4Q, accumulated in double in a fixed order) uses a
 * workspace of oopao_pwfs_slopes_workspace_bytes(B, w) bytes. */
int oopao_pwfs_slopes_workspace_bytes(int B, int w, size_t* bytes);
int oopao_pwfs_slopes(const float* frames, int B, int cam, const unsigned char* valid, int w,
                      const int* quad, const float* ref2d, float units, const int* vidx, int nsig,
                      int nvalid, void* workspace, size_t ws_bytes, float* sig2d, float* sig, void* stream);

/* Allocation-free variant of oopao_pwfs_slopes (same arguments without the workspace, same
 * results bit for bit). No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops). The partial sums live in a static
 * device array of 65536 doubles (512 KiB per device, held until the process ends): calls with
 * B * ceil(w*w / 256) above 65536 fail with ERR_WORKSPACE. Calls share that array on a device:
 * calls that overlap on one device (other streams, other threads) race. */
int oopao_pwfs_slopes_static(const float* frames, int B, int cam, const unsigned char* valid, int w,
                             const int* quad, const float* ref2d, float units, const int* vidx, int nsig,
                             int nvalid, float* sig2d, float* sig, void* stream);

/* ---------------------------------------------------------------------------
 * Deformable mirror (replaces DeformableMirror.modesComputation and the coefs
 * setter's modes @ coefs: DeformableMirror.py:572-600, 651-686)
 * ------------------------------------------------------------------------- */

/* Gaussian influence functions modes[n*n][nact] float32 on the grid
 * X = linspace(0,1,n)*n: exp(-(a dX^2 + 2 b dX dY + c dY^2)) * sign, with
 * actuator centres x0, y0 [nact] float32 in pixels (DeformableMirror.py:572-600). */
int oopao_dm_gaussian_modes(float* modes, int n, int nact, const float* x0, const float* y0, double a,
                            double b, double c, float sign, void* stream);

/* y[r] = alpha * sum_k M[r][k] x[k] + beta * y0[r] (y0 may be NULL); M row-major
 * [R][K] float32. Used for OPD = modes @ coefs (DeformableMirror.py:674) and the
 * reconstructor product (tutorials/AO_closed_loop_Pyramid_WFS.py:634). */
int oopao_gemv(const float* M, const float* x, float* y, const float* y0, int R, int K, float alpha,
               float beta, void* stream);

/* ---------------------------------------------------------------------------
 * BF16 contractions on the matrix cores (FP32 accumulation), for the dense
 * influence-function product OPD = modes @ coefs (DeformableMirror.py:651-686)
 * and the reconstructor product reconstructor @ signal
 * (tutorials/AO_closed_loop_Pyramid_WFS.py:620-624):
 *     out[b][r] = alpha * sum_k A[r][k] X[k][b].
 * Operands are stored in the MFMA fragment layout: for Q[i][k] (I x K),
 * IT = ceil(I/32), KT = ceil(K/16), element ((kt IT + it) 64 + l) 8 + j holds
 * Q[32 it + l%32][16 kt + 8 (l/32) + j] as BF16 (round to nearest even) for
 * kt < 8 ceil(KT/8) + 8, zero outside Q; lo (optional) holds bf16(Q - hi). The
 * layout has no counterpart in the reference (which keeps modes as a float64
 * [n^2, nAct] array).
 * ------------------------------------------------------------------------- */

/* Bytes of one packed BF16 operand (hi or lo) for an I x K matrix. */
int oopao_bf16_packed_bytes(int I, int K, size_t* bytes);

/* Pack Q[i][k] = src[i*si + k*sk] (float32, I x K) into hi (and lo if non-NULL). */
int oopao_bf16_pack(const float* src, long long si, long long sk, int I, int K, void* hi, void* lo, void* stream);

/* out[b*ldo + r] = alpha * sum_k A[r][k] X[k][b] for A packed as R x K and X
 * packed as B x K (Q[b][k] = X[k][b]). terms = 1: hi*hi (BF16); terms = 3:
 * hi*hi + hi*lo + lo*hi (split BF16, needs both lo parts). ldo >= R. */
int oopao_gemm_bf16(const void* a_hi, const void* a_lo, int R, int K, const void* x_hi, const void* x_lo, int B,
                    float* out, long long ldo, float alpha, int terms, void* stream);

/* y[r] = alpha * sum_k A_hi[r][k] x[k] + beta * y0[r] (y0 may be NULL) with A
 * packed as R x K (hi only), x float32 [K] (16-byte aligned); workspace of
 * oopao_gemv_bf16_workspace_bytes bytes (fixed-order partial sums). */
int oopao_gemv_bf16_workspace_bytes(int R, int K, size_t* bytes);
int oopao_gemv_bf16(const void* a_hi, int R, int K, const float* x, float* y, const float* y0, float alpha,
                    float beta, void* ws, size_t ws_bytes, void* stream);

/* Separable OPD of a Cartesian, zero-misregistration Gaussian DM for B
 * coefficient vectors coefs [nAct][B] (column j = one command vector):
 * opd[B][n][n] = gy C_j gx^T, where C_j[a][b] = coefs[grid[a*nA+b]][j] on the
 * nA x nA actuator grid (grid = actuator index, or -1 where there is no valid
 * actuator; row a along Y), gxT [nA][n] (transposed), gy [n][nA]; tmp scratch of
 * B (nA*nA + n*nA) floats. Equal to modes @ coefs for that DM
 * (DeformableMirror.py:572-600, 651-686). */
int oopao_dm_separable_opd(const float* coefs, const int* grid, const float* gxT, const float* gy, float* tmp,
                           float* opd, int n, int nA, int B, void* stream);

/* ---------------------------------------------------------------------------
 * 2-D FFT and gain-sensing camera (GainSensingCamera.py:78-199)
 * ------------------------------------------------------------------------- */

/* 1 if a batched N x N transform has a plan (N <= 2048). */
int oopao_fft2_supported(int N);
int oopao_fft2_bytes(int N, int B, size_t* bytes);

/* out = scale * S(FFT2_dir(S(in))) for B complex64 N x N arrays; dir = -1 forward
 * (exp(-2 pi i jk/N)), +1 inverse (unnormalised); S = fftshift when centred (even N),
 * identity otherwise: the reference's fftshift(fft2(fftshift(x))) / N convention
 * (GainSensingCamera.py:147-152) with scale = 1/N. Workspace oopao_fft2_bytes. */
int oopao_fft2(const void* in, void* out, int N, int B, int dir, int centred, float scale, void* ws,
               size_t ws_bytes, void* stream);

/* Workspace of the three GSC calls for an N x N mask, M modes, mb modes per basis batch. */
int oopao_gsc_workspace_bytes(int N, int M, int mb, size_t* bytes);

/* ir = 2 Im(conj(fmask) F(mask frame / sum(frame))) (GainSensingCamera.py:179-180), with
 * fmask = F(mask) = fftshift(fft2(fftshift(mask))) / N precomputed; complex64 N x N output
 * (zero imaginary part), frame float32 N x N. */
int oopao_gsc_impulse_response(const void* fmask, const void* mask, const float* frame, int N, void* ir, void* ws,
                               size_t ws_bytes, void* stream);

/* bp[f][m0 + j] = |F(pad_N(basis[m0 + j]))|^2 for j < mb: the reference's basis product
 * fft3(B) * ifft3(B) (GainSensingCamera.py:155-171), real for a real basis; basis float32
 * [M][n][n] (mode-major maps), bp float32 [N*N][M]. */
int oopao_gsc_basis_product(const float* basis, int n, int N, int M, int m0, int mb, float* bp, void* ws,
                            size_t ws_bytes, void* stream);

/* S[m] = sum_f (F(ir1) v2)[f] bp[f][m] (complex, interleaved doubles [M][2]) with
 * v2 = F'(ir2) = fftshift(ifft2(fftshift(ir2))) * N precomputed (GainSensingCamera.py:183-188). */
int oopao_gsc_sensitivity(const void* ir1, const void* v2, const float* bp, int N, int M, double* S, void* ws,
                          size_t ws_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Angular-spectrum propagation for the scintillation mode of the atmosphere
 * (Atmosphere.fill_scintillation_support / set_scintillation_support / ASM,
 * Atmosphere.py:644-706, 938-972)
 * ------------------------------------------------------------------------- */

/* field = (init ? 1 : field) * exp(i k opd) for count complex64 values. */
int oopao_field_phase(const float* opd, float k, void* field, int count, int init, void* stream);

int oopao_asm_bytes(int N, size_t* bytes);

/* In place: field = fftshift(ifft2(ifftshift(fftshift(fft2(ifftshift(field))) H))) for an even
 * N x N complex64 field and the centred-grid transfer function H (complex64 N x N). */
int oopao_asm_propagate(void* field, const void* H, int N, void* ws, size_t ws_bytes, void* stream);

/* intensity = |field|^2 (if non-NULL); opd = angle(field) / k, or with a guide geo (non-NULL)
 * opd = geo + angle(field exp(-i k geo)) / k (if opd non-NULL). */
int oopao_field_readout(const void* field, const float* geo, float k, float* intensity, float* opd, int count,
                        void* stream);

/* ---------------------------------------------------------------------------
 * Atmosphere frozen-flow update (replaces Atmosphere.add_row / updateLayer and
 * the footprint sum of fill_OPD_support: Atmosphere.py:410-421, 447-516,
 * 589-622; the warp of tools/tools.py:202-209)
 * ------------------------------------------------------------------------- */

/* One whole-pixel step of a layer: tmp = map shifted by (sx, sy) in {-1,0,1}
 * (tmp[r][c] = map[r-sy][c-sx]), z = tmp[inner_idx], X = A z + B xi,
 * tmp[outer_idx] = X, mm[0..1] = (min, max) of tmp. map/tmp [M*M], A [n_outer][n_inner],
 * B [n_outer][n_outer], xi [n_outer] float64 (the layer's RandomState draw); mm holds 4
 * doubles (mm[2..3] are working space). Three kernel launches. */
int oopao_atm_add_row(const double* map, double* tmp, int M, int sx, int sy, const int* inner_idx, int n_inner,
                      const int* outer_idx, int n_outer, const double* A, const double* Bm, const double* xi,
                      double* z, double* X, double* mm, void* stream);

/* mm[0] = min(map[0..count)), mm[1] = max; mm holds 4 doubles (mm[2..3] are working space). */
int oopao_atm_minmax(const double* map, int count, double* mm, void* stream);

/* Boiling (Atmosphere.apply_boiling, Atmosphere.py:387-408): for the layer's extended screen map
 * [M][M] float64 (M = res + 2), map = alpha map + beta phi, phi the phase screen of
 * ft_sh_phase_screen (nsub = 12) or ft_phase_screen (nsub = 0) (phaseStats.py:153-271):
 *   g   [2 M^2]   the RandomState(boiling_seed) normal draws (real parts, then imaginary parts)
 *   amp [M^2]     sqrt(PSD) df of the FFT grid (PSD[M/2][M/2] = 0)
 *   W   [M][M]    complex128 centred DFT matrix (W C W^T = fftshift(fft2(fftshift(C))))
 *   sub [nsub][4] (Re c, Im c, fx, fy) of the subharmonic terms, xc [M] their grid coordinates.
 * Workspace of oopao_atm_boil_bytes(M) bytes. */
int oopao_atm_boil_bytes(int M, size_t* bytes);
/* The subharmonic rows of oopao_atm_boil from draws on the device (device RNG mode):
 * sub[t] = (g[re_t] a_t, g[im_t] a_t, fx_t, fy_t) for table[t] = (re_t, im_t, a_t, fx_t, fy_t),
 * the indices re_t, im_t < 54 (the first draws of the screen's stream, phaseStats.py:225-263). */
int oopao_atm_boil_subharmonics(const double* g, long long ng, const double* table, int nsub, double* sub,
                                void* stream);

/* Device RNG mode of the atmosphere (no counterpart in the reference, which draws on the host with
 * numpy's RandomState): out[j] = normal j of draw `draw` on stream `stream` under the key
 * (key0, key1), Philox4x32-10 (Salmon et al., SC 2011) with counter (j / 2, draw, stream) and a
 * Box-Muller pair of 53-bit uniforms. */
int oopao_normal_philox(double* out, long long count, unsigned key0, unsigned key1, unsigned stream,
                        unsigned long long draw, void* hip_stream);
int oopao_atm_boil(double* map, int M, const double* g, const double* amp, const void* W, const double* sub,
                   int nsub, const double* xc, double alpha, double beta, void* ws, size_t ws_bytes, void* stream);

/* out[y][x] (+)= w * clip(cubic(map, row y + r0 - ty, col x + c0 - tx)) for an
 * n x n window of the M x M map: Catmull-Rom cubic convolution, 0 outside the
 * map, clipped to [mm[0], mm[1]] (scikit-image warp order 3, mode constant,
 * clip=True). accumulate = 0 overwrites out. */
int oopao_atm_layer_opd(const double* map, int M, const double* mm, double tx, double ty, int r0, int c0, double w,
                        float* out, int n, int accumulate, void* stream);

/* out[r][c] (+)= weight * bilinear(in; y = a r + by, x = a c + bx) for an H x W float32 image
 * and an h x w output; 0 outside the image (scikit-image warp order 1, mode constant, as the
 * reference's interpolate_image / interpolate_cube,
 * tools/interpolateGeometricalTransformation.py:152-265): the sub-pixel translation of an
 * off-axis footprint (a = 1) and the laser-guide-star cone magnification
 * (Atmosphere.fill_OPD_support, Atmosphere.py:589-622). */
int oopao_bilinear_affine(const float* in, int H, int W, float* out, int h, int w, double a, double by, double bx,
                          double weight, int accumulate, void* stream);

/* ---------------------------------------------------------------------------
 * Closed loop (tutorials/AO_closed_loop_Pyramid_WFS.py:601-638): residual OPD
 * opd = (turb + dm) * pupil (dm may be null) and the RMS wavefront errors over
 * the pupil in nm, np.std(tel.OPD[pupil > 0]) * 1e9 for the turbulence alone
 * and for the residual, written to device scalars wfe_turb / wfe_res (either
 * may be null). npix = number of pupil pixels; workspace of
 * oopao_loop_residual_workspace_bytes bytes.
 * ------------------------------------------------------------------------- */
int oopao_loop_residual_workspace_bytes(size_t* bytes);
int oopao_loop_residual(const float* turb, const float* dm, const float* pupil, float* opd, int count, double npix,
                        void* workspace, double* wfe_turb, double* wfe_res, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OOPAO_AMD_H */
