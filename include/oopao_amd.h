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
 *    allocates nothing per call (a per-(device, N) twiddle table is built once).
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

/* ---------------------------------------------------------------------------
 * Pyramid WFS (replaces Pyramid.pyramid_transform + the modulation loop and
 * detector binning of Pyramid.wfs_measure: Pyramid.py:612-646, 666-757,
 * 1178-1203; tools/tools.py:464-483)
 * ------------------------------------------------------------------------- */

/* 1 if the separable fast path is compiled for this geometry: N = 4 n,
 * 4 x 4 detector binning, N in {64,128,256,320,512,640,1024,1280,1536,2048}. */
int oopao_pwfs_supported(int n, int N, int bin);

/* Workspace needed by oopao_pwfs_frames for B screens of T modulation points
 * reduced in `nchunk` point chunks. */
int oopao_pwfs_workspace_bytes(int n, int N, int T, int B, int nchunk, size_t* bytes);

/* Modulated PWFS detector frames for B phase screens (Pyramid.py:612-646, 712-826).
 *   opd   [B][n][n] float32; phase = opd * scale (src.phase = OPD 2 pi / lambda,
 *         Source.py:294-296)
 *   amp   [n][n]    float32 field amplitude sqrt(src.intensity / nTheta) (Pyramid.py:674)
 *   tx,ty [T][n]    complex64 per-point tip (x) / tilt (y) phasors times the
 *                   PSF-centring phasor (Pyramid.py:334-344, 1129-1139)
 *   mx,my [N]       complex64 separable factors of the pyramid mask, with the
 *                   zero-padding offset shift and the 1/N of each inverse FFT
 *                   folded in (Pyramid.py:388, 453-480)
 *   frames[B][N/bin][N/bin] float32: sum over points of |ifft2(fft2(E) mask)|^2,
 *         sum-binned (tools/tools.py:464-483)
 * pts_per_task: modulation points per rows-pass task; nchunk: point chunks of
 * the columns pass (both are tuning knobs, any valid value gives the same
 * result up to float summation order within a chunk). */
int oopao_pwfs_frames(const float* opd, float scale, const float* amp, const void* tx,
                      const void* ty, const void* mx, const void* my, int n, int N, int T, int B,
                      int bin, int nchunk, int pts_per_task, void* workspace, size_t ws_bytes,
                      float* frames, void* stream);

/* slopesMaps post-processing (Pyramid.py:835-854, grabQuadrant 946-961):
 *   frames [B][cam][cam] float32, valid [w][w] uint8 (validI4Q),
 *   quad[8] int32 (row origins of quadrants 1..4, then column origins),
 *   ref2d [2w][w] float32 (referenceSignal_2D), units = slopesUnits,
 *   vidx [nsig] int32 flat indices of validSignal in the [2w][w] map,
 *   nvalid = number of valid pixels of validI4Q.
 * Outputs sig2d [B][2w][w] (signal_2D) and sig [B][nsig] (signal). */
int oopao_pwfs_slopes(const float* frames, int B, int cam, const unsigned char* valid, int w,
                      const int* quad, const float* ref2d, float units, const int* vidx, int nsig,
                      int nvalid, float* sig2d, float* sig, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OOPAO_AMD_H */
