/*
 * sarsolver_hip.h -- C ABI of the MI355X (gfx950) SAR forward-projection /
 * backprojection library, libsarsolver_hip.so.
 *
 * Two families of entry points:
 *
 *  1. Reference-compatible entry points.  `forward_evaluate` and
 *     `adjoint_evaluate` take the same 128-byte `SarCalculationInfo` record
 *     (host memory, FP64, interleaved complex) that the reference binds with
 *     ctypes, so the reference's loader can be pointed at this library
 *     unchanged.  They always use the reference's nearest-neighbour bin rules
 *     with FP64 ranges (interp = SARHIP_INTERP_NEAREST_REF,
 *     precision = SARHIP_PRECISION_FP64).
 *
 *  2. Plan-based device API (`sarhip_*`).  Geometry is uploaded once into a
 *     plan; forward / adjoint then run on device buffers (complex64, i.e.
 *     interleaved float pairs) on a caller-supplied HIP stream.  This is what
 *     the Python operator layer and the solvers call.
 *
 * Every entry point returns 0 on success and a negative SARHIP_E* code on
 * failure; `sarhip_last_error()` gives the message for the calling thread.
 * No C++ exception crosses this boundary.
 */
#ifndef SARSOLVER_HIP_H
#define SARSOLVER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order, types and size (3*8 + 8*8 + 5*8 = 128 bytes) as the
 * reference's record: src/sarsolver/cxx_binding/sarsolver_cxx.hpp:27-34,
 * mirrored in ctypes at src/sarsolver/cxx_binding/sarsolver_cxx.py:53-62.
 * Arrays are host memory: positions [P,3] / [M,3] FP64; complex arrays
 * interleaved (re, im) FP64; phase_history [P,N] in FFT frequency order;
 * waveform_fft [N] in FFT order. */
typedef struct SarCalculationInfo {
    size_t num_fast_times, num_slow_times, num_scatterers;
    double *transmit_posns, *receive_posns, *stab_ref_posns, *scat_posns;
    double *phase_history, *scattering_amplitudes, *waveform_fft, *slow_time_weighting;
    double centre_frequency, sample_frequency, c_eff, upsample_ratio, sign_multiplier;
} SarCalculationInfo;

/* Error codes. */
#define SARHIP_OK 0
#define SARHIP_EINVAL (-1)  /* bad argument / inconsistent sizes            */
#define SARHIP_EHIP (-2)    /* HIP runtime error                            */
#define SARHIP_EFFT (-3)    /* rocFFT error                                 */
#define SARHIP_ENOMEM (-4)  /* device or host allocation failed             */
#define SARHIP_ENODEV (-5)  /* no usable gfx950 device                       */

/* Fast-time interpolation rule. */
#define SARHIP_INTERP_NEAREST_REF 0 /* reference rules: forward bin round(u),
                                       adjoint bin round(0.5+u), adjoint does
                                       not conjugate the slow-time weights   */
#define SARHIP_INTERP_NEAREST 1     /* round(u) both ways; exact adjoint pair */
#define SARHIP_INTERP_LINEAR 2      /* linear interpolation; exact adjoint pair */

/* Range / phase arithmetic. */
#define SARHIP_PRECISION_FP64 0   /* platform-to-pixel ranges in FP64 exactly as
                                     the reference evaluates them; with fused
                                     spectral stages the forward's runs in
                                     FP64 and its phasors carry the phase's
                                     FP32 rounding (for scenes whose phase
                                     history cancels heavily)                 */
#define SARHIP_PRECISION_SERIES 1 /* FP64 tile-centred series for the range;
                                     exact FP64 re-evaluation at bin ties     */
#define SARHIP_PRECISION_HILO 2   /* FP64 tile-centre range + FP32 differential
                                     range from tile-relative coordinates     */
#define SARHIP_PRECISION_FP32 3   /* all-FP32 absolute-coordinate ranges (study
                                     only: fails at 10 km standoff)           */

/* Phase-history frequency order of device buffers. */
#define SARHIP_ORDER_FFT 0     /* FFT order (index 0 = zero frequency)       */
#define SARHIP_ORDER_CENTRED 1 /* centred order (fftshift of FFT order), the
                                  reference's public convention               */

/* Replaces the reference's `forward_evaluate(SarCalculationInfo&)`
 * (src/sarsolver/cxx_binding/sarsolver_cxx.hpp:113,
 *  sarsolver_cxx.cpp:26-30; bound at sarsolver_cxx.py:214-216).
 * Overwrites info->phase_history ([P,N], FFT order) with A x where
 * x = info->scattering_amplitudes. */
int forward_evaluate(SarCalculationInfo *info);

/* Replaces the reference's `adjoint_evaluate(SarCalculationInfo&)`
 * (src/sarsolver/cxx_binding/sarsolver_cxx.hpp:114,
 *  sarsolver_cxx.cpp:32-36; bound at sarsolver_cxx.py:218-220).
 * Overwrites info->scattering_amplitudes ([M]) with the reference's
 * backprojection of info->phase_history ([P,N], FFT order). */
int adjoint_evaluate(SarCalculationInfo *info);

/* Opaque plan: geometry, constants, FFT plans and scratch on one device. */
typedef struct sarhip_plan_s *sarhip_plan_t;

/* Create a plan on `device`.  Geometry is read from HOST memory (same layouts
 * as SarCalculationInfo; waveform in FFT order) and uploaded once.  Sizes and
 * constants follow the reference's SarWorker constructor
 * (sarsolver_cxx.cpp:219-252): N_w = (size_t)(upsample_ratio * N).
 * Replaces the per-call SarWorker construction of the reference; it has no
 * single counterpart there. */
int sarhip_plan_create(sarhip_plan_t *plan, size_t num_fast_times, size_t num_slow_times,
                       size_t num_scatterers, const double *transmit_posns,
                       const double *receive_posns, const double *stab_ref_posns,
                       const double *scat_posns, const double *waveform_fft,
                       const double *slow_time_weighting, double centre_frequency,
                       double sample_frequency, double c_eff, double upsample_ratio,
                       double sign_multiplier, int interp, int precision, int device);

/* As sarhip_plan_create, with the scatterer positions ([M][3] FP64) already
 * in DEVICE memory on `device` (read on the default stream; the caller's
 * writes must be complete).  The plan keeps its own copy in tile order.  It
 * replaces the upload of GridSarScene.positions / SimpleSarScene positions
 * (numpy/hypothesis.py:191-206) that the reference performs on every call. */
int sarhip_plan_create_device_positions(sarhip_plan_t *plan, size_t num_fast_times,
                                        size_t num_slow_times, size_t num_scatterers,
                                        const double *transmit_posns, const double *receive_posns,
                                        const double *stab_ref_posns, const void *d_scat_posns,
                                        const double *waveform_fft,
                                        const double *slow_time_weighting, double centre_frequency,
                                        double sample_frequency, double c_eff,
                                        double upsample_ratio, double sign_multiplier, int interp,
                                        int precision, int device);

/* As sarhip_plan_create, for the pixels of a GridSarScene generated on the
 * device (numpy/hypothesis.py:191-206): grid_shape and grid_dimensions in
 * (z, y, x) order, rotation row-major [3][3], centre [3]; M = product of the
 * shape.  Positions are bit-identical to numpy's linspace/meshgrid and to
 * R @ local + centre evaluated as an FMA chain in k order (how numpy's
 * OpenBLAS matmul rounds it; for rotations of 0 and +-1 entries any order
 * agrees). */
int sarhip_plan_create_grid(sarhip_plan_t *plan, size_t num_fast_times, size_t num_slow_times,
                            const size_t grid_shape[3], const double grid_dimensions[3],
                            const double grid_centre[3], const double rotation[9],
                            const double *transmit_posns, const double *receive_posns,
                            const double *stab_ref_posns, const double *waveform_fft,
                            const double *slow_time_weighting, double centre_frequency,
                            double sample_frequency, double c_eff, double upsample_ratio,
                            double sign_multiplier, int interp, int precision, int device);

/* GridSarScene.positions (numpy/hypothesis.py:191-206) into device memory
 * d_pos [M][3] FP64 (caller order), stream-ordered on `stream`. */
int sarhip_grid_positions(const size_t grid_shape[3], const double grid_dimensions[3],
                          const double grid_centre[3], const double rotation[9], void *d_pos,
                          void *stream);

/* generate_simple_basebanding_values (utils/k_space.py:109-134) on the device:
 * d_out[m] = exp(-i sign ((d_pos[m] - spot) . displacement)), complex128
 * interleaved [M]; d_pos [M][3] FP64 device memory; `displacement` is the
 * k-space box displacement the reference computes on the host. */
int sarhip_baseband_values(const void *d_pos, size_t num_scatterers, const double spot[3],
                           const double displacement[3], double sign_multiplier, void *d_out,
                           void *stream);

/* Decode a CPHD signal array on the device, replacing the host decode inside
 * sarpy's reader that io/cphd101.py:252-267 (io/cphd03.py) calls: d_raw holds
 * num_vectors x num_samples big-endian complex samples (format 0 = CF8 /
 * RE32F_IM32F, 1 = CI4 / RE16I_IM16I, 2 = CI2 / RE08I_IM08I), aligned to one
 * sample; d_amp_sf [num_vectors] float32 per-vector scale factors or NULL;
 * d_out [num_vectors][num_samples] complex64.  Stream-ordered on `stream`. */
int sarhip_decode_signal(const void *d_raw, int format, size_t num_vectors, size_t num_samples,
                         const float *d_amp_sf, void *d_out, void *stream);

int sarhip_plan_destroy(sarhip_plan_t plan);

/* Concurrency.  A plan owns mutable device scratch (profile rows, forward
 * window partials and bounds, row-span and poison flags, pulse-split partial
 * images) that every sarhip_forward / sarhip_adjoint call on it overwrites.
 * Calls on ONE plan must therefore be ordered: issue them on one stream, or
 * order the streams (events / synchronisation) and the host threads yourself;
 * two calls of one plan that overlap on different streams race and corrupt
 * each other's results.  Different plans share nothing and may run
 * concurrently on different streams from different host threads.  Error
 * messages are per host thread (sarhip_last_error). */

/* Forward projection on device buffers: amps [M] complex64 ->
 * phase_history [P,N] complex64 in `order`.  Stream-ordered on `stream`
 * (a hipStream_t, NULL = default stream).  Device counterpart of the
 * reference's SarWorker::execute_forward_evaluate (sarsolver_cxx.cpp:271-310). */
int sarhip_forward(sarhip_plan_t plan, const void *amps, void *phase_history, int order,
                   void *stream);

/* Backprojection on device buffers: phase_history [P,N] complex64 in
 * `order` -> amps [M] complex64.  Device counterpart of the reference's
 * SarWorker::execute_adjoint_evaluate (sarsolver_cxx.cpp:319-353). */
int sarhip_adjoint(sarhip_plan_t plan, const void *phase_history, void *amps, int order,
                   void *stream);

/* Plan introspection: N_w, bins per metre of bistatic range (by_dx) and the
 * signed centre wavenumber k_c (sarsolver_cxx.cpp:223-228). */
int sarhip_plan_info(sarhip_plan_t plan, size_t *working_num_fast_times, double *by_dx,
                     double *centre_wavenumber);

/* Plan layout: the precision mode in effect (a series request falls back to
 * SARHIP_PRECISION_FP64 when the geometry does not bound the series
 * truncation below 1e-9 bin and 1e-9 revolution), whether every valid pulse
 * is monostatic, the number of valid pulses, the forward LDS window length
 * (bins, 0 = global atomics), pulses per forward workgroup, the series
 * truncation bound (bins), the number of 1024-pixel tiles per forward region,
 * the order of the series the kernels use (3 when its bound meets the same
 * limits, else 5; nearest-rule bins within 1e-7 bin of a tie are decided on
 * the exact range either way) with its bound, and the upsampling
 * ratio U of the fused spectral kernels (N_w = U N; 0 = rocFFT stages).  Any
 * output pointer may be NULL.  No counterpart in the reference. */
int sarhip_plan_layout(sarhip_plan_t plan, int *precision, int *monostatic, int *valid_pulses,
                       int *fwd_window_bins, int *fwd_pulses_per_group, double *series_err_bins,
                       int *fwd_region_tiles, int *bp_series_order, double *bp_series_err_bins,
                       int *fused_fft_ratio);

/* The forward's adaptive spectral stage (fused kernels, precision other than
 * exact ranges): the transforms run in FP32, and the rows whose cancellation
 * factor kappa = sqrt(N ||h||^2 / ||X||^2) exceeds a threshold (||h|| the
 * profile row's norm, ||X|| its band's; SARHIP_FWD_KAPPA at plan creation,
 * default 2, 0 turns it off) are transformed again in FP64.  *kappa receives
 * the threshold (0 when the stage is not adaptive); with `rows` non-NULL,
 * *rows receives the number of rows of the most recent forward that were
 * redone (waits for the device).  No counterpart in the reference. */
int sarhip_plan_forward_fp64_rows(sarhip_plan_t plan, double *kappa, int *rows);

/* Whether the backprojection reads demodulated pair rows: after the spectral
 * stage a pass stores, over the bins the gather can read, (h[n], h[n+1] -
 * h[n]) e^{-2 pi i rho (n + 1/2)} (rho = k_c / (2 pi by_dx), the phase in
 * revolutions per bin), and the gather interpolates with two FMAs and forms
 * each pixel's phase rho f + phi' in FP32 (f the bin fraction); for the
 * nearest-bin rules it stores the first member alone.  Enabled for the
 * monostatic planar series plans with the third-order series and fused
 * spectral stages when rho is an FP32 number (to 2^-34) of magnitude at most
 * 2.5, in every interpolation mode; SARHIP_BP_PAIRS=0 at plan creation turns
 * it off.  No counterpart in the reference. */
int sarhip_plan_gather_pairs(sarhip_plan_t plan, int *enabled);

/* Stage timing with HIP events recorded on the caller's stream.  Enabling
 * (or re-enabling) clears the log; while enabled every forward / adjoint call
 * records its stage boundaries without synchronising.  sarhip_plan_timing
 * waits for the logged calls and returns the mean device time (ms) per call of
 * `which` (0 = forward, 1 = adjoint), split by stage: [0] spectral window
 * (window_pad / truncate), [1] FFT, [2] gather (adjoint) or scatter (forward),
 * [3] total; *ncalls receives the number of calls averaged.  No counterpart in
 * the reference (it has no instrumentation). */
int sarhip_plan_set_timing(sarhip_plan_t plan, int enable);
int sarhip_plan_timing(sarhip_plan_t plan, int which, float *ms4, int *ncalls);
/* Mean device time (ms) of the main kernel alone over the logged calls of
 * `which`: the scatter launch of the forward (without the region norms and
 * the window merge) or the gather launch of the adjoint (without the
 * pulse-split sum).  No counterpart in the reference. */
int sarhip_plan_kernel_timing(sarhip_plan_t plan, int which, float *ms, int *ncalls);

/* Message describing the most recent failure on the calling thread. */
const char *sarhip_last_error(void);

/* Library version string. */
const char *sarhip_version(void);

/* Dataset resampling, replacing the window loop of
 * NumpySimpleSarDataset.from_resampling (numpy/measurement.py:397-457), which
 * calls move_srps (utils/geodesy.py:33-73), fast_freq_downsample
 * (utils/signal.py:45-73) and the slow-time kernel of get_downsample_kernel
 * (utils/signal.py:7-25).  Output pulse i is
 *   sum_k kern[k] * fast_freq_downsample(move_srps(ph_in[starts[i] + k]))
 * with move_srps's phase exp(2 pi i dr (f_bb - sign fc) / c) per input pulse
 * (dr = new minus old bistatic SRP range, f_bb the baseband frequency of the
 * sample).  ph_in [P][N] and ph_out [P2][N2] are device complex64 arrays in
 * centred frequency order; dr [P], win [N2] (= fftshift(tukey(N2)) / N),
 * kern [K] (complex128 interleaved) and starts [P2] are host arrays.  Work
 * buffers and FFT plans are kept per device for the last (P, N, N2, P2) shape
 * (held until the process ends); the call returns once `stream` has finished.
 * Calls from several host threads are serialised. */
int sarhip_resample(const void *ph_in, size_t P, size_t N, const double *dr, double centre_freq,
                    double sample_freq, double sign_multiplier, double c_eff, size_t N2,
                    const double *win, const double *kern, int K, const int *starts, size_t P2,
                    void *ph_out, void *stream);

/* ---- 3x3 convolutions of the learned CG preconditioner ----------------------
 * No counterpart in the reference (it has no learned components); these
 * serve the preconditioner network that accelerates CG on the normal
 * equations of A.  Activations are NHWC bfloat16 with 32 channels (narrower
 * layers zero-padded), weights bfloat16, accumulation FP32, zero padding of
 * one pixel on every side.  Device pointers; `stream` is a hipStream_t (NULL =
 * default stream).  Return SARHIP_OK, SARHIP_EINVAL or SARHIP_EHIP;
 * sarhip_conv_last_error() gives the message. */

/* y = act(conv3x3(x, W) + bias): x [B,H,W,32] bf16; wfrag = W[9][32][out_c] in
 * matrix-core fragment order [tap][nb][lane][8], element (lane l, i) =
 * W[tap][ci = 8 (l / 16) + i][co = 16 nb + l % 16] (nb < 2 if out_c > 16,
 * else 1; zero beyond out_c); bias [out_c] f32; y [B,H,W,out_c], FP32 when
 * out_f32 else bf16; ReLU when relu != 0.  1 <= out_c <= 32. */
int sarhip_conv3x3_bf16(const void *x, const void *wfrag, const float *bias, int B, int H, int W,
                        int out_c, int relu, int out_f32, void *y, void *stream);

/* Workspace bytes needed by sarhip_conv3x3_wgrad for these sizes. */
size_t sarhip_conv3x3_wgrad_workspace(int B, int H, int W, int out_c);

/* Weight and bias gradient of a layer: x [B,H,W,32] bf16 (its input), dy
 * [B,H,W,co] f32 (gradient of the pre-activation output, co = 32 if out_c > 16
 * else 16, zero-padded), dwdb = [dW (9*32*co, [tap][ci][co]), db (co)] f32.
 * Deterministic (per-tile partials summed in a fixed order). */
int sarhip_conv3x3_wgrad(const void *x, const float *dy, int B, int H, int W, int out_c,
                         float *dwdb, void *workspace, size_t workspace_bytes, void *stream);

/* Inference of the preconditioner network in one call (no counterpart in the
 * reference): for each image b of z [B][H*W] complex64, s_b = max(sqrt(mean
 * |z_b|^2), 1e-30) (FP64 partial sums in a fixed order), x = bf16(z_b / s_b) in
 * channels 0 and 1 of a 32-channel input, `layers` convolutions as
 * sarhip_conv3x3_bf16 (wfrags[i], biases[i]: device pointers in host arrays;
 * out_c = 32, ..., 32, 2; ReLU after all but the last, the last in FP32), and
 * out_b = (z_b + s_b * (net_0 + i net_1)) * demod (demod [H*W] complex64, or
 * NULL for none).  `work`: device workspace of at least
 * sarhip_precond_net_workspace(B, H, W) bytes.  out may not alias z. */
size_t sarhip_precond_net_workspace(int B, int H, int W);
int sarhip_precond_net(const void *z, const void *demod, int B, int H, int W, int layers,
                       const void *const *wfrags, const float *const *biases, const int *out_c,
                       void *work, size_t work_bytes, void *out, void *stream);

/* Spectral filtering of the preconditioner's circulant base in one call (no
 * counterpart in the reference): for each of B images in [n][n] complex64,
 *   out = demod_out . crop_n( IFFT2_m( filter / m^2 . FFT2_m( pad_m(
 *         conj(demod_in) . in ) ) ) ),
 * i.e. the image zero-padded to m x m, filtered by the real filter [m][m]
 * (float32, indexed [ky][kx] in FFT order), and cropped back; demod_in /
 * demod_out are [n][n] complex64 phasors or NULL (none).  m a power of two in
 * [64, 1024], n <= m.  `work`: device workspace of at least
 * sarhip_spectral_filter_workspace(B, n, m) bytes.  out may alias in. */
size_t sarhip_spectral_filter_workspace(int B, int n, int m);
int sarhip_spectral_filter(const void *in, const void *demod_in, int B, int n, int m,
                           const float *filter, void *work, size_t work_bytes, void *out,
                           const void *demod_out, void *stream);

/* sarhip_precond_net without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight solver or
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments and result as sarhip_precond_net; the workspace
 * is a static device array of 34 MiB + 4 KiB in the library, enough for
 * B x H x W up to one 512 x 512 image (sizes whose sarhip_precond_net_workspace
 * exceeds it are rejected with SARHIP_EINVAL-style -1 and a message).  The
 * array's address is looked up on the first call on each device (make that
 * call before capturing a graph); the call itself allocates nothing.  Calls on
 * one device share the array: calls that overlap on one device (concurrent
 * streams or threads) race; no state carries over between calls.  The current
 * device must be the one that holds z, out and `stream`. */
int sarhip_precond_net_noalloc(const void *z, const void *demod, int B, int H, int W, int layers,
                               const void *const *wfrags, const float *const *biases,
                               const int *out_c, void *out, void *stream);

/* Message describing the most recent convolution failure on the calling thread. */
const char *sarhip_conv_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SARSOLVER_HIP_H */
