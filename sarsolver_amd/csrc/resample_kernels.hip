// resample_kernels.hip -- device stages of dataset resampling (gfx950).
//
// The reference (numpy/measurement.py:342-457) rebuilds each output pulse
// from a window of K input pulses: it moves their SRP (utils/geodesy.py:33-73),
// crops their range profiles (utils/signal.py:45-73) and sums them with a
// complex slow-time kernel (utils/signal.py:7-25).  The first two steps act on
// one input pulse at a time, so here they run once per input pulse instead of
// once per window:
//   rs_ramp_kernel : SRP phase ramp, ifftshift to FFT order         [P][N]
//   rocFFT inverse, length N (in place)
//   rs_crop_kernel : keep N2 profile samples, Tukey taper, 1/N       [P][N2]
//   rocFFT forward, length N2 (in place)
//   rs_fir_kernel  : slow-time kernel over each window, fftshift     [P2][N2]
// All three kernels are HBM-bound elementwise passes.
#include <hip/hip_runtime.h>

#include "sar_device.h"
#include "sar_launch.h"

namespace sarhip {

namespace {

__device__ __forceinline__ float2 cmulf(const float2 a, const float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// out[p][j] = in[p][(j + N/2) mod N] * exp(2 pi i dr_p (s(j) fs / N - sgn fc) / c),
// s(j) the signed FFT index of j (fftshift(fftfreq) order, geodesy.py:67-71).
// The phase is formed and reduced in FP64 revolutions; sin/cos take the rest.
__global__ __launch_bounds__(256) void rs_ramp_kernel(const float2 *__restrict__ in, int N,
                                                      const double *__restrict__ dr, double fs_over_n,
                                                      double sgn_fc, double inv_c,
                                                      float2 *__restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (j >= N) return;
    const int s = j < (N + 1) / 2 ? j : j - N;
    const double rev = dr[p] * ((double)s * fs_over_n - sgn_fc) * inv_c;
    float sn, cs;
    phasor_rev(rev, &sn, &cs);
    const float2 v = in[(size_t)p * N + (j + N / 2) % N];
    out[(size_t)p * N + j] = cmulf(v, make_float2(cs, sn));
}

// out[p][j2] = win[j2] * prof[p][src(j2)], src = j2 for j2 < N2/2 (floor), else
// N - (N2 - j2): the first floor(N2/2) and last N2 - floor(N2/2) samples
// (signal.py:64-68).  win holds fftshift(tukey) / N (numpy's ifft scaling).
__global__ __launch_bounds__(256) void rs_crop_kernel(const float2 *__restrict__ prof, int N,
                                                      int N2, const float *__restrict__ win,
                                                      float2 *__restrict__ out) {
    const int j2 = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (j2 >= N2) return;
    const int src = j2 < N2 / 2 ? j2 : N - (N2 - j2);
    const float2 v = prof[(size_t)p * N + src];
    const float w = win[j2];
    out[(size_t)p * N2 + j2] = make_float2(v.x * w, v.y * w);
}

// out[i][m] = sum_k kern[k] * spec[starts[i] + k][(m - N2/2) mod N2]: the
// slow-time kernel over window i, written in centred order (fftshift).
__global__ __launch_bounds__(256) void rs_fir_kernel(const float2 *__restrict__ spec, int N2,
                                                     const float2 *__restrict__ kern, int K,
                                                     const int *__restrict__ starts,
                                                     float2 *__restrict__ out) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (m >= N2) return;
    const int col = (m - N2 / 2 + N2) % N2;
    const float2 *base = spec + (size_t)starts[i] * N2 + col;
    float2 acc = make_float2(0.f, 0.f);
    for (int k = 0; k < K; ++k) {
        const float2 v = base[(size_t)k * N2];
        const float2 w = kern[k];
        acc.x = fmaf(v.x, w.x, fmaf(-v.y, w.y, acc.x));
        acc.y = fmaf(v.x, w.y, fmaf(v.y, w.x, acc.y));
    }
    out[(size_t)i * N2 + m] = acc;
}

}  // namespace

// Rows go on grid y, whose extent is at most 65535: launch in chunks of rows.
constexpr int kMaxRows = 65535;

hipError_t launch_rs_ramp(const float2 *in, int P, int N, const double *dr, double fs_over_n,
                          double sgn_fc, double inv_c, float2 *out, hipStream_t s) {
    for (int p0 = 0; p0 < P; p0 += kMaxRows) {
        const int rows = P - p0 < kMaxRows ? P - p0 : kMaxRows;
        rs_ramp_kernel<<<dim3((N + 255) / 256, rows), 256, 0, s>>>(
            in + (size_t)p0 * N, N, dr + p0, fs_over_n, sgn_fc, inv_c, out + (size_t)p0 * N);
    }
    return hipGetLastError();
}

hipError_t launch_rs_crop(const float2 *prof, int P, int N, int N2, const float *win, float2 *out,
                          hipStream_t s) {
    for (int p0 = 0; p0 < P; p0 += kMaxRows) {
        const int rows = P - p0 < kMaxRows ? P - p0 : kMaxRows;
        rs_crop_kernel<<<dim3((N2 + 255) / 256, rows), 256, 0, s>>>(
            prof + (size_t)p0 * N, N, N2, win, out + (size_t)p0 * N2);
    }
    return hipGetLastError();
}

hipError_t launch_rs_fir(const float2 *spec, int P2, int N2, const float2 *kern, int K,
                         const int *starts, float2 *out, hipStream_t s) {
    for (int i0 = 0; i0 < P2; i0 += kMaxRows) {
        const int rows = P2 - i0 < kMaxRows ? P2 - i0 : kMaxRows;
        rs_fir_kernel<<<dim3((N2 + 255) / 256, rows), 256, 0, s>>>(
            spec, N2, kern, K, starts + i0, out + (size_t)i0 * N2);
    }
    return hipGetLastError();
}

}  // namespace sarhip
