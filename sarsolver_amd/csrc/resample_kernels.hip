// resample_kernels.hip -- device stages of dataset resampling (gfx950).
//
// The reference (numpy/measurement.py:342-457) rebuilds each output pulse
// from a window of K input pulses: it moves their SRP (utils/geodesy.py:33-73),
// crops their range profiles (utils/signal.py:45-73) and sums them with a
// complex slow-time kernel (utils/signal.py:7-25).  The SRP move and the
// inverse FFT act on one input pulse at a time, so here they run once per
// input pulse instead of once per window.  The crop, the taper, the kernel sum
// and the forward FFT are all linear in the range samples, so the kernel sum
// is taken on the cropped range profiles and only the P2 output pulses are
// transformed back; the final fftshift is a phase applied before that FFT:
//   rs_ramp_kernel   : SRP phase ramp, ifftshift to FFT order         [P][N]
//   rocFFT inverse, length N, P rows (in place)
//   rs_window_kernel : crop N2 samples, slow-time kernel over each
//                      window, Tukey taper / N, fftshift phase         [P2][N2]
//   rocFFT forward, length N2, P2 rows (in place, on the output)
// Both kernels are HBM-bound elementwise passes.  decode_signal_kernel turns
// the raw big-endian samples of a CPHD signal array into complex64.
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

// out[i][j] = e^(2 pi i j h / N2) win[j] sum_k kern[k] prof[starts[i] + k][src(j)],
// h = floor(N2 / 2), src(j) = j for j < h, else N - (N2 - j): the first h and
// last N2 - h range samples of each profile (signal.py:64-68).  win holds
// fftshift(tukey) / N (numpy's ifft scaling).  The phase e^(2 pi i j h / N2)
// turns the forward FFT's output into centred order: fftshift(FFT(x))[m] =
// FFT(x)[(m - h) mod N2] = FFT(x e^(2 pi i j h / N2))[m].
__global__ __launch_bounds__(256) void rs_window_kernel(const float2 *__restrict__ prof, int N,
                                                        int N2, const float *__restrict__ win,
                                                        const float2 *__restrict__ kern, int K,
                                                        const int *__restrict__ starts,
                                                        float2 *__restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= N2) return;
    const int h = N2 / 2;
    const int src = j < h ? j : N - (N2 - j);
    const float2 *base = prof + (size_t)starts[i] * N + src;
    float2 acc = make_float2(0.f, 0.f);
    for (int k = 0; k < K; ++k) {
        const float2 v = base[(size_t)k * N];
        const float2 w = kern[k];
        acc.x = fmaf(v.x, w.x, fmaf(-v.y, w.y, acc.x));
        acc.y = fmaf(v.x, w.y, fmaf(v.y, w.x, acc.y));
    }
    // j h mod N2 in integers keeps the phase exact; 0 and N2/2 are +1 and -1.
    const long long r = ((long long)j * h) % N2;
    float2 ph = make_float2(1.f, 0.f);
    if (2 * r == N2) {
        ph = make_float2(-1.f, 0.f);
    } else if (r != 0) {
        phasor_rev((double)r / (double)N2, &ph.y, &ph.x);
    }
    const float wj = win[j];
    out[(size_t)i * N2 + j] = cmulf(make_float2(acc.x * wj, acc.y * wj), ph);
}

// CPHD signal samples (big-endian) to complex64, times the vector's amplitude
// scale when given: format 0 = two float32 ("CF8"), 1 = two int16 ("CI4"),
// 2 = two int8 ("CI2").  One thread per sample of row p.
__global__ __launch_bounds__(256) void decode_signal_kernel(const unsigned char *__restrict__ raw,
                                                            int format, int ns,
                                                            const float *__restrict__ sf,
                                                            float2 *__restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (j >= ns) return;
    const size_t i = (size_t)p * ns + j;
    float re, im;
    if (format == 0) {
        const uint2 v = ((const uint2 *)raw)[i];
        re = __uint_as_float(__builtin_bswap32(v.x));
        im = __uint_as_float(__builtin_bswap32(v.y));
    } else if (format == 1) {
        const unsigned v = ((const unsigned *)raw)[i];
        re = (float)(short)(((v & 0xffu) << 8) | ((v >> 8) & 0xffu));
        im = (float)(short)((((v >> 16) & 0xffu) << 8) | (v >> 24));
    } else {
        const unsigned short v = ((const unsigned short *)raw)[i];
        re = (float)(signed char)(v & 0xffu);
        im = (float)(signed char)(v >> 8);
    }
    if (sf) {
        const float a = sf[p];
        re *= a;
        im *= a;
    }
    out[i] = make_float2(re, im);
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

hipError_t launch_rs_window(const float2 *prof, int N, int N2, const float *win, const float2 *kern,
                            int K, const int *starts, int P2, float2 *out, hipStream_t s) {
    for (int i0 = 0; i0 < P2; i0 += kMaxRows) {
        const int rows = P2 - i0 < kMaxRows ? P2 - i0 : kMaxRows;
        rs_window_kernel<<<dim3((N2 + 255) / 256, rows), 256, 0, s>>>(
            prof, N, N2, win, kern, K, starts + i0, out + (size_t)i0 * N2);
    }
    return hipGetLastError();
}

hipError_t launch_decode_signal(const void *raw, int format, int nv, int ns, const float *sf,
                                float2 *out, hipStream_t s) {
    const size_t row = (size_t)ns * (format == 0 ? 8 : format == 1 ? 4 : 2);
    for (int p0 = 0; p0 < nv; p0 += kMaxRows) {
        const int rows = nv - p0 < kMaxRows ? nv - p0 : kMaxRows;
        decode_signal_kernel<<<dim3((ns + 255) / 256, rows), 256, 0, s>>>(
            (const unsigned char *)raw + (size_t)p0 * row, format, ns, sf ? sf + p0 : nullptr,
            out + (size_t)p0 * ns);
    }
    return hipGetLastError();
}

}  // namespace sarhip
