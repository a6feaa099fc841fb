// sar_kernels.hip -- gfx950 kernels of the SAR forward projection and
// backprojection (adjoint), plus their host-side launchers.
//
// Stage map against the reference's native core
// (src/sarsolver/cxx_binding/sarsolver_cxx.cpp, "CPP"):
//   window_pad   CPP:331-342  conj(W) * Y into the zero-padded N_w spectrum
//   (rocFFT)     CPP:344      unnormalised inverse FFT, batched over pulses
//   bp_gather    CPP:345-350  per pixel: range, bin, profile sample, phase
//   fp_bin       CPP:284-291  per pulse: scatter a * e^{+i k_c dR} into bins
//   (rocFFT)     CPP:292      unnormalised forward FFT, batched over pulses
//   truncate     CPP:295-307  keep N modes, multiply by W
// The slow-time weight w_p (CPP:288, CPP:349) is folded into window_pad /
// truncate, which is exact by linearity of the FFT.
#include <hip/hip_runtime.h>

#include "sar_device.h"
#include "sar_launch.h"

namespace sarhip {

// ---------------------------------------------------------------------------
// Spectral stages (HBM-bound elementwise kernels)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float2 cmul(const float2 a, const float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// K[p, j] for j in [0, N_w): conj(W[k]) * wfac[p] * Y[p, src(k)] on the N
// support positions, zero elsewhere.  Support: j < ceil(N/2) -> k = j;
// j >= N_w - floor(N/2) -> k = j - (N_w - N) (CPP:331-342; odd-N middle mode
// lands at j = floor(N/2)).  `order` selects FFT-order or centred input; the
// centred case applies ifftshift on read (adjoint.py:36).
__global__ __launch_bounds__(256) void window_pad_kernel(const float2 *__restrict__ Y,
                                                         const float2 *__restrict__ Wconj,
                                                         const float2 *__restrict__ wfac, int N,
                                                         int nw, int centred,
                                                         float2 *__restrict__ K) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (j >= nw) return;
    const int lo = (N + 1) / 2;
    int k = -1;
    if (j < lo) k = j;
    else if (j >= nw - (N - lo)) k = j - (nw - N);
    float2 v = make_float2(0.f, 0.f);
    if (k >= 0) {
        const int src = centred ? (k + N / 2) % N : k;
        const float2 y = Y[(size_t)p * N + src];
        v = cmul(cmul(Wconj[k], y), wfac[p]);
    }
    K[(size_t)p * nw + j] = v;
}

// Y[p, k] = W[k] * wfac[p] * G[p, src(k)] (CPP:295-307); centred output
// applies fftshift on write (forward.py:44).
__global__ __launch_bounds__(256) void truncate_kernel(const float2 *__restrict__ G,
                                                       const float2 *__restrict__ W,
                                                       const float2 *__restrict__ wfac, int N,
                                                       int nw, int centred,
                                                       float2 *__restrict__ Y) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (k >= N) return;
    const int lo = (N + 1) / 2;
    const int src = k < lo ? k : nw - N + k;
    const float2 g = G[(size_t)p * nw + src];
    const float2 v = cmul(cmul(W[k], g), wfac[p]);
    const int dst = centred ? (k + N / 2) % N : k;
    Y[(size_t)p * N + dst] = v;
}

__global__ __launch_bounds__(256) void fill_zero_kernel(float4 *__restrict__ buf, size_t n4) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < n4; i += stride) buf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// amps_perm[i] = amps[perm[i]]: pixel amplitudes into tile (Morton) order.
__global__ __launch_bounds__(256) void gather_perm_kernel(const float2 *__restrict__ src,
                                                          const int *__restrict__ perm, int M,
                                                          float2 *__restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < M) dst[i] = src[perm[i]];
}

// ---------------------------------------------------------------------------
// Backprojection gather, FP64 ranges (precision = fp64)
//
// One thread per pixel (tile order), a workgroup = 256 consecutive pixels of
// the Morton order.  The pulse loop is wave-uniform: per-pulse constants come
// from scalar loads.  Ranges are evaluated exactly as the reference does, so
// bin indices are bit-identical to it.  Profile samples are gathered from
// the upsampled range profiles (L1/L2-resident windows).  Phase: FP64
// revolutions, integer part removed in FP64, FP32 sin/cos.  Accumulation in
// FP32 over chunks of 64 pulses, chunks summed in FP64.
// ---------------------------------------------------------------------------
template <int INTERP, bool MONO>
__global__ __launch_bounds__(256) void bp_gather_fp64_kernel(
    const double *__restrict__ px, const double *__restrict__ py, const double *__restrict__ pz,
    const int *__restrict__ perm, int M, const PulseConst *__restrict__ pc, int P,
    const float2 *__restrict__ prof, int nw, double by_dx, double kc_rev, double inv_nw,
    float2 *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool active = i < M;
    const int ii = active ? i : M - 1;
    const double x = px[ii], y = py[ii], z = pz[ii];
    const bool finite = isfinite(x) && isfinite(y) && isfinite(z);
    const double nwd = (double)nw;
    double sr = 0.0, si = 0.0;
    float ar = 0.f, ai = 0.f;
    for (int p = 0; p < P; ++p) {
        const PulseConst c = pc[p];
        if (c.valid == 0.0) continue;
        const double dr = ref_dr<MONO>(c, x, y, z);
        const float2 *row = prof + (size_t)p * nw;
        const double u = dr * by_dx;
        float hr, hi;
        if (INTERP == kInterpLinear) {
            const double f0 = floor(u);
            const float fr = (float)(u - f0);
            const int n0 = wrap_bin(f0, nwd, inv_nw, nw);
            const int n1 = (n0 + 1 == nw) ? 0 : n0 + 1;
            const float2 h0 = row[n0], h1 = row[n1];
            hr = fmaf(fr, h1.x - h0.x, h0.x);
            hi = fmaf(fr, h1.y - h0.y, h0.y);
        } else {
            const double v = (INTERP == kInterpNearestRef) ? 0.5 + u : u;
            const int n = wrap_bin(round(v), nwd, inv_nw, nw);
            const float2 h = row[n];
            hr = h.x;
            hi = h.y;
        }
        float s, co;
        phasor_rev(dr * kc_rev, &s, &co);
        // h * e^{-i theta}
        ar += fmaf(hr, co, hi * s);
        ai += fmaf(hi, co, -(hr * s));
        if ((p & 63) == 63) {
            sr += (double)ar;
            si += (double)ai;
            ar = 0.f;
            ai = 0.f;
        }
    }
    sr += (double)ar;
    si += (double)ai;
    if (active) {
        if (!finite) sr = si = 0.0;
        out[perm[i]] = make_float2((float)sr, (float)si);
    }
}

// ---------------------------------------------------------------------------
// Forward binning, FP64 ranges (precision = fp64)
//
// Grid (forward tiles) x (pulse groups).  Each workgroup keeps, per pulse of
// its group, an LDS window of the range profile that covers every bin its
// tile can reach (tile bounding sphere: |dR - dR_centre| <= 2 radius), adds
// contributions with LDS float atomics, then flushes each non-zero window
// bin into the global profile with one float atomic per component.
// Contributions outside a window (never expected; kept for safety) go
// straight to global memory.
// ---------------------------------------------------------------------------
template <int INTERP, bool MONO>
__global__ __launch_bounds__(256) void fp_bin_fp64_kernel(
    const double *__restrict__ px, const double *__restrict__ py, const double *__restrict__ pz,
    const float2 *__restrict__ amps_perm, int M, int tile_len, const double4 *__restrict__ tiles,
    const PulseConst *__restrict__ pc, int P, int pg, int win, float2 *__restrict__ prof, int nw,
    double by_dx, double kc_rev, double inv_nw) {
    extern __shared__ double smem[];
    double *wbase = smem;                         // [pg] window base bin (unwrapped)
    float2 *wsum = (float2 *)(smem + pg);         // [pg][win]
    const int t = blockIdx.x, g = blockIdx.y;
    const int p0 = g * pg;
    const int np = min(pg, P - p0);
    const int m0 = t * tile_len, m1 = min(M, m0 + tile_len);
    const double nwd = (double)nw;
    for (int j = threadIdx.x; j < pg * win; j += 256) wsum[j] = make_float2(0.f, 0.f);
    if ((int)threadIdx.x < np) {
        const PulseConst c = pc[p0 + threadIdx.x];
        const double4 tl = tiles[t];
        const double dr0 = ref_dr<MONO>(c, tl.x, tl.y, tl.z);
        wbase[threadIdx.x] = floor((dr0 - 2.0 * tl.w) * by_dx) - 2.0;
    }
    __syncthreads();
    for (int m = m0 + (int)threadIdx.x; m < m1; m += 256) {
        const double x = px[m], y = py[m], z = pz[m];
        if (!(isfinite(x) && isfinite(y) && isfinite(z))) continue;
        const float2 a = amps_perm[m];
        for (int q = 0; q < np; ++q) {
            const PulseConst c = pc[p0 + q];
            if (c.valid == 0.0) continue;
            const double dr = ref_dr<MONO>(c, x, y, z);
            float s, co;
            phasor_rev(dr * kc_rev, &s, &co);
            // e^{+i theta} * a
            const float cr = fmaf(co, a.x, -(s * a.y));
            const float ci = fmaf(co, a.y, s * a.x);
            const double u = dr * by_dx;
            const double wb = wbase[q];
            float2 *row = prof + (size_t)(p0 + q) * nw;
            if (INTERP == kInterpLinear) {
                const double f0 = floor(u);
                const float fr = (float)(u - f0);
                const float gr = 1.0f - fr;
                const int l = (int)(f0 - wb);
                if (l >= 0 && l + 1 < win) {
                    float2 *w0 = wsum + q * win + l;
                    atomicAdd(&w0[0].x, gr * cr);
                    atomicAdd(&w0[0].y, gr * ci);
                    atomicAdd(&w0[1].x, fr * cr);
                    atomicAdd(&w0[1].y, fr * ci);
                } else {
                    const int n0 = wrap_bin(f0, nwd, inv_nw, nw);
                    const int n1 = (n0 + 1 == nw) ? 0 : n0 + 1;
                    atomicAdd(&row[n0].x, gr * cr);
                    atomicAdd(&row[n0].y, gr * ci);
                    atomicAdd(&row[n1].x, fr * cr);
                    atomicAdd(&row[n1].y, fr * ci);
                }
            } else {
                const double nd = round(u);
                const int l = (int)(nd - wb);
                if (l >= 0 && l < win) {
                    float2 *w0 = wsum + q * win + l;
                    atomicAdd(&w0->x, cr);
                    atomicAdd(&w0->y, ci);
                } else {
                    const int n = wrap_bin(nd, nwd, inv_nw, nw);
                    atomicAdd(&row[n].x, cr);
                    atomicAdd(&row[n].y, ci);
                }
            }
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < np * win; j += 256) {
        const int q = j / win, l = j - q * win;
        const float2 v = wsum[j];
        if (v.x != 0.f || v.y != 0.f) {
            const int n = wrap_bin(wbase[q] + (double)l, nwd, inv_nw, nw);
            float2 *row = prof + (size_t)(p0 + q) * nw;
            atomicAdd(&row[n].x, v.x);
            atomicAdd(&row[n].y, v.y);
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_window_pad(const float2 *Y, const float2 *Wconj, const float2 *wfac, int N,
                             int nw, int P, int centred, float2 *K, hipStream_t s) {
    dim3 grid((nw + 255) / 256, P);
    window_pad_kernel<<<grid, 256, 0, s>>>(Y, Wconj, wfac, N, nw, centred, K);
    return hipGetLastError();
}

hipError_t launch_truncate(const float2 *G, const float2 *W, const float2 *wfac, int N, int nw,
                           int P, int centred, float2 *Y, hipStream_t s) {
    dim3 grid((N + 255) / 256, P);
    truncate_kernel<<<grid, 256, 0, s>>>(G, W, wfac, N, nw, centred, Y);
    return hipGetLastError();
}

hipError_t launch_fill_zero(void *buf, size_t bytes, hipStream_t s) {
    const size_t n4 = bytes / 16;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    fill_zero_kernel<<<(unsigned)blocks, 256, 0, s>>>((float4 *)buf, n4);
    return hipGetLastError();
}

hipError_t launch_gather_perm(const float2 *src, const int *perm, int M, float2 *dst,
                              hipStream_t s) {
    gather_perm_kernel<<<(M + 255) / 256, 256, 0, s>>>(src, perm, M, dst);
    return hipGetLastError();
}

hipError_t launch_bp_gather(const GatherArgs &a, hipStream_t s) {
    const int blocks = (a.M + 255) / 256;
#define SARHIP_BP(I, MO)                                                                      \
    bp_gather_fp64_kernel<I, MO><<<blocks, 256, 0, s>>>(a.px, a.py, a.pz, a.perm, a.M, a.pc,   \
                                                        a.P, a.prof, a.nw, a.by_dx, a.kc_rev, \
                                                        a.inv_nw, a.out)
    if (a.mono) {
        if (a.interp == kInterpLinear) SARHIP_BP(kInterpLinear, true);
        else if (a.interp == kInterpNearest) SARHIP_BP(kInterpNearest, true);
        else SARHIP_BP(kInterpNearestRef, true);
    } else {
        if (a.interp == kInterpLinear) SARHIP_BP(kInterpLinear, false);
        else if (a.interp == kInterpNearest) SARHIP_BP(kInterpNearest, false);
        else SARHIP_BP(kInterpNearestRef, false);
    }
#undef SARHIP_BP
    return hipGetLastError();
}

hipError_t launch_fp_bin(const BinArgs &a, hipStream_t s) {
    dim3 grid(a.num_tiles, (a.P + a.pg - 1) / a.pg);
    const size_t lds = sizeof(double) * a.pg + sizeof(float2) * (size_t)a.pg * a.win;
#define SARHIP_FP(I, MO)                                                                       \
    fp_bin_fp64_kernel<I, MO><<<grid, 256, lds, s>>>(a.px, a.py, a.pz, a.amps_perm, a.M,        \
                                                     a.tile_len, a.tiles, a.pc, a.P, a.pg,      \
                                                     a.win, a.prof, a.nw, a.by_dx, a.kc_rev,    \
                                                     a.inv_nw)
    // nearest_ref and nearest share the forward bin rule round(u) (CPP:289).
    if (a.mono) {
        if (a.interp == kInterpLinear) SARHIP_FP(kInterpLinear, true);
        else SARHIP_FP(kInterpNearest, true);
    } else {
        if (a.interp == kInterpLinear) SARHIP_FP(kInterpLinear, false);
        else SARHIP_FP(kInterpNearest, false);
    }
#undef SARHIP_FP
    return hipGetLastError();
}

}  // namespace sarhip
