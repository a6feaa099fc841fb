// sar_kernels.hip -- gfx950 kernels of the SAR forward projection and
// backprojection (adjoint), plus their host-side launchers.
//
// Stage map against the reference's native core
// (src/sarsolver/cxx_binding/sarsolver_cxx.cpp, "CPP"):
//   window_pad   CPP:331-342  conj(W) * Y into the zero-padded N_w spectrum
//   (rocFFT)     CPP:344      unnormalised inverse FFT, batched over pulses
//   bp_gather    CPP:345-350  per pixel: range, bin, profile sample, phase
//   fp_scatter   CPP:284-291  per pulse: scatter a * e^{+i k_c dR} into bins
//   fp_reduce    (none)       merge the per-tile LDS windows of fp_scatter
//   (rocFFT)     CPP:292      unnormalised forward FFT, batched over pulses
//   truncate     CPP:295-307  keep N modes, multiply by W
// The slow-time weight w_p (CPP:288, CPP:349) is folded into window_pad /
// truncate, which is exact by linearity of the FFT.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sar_device.h"
#include "sar_launch.h"

namespace sarhip {

// ---------------------------------------------------------------------------
// Spectral stages (HBM-bound elementwise kernels)
// ---------------------------------------------------------------------------

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format,
// no swizzle, no stride).
constexpr int kBufferWord3 = 0x00020000;

// Structured-buffer load (LLVM's struct buffer intrinsic; HIP has no builtin
// for it): 16 bytes at base + stride * vindex + voffset, range-checked on the
// index alone -- vindex >= num_records reads zeros, any smaller index reads all
// 16 bytes (checked on MI355X with stride 8: index num_records - 1 returns its
// record and the next one).
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ i32x4 llvm_struct_buffer_load_b128(i32x4 rsrc, int vindex, int voffset, int soffset,
                                              int aux) __asm("llvm.amdgcn.struct.buffer.load.v4i32");
// The same, 8 bytes (the first half of a 16-byte pair entry).
typedef int i32x2 __attribute__((ext_vector_type(2)));
__device__ i32x2 llvm_struct_buffer_load_b64(i32x4 rsrc, int vindex, int voffset, int soffset,
                                             int aux) __asm("llvm.amdgcn.struct.buffer.load.v2i32");
// Resource over one profile row of `ld` complex64 bins as records of 8 bytes,
// num_records = ld - 1: a load at bin n reads bins n and n + 1 for n <= ld - 2
// and zeros beyond, so it never reads past the row.  The bin index goes to the
// load as is (no byte-offset shift per sample).
__device__ __forceinline__ i32x4 row_pair_rsrc(const float2 *row, const int ld) {
    const unsigned long long a = (unsigned long long)row;
    i32x4 r;
    r.x = (int)(unsigned)a;
    r.y = (int)(((unsigned)(a >> 32) & 0xffffu) | (8u << 16));  // stride 8 bytes
    r.z = ld - 1;
    r.w = kBufferWord3;
    return r;
}

// The same over a pair row (16-byte entries, the last one never valid).
__device__ __forceinline__ i32x4 pair_rsrc(const float4 *row, const int ld) {
    const unsigned long long a = (unsigned long long)row;
    i32x4 r;
    r.x = (int)(unsigned)a;
    r.y = (int)(((unsigned)(a >> 32) & 0xffffu) | (16u << 16));  // stride 16 bytes
    r.z = ld - 1;
    r.w = kBufferWord3;
    return r;
}

// Read-only elements at wave-uniform indices, through the constant address
// space so that the compiler uses scalar loads (for global pointers it must
// first prove that no store of the kernel reaches the memory, a proof it
// abandons in the larger loops).
typedef const __attribute__((address_space(4))) double *const_f64_ptr;
typedef const __attribute__((address_space(4))) int *const_i32_ptr;
static_assert(offsetof(SeriesGather, phi) == 48 && offsetof(SeriesGather, phi_pair) == 56 &&
                  sizeof(SeriesGather) == 64,
              "SeriesGather field order");
__device__ __forceinline__ SeriesGather load_uniform(const SeriesGather *p, size_t i) {
    const const_f64_ptr d = (const_f64_ptr)(p + i);
    SeriesGather g;
    g.bx = d[0];
    g.by = d[1];
    g.k1 = d[2];
    g.k2 = d[3];
    g.k3 = d[4];
    g.cu = d[5];
    g.phi = d[6];
    const long long w = __double_as_longlong(d[7]);
    g.phi_pair = __uint_as_float((unsigned)w);
    g.pad = __uint_as_float((unsigned)(w >> 32));
    return g;
}
__device__ __forceinline__ int load_uniform(const int *p, size_t i) {
    return ((const_i32_ptr)(p + i))[0];
}
static_assert(offsetof(PulseConst, rx) == 24 && offsetof(PulseConst, r_srp) == 48 &&
                  sizeof(PulseConst) == 64,
              "PulseConst field order");
__device__ __forceinline__ PulseConst load_uniform(const PulseConst *p, size_t i) {
    const const_f64_ptr d = (const_f64_ptr)(p + i);
    PulseConst c;
    c.tx[0] = d[0];
    c.tx[1] = d[1];
    c.tx[2] = d[2];
    c.rx[0] = d[3];
    c.rx[1] = d[4];
    c.rx[2] = d[5];
    c.r_srp = d[6];
    c.valid = d[7];
    return c;
}

// Two consecutive profile samples (8-byte aligned) read as one 16-byte load.
struct alignas(8) Float2x2 {
    float a, b, c, d;
};

__device__ __forceinline__ float2 cmul(const float2 a, const float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// K[p, j] for j in [0, N_w): conj(W[k]) * wfac[p] * Y[p, src(k)] on the N
// support positions, zero elsewhere.  Support: j < ceil(N/2) -> k = j;
// j >= N_w - floor(N/2) -> k = j - (N_w - N) (CPP:331-342; odd-N middle mode
// lands at j = floor(N/2)).  `centred` input applies ifftshift on read
// (adjoint.py:36).
__global__ __launch_bounds__(256) void window_pad_kernel(const float2 *__restrict__ Y,
                                                         const float2 *__restrict__ Wconj,
                                                         const float2 *__restrict__ wfac, int N,
                                                         int nw, int ld, int centred,
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
    K[(size_t)p * ld + j] = v;
}

// Row tails of the upsampled profiles: prof[p][nw + j] = prof[p][j mod nw] for
// j < ld - nw, so a gather can read bins n and n + 1 for any n < ld - 1
// without wrapping.
__global__ __launch_bounds__(256) void wrap_pad_kernel(float2 *__restrict__ prof, int nw, int ld) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int pad = ld - nw;
    if (j >= pad) return;
    float2 *row = prof + (size_t)blockIdx.y * ld;
    row[nw + j] = row[j % nw];
}

// Y[p, k] = W[k] * wfac[p] * G[p, src(k)] (CPP:295-307); `centred` output
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

// The same from FP64 rows (the exact-range mode's rocFFT forward), products
// in FP64 and one rounding of each output.
__global__ __launch_bounds__(256) void truncate64_kernel(const double2 *__restrict__ G,
                                                         const float2 *__restrict__ W,
                                                         const float2 *__restrict__ wfac, int N,
                                                         int nw, int centred,
                                                         float2 *__restrict__ Y) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (k >= N) return;
    const int lo = (N + 1) / 2;
    const int src = k < lo ? k : nw - N + k;
    const double2 g = G[(size_t)p * nw + src];
    const double wr = W[k].x, wi = W[k].y, pr = wfac[p].x, pi = wfac[p].y;
    const double ar = fma(wr, g.x, -wi * g.y), ai = fma(wr, g.y, wi * g.x);
    const int dst = centred ? (k + N / 2) % N : k;
    Y[(size_t)p * N + dst] = make_float2((float)fma(ar, pr, -ai * pi), (float)fma(ar, pi, ai * pr));
}

// dst[i] = src[i] in FP64 (profile rows into the FP64 transform's buffer).
__global__ __launch_bounds__(256) void widen_kernel(const float2 *__restrict__ src,
                                                    double2 *__restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        dst[i] = make_double2((double)src[i].x, (double)src[i].y);
}

// Zero `bytes` bytes at a 16-byte aligned `buf`: 16-byte stores, then the
// remaining bytes % 16 (e.g. the last complex64 of an odd element count).
__global__ __launch_bounds__(256) void fill_zero_kernel(float4 *__restrict__ buf, size_t bytes) {
    const size_t n4 = bytes / 16;
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < n4; i += stride) buf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (blockIdx.x == 0 && threadIdx.x < bytes - 16 * n4)
        ((unsigned char *)(buf + n4))[threadIdx.x] = 0;
}

// Demodulated pairs of the adjoint profile rows (the gather's pair path).
// With u = n + f (n the bin, f in [0, 1)), the gather's sample times its phasor
//   (h[n] + f (h[n+1] - h[n])) e^{-2 pi i (rho u + phi)}
//     = (A_n + f D_n) e^{-2 pi i (rho f + phi - rho / 2)}
// with (A_n, D_n) = (h[n], h[n+1] - h[n]) e^{-2 pi i rho (n + 1/2)}, so the
// gather forms the pixel phase rho f + phi' in FP32 (|.| < rho / 2 + 1/2)
// instead of reducing rho u + phi in FP64, and interpolates with two FMAs.
// Each pair is formed in FP64 from the FP32 row and rounded once; the phasor
// is sincospi of the phase reduced in FP64.  One workgroup per valid row: the
// bins of its span (wrapping at N_w; the bins the gather can read, with
// margin) and their tail copies, as the spectral stage stores them.
// AONLY (the nearest rules): A_n alone, rows of `ld` float2 entries.
template <bool AONLY>
__global__ __launch_bounds__(256) void pair_rows_kernel(const float2 *__restrict__ prof,
                                                        const int2 *__restrict__ spans,
                                                        const int *__restrict__ rows, int nw,
                                                        int ld, double rho,
                                                        void *__restrict__ pairs) {
    const int p = rows[blockIdx.x];
    const int2 sp = spans[p];  // first bin in [0, nw), count <= nw
    const float2 *row = prof + (size_t)p * ld;
    for (int d = (int)threadIdx.x; d < sp.y; d += 256) {
        int n = sp.x + d;
        n -= n >= nw ? nw : 0;
        for (int i = n; i < ld - 1; i += nw) {
            const float2 a = row[i];
            double ph = rho * ((double)i + 0.5);
            ph -= floor(ph);
            double sn, cs;
            sincospi(2.0 * ph, &sn, &cs);
            const double ar = a.x, ai = a.y;
            const float2 A = make_float2((float)(ar * cs + ai * sn), (float)(ai * cs - ar * sn));
            if constexpr (AONLY) {
                ((float2 *)pairs)[(size_t)p * ld + i] = A;
            } else {
                const float2 b = row[i + 1];
                const double dr = (double)b.x - ar, di = (double)b.y - ai;
                ((float4 *)pairs)[(size_t)p * ld + i] =
                    make_float4(A.x, A.y, (float)(dr * cs + di * sn), (float)(di * cs - dr * sn));
            }
        }
    }
}

// amps_perm[i] = amps[perm[i]]: pixel amplitudes into tile (Morton) order.
__global__ __launch_bounds__(256) void gather_perm_kernel(const float2 *__restrict__ src,
                                                          const int *__restrict__ perm, int M,
                                                          float2 *__restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < M) dst[i] = src[perm[i]];
}

// ---------------------------------------------------------------------------
// Series constants per (tile, pulse), evaluated once per plan in FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void leg_consts(const double *T, const double4 X0, double &bx,
                                           double &by, double &bz, double &c, double &A) {
    const double ax = T[0] - X0.x, ay = T[1] - X0.y, az = T[2] - X0.z;
    const double A2 = fma(ax, ax, fma(ay, ay, az * az));
    A = sqrt(A2);
    c = 1.0 / A2;
    bx = -2.0 * ax * c;
    by = -2.0 * ay * c;
    bz = -2.0 * az * c;
}

template <bool MONO>
__global__ __launch_bounds__(256) void series_consts_kernel(const double4 *__restrict__ tile_c,
                                                            int num_tiles,
                                                            const PulseConst *__restrict__ pc,
                                                            const int *__restrict__ rows, int Pv,
                                                            double by_dx, double kc_rev,
                                                            void *__restrict__ out) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)num_tiles * Pv) return;
    const int t = (int)(idx / Pv), j = (int)(idx % Pv);
    const PulseConst c = pc[rows[j]];
    const double4 X0 = tile_c[t];
    if (MONO) {
        SeriesMono s;
        double A;
        leg_consts(c.tx, X0, s.bx, s.by, s.bz, s.c, A);
        const double C = (A + A) - c.r_srp;  // dR at the tile centre
        s.ku = 2.0 * A * by_dx;
        s.cu = C * by_dx;
        s.kt = 2.0 * A * kc_rev;
        const double th = C * kc_rev;
        s.ct = th - rint(th);
        ((SeriesMono *)out)[idx] = s;
    } else {
        SeriesBi s;
        double At, Ar;
        leg_consts(c.tx, X0, s.tbx, s.tby, s.tbz, s.tc, At);
        leg_consts(c.rx, X0, s.rbx, s.rby, s.rbz, s.rc, Ar);
        const double C = (At + Ar) - c.r_srp;
        s.kut = At * by_dx;
        s.kur = Ar * by_dx;
        s.cu = C * by_dx;
        s.ktt = At * kc_rev;
        s.ktr = Ar * kc_rev;
        const double th = C * kc_rev;
        s.ct = th - rint(th);
        s.pad0 = s.pad1 = 0.0;
        ((SeriesBi *)out)[idx] = s;
    }
}

// ---------------------------------------------------------------------------
// Per-pixel range evaluation for the four precision modes.
// ---------------------------------------------------------------------------
struct Pix {
    double x, y, z;     // absolute position (exact input; tile centre if non-finite)
    double dx, dy, dz;  // offset from the tile centre (exact)
    double d2;          // |d|^2
    bool ok;            // position was finite
};

__device__ __forceinline__ Pix load_pix(const KernelParams &k, int i, const double4 X0) {
    Pix p;
    const bool act = i < k.M;
    const int ii = act ? i : k.M - 1;
    double x = k.px[ii], y = k.py[ii], z = k.pz[ii];
    p.ok = act && isfinite(x) && isfinite(y) && isfinite(z);
    if (!p.ok) {
        x = X0.x;
        y = X0.y;
        z = X0.z;
    }
    p.x = x;
    p.y = y;
    p.z = z;
    p.dx = x - X0.x;
    p.dy = y - X0.y;
    p.dz = z - X0.z;
    p.d2 = fma(p.dx, p.dx, fma(p.dy, p.dy, p.dz * p.dz));
    return p;
}

// Uniform per-pulse constants: the pulse record and (series / hilo) the
// (tile, pulse) series record.  All loads are wave-uniform (scalar).
template <bool MONO>
struct PulseData {
    PulseConst pc;
    typename std::conditional<MONO, SeriesMono, SeriesBi>::type s;
};

template <int PREC, bool MONO>
__device__ __forceinline__ void load_pulse(const KernelParams &k, int tile, int j, int p,
                                           PulseData<MONO> &d) {
    if constexpr (PREC == kPrecFp64 || PREC == kPrecFp32) {
        d.pc = k.pc[p];
    } else {
        using S = typename std::conditional<MONO, SeriesMono, SeriesBi>::type;
        d.s = ((const S *)k.series)[(size_t)tile * k.Pv + j];
    }
}

// u = dR * by_dx (bins), th = dR * k_c / 2 pi (revolutions).
// PLANAR: every pixel offset has dz == 0 (the z terms are dropped).
template <int PREC, bool MONO, bool PLANAR = false, int SORD = 5>
__device__ __forceinline__ void eval_range(const KernelParams &k, const PulseData<MONO> &d,
                                           const Pix &q, double &u, double &th,
                                           const PolyRegs &pr) {
    if constexpr (PREC == kPrecFp64) {
        const double dr = ref_dr<MONO>(d.pc, q.x, q.y, q.z);
        u = dr * k.by_dx;
        th = dr * k.kc_rev;
    } else if constexpr (PREC == kPrecSeries) {
        if constexpr (MONO) {
            const double t = PLANAR ? fma(d.s.bx, q.dx, fma(d.s.by, q.dy, d.s.c * q.d2))
                                    : fma(d.s.bx, q.dx, fma(d.s.by, q.dy, fma(d.s.bz, q.dz, d.s.c * q.d2)));
            const double g = sqrt1pm1_series<SORD>(t, pr);
            u = fma(d.s.ku, g, d.s.cu);
            th = fma(d.s.kt, g, d.s.ct);
        } else {
            const double zt = PLANAR ? d.s.tc * q.d2 : fma(d.s.tbz, q.dz, d.s.tc * q.d2);
            const double zr = PLANAR ? d.s.rc * q.d2 : fma(d.s.rbz, q.dz, d.s.rc * q.d2);
            const double tt = fma(d.s.tbx, q.dx, fma(d.s.tby, q.dy, zt));
            const double tr = fma(d.s.rbx, q.dx, fma(d.s.rby, q.dy, zr));
            const double gt = sqrt1pm1_series<SORD>(tt, pr), gr = sqrt1pm1_series<SORD>(tr, pr);
            u = fma(d.s.kut, gt, fma(d.s.kur, gr, d.s.cu));
            th = fma(d.s.ktt, gt, fma(d.s.ktr, gr, d.s.ct));
        }
    } else if constexpr (PREC == kPrecHilo) {
        // FP64 tile-centre range (cu, ct) + FP32 differential range.
        const float dx = (float)q.dx, dy = (float)q.dy, dz = (float)q.dz, d2 = (float)q.d2;
        if constexpr (MONO) {
            const float tz = PLANAR ? (float)d.s.c * d2 : fmaf((float)d.s.bz, dz, (float)d.s.c * d2);
            const float t = fmaf((float)d.s.bx, dx, fmaf((float)d.s.by, dy, tz));
            const float g = sqrt1pm1_series_f(t);
            u = d.s.cu + (double)((float)d.s.ku * g);
            th = d.s.ct + (double)((float)d.s.kt * g);
        } else {
            const float zt = PLANAR ? (float)d.s.tc * d2 : fmaf((float)d.s.tbz, dz, (float)d.s.tc * d2);
            const float zr = PLANAR ? (float)d.s.rc * d2 : fmaf((float)d.s.rbz, dz, (float)d.s.rc * d2);
            const float tt = fmaf((float)d.s.tbx, dx, fmaf((float)d.s.tby, dy, zt));
            const float tr = fmaf((float)d.s.rbx, dx, fmaf((float)d.s.rby, dy, zr));
            const float gt = sqrt1pm1_series_f(tt), gr = sqrt1pm1_series_f(tr);
            u = d.s.cu + (double)fmaf((float)d.s.kut, gt, (float)d.s.kur * gr);
            th = d.s.ct + (double)fmaf((float)d.s.ktt, gt, (float)d.s.ktr * gr);
        }
    } else {
        // all-FP32 absolute coordinates
        const float x = (float)q.x, y = (float)q.y, z = (float)q.z;
        const float ax = (float)d.pc.tx[0] - x, ay = (float)d.pc.tx[1] - y, az = (float)d.pc.tx[2] - z;
        const float d1 = sqrtf(fmaf(ax, ax, fmaf(ay, ay, az * az)));
        float dr;
        if constexpr (MONO) {
            dr = 2.0f * d1 - (float)d.pc.r_srp;
        } else {
            const float bx = (float)d.pc.rx[0] - x, by = (float)d.pc.rx[1] - y,
                        bz = (float)d.pc.rx[2] - z;
            dr = d1 + sqrtf(fmaf(bx, bx, fmaf(by, by, bz * bz))) - (float)d.pc.r_srp;
        }
        u = (double)(dr * (float)k.by_dx);
        th = (double)(dr * (float)k.kc_rev);
    }
}

// Series ranges with the bin scale folded into the polynomial, for kernels
// that amortise per-(tile, pulse) constants over many pixels (the scatter's
// lanes each own a pulse): per leg k_i = ku * c_i with c = (1/2, -1/8, 1/16,
// -5/128, 7/256) the coefficients of g(t) = sqrt(1 + t) - 1, so
// u = cu + sum_i k_i t^i by Horner's rule, and theta = rho u + phi with
// rho = kc_rev / by_dx (u and theta are both dR times a constant) and
// phi = ct - rho cu.  One FP64 instruction fewer per leg than forming g(t)
// and then u and theta from it.
template <bool MONO, int SORD>
struct SeriesHorner {
    double kt[SORD];             // leg T (monostatic: the doubled leg)
    double kr[MONO ? 1 : SORD];  // leg R
    double cu, phi;
};

template <bool MONO, int SORD>
__device__ __forceinline__ SeriesHorner<MONO, SORD> make_horner(const PulseData<MONO> &d,
                                                                const double rho) {
    constexpr double c[5] = {0.5, -0.125, 0.0625, -0.0390625, 0.02734375};
    SeriesHorner<MONO, SORD> h;
#pragma unroll
    for (int i = 0; i < SORD; ++i) {
        if constexpr (MONO) {
            h.kt[i] = d.s.ku * c[i];
        } else {
            h.kt[i] = d.s.kut * c[i];
            h.kr[i] = d.s.kur * c[i];
        }
    }
    h.cu = d.s.cu;
    h.phi = fma(-rho, d.s.cu, d.s.ct);
    return h;
}

// add + sum_i k_i t^(i+1), i < SORD
template <int SORD>
__device__ __forceinline__ double horner(const double t, const double (&k)[SORD], const double add) {
    double v = k[SORD - 1];
#pragma unroll
    for (int i = SORD - 2; i >= 0; --i) v = fma(t, v, k[i]);
    return fma(t, v, add);
}

template <bool MONO, bool PLANAR, int SORD>
__device__ __forceinline__ void eval_range_horner(const PulseData<MONO> &d,
                                                  const SeriesHorner<MONO, SORD> &h, const Pix &q,
                                                  const double rho, double &u, double &th) {
    if constexpr (MONO) {
        const double t = PLANAR ? fma(d.s.bx, q.dx, fma(d.s.by, q.dy, d.s.c * q.d2))
                                : fma(d.s.bx, q.dx, fma(d.s.by, q.dy, fma(d.s.bz, q.dz, d.s.c * q.d2)));
        u = horner<SORD>(t, h.kt, h.cu);
    } else {
        const double zt = PLANAR ? d.s.tc * q.d2 : fma(d.s.tbz, q.dz, d.s.tc * q.d2);
        const double zr = PLANAR ? d.s.rc * q.d2 : fma(d.s.rbz, q.dz, d.s.rc * q.d2);
        const double tt = fma(d.s.tbx, q.dx, fma(d.s.tby, q.dy, zt));
        const double tr = fma(d.s.rbx, q.dx, fma(d.s.rby, q.dy, zr));
        u = horner<SORD>(tr, h.kr, horner<SORD>(tt, h.kt, h.cu));
    }
    th = fma(rho, u, h.phi);
}

// Nearest-neighbour bin (unwrapped, integer-valued double) with the
// reference's rule; the series path re-evaluates exactly near a tie.
template <int INTERP, int PREC, bool MONO>
__device__ __forceinline__ double nearest_bin(const KernelParams &k, const Pix &q, int p, double &u,
                                              double &th) {
    constexpr bool REF = INTERP == kInterpNearestRef;
    double v = REF ? 0.5 + u : u;
    if constexpr (PREC == kPrecSeries) {
        const double fl = floor(v);
        const double fr = v - fl;
        if (fabs(fr - 0.5) >= kTieMargin) return fr > 0.5 ? fl + 1.0 : fl;
        // rare: exact reference arithmetic (CPP:287-289 / CPP:346-347)
        const PulseConst c = k.pc[p];
        const double dr = ref_dr<MONO>(c, q.x, q.y, q.z);
        u = dr * k.by_dx;
        th = dr * k.kc_rev;
        v = REF ? 0.5 + u : u;
    }
    return round(v);
}

// Bin index in [0, nw) of an unwrapped integer-valued bin; `span` = 1 for
// nearest, 2 for linear (so n + 1 stays in range on the fast path).
__device__ __forceinline__ int bin_index(const KernelParams &k, double n, int span) {
    int ni = (int)n;
    if ((unsigned)ni > (unsigned)(k.nw - span)) {
        ni = wrap_bin(n, k.nwd, k.inv_nw, k.nw);
        ni = min(max(ni, 0), k.nw - 1);
    }
    return ni;
}

// ---------------------------------------------------------------------------
// Backprojection gather
//
// A workgroup owns one tile of kTilePixels consecutive pixels of the Morton
// order (kPixPerThread per thread).  The pulse loop is wave-uniform: per-pulse
// or per-(tile, pulse) constants come through scalar loads.  Profile samples
// are gathered from the upsampled range profiles (L1/L2-resident windows).
// Phase in FP64 revolutions, integer part removed, FP32 sin/cos.  Products and
// sums in FP32; every 64 pulses the per-pixel FP32 sums are folded into FP64
// sums kept in LDS (out of registers, for occupancy).
// ---------------------------------------------------------------------------
template <int INTERP, bool MONO, int PREC, bool PLANAR, int SORD = 5, bool PAIRS = false>
__global__ __launch_bounds__(kThreads) void bp_gather_kernel(const KernelParams k) {
    __shared__ double acc[2 * kPixPerThread][kThreads];
    const PolyRegs pr = make_poly_regs<SORD>();
    // grid: one workgroup per (tile, pulse split), split s taking the valid
    // pulses [j_begin, j_end).  XCD-aware order: workgroups are dealt
    // round-robin to the 8 XCDs, so XCD x gets the x-th eighth of the pairs
    // in (split, tile) order, tiles in Morton order: the workgroups an XCD
    // runs together are neighbouring tiles on the same pulses, whose profile
    // bins its L2 then serves to all of them.  (Blocks past the last pair
    // exit.)
    const int nsplit = k.bp_splits;
    const long long npairs = (long long)k.num_tiles * nsplit;
    const long long chunk = (npairs + 7) >> 3;
    const long long f = (long long)(blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
    if (f >= npairs) return;
    const int tile = (int)(f % k.num_tiles);
    const int split = (int)(f / k.num_tiles);
    const int j_begin = (int)((long long)k.Pv * split / nsplit);
    const int j_end = (int)((long long)k.Pv * (split + 1) / nsplit);
    const int tid = (int)threadIdx.x;
    const double4 X0 = k.tile_c[tile];
    Pix q[kPixPerThread];
    float ar[kPixPerThread], ai[kPixPerThread];
#pragma unroll
    for (int m = 0; m < kPixPerThread; ++m) {
        q[m] = load_pix(k, tile * kTilePixels + m * kThreads + tid, X0);
        ar[m] = ai[m] = 0.f;
        acc[2 * m][tid] = acc[2 * m + 1][tid] = 0.0;
    }
    constexpr bool kShift = INTERP == kInterpLinear && (PREC == kPrecSeries || PREC == kPrecHilo);
    // The next pulse's row index and constants are requested one iteration
    // ahead (scalar loads: a wait at the top of the loop would otherwise
    // expose their latency on every pulse).  With kShift, u of the tile
    // centre comes moved by a multiple of nw (bp_shift_kernel, once per
    // plan) so that every pixel of the tile lands in [2, nw + 2 shift_lo - 2):
    // bins are unchanged modulo nw and the padded row tail makes bins
    // n + 1 <= ld - 1 valid without a wrap.
    // kHG: one SeriesGather record per (tile, pulse) replaces the series
    // constants and the row shift (Horner form, phase from u); the nearest
    // rules read them on the pair path only, with the rule's offset in cu.
    constexpr bool kHG = (kShift || PAIRS) && MONO && PLANAR && PREC == kPrecSeries && SORD == 3;
    const double rho = k.rho;
    const float rho32 = (float)rho;
    int p_next = 0;
    PulseData<MONO> d_next;
    SeriesGather g_next;
    if (j_begin < j_end) {
        if constexpr (kHG) {
            p_next = load_uniform(k.rows, (size_t)j_begin);
            g_next = load_uniform(k.series_g, (size_t)tile * k.Pv + j_begin);
        } else {
            p_next = k.rows[j_begin];
            load_pulse<PREC, MONO>(k, tile, j_begin, p_next, d_next);
            if constexpr (kShift) d_next.s.cu = k.bp_cu[(size_t)tile * k.Pv + j_begin];
        }
    }
    for (int j = j_begin; j < j_end; ++j) {
        const int p = p_next;
        PulseData<MONO> d = d_next;
        const SeriesGather g = g_next;
        if (j + 1 < j_end) {
            if constexpr (kHG) {
                p_next = load_uniform(k.rows, (size_t)j + 1);
                g_next = load_uniform(k.series_g, (size_t)tile * k.Pv + j + 1);
            } else {
                p_next = k.rows[j + 1];
                load_pulse<PREC, MONO>(k, tile, j + 1, p_next, d_next);
                if constexpr (kShift) d_next.s.cu = k.bp_cu[(size_t)tile * k.Pv + j + 1];
            }
        }
        const float2 *row = k.prof + (size_t)p * k.ld;
        // 16-byte loads of bins (n, n + 1) through a structured buffer over the
        // padded row: indices beyond it read zeros instead of faulting.
        const i32x4 rsrc = row_pair_rsrc(row, k.ld);
        if constexpr (PAIRS && INTERP != kInterpLinear) {
            // nearest rules on the demodulated rows (A_n alone, pair_rows_kernel
            // with AONLY): the record's cu carries the rule's offset o (1 for
            // round(1/2 + u), 1/2 for round(u)), so n = floor(u + o) and f =
            // u + o - n; one 8-byte load of A_n = h[n] e^{-2 pi i rho (n + 1/2)},
            // whose pixel phase is rho f + phi' as on the linear pair path.
            // Within 2e-7 bin of a bin edge the reference's rule decides on
            // the exact range (CPP:346-347, CPP:289), a rare lane-divergent
            // branch.
            static_assert(kHG, "the pair path reads SeriesGather records");
            constexpr bool REF = INTERP == kInterpNearestRef;
            const i32x4 psrc = row_pair_rsrc((const float2 *)k.bp_pairs + (size_t)p * k.ld, k.ld);
            int n0[kPixPerThread];
            float f[kPixPerThread];
            bool tie = false;
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                const double t = fma(g.bx, q[m].dx, fma(g.by, q[m].dy, q[m].d2));
                const double u = fma(t, fma(t, fma(t, g.k3, g.k2), g.k1), g.cu);
                n0[m] = (int)u;
                f[m] = (float)__builtin_amdgcn_fract(u);
                tie = tie || fabsf(f[m] - 0.5f) > 0.5f - 2.0f * (float)kTieMargin;
            }
            if (__builtin_amdgcn_read_exec() & __ballot(tie)) {
#pragma unroll
                for (int m = 0; m < kPixPerThread; ++m) {
                    // (pixels that are not ok contribute nothing: no exact range)
                    if (q[m].ok && fabsf(f[m] - 0.5f) > 0.5f - 2.0f * (float)kTieMargin) {
                        // the exact range of the pixel
                        int i = tile * kTilePixels + m * kThreads + tid;
                        // formed here, not kept in registers across the pulse loop
                        asm volatile("" : "+v"(i));
                        const double x = k.px[i], y = k.py[i], z = k.pz[i];
                        const double ue = ref_dr<MONO>(load_uniform(k.pc, (size_t)p), x, y, z) * k.by_dx;
                        const double ne = round(REF ? 0.5 + ue : ue);
                        // u - o - ue is the row shift, a multiple of nw (u to
                        // 1e-7 bin from n0 + f)
                        const double u = (double)n0[m] + (double)f[m];
                        const int ns = (int)(ne + rint(u - (REF ? 1.0 : 0.5) - ue));
                        f[m] -= (float)(ns - n0[m]);
                        n0[m] = ns;
                    }
                    // one pixel at a time (keeps the registers of the rare path few)
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            i32x2 h[kPixPerThread];
            float sn[kPixPerThread], cs[kPixPerThread];
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                h[m] = llvm_struct_buffer_load_b64(psrc, n0[m], 0, 0, 0);
                const float th = fmaf(rho32, f[m], g.phi_pair);
                sn[m] = __builtin_amdgcn_sinf(th);
                cs[m] = __builtin_amdgcn_cosf(th);
            }
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                const float hr = __int_as_float(h[m].x), hi = __int_as_float(h[m].y);
                ar[m] = fmaf(hr, cs[m], fmaf(hi, sn[m], ar[m]));
                ai[m] = fmaf(hi, cs[m], fmaf(-hr, sn[m], ai[m]));
            }
        } else if constexpr (PAIRS) {
            // demodulated pair rows (pair_rows_kernel): one 16-byte load of
            // (A_n, D_n), the phase rho f + phi' in FP32
            static_assert(kHG, "the pair path reads SeriesGather records");
            const i32x4 psrc = pair_rsrc(k.bp_pairs + (size_t)p * k.ld, k.ld);
            u32x4 h[kPixPerThread];
            float fr[kPixPerThread], sn[kPixPerThread], cs[kPixPerThread];
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                const double t = fma(g.bx, q[m].dx, fma(g.by, q[m].dy, q[m].d2));
                const double u = fma(t, fma(t, fma(t, g.k3, g.k2), g.k1), g.cu);
                const int n0 = (int)u;
                fr[m] = (float)__builtin_amdgcn_fract(u);
                h[m] = (u32x4)llvm_struct_buffer_load_b128(psrc, n0, 0, 0, 0);
                const float th = fmaf(rho32, fr[m], g.phi_pair);
                sn[m] = __builtin_amdgcn_sinf(th);
                cs[m] = __builtin_amdgcn_cosf(th);
            }
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                const float hr = fmaf(fr[m], __uint_as_float(h[m].z), __uint_as_float(h[m].x));
                const float hi = fmaf(fr[m], __uint_as_float(h[m].w), __uint_as_float(h[m].y));
                ar[m] = fmaf(hr, cs[m], fmaf(hi, sn[m], ar[m]));
                ai[m] = fmaf(hi, cs[m], fmaf(-hr, sn[m], ai[m]));
            }
        } else if constexpr (INTERP == kInterpLinear && kShift) {
            // all pixels' bins and phases first, so that their profile loads
            // are in flight together, then interpolation and accumulation
            u32x4 h[kPixPerThread];
            float fr[kPixPerThread], sn[kPixPerThread], cs[kPixPerThread];
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                double u, th;
                if constexpr (kHG) {
                    const double t = fma(g.bx, q[m].dx, fma(g.by, q[m].dy, q[m].d2));
                    u = fma(t, fma(t, fma(t, g.k3, g.k2), g.k1), g.cu);
                    th = fma(rho, u, g.phi);
                } else {
                    eval_range<PREC, MONO, PLANAR, SORD>(k, d, q[m], u, th, pr);
                }
                // u > 0: truncation is floor, v_fract_f64 the exact rest
                const int n0 = (int)u;
                fr[m] = (float)__builtin_amdgcn_fract(u);
                h[m] = (u32x4)llvm_struct_buffer_load_b128(rsrc, n0, 0, 0, 0);
                phasor_fract(th, &sn[m], &cs[m]);
            }
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                const float h0x = __uint_as_float(h[m].x), h0y = __uint_as_float(h[m].y);
                const float hr = fmaf(fr[m], __uint_as_float(h[m].z) - h0x, h0x);
                const float hi = fmaf(fr[m], __uint_as_float(h[m].w) - h0y, h0y);
                // h * e^{-i theta}
                ar[m] = fmaf(hr, cs[m], fmaf(hi, sn[m], ar[m]));
                ai[m] = fmaf(hi, cs[m], fmaf(-hr, sn[m], ai[m]));
            }
        } else {
#pragma unroll
        for (int m = 0; m < kPixPerThread; ++m) {
            double u, th;
            eval_range<PREC, MONO, PLANAR, SORD>(k, d, q[m], u, th, pr);
            float hr, hi;
            if constexpr (INTERP == kInterpLinear) {
                const double fl = floor(u);
                const float fr = (float)(u - fl);
                const int n0 = bin_index(k, fl, 1);
                const u32x4 h = (u32x4)llvm_struct_buffer_load_b128(rsrc, n0, 0, 0, 0);
                const float h0x = __uint_as_float(h.x), h0y = __uint_as_float(h.y);
                hr = fmaf(fr, __uint_as_float(h.z) - h0x, h0x);
                hi = fmaf(fr, __uint_as_float(h.w) - h0y, h0y);
            } else {
                const double n = nearest_bin<INTERP, PREC, MONO>(k, q[m], p, u, th);
                const float2 h = row[bin_index(k, n, 1)];
                hr = h.x;
                hi = h.y;
            }
            float s, c;
            phasor_rev(th, &s, &c);
            // h * e^{-i theta}
            ar[m] = fmaf(hr, c, fmaf(hi, s, ar[m]));
            ai[m] = fmaf(hi, c, fmaf(-hr, s, ai[m]));
        }
        }
        if (((j - j_begin) & 63) == 63) {
#pragma unroll
            for (int m = 0; m < kPixPerThread; ++m) {
                acc[2 * m][tid] += (double)ar[m];
                acc[2 * m + 1][tid] += (double)ai[m];
                ar[m] = ai[m] = 0.f;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < kPixPerThread; ++m) {
        const int i = tile * kTilePixels + m * kThreads + tid;
        if (i < k.M) {
            const double vr = q[m].ok ? acc[2 * m][tid] + (double)ar[m] : 0.0;
            const double vi = q[m].ok ? acc[2 * m + 1][tid] + (double)ai[m] : 0.0;
            if (nsplit == 1) k.out[k.perm[i]] = make_float2((float)vr, (float)vi);
            else k.bp_partial[(size_t)split * k.M + i] = make_double2(vr, vi);
        }
    }
}

// Sum of the pulse-split partial images in split order (FP64), scattered back
// to the caller's pixel order.
__global__ __launch_bounds__(256) void bp_reduce_kernel(const double2 *__restrict__ part, int M,
                                                        int nsplit, const int *__restrict__ perm,
                                                        float2 *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M) return;
    double vr = 0.0, vi = 0.0;
    for (int s = 0; s < nsplit; ++s) {
        const double2 v = part[(size_t)s * M + i];
        vr += v.x;
        vi += v.y;
    }
    out[perm[i]] = make_float2((float)vr, (float)vi);
}

// ---------------------------------------------------------------------------
// Forward scatter
//
// Grid (regions) x (groups of G valid pulses).  A region is R consecutive
// series tiles of the Morton order.  Lane q of the workgroup owns pulse
// j0 + (q mod G): each pulse has its own LDS window covering every bin the
// region's pixels can reach for that pulse (bounding sphere, |dR - dR_centre|
// <= 2 radius, two-bin margins), and the 256/G thread streams walk disjoint
// pixels of the region (pixel records are broadcast loads).  Lanes of one
// stream never share a window; different streams only collide on the same
// bin at the same instant.
//
// Window bins accumulate in 32-bit fixed point with LDS integer atomics: on
// gfx950 ds_add_u32 runs ~40x faster than ds_add_f32.  The scale is
// 2^31 / (1.001 B) (FixScale) for the region's bound B on every bin
// sum (region_l1_kernel), and the conversion to integers truncates toward
// zero, which never increases a magnitude, so every sum stays inside int32.
// FP32 pre-scaling and products carry ~6e-8 relative error and the
// truncation less than one unit per contribution.  The
// integer parts of the two linear-interpolation shares sum to the truncated
// total.  A region whose amplitudes are not finite (and any contribution
// outside its window, which the window bounds rule out) cannot be windowed:
// with cleared rows it goes to float global atomics, on the span path the
// row is poisoned instead.
//
// Windows are written whole (zeros included) to `partial` as their integer
// sums, with liveness flags (the bases are fixed per plan), then
// merged per pulse by fp_reduce_kernel in FP64 in a fixed order: no global
// atomics on the main path, and a result that does not depend on scheduling.
// ---------------------------------------------------------------------------
// Toward zero (one v_cvt_i32_f32).  Its error is odd in v, so bin sums of
// contributions with mixed signs carry no bias; rounding halves always up
// (floor(v + 0.5)) biased every sum by about n/32 units, since the scaled
// values keep only a few fractional bits.  At the 2^31 scale the truncation
// error matches that of round-to-nearest at 2^30.
__device__ __forceinline__ int fix_tz(const float v) { return (int)v; }
// Fixed-point scale 2^31 / (1.001 b) for a bound b on every bin sum, held as
// m 2^e with m in (1, 2]: values are scaled by ldexp(v m, e) and sums taken
// back as s (1 / m) 2^-e, 1 / m = mm the mantissa of 1.001 b, in [1/2, 1).
// 2^31 / b itself leaves the FP32 range for b below about 6e-30 (and 1 / that
// scale for b below 1e-29 is subnormal); the split form stays exact there, so
// the forward stays linear for tiny inputs.  (A power-of-two scale would make
// the pre-scaling exact but leave up to one bit of the int32 range unused,
// doubling the truncation unit, the larger error of the two.)
struct FixScale {
    float m, inv_m;
    int e;
};
// b finite and > 0 (and 1.001 b finite).
__device__ __forceinline__ FixScale fix_scale(const float b) {
    int ex;
    const float mm = frexpf(b * 1.001f, &ex);  // 1.001 b = mm 2^ex, mm in [1/2, 1)
    FixScale f;
    f.m = 1.0f / mm;
    f.inv_m = mm;
    f.e = 31 - ex;
    return f;
}
__device__ __forceinline__ float fix_apply(const FixScale &f, const float v) {
    return ldexpf(v * f.m, f.e);
}
__device__ __forceinline__ float fix_undo(const FixScale &f, const float v) {
    return ldexpf(v * f.inv_m, -f.e);
}
// Whether a bound admits fixed-point windows: finite with 1.001 b finite.
__device__ __forceinline__ bool fix_bound_ok(const float b) { return isfinite(b * 1.001f); }
static_assert(offsetof(PixRec, dx) == 0 && offsetof(PixRec, d2) == 24 &&
                  offsetof(PixRec, x) == 32 && offsetof(PixRec, ok) == 56,
              "PixRec field order");
__device__ __forceinline__ PixRec load_uniform(const PixRec *p, size_t i) {
    const const_f64_ptr d = (const_f64_ptr)(p + i);
    PixRec r;
    r.dx = d[0];
    r.dy = d[1];
    r.dz = d[2];
    r.d2 = d[3];
    r.x = d[4];
    r.y = d[5];
    r.z = d[6];
    r.ok = d[7];
    return r;
}
__device__ __forceinline__ float2 load_uniform(const float2 *p, size_t i) {
    const double v = ((const_f64_ptr)(p + i))[0];
    return make_float2(__uint_as_float((unsigned)__double_as_longlong(v)),
                       __uint_as_float((unsigned)(__double_as_longlong(v) >> 32)));
}
// Records i and i + 1 at a wave-uniform index: one 64-byte scalar load.
__device__ __forceinline__ void load_srec2(const ScatRec *p, size_t i, ScatRec &a, ScatRec &b) {
    const const_f64_ptr d = (const_f64_ptr)(p + i);
    double v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = d[e];
    a.dx = v[0];
    a.dy = v[1];
    a.d2 = v[2];
    long long w = __double_as_longlong(v[3]);
    a.a = make_float2(__uint_as_float((unsigned)w), __uint_as_float((unsigned)(w >> 32)));
    b.dx = v[4];
    b.dy = v[5];
    b.d2 = v[6];
    w = __double_as_longlong(v[7]);
    b.a = make_float2(__uint_as_float((unsigned)w), __uint_as_float((unsigned)(w >> 32)));
}
// Fixed-point window bin: two 32-bit words (re, im), each accumulated by its
// own 32-bit LDS add (no carry between the components).
__device__ __forceinline__ void add_fix(unsigned int *bin, const int re, const int im) {
    atomicAdd(bin, (unsigned)re);
    atomicAdd(bin + 1, (unsigned)im);
}

// Window base (bin) of forward region `region` for pulse row p: for a leg with
// phase centre T, unit vector e from the region centre Xc to T and a pixel
// X = Xc + d in the region's box of half-extents h, |T - X| - |T - Xc| lies in
// [-s, s + r^2 / (2 (|T - Xc| - r))] with s = sum_i |e_i| h_i and r the region
// radius, so dR - dR(Xc) >= -(s_T + s_R); the base sits two bins below.
template <bool MONO>
__device__ __forceinline__ double window_base(const KernelParams &k, const int region,
                                              const int p) {
    const double4 rc = k.region_c[region];
    const double4 rh = k.region_h[region];
    double sleg;
    const PulseConst &c = k.pc[p];
    const double ax = c.tx[0] - rc.x, ay = c.tx[1] - rc.y, az = c.tx[2] - rc.z;
    const double inv = 1.0 / sqrt(fma(ax, ax, fma(ay, ay, az * az)));
    sleg = (fabs(ax) * rh.x + fabs(ay) * rh.y + fabs(az) * rh.z) * inv;
    if (!MONO) {
        const double bx = c.rx[0] - rc.x, by = c.rx[1] - rc.y, bz = c.rx[2] - rc.z;
        const double invr = 1.0 / sqrt(fma(bx, bx, fma(by, by, bz * bz)));
        sleg += (fabs(bx) * rh.x + fabs(by) * rh.y + fabs(bz) * rh.z) * invr;
    } else {
        sleg *= 2.0;
    }
    return floor((ref_dr<MONO>(c, rc.x, rc.y, rc.z) - sleg * (1.0 + 1e-9) - 1e-9) * k.by_dx) - 2.0;
}

// Once per plan: for each forward region, the largest number of valid pixels
// whose floor(u) is n - 1 or n, over the bins n and valid pulses (the count C
// of region_l1_kernel's bound).  u is the exact FP64 range (the reference's
// formula) relative to the scatter's window base; a pixel within 1e-6 bin of a
// bin edge counts on both sides, which covers the series modes' deviation
// from it (below 1e-9 bin).  One workgroup per (region, PB pulses), lane
// q % PB owns a pulse's histogram in LDS; a u outside the window (which the
// window bounds rule out) gives -1: no count bound for that region.
// cmax[region] must start at 0.
template <bool MONO>
__global__ __launch_bounds__(256) void region_count_kernel(const KernelParams k, int PB,
                                                           int *__restrict__ cmax) {
    extern __shared__ int hist[];  // [PB][wst]
    __shared__ int s_over;
    const int region = blockIdx.x;
    const int j0 = blockIdx.y * PB;
    const int tid = (int)threadIdx.x;
    const int q = tid % PB, stream = tid / PB, nstreams = 256 / PB;
    const int np = min(PB, k.Pv - j0);
    const bool qok = q < np;
    const int p = k.rows[j0 + (qok ? q : np - 1)];
    const int win = k.win, wst = k.win_stride;
    if (tid == 0) s_over = 0;
    for (int e = tid; e < PB * wst; e += 256) hist[e] = 0;
    __syncthreads();
    const double wb = window_base<MONO>(k, region, p);
    int *h = hist + q * wst;
    const int pb = region * k.region_tiles * kTilePixels;
    const int pe = min(k.M, pb + k.region_tiles * kTilePixels);
    bool over = false;
    for (int i = pb + stream; i < pe; i += nstreams) {
        const PixRec r = k.pixrec[i];
        if (r.ok == 0.0 || !qok) continue;
        const double u = ref_dr<MONO>(k.pc[p], r.x, r.y, r.z) * k.by_dx - wb;
        const double a = floor(u - 1e-6), b = floor(u + 1e-6);
        if (!(a >= 0.0) || !(b < (double)win)) {
            over = true;
            continue;
        }
        atomicAdd(&h[(int)a], 1);
        if (b != a) atomicAdd(&h[(int)b], 1);
    }
    if (over) s_over = 1;
    __syncthreads();
    if (tid < np) {
        int best = 0, prev = 0;
        for (int n = 0; n < win; ++n) {
            const int c = hist[tid * wst + n];
            best = max(best, prev + c);
            prev = c;
        }
        atomicMax(&cmax[region], s_over ? 0x7fffffff : best);
    }
}

// Per-region bound on every window bin sum of the forward, and the amplitudes
// scaled in place by the region's fixed-point scale (when windows are used and
// the bound is finite and non-zero), so the scatter forms fixed-point
// contributions without a multiply.
//
// A bin n of one pulse's window receives shares of weight <= 1 from the pixels
// whose floor(u) is n - 1 or n, each of magnitude <= |Re a| + |Im a| per
// component, so |sum| <= C a_max with C the largest number of such pixels over
// the region's bins and pulses (`cmax`, counted once per plan by
// region_count_kernel; NULL or negative: not counted) and a_max the largest
// |Re a| + |Im a| of the region.  The L1 norm sum(|Re a| + |Im a|) is a bound
// too; the smaller one is taken.  (For a 2048-pixel region C is about 64, and
// C a_max is about a ninth of the L1 norm of a dense image: the fixed-point
// unit, bound / 2^31, is as much finer.)  Sums in FP32 in a fixed order;
// a non-finite L1 norm is kept (the scatter then takes no windows).
__global__ __launch_bounds__(kThreads) void region_l1_kernel(float2 *__restrict__ amps,
                                                             int M, int region_pixels,
                                                             int windows,
                                                             float *__restrict__ out,
                                                             const int *__restrict__ cmax,
                                                             const PixRec *__restrict__ pixrec,
                                                             ScatRec *__restrict__ srec) {
    __shared__ float red[kThreads], redm[kThreads];
    const int base = blockIdx.x * region_pixels;
    const int n = min(region_pixels, M - base);
    float l1 = 0.f, am = 0.f;
    for (int i = threadIdx.x; i < n; i += kThreads) {
        const float2 a = amps[base + i];
        const float v = fabsf(a.x) + fabsf(a.y);
        l1 += v;
        am = fmaxf(am, v);
    }
    red[threadIdx.x] = l1;
    redm[threadIdx.x] = am;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[threadIdx.x] += red[threadIdx.x + s];
            redm[threadIdx.x] = fmaxf(redm[threadIdx.x], redm[threadIdx.x + s]);
        }
        __syncthreads();
    }
    float bound = red[0];
    if (cmax && isfinite(bound)) {
        const int c = cmax[blockIdx.x];
        if (c >= 0) bound = fminf(bound, (float)c * redm[0]);
    }
    if (threadIdx.x == 0) out[blockIdx.x] = bound;
    const bool scale = windows && fix_bound_ok(bound) && bound > 0.f;
    FixScale fs{1.f, 1.f, 0};
    if (scale) fs = fix_scale(bound);
    if (scale || srec) {
        for (int i = threadIdx.x; i < n; i += kThreads) {
            float2 a = amps[base + i];
            if (scale) {
                a = make_float2(fix_apply(fs, a.x), fix_apply(fs, a.y));
                amps[base + i] = a;
            }
            if (srec) {
                // planar scatter records: offsets from the tile centre and the
                // (scaled) amplitude, zero for invalid pixels
                const PixRec &q = pixrec[base + i];
                ScatRec r;
                r.dx = q.dx;
                r.dy = q.dy;
                r.d2 = q.d2;
                r.a = q.ok != 0.0 ? a : make_float2(0.f, 0.f);
                srec[base + i] = r;
            }
        }
    }
}

constexpr int kScatGroup = 4;  // pixel records per scalar-load group (compact path)
// s_waitcnt lgkmcnt(0): scalar loads return out of order, so their only wait
// is for every outstanding scalar and LDS operation.
__device__ __forceinline__ void lgkm_wait0() { __builtin_amdgcn_s_waitcnt(0xc07f); }
template <int INTERP, bool MONO, int PREC, bool PLANAR, int NT, int SORD = 5>
__global__ __launch_bounds__(NT) void fp_scatter_kernel(const KernelParams k) {
    // Fixed-point windows [G][wst] of bins, a bin being two 32-bit words
    // (re, im) accumulated by 32-bit LDS adds: the four adds of a linear
    // contribution share one address (immediate offsets 0, 4, 8, 12), and
    // each component sum stays below 2^31 in magnitude.  (One 64-bit add of
    // im * 2^32 + re needs a sign extension and a 64-bit shift-add per
    // contribution, five VALU instructions more.)  The per-pulse stride wst is
    // odd, spreading a wave's lanes (distinct pulses, nearby bins) over the
    // LDS banks.
    extern __shared__ unsigned long long wsum[];
    const int G = NT == 512 ? 64 : k.fwd_pulses, R = k.region_tiles;
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so the
    // groups of one region are given ids congruent mod 8 and run back to back
    // on one XCD, whose L2 then serves the region's pixel records to all of them.
    const int bid = (int)blockIdx.x, npg = k.fwd_groups;
    const int slot = bid >> 3;
    const int region = (bid & 7) + 8 * (slot / npg);
    const int j0 = (slot % npg) * G;
    if (region >= k.num_regions) return;
    const int tid = (int)threadIdx.x;
    const int q = tid & (G - 1);
    // G = 64: a wave's lanes are the 64 pulses and its pixel stream is
    // wave-uniform, so pixel records and amplitudes come through scalar loads
    const int stream = NT == 512 ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid / G;
    const int nstreams = NT / G;
    const int np = min(G, k.Pv - j0);
    const bool qok = q < np;
    const int j = j0 + (qok ? q : np - 1);
    const int p = k.rows[j];
    const int win = k.win;
    const int t0 = region * R;
    const int t1 = min(k.num_tiles, t0 + R);

    const PolyRegs pr = make_poly_regs<SORD>();
    const double rho = k.rho;  // revolutions per bin
    // region bound (region_l1_kernel) -> fixed-point scale
    const float bound = k.region_l1[region];
    // region-uniform flags, made scalar so that they branch on the scalar unit
    const bool fixed_ok =
        __builtin_amdgcn_readfirstlane((int)(fix_bound_ok(bound) && win > 0)) != 0;
    const bool zero_region = fixed_ok && __builtin_amdgcn_readfirstlane((int)(bound == 0.f)) != 0;
    // amplitudes arrive pre-scaled by `fs` when `scaled`
    const bool scaled = fixed_ok && !zero_region;
    FixScale fs{1.f, 1.f, 0};
    if (scaled) fs = fix_scale(bound);
    auto unscale = [&](const float v) { return scaled ? fix_undo(fs, v) : v; };
    // linear series / hilo, and the nearest rules with series ranges: u is
    // evaluated relative to the window base (for the nearest rules plus 1/2,
    // so that truncation gives round(u) away from ties)
    constexpr bool kNear = INTERP != kInterpLinear && PREC == kPrecSeries;
    constexpr bool kRel =
        (INTERP == kInterpLinear && (PREC == kPrecSeries || PREC == kPrecHilo)) || kNear;
    constexpr bool kCompact = kRel && PLANAR && NT == 512;
    // window base of (region, pulse): fixed by the geometry, computed once per
    // plan (window_base_kernel)
    const double wb = k.wbases[(size_t)j * k.num_regions + region];
    const int wst = k.win_stride;
    for (int e = tid; e < G * wst; e += NT) wsum[e] = 0ull;
    __syncthreads();
    unsigned int *ws = (unsigned int *)(wsum + q * wst);
    float2 *row = k.prof + (size_t)p * k.nw;
    bool oob = false;  // a contribution of this lane fell outside its window
    // The pixel loops, instantiated for regions with windows (kFixed) and
    // without, so that the per-pixel code of each has no region test.
    auto scan = [&](auto fixed_tag) {
        constexpr bool kFixed = decltype(fixed_tag)::value;
            for (int tile = t0; tile < t1; ++tile) {
                PulseData<MONO> d;
                load_pulse<PREC, MONO>(k, tile, j, p, d);
                if constexpr (kNear)
                    d.s.cu = (d.s.cu - wb) + 0.5;
                else if constexpr (kRel)
                    d.s.cu -= wb;
                SeriesHorner<MONO, SORD> hc;
                if constexpr (PREC == kPrecSeries) hc = make_horner<MONO, SORD>(d, rho);
                const int tb = tile * kTilePixels;
                const int tn = min(kTilePixels, k.M - tb);
                // NT == 512: the next pixel's record is requested one iteration ahead.
                // kCompact: each wave walks a contiguous block of the tile's
                // pixels through the 32-byte records (one scalar load per pixel,
                // two records per cache line); otherwise the streams interleave.
                PixRec r_next;
                float2 a_next;
                // kCompact: blocks of whole record groups, so that no group
                // straddles two streams (only the image's last group may end
                // past its last pixel, in the record array's spare records)
                const int per = kCompact ? (tn + nstreams * kScatGroup - 1) / (nstreams * kScatGroup) * kScatGroup
                                         : (tn + nstreams - 1) / nstreams;
                const int ibeg = kCompact ? stream * per : stream;
                const int iend = kCompact ? min(tn, ibeg + per) : tn;
                const int istep = kCompact ? 1 : nstreams;
                if constexpr (kCompact) {
                    // (loaded in pairs below)
                } else if constexpr (NT == 512) {
                    if (stream < tn) {
                        r_next = load_uniform(k.pixrec, (size_t)(tb + stream));
                        a_next = load_uniform((const float2 *)k.amps_perm, (size_t)(tb + stream));
                    }
                }
                // one pixel's contribution to this lane's pulse window (ri:
                // the pixel's record index, for the exact range on kCompact)
                auto contribute = [&](const Pix &px, const float2 a, const int ri) {
                    double u, th;
                    if constexpr (PREC == kPrecSeries)
                        eval_range_horner<MONO, PLANAR, SORD>(d, hc, px, rho, u, th);
                    else
                        eval_range<PREC, MONO, PLANAR, SORD>(k, d, px, u, th, pr);
                    float sn, cs;
                    double fl = 0.0;
                    int ln = 0;  // kNear: the window bin
                    if constexpr (INTERP == kInterpLinear && kRel) {
                        // u is taken from the window base, at least 2 bins below
                        // every pixel of the region: truncation is floor
                        fl = (double)(int)u;
                    } else if constexpr (kNear) {
                        // u - 1/2 from the window base: truncation is the
                        // forward rule round(u) (CPP:289) unless u is within
                        // 1e-7 bin of a tie, where the rule decides on the
                        // exact range (rare, lane-divergent)
                        ln = (int)u;
                        // (the spare records past the image's last pixel add
                        // zero and have no position record)
                        if ((!kCompact || ri < k.M) &&
                            fabs(__builtin_amdgcn_fract(u) - 0.5) > 0.5 - kTieMargin) {
                            double x = px.x, y = px.y, z = px.z;
                            if constexpr (kCompact) {
                                const PixRec r = load_uniform(k.pixrec, (size_t)ri);
                                x = r.x;
                                y = r.y;
                                z = r.z;
                            }
                            ln = (int)(round(ref_dr<MONO>(k.pc[p], x, y, z) * k.by_dx) - wb);
                        }
                    } else if constexpr (INTERP == kInterpLinear) {
                        fl = floor(u);
                    } else {
                        // forward bin rule round(u) for both nearest modes (CPP:289)
                        fl = nearest_bin<kInterpNearest, PREC, MONO>(k, px, p, u, th);
                    }
                    // exact ranges: the phase argument's FP32 rounding is corrected
                    if constexpr (PREC == kPrecFp64)
                        phasor_fract_fine(th, &sn, &cs);
                    else
                        phasor_fract(th, &sn, &cs);
                    // contribution a e^{i theta} (times `scale` when pre-scaled)
                    const float cr = fmaf(cs, a.x, -(sn * a.y));
                    const float ci = fmaf(cs, a.y, sn * a.x);
                    const int l = kNear ? ln : kRel ? (int)fl : (int)(fl - wb);
                    if constexpr (kFixed) {
                        // Every lane adds into its window without a branch: a lane
                        // whose bin falls outside the window adds at the spare
                        // bins at its end (for linear at l = win - 1 into the last
                        // bin and a spare one; the row is poisoned then anyway)
                        // and is flagged (oob, handled after the pixel loops);
                        // lanes without a pulse (q >= np) add into windows never
                        // written out.
                        if constexpr (INTERP == kInterpLinear) {
                            const float fr = kRel ? (float)__builtin_amdgcn_fract(u) : (float)(u - fl);
                            const bool in = (unsigned)l < (unsigned)(win - 1);
                            const int lt = (int)min((unsigned)l, (unsigned)win);
                            const int tr = fix_tz(cr), ti = fix_tz(ci);
                            const int r1 = fix_tz(fr * cr), i1 = fix_tz(fr * ci);
                            add_fix(ws + 2 * lt, tr - r1, ti - i1);
                            add_fix(ws + 2 * lt + 2, r1, i1);
                            oob = oob || !in;
                        } else {
                            const bool in = (unsigned)l < (unsigned)win;
                            add_fix(ws + 2 * (int)min((unsigned)l, (unsigned)win), fix_tz(cr), fix_tz(ci));
                            oob = oob || !in;
                        }
                    } else if (qok) {
                        // a region without windows (non-finite amplitudes, or no
                        // window fits the LDS): float global atomics into the
                        // cleared rows, or on the span path (rows not cleared) a
                        // poisoned row (NaN output row)
                        if (k.fwd_spans) {
                            atomicOr(&k.fpoison[p], 1);
                        } else if constexpr (INTERP == kInterpLinear) {
                            const float fr = kRel ? (float)__builtin_amdgcn_fract(u) : (float)(u - fl);
                            const float gr = 1.0f - fr;
                            const int n0 = bin_index(k, kRel ? floor(u) + wb : fl, 1);
                            const int n1 = n0 + 1 == k.nw ? 0 : n0 + 1;
                            atomicAdd(&row[n0].x, unscale(gr * cr));
                            atomicAdd(&row[n0].y, unscale(gr * ci));
                            atomicAdd(&row[n1].x, unscale(fr * cr));
                            atomicAdd(&row[n1].y, unscale(fr * ci));
                        } else {
                            const int nb = bin_index(k, kNear ? (double)ln + wb : fl, 1);
                            atomicAdd(&row[nb].x, unscale(cr));
                            atomicAdd(&row[nb].y, unscale(ci));
                        }
                    }
                };
                if constexpr (kCompact) {
                    // records in groups of four (two 64-byte scalar loads), the
                    // next group requested while this one is processed: one wait
                    // per four pixels; invalid pixels carry a zero amplitude and
                    // a zero offset (the record array has four spare records at
                    // its end, so a group never reads past it, and a spare
                    // record adds zero).  Two group buffers alternate, so no
                    // registers are copied between groups.
                    constexpr int KB = kScatGroup;
                    ScatRec ca[KB], cb[KB];
                    auto load_group = [&](ScatRec(&c)[KB], const int i) {
                        load_srec2(k.srec, (size_t)(tb + i), c[0], c[1]);
                        load_srec2(k.srec, (size_t)(tb + i + 2), c[2], c[3]);
                    };
                    auto run_group = [&](const ScatRec(&c)[KB], const int i) {
                        Pix px;
                        px.x = px.y = px.z = 0.0;  // unused on this path
                        px.dz = 0.0;
                        px.ok = true;
    #pragma unroll
                        for (int e = 0; e < KB; ++e) {
                            px.dx = c[e].dx;
                            px.dy = c[e].dy;
                            px.d2 = c[e].d2;
                            contribute(px, c[e].a, tb + i + e);
                        }
                    };
                    // The next group is requested after the wait for this one:
                    // a scalar load's only wait is lgkmcnt(0), so the wait the
                    // compiler places before this group's first use would also
                    // wait for a request issued just before it (1.75 -> 1.72 ms
                    // at C2, A/B on one box).
                    if (ibeg < iend) load_group(ca, ibeg);
                    for (int i = ibeg; i < iend; i += 2 * KB) {
                        lgkm_wait0();
                        if (i + KB < iend) load_group(cb, i + KB);
                        run_group(ca, i);
                        if (i + KB >= iend) break;
                        lgkm_wait0();
                        if (i + 2 * KB < iend) load_group(ca, i + 2 * KB);
                        run_group(cb, i + KB);
                    }
                } else {
                    for (int i = ibeg; i < iend; i += istep) {
                        PixRec r;
                        float2 a;
                        if constexpr (NT == 512) {
                            // wave-uniform index: scalar loads through the constant address space
                            r = r_next;
                            a = a_next;
                            if (i + nstreams < tn) {
                                r_next = load_uniform(k.pixrec, (size_t)(tb + i + nstreams));
                                a_next = load_uniform((const float2 *)k.amps_perm,
                                                      (size_t)(tb + i + nstreams));
                            }
                            // ok is exactly 0.0 or 1.0: an integer test keeps it on the scalar unit
                            if (__double_as_longlong(r.ok) == 0) continue;
                        } else {
                            r = k.pixrec[tb + i];
                            if (r.ok == 0.0) continue;
                            a = k.amps_perm[tb + i];
                        }
                        Pix px;
                        px.x = r.x;
                        px.y = r.y;
                        px.z = r.z;
                        px.dx = r.dx;
                        px.dy = r.dy;
                        px.dz = r.dz;
                        px.d2 = r.d2;
                        px.ok = true;
                        contribute(px, a, 0);
                    }
                }
            }
    };
    if (fixed_ok && !zero_region)
        scan(std::true_type{});
    else if (!fixed_ok)
        scan(std::false_type{});  // zero regions are skipped (exact)
    // A bin outside the window is ruled out by the window bounds; should one
    // occur, the row is made NaN (poisoned on the span path, a NaN added to the
    // cleared row otherwise) rather than left silently wrong.
    if (oob && qok) {
        if (k.fwd_spans)
            atomicOr(&k.fpoison[p], 1);
        else
            atomicAdd(&row[0].x, __builtin_nanf(""));
    }
    __syncthreads();
    // Partials and liveness: wave w writes the integer windows of pulses w,
    // w + NT/64, ... (coalesced rows) and whether any of their sums is
    // non-zero (one ballot; the merge skips windows that are all zero).
    const int NR = k.num_regions;
    const int lane = tid & 63, wave = tid >> 6;
    for (int qq = wave; qq < np; qq += NT / 64) {
        bool nz = false;
        int2 *dst = k.partial + ((size_t)(j0 + qq) * NR + region) * win;
        for (int l = lane; l < win; l += 64) {
            const unsigned long long w = wsum[qq * wst + l];
            const int2 v = scaled ? make_int2((int)(unsigned)(w & 0xffffffffull),
                                              (int)(unsigned)(w >> 32))
                                  : make_int2(0, 0);
            dst[l] = v;
            nz = nz || (v.x | v.y) != 0;
        }
        const bool live = __ballot(nz) != 0ull;
        if (lane == 0) k.wlive[(size_t)(j0 + qq) * NR + region] = live ? 1.f : 0.f;
    }
}

// Once per plan: the window base of every (valid pulse j, forward region),
// wbases[j][region] = window_base(region, rows[j]).
template <bool MONO>
__global__ __launch_bounds__(256) void window_base_kernel(const KernelParams k) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t NR = (size_t)k.num_regions;
    if (i >= (size_t)k.Pv * NR) return;
    const int j = (int)(i / NR), region = (int)(i % NR);
    k.wbases[i] = window_base<MONO>(k, region, k.rows[j]);
}

// One workgroup per valid pulse: merge the pulse's window partials into its
// profile row, with no atomics.  The span [lo, hi) of the row's non-zero
// windows is cut into 64-bin blocks; wave w takes blocks w, w + 8, ...: it
// lists the windows that meet the block, in window order (a ballot over 64
// windows at a time, each lane testing one window's block range), and lane l
// sums bin 64 b + l over them, kRedBatch loads in flight (a window's part of
// the block is one coalesced load).  Every partial is read once.
//
// A window sum s (an integer, |s| < 2^31, at the window's fixed-point scale
// m 2^e) enters as s * ((1 / m) 2^-e); the factor is exact in FP64, so each
// product and each sum rounds at 2^-53 relative, in a fixed order (the result
// does not depend on scheduling), and a power-of-two scaling of the image
// scales every term and sum exactly.  Each bin is rounded once to FP32 and
// stored into the row (span path) or added to it (rows cleared before the
// scatter).  The span is merged in passes of at most min(N_w, 64 kRedBlocks)
// bins, so that no two bins of one pass share a row bin when it is longer
// than the row.  (The first exact merge summed 64-bit integers at a row-wide
// unit: about 35 VALU instructions per window and block against 11 here.)
constexpr int kRedThreads = 512;
constexpr int kRedWaves = kRedThreads / 64;
constexpr int kRedBatch = 8;                   // loads in flight per lane
constexpr int kRedList = 128;                  // per-wave list capacity
constexpr int kRedBlocks = 4 * kRedThreads;    // 64-bin blocks per pass
__global__ __launch_bounds__(kRedThreads) void fp_reduce_kernel(const KernelParams k) {
    // [NR] window factors (1 / m) 2^-e (double), then [NR] window offsets from
    // lo in bins (int, -1 for a window that is all zero)
    extern __shared__ double sc[];
    // per-wave lists: {window offset r, index of its bin 0 less r}, factor
    __shared__ int2 lists[kRedWaves][kRedList];
    __shared__ double listf[kRedWaves][kRedList];
    __shared__ double red_lo[kRedWaves], red_hi[kRedWaves];
    const int j = blockIdx.x;
    const int p = k.rows[j];
    const int NR = k.num_regions, win = k.win;
    int *sr = (int *)(sc + NR);
    const double *b = k.wbases + (size_t)j * NR;
    const float *wm = k.wlive + (size_t)j * NR;
    const int2 *part = k.partial + (size_t)j * NR * win;
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // a row the scatter poisoned: no span, NaN output row (flag cleared here)
    __shared__ int s_poison;
    if (tid == 0) s_poison = k.fwd_spans ? k.fpoison[p] : 0;
    __syncthreads();
    if (s_poison) {
        if (tid == 0) {
            k.fspan[p] = make_int2(0, -1);
            k.fpoison[p] = 0;
        }
        return;
    }
    double lo = INFINITY, hi = -INFINITY;
    for (int t = tid; t < NR; t += kRedThreads) {
        const double bt = b[t];
        if (wm[t] > 0.f) {  // a window with a non-zero sum
            lo = fmin(lo, bt);
            hi = fmax(hi, bt);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    if (lane == 0) {
        red_lo[wave] = lo;
        red_hi[wave] = hi;
    }
    __syncthreads();
    lo = red_lo[0];
    hi = red_hi[0];
    for (int w = 1; w < kRedWaves; ++w) {
        lo = fmin(lo, red_lo[w]);
        hi = fmax(hi, red_hi[w]);
    }
    hi += (double)win;
    // With fwd_spans the covered bins [lo, hi) are written (stored) once each,
    // and their span is recorded for the spectral stage; the rest of the row
    // is not touched (a span longer than the row wraps onto itself: the row is
    // zeroed first and the passes add into it).  Otherwise the row was cleared
    // before the scatter, which may have added global atomics, and the passes
    // add into it.
    const bool store = k.fwd_spans != 0;
    const bool empty = !(hi > lo);  // every window zero
    const bool full = !empty && hi - lo > (double)k.nw;
    if (tid == 0)
        k.fspan[p] = empty ? make_int2(0, 0)
                           : full ? make_int2(0, k.nw) : make_int2(bin_index(k, lo, 1), (int)(hi - lo));
    if (empty) return;
    // window bases are integers (bins), so the offsets are exact
    for (int t = tid; t < NR; t += kRedThreads) {
        const bool live = wm[t] > 0.f;
        double c = 0.0;
        if (live) {
            const FixScale fs = fix_scale(k.region_l1[t]);
            c = ldexp((double)fs.inv_m, -fs.e);
        }
        sc[t] = c;
        sr[t] = live ? (int)(b[t] - lo) : -1;
    }
    float2 *row = k.prof + (size_t)p * k.nw;
    if (full && store) {
        for (int e = tid; e < k.nw; e += kRedThreads) row[e] = make_float2(0.f, 0.f);
        __threadfence_block();
    }
    __syncthreads();
    const int S = (int)(hi - lo);  // bins of the span
    const int W = min(k.nw, 64 * kRedBlocks);
    int2 *list = lists[wave];
    double *listf_w = listf[wave];
    const unsigned long long below = (1ull << lane) - 1ull;  // lanes under this one
    for (int c0 = 0; c0 < S; c0 += W) {
        const int len = min(W, S - c0);
        const int nb = (len + 63) / 64;
        for (int blk = wave; blk < nb; blk += kRedWaves) {
            const int ab = (c0 >> 6) + blk;  // the block's index from lo
            const int rel = c0 + 64 * blk + lane;  // the lane's bin from lo
            double ar = 0.0, ai = 0.0;
            int n = 0;  // listed windows (wave-uniform)
            for (int t0 = 0; t0 < NR; t0 += 64) {
                const int t = t0 + lane;
                bool meet = false;
                int r = -1;
                if (t < NR) {
                    r = sr[t];
                    meet = r >= 0 && (r >> 6) <= ab && ((r + win - 1) >> 6) >= ab;
                }
                const unsigned long long mask = __ballot(meet);
                if (meet) {
                    const int at = n + __popcll(mask & below);
                    list[at] = make_int2(r, t * win - r);
                    listf_w[at] = sc[t];
                }
                n += __popcll(mask);
                if (n > kRedList - 64 || (t0 + 64 >= NR && n > 0)) {
                    for (int i = 0; i < n; i += kRedBatch) {
                        int2 v[kRedBatch];
                        double f[kRedBatch];
#pragma unroll
                        for (int q = 0; q < kRedBatch; ++q) {
                            v[q] = make_int2(0, 0);
                            f[q] = 0.0;
                            if (i + q < n) {
                                const int2 e = list[i + q];
                                const int o = rel - e.x;
                                f[q] = listf_w[i + q];
                                if (o >= 0 && o < win) v[q] = part[e.y + rel];
                            }
                        }
#pragma unroll
                        for (int q = 0; q < kRedBatch; ++q) {
                            ar = fma((double)v[q].x, f[q], ar);
                            ai = fma((double)v[q].y, f[q], ai);
                        }
                    }
                    n = 0;
                }
            }
            const int nrel = 64 * blk + lane;  // the lane's bin in the pass
            if (nrel < len) {
                const int nn = bin_index(k, lo + (double)(c0 + nrel), 1);
                float2 r = (full || !store) ? row[nn] : make_float2(0.f, 0.f);
                r.x += (float)ar;
                r.y += (float)ai;
                row[nn] = r;
            }
        }
        __threadfence_block();
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_window_pad(const float2 *Y, const float2 *Wconj, const float2 *wfac, int N,
                             int nw, int ld, int P, int centred, float2 *K, hipStream_t s) {
    dim3 grid((nw + 255) / 256, P);
    window_pad_kernel<<<grid, 256, 0, s>>>(Y, Wconj, wfac, N, nw, ld, centred, K);
    return hipGetLastError();
}

hipError_t launch_wrap_pad(float2 *prof, int nw, int ld, int P, hipStream_t s) {
    if (ld <= nw || P <= 0) return hipSuccess;
    dim3 grid((ld - nw + 255) / 256, P);
    wrap_pad_kernel<<<grid, 256, 0, s>>>(prof, nw, ld);
    return hipGetLastError();
}

hipError_t launch_truncate(const float2 *G, const float2 *W, const float2 *wfac, int N, int nw,
                           int P, int centred, float2 *Y, hipStream_t s) {
    dim3 grid((N + 255) / 256, P);
    truncate_kernel<<<grid, 256, 0, s>>>(G, W, wfac, N, nw, centred, Y);
    return hipGetLastError();
}

hipError_t launch_truncate64(const double2 *G, const float2 *W, const float2 *wfac, int N, int nw,
                             int P, int centred, float2 *Y, hipStream_t s) {
    dim3 grid((N + 255) / 256, P);
    truncate64_kernel<<<grid, 256, 0, s>>>(G, W, wfac, N, nw, centred, Y);
    return hipGetLastError();
}

hipError_t launch_widen(const float2 *src, double2 *dst, size_t n, hipStream_t s) {
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    if (blocks == 0) return hipSuccess;
    widen_kernel<<<(unsigned)blocks, 256, 0, s>>>(src, dst, n);
    return hipGetLastError();
}

hipError_t launch_pair_rows(const float2 *prof, const int2 *spans, const int *rows, int Pv,
                            int nw, int ld, double rho, bool a_only, void *pairs, hipStream_t s) {
    if (Pv <= 0) return hipSuccess;
    if (a_only)
        pair_rows_kernel<true><<<Pv, 256, 0, s>>>(prof, spans, rows, nw, ld, rho, pairs);
    else
        pair_rows_kernel<false><<<Pv, 256, 0, s>>>(prof, spans, rows, nw, ld, rho, pairs);
    return hipGetLastError();
}

hipError_t launch_fill_zero(void *buf, size_t bytes, hipStream_t s) {
    const size_t n4 = bytes / 16;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    fill_zero_kernel<<<(unsigned)blocks, 256, 0, s>>>((float4 *)buf, bytes);
    return hipGetLastError();
}

hipError_t launch_gather_perm(const float2 *src, const int *perm, int M, float2 *dst,
                              hipStream_t s) {
    gather_perm_kernel<<<(M + 255) / 256, 256, 0, s>>>(src, perm, M, dst);
    return hipGetLastError();
}

// out[i] = cu_i moved by a multiple of nw into [shift_lo, shift_lo + nw):
// the backprojection's row offset of the tile centre for (tile, pulse) i,
// cu_i the `cu` field of that record (index cu_at of `stride` doubles).
__global__ __launch_bounds__(256) void bp_shift_kernel(const double *__restrict__ series,
                                                       int stride, int cu_at, size_t n,
                                                       double shift_lo, double inv_nw, double nwd,
                                                       double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double cu = series[i * stride + cu_at];
    out[i] = fma(-floor((cu - shift_lo) * inv_nw), nwd, cu);
}

// Bins of each adjoint profile row that the backprojection can read, one
// workgroup per valid pulse: {first bin in [0, nw), count}, the whole row when
// the count reaches nw.  Per leg |T - X| moves by at most |X - X0| <= r from
// the tile centre X0 (triangle inequality), so every pixel of the tile has u
// within 2 r by_dx of u(X0); the row is read at floor(u) to floor(u) + 2
// (linear neighbours, nearest rules), and 8 bins on each side cover rounding
// in every precision mode.
__global__ __launch_bounds__(256) void adjoint_span_kernel(const PulseConst *__restrict__ pc,
                                                           const int *__restrict__ rows,
                                                           const double4 *__restrict__ tiles,
                                                           int nt, int mono, double by_dx, int nw,
                                                           int2 *__restrict__ span) {
    __shared__ double s_lo[256], s_hi[256];
    const int p = rows[blockIdx.x];
    const int tid = (int)threadIdx.x;
    const PulseConst c = pc[p];
    double lo = INFINITY, hi = -INFINITY;
    for (int t = tid; t < nt; t += 256) {
        const double4 x = tiles[t];
        const double ax = c.tx[0] - x.x, ay = c.tx[1] - x.y, az = c.tx[2] - x.z;
        const double d1 = sqrt(fma(ax, ax, fma(ay, ay, az * az)));
        double d2 = d1;
        if (!mono) {
            const double bx = c.rx[0] - x.x, by = c.rx[1] - x.y, bz = c.rx[2] - x.z;
            d2 = sqrt(fma(bx, bx, fma(by, by, bz * bz)));
        }
        const double u = (d1 + d2 - c.r_srp) * by_dx, du = 2.0 * x.w * by_dx;
        lo = fmin(lo, u - du);
        hi = fmax(hi, u + du);
    }
    s_lo[tid] = lo;
    s_hi[tid] = hi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s_lo[tid] = fmin(s_lo[tid], s_lo[tid + h]);
            s_hi[tid] = fmax(s_hi[tid], s_hi[tid + h]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double first = floor(s_lo[0]) - 8.0;
        const double count = floor(s_hi[0]) - floor(s_lo[0]) + 19.0;
        if (!(count < (double)nw) || !isfinite(first)) {
            span[p] = make_int2(0, nw);
        } else {
            const double w = first - floor(first / (double)nw) * (double)nw;
            span[p] = make_int2(min((int)w, nw - 1), (int)count);
        }
    }
}

hipError_t launch_adjoint_spans(const PulseConst *pc, const int *rows, int Pv, const double4 *tiles,
                                int nt, bool mono, double by_dx, int nw, int2 *span,
                                hipStream_t s) {
    if (Pv <= 0) return hipSuccess;
    adjoint_span_kernel<<<Pv, 256, 0, s>>>(pc, rows, tiles, nt, mono ? 1 : 0, by_dx, nw, span);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void series_gather_kernel(const SeriesMono *__restrict__ series,
                                                            size_t n, double shift_lo,
                                                            double inv_nw, double nwd, double rho,
                                                            double off,
                                                            SeriesGather *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const SeriesMono s = series[i];
    SeriesGather g;
    const double inv_c = 1.0 / s.c, c2 = s.c * s.c;
    g.bx = s.bx * inv_c;
    g.by = s.by * inv_c;
    g.k1 = s.ku * 0.5 * s.c;
    g.k2 = s.ku * -0.125 * c2;
    g.k3 = s.ku * 0.0625 * (c2 * s.c);
    g.pad = 0.f;
    g.cu = fma(-floor((s.cu - shift_lo) * inv_nw), nwd, s.cu);  // as bp_shift_kernel
    // the nearest rules' offset (floor(u + off) is their bin away from ties);
    // phi then refers to u + off
    g.cu += off;
    g.phi = fma(-rho, g.cu, s.ct);
    // the pair path's phase constant: phi reduced to [-1/2, 1/2], less rho / 2
    // (the pair rows carry e^{-2 pi i rho (n + 1/2)}), so that the pixel phase
    // rho f + phi_pair stays within rho / 2 + 1/2 of 0
    g.phi_pair = (float)((g.phi - rint(g.phi)) - 0.5 * rho);
    out[i] = g;
}

hipError_t launch_series_gather(const void *series, size_t n, double shift_lo, double inv_nw,
                                double nwd, double rho, double off, SeriesGather *out,
                                hipStream_t s) {
    series_gather_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
        (const SeriesMono *)series, n, shift_lo, inv_nw, nwd, rho, off, out);
    return hipGetLastError();
}

hipError_t launch_bp_shift(const void *series, bool mono, size_t n, double shift_lo, double inv_nw,
                           double nwd, double *out, hipStream_t s) {
    const int stride = mono ? (int)(sizeof(SeriesMono) / 8) : (int)(sizeof(SeriesBi) / 8);
    const int cu_at = mono ? (int)(offsetof(SeriesMono, cu) / 8) : (int)(offsetof(SeriesBi, cu) / 8);
    bp_shift_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>((const double *)series, stride,
                                                               cu_at, n, shift_lo, inv_nw, nwd, out);
    return hipGetLastError();
}

hipError_t launch_series_consts(const double4 *tile_c, int num_tiles, const PulseConst *pc,
                                const int *rows, int Pv, double by_dx, double kc_rev, bool mono,
                                void *out, hipStream_t s) {
    const size_t n = (size_t)num_tiles * Pv;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (mono)
        series_consts_kernel<true><<<blocks, 256, 0, s>>>(tile_c, num_tiles, pc, rows, Pv, by_dx,
                                                           kc_rev, out);
    else
        series_consts_kernel<false><<<blocks, 256, 0, s>>>(tile_c, num_tiles, pc, rows, Pv, by_dx,
                                                            kc_rev, out);
    return hipGetLastError();
}

template <int I, bool MO>
static void bp_dispatch_prec(const KernelParams &k, hipStream_t s) {
    // (tile, split) pairs in the kernel's XCD-aware order
    const dim3 grid((unsigned)(8 * (((long long)k.num_tiles * k.bp_splits + 7) / 8)));
    // planar specialisations for the linear series / hilo paths only
    const bool pl = I == kInterpLinear && k.planar;
    switch (k.precision) {
        case kPrecSeries:
            if constexpr (I != kInterpLinear && MO) {
                // nearest rules: the pair path when the plan built pair rows
                // (monostatic planar tiles, third-order series)
                if (k.bp_pairs && k.planar && k.series_order == 3) {
                    bp_gather_kernel<I, MO, kPrecSeries, true, 3, true><<<grid, kThreads, 0, s>>>(k);
                    break;
                }
            }
            if (k.series_order == 3) {
                if (pl) {
                    if constexpr (I == kInterpLinear && MO) {
                        if (k.bp_pairs) {
                            bp_gather_kernel<I, MO, kPrecSeries, true, 3, true>
                                <<<grid, kThreads, 0, s>>>(k);
                            break;
                        }
                    }
                    bp_gather_kernel<I, MO, kPrecSeries, true, 3><<<grid, kThreads, 0, s>>>(k);
                } else {
                    bp_gather_kernel<I, MO, kPrecSeries, false, 3><<<grid, kThreads, 0, s>>>(k);
                }
            } else {
                if (pl) bp_gather_kernel<I, MO, kPrecSeries, true><<<grid, kThreads, 0, s>>>(k);
                else bp_gather_kernel<I, MO, kPrecSeries, false><<<grid, kThreads, 0, s>>>(k);
            }
            break;
        case kPrecHilo:
            if (pl) bp_gather_kernel<I, MO, kPrecHilo, true><<<grid, kThreads, 0, s>>>(k);
            else bp_gather_kernel<I, MO, kPrecHilo, false><<<grid, kThreads, 0, s>>>(k);
            break;
        case kPrecFp32: bp_gather_kernel<I, MO, kPrecFp32, false><<<grid, kThreads, 0, s>>>(k); break;
        default: bp_gather_kernel<I, MO, kPrecFp64, false><<<grid, kThreads, 0, s>>>(k); break;
    }
}

hipError_t launch_bp_gather(const KernelParams &k, hipStream_t s) {
    if (k.mono) {
        if (k.interp == kInterpLinear) bp_dispatch_prec<kInterpLinear, true>(k, s);
        else if (k.interp == kInterpNearest) bp_dispatch_prec<kInterpNearest, true>(k, s);
        else bp_dispatch_prec<kInterpNearestRef, true>(k, s);
    } else {
        if (k.interp == kInterpLinear) bp_dispatch_prec<kInterpLinear, false>(k, s);
        else if (k.interp == kInterpNearest) bp_dispatch_prec<kInterpNearest, false>(k, s);
        else bp_dispatch_prec<kInterpNearestRef, false>(k, s);
    }
    return hipGetLastError();
}

hipError_t launch_bp_reduce(const KernelParams &k, hipStream_t s) {
    if (k.bp_splits > 1)
        bp_reduce_kernel<<<(k.M + 255) / 256, 256, 0, s>>>(k.bp_partial, k.M, k.bp_splits, k.perm,
                                                           k.out);
    return hipGetLastError();
}

template <int I, bool MO, int NT>
static void fp_dispatch_prec(const KernelParams &k, hipStream_t s) {
    const int nr8 = (k.num_regions + 7) / 8 * 8;
    const dim3 grid((unsigned)(nr8 * k.fwd_groups));
    const size_t lds = sizeof(unsigned long long) * (size_t)k.fwd_pulses * k.win_stride;
    const bool pl = k.planar;
    switch (k.precision) {
        case kPrecSeries:
            if (k.series_order == 3) {
                if (pl) fp_scatter_kernel<I, MO, kPrecSeries, true, NT, 3><<<grid, NT, lds, s>>>(k);
                else fp_scatter_kernel<I, MO, kPrecSeries, false, NT, 3><<<grid, NT, lds, s>>>(k);
            } else {
                if (pl) fp_scatter_kernel<I, MO, kPrecSeries, true, NT><<<grid, NT, lds, s>>>(k);
                else fp_scatter_kernel<I, MO, kPrecSeries, false, NT><<<grid, NT, lds, s>>>(k);
            }
            break;
        case kPrecHilo:
            if (pl) fp_scatter_kernel<I, MO, kPrecHilo, true, NT><<<grid, NT, lds, s>>>(k);
            else fp_scatter_kernel<I, MO, kPrecHilo, false, NT><<<grid, NT, lds, s>>>(k);
            break;
        case kPrecFp32: fp_scatter_kernel<I, MO, kPrecFp32, false, NT><<<grid, NT, lds, s>>>(k); break;
        default: fp_scatter_kernel<I, MO, kPrecFp64, false, NT><<<grid, NT, lds, s>>>(k); break;
    }
}

// 64 pulses per workgroup run with 512 threads (8 pixel streams), fewer with 256.
hipError_t launch_region_l1(const KernelParams &k, hipStream_t s) {
    region_l1_kernel<<<k.num_regions, kThreads, 0, s>>>(const_cast<float2 *>(k.amps_perm), k.M,
                                                        k.region_tiles * kTilePixels, k.win > 0,
                                                        k.region_l1, k.region_cmax, k.pixrec,
                                                        const_cast<ScatRec *>(k.srec));
    return hipGetLastError();
}

hipError_t launch_window_bases(const KernelParams &k, hipStream_t s) {
    const size_t n = (size_t)k.Pv * (size_t)k.num_regions;
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (k.mono) window_base_kernel<true><<<blocks, 256, 0, s>>>(k);
    else window_base_kernel<false><<<blocks, 256, 0, s>>>(k);
    return hipGetLastError();
}

hipError_t launch_region_count(const KernelParams &k, int *cmax, hipStream_t s) {
    if (k.win <= 0 || k.num_regions <= 0 || k.Pv <= 0) return hipSuccess;
    // pulses per workgroup: histograms of 64, 32 or 16 pulses within 48 KiB
    int PB = 64;
    while (PB > 16 && (size_t)PB * k.win_stride * sizeof(int) > 48 * 1024) PB /= 2;
    if ((size_t)PB * k.win_stride * sizeof(int) > 48 * 1024) return hipErrorInvalidValue;
    const dim3 grid((unsigned)k.num_regions, (unsigned)((k.Pv + PB - 1) / PB));
    const size_t lds = (size_t)PB * k.win_stride * sizeof(int);
    if (k.mono) region_count_kernel<true><<<grid, 256, lds, s>>>(k, PB, cmax);
    else region_count_kernel<false><<<grid, 256, lds, s>>>(k, PB, cmax);
    return hipGetLastError();
}

template <int I, bool MO>
static void fp_dispatch_threads(const KernelParams &k, hipStream_t s) {
    if (k.fwd_pulses == 64) fp_dispatch_prec<I, MO, 512>(k, s);
    else fp_dispatch_prec<I, MO, kThreads>(k, s);
}

hipError_t launch_fp_bin(const KernelParams &k, hipStream_t s) {
    // nearest_ref and nearest share the forward bin rule round(u) (CPP:289)
    if (k.mono) {
        if (k.interp == kInterpLinear) fp_dispatch_threads<kInterpLinear, true>(k, s);
        else fp_dispatch_threads<kInterpNearest, true>(k, s);
    } else {
        if (k.interp == kInterpLinear) fp_dispatch_threads<kInterpLinear, false>(k, s);
        else fp_dispatch_threads<kInterpNearest, false>(k, s);
    }
    return hipGetLastError();
}

size_t fp_reduce_lds(int num_regions, int win) {
    (void)win;
    const size_t nr = (size_t)std::max(num_regions, 0);
    return nr * (sizeof(double) + sizeof(int));
}

hipError_t launch_fp_reduce(const KernelParams &k, hipStream_t s) {
    const size_t lds = fp_reduce_lds(k.num_regions, k.win);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)fp_reduce_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    fp_reduce_kernel<<<k.Pv, kRedThreads, lds, s>>>(k);
    return hipGetLastError();
}

}  // namespace sarhip
