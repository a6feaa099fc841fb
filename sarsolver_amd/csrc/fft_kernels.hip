// fft_kernels.hip -- fused spectral stages of the operators (gfx950).
//
// The operators move between an N-sample band (the phase-history row) and an
// N_w = U N sample range profile (CPP:292, CPP:331-344).  Zero-padding the
// band to N_w and running a length-N_w FFT reads and writes the padded row
// several times.  With U an integer the transform splits into U polyphase
// components, each an N-point DFT:
//   adjoint  h[U m + r] = sum_k (B[k] e^{+2 pi i s_k r / N_w}) e^{+2 pi i k m / N}
//   forward  X[k]       = sum_r e^{-2 pi i s_k r / N_w} sum_m h[U m + r] e^{-2 pi i k m / N}
// (k the FFT-order band index, s_k its signed frequency).  One workgroup
// takes one pulse row; G phases r are transformed at a time, one 256-thread
// group each, with a Stockham FFT in LDS (radix-16 passes, a radix-2/4/8 pass
// when N is not a power of 16) and coalesced G-phase loads / stores of the
// profile row.  Band weights (conj(W) w_p on the way in, W w_p on the way
// out), the centred / FFT order and the adjoint row tail are fused.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <type_traits>

#include "sar_device.h"
#include "sar_launch.h"

namespace sarhip {

namespace {

__device__ __forceinline__ float2 cmulf2(const float2 a, const float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 conjf2(const float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 addf2(const float2 a, const float2 b) {
    return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 subf2(const float2 a, const float2 b) {
    return make_float2(a.x - b.x, a.y - b.y);
}
// The transform core below is written once for FP32 (float2) and FP64
// (double2) complex values through these overloads.
__device__ __forceinline__ float2 cmul(const float2 a, const float2 b) { return cmulf2(a, b); }
__device__ __forceinline__ float2 cconj(const float2 a) { return conjf2(a); }
__device__ __forceinline__ float2 cadd(const float2 a, const float2 b) { return addf2(a, b); }
__device__ __forceinline__ float2 csub(const float2 a, const float2 b) { return subf2(a, b); }
__device__ __forceinline__ double2 cmul(const double2 a, const double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(const double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(const double2 a, const double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(const double2 a, const double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
template <typename C>
__device__ __forceinline__ C cmake(const double x, const double y) {
    C c;
    c.x = x;
    c.y = y;
    return c;
}
// i * a (INV) or -i * a
template <bool INV, typename C>
__device__ __forceinline__ C rot(const C a) {
    return INV ? cmake<C>(-a.y, a.x) : cmake<C>(a.y, -a.x);
}

// LDS index with one pad slot per 16 (spreads the radix-16 writes over banks).
__device__ __forceinline__ int phys(const int i) { return i + (i >> 4); }

// In-register DFTs, out[q] = sum_r v[r] e^{sign 2 pi i r q / R}, sign = +1 for INV.
template <bool INV, typename C>
__device__ __forceinline__ void dft2(C &a, C &b) {
    const C t = a;
    a = cadd(t, b);
    b = csub(t, b);
}
template <bool INV, typename C>
__device__ __forceinline__ void dft4(C &a, C &b, C &c, C &d) {
    const C s0 = cadd(a, c), d0 = csub(a, c);
    const C s1 = cadd(b, d), d1 = rot<INV>(csub(b, d));
    a = cadd(s0, s1);
    c = csub(s0, s1);
    b = cadd(d0, d1);
    d = csub(d0, d1);
}

// e^{sign 2 pi i j / 16}, j = 0..9 (the products b c of the 4 x 4 split).
template <typename C>
__device__ __forceinline__ C w16(const int j, const bool inv) {
    constexpr double C1 = 0.92387953251128674, S1 = 0.38268343236508978, H = 0.70710678118654752;
    C w;
    switch (j) {
        case 0: w = cmake<C>(1.0, 0.0); break;
        case 1: w = cmake<C>(C1, S1); break;
        case 2: w = cmake<C>(H, H); break;
        case 3: w = cmake<C>(S1, C1); break;
        case 4: w = cmake<C>(0.0, 1.0); break;
        case 6: w = cmake<C>(-H, H); break;
        default: w = cmake<C>(-C1, -S1); break;  // j = 9
    }
    return inv ? w : cconj(w);
}

template <int R, bool INV, typename C>
__device__ __forceinline__ void dft(C (&v)[R]) {
    if constexpr (R == 2) {
        dft2<INV>(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // r = 2a + b, q = c + 4d: DFT4 over a, twiddle e^{sign 2 pi i b c / 8}, DFT2 over b
        dft4<INV>(v[0], v[2], v[4], v[6]);
        dft4<INV>(v[1], v[3], v[5], v[7]);
        constexpr double H = 0.70710678118654752;
        v[3] = cmul(v[3], INV ? cmake<C>(H, H) : cmake<C>(H, -H));
        v[5] = rot<INV>(v[5]);
        v[7] = cmul(v[7], INV ? cmake<C>(-H, H) : cmake<C>(-H, -H));
        C o[8];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            C a = v[2 * c], b = v[2 * c + 1];
            dft2<INV>(a, b);
            o[c] = a;
            o[c + 4] = b;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = o[i];
    } else {
        static_assert(R == 16, "radix 2, 4, 8 or 16");
        // r = 4a + b, q = c + 4d: DFT4 over a (per b), twiddle w16^{b c}, DFT4 over b (per c)
#pragma unroll
        for (int b = 0; b < 4; ++b) dft4<INV>(v[b], v[b + 4], v[b + 8], v[b + 12]);
        // after the first DFT4, v[b + 4 c] holds A_b[c]
#pragma unroll
        for (int b = 1; b < 4; ++b)
#pragma unroll
            for (int c = 1; c < 4; ++c) v[b + 4 * c] = cmul(v[b + 4 * c], w16<C>(b * c, INV));
        C o[16];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            C a0 = v[4 * c], a1 = v[4 * c + 1], a2 = v[4 * c + 2], a3 = v[4 * c + 3];
            dft4<INV>(a0, a1, a2, a3);
            o[c] = a0;
            o[c + 4] = a1;
            o[c + 8] = a2;
            o[c + 12] = a3;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = o[i];
    }
}

// Twiddle tables in LDS, two levels each: e^{-2 pi i t / N} =
// nlo[t & 15] nhi[t >> 4] (t < N) and e^{-2 pi i t / N_w} = wlo[t & 255]
// whi[t >> 8] (t < N_w); every entry is rounded once from FP64, so an FP32
// twiddle carries about 1.5 ulp, and one that is a multiple of 16 (the second
// radix-16 pass at N >= 4096) is one entry of nhi (0.5 ulp).
template <typename C>
struct Twiddles {
    const C *nlo, *nhi, *wlo, *whi;
    __device__ __forceinline__ C n(const int t) const { return cmul(nhi[t >> 4], nlo[t & 15]); }
    __device__ __forceinline__ C w(const int t) const { return cmul(whi[t >> 8], wlo[t & 255]); }
};
// Host-side layout of the table array: [16 | N/16 | 256 | N_w/256] entries.
template <int N, typename C>
__device__ __forceinline__ Twiddles<C> load_twiddles(C *dst, const C *__restrict__ src, int nw) {
    const int total = 16 + N / 16 + 256 + nw / 256;
    for (int i = threadIdx.x; i < total; i += blockDim.x) dst[i] = src[i];
    return Twiddles<C>{dst, dst + 16, dst + 16 + N / 16, dst + 16 + N / 16 + 256};
}

// One Stockham pass of radix R with span Ns, in place on the group's LDS
// buffer sh, by T threads (gt < T; all threads of the workgroup reach both
// barriers).  FP32 takes every twiddle from the tables; FP64 takes the pass
// twiddle w from them and forms w^2 .. w^{R-1} by products (a few FP64 ulps),
// which keeps the registers of fifteen table products free.
template <int N, int R, int Ns, bool INV, int T, typename C>
__device__ __forceinline__ void stockham_pass(C *sh, const Twiddles<C> &tw, int gt) {
    constexpr int NB = N / R;
    constexpr int PER = (NB + T - 1) / T;
    C v[PER][R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = gt + T * b;
        if (j < NB)
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = sh[phys(j + r * NB)];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = gt + T * b;
        if (j < NB) {
            const int jm = j % Ns;
            if constexpr (Ns > 1) {
                // pass twiddle stride: a multiple of 16 reads one table entry
                constexpr int ST = N / (Ns * R);
                if constexpr (std::is_same<C, double2>::value) {
                    const C w1 = ST % 16 == 0 ? tw.nhi[(jm * ST) >> 4] : tw.n(jm * ST);
                    C w = w1;
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        v[b][r] = cmul(v[b][r], INV ? cconj(w) : w);
                        if (r + 1 < R) w = cmul(w, w1);
                    }
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        const C t = ST % 16 == 0 ? tw.nhi[(jm * r * ST) >> 4] : tw.n(jm * r * ST);
                        v[b][r] = cmul(v[b][r], INV ? cconj(t) : t);
                    }
                }
            }
            dft<R, INV>(v[b]);
            const int idx = (j / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) sh[phys(idx + r * Ns)] = v[b][r];
        }
    }
    __syncthreads();
}

// All passes of an N-point transform on sh (natural order in and out) by T
// threads: radix 16 while a factor 16 is left, then the remaining 2, 4 or 8.
template <int N, int Ns, bool INV, int T, typename C>
__device__ __forceinline__ void fft_passes(C *sh, const Twiddles<C> &tw, int gt) {
    if constexpr (Ns < N) {
        constexpr int rest = N / Ns;
        constexpr int R = rest >= 16 ? 16 : rest;
        stockham_pass<N, R, Ns, INV, T>(sh, tw, gt);
        fft_passes<N, Ns * R, INV, T>(sh, tw, gt);
    }
}
template <int N, bool INV, int T = 256, typename C>
__device__ __forceinline__ void fft_lds(C *sh, const Twiddles<C> &tw, int gt) {
    static_assert(N >= 64 && N <= 8192 && (N & (N - 1)) == 0, "N: power of two in [64, 8192]");
    fft_passes<N, 1, INV, T>(sh, tw, gt);
}

__device__ __forceinline__ int signed_freq(const int k, const int N) { return k < N / 2 ? k : k - N; }

// Residue of s r modulo N_w for a signed band frequency s in [-N/2, N/2) and a
// phase r < U: s r lies in (-N_w/2, N_w/2), so it is s r or s r + N_w (no
// division; |s| < 2^13 and r < 2^11 fit the 24-bit multiply).
__device__ __forceinline__ int band_phase(const int s, const int r, const int nw) {
    const int t = __mul24(s, r);
    return t < 0 ? t + nw : t;
}

// Q phases per batch share the LDS (Q buffers): two 256-thread groups (one
// for Q = 1) take Q / 2 phases each in turn, then every m's Q consecutive
// phases are one contiguous store / load of 8 Q bytes of the profile row.
template <int Q>
struct Batch {
    static constexpr int G = Q >= 2 ? 2 : 1;  // thread groups
    static constexpr int PPG = Q / G;         // phases per group
    static constexpr int NT = 256 * G;        // threads
};

template <int Q>
__device__ __forceinline__ void store_phases(float2 *dst, const float2 (&v)[Q]) {
    if constexpr (Q == 1) {
        dst[0] = v[0];
    } else {
#pragma unroll
        for (int q = 0; q < Q; q += 2)
            *(float4 *)(dst + q) = make_float4(v[q].x, v[q].y, v[q + 1].x, v[q + 1].y);
    }
}
template <int Q>
__device__ __forceinline__ void load_phases(const float2 *src, float2 (&v)[Q]) {
    if constexpr (Q == 1) {
        v[0] = src[0];
    } else {
#pragma unroll
        for (int q = 0; q < Q; q += 2) {
            const float4 a = *(const float4 *)(src + q);
            v[q] = make_float2(a.x, a.y);
            v[q + 1] = make_float2(a.z, a.w);
        }
    }
}

// Adjoint: prof[p][n] (row stride ld) = sum_k B[k] e^{2 pi i s_k n / N_w},
// B[k] = conj(W[k]) wfac[p] Y[p][src(k)]; bin N_w + j of the tail repeats bin j mod N_w.
// B is formed once per row, in registers (k = gt + 256 i of each group).
// With `spans`, only the spans[p].y bins from spans[p].x on (wrapping at N_w)
// and their tail copies are stored.
// Waves per SIMD the register allocation must allow: 3 (168 VGPRs) where
// that needs no spills (N <= 4096, Q <= 2; measured 0.187 -> 0.172 ms at
// N = 4096), else no constraint.
template <int N, int Q>
constexpr int kUpWaves = (N <= 4096 && Q <= 2) ? 3 : 1;

// Thread groups of the adjoint: as Batch, with T = 512 threads per transform
// at N = 8192 with one phase per batch (one radix-16 butterfly per thread and
// pass; two groups of 512 would leave 128 registers per thread), else 256.
template <int N, int Q>
struct UpBatch {
    static constexpr int T = N >= 8192 && Q == 1 ? 512 : 256;
    static constexpr int G = Q >= 2 ? 2 : 1;
    static constexpr int PPG = Q / G;
    static constexpr int NT = T * G;
};
template <int N, int Q>
__global__ __launch_bounds__((UpBatch<N, Q>::NT)) __attribute__((amdgpu_waves_per_eu(kUpWaves<N, Q>)))
void fft_up_kernel(const int2 *__restrict__ spans, const float2 *__restrict__ Y,
                   const float2 *__restrict__ Wconj, const float2 *__restrict__ wfac, int U, int ld,
                   int centred, const float2 *__restrict__ tables, float2 *__restrict__ prof,
                   const UpPairs pairs) {
    using Bt = UpBatch<N, Q>;
    constexpr int T = Bt::T;
    extern __shared__ float2 lds[];
    constexpr int SZ = N + N / 16;
    constexpr int NI = (N + T - 1) / T;
    const int p = blockIdx.x;
    const int tid = threadIdx.x, g = tid / T, gt = tid % T;
    const int nw = U * N;
    const Twiddles<float2> tw = load_twiddles<N>(lds + Q * SZ, tables, nw);
    const float2 wp = wfac[p];
    const float2 *y = Y + (size_t)p * N;
    float2 *row = prof + (size_t)p * ld;
    const int2 sp = spans ? spans[p] : make_int2(0, nw);
    if (sp.y == 0) return;  // row never read
    float2 B[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int k = gt + T * i;
        if (k < N) {
            const int src = centred ? (k + N / 2) % N : k;
            B[i] = cmulf2(cmulf2(Wconj[k], y[src]), wp);
        }
    }
    for (int r0 = 0; r0 < U; r0 += Q) {
        __syncthreads();  // tables loaded / previous batch written out
#pragma unroll 1
        for (int pp = 0; pp < Bt::PPG; ++pp) {
            const int buf = g + Bt::G * pp, rp = r0 + buf;
            // the thread index passed through an empty asm: the passes' LDS
            // and table addresses are formed per phase instead of being kept
            // in registers across phases (N = 4096: 164 -> 114 VGPRs, four
            // waves per SIMD)
            int gt_l = gt;
            asm volatile("" : "+v"(gt_l));
            float2 *sh = lds + buf * SZ;
            if constexpr (N >= 16 * T) {
                // the first radix-16 pass straight from registers: butterfly
                // j = gt + T b takes k = j + (N / 16) r, held in B[b + PB r].
                // Its signed frequency is j + (N / 16) r' (r' = r, or r - 16
                // past N / 2), so the phase factor e^{2 pi i s_k rp / N_w}
                // splits into e^{2 pi i j rp / N_w} (j rp < N_w / 16), one per
                // butterfly, and e^{2 pi i (N / 16) r' rp / N_w}, the same for
                // the whole wave and, N / 16 being a multiple of 256, one
                // entry of the coarse table.
                constexpr int PB = N / (16 * T);
                const int rpu = __builtin_amdgcn_readfirstlane(rp);
#pragma unroll
                for (int b = 0; b < PB; ++b) {
                    const int j = gt + T * b;
                    const float2 c1 = conjf2(tw.w(j * rpu));
                    float2 v[16];
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int t2 = band_phase((r < 8 ? r : r - 16) * (N / 16), rpu, nw);
                        v[r] = cmulf2(B[b + PB * r], cmulf2(c1, conjf2(tw.whi[t2 >> 8])));
                    }
                    dft<16, true>(v);
#pragma unroll
                    for (int r = 0; r < 16; ++r) sh[phys(16 * j + r)] = v[r];
                }
                __syncthreads();
                fft_passes<N, 16, true, T>(sh, tw, gt_l);
            } else {
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    const int k = gt + T * i;
                    if (k < N) {
                        const int t = band_phase(signed_freq(k, N), rp, nw);
                        sh[phys(k)] = cmulf2(B[i], conjf2(tw.w(t)));
                    }
                }
                __syncthreads();
                fft_lds<N, true, T>(sh, tw, gt_l);
            }
        }
        // Q consecutive phases of each m are contiguous in the row
#pragma unroll 2
        for (int m = tid; m < N; m += Bt::NT) {
            const int n = U * m + r0;
            bool in[Q], all = true;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                int d = n + q - sp.x;
                d += d < 0 ? nw : 0;
                in[q] = d < sp.y;
                all &= in[q];
            }
            float2 v[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) v[q] = lds[q * SZ + phys(m)];
            if (all) {
                store_phases<Q>(row + n, v);
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (in[q]) row[n + q] = v[q];
            }
            // the tail [N_w, ld) can be longer than N_w when tiles are wide
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (in[q])
                    for (int t = nw + n + q; t < ld; t += nw) row[t] = v[q];
        }
    }
    if (pairs.mode != 0) {
        // the demodulated rows (integer rho: the phasor is the sign) from the
        // row this workgroup has just stored: its global stores are visible
        // to the whole workgroup after the barrier.  The same entries as
        // pair_rows_kernel (span bins and their tail copies), and the same
        // values: the difference is rounded once, the sign is exact.
        __syncthreads();
        const float sg = pairs.sgn;
        for (int d = tid; d < sp.y; d += Bt::NT) {
            int n = sp.x + d;
            n -= n >= nw ? nw : 0;
            for (int i = n; i < ld - 1; i += nw) {
                const float2 a = row[i];
                if (pairs.mode == 2) {
                    ((float2 *)pairs.out)[(size_t)p * ld + i] = make_float2(sg * a.x, sg * a.y);
                } else {
                    const float2 b = row[i + 1];
                    ((float4 *)pairs.out)[(size_t)p * ld + i] =
                        make_float4(sg * a.x, sg * a.y, sg * (b.x - a.x), sg * (b.y - a.y));
                }
            }
        }
    }
}

// Thread groups of the forward: one transform per group of T threads, one or
// two groups per workgroup as for Batch.  With one phase per batch, T = 512 at
// N = 8192 (one radix-16 butterfly per thread and pass) and for FP64 at
// N = 4096 (twice the waves on the one 68 KiB buffer a CU can hold two of),
// else 256.  Measured at
// 2048 / 4096 rows (forward spectral stage): FP32 N = 8192 1.49 -> 0.99 ms,
// FP64 N = 4096 0.49 -> 0.33 ms; 512 threads made FP32 N = 4096 (0.20 ->
// 0.31 ms) and N = 2048 slower.
template <int N, int Q, typename C>
struct DownBatch {
    static constexpr bool kF64 = std::is_same<C, double2>::value;
    static constexpr int T = Q == 1 && (N >= 8192 || (kF64 && N >= 4096)) ? 512 : 256;
    static constexpr int G = Q >= 2 ? 2 : 1;
    static constexpr int PPG = Q / G;
    static constexpr int NT = T * G;
};

// Forward: Y[p][dst(k)] = W[k] wfac[p] sum_n prof[p][n] e^{-2 pi i s_k n / N_w}
// (row stride nw = U N), k over the N band bins.  With `spans`, row p holds
// data only in the spans[p].y bins from spans[p].x on (wrapping at N_w); the
// other bins are taken as zero without being read; a count of -1 marks a
// poisoned row, whose output is NaN.
//
// C is the arithmetic of the transforms and of the sum over phases (LDS
// buffers, twiddles, accumulators): float2 or double2; the profile is read and
// the band written in FP32 either way.  Every phase transform carries the
// whole row's energy, most of which lies outside the band and cancels in the
// sum over phases.  For a smooth scene the band keeps a small fraction of it
// (about 1/20 of the norm in the C4 looks that cancel most), and FP32
// rounding inside the transforms, relative to the whole energy, is amplified
// as much: about 8e-6 of the band there.  In FP64 what is left is the FP32
// rounding of the profile itself.
//
// Cancelling rows are found as they are transformed: the FP32 transforms'
// rounding, relative to the band, grows with kappa = sqrt(N ||h||^2 /
// ||X||^2) (||h|| the row's norm, ||X|| the band's before the weights; about
// 0.8 for a dense random scene, 0.4 to 120 over the C4 looks).  With `redo`
// the FP32 kernel sets redo[p] when kappa^2 > kappa2 (the sums in FP64, so
// the test is invariant under power-of-two scalings of the input), and the
// FP64 kernel, launched after it with the same `redo`, transforms those rows
// only and overwrites their band.
template <int N, int Q, typename C>
__global__ __launch_bounds__((DownBatch<N, Q, C>::NT)) void fft_down_kernel(
    const float2 *__restrict__ prof, const int2 *__restrict__ spans, const float2 *__restrict__ W,
    const float2 *__restrict__ wfac, int U, int centred, const C *__restrict__ tables,
    float2 *__restrict__ Y, int *__restrict__ redo, double kappa2) {
    using Bt = DownBatch<N, Q, C>;
    extern __shared__ double2 lds64[];
    C *lds = reinterpret_cast<C *>(lds64);
    constexpr int SZ = N + N / 16;
    constexpr int PER = (N + Bt::NT - 1) / Bt::NT;
    const int p = blockIdx.x;
    // the FP64 pass of the adaptive mode: only the rows the FP32 pass flagged
    if (Bt::kF64 && redo && redo[p] == 0) return;
    const int tid = threadIdx.x, g = tid / Bt::T, gt = tid % Bt::T;
    const int nw = U * N;
    const Twiddles<C> tw = load_twiddles<N>(lds + Q * SZ, tables, nw);
    const float2 *row = prof + (size_t)p * nw;
    const int2 sp = spans ? spans[p] : make_int2(0, nw);
    if (sp.y < 0) {  // poisoned row (see the merge)
        for (int k = tid; k < N; k += Bt::NT) Y[(size_t)p * N + k] = make_float2(NAN, NAN);
        if (!Bt::kF64 && redo && tid == 0) redo[p] = 0;
        return;
    }
    constexpr bool kMeasure = !Bt::kF64;
    double e_row = 0.0;  // this thread's part of ||h||^2 (kMeasure with redo)
    C acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = cmake<C>(0.0, 0.0);
    for (int r0 = 0; r0 < U; r0 += Q) {
#pragma unroll 2
        for (int m = tid; m < N; m += Bt::NT) {
            const int n = U * m + r0;
            bool in[Q], any = false;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                int d = n + q - sp.x;
                d += d < 0 ? nw : 0;
                in[q] = d < sp.y;
                any |= in[q];
            }
            float2 v[Q];
            if (any) {
                load_phases<Q>(row + n, v);
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    if (!in[q]) v[q] = make_float2(0.f, 0.f);
                    if (kMeasure && redo)
                        e_row = fma((double)v[q].x, (double)v[q].x,
                                    fma((double)v[q].y, (double)v[q].y, e_row));
                }
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) v[q] = make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) lds[q * SZ + phys(m)] = cmake<C>(v[q].x, v[q].y);
        }
        __syncthreads();
        // addresses formed per phase, as in the adjoint (N = 4096: 208 -> 162
        // VGPRs, three waves per SIMD instead of two)
        int gt_l = gt;
        asm volatile("" : "+v"(gt_l));
#pragma unroll 1
        for (int pp = 0; pp < Bt::PPG; ++pp)
            fft_lds<N, false, Bt::T>(lds + (g + Bt::G * pp) * SZ, tw, gt_l);
        if constexpr (Bt::kF64) {
#pragma unroll
            for (int qq = 0; qq < Q; ++qq) {
                const int r = r0 + qq;
                // e^{-2 pi i s_k r / N_w} for k = tid + NT q: from the tables at
                // the first k of each half of the band (s_k = k, then k - N), and
                // from there on times e^{-2 pi i NT r / N_w}
                const C step = tw.w((Bt::NT * r) % nw);
                C w = cmake<C>(1.0, 0.0);
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int k = tid + Bt::NT * q;
                    if (k >= N) break;
                    if (q == 0 || q == PER / 2) {
                        w = tw.w(band_phase(signed_freq(k, N), r, nw));
                    } else {
                        w = cmul(w, step);
                    }
                    acc[q] = cadd(acc[q], cmul(lds[qq * SZ + phys(k)], w));
                }
            }
        } else if constexpr (N >= 2 * Bt::NT) {
            // k = tid + NT q has the signed frequency tid + NT q' (q' = q, or
            // q - PER past N / 2), so e^{-2 pi i s_k r / N_w} splits into
            // e^{-2 pi i tid r / N_w} (tid r < N_w / 2), one per thread and
            // phase, and e^{-2 pi i NT q' r / N_w}, the same for the whole
            // wave and, NT being a multiple of 256, one entry of the coarse table.
#pragma unroll
            for (int qq = 0; qq < Q; ++qq) {
                const int r = __builtin_amdgcn_readfirstlane(r0 + qq);
                const C c1 = tw.w(tid * r);
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int t2 = band_phase((q < PER / 2 ? q : q - PER) * Bt::NT, r, nw);
                    const C w = cmul(c1, tw.whi[t2 >> 8]);
                    acc[q] = cadd(acc[q], cmul(lds[qq * SZ + phys(tid + Bt::NT * q)], w));
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int k = tid + Bt::NT * q;
                if (k >= N) break;
                const int s = signed_freq(k, N);
#pragma unroll
                for (int qq = 0; qq < Q; ++qq)
                    acc[q] = cadd(acc[q], cmul(lds[qq * SZ + phys(k)], tw.w(band_phase(s, r0 + qq, nw))));
            }
        }
        __syncthreads();
    }
    const C wp = cmake<C>(wfac[p].x, wfac[p].y);
    double e_band = 0.0;  // this thread's part of ||X||^2 (kMeasure with redo)
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int k = tid + Bt::NT * q;
        if (k >= N) break;
        if (kMeasure && redo)
            e_band = fma((double)acc[q].x, (double)acc[q].x,
                         fma((double)acc[q].y, (double)acc[q].y, e_band));
        const int dst = centred ? (k + N / 2) % N : k;
        const C y = cmul(cmul(cmake<C>(W[k].x, W[k].y), acc[q]), wp);
        Y[(size_t)p * N + dst] = make_float2((float)y.x, (float)y.y);
    }
    if (kMeasure && redo) {
        // block sums in a fixed order: wave sums by butterfly, then wave 0
        // adds the waves' sums from LDS (free after the last batch's barrier)
        for (int off = 32; off > 0; off >>= 1) {
            e_row += __shfl_xor(e_row, off, 64);
            e_band += __shfl_xor(e_band, off, 64);
        }
        double2 *red = lds64;
        const int lane = tid & 63, wave = tid >> 6;
        if (lane == 0) red[wave] = make_double2(e_row, e_band);
        __syncthreads();
        if (tid == 0) {
            double r = 0.0, b = 0.0;
            for (int w = 0; w < Bt::NT / 64; ++w) {
                r += red[w].x;
                b += red[w].y;
            }
            redo[p] = (double)N * r > kappa2 * b ? 1 : 0;
        }
    }
}

// LDS: Q phase buffers, then the twiddle tables (elements of `elem` bytes:
// float2 for the adjoint, double2 for the forward).
constexpr size_t kMaxLds = 160 * 1024;  // LDS of one gfx950 workgroup

#define SARHIP_FFT_SIZES(X) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)
size_t fused_lds_bytes(int N, int U, int Q, size_t elem) {
    return elem * ((size_t)Q * (N + N / 16) + 16 + N / 16 + 256 + (size_t)U * N / 256);
}

// ---- 2D spectral filter of the learned preconditioner ----------------------
// out = Dout . crop_n( IFFT2_M( f / M^2 . FFT2_M( pad_M( conj(Din) . in ) ) ) )
// for B images of n x n complex64 and an M x M real filter (M >= n, a power of
// two): the circulant preconditioner's zero-padded filtering in three passes
// -- rows (demodulation, zero padding, forward FFT), columns (forward FFT,
// filter, inverse FFT, crop), rows (inverse FFT, crop, phasor) -- with one
// 64-lane wave per length-M transform (the radix-16 Stockham passes above, in
// LDS) and four transforms per 256-thread workgroup.  Rows of the padding are
// never transformed, and the intermediate [B][n][M] array holds only the n
// rows that can be non-zero.

// e^{-2 pi i t / M} = hi[t >> 6] lo[t & 63], each entry rounded once from FP64.
template <int M>
struct WaveTw {
    const float2 *lo, *hi;
    __device__ __forceinline__ float2 operator()(const int t) const {
        return cmulf2(hi[t >> 6], lo[t & 63]);
    }
};
template <int M>
__device__ __forceinline__ WaveTw<M> wave_twiddles(float2 *dst) {
    for (int i = threadIdx.x; i < 64 + M / 64; i += blockDim.x) {
        const int t = i < 64 ? i : 64 * (i - 64);
        double sn, cs;
        sincospi(-2.0 * (double)t / (double)M, &sn, &cs);
        dst[i] = make_float2((float)cs, (float)sn);
    }
    return WaveTw<M>{dst, dst + 64};
}

// One Stockham pass of radix R and span Ns of the wave's transform on sh
// (lanes 0..63 of the wave; every wave of the workgroup reaches both barriers).
template <int M, int R, int Ns, bool INV>
__device__ __forceinline__ void wave_pass(float2 *sh, const WaveTw<M> &tw, const int lane) {
    constexpr int NB = M / R;
    constexpr int PER = (NB + 63) / 64;
    float2 v[PER][R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = lane + 64 * b;
        if (j < NB)
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = sh[phys(j + r * NB)];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = lane + 64 * b;
        if (j < NB) {
            const int jm = j % Ns;
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 t = tw(jm * r * (M / (Ns * R)));
                    v[b][r] = cmulf2(v[b][r], INV ? conjf2(t) : t);
                }
            }
            dft<R, INV>(v[b]);
            const int idx = (j / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) sh[phys(idx + r * Ns)] = v[b][r];
        }
    }
    __syncthreads();
}
template <int M, int Ns, bool INV>
__device__ __forceinline__ void wave_fft(float2 *sh, const WaveTw<M> &tw, const int lane) {
    if constexpr (Ns < M) {
        constexpr int rest = M / Ns;
        constexpr int R = rest >= 16 ? 16 : rest;
        wave_pass<M, R, Ns, INV>(sh, tw, lane);
        wave_fft<M, Ns * R, INV>(sh, tw, lane);
    }
}
template <int M>
constexpr int kSfSlot = M + M / 16;  // LDS entries of one transform (padded)
template <int M>
constexpr size_t kSfLds = sizeof(float2) * (4 * (size_t)kSfSlot<M> + 64 + M / 64);

// rows: T[b][y][kx] = FFT_M(pad_M(conj(Din) in[b][y]))[kx], y < n
template <int M>
__global__ __launch_bounds__(256) void sf_rows_fwd_kernel(const float2 *__restrict__ in,
                                                          const float2 *__restrict__ din, int n,
                                                          float2 *__restrict__ T) {
    extern __shared__ float2 lds[];
    const WaveTw<M> tw = wave_twiddles<M>(lds + 4 * kSfSlot<M>);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y = blockIdx.x * 4 + w, b = blockIdx.y;
    const bool live = y < n;
    float2 *sh = lds + w * kSfSlot<M>;
    const float2 *row = in + ((size_t)b * n + (live ? y : 0)) * n;
    for (int x = lane; x < M; x += 64) {
        float2 v = make_float2(0.f, 0.f);
        if (live && x < n) {
            v = row[x];
            if (din) v = cmulf2(v, conjf2(din[(size_t)y * n + x]));
        }
        sh[phys(x)] = v;
    }
    __syncthreads();
    wave_fft<M, 1, false>(sh, tw, lane);
    if (live) {
        float2 *dst = T + ((size_t)b * n + y) * M;
        for (int k = lane; k < M; k += 64) dst[k] = sh[phys(k)];
    }
}

// columns: T[b][y][kx] (y < n) <- crop_n IFFT_M( f[.][kx] scale FFT_M(pad_M T[b][.][kx]) )
template <int M>
__global__ __launch_bounds__(256) void sf_cols_kernel(float2 *__restrict__ T, int n,
                                                      const float *__restrict__ f, float scale) {
    extern __shared__ float2 lds[];
    const WaveTw<M> tw = wave_twiddles<M>(lds + 4 * kSfSlot<M>);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kx = blockIdx.x * 4 + w, b = blockIdx.y;
    float2 *sh = lds + w * kSfSlot<M>;
    float2 *col = T + (size_t)b * n * M + kx;
    for (int y = lane; y < M; y += 64) sh[phys(y)] = y < n ? col[(size_t)y * M] : make_float2(0.f, 0.f);
    __syncthreads();
    wave_fft<M, 1, false>(sh, tw, lane);
    for (int k = lane; k < M; k += 64) {
        const float g = f[(size_t)k * M + kx] * scale;
        const float2 v = sh[phys(k)];
        sh[phys(k)] = make_float2(v.x * g, v.y * g);
    }
    __syncthreads();
    wave_fft<M, 1, true>(sh, tw, lane);
    for (int y = lane; y < n; y += 64) col[(size_t)y * M] = sh[phys(y)];
}

// rows back: out[b][y][x] = Dout[y][x] IFFT_M(T[b][y])[x], x < n
template <int M>
__global__ __launch_bounds__(256) void sf_rows_inv_kernel(const float2 *__restrict__ T, int n,
                                                          const float2 *__restrict__ dout,
                                                          float2 *__restrict__ out) {
    extern __shared__ float2 lds[];
    const WaveTw<M> tw = wave_twiddles<M>(lds + 4 * kSfSlot<M>);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y = blockIdx.x * 4 + w, b = blockIdx.y;
    const bool live = y < n;
    float2 *sh = lds + w * kSfSlot<M>;
    const float2 *src = T + ((size_t)b * n + (live ? y : 0)) * M;
    for (int k = lane; k < M; k += 64) sh[phys(k)] = live ? src[k] : make_float2(0.f, 0.f);
    __syncthreads();
    wave_fft<M, 1, true>(sh, tw, lane);
    if (live) {
        float2 *dst = out + ((size_t)b * n + y) * n;
        for (int x = lane; x < n; x += 64) {
            float2 v = sh[phys(x)];
            if (dout) v = cmulf2(v, dout[(size_t)y * n + x]);
            dst[x] = v;
        }
    }
}

template <int M>
hipError_t spectral_filter(const float2 *in, const float2 *din, int B, int n, const float *f,
                           float2 *T, float2 *out, const float2 *dout, hipStream_t s) {
    const dim3 rows((unsigned)((n + 3) / 4), (unsigned)B), cols((unsigned)(M / 4), (unsigned)B);
    sf_rows_fwd_kernel<M><<<rows, 256, kSfLds<M>, s>>>(in, din, n, T);
    sf_cols_kernel<M><<<cols, 256, kSfLds<M>, s>>>(T, n, f, 1.0f / ((float)M * (float)M));
    sf_rows_inv_kernel<M><<<rows, 256, kSfLds<M>, s>>>(T, n, dout, out);
    return hipGetLastError();
}

template <int N, int Q>
hipError_t up(const int2 *spans, const float2 *Y, const float2 *Wconj, const float2 *wfac, int U,
              int ld, int P, int centred, const float2 *tables, float2 *prof, UpPairs pairs,
              hipStream_t s) {
    const size_t lds = fused_lds_bytes(N, U, Q, sizeof(float2));
    hipError_t e = hipFuncSetAttribute((const void *)fft_up_kernel<N, Q>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    fft_up_kernel<N, Q><<<P, UpBatch<N, Q>::NT, lds, s>>>(spans, Y, Wconj, wfac, U, ld, centred,
                                                           tables, prof, pairs);
    return hipGetLastError();
}

template <int N, int Q, typename C>
hipError_t down(const float2 *prof, const int2 *spans, const float2 *W, const float2 *wfac, int U,
                int P, int centred, const C *tables, float2 *Y, int *redo, double kappa2,
                hipStream_t s) {
    const size_t lds = fused_lds_bytes(N, U, Q, sizeof(C));
    hipError_t e = hipFuncSetAttribute((const void *)fft_down_kernel<N, Q, C>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    fft_down_kernel<N, Q, C><<<P, DownBatch<N, Q, C>::NT, lds, s>>>(
        prof, spans, W, wfac, U, centred, tables, Y, redo, kappa2);
    return hipGetLastError();
}

// The forward's instantiations whose buffers fit the LDS for every admitted
// N_w (tables of at most 16 + N/16 + 256 + 512 entries); the others are never
// selected by fused_fft_phases and are not built.
template <int N, int Q, typename C>
hipError_t down_fit(const float2 *prof, const int2 *spans, const float2 *W, const float2 *wfac,
                    int U, int P, int centred, const C *tables, float2 *Y, int *redo,
                    double kappa2, hipStream_t s) {
    constexpr size_t need = sizeof(C) * ((size_t)Q * (N + N / 16) + 16 + N / 16 + 256 + 512);
    if constexpr (need <= kMaxLds)
        return down<N, Q, C>(prof, spans, W, wfac, U, P, centred, tables, Y, redo, kappa2, s);
    else
        return hipErrorInvalidValue;
}

template <typename C>
hipError_t down_n(const float2 *prof, const int2 *spans, const float2 *W, const float2 *wfac,
                  int N, int U, int P, int Q, int centred, const C *tables, float2 *Y, int *redo,
                  double kappa2, hipStream_t s) {
#define DOWN_ARGS prof, spans, W, wfac, U, P, centred, tables, Y, redo, kappa2, s
#define DOWN_CASE(n)                                        \
    if (N == n) {                                           \
        if (Q == 8) return down_fit<n, 8, C>(DOWN_ARGS);    \
        if (Q == 4) return down_fit<n, 4, C>(DOWN_ARGS);    \
        if (Q == 2) return down_fit<n, 2, C>(DOWN_ARGS);    \
        return down_fit<n, 1, C>(DOWN_ARGS);                \
    }
    SARHIP_FFT_SIZES(DOWN_CASE)
#undef DOWN_CASE
#undef DOWN_ARGS
    return hipErrorInvalidValue;
}

}  // namespace

// Phases per batch: the largest Q of 8, 4, 2, 1 that divides U, whose Q
// buffers and the tables fit the LDS, and not above a cap.  When the rows are
// read or written whole, batches make the row accesses contiguous: cap 8 for
// the adjoint, 2 for the forward (N = 4096, U = 8: 0.39 ms against 0.41 ms
// with 4).  When only sparse spans of the rows are touched the transforms
// dominate and occupancy wins: cap 1 (0.19 / 0.21 ms against 0.24 / 0.23 ms
// with the dense caps, spans of ~9% of the row).  SARHIP_FFT_PHASES
// overrides every cap.
int fused_fft_phases(int N, int U, bool up, bool sparse, bool fp64) {
    const size_t elem = fp64 ? sizeof(double2) : sizeof(float2);
    int cap = sparse ? 1 : up ? 8 : 2;
    if (const char *e = std::getenv("SARHIP_FFT_PHASES")) cap = std::atoi(e);
    for (int Q : {8, 4, 2}) {
        if (Q <= cap && U % Q == 0 && fused_lds_bytes(N, U, Q, elem) <= kMaxLds) return Q;
    }
    return 1;
}

// N a power of two in [64, 8192], N_w = U N a multiple of 256 up to 2^17
// (the two-level twiddle tables stay small; one FP64 phase buffer and the
// tables then fit the LDS: 156 KiB at N = 8192, N_w = 2^17).
bool fused_fft_supported(int N, int U) {
    const long long nw = (long long)U * N;
    return N >= 64 && N <= 8192 && (N & (N - 1)) == 0 && U >= 1 && nw % 256 == 0 &&
           nw <= (1 << 17) && fused_lds_bytes(N, U, 1, sizeof(double2)) <= kMaxLds;
}

hipError_t launch_fft_up(const int2 *spans, const float2 *Y, const float2 *Wconj, const float2 *wfac,
                         int N, int U, int ld, int P, int centred, const float2 *tables,
                         float2 *prof, bool sparse, UpPairs pairs, hipStream_t s) {
    if (pairs.mode != 0 && !spans) return hipErrorInvalidValue;
    const int Q = fused_fft_phases(N, U, true, sparse, false);
#define UP_ARGS spans, Y, Wconj, wfac, U, ld, P, centred, tables, prof, pairs, s
#define UP_CASE(n)                                     \
    if (N == n) {                                      \
        if (Q == 8) return up<n, 8>(UP_ARGS);          \
        if (Q == 4) return up<n, 4>(UP_ARGS);          \
        if (Q == 2) return up<n, 2>(UP_ARGS);          \
        return up<n, 1>(UP_ARGS);                      \
    }
    SARHIP_FFT_SIZES(UP_CASE)
#undef UP_CASE
#undef UP_ARGS
    return hipErrorInvalidValue;
}

hipError_t launch_fft_down(const float2 *prof, const int2 *spans, const float2 *W, const float2 *wfac,
                           int N, int U, int P, int centred, const float2 *tables,
                           const double2 *tables64, float2 *Y, bool sparse, int *redo,
                           double kappa2, hipStream_t s) {
    const int Q64 = fused_fft_phases(N, U, false, sparse, true);
    if (tables64 && !redo)
        return down_n<double2>(prof, spans, W, wfac, N, U, P, Q64, centred, tables64, Y, nullptr,
                               0.0, s);
    const int Q = fused_fft_phases(N, U, false, sparse, false);
    const bool adaptive = tables64 && redo;
    hipError_t e = down_n<float2>(prof, spans, W, wfac, N, U, P, Q, centred, tables, Y,
                                  adaptive ? redo : nullptr, kappa2, s);
    if (e != hipSuccess || !adaptive) return e;
    return down_n<double2>(prof, spans, W, wfac, N, U, P, Q64, centred, tables64, Y, redo, kappa2,
                           s);
}

bool spectral_filter_supported(int n, int M) {
    return n >= 1 && n <= M && M >= 64 && M <= 1024 && (M & (M - 1)) == 0;
}

hipError_t launch_spectral_filter(const float2 *in, const float2 *din, int B, int n, int M,
                                  const float *f, float2 *work, float2 *out, const float2 *dout,
                                  hipStream_t s) {
    switch (M) {
        case 64: return spectral_filter<64>(in, din, B, n, f, work, out, dout, s);
        case 128: return spectral_filter<128>(in, din, B, n, f, work, out, dout, s);
        case 256: return spectral_filter<256>(in, din, B, n, f, work, out, dout, s);
        case 512: return spectral_filter<512>(in, din, B, n, f, work, out, dout, s);
        case 1024: return spectral_filter<1024>(in, din, B, n, f, work, out, dout, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sarhip
