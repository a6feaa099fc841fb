// sar_device.h -- device-side types and helpers shared by the gfx950 SAR kernels.
//
// All HIP code in this library is compiled with -ffp-contract=off so that the
// FP64 range arithmetic is evaluated with the same roundings as the
// reference's x86-64 build (src/sarsolver/cxx_binding/sarsolver_cxx.cpp:10-20);
// fused multiply-adds are written explicitly (fma / fmaf) where wanted.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sarhip {

constexpr int kInterpNearestRef = 0;
constexpr int kInterpNearest = 1;
constexpr int kInterpLinear = 2;

constexpr int kPrecFp64 = 0;
constexpr int kPrecSeries = 1;
constexpr int kPrecHilo = 2;
constexpr int kPrecFp32 = 3;

// Pixels per tile: a workgroup of the gather kernel owns one tile (256 threads,
// kPixPerThread pixels each); the range series is expanded about the tile's
// centre.  A forward workgroup owns a region of consecutive tiles and a group
// of pulses (both chosen per plan from the LDS budget).
constexpr int kThreads = 256;
constexpr int kPixPerThread = 4;
constexpr int kTilePixels = kThreads * kPixPerThread;  // 1024
// |frac(v) - 1/2| below this (in bins) sends a nearest-neighbour decision of
// the series path to an exact FP64 re-evaluation with the reference's formula.
constexpr double kTieMargin = 1.0e-7;

// One record per pulse (64 bytes), indexed with a wave-uniform pulse index so
// the compiler reads it with scalar loads into SGPRs.
struct PulseConst {
    double tx[3];   // transmitter phase centre
    double rx[3];   // receiver phase centre
    double r_srp;   // bistatic range of the SRP, evaluated as CPP:283
    double valid;   // 1.0 if the pulse geometry is finite (invalid pulses are skipped)
};
static_assert(sizeof(PulseConst) == 64, "PulseConst must be 64 bytes");

// Series constants of one (tile, pulse).  For a pixel X = X0 + d (X0 the tile
// centre, d exact) and a leg with phase centre T, a = T - X0, A = |a|:
//     |T - X| = A sqrt(1 + t),  t = b.d + c |d|^2,  b = -2a/A^2,  c = 1/A^2,
// and with g(t) = sqrt(1+t) - 1 (5th-order series; the plan bounds the rest):
//     u     = dR * by_dx   = ku_T g_T + ku_R g_R + cu
//     theta = dR * kc/2pi  = kt_T g_T + kt_R g_R + ct      (ct reduced mod 1)
// Monostatic pulses use one leg with the factors doubled (8 doubles).
struct SeriesMono {
    double bx, by, bz, c, ku, cu, kt, ct;
};
struct SeriesBi {
    double tbx, tby, tbz, tc, rbx, rby, rbz, rc;
    double kut, kur, cu, ktt, ktr, ct, pad0, pad1;
};
static_assert(sizeof(SeriesMono) == 64, "SeriesMono must be 64 bytes");
// Backprojection record of one (tile, pulse) for monostatic planar tiles with
// the third-order series, the argument scaled by 1/c so that it takes two
// FMAs: s = t / c = bx' dx + by' dy + |d|^2 with bx' = bx / c, by' = by / c,
// u = cu + s (k1 + s (k2 + s k3)) with k_i = ku (1/2, -1/8, 1/16)_i c^i and cu
// moved by a multiple of nw into the padded row (as the linear gather needs),
// theta = rho u + phi with rho = kc_rev / by_dx and phi = ct - rho cu.
struct SeriesGather {
    double bx, by, k1, k2, k3, cu, phi;
    float phi_pair;  // phi - rho / 2 reduced to [-1/2, 1/2) (demodulated pair rows)
    float pad;
};
static_assert(sizeof(SeriesGather) == 64, "SeriesGather must be 64 bytes");
static_assert(sizeof(SeriesBi) == 128, "SeriesBi must be 128 bytes");

// ||a - x|| with the reference's evaluation order: ((dx^2 + dy^2) + dz^2), sqrt.
__device__ __forceinline__ double ref_distance(const double a0, const double a1, const double a2,
                                               const double x, const double y, const double z) {
    const double d0 = a0 - x, d1 = a1 - y, d2 = a2 - z;
    return sqrt((d0 * d0 + d1 * d1) + d2 * d2);
}

// Differential bistatic range dR = (|Tx-X| + |Rx-X|) - R_srp (CPP:287, CPP:346).
// For a monostatic pulse (Tx == Rx bitwise) d + d is exactly the reference's sum.
template <bool MONO>
__device__ __forceinline__ double ref_dr(const PulseConst &c, const double x, const double y,
                                         const double z) {
    const double d1 = ref_distance(c.tx[0], c.tx[1], c.tx[2], x, y, z);
    if (MONO) return (d1 + d1) - c.r_srp;
    const double d2 = ref_distance(c.rx[0], c.rx[1], c.rx[2], x, y, z);
    return (d1 + d2) - c.r_srp;
}

// g(t) = sqrt(1 + t) - 1 = t/2 - t^2/8 + t^3/16 - 5 t^4/128 + 7 t^5/256 - O(21 t^6/1024).
// The plan checks that the truncation term stays below 1e-9 bin and 1e-9
// revolution for every (tile, pulse) before it selects a series mode.
__device__ __forceinline__ double sqrt1pm1_series(const double t) {
    return t * fma(t, fma(t, fma(t, fma(t, 0.02734375, -0.0390625), 0.0625), -0.125), 0.5);
}
// The two innermost coefficients held in registers the compiler cannot
// rematerialise: gfx9 VOP3 has no literal operand, so with literals the
// innermost step is a VOP2 fmac whose accumulator is re-initialised by two
// v_mov per evaluation.  An identity DPP move (quad_perm [0,1,2,3]) yields
// the same bits in a register the compiler treats as opaque, computed once.
// ORD 5: (7/256, -5/128); ORD 3: (1/16, -1/8).
struct PolyRegs {
    double hi, lo;
};
__device__ __forceinline__ double opaque_reg(const double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xE4, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xE4, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int ORD = 5>
__device__ __forceinline__ PolyRegs make_poly_regs() {
    static_assert(ORD == 3 || ORD == 5, "series order 3 or 5");
    if constexpr (ORD == 5) return PolyRegs{opaque_reg(0.02734375), opaque_reg(-0.0390625)};
    else return PolyRegs{opaque_reg(0.0625), opaque_reg(-0.125)};
}
// sqrt(1 + t) - 1 truncated after t^ORD; pr from make_poly_regs<ORD>().
template <int ORD = 5>
__device__ __forceinline__ double sqrt1pm1_series(const double t, const PolyRegs &pr) {
    if constexpr (ORD == 5) return t * fma(t, fma(t, fma(t, fma(t, pr.hi, pr.lo), 0.0625), -0.125), 0.5);
    else return t * fma(t, fma(t, pr.hi, pr.lo), 0.5);
}
__device__ __forceinline__ float sqrt1pm1_series_f(const float t) {
    return t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 0.02734375f, -0.0390625f), 0.0625f), -0.125f), 0.5f);
}

// Mathematical modulo of an integer-valued double into [0, nw), equal to the
// reference's math_modulo((long)n, nw) (CPP:22-24) for |n| < 2^52.
__device__ __forceinline__ int wrap_bin(const double n, const double nwd, const double inv_nw,
                                        const int nw) {
    const double q = floor(n * inv_nw);
    int r = (int)(n - q * nwd);
    r = r < 0 ? r + nw : r;
    r = r >= nw ? r - nw : r;
    return r;
}

// Phasor e^{i 2 pi theta} for an FP64 phase in revolutions: the integer part
// is removed in FP64 and the remaining |r| <= 1/2 goes to the hardware
// v_sin_f32 / v_cos_f32, which take their argument in revolutions (measured
// on MI355X over [-1/2, 1/2]: max abs error 1.2e-7, rms 3.5e-8).
// Same for a phase reduced to [0, 1) by v_fract_f64 (one instruction instead
// of a round and a subtraction; the FP32 value then carries up to 3e-8 rev of
// rounding instead of 1.5e-8).
__device__ __forceinline__ void phasor_fract(const double theta, float *s, float *c) {
    const float r = (float)__builtin_amdgcn_fract(theta);
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}
// As phasor_fract, with the argument's FP32 rounding residual d (radians)
// carried by a first-order correction: sin(r + d) = s + d c, cos(r + d) = c - d s
// (|d| <= pi 2^-24, so the neglected d^2 / 2 is below 2^-45).
__device__ __forceinline__ void phasor_fract_fine(const double theta, float *s, float *c) {
    const double f = __builtin_amdgcn_fract(theta);
    const float r = (float)f;
    const float d = (float)(f - (double)r) * 6.28318530717958648f;
    const float s0 = __builtin_amdgcn_sinf(r), c0 = __builtin_amdgcn_cosf(r);
    *s = fmaf(d, c0, s0);
    *c = fmaf(-d, s0, c0);
}
__device__ __forceinline__ void phasor_rev(const double theta, float *s, float *c) {
    const float r = (float)(theta - rint(theta));
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// Pixel record for scatter kernels (64 bytes, tile order): offset from the
// series tile centre (exact), |offset|^2, absolute position, finite flag.
struct PixRec {
    double dx, dy, dz, d2, x, y, z, ok;
};
static_assert(sizeof(PixRec) == 64, "PixRec must be 64 bytes");

// Pixel record of the planar linear scatter (32 bytes, tile order): offsets
// from the tile centre, |offset|^2 and the pixel's (scaled) amplitude.
struct ScatRec {
    double dx, dy, d2;
    float2 a;
};
static_assert(sizeof(ScatRec) == 32, "ScatRec must be 32 bytes");

}  // namespace sarhip
