// sar_device.h -- device-side types and helpers shared by the gfx950 SAR kernels.
//
// All HIP code in this library is compiled with -ffp-contract=off so that the
// FP64 range arithmetic is evaluated with the same roundings as the
// reference's x86-64 build (src/sarsolver/cxx_binding/sarsolver_cxx.cpp:10-20);
// fused multiply-adds are written explicitly (fma / fmaf) where wanted.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sarhip {

constexpr int kInterpNearestRef = 0;
constexpr int kInterpNearest = 1;
constexpr int kInterpLinear = 2;

constexpr int kPrecFp64 = 0;
constexpr int kPrecSeries = 1;
constexpr int kPrecHilo = 2;
constexpr int kPrecFp32 = 3;

// One record per pulse (64 bytes).  Kernels index it with a wave-uniform pulse
// index, so the compiler reads it with scalar loads into SGPRs.
struct PulseConst {
    double tx[3];   // transmitter phase centre
    double rx[3];   // receiver phase centre
    double r_srp;   // bistatic range of the SRP, evaluated as CPP:283
    double valid;   // 1.0 if the pulse geometry is finite, else 0.0 (pulse skipped)
};
static_assert(sizeof(PulseConst) == 64, "PulseConst must be 64 bytes");

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

// sin(2 pi r), cos(2 pi r) for |r| <= 0.5 revolutions.  Quadrant reduction
// then minimax-style Taylor polynomials in t = 4r - q, |t| <= 1/2 (quarter
// turns), i.e. angle |pi t / 2| <= pi / 4.  Max abs error ~2e-8 in FP32.
__device__ __forceinline__ void sincos_rev(const float r, float *s, float *c) {
    const float x = 4.0f * r;
    const float q = rintf(x);
    const float t = x - q;                     // exact (Sterbenz)
    const float a = t * 1.57079632679489662f;  // angle in radians, |a| <= pi/4
    const float a2 = a * a;
    // sin a = a (1 - a^2/6 + a^4/120 - a^6/5040 + a^8/362880)
    float sp = fmaf(a2, 2.75573192e-6f, -1.98412698e-4f);
    sp = fmaf(a2, sp, 8.33333333e-3f);
    sp = fmaf(a2, sp, -1.66666667e-1f);
    const float sa = fmaf(a * a2, sp, a);
    // cos a = 1 - a^2/2 + a^4/24 - a^6/720 + a^8/40320
    float cp = fmaf(a2, 2.48015873e-5f, -1.38888889e-3f);
    cp = fmaf(a2, cp, 4.16666667e-2f);
    cp = fmaf(a2, cp, -0.5f);
    const float ca = fmaf(a2, cp, 1.0f);
    const int qi = ((int)q) & 3;
    // rotate by q quarter turns
    float so = sa, co = ca;
    if (qi == 1) { so = ca; co = -sa; }
    else if (qi == 2) { so = -sa; co = -ca; }
    else if (qi == 3) { so = -ca; co = sa; }
    *s = so;
    *c = co;
}

// Phasor e^{i 2 pi theta} for an FP64 phase in revolutions: the integer part
// is removed in FP64, the remaining |r| <= 1/2 handled in FP32.
__device__ __forceinline__ void phasor_rev(const double theta, float *s, float *c) {
    const float r = (float)(theta - rint(theta));
    sincos_rev(r, s, c);
}

}  // namespace sarhip
