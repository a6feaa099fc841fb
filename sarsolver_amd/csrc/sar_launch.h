// sar_launch.h -- host-side launcher declarations for the gfx950 SAR kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "sar_device.h"

namespace sarhip {

struct GatherArgs {
    const double *px, *py, *pz;  // tile-ordered pixel positions (SoA, FP64)
    const int *perm;             // tile order -> caller's pixel index
    int M;
    const PulseConst *pc;
    int P;
    const float2 *prof;  // [P][nw] upsampled range profiles (w_p folded in)
    int nw;
    double by_dx, kc_rev, inv_nw;
    float2 *out;  // [M] caller order
    int interp;
    bool mono;
};

struct BinArgs {
    const double *px, *py, *pz;
    const float2 *amps_perm;  // [M] tile order
    int M, tile_len, num_tiles;
    const double4 *tiles;  // per forward tile: centre xyz, bounding radius
    const PulseConst *pc;
    int P, pg, win;
    float2 *prof;  // [P][nw], zeroed before the launch
    int nw;
    double by_dx, kc_rev, inv_nw;
    int interp;
    bool mono;
};

hipError_t launch_window_pad(const float2 *Y, const float2 *Wconj, const float2 *wfac, int N,
                             int nw, int P, int centred, float2 *K, hipStream_t s);
hipError_t launch_truncate(const float2 *G, const float2 *W, const float2 *wfac, int N, int nw,
                           int P, int centred, float2 *Y, hipStream_t s);
hipError_t launch_fill_zero(void *buf, size_t bytes, hipStream_t s);
hipError_t launch_gather_perm(const float2 *src, const int *perm, int M, float2 *dst,
                              hipStream_t s);
hipError_t launch_bp_gather(const GatherArgs &a, hipStream_t s);
hipError_t launch_fp_bin(const BinArgs &a, hipStream_t s);

}  // namespace sarhip
