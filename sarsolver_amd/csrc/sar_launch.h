// sar_launch.h -- host-side launcher declarations for the gfx950 SAR kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "sar_device.h"

namespace sarhip {

// Everything a gather / scatter kernel reads, passed by value.
// Per-pulse stride (in bins) of the forward's LDS windows: the window's bins,
// two spare bins that out-of-window contributions are sent to (never read),
// rounded up to an odd count (consecutive pulses start in different banks).
inline int fwd_window_stride(int win) { return (win + 2) | 1; }

struct KernelParams {
    // pixels, in tile (Morton) order
    const double *px, *py, *pz;  // FP64 SoA positions
    const int *perm;             // tile order -> caller's pixel index
    int M;
    int num_tiles;               // ceil(M / kTilePixels)
    const double4 *tile_c;       // per tile: series centre (multiple of 2^-10 m), radius
    const PixRec *pixrec;        // [M] per-pixel records (scatter kernels)
    const ScatRec *srec;         // [M] planar scatter records, rewritten per forward call (or NULL)
    // pulses
    const PulseConst *pc;  // [P]
    const int *rows;       // valid pulse rows [Pv]
    int Pv;
    const void *series;  // [num_tiles][Pv] SeriesMono / SeriesBi (series, hilo modes)
    // profiles [P][nw] complex64
    float2 *prof;
    int nw;
    double nwd, inv_nw, by_dx, kc_rev;
    double rho;  // kc_rev / by_dx: phase revolutions per bin
    // adjoint profile rows: stride ld = nw + pad, bins [nw, ld) repeat bins
    // [0, ld - nw) (mod nw); shift_lo bounds |u - u(tile centre)| in bins
    int ld;
    double shift_lo;
    // adjoint output (caller order)
    float2 *out;
    // forward
    const float2 *amps_perm;  // [M] tile order
    int num_regions;          // forward regions of region_tiles consecutive tiles
    int region_tiles;
    const double4 *region_c;  // per region: bounding-box centre, radius
    const double4 *region_h;  // per region: bounding-box half-extents (x, y, z)
    int fwd_pulses;           // pulses (LDS windows) per forward workgroup, power of 2
    int fwd_groups;           // ceil(Pv / fwd_pulses)
    int win;                  // LDS window length per pulse (0 = no windows)
    int win_stride;           // odd per-pulse stride of the LDS windows (fwd_window_stride)
    int2 *partial;            // [Pv][num_regions][win] integer window sums
    double *wbases;           // [Pv][num_regions] window bases (once per plan, window_base_kernel)
    float *wlive;             // [Pv][num_regions] 1 for a window with a non-zero sum, else 0
    int2 *fspan;              // [P] covered bins of each forward row: {first bin, count}; count
                              // -1: the row is poisoned (a contribution could not be windowed)
    int *fpoison;             // [P] set by the scatter, read and cleared by the merge
    int fwd_spans;            // 1: the merge stores spans (fused spectral stage, rows not cleared)
    float *region_l1;         // [num_regions] bound on the region's window bin sums (per call)
    const int *region_cmax;   // [num_regions] pixels per adjacent bin pair, max over pulses
                              // (once per plan), or NULL: the bound is the region's L1 norm
    int bp_splits;            // backprojection: pulse splits per tile (workgroups per tile)
    double2 *bp_partial;      // [bp_splits][M] partial images (tile order) when bp_splits > 1
    const double *bp_cu;      // [num_tiles][Pv] series `cu` moved into [shift_lo, shift_lo + nw)
    const SeriesGather *series_g;  // [num_tiles][Pv] gather records (monostatic planar, order 3) or NULL
    const float4 *bp_pairs;   // [P][ld] demodulated adjoint row pairs (see pair_rows_kernel) or NULL
    int series_order;         // series order of the linear gather / scatter (3 or 5)
    // variant
    int interp, precision;
    bool mono;
    bool planar;  // every pixel's offset from its tile centre has dz == 0
};

// A regular grid filling a rotated, translated cuboid (GridSarScene):
// sizes and edge lengths in (z, y, x) order, rotation row-major.
struct GridSpec {
    int nz, ny, nx;
    double lz, ly, lx;
    double rot[9];
    double centre[3];
};

// scene set-up (scene_kernels.hip)
hipError_t launch_grid_positions(const GridSpec &g, double *pos, hipStream_t s);
hipError_t launch_bbox(const double *pos, int M, double *part, int part_blocks, double *out,
                       hipStream_t s);
// scale: per-coordinate factor on the quantisation; axes: the coordinate of
// each Morton bit position (axes[0] takes the lowest bits)
hipError_t launch_morton(const double *pos, int M, const double lo[3], double qscale,
                         const double scale[3], const int axes[3], unsigned long long *keys,
                         int *idx, hipStream_t s);
hipError_t sort_pairs(void *temp, size_t &temp_bytes, const unsigned long long *keys_in,
                      unsigned long long *keys_out, const int *vals_in, int *vals_out, int M,
                      hipStream_t s);
hipError_t launch_gather_sorted(const double *pos, const int *perm, int M, double *sx, double *sy,
                                double *sz, hipStream_t s);
hipError_t launch_tile_stats(const double *sx, const double *sy, const double *sz, int M, int len,
                             bool snap, double4 *centre_r, double4 *half_planar, hipStream_t s);
hipError_t launch_pixrec(const double *sx, const double *sy, const double *sz, int M,
                         const double4 *tiles, PixRec *recs, hipStream_t s);
hipError_t launch_baseband(const double *pos, int M, const double spot[3], const double disp[3],
                           double sgn, double2 *out, hipStream_t s);

hipError_t launch_window_pad(const float2 *Y, const float2 *Wconj, const float2 *wfac, int N,
                             int nw, int ld, int P, int centred, float2 *K, hipStream_t s);
hipError_t launch_wrap_pad(float2 *prof, int nw, int ld, int P, hipStream_t s);
hipError_t launch_truncate(const float2 *G, const float2 *W, const float2 *wfac, int N, int nw,
                           int P, int centred, float2 *Y, hipStream_t s);
// the same from FP64 rows, and FP64 copies of complex64 rows
hipError_t launch_truncate64(const double2 *G, const float2 *W, const float2 *wfac, int N, int nw,
                             int P, int centred, float2 *Y, hipStream_t s);
hipError_t launch_widen(const float2 *src, double2 *dst, size_t n, hipStream_t s);
hipError_t launch_fill_zero(void *buf, size_t bytes, hipStream_t s);
// demodulated pairs of the adjoint profile rows over each valid row's span and
// its tail copies: pairs[i] = (h[i], h[i+1] - h[i]) e^{-2 pi i rho (i + 1/2)}
// (float4 entries), or with a_only the first of them alone (float2 entries)
hipError_t launch_pair_rows(const float2 *prof, const int2 *spans, const int *rows, int Pv,
                            int nw, int ld, double rho, bool a_only, void *pairs, hipStream_t s);
hipError_t launch_gather_perm(const float2 *src, const int *perm, int M, float2 *dst,
                              hipStream_t s);
hipError_t launch_series_consts(const double4 *tile_c, int num_tiles, const PulseConst *pc,
                                const int *rows, int Pv, double by_dx, double kc_rev, bool mono,
                                void *out, hipStream_t s);
hipError_t launch_bp_gather(const KernelParams &k, hipStream_t s);
hipError_t launch_bp_reduce(const KernelParams &k, hipStream_t s);  // pulse-split images
hipError_t launch_region_l1(const KernelParams &k, hipStream_t s);  // before launch_fp_bin
// Once per plan, after the pixel records and window sizes: cmax[r] (zeroed by
// the caller) = largest number of pixels of region r on two adjacent bins of
// one pulse; 0x7fffffff where a pixel falls outside its window.
hipError_t launch_region_count(const KernelParams &k, int *cmax, hipStream_t s);
// Per valid pulse, the bins {first, count} of its adjoint row that the
// backprojection can read (span[p] for p = rows[j]; other entries untouched).
hipError_t launch_adjoint_spans(const PulseConst *pc, const int *rows, int Pv, const double4 *tiles,
                                int nt, bool mono, double by_dx, int nw, int2 *span, hipStream_t s);
// Gather records from the monostatic series constants (cu moved as launch_bp_shift,
// plus `off`: 0 for linear interpolation, the nearest rules' offset otherwise).
hipError_t launch_series_gather(const void *series, size_t n, double shift_lo, double inv_nw,
                                double nwd, double rho, double off, SeriesGather *out,
                                hipStream_t s);
hipError_t launch_bp_shift(const void *series, bool mono, size_t n, double shift_lo, double inv_nw,
                           double nwd, double *out, hipStream_t s);
// dataset resampling stages (resample_kernels.hip)
hipError_t launch_rs_ramp(const float2 *in, int P, int N, const double *dr, double fs_over_n,
                          double sgn_fc, double inv_c, float2 *out, hipStream_t s);
hipError_t launch_rs_window(const float2 *prof, int N, int N2, const float *win, const float2 *kern,
                            int K, const int *starts, int P2, float2 *out, hipStream_t s);
hipError_t launch_decode_signal(const void *raw, int format, int nv, int ns, const float *sf,
                                float2 *out, hipStream_t s);
// fused spectral stages (fft_kernels.hip): N-sample band <-> U N-sample rows
bool fused_fft_supported(int N, int U);
int fused_fft_phases(int N, int U, bool up, bool sparse, bool fp64);
// tables: [16 | N/16 | 256 | U N/256] two-level twiddle tables,
// e^{-2 pi i q / N}, e^{-2 pi i 16 q / N}, e^{-2 pi i q / N_w}, e^{-2 pi i 256 q / N_w}
// (tables64: the same in FP64)
// spans: per-row bins to store {first, count} (and their tail copies), or NULL for whole rows;
// sparse: the spans cover a small part of the rows (picks the phase batching)
// Demodulated rows formed by the adjoint spectral kernel itself (an integer
// rho, whose phasor e^{-2 pi i rho (n + 1/2)} is sgn = (-1)^rho): mode 0 none,
// 1 the pairs sgn (h[n], h[n+1] - h[n]) (float4 entries), 2 sgn h[n] alone
// (float2 entries), over the spans (which must be given) and their tails.
struct UpPairs {
    void *out;
    int mode;
    float sgn;
};
hipError_t launch_fft_up(const int2 *spans, const float2 *Y, const float2 *Wconj, const float2 *wfac,
                         int N, int U, int ld, int P, int centred, const float2 *tables,
                         float2 *prof, bool sparse, UpPairs pairs, hipStream_t s);
// spans: per-row covered bins {first, count} (only those are read), or NULL for whole rows;
// sparse: as for launch_fft_up.  tables64 non-NULL and redo NULL: transforms and
// phase sums in FP64 for every row; both non-NULL: FP32, then FP64 again for the
// rows whose cancellation factor kappa^2 = N ||h||^2 / ||X||^2 exceeds kappa2
// (redo [P] receives the flags); tables64 NULL: FP32
hipError_t launch_fft_down(const float2 *prof, const int2 *spans, const float2 *W, const float2 *wfac,
                           int N, int U, int P, int centred, const float2 *tables,
                           const double2 *tables64, float2 *Y, bool sparse, int *redo,
                           double kappa2, hipStream_t s);
// spectral filter of the learned preconditioner (fft_kernels.hip): out =
// dout . crop_n IFFT2_M(f / M^2 . FFT2_M(pad_M(conj(din) . in))), B images n x n,
// filter f [M][M] float (ky, kx), work >= B n M complex64, din / dout optional
bool spectral_filter_supported(int n, int M);
hipError_t launch_spectral_filter(const float2 *in, const float2 *din, int B, int n, int M,
                                  const float *f, float2 *work, float2 *out, const float2 *dout,
                                  hipStream_t s);
hipError_t launch_fp_bin(const KernelParams &k, hipStream_t s);
// once per plan: k.wbases[j][region] for every valid pulse j and forward region
hipError_t launch_window_bases(const KernelParams &k, hipStream_t s);
// dynamic LDS bytes of the merge for the number of forward regions and the window length
size_t fp_reduce_lds(int num_regions, int win);
hipError_t launch_fp_reduce(const KernelParams &k, hipStream_t s);

}  // namespace sarhip
