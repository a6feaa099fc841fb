// scene_kernels.hip -- pixel-side set-up of a plan on the device (gfx950).
//
// A plan orders its pixels along a 63-bit Morton curve and cuts the order
// into tiles of kTilePixels pixels (series centres, forward regions).  All of
// that is O(M) work on FP64 positions, done here on the device so that a
// plan never needs the positions in host memory:
//   grid_positions_kernel : GridSarScene.positions (numpy/hypothesis.py:191-206)
//   bbox_*_kernel          : bounding box of the finite pixels
//   morton_kernel          : quantised Morton key per pixel (+ its index)
//   (hipcub radix sort of (key, index): the order of sorting the pairs)
//   gather_sorted_kernel   : positions in tile order, SoA
//   tile_stats_kernel      : per-tile centre, radius, half-extents, planarity
//   pixrec_kernel          : per-pixel records relative to the tile centre
//   baseband_kernel        : generate_simple_basebanding_values
//                            (utils/k_space.py:109-134)
// Every floating-point step is written out with the rounding it needs
// (-ffp-contract=off; fma only where written), so results do not depend on
// the thread layout.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "sar_device.h"
#include "sar_launch.h"

namespace sarhip {

namespace {

constexpr int kBlock = 256;

// numpy.linspace(-L/2, L/2, n)[i], step by step as numpy 2.x evaluates it:
// delta = stop - start; for n > 1 step = delta / (n - 1) and y = i * step
// (or (i / (n - 1)) * delta when step underflows to 0); for n = 1 y = 0 * delta;
// then y + start, and the last sample is `stop` itself.
__device__ double grid_axis(int i, int n, double L) {
    const double start = -0.5 * L, stop = 0.5 * L;
    const double delta = stop - start;
    double y;
    if (n > 1) {
        const double div = (double)(n - 1);
        const double step = delta / div;
        y = step == 0.0 ? ((double)i / div) * delta : (double)i * step;
    } else {
        y = (double)i * delta;
    }
    y = y + start;
    if (n > 1 && i == n - 1) y = stop;
    return y;
}

// Flattened in C order over (z, y, x); local point (x, y, z); position =
// R @ local + centre, the 3-term product as an FMA chain in k order
// (fma(R2, z, fma(R1, y, R0 * x))), which is how numpy's matmul rounds it on
// OpenBLAS's FMA kernels.
__global__ __launch_bounds__(kBlock) void grid_positions_kernel(GridSpec g, double *__restrict__ pos) {
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    const long long M = (long long)g.nz * g.ny * g.nx;
    if (m >= M) return;
    const int ix = (int)(m % g.nx);
    const int iy = (int)((m / g.nx) % g.ny);
    const int iz = (int)(m / ((long long)g.nx * g.ny));
    const double lx = grid_axis(ix, g.nx, g.lx);
    const double ly = grid_axis(iy, g.ny, g.ly);
    const double lz = grid_axis(iz, g.nz, g.lz);
    for (int r = 0; r < 3; ++r) {
        const double v = fma(g.rot[3 * r + 2], lz, fma(g.rot[3 * r + 1], ly, g.rot[3 * r] * lx));
        pos[3 * m + r] = v + g.centre[r];
    }
}

__device__ __forceinline__ bool finite3(double x, double y, double z) {
    return isfinite(x) && isfinite(y) && isfinite(z);
}

// Block-wide min (k < 3) / max (k >= 3) of six values held by every thread.
__device__ void block_minmax6(double v[6], double *sh /* [6][kBlock] */) {
    const int t = threadIdx.x;
    for (int k = 0; k < 6; ++k) sh[k * kBlock + t] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 6; ++k) {
                const double a = sh[k * kBlock + t], b = sh[k * kBlock + t + s];
                sh[k * kBlock + t] = k < 3 ? fmin(a, b) : fmax(a, b);
            }
        __syncthreads();
    }
    for (int k = 0; k < 6; ++k) v[k] = sh[k * kBlock];
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void bbox_partial_kernel(const double *__restrict__ pos, int M,
                                                              double *__restrict__ part) {
    __shared__ double sh[6 * kBlock];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int m = blockIdx.x * kBlock + threadIdx.x; m < M; m += gridDim.x * kBlock) {
        const double x = pos[3 * (size_t)m], y = pos[3 * (size_t)m + 1], z = pos[3 * (size_t)m + 2];
        if (!finite3(x, y, z)) continue;
        v[0] = fmin(v[0], x);
        v[1] = fmin(v[1], y);
        v[2] = fmin(v[2], z);
        v[3] = fmax(v[3], x);
        v[4] = fmax(v[4], y);
        v[5] = fmax(v[5], z);
    }
    block_minmax6(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = v[k];
}

__global__ __launch_bounds__(kBlock) void bbox_final_kernel(const double *__restrict__ part, int n,
                                                            double *__restrict__ out) {
    __shared__ double sh[6 * kBlock];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < n; b += kBlock)
        for (int k = 0; k < 6; ++k)
            v[k] = k < 3 ? fmin(v[k], part[6 * b + k]) : fmax(v[k], part[6 * b + k]);
    block_minmax6(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) out[k] = v[k];
}

// Interleave the low 21 bits of three integers (3-D Morton code).
__device__ __forceinline__ unsigned long long spread21(unsigned long long v) {
    v &= 0x1fffffULL;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

// key = Morton code of ((v - lo) * qscale) clamped to [0, 2^21 - 1] and
// truncated; non-finite pixels get the largest key (they sort last).
__global__ __launch_bounds__(kBlock) void morton_kernel(const double *__restrict__ pos, int M,
                                                        double lo0, double lo1, double lo2,
                                                        double qscale, double f0, double f1,
                                                        double f2, int a0, int a1, int a2,
                                                        unsigned long long *__restrict__ keys,
                                                        int *__restrict__ idx) {
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= M) return;
    const double v[3] = {pos[3 * (size_t)m], pos[3 * (size_t)m + 1], pos[3 * (size_t)m + 2]};
    const double lo[3] = {lo0, lo1, lo2};
    unsigned long long code = ~0ULL;
    if (finite3(v[0], v[1], v[2])) {
        unsigned long long q[3];
        const double f[3] = {f0, f1, f2};
        for (int k = 0; k < 3; ++k) {
            double t = (v[k] - lo[k]) * qscale * f[k];
            t = fmin(fmax(t, 0.0), (double)((1 << 21) - 1));
            q[k] = (unsigned long long)t;
        }
        code = spread21(q[a0]) | (spread21(q[a1]) << 1) | (spread21(q[a2]) << 2);
    }
    keys[m] = code;
    idx[m] = m;
}

__global__ __launch_bounds__(kBlock) void gather_sorted_kernel(const double *__restrict__ pos,
                                                               const int *__restrict__ perm, int M,
                                                               double *__restrict__ sx,
                                                               double *__restrict__ sy,
                                                               double *__restrict__ sz) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M) return;
    const size_t m = (size_t)perm[i];
    sx[i] = pos[3 * m];
    sy[i] = pos[3 * m + 1];
    sz[i] = pos[3 * m + 2];
}

// One workgroup per tile [t len, min(M, (t + 1) len)).  Centre = bounding-box
// midpoint of the finite pixels, rounded to a multiple of 2^-10 m when `snap`
// (pixel offsets from it are then exact in FP64); radius = largest distance to
// the centre, inflated by 1e-9 relative + 1e-9 m; half-extents likewise.
// planar = 1 when every pixel's z offset from the centre is exactly zero
// (non-finite pixels count as the centre).  Tiles without a finite pixel get
// zeros and planar = 1.
__global__ __launch_bounds__(kBlock) void tile_stats_kernel(const double *__restrict__ sx,
                                                            const double *__restrict__ sy,
                                                            const double *__restrict__ sz, int M,
                                                            int len, int snap,
                                                            double4 *__restrict__ centre_r,
                                                            double4 *__restrict__ half_planar) {
    __shared__ double sh[6 * kBlock];
    __shared__ double c_sh[3];
    const int t = blockIdx.x;
    const long long a = (long long)t * len;
    const long long b = a + len < M ? a + len : M;
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (long long i = a + threadIdx.x; i < b; i += kBlock) {
        const double x = sx[i], y = sy[i], z = sz[i];
        if (!finite3(x, y, z)) continue;
        v[0] = fmin(v[0], x);
        v[1] = fmin(v[1], y);
        v[2] = fmin(v[2], z);
        v[3] = fmax(v[3], x);
        v[4] = fmax(v[4], y);
        v[5] = fmax(v[5], z);
    }
    block_minmax6(v, sh);
    const bool any = v[0] <= v[3];
    if (!any) {
        if (threadIdx.x == 0) {
            centre_r[t] = make_double4(0.0, 0.0, 0.0, 0.0);
            half_planar[t] = make_double4(0.0, 0.0, 0.0, 1.0);
        }
        return;
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) {
            double c = 0.5 * (v[k] + v[k + 3]);
            if (snap) c = ldexp(rint(ldexp(c, 10)), -10);
            c_sh[k] = c;
        }
    }
    __syncthreads();
    const double cx = c_sh[0], cy = c_sh[1], cz = c_sh[2];
    double r = 0.0;
    for (long long i = a + threadIdx.x; i < b; i += kBlock) {
        const double dx = sx[i] - cx, dy = sy[i] - cy, dz = sz[i] - cz;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (isfinite(d)) r = fmax(r, d);
    }
    double w[6] = {0.0, 0.0, 0.0, r, 0.0, 0.0};
    block_minmax6(w, sh);
    if (threadIdx.x == 0) {
        const double rad = w[3] * (1.0 + 1e-9) + 1e-9;
        double h[3];
        const double c[3] = {cx, cy, cz};
        for (int k = 0; k < 3; ++k)
            h[k] = fmax(v[k + 3] - c[k], c[k] - v[k]) * (1.0 + 1e-9) + 1e-9;
        const bool planar = v[2] == cz && v[5] == cz;
        centre_r[t] = make_double4(cx, cy, cz, rad);
        half_planar[t] = make_double4(h[0], h[1], h[2], planar ? 1.0 : 0.0);
    }
}

__global__ __launch_bounds__(kBlock) void pixrec_kernel(const double *__restrict__ sx,
                                                        const double *__restrict__ sy,
                                                        const double *__restrict__ sz, int M,
                                                        const double4 *__restrict__ tiles,
                                                        PixRec *__restrict__ recs) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M) return;
    const double4 c = tiles[i / kTilePixels];
    const double x = sx[i], y = sy[i], z = sz[i];
    const bool fin = finite3(x, y, z);
    PixRec r;
    r.x = fin ? x : c.x;
    r.y = fin ? y : c.y;
    r.z = fin ? z : c.z;
    r.dx = r.x - c.x;
    r.dy = r.y - c.y;
    r.dz = r.z - c.z;
    r.d2 = fma(r.dx, r.dx, fma(r.dy, r.dy, r.dz * r.dz));
    r.ok = fin ? 1.0 : 0.0;
    recs[i] = r;
}

// out[m] = exp(-i sgn ((pos[m] - spot) . disp)), complex128 interleaved; the
// dot product as an FMA chain in k order, the phase reduced by FP64 sin/cos.
__global__ __launch_bounds__(kBlock) void baseband_kernel(const double *__restrict__ pos, int M,
                                                          double s0, double s1, double s2,
                                                          double d0, double d1, double d2,
                                                          double sgn, double2 *__restrict__ out) {
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= M) return;
    const double r0 = pos[3 * (size_t)m] - s0, r1 = pos[3 * (size_t)m + 1] - s1,
                 r2 = pos[3 * (size_t)m + 2] - s2;
    const double ph = fma(r2, d2, fma(r1, d1, r0 * d0));
    const double th = -sgn * ph;
    double sn, cs;
    sincos(th, &sn, &cs);
    out[m] = make_double2(cs, sn);
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_grid_positions(const GridSpec &g, double *pos, hipStream_t s) {
    const long long M = (long long)g.nz * g.ny * g.nx;
    if (M <= 0) return hipSuccess;
    grid_positions_kernel<<<blocks_for(M), kBlock, 0, s>>>(g, pos);
    return hipGetLastError();
}

hipError_t launch_bbox(const double *pos, int M, double *part, int part_blocks, double *out,
                       hipStream_t s) {
    const int nb = (int)std::min<long long>(part_blocks, std::max<long long>(1, blocks_for(M)));
    bbox_partial_kernel<<<nb, kBlock, 0, s>>>(pos, M, part);
    bbox_final_kernel<<<1, kBlock, 0, s>>>(part, nb, out);
    return hipGetLastError();
}

hipError_t launch_morton(const double *pos, int M, const double lo[3], double qscale,
                         const double scale[3], const int axes[3], unsigned long long *keys,
                         int *idx, hipStream_t s) {
    morton_kernel<<<blocks_for(M), kBlock, 0, s>>>(pos, M, lo[0], lo[1], lo[2], qscale, scale[0],
                                                   scale[1], scale[2], axes[0], axes[1], axes[2],
                                                   keys, idx);
    return hipGetLastError();
}

hipError_t sort_pairs(void *temp, size_t &temp_bytes, const unsigned long long *keys_in,
                      unsigned long long *keys_out, const int *vals_in, int *vals_out, int M,
                      hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, vals_in,
                                              vals_out, M, 0, 64, s);
}

hipError_t launch_gather_sorted(const double *pos, const int *perm, int M, double *sx, double *sy,
                                double *sz, hipStream_t s) {
    gather_sorted_kernel<<<blocks_for(M), kBlock, 0, s>>>(pos, perm, M, sx, sy, sz);
    return hipGetLastError();
}

hipError_t launch_tile_stats(const double *sx, const double *sy, const double *sz, int M, int len,
                             bool snap, double4 *centre_r, double4 *half_planar, hipStream_t s) {
    const int nt = (int)(((long long)M + len - 1) / len);
    tile_stats_kernel<<<nt, kBlock, 0, s>>>(sx, sy, sz, M, len, snap ? 1 : 0, centre_r,
                                            half_planar);
    return hipGetLastError();
}

hipError_t launch_pixrec(const double *sx, const double *sy, const double *sz, int M,
                         const double4 *tiles, PixRec *recs, hipStream_t s) {
    pixrec_kernel<<<blocks_for(M), kBlock, 0, s>>>(sx, sy, sz, M, tiles, recs);
    return hipGetLastError();
}

hipError_t launch_baseband(const double *pos, int M, const double spot[3], const double disp[3],
                           double sgn, double2 *out, hipStream_t s) {
    if (M <= 0) return hipSuccess;
    baseband_kernel<<<blocks_for(M), kBlock, 0, s>>>(pos, M, spot[0], spot[1], spot[2], disp[0],
                                                     disp[1], disp[2], sgn, out);
    return hipGetLastError();
}

}  // namespace sarhip
