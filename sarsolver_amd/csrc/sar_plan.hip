// sar_plan.hip -- plans, rocFFT stages and the C ABI of libsarsolver_hip.so.
//
// A plan holds everything that depends only on the geometry: per-pulse
// constants, the pixel positions in tile (Morton) order, forward tiles and
// their LDS window sizes, the waveform / slow-time weights and the rocFFT
// plans.  forward / adjoint then only touch the operand buffers and one
// [P, N_w] complex64 profile buffer.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sarsolver_hip.h"
#include "sar_device.h"
#include "sar_launch.h"

using namespace sarhip;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return fail(SARHIP_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define FFT_TRY(expr)                                                                     \
    do {                                                                                  \
        rocfft_status _s = (expr);                                                        \
        if (_s != rocfft_status_success)                                                  \
            return fail(SARHIP_EFFT, std::string(#expr) + ": rocfft status " +            \
                                         std::to_string((int)_s));                        \
    } while (0)

std::once_flag g_fft_once;
rocfft_status g_fft_setup_status = rocfft_status_success;

// Host FP64 range with the reference's evaluation order (CPP:10-20).
double host_distance(const double *a, const double *x) {
    double dsq = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double d = a[k] - x[k];
        dsq += d * d;
    }
    return std::sqrt(dsq);
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

template <class T>
void dev_free(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Device buffer freed when it goes out of scope (plan set-up temporaries).
template <class T>
struct DevBuf {
    T *p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, sizeof(T) * n); }
    ~DevBuf() {
        if (p) (void)hipFree((void *)p);
    }
};

// Where a plan's pixel positions come from: a host [M][3] array, a device
// [M][3] array, or a grid generated on the device.
struct PosSource {
    enum Kind { HOST, DEVICE, GRID } kind = HOST;
    const double *ptr = nullptr;
    GridSpec grid{};
};

}  // namespace

struct sarhip_plan_s {
    int device = 0;
    size_t N = 0, P = 0, M = 0, nw = 0;
    double fc = 0, fs = 0, c_eff = 0, U = 0, sgn = 1;
    double by_dx = 0, kc = 0, kc_rev = 0, inv_nw = 0;
    int interp = 0, precision = 0, requested_precision = 0;
    bool mono = false;
    bool planar = false;
    int Pv = 0;                  // valid pulses
    int num_tiles = 0, win = 0;
    int num_regions = 0, region_tiles = 1, fwd_pulses = 32, fwd_groups = 1;
    double series_err_bins = 0;  // bound on the series truncation term (bins)
    double bp_series_err_bins = 0;  // the same for the linear kernels' series order
    int series_order = 5;           // series order of the linear kernels (3 or 5)
    size_t ld = 0;               // adjoint profile row stride (nw + padded tail)
    double shift_lo = 0;         // bins: bound on |u - u(tile centre)| plus one

    PulseConst *d_pc = nullptr;
    int *d_rows = nullptr;
    double *d_px = nullptr, *d_py = nullptr, *d_pz = nullptr;
    int *d_perm = nullptr;
    double4 *d_tile_c = nullptr, *d_region_c = nullptr, *d_region_h = nullptr;
    PixRec *d_pixrec = nullptr;
    void *d_series = nullptr;
    int fused_U = 0;                                // > 0: fused spectral kernels, N_w = U N
    bool fft_sparse = false;  // adjoint spans cover under half of the valid rows' bins
    float2 *d_fft_tables = nullptr;                 // twiddle tables of the fused kernels
    double2 *d_fft_tables64 = nullptr;              // the same in FP64 (forward: exact
                                                    // ranges, or the cancelling rows)
    int *d_fft_redo = nullptr;  // [P] forward rows redone in FP64 (adaptive spectral stage)
    double fwd_kappa = 0.0, fwd_kappa2 = 0.0;  // > 0: adaptive forward spectral stage
                                               // (threshold on kappa, its square)
    double *d_bp_cu = nullptr;  // [num_tiles][Pv] backprojection row offsets of tile centres
    SeriesGather *d_series_g = nullptr;  // [num_tiles][Pv] gather records (see SeriesGather)
    float4 *d_bp_pairs = nullptr;  // [P][ld] demodulated adjoint row pairs (gather pair path)
    // (-1)^rho when rho is an integer (the adjoint spectral kernel then forms
    // the demodulated rows itself), else 0 (pair_rows_kernel after it)
    int pairs_sgn = 0;
    ScatRec *d_srec = nullptr;  // [M] planar scatter records (64-pulse linear series / hilo layout)
    float2 *d_W = nullptr, *d_Wconj = nullptr, *d_wfac_fwd = nullptr, *d_wfac_adj = nullptr;
    float2 *d_prof = nullptr;
    float2 *d_amps_perm = nullptr;
    int2 *d_partial = nullptr;
    double *d_wbases = nullptr;
    float *d_wlive = nullptr;
    int2 *d_fspan = nullptr;  // [P] forward row spans written by the merge
    int *d_fpoison = nullptr;  // [P] rows the scatter could not window (cleared by the merge)
    int2 *d_aspan = nullptr;  // [P] adjoint row spans the backprojection reads (fused path)
    float *d_region_l1 = nullptr;
    int *d_region_cmax = nullptr;  // [num_regions] see launch_region_count (or NULL)
    double2 *d_bp_partial = nullptr;
    int bp_splits = 1;

    rocfft_plan fft_fwd = nullptr, fft_bwd = nullptr;
    rocfft_plan fft_fwd64 = nullptr;  // exact ranges requested: the forward in FP64
    double2 *d_prof64 = nullptr;      // [P][N_w] its rows
    rocfft_execution_info fft_info = nullptr;
    void *fft_work = nullptr;
    size_t fft_work_size = 0;

    // Stage timing: one set of 4 events per call while enabled, read back lazily.
    bool timing = false;
    std::vector<hipEvent_t> ev_pool[2];  // [0] forward, [1] adjoint
    size_t ev_used[2] = {0, 0};
    hipEvent_t *ev = nullptr;            // the current call's 4 events

    ~sarhip_plan_s() {
        DeviceGuard g(device);
        dev_free(d_pc);
        dev_free(d_rows);
        dev_free(d_px);
        dev_free(d_py);
        dev_free(d_pz);
        dev_free(d_perm);
        dev_free(d_tile_c);
        dev_free(d_region_c);
        dev_free(d_region_h);
        dev_free(d_pixrec);
        dev_free(d_series);
        dev_free(d_bp_cu);
        dev_free(d_series_g);
        dev_free(d_bp_pairs);
        dev_free(d_srec);
        dev_free(d_fft_tables);
        dev_free(d_fft_tables64);
        dev_free(d_fft_redo);
        dev_free(d_W);
        dev_free(d_Wconj);
        dev_free(d_wfac_fwd);
        dev_free(d_wfac_adj);
        dev_free(d_prof);
        dev_free(d_amps_perm);
        dev_free(d_partial);
        dev_free(d_wbases);
        dev_free(d_wlive);
        dev_free(d_fspan);
        dev_free(d_fpoison);
        dev_free(d_aspan);
        dev_free(d_region_l1);
        dev_free(d_region_cmax);
        dev_free(d_bp_partial);
        if (fft_work) (void)hipFree(fft_work);
        if (fft_fwd) rocfft_plan_destroy(fft_fwd);
        if (fft_bwd) rocfft_plan_destroy(fft_bwd);
        if (fft_fwd64) rocfft_plan_destroy(fft_fwd64);
        dev_free(d_prof64);
        if (fft_info) rocfft_execution_info_destroy(fft_info);
        for (auto &pool : ev_pool)
            for (auto &e : pool)
                if (e) (void)hipEventDestroy(e);
    }

    KernelParams params() const {
        KernelParams k;
        k.px = d_px;
        k.py = d_py;
        k.pz = d_pz;
        k.perm = d_perm;
        k.M = (int)M;
        k.num_tiles = num_tiles;
        k.tile_c = d_tile_c;
        k.pixrec = d_pixrec;
        k.srec = d_srec;
        k.pc = d_pc;
        k.rows = d_rows;
        k.Pv = Pv;
        k.series = d_series;
        k.bp_cu = d_bp_cu;
        k.series_g = d_series_g;
        k.bp_pairs = d_bp_pairs;
        k.series_order = series_order;
        k.prof = d_prof;
        k.nw = (int)nw;
        k.ld = (int)ld;
        k.shift_lo = shift_lo;
        k.nwd = (double)nw;
        k.inv_nw = inv_nw;
        k.by_dx = by_dx;
        k.kc_rev = kc_rev;
        k.rho = kc_rev / by_dx;
        k.out = nullptr;
        k.amps_perm = d_amps_perm;
        k.num_regions = num_regions;
        k.region_tiles = region_tiles;
        k.region_c = d_region_c;
        k.region_h = d_region_h;
        k.fwd_pulses = fwd_pulses;
        k.fwd_groups = fwd_groups;
        k.win = win;
        k.win_stride = fwd_window_stride(win);
        k.partial = d_partial;
        k.wbases = d_wbases;
        k.wlive = d_wlive;
        k.fspan = d_fspan;
        k.fpoison = d_fpoison;
        k.fwd_spans = 0;
        k.region_l1 = d_region_l1;
        k.region_cmax = d_region_cmax;
        k.bp_splits = bp_splits;
        k.bp_partial = d_bp_partial;
        k.interp = interp;
        k.precision = precision;
        k.mono = mono;
        k.planar = planar;
        return k;
    }
};

namespace {

int run_fft(sarhip_plan_t pl, rocfft_plan fp, void *buf, hipStream_t s) {
    FFT_TRY(rocfft_execution_info_set_stream(pl->fft_info, s));
    void *bufs[1] = {buf};
    FFT_TRY(rocfft_execute(fp, bufs, nullptr, pl->fft_info));
    return SARHIP_OK;
}

constexpr size_t kMaxTimedCalls = 4096;
// Events per timed call: stage boundaries 0..3, then 4 / 5 around the main
// kernel's launch (scatter or gather alone).
constexpr size_t kEventsPerCall = 6;

// Reserve the next event set of `which` (0 forward, 1 adjoint) for this call.
int begin_timing(sarhip_plan_t pl, int which) {
    pl->ev = nullptr;
    if (!pl->timing || pl->ev_used[which] >= kMaxTimedCalls) return SARHIP_OK;
    auto &pool = pl->ev_pool[which];
    const size_t base = kEventsPerCall * pl->ev_used[which];
    while (pool.size() < base + kEventsPerCall) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreate(&e));
        pool.push_back(e);
    }
    pl->ev = pool.data() + base;
    pl->ev_used[which] += 1;
    return SARHIP_OK;
}

int record(sarhip_plan_t pl, int i, hipStream_t s) {
    if (pl->ev) HIP_TRY(hipEventRecord(pl->ev[i], s));
    return SARHIP_OK;
}

// max over i < n of f(i) (>= 0; a non-finite value gives INFINITY), on up to
// 16 host threads; max is exact, so the result does not depend on the split.
template <class F>
double parallel_max(size_t n, F f) {
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t T = std::min(hw, std::max<size_t>(1, n / 8));
    std::vector<double> part(T, 0.0);
    auto run = [&](size_t t) {
        double w = 0.0;
        for (size_t i = t; i < n; i += T) {
            const double v = f(i);
            w = std::isfinite(v) ? std::max(w, v) : INFINITY;
            if (!std::isfinite(w)) break;
        }
        part[t] = w;
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < T; ++t) pool.emplace_back(run, t);
    run(0);
    for (std::thread &th : pool) th.join();
    double w = 0.0;
    for (const double v : part) w = std::max(w, v);
    return w;
}

// Forward LDS window length (bins) for regions with bounding boxes `reg` /
// half-extents `half`: per leg |T - X| - |T - Xc| lies in
// [-s, s + r^2 / (2 (A - r))], s = sum_i |e_i| h_i, so the bins of one
// (region, pulse) span at most y = (2 S + Q) by_dx.  The scatter's window
// base sits 2 bins below the lowest bin (floored), so a pixel's relative u
// lies in [2, y + 3 + e), e the range rounding (far below one bin in every
// precision mode); the linear rule touches floor(u) + 1 <= ceil(y) + 4 and
// the nearest rules round(u + 1/2) <= ceil(y) + 4, hence ceil(y) + 5 bins.
// A contribution outside its window poisons the row (it is never dropped).
// Returns INFINITY if a leg is inside a region's bounding sphere.
double window_bins(const std::vector<double4> &reg, const std::vector<double4> &half,
                   const std::vector<PulseConst> &pc, const std::vector<int> &rows, bool mono,
                   double by_dx) {
    const double worst = parallel_max(reg.size(), [&](size_t t) {
        const double4 &c = reg[t];
        const double4 &h = half[t];
        double w = 0.0;
        for (int p : rows) {
            double S = 0.0, Q = 0.0;
            for (int leg = 0; leg < 2; ++leg) {
                const double *T = leg == 0 ? pc[p].tx : pc[p].rx;
                const double ax = T[0] - c.x, ay = T[1] - c.y, az = T[2] - c.z;
                const double A = std::sqrt(ax * ax + ay * ay + az * az);
                if (!(A > c.w)) return (double)INFINITY;
                S += (std::fabs(ax) * h.x + std::fabs(ay) * h.y + std::fabs(az) * h.z) / A;
                Q += c.w * c.w / (2.0 * (A - c.w));
                if (mono) {
                    S *= 2.0;
                    Q *= 2.0;
                    break;
                }
            }
            w = std::max(w, (2.0 * S + Q) * by_dx);
        }
        return w;
    });
    return std::ceil(worst * (1.0 + 1e-9)) + 5.0;
}

// Bound (bins) on the truncation of sqrt(1 + t) - 1 after t^order over every
// (tile, pulse): |t| <= (2 s + r^2) / A^2 for a tile of radius r and extent s
// projected on the leg at range A, and the remainder is at most the first
// omitted term / (1 - |t|) (the terms
// alternate for t > 0 and shrink geometrically for t < 0).
double series_error_bound(const std::vector<double4> &tiles, const std::vector<double4> &half,
                          const std::vector<PulseConst> &pc, const std::vector<int> &rows,
                          bool mono, double by_dx, int order) {
    const double coef = order == 3 ? 5.0 / 128.0 : 21.0 / 1024.0;
    return parallel_max(tiles.size(), [&](size_t i) {
        const double4 &tl = tiles[i];
        const double4 &h = half[i];
        const double r = tl.w;
        double w = 0.0;
        for (int p : rows) {
            double legs = 0.0;
            for (int leg = 0; leg < (mono ? 1 : 2); ++leg) {
                const double *T = leg == 0 ? pc[p].tx : pc[p].rx;
                const double dx = T[0] - tl.x, dy = T[1] - tl.y, dz = T[2] - tl.z;
                const double A = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (!(A > 2.0 * r)) return (double)INFINITY;
                // |b.d| = 2 |a.d| / A^2 <= 2 sum_i |a_i| h_i / A^2 (|d_i| <= h_i) and
                // c |d|^2 <= r^2 / A^2: the tile's extent projected on the leg
                const double s = std::fabs(dx) * h.x + std::fabs(dy) * h.y + std::fabs(dz) * h.z;
                const double t = (2.0 * s + r * r) / (A * A);
                if (!(t < 0.5)) return (double)INFINITY;
                legs += coef * std::pow(t, order + 1) / (1.0 - t) * A * (mono ? 2.0 : 1.0);
            }
            w = std::max(w, legs * by_dx);
        }
        return w;
    });
}

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return std::atoi(v);
}

double env_double(const char *name, double dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return std::atof(v);
}

}  // namespace

extern "C" {

const char *sarhip_last_error(void) { return g_last_error.c_str(); }

// Dataset resampling (numpy/measurement.py:342-457): see resample_kernels.hip.
// Host arrays: dr [P] (new minus old SRP range per input pulse), win [N2]
// (fftshift(tukey) / N), kern [K] complex128 interleaved, starts [P2] (first
// input pulse of each output pulse).  Device arrays: ph_in [P][N], ph_out
// [P2][N2], complex64, centred frequency order.  Work buffers and FFT plans
// are kept per device for the last (P, N, N2, P2) seen, so repeated calls of one
// shape allocate nothing; the call returns after the stream has finished.
namespace {
struct ResampleCache {
    size_t P = 0, N = 0, N2 = 0, P2 = 0;
    float2 *prof = nullptr;
    void *work = nullptr;
    size_t work_size = 0;
    rocfft_plan inv = nullptr, fwd = nullptr;
    rocfft_execution_info info = nullptr;
    void release() {
        dev_free(prof);
        if (work) (void)hipFree(work);
        work = nullptr;
        work_size = 0;
        if (inv) rocfft_plan_destroy(inv);
        if (fwd) rocfft_plan_destroy(fwd);
        if (info) rocfft_execution_info_destroy(info);
        inv = fwd = nullptr;
        info = nullptr;
        P = N = N2 = P2 = 0;
    }
};
std::mutex g_rs_mutex;
ResampleCache g_rs_cache[64];  // per device index

// Inverse FFT of length N over the P input rows (in place on the profile
// buffer); forward FFT of length N2 over the P2 output rows (in place on the
// caller's output).
int resample_setup(ResampleCache &c, size_t P, size_t N, size_t N2, size_t P2) {
    if (c.P == P && c.N == N && c.N2 == N2 && c.P2 == P2 && c.prof) return SARHIP_OK;
    c.release();
    HIP_TRY(hipMalloc((void **)&c.prof, sizeof(float2) * P * N));
    FFT_TRY(rocfft_plan_create(&c.inv, rocfft_placement_inplace,
                               rocfft_transform_type_complex_inverse, rocfft_precision_single, 1, &N,
                               P, nullptr));
    size_t w1 = 0, w2 = 0;
    FFT_TRY(rocfft_plan_get_work_buffer_size(c.inv, &w1));
    if (P2) {
        FFT_TRY(rocfft_plan_create(&c.fwd, rocfft_placement_inplace,
                                   rocfft_transform_type_complex_forward, rocfft_precision_single,
                                   1, &N2, P2, nullptr));
        FFT_TRY(rocfft_plan_get_work_buffer_size(c.fwd, &w2));
    }
    FFT_TRY(rocfft_execution_info_create(&c.info));
    c.work_size = std::max(w1, w2);
    if (c.work_size) {
        HIP_TRY(hipMalloc(&c.work, c.work_size));
        FFT_TRY(rocfft_execution_info_set_work_buffer(c.info, c.work, c.work_size));
    }
    c.P = P;
    c.N = N;
    c.N2 = N2;
    c.P2 = P2;
    return SARHIP_OK;
}
}  // namespace

int sarhip_resample(const void *ph_in, size_t P, size_t N, const double *dr, double centre_freq,
                    double sample_freq, double sign_multiplier, double c_eff, size_t N2,
                    const double *win, const double *kern, int K, const int *starts, size_t P2,
                    void *ph_out, void *stream) {
    if (!ph_in || !ph_out || !dr || !win || !kern || !starts || P == 0 || N == 0 || N2 == 0 ||
        N2 > N || K <= 0 || (size_t)K > P || P > (size_t)INT32_MAX || N > (size_t)INT32_MAX ||
        P2 > (size_t)INT32_MAX || !(c_eff > 0.0))
        return fail(SARHIP_EINVAL, "sarhip_resample: bad argument");
    for (size_t i = 0; i < P2; ++i)
        if (starts[i] < 0 || (size_t)starts[i] + (size_t)K > P)
            return fail(SARHIP_EINVAL, "sarhip_resample: window outside the input pulses");
    std::call_once(g_fft_once, [] { g_fft_setup_status = rocfft_setup(); });
    if (g_fft_setup_status != rocfft_status_success) return fail(SARHIP_EFFT, "rocfft_setup failed");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(SARHIP_ENODEV, "device index out of range");
    std::lock_guard<std::mutex> lock(g_rs_mutex);
    ResampleCache &c = g_rs_cache[dev];
    int rc = resample_setup(c, P, N, N2, P2);
    if (rc) {
        c.release();
        return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    struct Small {
        double *dr = nullptr;
        float *win = nullptr;
        float2 *kern = nullptr;
        int *starts = nullptr;
        ~Small() {
            dev_free(dr);
            dev_free(win);
            dev_free(kern);
            dev_free(starts);
        }
    } b;
    std::vector<float> wf(N2);
    for (size_t j = 0; j < N2; ++j) wf[j] = (float)win[j];
    std::vector<float2> kf(K);
    for (int k = 0; k < K; ++k) kf[k] = make_float2((float)kern[2 * k], (float)kern[2 * k + 1]);
    HIP_TRY(hipMalloc((void **)&b.dr, sizeof(double) * P));
    HIP_TRY(hipMalloc((void **)&b.win, sizeof(float) * N2));
    HIP_TRY(hipMalloc((void **)&b.kern, sizeof(float2) * K));
    HIP_TRY(hipMalloc((void **)&b.starts, sizeof(int) * std::max<size_t>(P2, 1)));
    HIP_TRY(hipMemcpyAsync(b.dr, dr, sizeof(double) * P, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b.win, wf.data(), sizeof(float) * N2, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b.kern, kf.data(), sizeof(float2) * K, hipMemcpyHostToDevice, s));
    if (P2) HIP_TRY(hipMemcpyAsync(b.starts, starts, sizeof(int) * P2, hipMemcpyHostToDevice, s));
    FFT_TRY(rocfft_execution_info_set_stream(c.info, s));
    HIP_TRY(launch_rs_ramp((const float2 *)ph_in, (int)P, (int)N, b.dr,
                           sample_freq / (double)N, sign_multiplier * centre_freq, 1.0 / c_eff,
                           c.prof, s));
    void *bp[1] = {c.prof};
    FFT_TRY(rocfft_execute(c.inv, bp, nullptr, c.info));
    if (P2) {
        HIP_TRY(launch_rs_window(c.prof, (int)N, (int)N2, b.win, b.kern, K, b.starts, (int)P2,
                                 (float2 *)ph_out, s));
        void *bo[1] = {ph_out};
        FFT_TRY(rocfft_execute(c.fwd, bo, nullptr, c.info));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return SARHIP_OK;
}

const char *sarhip_version(void) { return "sarsolver_hip 0.1.0 (gfx950)"; }

namespace {
int plan_create_impl(sarhip_plan_t *out, size_t num_fast_times, size_t num_slow_times,
                     size_t num_scatterers, const double *tx, const double *rx, const double *srp,
                     const PosSource &src, const double *waveform_fft,
                     const double *slow_time_weighting, double centre_frequency,
                     double sample_frequency, double c_eff, double upsample_ratio,
                     double sign_multiplier, int interp, int precision, int device) {
    if (!out) return fail(SARHIP_EINVAL, "plan output pointer is NULL");
    *out = nullptr;
    if (interp < 0 || interp > 2) return fail(SARHIP_EINVAL, "interp must be 0, 1 or 2");
    if (precision < SARHIP_PRECISION_FP64 || precision > SARHIP_PRECISION_FP32)
        return fail(SARHIP_EINVAL, "precision must be 0 (fp64), 1 (series), 2 (hilo) or 3 (fp32)");
    const size_t N = num_fast_times, P = num_slow_times, M = num_scatterers;
    if (N == 0 || P == 0 || M == 0) return fail(SARHIP_EINVAL, "N, P and M must be positive");
    if (!tx || !rx || !srp || (src.kind != PosSource::GRID && !src.ptr) || !waveform_fft ||
        !slow_time_weighting)
        return fail(SARHIP_EINVAL, "NULL geometry pointer");
    if (!(upsample_ratio > 0.0) || !std::isfinite(upsample_ratio))
        return fail(SARHIP_EINVAL, "upsample_ratio must be positive and finite");
    // CPP:223
    const size_t nw = (size_t)(upsample_ratio * (double)N);
    if (nw < N) return fail(SARHIP_EINVAL, "upsample_ratio * N must be >= N");
    if (nw > (size_t)1 << 30 || M > (size_t)INT32_MAX || P > (size_t)INT32_MAX ||
        N > (size_t)INT32_MAX)
        return fail(SARHIP_EINVAL, "sizes exceed the 32-bit index range of the kernels");
    if (!(c_eff > 0.0) || !(sample_frequency > 0.0))
        return fail(SARHIP_EINVAL, "c_eff and sample_frequency must be positive");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(SARHIP_ENODEV, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(SARHIP_ENODEV, "device index out of range");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SARHIP_ENODEV, std::string("library is built for gfx950, device is ") +
                                       prop.gcnArchName);
    DeviceGuard guard(device);

    std::call_once(g_fft_once, [] { g_fft_setup_status = rocfft_setup(); });
    if (g_fft_setup_status != rocfft_status_success) return fail(SARHIP_EFFT, "rocfft_setup failed");

    sarhip_plan_s *pl = new (std::nothrow) sarhip_plan_s();
    if (!pl) return fail(SARHIP_ENOMEM, "plan allocation failed");
    struct Owner {
        sarhip_plan_s *p;
        ~Owner() { delete p; }
    } owner{pl};

    pl->device = device;
    pl->N = N;
    pl->P = P;
    pl->M = M;
    pl->nw = nw;
    pl->fc = centre_frequency;
    pl->fs = sample_frequency;
    pl->c_eff = c_eff;
    pl->U = upsample_ratio;
    pl->sgn = sign_multiplier;
    pl->interp = interp;
    pl->precision = precision;
    // CPP:225-228
    pl->by_dx = (sample_frequency * (double)nw) / ((double)N * c_eff);
    pl->kc = 2.0 * sign_multiplier * M_PI * centre_frequency / c_eff;
    pl->kc_rev = sign_multiplier * centre_frequency / c_eff;  // k_c / (2 pi), revolutions per metre
    pl->inv_nw = 1.0 / (double)nw;

    // Per-pulse constants; pulses with non-finite geometry are skipped.
    std::vector<PulseConst> pc(P);
    std::vector<int> rows;
    bool mono = true;
    for (size_t p = 0; p < P; ++p) {
        PulseConst &c = pc[p];
        bool fin = true;
        for (int k = 0; k < 3; ++k) {
            c.tx[k] = tx[3 * p + k];
            c.rx[k] = rx[3 * p + k];
            fin = fin && std::isfinite(c.tx[k]) && std::isfinite(c.rx[k]) &&
                  std::isfinite(srp[3 * p + k]);
        }
        c.r_srp = host_distance(c.tx, srp + 3 * p) + host_distance(c.rx, srp + 3 * p);
        c.valid = (fin && std::isfinite(c.r_srp)) ? 1.0 : 0.0;
        if (c.valid != 0.0) {
            rows.push_back((int)p);
            if (std::memcmp(c.tx, c.rx, sizeof(c.tx)) != 0) mono = false;
        }
    }
    pl->mono = mono;
    pl->Pv = (int)rows.size();
    if (rows.empty()) rows.push_back(0);  // keeps device arrays non-empty; Pv stays 0

    // Pixel positions on the device, [M][3] FP64: uploaded from the host, the
    // caller's device array, or generated from a grid description.
    DevBuf<double> pos_buf;
    const double *d_pos = nullptr;
    if (src.kind == PosSource::HOST) {
        HIP_TRY(pos_buf.alloc(3 * M));
        HIP_TRY(hipMemcpy(pos_buf.p, src.ptr, sizeof(double) * 3 * M, hipMemcpyHostToDevice));
        d_pos = pos_buf.p;
    } else if (src.kind == PosSource::DEVICE) {
        d_pos = src.ptr;
    } else {
        HIP_TRY(pos_buf.alloc(3 * M));
        HIP_TRY(launch_grid_positions(src.grid, pos_buf.p, nullptr));
        d_pos = pos_buf.p;
    }

    // Tile (Morton) order: bounding box of the finite pixels, a 21-bit
    // quantisation per axis, and a stable radix sort of (key, index), which
    // orders the pixels as sorting the pairs would.
    double lo[3], hi[3];
    {
        DevBuf<double> part, bb;
        HIP_TRY(part.alloc(6 * 1024));
        HIP_TRY(bb.alloc(6));
        HIP_TRY(launch_bbox(d_pos, (int)M, part.p, 1024, bb.p, nullptr));
        double h[6];
        HIP_TRY(hipMemcpy(h, bb.p, sizeof(h), hipMemcpyDeviceToHost));
        for (int k = 0; k < 3; ++k) {
            lo[k] = h[k];
            hi[k] = h[k + 3];
        }
    }
    double span = 0.0;
    for (int k = 0; k < 3; ++k)
        if (hi[k] > lo[k]) span = std::max(span, hi[k] - lo[k]);
    const double qscale = span > 0.0 ? (double)((1 << 21) - 1) / span : 0.0;
    // Morton axis order: the lowest bits go to the axis along which the
    // platform is seen least (mean |e_k| over valid pulses and legs, e the
    // unit vector from the scene centre to the phase centre).  Tiles stay
    // compact either way; pairs of consecutive tiles (2-tile forward
    // regions) then extend across range, where a region's projected range
    // extent, and so its LDS window, grows least.
    int axes[3] = {0, 1, 2};
    double qs_axis[3] = {1.0, 1.0, 1.0};
    {
        double m[3] = {0.0, 0.0, 0.0};
        const double ctr[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
        if (pl->Pv > 0 && std::isfinite(ctr[0] + ctr[1] + ctr[2])) {
            for (int p : rows) {
                for (int leg = 0; leg < 2; ++leg) {
                    const double *T = leg == 0 ? pc[p].tx : pc[p].rx;
                    const double d[3] = {T[0] - ctr[0], T[1] - ctr[1], T[2] - ctr[2]};
                    const double A = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    if (A > 0.0)
                        for (int k = 0; k < 3; ++k) m[k] += std::fabs(d[k]) / A;
                }
            }
            std::stable_sort(axes, axes + 3, [&](int a, int b) { return m[a] < m[b]; });
            // strongly directional looks: the cross-range axis is quantised 4x
            // coarser, so tiles are 4x longer across range than along it
            // (C2: 16 x 64 pixels; windows of 2-tile regions 63 bins, against
            // 101 for square tiles)
            const double f = env_double("SARHIP_TILE_ASPECT", 0.25);
            if (m[axes[0]] <= 0.25 * m[axes[2]] && f > 0.0 && f <= 1.0) qs_axis[axes[0]] = f;
        }
    }
    {
        DevBuf<unsigned long long> keys_in, keys_out;
        DevBuf<int> idx_in;
        HIP_TRY(keys_in.alloc(M));
        HIP_TRY(keys_out.alloc(M));
        HIP_TRY(idx_in.alloc(M));
        HIP_TRY(hipMalloc((void **)&pl->d_perm, sizeof(int) * M));
        HIP_TRY(launch_morton(d_pos, (int)M, lo, qscale, qs_axis, axes, keys_in.p, idx_in.p,
                              nullptr));
        size_t temp_bytes = 0;
        HIP_TRY(sort_pairs(nullptr, temp_bytes, keys_in.p, keys_out.p, idx_in.p, pl->d_perm,
                           (int)M, nullptr));
        DevBuf<unsigned char> temp;
        HIP_TRY(temp.alloc(std::max<size_t>(temp_bytes, 1)));
        HIP_TRY(sort_pairs(temp.p, temp_bytes, keys_in.p, keys_out.p, idx_in.p, pl->d_perm,
                           (int)M, nullptr));
        HIP_TRY(hipStreamSynchronize(nullptr));  // before the key buffers are freed
    }
    HIP_TRY(hipMalloc((void **)&pl->d_px, sizeof(double) * M));
    HIP_TRY(hipMalloc((void **)&pl->d_py, sizeof(double) * M));
    HIP_TRY(hipMalloc((void **)&pl->d_pz, sizeof(double) * M));
    HIP_TRY(launch_gather_sorted(d_pos, pl->d_perm, (int)M, pl->d_px, pl->d_py, pl->d_pz, nullptr));

    // Tiles of `len` consecutive pixels (centre, radius; half-extents,
    // planarity), computed on the device and copied back (O(tiles)).
    auto tile_stats = [&](size_t len, bool snap, std::vector<double4> &c,
                          std::vector<double4> &h) -> int {
        const size_t nt = (M + len - 1) / len;
        DevBuf<double4> dc, dh;
        HIP_TRY(dc.alloc(nt));
        HIP_TRY(dh.alloc(nt));
        HIP_TRY(launch_tile_stats(pl->d_px, pl->d_py, pl->d_pz, (int)M, (int)len, snap, dc.p,
                                  dh.p, nullptr));
        c.resize(nt);
        h.resize(nt);
        HIP_TRY(hipMemcpy(c.data(), dc.p, sizeof(double4) * nt, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(h.data(), dh.p, sizeof(double4) * nt, hipMemcpyDeviceToHost));
        return SARHIP_OK;
    };

    // Series tiles (also the forward workgroup tiles) and per-pixel records.
    std::vector<double4> tiles, tile_half;
    int rc;
    if ((rc = tile_stats(kTilePixels, true, tiles, tile_half))) return rc;
    double rad_t = 0.0;
    pl->planar = true;
    for (size_t t = 0; t < tiles.size(); ++t) {
        rad_t = std::max(rad_t, tiles[t].w);
        pl->planar = pl->planar && tile_half[t].w != 0.0;
    }
    pl->num_tiles = (int)tiles.size();
    // Within a tile of radius r each leg's range moves by at most r (triangle
    // inequality), so |u - u(tile centre)| <= 2 r by_dx; series truncation and
    // rounding add far less than the one-bin slack.  Adjoint rows get a tail
    // of at least ceil(2 shift_lo) + 2 bins (see bp_gather_kernel).
    pl->shift_lo = 2.0 * rad_t * pl->by_dx * (1.0 + 1e-6) + 2.0;
    {
        const double tail = std::ceil(2.0 * pl->shift_lo) + 2.0;
        if (!(tail < 1e9) || (double)nw + tail > 2147483647.0 / 8.0)
            return fail(SARHIP_EINVAL, "scene tiles too large for the padded profile rows");
        // rows start on 4 KiB boundaries: the strided batched FFT is as fast
        // as the contiguous one only for aligned row starts
        pl->ld = (nw + (size_t)tail + 511) / 512 * 512;
    }
    HIP_TRY(hipMalloc((void **)&pl->d_tile_c, sizeof(double4) * tiles.size()));
    HIP_TRY(hipMemcpy(pl->d_tile_c, tiles.data(), sizeof(double4) * tiles.size(),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc((void **)&pl->d_pixrec, sizeof(PixRec) * M));
    HIP_TRY(launch_pixrec(pl->d_px, pl->d_py, pl->d_pz, (int)M, pl->d_tile_c, pl->d_pixrec,
                          nullptr));

    // Series modes need the truncation term far below the tie margin; if the
    // geometry does not allow it the series mode falls back to exact FP64.
    pl->requested_precision = precision;
    pl->precision = precision;
    // The kernels of a series plan use the third-order series when its bound
    // meets the same limits.  (The nearest rules stay exact with it: a bin
    // decision within 1e-7 bin of a tie is taken on the exact range.)
    pl->series_order = 5;
    if (precision == SARHIP_PRECISION_SERIES || precision == SARHIP_PRECISION_HILO) {
        auto ok = [&](double bins) {
            return bins <= 1e-9 && std::fabs(bins * pl->kc_rev / pl->by_dx) <= 1e-9;
        };
        pl->series_err_bins = series_error_bound(tiles, tile_half, pc, rows, mono, pl->by_dx, 5);
        if (precision == SARHIP_PRECISION_SERIES && !ok(pl->series_err_bins))
            pl->precision = SARHIP_PRECISION_FP64;
        if (pl->precision == SARHIP_PRECISION_SERIES && env_int("SARHIP_SERIES_ORDER", 3) == 3) {
            const double e3 = series_error_bound(tiles, tile_half, pc, rows, mono, pl->by_dx, 3);
            if (ok(e3)) {
                pl->series_order = 3;
                pl->bp_series_err_bins = e3;
            }
        }
    }
    if (pl->series_order == 5) pl->bp_series_err_bins = pl->series_err_bins;
    const bool need_series =
        pl->precision == SARHIP_PRECISION_SERIES || pl->precision == SARHIP_PRECISION_HILO;

    // Forward regions: R consecutive tiles share LDS windows of G pulses.
    // Each window covers every bin the region can reach for one pulse.  The
    // choice minimises the window partials written and read per call
    // (proportional to win_R / R) under a 40 KiB LDS budget, which keeps four
    // workgroups per CU resident (measured on MI355X: occupancy matters more
    // than partial traffic), preferring G = 32 pulses.  Windows that do not
    // fit at all fall back to global atomics into the profile.
    std::vector<double4> regions, region_half;
    {
        const size_t budget = (size_t)std::min(std::max(env_int("SARHIP_FP_LDS_KB", 40), 16), 64) * 1024;
        const int forced_r = env_int("SARHIP_FP_REGION", 0);
        const int forced_g = env_int("SARHIP_FP_PULSES", 0);
        double best_cost = INFINITY;
        int best_r = 1, best_g = 32, best_w = 0;
        // the merge's window tables beside its 26 KiB of static LDS
        auto merge_fits = [&](int R, int w) {
            const size_t nreg = (pl->num_tiles + (size_t)R - 1) / (size_t)R;
            return nreg <= (size_t)INT_MAX && fp_reduce_lds((int)nreg, w) <= 128 * 1024;
        };
        // 64 pulses per workgroup (512 threads, wave-uniform pixel streams)
        // whenever its windows fit in 53 KiB (three workgroups per CU; with
        // two, C2's 8-tile regions ran slower than 2-tile ones in 32 KiB);
        // otherwise 32 or 16 pulses under `budget`.  Among region sizes,
        // fewest window bins per tile.
        const size_t budget64 =
            (size_t)std::min(std::max(env_int("SARHIP_FP_LDS64_KB", 53), 16), 160) * 1024;
        for (int pass = 0; pass < 2 && !std::isfinite(best_cost); ++pass) {
            for (int R : {1, 2, 4, 8, 16}) {
                if (forced_r > 0 && R != forced_r) continue;
                // The fixed-point unit is (region L1 norm) / 2^31, so the
                // windows' rounding grows with the region's pixel count
                // (dense-image forward against the FP64 oracle: 2.2e-7 with
                // float atomics, 5e-7 / 9e-7 / 1.8e-6 / 3.6e-6 for regions of
                // 1 / 2 / 4 / 8 tiles); regions stay within 2 tiles unless
                // forced or the merge's LDS tables need fewer regions.
                if (forced_r == 0 && R > 2 && merge_fits(2, 0)) continue;
                std::vector<double4> reg, hal;
                if ((rc = tile_stats((size_t)R * kTilePixels, false, reg, hal))) return rc;
                const double wd = window_bins(reg, hal, pc, rows, mono, pl->by_dx);
                if (wd > 1e6) continue;
                const int w = (int)wd;
                // the merge stages every region's window base, scale and block
                // lists in LDS
                if (!merge_fits(R, w)) continue;
                for (int G : {64, 32, 16}) {
                    if (forced_g > 0 && G != forced_g) continue;
                    if ((pass == 0) != (G == 64)) continue;
                    const size_t lim = G == 64 ? budget64 : budget;
                    if ((size_t)G * fwd_window_stride(w) * sizeof(unsigned long long) > lim) continue;
                    const double cost = (double)w / R * (G >= 32 ? 1.0 : 4.0);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_r = R;
                        best_g = G;
                        best_w = w;
                    }
                    break;  // the larger G that fits is preferred for this R
                }
            }
        }
        if (env_int("SARHIP_FP_NO_WINDOWS", 0)) best_cost = INFINITY;
        if (!std::isfinite(best_cost)) {  // no window fits: global atomics
            best_r = 1;
            best_g = 32;
            best_w = 0;
        }
        if ((rc = tile_stats((size_t)best_r * kTilePixels, false, regions, region_half))) return rc;
        pl->region_tiles = best_r;
        pl->fwd_pulses = best_g;
        pl->fwd_groups = (std::max(pl->Pv, 1) + best_g - 1) / best_g;
        pl->win = best_w;
        pl->num_regions = (int)regions.size();
    }

    // Waveform and weights (complex64).
    std::vector<float2> W(N), Wc(N), wf(P), wa(P);
    for (size_t k = 0; k < N; ++k) {
        W[k] = make_float2((float)waveform_fft[2 * k], (float)waveform_fft[2 * k + 1]);
        Wc[k] = make_float2(W[k].x, -W[k].y);
    }
    for (size_t p = 0; p < P; ++p) {
        const float re = (float)slow_time_weighting[2 * p], im = (float)slow_time_weighting[2 * p + 1];
        wf[p] = make_float2(re, im);
        // reference adjoint multiplies by w_p unconjugated (CPP:349)
        wa[p] = interp == SARHIP_INTERP_NEAREST_REF ? make_float2(re, im) : make_float2(re, -im);
    }

    auto upload = [&](void **dst, const void *src, size_t bytes) -> int {
        HIP_TRY(hipMalloc(dst, bytes));
        HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return SARHIP_OK;
    };
    if ((rc = upload((void **)&pl->d_pc, pc.data(), sizeof(PulseConst) * P))) return rc;
    if ((rc = upload((void **)&pl->d_rows, rows.data(), sizeof(int) * rows.size()))) return rc;
    if ((rc = upload((void **)&pl->d_region_c, regions.data(), sizeof(double4) * regions.size())))
        return rc;
    if ((rc = upload((void **)&pl->d_region_h, region_half.data(),
                     sizeof(double4) * region_half.size())))
        return rc;
    if ((rc = upload((void **)&pl->d_W, W.data(), sizeof(float2) * N))) return rc;
    if ((rc = upload((void **)&pl->d_Wconj, Wc.data(), sizeof(float2) * N))) return rc;
    if ((rc = upload((void **)&pl->d_wfac_fwd, wf.data(), sizeof(float2) * P))) return rc;
    if ((rc = upload((void **)&pl->d_wfac_adj, wa.data(), sizeof(float2) * P))) return rc;
    HIP_TRY(hipMalloc((void **)&pl->d_prof, sizeof(float2) * P * pl->ld));
    HIP_TRY(hipMalloc((void **)&pl->d_amps_perm, sizeof(float2) * M));
    const size_t nrows = std::max<size_t>(1, (size_t)pl->Pv);
    HIP_TRY(hipMalloc((void **)&pl->d_wbases, sizeof(double) * nrows * pl->num_regions));
    HIP_TRY(hipMalloc((void **)&pl->d_wlive, sizeof(float) * nrows * pl->num_regions));
    if (pl->Pv > 0) HIP_TRY(launch_window_bases(pl->params(), nullptr));
    // rows the merge never writes (invalid pulses) keep the empty span
    HIP_TRY(hipMalloc((void **)&pl->d_fspan, sizeof(int2) * std::max<size_t>(P, 1)));
    HIP_TRY(hipMemset(pl->d_fspan, 0, sizeof(int2) * std::max<size_t>(P, 1)));
    HIP_TRY(hipMalloc((void **)&pl->d_fpoison, sizeof(int) * std::max<size_t>(P, 1)));
    HIP_TRY(hipMemset(pl->d_fpoison, 0, sizeof(int) * std::max<size_t>(P, 1)));
    HIP_TRY(hipMalloc((void **)&pl->d_region_l1, sizeof(float) * std::max(1, pl->num_regions)));
    // Pixels per adjacent bin pair of each region (the fixed-point bound of the
    // forward windows, see region_l1_kernel), for the modes whose ranges agree
    // with the exact FP64 range far inside the count's 1e-6-bin margin.
    if (pl->win > 0 && pl->Pv > 0 &&
        (pl->precision == SARHIP_PRECISION_FP64 || pl->precision == SARHIP_PRECISION_SERIES) &&
        !env_int("SARHIP_FP_L1_BOUND", 0)) {
        HIP_TRY(hipMalloc((void **)&pl->d_region_cmax, sizeof(int) * pl->num_regions));
        HIP_TRY(hipMemset(pl->d_region_cmax, 0, sizeof(int) * pl->num_regions));
        HIP_TRY(launch_region_count(pl->params(), pl->d_region_cmax, nullptr));
    }
    // The planar 64-pulse scatter of the linear series / hilo modes and of the
    // nearest rules with series ranges streams 32-byte records that the
    // region-norm kernel writes on every forward call.
    // The image's last record group may run into four spare records, which
    // are zero (zero offset and amplitude: a zero contribution).
    if (pl->planar && pl->fwd_pulses == 64 &&
        ((interp == SARHIP_INTERP_LINEAR && (pl->precision == SARHIP_PRECISION_SERIES ||
                                             pl->precision == SARHIP_PRECISION_HILO)) ||
         (interp != SARHIP_INTERP_LINEAR && pl->precision == SARHIP_PRECISION_SERIES))) {
        HIP_TRY(hipMalloc((void **)&pl->d_srec, sizeof(ScatRec) * (M + 4)));
        HIP_TRY(hipMemset(pl->d_srec + M, 0, sizeof(ScatRec) * 4));
    }
    // Backprojection: split each tile's pulses over several workgroups until
    // the grid has about 32 workgroups per CU, keeping at least 128 pulses per
    // split; partial images are summed in FP64 in split order.  (C2, 1024
    // tiles: 2 / 4 / 6 / 8 / 12 / 16 splits took 1.64 / 1.56 / 1.54 / 1.49 /
    // 1.47 / 1.48 ms; the smaller counts leave the last wave of workgroups
    // short of work.)
    {
        int S = 1;
        while ((long long)pl->num_tiles * S < 8192 && pl->Pv / (2 * S) >= 128) S *= 2;
        const int forced = env_int("SARHIP_BP_SPLITS", 0);
        if (forced > 0) S = std::min(forced, std::max(pl->Pv, 1));
        pl->bp_splits = S;
        if (S > 1) HIP_TRY(hipMalloc((void **)&pl->d_bp_partial, sizeof(double2) * S * M));
    }
    if (pl->win > 0)
        HIP_TRY(hipMalloc((void **)&pl->d_partial,
                          sizeof(int2) * nrows * pl->num_regions * (size_t)pl->win));
    if (need_series && pl->Pv > 0) {
        const size_t rec = mono ? sizeof(SeriesMono) : sizeof(SeriesBi);
        HIP_TRY(hipMalloc(&pl->d_series, rec * pl->num_tiles * (size_t)pl->Pv));
        HIP_TRY(launch_series_consts(pl->d_tile_c, pl->num_tiles, pl->d_pc, pl->d_rows, pl->Pv,
                                     pl->by_dx, pl->kc_rev, mono, pl->d_series, nullptr));
        if (interp == SARHIP_INTERP_LINEAR) {
            const size_t n = (size_t)pl->num_tiles * (size_t)pl->Pv;
            // the monostatic planar third-order gather reads SeriesGather
            // records (the kernel that selects it: bp_dispatch_prec)
            if (mono && pl->planar && pl->precision == SARHIP_PRECISION_SERIES &&
                pl->series_order == 3) {
                HIP_TRY(hipMalloc((void **)&pl->d_series_g, sizeof(SeriesGather) * n));
                HIP_TRY(launch_series_gather(pl->d_series, n, pl->shift_lo, pl->inv_nw, (double)nw,
                                             pl->kc_rev / pl->by_dx, 0.0, pl->d_series_g,
                                             nullptr));
            } else {
                HIP_TRY(hipMalloc((void **)&pl->d_bp_cu, sizeof(double) * n));
                HIP_TRY(launch_bp_shift(pl->d_series, mono, n, pl->shift_lo, pl->inv_nw,
                                        (double)nw, pl->d_bp_cu, nullptr));
            }
        }
    }

    // Spectral stages: the fused polyphase kernels when N_w = U N with an
    // integer U and a supported N (twiddle tables from FP64), else rocFFT.
    {
        const size_t U = nw / N;
        pl->fused_U = 0;
        if (U * N == nw && fused_fft_supported((int)N, (int)U) && env_int("SARHIP_FUSED_FFT", 1)) {
            // two-level twiddle tables in FP64, and each entry rounded once to FP32
            std::vector<double2> tab64;
            auto put = [&](size_t count, double step, size_t len) {
                for (size_t q = 0; q < count; ++q) {
                    const double a = -2.0 * M_PI * (double)(q * step) / (double)len;
                    tab64.push_back(make_double2(std::cos(a), std::sin(a)));
                }
            };
            put(16, 1, N);
            put(N / 16, 16, N);
            put(256, 1, nw);
            put(nw / 256, 256, nw);
            std::vector<float2> tab;
            for (const double2 &t : tab64) tab.push_back(make_float2((float)t.x, (float)t.y));
            if ((rc = upload((void **)&pl->d_fft_tables, tab.data(), sizeof(float2) * tab.size())))
                return rc;
            // exact ranges requested: the forward's transforms in FP64 as well;
            // otherwise FP32 transforms, and FP64 again for the rows that cancel
            // (kappa above SARHIP_FWD_KAPPA, default 2; 0 turns this off)
            const double kappa = env_double("SARHIP_FWD_KAPPA", 2.0);
            if (pl->requested_precision != SARHIP_PRECISION_FP64 && kappa > 0.0) {
                pl->fwd_kappa = kappa;
                pl->fwd_kappa2 = kappa * kappa;
                HIP_TRY(hipMalloc((void **)&pl->d_fft_redo, sizeof(int) * std::max<size_t>(P, 1)));
                HIP_TRY(hipMemset(pl->d_fft_redo, 0, sizeof(int) * std::max<size_t>(P, 1)));
            }
            if ((pl->requested_precision == SARHIP_PRECISION_FP64 || pl->fwd_kappa2 > 0.0) &&
                (rc = upload((void **)&pl->d_fft_tables64, tab64.data(),
                             sizeof(double2) * tab64.size())))
                return rc;
            pl->fused_U = (int)U;
            // rows of invalid pulses are never read: empty spans
            HIP_TRY(hipMalloc((void **)&pl->d_aspan, sizeof(int2) * std::max<size_t>(P, 1)));
            HIP_TRY(hipMemset(pl->d_aspan, 0, sizeof(int2) * std::max<size_t>(P, 1)));
            HIP_TRY(launch_adjoint_spans(pl->d_pc, pl->d_rows, pl->Pv, pl->d_tile_c,
                                         pl->num_tiles, mono, pl->by_dx, (int)nw, pl->d_aspan,
                                         nullptr));
            std::vector<int2> sp(P);
            HIP_TRY(hipMemcpy(sp.data(), pl->d_aspan, sizeof(int2) * P, hipMemcpyDeviceToHost));
            double covered = 0.0;
            for (const int2 &v : sp) covered += v.y;
            pl->fft_sparse = covered < 0.5 * (double)pl->Pv * (double)nw;
        }
    }
    if (pl->fused_U > 0) {
        // The gather's pair path (pair_rows_kernel): its pixel phase
        // rho f + phi' is formed in FP32, so rho must be an FP32 number (to
        // 2^-34 revolutions per bin) and at most 2.5 (|phase| < 1.75).
        // SARHIP_BP_PAIRS=0 turns it off.
        const double rho = pl->kc_rev / pl->by_dx;
        const bool rho_ok = std::fabs(rho) <= 2.5 &&
                            std::fabs(rho - (double)(float)rho) <= 0x1p-34 &&
                            env_int("SARHIP_BP_PAIRS", 1);
        // the nearest rules' pair path reads gather records whose cu carries
        // the adjoint rule's offset: 1 for round(1/2 + u) (CPP:346-347), 1/2
        // for round(u)
        if (interp != SARHIP_INTERP_LINEAR && rho_ok && mono && pl->planar && pl->d_series &&
            pl->precision == SARHIP_PRECISION_SERIES && pl->series_order == 3 && pl->Pv > 0) {
            const size_t n = (size_t)pl->num_tiles * (size_t)pl->Pv;
            HIP_TRY(hipMalloc((void **)&pl->d_series_g, sizeof(SeriesGather) * n));
            HIP_TRY(launch_series_gather(pl->d_series, n, pl->shift_lo, pl->inv_nw, (double)nw, rho,
                                         interp == SARHIP_INTERP_NEAREST_REF ? 1.0 : 0.5,
                                         pl->d_series_g, nullptr));
        }
        if (pl->d_series_g && pl->Pv > 0 && rho_ok)
            // (the nearest rules read A_n alone: float2 entries)
            HIP_TRY(hipMalloc((void **)&pl->d_bp_pairs,
                              (interp == SARHIP_INTERP_LINEAR ? sizeof(float4) : sizeof(float2)) *
                                  P * pl->ld));
            // an integer rho makes the demodulating phasor a sign, which the
            // adjoint spectral kernel applies as it stores the row
            // (SARHIP_PAIRS_FUSED=0: pair_rows_kernel after it instead)
            if (rho == std::rint(rho) && env_int("SARHIP_PAIRS_FUSED", 1))
                pl->pairs_sgn = ((long long)std::llround(std::fabs(rho)) % 2 == 0) ? 1 : -1;
        HIP_TRY(hipDeviceSynchronize());
        owner.p = nullptr;
        *out = pl;
        return SARHIP_OK;
    }

    // rocFFT: batched in-place complex64 transforms of length N_w.
    const size_t len = nw;
    FFT_TRY(rocfft_plan_create(&pl->fft_fwd, rocfft_placement_inplace,
                               rocfft_transform_type_complex_forward, rocfft_precision_single, 1,
                               &len, P, nullptr));
    {
        // adjoint: rows of stride ld (padded tails), in place
        rocfft_plan_description desc = nullptr;
        FFT_TRY(rocfft_plan_description_create(&desc));
        const size_t stride = 1, dist = pl->ld;
        rocfft_status st = rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved,
            nullptr, nullptr, 1, &stride, dist, 1, &stride, dist);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&pl->fft_bwd, rocfft_placement_inplace,
                                    rocfft_transform_type_complex_inverse, rocfft_precision_single,
                                    1, &len, P, desc);
        rocfft_plan_description_destroy(desc);
        FFT_TRY(st);
    }
    // exact ranges requested: the forward's transform in FP64 as in the fused
    // kernels (the rows are widened before it)
    size_t w3 = 0;
    if (pl->requested_precision == SARHIP_PRECISION_FP64) {
        FFT_TRY(rocfft_plan_create(&pl->fft_fwd64, rocfft_placement_inplace,
                                   rocfft_transform_type_complex_forward, rocfft_precision_double,
                                   1, &len, P, nullptr));
        FFT_TRY(rocfft_plan_get_work_buffer_size(pl->fft_fwd64, &w3));
        HIP_TRY(hipMalloc((void **)&pl->d_prof64, sizeof(double2) * P * nw));
    }
    size_t w1 = 0, w2 = 0;
    FFT_TRY(rocfft_plan_get_work_buffer_size(pl->fft_fwd, &w1));
    FFT_TRY(rocfft_plan_get_work_buffer_size(pl->fft_bwd, &w2));
    pl->fft_work_size = std::max(std::max(w1, w2), w3);
    FFT_TRY(rocfft_execution_info_create(&pl->fft_info));
    if (pl->fft_work_size) {
        HIP_TRY(hipMalloc(&pl->fft_work, pl->fft_work_size));
        FFT_TRY(rocfft_execution_info_set_work_buffer(pl->fft_info, pl->fft_work,
                                                      pl->fft_work_size));
    }
    HIP_TRY(hipDeviceSynchronize());

    owner.p = nullptr;
    *out = pl;
    return SARHIP_OK;
}
}  // namespace

int sarhip_plan_create(sarhip_plan_t *out, size_t num_fast_times, size_t num_slow_times,
                       size_t num_scatterers, const double *tx, const double *rx,
                       const double *srp, const double *pos, const double *waveform_fft,
                       const double *slow_time_weighting, double centre_frequency,
                       double sample_frequency, double c_eff, double upsample_ratio,
                       double sign_multiplier, int interp, int precision, int device) {
    PosSource src;
    src.kind = PosSource::HOST;
    src.ptr = pos;
    return plan_create_impl(out, num_fast_times, num_slow_times, num_scatterers, tx, rx, srp, src,
                            waveform_fft, slow_time_weighting, centre_frequency, sample_frequency,
                            c_eff, upsample_ratio, sign_multiplier, interp, precision, device);
}

int sarhip_plan_create_device_positions(sarhip_plan_t *out, size_t num_fast_times,
                                        size_t num_slow_times, size_t num_scatterers,
                                        const double *tx, const double *rx, const double *srp,
                                        const void *d_pos, const double *waveform_fft,
                                        const double *slow_time_weighting, double centre_frequency,
                                        double sample_frequency, double c_eff,
                                        double upsample_ratio, double sign_multiplier, int interp,
                                        int precision, int device) {
    PosSource src;
    src.kind = PosSource::DEVICE;
    src.ptr = (const double *)d_pos;
    return plan_create_impl(out, num_fast_times, num_slow_times, num_scatterers, tx, rx, srp, src,
                            waveform_fft, slow_time_weighting, centre_frequency, sample_frequency,
                            c_eff, upsample_ratio, sign_multiplier, interp, precision, device);
}

namespace {
int grid_spec(const size_t shape[3], const double dims[3], const double centre[3],
              const double rotation[9], GridSpec &g, size_t &M) {
    if (!shape || !dims || !centre || !rotation) return fail(SARHIP_EINVAL, "NULL grid argument");
    if (shape[0] == 0 || shape[1] == 0 || shape[2] == 0 || shape[0] > (size_t)INT32_MAX ||
        shape[1] > (size_t)INT32_MAX || shape[2] > (size_t)INT32_MAX)
        return fail(SARHIP_EINVAL, "grid shape entries must be positive");
    const double m = (double)shape[0] * (double)shape[1] * (double)shape[2];
    if (m > (double)INT32_MAX) return fail(SARHIP_EINVAL, "grid has more than 2^31 - 1 points");
    g.nz = (int)shape[0];
    g.ny = (int)shape[1];
    g.nx = (int)shape[2];
    g.lz = dims[0];
    g.ly = dims[1];
    g.lx = dims[2];
    for (int k = 0; k < 9; ++k) g.rot[k] = rotation[k];
    for (int k = 0; k < 3; ++k) g.centre[k] = centre[k];
    M = shape[0] * shape[1] * shape[2];
    return SARHIP_OK;
}
}  // namespace

int sarhip_plan_create_grid(sarhip_plan_t *out, size_t num_fast_times, size_t num_slow_times,
                            const size_t grid_shape[3], const double grid_dimensions[3],
                            const double grid_centre[3], const double rotation[9],
                            const double *tx, const double *rx, const double *srp,
                            const double *waveform_fft, const double *slow_time_weighting,
                            double centre_frequency, double sample_frequency, double c_eff,
                            double upsample_ratio, double sign_multiplier, int interp,
                            int precision, int device) {
    if (out) *out = nullptr;
    PosSource src;
    src.kind = PosSource::GRID;
    size_t M = 0;
    int rc = grid_spec(grid_shape, grid_dimensions, grid_centre, rotation, src.grid, M);
    if (rc) return rc;
    return plan_create_impl(out, num_fast_times, num_slow_times, M, tx, rx, srp, src,
                            waveform_fft, slow_time_weighting, centre_frequency, sample_frequency,
                            c_eff, upsample_ratio, sign_multiplier, interp, precision, device);
}

int sarhip_grid_positions(const size_t grid_shape[3], const double grid_dimensions[3],
                          const double grid_centre[3], const double rotation[9], void *d_pos,
                          void *stream) {
    GridSpec g;
    size_t M = 0;
    int rc = grid_spec(grid_shape, grid_dimensions, grid_centre, rotation, g, M);
    if (rc) return rc;
    if (!d_pos) return fail(SARHIP_EINVAL, "NULL output pointer");
    HIP_TRY(launch_grid_positions(g, (double *)d_pos, (hipStream_t)stream));
    return SARHIP_OK;
}

int sarhip_decode_signal(const void *d_raw, int format, size_t num_vectors, size_t num_samples,
                         const float *d_amp_sf, void *d_out, void *stream) {
    if (!d_raw || !d_out) return fail(SARHIP_EINVAL, "NULL argument");
    if (format < 0 || format > 2) return fail(SARHIP_EINVAL, "format must be 0 (CF8), 1 (CI4) or 2 (CI2)");
    if (num_vectors > (size_t)INT32_MAX || num_samples > (size_t)INT32_MAX / 8)
        return fail(SARHIP_EINVAL, "signal array too large");
    if (num_vectors == 0 || num_samples == 0) return SARHIP_OK;
    const size_t align = format == 0 ? 8 : format == 1 ? 4 : 2;
    if ((uintptr_t)d_raw % align) return fail(SARHIP_EINVAL, "raw samples must be aligned to one sample");
    HIP_TRY(launch_decode_signal(d_raw, format, (int)num_vectors, (int)num_samples, d_amp_sf,
                                 (float2 *)d_out, (hipStream_t)stream));
    return SARHIP_OK;
}

int sarhip_baseband_values(const void *d_pos, size_t num_scatterers, const double spot[3],
                           const double displacement[3], double sign_multiplier, void *d_out,
                           void *stream) {
    if (!d_pos || !d_out || !spot || !displacement) return fail(SARHIP_EINVAL, "NULL argument");
    if (num_scatterers > (size_t)INT32_MAX) return fail(SARHIP_EINVAL, "too many scatterers");
    HIP_TRY(launch_baseband((const double *)d_pos, (int)num_scatterers, spot, displacement,
                            sign_multiplier, (double2 *)d_out, (hipStream_t)stream));
    return SARHIP_OK;
}

int sarhip_plan_destroy(sarhip_plan_t plan) {
    delete plan;
    return SARHIP_OK;
}

int sarhip_plan_info(sarhip_plan_t plan, size_t *nw, double *by_dx, double *kc) {
    if (!plan) return fail(SARHIP_EINVAL, "NULL plan");
    if (nw) *nw = plan->nw;
    if (by_dx) *by_dx = plan->by_dx;
    if (kc) *kc = plan->kc;
    return SARHIP_OK;
}

int sarhip_plan_layout(sarhip_plan_t plan, int *precision, int *monostatic, int *valid_pulses,
                       int *fwd_window_bins, int *fwd_pulses_per_group, double *series_err_bins,
                       int *fwd_region_tiles, int *bp_series_order, double *bp_series_err_bins,
                       int *fused_fft_ratio) {
    if (!plan) return fail(SARHIP_EINVAL, "NULL plan");
    if (fused_fft_ratio) *fused_fft_ratio = plan->fused_U;
    if (bp_series_order) *bp_series_order = plan->series_order;
    if (bp_series_err_bins) *bp_series_err_bins = plan->bp_series_err_bins;
    if (fwd_region_tiles) *fwd_region_tiles = plan->region_tiles;
    if (precision) *precision = plan->precision;
    if (monostatic) *monostatic = plan->mono ? 1 : 0;
    if (valid_pulses) *valid_pulses = plan->Pv;
    if (fwd_window_bins) *fwd_window_bins = plan->win;
    if (fwd_pulses_per_group) *fwd_pulses_per_group = plan->fwd_pulses;
    if (series_err_bins) *series_err_bins = plan->series_err_bins;
    return SARHIP_OK;
}

int sarhip_plan_forward_fp64_rows(sarhip_plan_t plan, double *kappa, int *rows) {
    if (!plan) return fail(SARHIP_EINVAL, "NULL plan");
    const bool adaptive = plan->fused_U > 0 && plan->d_fft_redo;
    if (kappa) *kappa = adaptive ? plan->fwd_kappa : 0.0;
    if (rows) {
        *rows = 0;
        if (adaptive) {
            DeviceGuard g(plan->device);
            std::vector<int> f(plan->P);
            HIP_TRY(hipDeviceSynchronize());
            HIP_TRY(hipMemcpy(f.data(), plan->d_fft_redo, sizeof(int) * plan->P,
                              hipMemcpyDeviceToHost));
            for (const int v : f) *rows += v != 0;
        }
    }
    return SARHIP_OK;
}

int sarhip_plan_gather_pairs(sarhip_plan_t plan, int *enabled) {
    if (!plan || !enabled) return fail(SARHIP_EINVAL, "NULL argument");
    *enabled = plan->d_bp_pairs ? 1 : 0;
    return SARHIP_OK;
}

int sarhip_plan_set_timing(sarhip_plan_t plan, int enable) {
    if (!plan) return fail(SARHIP_EINVAL, "NULL plan");
    plan->timing = enable != 0;
    plan->ev_used[0] = plan->ev_used[1] = 0;
    return SARHIP_OK;
}

int sarhip_plan_timing(sarhip_plan_t plan, int which, float *ms4, int *ncalls) {
    if (!plan || !ms4 || (which != 0 && which != 1)) return fail(SARHIP_EINVAL, "bad argument");
    DeviceGuard g(plan->device);
    double acc[3] = {0, 0, 0};
    const size_t n = plan->ev_used[which];
    const hipEvent_t *e = plan->ev_pool[which].data();
    for (size_t c = 0; c < n; ++c, e += kEventsPerCall) {
        HIP_TRY(hipEventSynchronize(e[3]));
        for (int k = 0; k < 3; ++k) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e[k], e[k + 1]));
            acc[k] += ms;
        }
    }
    const double d = n ? (double)n : 1.0;
    // forward events bracket (scatter, FFT, truncate); adjoint (window, FFT, gather)
    const int order_fwd[3] = {2, 1, 0};
    for (int k = 0; k < 3; ++k) {
        const int slot = which == 0 ? order_fwd[k] : k;
        ms4[slot] = (float)(acc[k] / d);
    }
    ms4[3] = (float)((acc[0] + acc[1] + acc[2]) / d);
    if (ncalls) *ncalls = (int)n;
    return SARHIP_OK;
}

int sarhip_plan_kernel_timing(sarhip_plan_t plan, int which, float *ms, int *ncalls) {
    if (!plan || !ms || (which != 0 && which != 1)) return fail(SARHIP_EINVAL, "bad argument");
    DeviceGuard g(plan->device);
    double acc = 0.0;
    const size_t n = plan->ev_used[which];
    const hipEvent_t *e = plan->ev_pool[which].data();
    for (size_t c = 0; c < n; ++c, e += kEventsPerCall) {
        HIP_TRY(hipEventSynchronize(e[3]));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, e[4], e[5]));
        acc += t;
    }
    *ms = (float)(n ? acc / (double)n : 0.0);
    if (ncalls) *ncalls = (int)n;
    return SARHIP_OK;
}

int sarhip_forward(sarhip_plan_t pl, const void *amps, void *phase_history, int order,
                   void *stream) {
    if (!pl || !amps || !phase_history) return fail(SARHIP_EINVAL, "NULL argument");
    if (order != SARHIP_ORDER_FFT && order != SARHIP_ORDER_CENTRED)
        return fail(SARHIP_EINVAL, "order must be SARHIP_ORDER_FFT or SARHIP_ORDER_CENTRED");
    DeviceGuard g(pl->device);
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = begin_timing(pl, 0))) return rc;
    if ((rc = record(pl, 0, s))) return rc;
    // The merge writes each row's covered span and records it; the fused
    // spectral stage reads only those spans, so the rows need no clearing.
    // Global-atomic scatter (no windows) and rocFFT need zeroed rows.
    const bool spans = pl->fused_U > 0 && pl->Pv > 0 && pl->win > 0;
    if (!spans) HIP_TRY(launch_fill_zero(pl->d_prof, sizeof(float2) * pl->P * pl->nw, s));
    if (pl->Pv > 0) {
        HIP_TRY(launch_gather_perm((const float2 *)amps, pl->d_perm, (int)pl->M, pl->d_amps_perm,
                                   s));
        KernelParams k = pl->params();
        k.fwd_spans = spans ? 1 : 0;
        HIP_TRY(launch_region_l1(k, s));
        if ((rc = record(pl, 4, s))) return rc;
        HIP_TRY(launch_fp_bin(k, s));
        if ((rc = record(pl, 5, s))) return rc;
        if (pl->win > 0) HIP_TRY(launch_fp_reduce(k, s));
    } else {
        if ((rc = record(pl, 4, s))) return rc;
        if ((rc = record(pl, 5, s))) return rc;
    }
    if ((rc = record(pl, 1, s))) return rc;
    if (pl->fused_U > 0) {
        // FFT and truncation in one pass over the profile rows
        HIP_TRY(launch_fft_down(pl->d_prof, spans ? pl->d_fspan : nullptr, pl->d_W, pl->d_wfac_fwd,
                                (int)pl->N, pl->fused_U, (int)pl->P, order == SARHIP_ORDER_CENTRED,
                                pl->d_fft_tables, pl->d_fft_tables64, (float2 *)phase_history,
                                spans && pl->fft_sparse, pl->d_fft_redo, pl->fwd_kappa2, s));
        if ((rc = record(pl, 2, s))) return rc;
        if ((rc = record(pl, 3, s))) return rc;
        return SARHIP_OK;
    }
    if (pl->fft_fwd64) {
        HIP_TRY(launch_widen(pl->d_prof, pl->d_prof64, pl->P * pl->nw, s));
        if ((rc = run_fft(pl, pl->fft_fwd64, pl->d_prof64, s))) return rc;
        if ((rc = record(pl, 2, s))) return rc;
        HIP_TRY(launch_truncate64(pl->d_prof64, pl->d_W, pl->d_wfac_fwd, (int)pl->N, (int)pl->nw,
                                  (int)pl->P, order == SARHIP_ORDER_CENTRED,
                                  (float2 *)phase_history, s));
        if ((rc = record(pl, 3, s))) return rc;
        return SARHIP_OK;
    }
    if ((rc = run_fft(pl, pl->fft_fwd, pl->d_prof, s))) return rc;
    if ((rc = record(pl, 2, s))) return rc;
    HIP_TRY(launch_truncate(pl->d_prof, pl->d_W, pl->d_wfac_fwd, (int)pl->N, (int)pl->nw,
                            (int)pl->P, order == SARHIP_ORDER_CENTRED, (float2 *)phase_history, s));
    if ((rc = record(pl, 3, s))) return rc;
    return SARHIP_OK;
}

int sarhip_adjoint(sarhip_plan_t pl, const void *phase_history, void *amps, int order,
                   void *stream) {
    if (!pl || !amps || !phase_history) return fail(SARHIP_EINVAL, "NULL argument");
    if (order != SARHIP_ORDER_FFT && order != SARHIP_ORDER_CENTRED)
        return fail(SARHIP_EINVAL, "order must be SARHIP_ORDER_FFT or SARHIP_ORDER_CENTRED");
    DeviceGuard g(pl->device);
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = begin_timing(pl, 1))) return rc;
    if ((rc = record(pl, 0, s))) return rc;
    if (pl->fused_U > 0) {
        // band weights, zero padding, inverse FFT and row tails in one pass
        if ((rc = record(pl, 1, s))) return rc;
        UpPairs up{nullptr, 0, 1.f};
        if (pl->d_bp_pairs && pl->pairs_sgn != 0)
            up = UpPairs{pl->d_bp_pairs, pl->interp == SARHIP_INTERP_LINEAR ? 1 : 2,
                         (float)pl->pairs_sgn};
        HIP_TRY(launch_fft_up(pl->d_aspan, (const float2 *)phase_history, pl->d_Wconj,
                              pl->d_wfac_adj, (int)pl->N, pl->fused_U, (int)pl->ld, (int)pl->P,
                              order == SARHIP_ORDER_CENTRED, pl->d_fft_tables, pl->d_prof,
                              pl->fft_sparse, up, s));
        if (pl->d_bp_pairs && pl->pairs_sgn == 0)
            HIP_TRY(launch_pair_rows(pl->d_prof, pl->d_aspan, pl->d_rows, pl->Pv, (int)pl->nw,
                                     (int)pl->ld, pl->kc_rev / pl->by_dx,
                                     pl->interp != SARHIP_INTERP_LINEAR, pl->d_bp_pairs, s));
    } else {
        HIP_TRY(launch_window_pad((const float2 *)phase_history, pl->d_Wconj, pl->d_wfac_adj,
                                  (int)pl->N, (int)pl->nw, (int)pl->ld, (int)pl->P,
                                  order == SARHIP_ORDER_CENTRED, pl->d_prof, s));
        if ((rc = record(pl, 1, s))) return rc;
        if ((rc = run_fft(pl, pl->fft_bwd, pl->d_prof, s))) return rc;
        HIP_TRY(launch_wrap_pad(pl->d_prof, (int)pl->nw, (int)pl->ld, (int)pl->P, s));
    }
    if ((rc = record(pl, 2, s))) return rc;
    KernelParams k = pl->params();
    k.out = (float2 *)amps;
    if ((rc = record(pl, 4, s))) return rc;
    HIP_TRY(launch_bp_gather(k, s));
    if ((rc = record(pl, 5, s))) return rc;
    HIP_TRY(launch_bp_reduce(k, s));
    if ((rc = record(pl, 3, s))) return rc;
    return SARHIP_OK;
}

// Reference-compatible entry points (host FP64 record in, host FP64 out).
static int evaluate_host(SarCalculationInfo *info, bool forward) {
    if (!info) return fail(SARHIP_EINVAL, "NULL SarCalculationInfo");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    sarhip_plan_t pl = nullptr;
    int rc = sarhip_plan_create(&pl, info->num_fast_times, info->num_slow_times,
                                info->num_scatterers, info->transmit_posns, info->receive_posns,
                                info->stab_ref_posns, info->scat_posns, info->waveform_fft,
                                info->slow_time_weighting, info->centre_frequency,
                                info->sample_frequency, info->c_eff, info->upsample_ratio,
                                info->sign_multiplier, SARHIP_INTERP_NEAREST_REF,
                                SARHIP_PRECISION_FP64, dev);
    if (rc) return rc;
    struct PlanOwner {
        sarhip_plan_t p;
        ~PlanOwner() { sarhip_plan_destroy(p); }
    } po{pl};
    const size_t M = info->num_scatterers, P = info->num_slow_times, N = info->num_fast_times;
    if (!info->phase_history || !info->scattering_amplitudes)
        return fail(SARHIP_EINVAL, "NULL operand pointer");
    std::vector<float2> hx(M), hy(P * N);
    float2 *dx = nullptr, *dy = nullptr;
    struct Bufs {
        float2 *&a, *&b;
        ~Bufs() {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
        }
    } bufs{dx, dy};
    HIP_TRY(hipMalloc((void **)&dx, sizeof(float2) * M));
    HIP_TRY(hipMalloc((void **)&dy, sizeof(float2) * P * N));
    if (forward) {
        for (size_t m = 0; m < M; ++m)
            hx[m] = make_float2((float)info->scattering_amplitudes[2 * m],
                                (float)info->scattering_amplitudes[2 * m + 1]);
        HIP_TRY(hipMemcpy(dx, hx.data(), sizeof(float2) * M, hipMemcpyHostToDevice));
        if ((rc = sarhip_forward(pl, dx, dy, SARHIP_ORDER_FFT, nullptr))) return rc;
        HIP_TRY(hipMemcpy(hy.data(), dy, sizeof(float2) * P * N, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < P * N; ++i) {
            info->phase_history[2 * i] = hy[i].x;
            info->phase_history[2 * i + 1] = hy[i].y;
        }
    } else {
        for (size_t i = 0; i < P * N; ++i)
            hy[i] = make_float2((float)info->phase_history[2 * i],
                                (float)info->phase_history[2 * i + 1]);
        HIP_TRY(hipMemcpy(dy, hy.data(), sizeof(float2) * P * N, hipMemcpyHostToDevice));
        if ((rc = sarhip_adjoint(pl, dy, dx, SARHIP_ORDER_FFT, nullptr))) return rc;
        HIP_TRY(hipMemcpy(hx.data(), dx, sizeof(float2) * M, hipMemcpyDeviceToHost));
        for (size_t m = 0; m < M; ++m) {
            info->scattering_amplitudes[2 * m] = hx[m].x;
            info->scattering_amplitudes[2 * m + 1] = hx[m].y;
        }
    }
    return SARHIP_OK;
}

int forward_evaluate(SarCalculationInfo *info) { return evaluate_host(info, true); }

int adjoint_evaluate(SarCalculationInfo *info) { return evaluate_host(info, false); }

}  // extern "C"
