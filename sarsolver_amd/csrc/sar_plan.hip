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

// Interleave the low 21 bits of three integers (3-D Morton code).
uint64_t spread21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
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

constexpr size_t kLdsBudget = 64 * 1024;

}  // namespace

struct sarhip_plan_s {
    int device = 0;
    size_t N = 0, P = 0, M = 0, nw = 0;
    double fc = 0, fs = 0, c_eff = 0, U = 0, sgn = 1;
    double by_dx = 0, kc = 0, kc_rev = 0, inv_nw = 0;
    int interp = 0, precision = 0;
    bool mono = false;

    PulseConst *d_pc = nullptr;
    double *d_px = nullptr, *d_py = nullptr, *d_pz = nullptr;
    int *d_perm = nullptr;
    float2 *d_W = nullptr, *d_Wconj = nullptr, *d_wfac_fwd = nullptr, *d_wfac_adj = nullptr;
    double4 *d_ftiles = nullptr;
    int f_tile_len = 0, f_num_tiles = 0, f_pg = 0, f_win = 0;
    float2 *d_prof = nullptr;
    float2 *d_amps_perm = nullptr;

    rocfft_plan fft_fwd = nullptr, fft_bwd = nullptr;
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
        dev_free(d_px);
        dev_free(d_py);
        dev_free(d_pz);
        dev_free(d_perm);
        dev_free(d_W);
        dev_free(d_Wconj);
        dev_free(d_wfac_fwd);
        dev_free(d_wfac_adj);
        dev_free(d_ftiles);
        dev_free(d_prof);
        dev_free(d_amps_perm);
        if (fft_work) (void)hipFree(fft_work);
        if (fft_fwd) rocfft_plan_destroy(fft_fwd);
        if (fft_bwd) rocfft_plan_destroy(fft_bwd);
        if (fft_info) rocfft_execution_info_destroy(fft_info);
        for (auto &pool : ev_pool)
            for (auto &e : pool)
                if (e) (void)hipEventDestroy(e);
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

// Reserve the next event set of `which` (0 forward, 1 adjoint) for this call.
int begin_timing(sarhip_plan_t pl, int which) {
    pl->ev = nullptr;
    if (!pl->timing || pl->ev_used[which] >= kMaxTimedCalls) return SARHIP_OK;
    auto &pool = pl->ev_pool[which];
    const size_t base = 4 * pl->ev_used[which];
    while (pool.size() < base + 4) {
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

// Forward tiling: consecutive runs of `len` pixels in Morton order; centre =
// bounding-box midpoint, radius = max distance to it (inflated slightly).
void make_tiles(const std::vector<double> &sx, const std::vector<double> &sy,
                const std::vector<double> &sz, size_t M, int len, std::vector<double4> &tiles,
                double &rad_max) {
    const size_t nt = (M + len - 1) / len;
    tiles.assign(nt, double4{0, 0, 0, 0});
    rad_max = 0.0;
    for (size_t t = 0; t < nt; ++t) {
        const size_t a = t * len, b = std::min(M, a + len);
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool any = false;
        for (size_t i = a; i < b; ++i) {
            const double v[3] = {sx[i], sy[i], sz[i]};
            if (!(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]))) continue;
            any = true;
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], v[k]);
                hi[k] = std::max(hi[k], v[k]);
            }
        }
        if (!any) continue;
        const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]),
                     cz = 0.5 * (lo[2] + hi[2]);
        double r = 0.0;
        for (size_t i = a; i < b; ++i) {
            const double dx = sx[i] - cx, dy = sy[i] - cy, dz = sz[i] - cz;
            const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (std::isfinite(d)) r = std::max(r, d);
        }
        r = r * (1.0 + 1e-9) + 1e-9;
        tiles[t] = double4{cx, cy, cz, r};
        rad_max = std::max(rad_max, r);
    }
}

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return std::atoi(v);
}

}  // namespace

extern "C" {

const char *sarhip_last_error(void) { return g_last_error.c_str(); }

const char *sarhip_version(void) { return "sarsolver_hip 0.1.0 (gfx950)"; }

int sarhip_plan_create(sarhip_plan_t *out, size_t num_fast_times, size_t num_slow_times,
                       size_t num_scatterers, const double *tx, const double *rx,
                       const double *srp, const double *pos, const double *waveform_fft,
                       const double *slow_time_weighting, double centre_frequency,
                       double sample_frequency, double c_eff, double upsample_ratio,
                       double sign_multiplier, int interp, int precision, int device) {
    if (!out) return fail(SARHIP_EINVAL, "plan output pointer is NULL");
    *out = nullptr;
    if (interp < 0 || interp > 2) return fail(SARHIP_EINVAL, "interp must be 0, 1 or 2");
    if (precision != SARHIP_PRECISION_FP64)
        return fail(SARHIP_EINVAL, "precision mode not available in this build");
    const size_t N = num_fast_times, P = num_slow_times, M = num_scatterers;
    if (N == 0 || P == 0 || M == 0) return fail(SARHIP_EINVAL, "N, P and M must be positive");
    if (!tx || !rx || !srp || !pos || !waveform_fft || !slow_time_weighting)
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

    // Per-pulse constants.
    std::vector<PulseConst> pc(P);
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
        if (std::memcmp(c.tx, c.rx, sizeof(c.tx)) != 0) mono = false;
    }
    pl->mono = mono;

    // Tile (Morton) order of the pixels.
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t m = 0; m < M; ++m) {
        const double *v = pos + 3 * m;
        if (!(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]))) continue;
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], v[k]);
            hi[k] = std::max(hi[k], v[k]);
        }
    }
    double span = 0.0;
    for (int k = 0; k < 3; ++k)
        if (hi[k] > lo[k]) span = std::max(span, hi[k] - lo[k]);
    const double qscale = span > 0.0 ? (double)((1 << 21) - 1) / span : 0.0;
    std::vector<std::pair<uint64_t, int>> keys(M);
    for (size_t m = 0; m < M; ++m) {
        const double *v = pos + 3 * m;
        uint64_t code = ~0ULL;
        if (std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2])) {
            uint64_t q[3];
            for (int k = 0; k < 3; ++k) {
                double t = (v[k] - lo[k]) * qscale;
                t = std::min(std::max(t, 0.0), (double)((1 << 21) - 1));
                q[k] = (uint64_t)t;
            }
            code = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
        }
        keys[m] = {code, (int)m};
    }
    std::sort(keys.begin(), keys.end());
    std::vector<int> perm(M);
    std::vector<double> sx(M), sy(M), sz(M);
    for (size_t i = 0; i < M; ++i) {
        const int m = keys[i].second;
        perm[i] = m;
        sx[i] = pos[3 * (size_t)m];
        sy[i] = pos[3 * (size_t)m + 1];
        sz[i] = pos[3 * (size_t)m + 2];
    }

    // Forward tiling: the largest tile whose LDS window still leaves room for
    // at least 16 pulses per workgroup.
    std::vector<double4> tiles;
    int tile_len = 256, pg = 1, win = 0;
    {
        const int forced = env_int("SARHIP_FP_TILE", 0);
        const int cands[] = {4096, 2048, 1024, 512, 256};
        bool chosen = false;
        for (int len : cands) {
            if (forced > 0 && len != forced) continue;
            double rad = 0.0;
            make_tiles(sx, sy, sz, M, len, tiles, rad);
            const double wd = std::ceil(4.0 * rad * pl->by_dx) + 8.0;
            const int w = wd > 1e7 ? (1 << 24) : (int)wd;
            const int g = (int)((kLdsBudget - 64) / (8 * (size_t)w + 8));
            const int need = (int)std::min<size_t>(16, P);
            if (g >= need || len == 256 || forced > 0) {
                tile_len = len;
                win = g >= 1 ? w : 0;
                pg = g >= 1 ? std::min<int>(g, 64) : 16;
                chosen = true;
                break;
            }
        }
        if (!chosen) {
            double rad = 0.0;
            make_tiles(sx, sy, sz, M, 256, tiles, rad);
            tile_len = 256;
            win = 0;
            pg = 16;
        }
        const int forced_pg = env_int("SARHIP_FP_PG", 0);
        if (forced_pg > 0 && win > 0) {
            const int gmax = (int)((kLdsBudget - 64) / (8 * (size_t)win + 8));
            pg = std::max(1, std::min(forced_pg, gmax));
        }
        pg = (int)std::min<size_t>((size_t)pg, P);
        if (pg < 1) pg = 1;
    }
    pl->f_tile_len = tile_len;
    pl->f_num_tiles = (int)tiles.size();
    pl->f_pg = pg;
    pl->f_win = win;

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
    int rc;
    if ((rc = upload((void **)&pl->d_pc, pc.data(), sizeof(PulseConst) * P))) return rc;
    if ((rc = upload((void **)&pl->d_px, sx.data(), sizeof(double) * M))) return rc;
    if ((rc = upload((void **)&pl->d_py, sy.data(), sizeof(double) * M))) return rc;
    if ((rc = upload((void **)&pl->d_pz, sz.data(), sizeof(double) * M))) return rc;
    if ((rc = upload((void **)&pl->d_perm, perm.data(), sizeof(int) * M))) return rc;
    if ((rc = upload((void **)&pl->d_W, W.data(), sizeof(float2) * N))) return rc;
    if ((rc = upload((void **)&pl->d_Wconj, Wc.data(), sizeof(float2) * N))) return rc;
    if ((rc = upload((void **)&pl->d_wfac_fwd, wf.data(), sizeof(float2) * P))) return rc;
    if ((rc = upload((void **)&pl->d_wfac_adj, wa.data(), sizeof(float2) * P))) return rc;
    if ((rc = upload((void **)&pl->d_ftiles, tiles.data(), sizeof(double4) * tiles.size())))
        return rc;
    HIP_TRY(hipMalloc((void **)&pl->d_prof, sizeof(float2) * P * nw));
    HIP_TRY(hipMalloc((void **)&pl->d_amps_perm, sizeof(float2) * M));

    // rocFFT: batched in-place complex64 transforms of length N_w.
    const size_t len = nw;
    FFT_TRY(rocfft_plan_create(&pl->fft_fwd, rocfft_placement_inplace,
                               rocfft_transform_type_complex_forward, rocfft_precision_single, 1,
                               &len, P, nullptr));
    FFT_TRY(rocfft_plan_create(&pl->fft_bwd, rocfft_placement_inplace,
                               rocfft_transform_type_complex_inverse, rocfft_precision_single, 1,
                               &len, P, nullptr));
    size_t w1 = 0, w2 = 0;
    FFT_TRY(rocfft_plan_get_work_buffer_size(pl->fft_fwd, &w1));
    FFT_TRY(rocfft_plan_get_work_buffer_size(pl->fft_bwd, &w2));
    pl->fft_work_size = std::max(w1, w2);
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
    for (size_t c = 0; c < n; ++c, e += 4) {
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
    HIP_TRY(launch_gather_perm((const float2 *)amps, pl->d_perm, (int)pl->M, pl->d_amps_perm, s));
    HIP_TRY(launch_fill_zero(pl->d_prof, sizeof(float2) * pl->P * pl->nw, s));
    BinArgs a;
    a.px = pl->d_px;
    a.py = pl->d_py;
    a.pz = pl->d_pz;
    a.amps_perm = pl->d_amps_perm;
    a.M = (int)pl->M;
    a.tile_len = pl->f_tile_len;
    a.num_tiles = pl->f_num_tiles;
    a.tiles = pl->d_ftiles;
    a.pc = pl->d_pc;
    a.P = (int)pl->P;
    a.pg = pl->f_pg;
    a.win = pl->f_win;
    a.prof = pl->d_prof;
    a.nw = (int)pl->nw;
    a.by_dx = pl->by_dx;
    a.kc_rev = pl->kc_rev;
    a.inv_nw = pl->inv_nw;
    a.interp = pl->interp;
    a.mono = pl->mono;
    HIP_TRY(launch_fp_bin(a, s));
    if ((rc = record(pl, 1, s))) return rc;
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
    HIP_TRY(launch_window_pad((const float2 *)phase_history, pl->d_Wconj, pl->d_wfac_adj,
                              (int)pl->N, (int)pl->nw, (int)pl->P,
                              order == SARHIP_ORDER_CENTRED, pl->d_prof, s));
    if ((rc = record(pl, 1, s))) return rc;
    if ((rc = run_fft(pl, pl->fft_bwd, pl->d_prof, s))) return rc;
    if ((rc = record(pl, 2, s))) return rc;
    GatherArgs a;
    a.px = pl->d_px;
    a.py = pl->d_py;
    a.pz = pl->d_pz;
    a.perm = pl->d_perm;
    a.M = (int)pl->M;
    a.pc = pl->d_pc;
    a.P = (int)pl->P;
    a.prof = pl->d_prof;
    a.nw = (int)pl->nw;
    a.by_dx = pl->by_dx;
    a.kc_rev = pl->kc_rev;
    a.inv_nw = pl->inv_nw;
    a.out = (float2 *)amps;
    a.interp = pl->interp;
    a.mono = pl->mono;
    HIP_TRY(launch_bp_gather(a, s));
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
