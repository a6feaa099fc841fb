// sar_asan_driver.cpp -- drives the C ABI of libsarsolver_hip.so (include/
// sarsolver_hip.h) on small problems, for a build whose HOST code carries
// AddressSanitizer (tests/asan/build.sh): the plan set-up, the record entry
// points and the argument checks are host code with non-trivial indexing, so
// a heap overflow or use after free there surfaces as an ASan report and a
// non-zero exit.  Device code is built as usual.  Checks made along the way:
// plans from host, device and grid positions give bit-identical outputs;
// every output is finite; bad arguments return an error code and a message.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sarsolver_hip.h"

namespace {

int g_fail = 0;

#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);            \
            std::fprintf(stderr, "\n");                   \
            ++g_fail;                                     \
        }                                                 \
    } while (0)

#define HIP_OK(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                           \
            std::exit(2);                                                     \
        }                                                                     \
    } while (0)

// Deterministic pseudo-random numbers in [-1, 1).
struct Lcg {
    unsigned long long s;
    double next() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
};

struct Problem {
    size_t P, N, nx, ny, M;
    std::vector<double> tx, srp, pos, W, w, amps, ph;
};

// Circular monostatic path (R = 10 km, h = 5 km, 2 deg), grid ny x nx at 0.25 m
// in the plane z = 0 centred on the origin, as GridSarScene([1, ny, nx], ...).
Problem make_problem(size_t P, size_t N, size_t ny, size_t nx, unsigned long long seed) {
    Problem p{P, N, nx, ny, nx * ny, {}, {}, {}, {}, {}, {}, {}};
    const double R = 10e3, h = 5e3, ap = 2.0 * M_PI / 180.0;
    for (size_t k = 0; k < P; ++k) {
        const double th = ap * ((double)k / (double)(P - 1) - 0.5);
        p.tx.insert(p.tx.end(), {R * std::cos(th), R * std::sin(th), h});
        p.srp.insert(p.srp.end(), {0.0, 0.0, 0.0});
    }
    const double Ly = 0.25 * (double)(ny - 1), Lx = 0.25 * (double)(nx - 1);
    for (size_t iy = 0; iy < ny; ++iy)
        for (size_t ix = 0; ix < nx; ++ix) {
            const double y = ny > 1 ? -0.5 * Ly + Ly * (double)iy / (double)(ny - 1) : 0.0;
            const double x = nx > 1 ? -0.5 * Lx + Lx * (double)ix / (double)(nx - 1) : 0.0;
            p.pos.insert(p.pos.end(), {x, y, 0.0});
        }
    p.W.assign(2 * N, 0.0);
    for (size_t k = 0; k < N; ++k) p.W[2 * k] = 1.0;
    p.w.assign(2 * P, 0.0);
    for (size_t k = 0; k < P; ++k) p.w[2 * k] = 1.0;
    Lcg g{seed};
    p.amps.resize(2 * p.M);
    for (double &v : p.amps) v = g.next();
    p.ph.resize(2 * P * N);
    for (double &v : p.ph) v = g.next();
    return p;
}

std::vector<float> to_f32(const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); }

bool all_finite(const std::vector<float> &v) {
    for (float x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

// Forward and adjoint of a plan on the problem's data; outputs to host.
void run_plan(sarhip_plan_t plan, const Problem &p, std::vector<float> &Y, std::vector<float> &X) {
    const std::vector<float> a = to_f32(p.amps), y = to_f32(p.ph);
    void *da = nullptr, *dy = nullptr, *dY = nullptr, *dX = nullptr;
    HIP_OK(hipMalloc(&da, a.size() * 4));
    HIP_OK(hipMalloc(&dy, y.size() * 4));
    HIP_OK(hipMalloc(&dY, y.size() * 4));
    HIP_OK(hipMalloc(&dX, a.size() * 4));
    HIP_OK(hipMemcpy(da, a.data(), a.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dy, y.data(), y.size() * 4, hipMemcpyHostToDevice));
    for (int order = 0; order < 2; ++order) {
        CHECK(sarhip_forward(plan, da, dY, order, nullptr) == SARHIP_OK, "forward: %s",
              sarhip_last_error());
        CHECK(sarhip_adjoint(plan, dy, dX, order, nullptr) == SARHIP_OK, "adjoint: %s",
              sarhip_last_error());
    }
    HIP_OK(hipDeviceSynchronize());
    Y.resize(y.size());
    X.resize(a.size());
    HIP_OK(hipMemcpy(Y.data(), dY, Y.size() * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(X.data(), dX, X.size() * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(da));
    HIP_OK(hipFree(dy));
    HIP_OK(hipFree(dY));
    HIP_OK(hipFree(dX));
}

void plans(const Problem &p, int interp, int precision) {
    const double fc = 9.6e9, fs = 600e6, c = 299792458.0, U = 8.0;
    std::vector<float> Y[3], X[3];
    // host positions
    sarhip_plan_t pl = nullptr;
    CHECK(sarhip_plan_create(&pl, p.N, p.P, p.M, p.tx.data(), p.tx.data(), p.srp.data(),
                             p.pos.data(), p.W.data(), p.w.data(), fc, fs, c, U, 1.0, interp,
                             precision, 0) == SARHIP_OK,
          "plan_create: %s", sarhip_last_error());
    if (!pl) return;
    run_plan(pl, p, Y[0], X[0]);
    size_t nw = 0;
    double by_dx = 0, kc = 0;
    CHECK(sarhip_plan_info(pl, &nw, &by_dx, &kc) == SARHIP_OK && nw == (size_t)(U * p.N),
          "plan_info");
    int prec = -1, mono = -1, valid = -1, win = -1, G = -1, rt = -1, so = -1, fu = -1;
    double se = -1, bse = -1;
    CHECK(sarhip_plan_layout(pl, &prec, &mono, &valid, &win, &G, &se, &rt, &so, &bse, &fu) ==
              SARHIP_OK && mono == 1 && valid == (int)p.P,
          "plan_layout");
    // adaptive spectral stage: a threshold unless exact ranges (every row in
    // FP64) or the rocFFT path; the row count reads the plan's flags back
    double kappa = -1.0;
    int redone = -1;
    CHECK(sarhip_plan_forward_fp64_rows(pl, &kappa, &redone) == SARHIP_OK && kappa >= 0.0 &&
              redone >= 0 && redone <= (int)p.P && (precision != 0 || kappa == 0.0),
          "forward_fp64_rows");
    int pairs = -1;
    CHECK(sarhip_plan_gather_pairs(pl, &pairs) == SARHIP_OK && (pairs == 0 || pairs == 1) &&
              (pairs == 0 || (precision == 1 && interp == 2)),
          "gather_pairs");
    CHECK(sarhip_plan_set_timing(pl, 1) == SARHIP_OK, "set_timing");
    std::vector<float> Yt, Xt;
    run_plan(pl, p, Yt, Xt);
    float ms[4];
    int calls = 0;
    CHECK(sarhip_plan_timing(pl, 0, ms, &calls) == SARHIP_OK && calls == 2, "timing");
    CHECK(sarhip_plan_destroy(pl) == SARHIP_OK, "destroy");
    // device positions
    void *dpos = nullptr;
    HIP_OK(hipMalloc(&dpos, p.pos.size() * 8));
    HIP_OK(hipMemcpy(dpos, p.pos.data(), p.pos.size() * 8, hipMemcpyHostToDevice));
    pl = nullptr;
    CHECK(sarhip_plan_create_device_positions(&pl, p.N, p.P, p.M, p.tx.data(), p.tx.data(),
                                              p.srp.data(), dpos, p.W.data(), p.w.data(), fc, fs,
                                              c, U, 1.0, interp, precision, 0) == SARHIP_OK,
          "plan_create_device_positions: %s", sarhip_last_error());
    if (pl) {
        run_plan(pl, p, Y[1], X[1]);
        sarhip_plan_destroy(pl);
    }
    HIP_OK(hipFree(dpos));
    // grid description
    const size_t shape[3] = {1, p.ny, p.nx};
    const double dims[3] = {0.0, 0.25 * (double)(p.ny - 1), 0.25 * (double)(p.nx - 1)};
    const double centre[3] = {0, 0, 0}, rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    pl = nullptr;
    CHECK(sarhip_plan_create_grid(&pl, p.N, p.P, shape, dims, centre, rot, p.tx.data(),
                                  p.tx.data(), p.srp.data(), p.W.data(), p.w.data(), fc, fs, c, U,
                                  1.0, interp, precision, 0) == SARHIP_OK,
          "plan_create_grid: %s", sarhip_last_error());
    if (pl) {
        run_plan(pl, p, Y[2], X[2]);
        sarhip_plan_destroy(pl);
    }
    for (int s = 0; s < 3; ++s)
        CHECK(all_finite(Y[s]) && all_finite(X[s]), "non-finite output (source %d)", s);
    CHECK(Y[0] == Y[1] && Y[0] == Y[2] && X[0] == X[1] && X[0] == X[2],
          "position sources disagree (interp %d precision %d)", interp, precision);
    CHECK(Y[0] == Yt && X[0] == Xt, "timed call differs");
    std::printf("plans interp %d precision %d (M %zu, P %zu, N %zu): ok\n", interp, precision, p.M,
                p.P, p.N);
}

void record_entry_points(const Problem &p) {
    std::vector<double> ph(2 * p.P * p.N, 0.0), amps = p.amps;
    SarCalculationInfo info{p.N, p.P, p.M,
                            const_cast<double *>(p.tx.data()), const_cast<double *>(p.tx.data()),
                            const_cast<double *>(p.srp.data()), const_cast<double *>(p.pos.data()),
                            ph.data(), amps.data(), const_cast<double *>(p.W.data()),
                            const_cast<double *>(p.w.data()), 9.6e9, 600e6, 299792458.0, 16.0, 1.0};
    CHECK(forward_evaluate(&info) == SARHIP_OK, "forward_evaluate: %s", sarhip_last_error());
    info.phase_history = const_cast<double *>(p.ph.data());
    CHECK(adjoint_evaluate(&info) == SARHIP_OK, "adjoint_evaluate: %s", sarhip_last_error());
    bool fin = true;
    for (double v : ph) fin = fin && std::isfinite(v);
    for (double v : amps) fin = fin && std::isfinite(v);
    CHECK(fin, "record outputs not finite");
    std::printf("record entry points: ok\n");
}

void bad_arguments(const Problem &p) {
    sarhip_plan_t pl = nullptr;
    CHECK(sarhip_plan_create(&pl, 0, p.P, p.M, p.tx.data(), p.tx.data(), p.srp.data(),
                             p.pos.data(), p.W.data(), p.w.data(), 9.6e9, 600e6, 3e8, 8.0, 1.0, 2,
                             1, 0) != SARHIP_OK && std::strlen(sarhip_last_error()) > 0,
          "N = 0 accepted");
    CHECK(sarhip_plan_create(&pl, p.N, p.P, p.M, nullptr, p.tx.data(), p.srp.data(), p.pos.data(),
                             p.W.data(), p.w.data(), 9.6e9, 600e6, 3e8, 8.0, 1.0, 2, 1, 0) !=
              SARHIP_OK,
          "NULL positions accepted");
    CHECK(sarhip_plan_create(&pl, p.N, p.P, p.M, p.tx.data(), p.tx.data(), p.srp.data(),
                             p.pos.data(), p.W.data(), p.w.data(), 9.6e9, 600e6, 3e8, 8.0, 1.0, 7,
                             1, 0) != SARHIP_OK,
          "interp 7 accepted");
    CHECK(sarhip_forward(nullptr, nullptr, nullptr, 0, nullptr) != SARHIP_OK, "NULL plan accepted");
    SarCalculationInfo bad{};
    CHECK(forward_evaluate(&bad) != SARHIP_OK, "empty record accepted");
    CHECK(forward_evaluate(nullptr) != SARHIP_OK, "NULL record accepted");
    CHECK(sarhip_spectral_filter(nullptr, nullptr, 1, 8, 100, nullptr, nullptr, 0, nullptr,
                                 nullptr, nullptr) != 0,
          "bad spectral filter accepted");
    std::printf("bad arguments: rejected\n");
}

void resample_and_decode() {
    // resampling: 64 x 128 -> slow factor 2 (kernel of 5 taps), 64 samples
    const size_t P = 64, N = 128, N2 = 64;
    const int K = 5;
    std::vector<float> ph(2 * P * N);
    Lcg g{7};
    for (float &v : ph) v = (float)g.next();
    std::vector<double> dr(P, 0.01), win(N2, 1.0 / N), kern(2 * K, 0.0);
    for (int k = 0; k < K; ++k) kern[2 * k] = 0.2;
    std::vector<int> starts;
    for (size_t s = 0; s + K <= P; s += 2) starts.push_back((int)s);
    void *din = nullptr, *dout = nullptr;
    HIP_OK(hipMalloc(&din, ph.size() * 4));
    HIP_OK(hipMalloc(&dout, starts.size() * N2 * 8));
    HIP_OK(hipMemcpy(din, ph.data(), ph.size() * 4, hipMemcpyHostToDevice));
    CHECK(sarhip_resample(din, P, N, dr.data(), 9.6e9, 600e6, 1.0, 299792458.0, N2, win.data(),
                          kern.data(), K, starts.data(), starts.size(), dout, nullptr) == SARHIP_OK,
          "resample: %s", sarhip_last_error());
    std::vector<float> out(starts.size() * N2 * 2);
    HIP_OK(hipMemcpy(out.data(), dout, out.size() * 4, hipMemcpyDeviceToHost));
    CHECK(all_finite(out), "resample output not finite");
    HIP_OK(hipFree(din));
    HIP_OK(hipFree(dout));
    // decode: 3 vectors x 5 samples of big-endian CI2 (int8 pairs)
    std::vector<signed char> raw(3 * 5 * 2);
    for (size_t i = 0; i < raw.size(); ++i) raw[i] = (signed char)(i - 10);
    void *draw = nullptr, *ddec = nullptr;
    HIP_OK(hipMalloc(&draw, raw.size()));
    HIP_OK(hipMalloc(&ddec, 3 * 5 * 8));
    HIP_OK(hipMemcpy(draw, raw.data(), raw.size(), hipMemcpyHostToDevice));
    CHECK(sarhip_decode_signal(draw, 2, 3, 5, nullptr, ddec, nullptr) == SARHIP_OK, "decode: %s",
          sarhip_last_error());
    std::vector<float> dec(3 * 5 * 2);
    HIP_OK(hipMemcpy(dec.data(), ddec, dec.size() * 4, hipMemcpyDeviceToHost));
    bool same = true;
    for (size_t i = 0; i < dec.size(); ++i) same = same && dec[i] == (float)raw[i];
    CHECK(same, "CI2 decode mismatch");
    HIP_OK(hipFree(draw));
    HIP_OK(hipFree(ddec));
    std::printf("resample and decode: ok\n");
}

}  // namespace

int main() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        std::fprintf(stderr, "no GPU\n");
        return 3;
    }
    std::printf("libsarsolver_hip %s\n", sarhip_version());
    const Problem a = make_problem(64, 128, 48, 40, 11);   // fused spectral stages (N pow2)
    const Problem b = make_problem(37, 33, 21, 19, 12);    // rocFFT stages, odd sizes
    for (const Problem *p : {&a, &b}) {
        plans(*p, SARHIP_INTERP_NEAREST_REF, SARHIP_PRECISION_FP64);
        plans(*p, SARHIP_INTERP_LINEAR, SARHIP_PRECISION_SERIES);
        plans(*p, SARHIP_INTERP_NEAREST, SARHIP_PRECISION_SERIES);
        plans(*p, SARHIP_INTERP_LINEAR, SARHIP_PRECISION_HILO);
    }
    record_entry_points(b);
    bad_arguments(a);
    resample_and_decode();
    if (g_fail) {
        std::fprintf(stderr, "%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("all checks passed\n");
    return 0;
}
