/*
 * sar_oracle.c -- FP64 CPU restatement of the reference SAR operators.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it, and never as the thing measured or shipped.
 *
 * It follows the reference native core src/sarsolver/cxx_binding/sarsolver_cxx.cpp
 * (cited as CPP:<line>) without FFTW, which is absent here: the FFT is this
 * file's own (iterative radix-2 for powers of two, Bluestein otherwise).
 *
 * Pins (see tests/test_oracle_*.py):
 *   - FFT: numpy.fft at powers of two and at 2^7*251 (Bluestein path).
 *   - Operators, all three interpolation modes: a brute-force direct sum with
 *     no FFT (oracle/brute.py) on tiny inputs.
 *   - Reference adjoint: the reference's own fixture little_backproj.npy
 *     (A^H of little_noisy_meas.npy at U = 128), elementwise rel 1e-3.
 *   - Reference forward + adjoint together: a 3-iteration PDHG against
 *     little_soln.npy and a power iteration against the reference's preset
 *     operator norm 756.6231381576064.
 *   - Bin rule: hand-built ties (u = k +- 1/2) with exact arithmetic.
 *   - Closed forms: scatterer at the SRP gives a flat spectrum; (A^H A e_m)_m.
 *   - Blocking invariance over (threads, pulse blocks).
 */
#include "sar_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* FFT (unnormalised; sign = -1 forward e^{-i}, +1 backward e^{+i}), matching
 * FFTW's FFTW_FORWARD / FFTW_BACKWARD conventions used at CPP:244-251.      */
/* ------------------------------------------------------------------------ */

typedef struct {
    size_t n;       /* transform length                                    */
    int sign;       /* -1 or +1                                            */
    int pow2;       /* 1 if n is a power of two                            */
    double *tw;     /* pow2: n/2 twiddles (interleaved)                    */
    size_t m;       /* Bluestein: padded power-of-two length               */
    double *chirp;  /* Bluestein: c_k = e^{sign*pi*i*k^2/n}, k<n           */
    double *bfft;   /* Bluestein: FFT_m of conj-chirp kernel               */
    double *work;   /* Bluestein: m complex work                           */
    struct oracle_fft_plan *sub_fwd, *sub_bwd; /* length-m pow2 plans     */
} fft_plan_t;

static int is_pow2(size_t n) { return n && !(n & (n - 1)); }

static void fft_pow2_exec(const fft_plan_t *p, double *x) {
    size_t n = p->n;
    /* bit reversal */
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j];
            x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr;
            x[2 * j + 1] = ti;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1, step = n / len;
        for (size_t s = 0; s < n; s += len) {
            for (size_t k = 0; k < half; ++k) {
                double wr = p->tw[2 * k * step], wi = p->tw[2 * k * step + 1];
                double *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                double br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br;
                b[1] = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
        }
    }
}

static fft_plan_t *fft_plan_create(size_t n, int sign);

static void fft_plan_free(fft_plan_t *p) {
    if (!p) return;
    free(p->tw);
    free(p->chirp);
    free(p->bfft);
    free(p->work);
    fft_plan_free((fft_plan_t *)p->sub_fwd);
    fft_plan_free((fft_plan_t *)p->sub_bwd);
    free(p);
}

static void fft_exec(const fft_plan_t *p, double *x);

static fft_plan_t *fft_plan_create(size_t n, int sign) {
    fft_plan_t *p = (fft_plan_t *)calloc(1, sizeof(fft_plan_t));
    if (!p) return NULL;
    p->n = n;
    p->sign = sign;
    p->pow2 = is_pow2(n);
    if (p->pow2) {
        size_t h = n / 2 ? n / 2 : 1;
        p->tw = (double *)malloc(sizeof(double) * 2 * h);
        if (!p->tw) { fft_plan_free(p); return NULL; }
        for (size_t k = 0; k < h; ++k) {
            double ang = 2.0 * M_PI * (double)k / (double)n;
            p->tw[2 * k] = cos(ang);
            p->tw[2 * k + 1] = sign * sin(ang);
        }
        return p;
    }
    /* Bluestein: X_j = c_j sum_k (x_k c_k) conj(c_{j-k}), c_k = e^{s pi i k^2/n} */
    size_t m = 1;
    while (m < 2 * n - 1) m <<= 1;
    p->m = m;
    p->chirp = (double *)malloc(sizeof(double) * 2 * n);
    p->bfft = (double *)calloc(2 * m, sizeof(double));
    p->work = (double *)malloc(sizeof(double) * 2 * m);
    p->sub_fwd = (struct oracle_fft_plan *)fft_plan_create(m, -1);
    p->sub_bwd = (struct oracle_fft_plan *)fft_plan_create(m, +1);
    if (!p->chirp || !p->bfft || !p->work || !p->sub_fwd || !p->sub_bwd) {
        fft_plan_free(p);
        return NULL;
    }
    uint64_t two_n = 2 * (uint64_t)n;
    for (size_t k = 0; k < n; ++k) {
        uint64_t k2 = ((uint64_t)k * (uint64_t)k) % two_n; /* exact argument reduction */
        double ang = M_PI * (double)k2 / (double)n;
        p->chirp[2 * k] = cos(ang);
        p->chirp[2 * k + 1] = sign * sin(ang);
    }
    /* kernel b_k = conj(c_k) for k in (-n, n), wrapped into length m */
    p->bfft[0] = p->chirp[0];
    p->bfft[1] = -p->chirp[1];
    for (size_t k = 1; k < n; ++k) {
        p->bfft[2 * k] = p->chirp[2 * k];
        p->bfft[2 * k + 1] = -p->chirp[2 * k + 1];
        p->bfft[2 * (m - k)] = p->chirp[2 * k];
        p->bfft[2 * (m - k) + 1] = -p->chirp[2 * k + 1];
    }
    fft_exec((fft_plan_t *)p->sub_fwd, p->bfft);
    return p;
}

/* Not re-entrant for Bluestein plans (shared work buffer): one plan per thread. */
static void fft_exec(const fft_plan_t *p, double *x) {
    if (p->n <= 1) return;
    if (p->pow2) {
        fft_pow2_exec(p, x);
        return;
    }
    size_t n = p->n, m = p->m;
    double *w = p->work;
    memset(w, 0, sizeof(double) * 2 * m);
    for (size_t k = 0; k < n; ++k) {
        double cr = p->chirp[2 * k], ci = p->chirp[2 * k + 1];
        w[2 * k] = x[2 * k] * cr - x[2 * k + 1] * ci;
        w[2 * k + 1] = x[2 * k] * ci + x[2 * k + 1] * cr;
    }
    fft_exec((fft_plan_t *)p->sub_fwd, w);
    for (size_t j = 0; j < m; ++j) {
        double ar = w[2 * j], ai = w[2 * j + 1];
        double br = p->bfft[2 * j], bi = p->bfft[2 * j + 1];
        w[2 * j] = ar * br - ai * bi;
        w[2 * j + 1] = ar * bi + ai * br;
    }
    fft_exec((fft_plan_t *)p->sub_bwd, w);
    double inv_m = 1.0 / (double)m;
    for (size_t j = 0; j < n; ++j) {
        double vr = w[2 * j] * inv_m, vi = w[2 * j + 1] * inv_m;
        double cr = p->chirp[2 * j], ci = p->chirp[2 * j + 1];
        x[2 * j] = vr * cr - vi * ci;
        x[2 * j + 1] = vr * ci + vi * cr;
    }
}

int oracle_fft(double *data, size_t n, int sign) {
    if (!data || n == 0 || (sign != 1 && sign != -1)) return -1;
    fft_plan_t *p = fft_plan_create(n, sign);
    if (!p) return -4;
    fft_exec(p, data);
    fft_plan_free(p);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Geometry and bin rules                                                    */
/* ------------------------------------------------------------------------ */

/* CPP:10-16: explicit three-term sum of squared differences, then sqrt.
 * Evaluated without FMA contraction (this file is built with
 * -ffp-contract=off), as the reference's x86-64 build evaluates it. */
static double distance3(const double *x, const double *y) {
    double dsq = 0.0;
    for (int a = 0; a < 3; ++a) {
        double d = x[a] - y[a];
        dsq += d * d;
    }
    return sqrt(dsq);
}

/* CPP:18-20 */
double oracle_bistatic_range(const double *tx, const double *rx, const double *x) {
    return distance3(tx, x) + distance3(rx, x);
}

/* CPP:22-24: ((a % b) + b) % b on signed 64-bit integers. */
static size_t math_modulo(long a, long b) { return (size_t)(((a % b) + b) % b); }

/* CPP:223: N_w = (size_t)(U * N), the product taken in double. */
size_t oracle_working_len(size_t n, double upsample_ratio) {
    return (size_t)(upsample_ratio * (double)n);
}

/* Bin index used at one pixel-pulse.
 *   nearest_ref forward  (CPP:289): mod(round(u), N_w)
 *   nearest_ref adjoint  (CPP:347): mod(round(0.5 + u), N_w)
 *   nearest (both ways)           : mod(round(u), N_w)
 *   linear                        : mod(floor(u), N_w), frac in *frac
 * round() is C's round-half-away-from-zero, as std::round. */
long oracle_bin(double dr, double by_dx, size_t nw, int interp, int adjoint, double *frac) {
    double u = dr * by_dx;
    long n;
    if (interp == ORACLE_LINEAR) {
        double f0 = floor(u);
        if (frac) *frac = u - f0;
        n = (long)f0;
    } else if (interp == ORACLE_NEAREST_REF && adjoint) {
        n = (long)round(0.5 + u);
    } else {
        n = (long)round(u);
    }
    return (long)math_modulo(n, (long)nw);
}

/* ------------------------------------------------------------------------ */
/* Operators                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    const SarCalculationInfo *info;
    int interp;
    size_t m0, m1;       /* pixel block [m0, m1)                          */
    double *ph_out;      /* forward: private [P,N] interleaved partial    */
    int status;
} worker_arg_t;

static void constants(const SarCalculationInfo *info, size_t *nw, double *by_dx, double *kc) {
    size_t N = info->num_fast_times;
    *nw = oracle_working_len(N, info->upsample_ratio);
    /* CPP:225-228 */
    *by_dx = (info->sample_frequency * (double)(*nw)) / ((double)N * info->c_eff);
    *kc = 2.0 * info->sign_multiplier * M_PI * info->centre_frequency / info->c_eff;
}

/* Forward for pixels [m0,m1) over all pulses: CPP:271-310. */
static void *forward_worker(void *argp) {
    worker_arg_t *arg = (worker_arg_t *)argp;
    const SarCalculationInfo *info = arg->info;
    size_t N = info->num_fast_times, P = info->num_slow_times;
    size_t nw;
    double by_dx, kc;
    constants(info, &nw, &by_dx, &kc);
    double *prof = (double *)calloc(2 * nw, sizeof(double));
    fft_plan_t *plan = fft_plan_create(nw, -1);
    if (!prof || !plan) {
        arg->status = -4;
        free(prof);
        fft_plan_free(plan);
        return NULL;
    }
    const double *W = info->waveform_fft, *w = info->slow_time_weighting;
    const double *amps = info->scattering_amplitudes;
    size_t half = N / 2;
    for (size_t p = 0; p < P; ++p) {
        const double *tx = info->transmit_posns + 3 * p, *rx = info->receive_posns + 3 * p;
        double r_srp = oracle_bistatic_range(tx, rx, info->stab_ref_posns + 3 * p);
        double wr = w[2 * p], wi = w[2 * p + 1];
        memset(prof, 0, sizeof(double) * 2 * nw);
        for (size_t m = arg->m0; m < arg->m1; ++m) {
            double dr = oracle_bistatic_range(tx, rx, info->scat_posns + 3 * m) - r_srp;
            double th = dr * kc;
            double er = cos(th), ei = sin(th);
            double ar = amps[2 * m], ai = amps[2 * m + 1];
            /* exp(i dR k_c) * a * w_p, left to right (CPP:288) */
            double t_r = er * ar - ei * ai, t_i = er * ai + ei * ar;
            double cr = t_r * wr - t_i * wi, ci = t_r * wi + t_i * wr;
            if (!isfinite(dr)) continue;
            if (arg->interp == ORACLE_LINEAR) {
                double f;
                long n0 = oracle_bin(dr, by_dx, nw, ORACLE_LINEAR, 0, &f);
                long n1 = (long)(((size_t)n0 + 1) % nw);
                prof[2 * n0] += (1.0 - f) * cr;
                prof[2 * n0 + 1] += (1.0 - f) * ci;
                prof[2 * n1] += f * cr;
                prof[2 * n1 + 1] += f * ci;
            } else {
                long n = oracle_bin(dr, by_dx, nw, arg->interp, 0, NULL);
                prof[2 * n] += cr;
                prof[2 * n + 1] += ci;
            }
        }
        fft_exec(plan, prof);
        /* CPP:295-307: keep N modes, multiply by the waveform spectrum */
        double *y = arg->ph_out + 2 * N * p;
        for (size_t i = 0; i < half; ++i) {
            size_t k = N - i - 1, g = nw - i - 1;
            y[2 * k] = W[2 * k] * prof[2 * g] - W[2 * k + 1] * prof[2 * g + 1];
            y[2 * k + 1] = W[2 * k] * prof[2 * g + 1] + W[2 * k + 1] * prof[2 * g];
            y[2 * i] = W[2 * i] * prof[2 * i] - W[2 * i + 1] * prof[2 * i + 1];
            y[2 * i + 1] = W[2 * i] * prof[2 * i + 1] + W[2 * i + 1] * prof[2 * i];
        }
        if (N % 2 == 1) {
            y[2 * half] = W[2 * half] * prof[2 * half] - W[2 * half + 1] * prof[2 * half + 1];
            y[2 * half + 1] = W[2 * half] * prof[2 * half + 1] + W[2 * half + 1] * prof[2 * half];
        }
    }
    free(prof);
    fft_plan_free(plan);
    arg->status = 0;
    return NULL;
}

/* Adjoint for pixels [m0,m1) over all pulses: CPP:319-353. */
static void *adjoint_worker(void *argp) {
    worker_arg_t *arg = (worker_arg_t *)argp;
    const SarCalculationInfo *info = arg->info;
    size_t N = info->num_fast_times, P = info->num_slow_times;
    size_t nw;
    double by_dx, kc;
    constants(info, &nw, &by_dx, &kc);
    double *modes = (double *)calloc(2 * nw, sizeof(double));
    fft_plan_t *plan = fft_plan_create(nw, +1);
    if (!modes || !plan) {
        arg->status = -4;
        free(modes);
        fft_plan_free(plan);
        return NULL;
    }
    const double *W = info->waveform_fft, *w = info->slow_time_weighting;
    double *out = info->scattering_amplitudes;
    size_t half = N / 2;
    int conj_w = arg->interp != ORACLE_NEAREST_REF; /* reference: CPP:349, w_p unconjugated */
    for (size_t m = arg->m0; m < arg->m1; ++m) out[2 * m] = out[2 * m + 1] = 0.0;
    for (size_t p = 0; p < P; ++p) {
        const double *tx = info->transmit_posns + 3 * p, *rx = info->receive_posns + 3 * p;
        double r_srp = oracle_bistatic_range(tx, rx, info->stab_ref_posns + 3 * p);
        const double *y = info->phase_history + 2 * N * p;
        memset(modes, 0, sizeof(double) * 2 * nw);
        /* CPP:331-342: conj(W) * Y placed at the FFT-order positions of N_w */
        for (size_t i = 0; i < half; ++i) {
            size_t k = N - i - 1, g = nw - i - 1;
            modes[2 * i] = W[2 * i] * y[2 * i] + W[2 * i + 1] * y[2 * i + 1];
            modes[2 * i + 1] = W[2 * i] * y[2 * i + 1] - W[2 * i + 1] * y[2 * i];
            modes[2 * g] = W[2 * k] * y[2 * k] + W[2 * k + 1] * y[2 * k + 1];
            modes[2 * g + 1] = W[2 * k] * y[2 * k + 1] - W[2 * k + 1] * y[2 * k];
        }
        if (N % 2 == 1) {
            modes[2 * half] = W[2 * half] * y[2 * half] + W[2 * half + 1] * y[2 * half + 1];
            modes[2 * half + 1] = W[2 * half] * y[2 * half + 1] - W[2 * half + 1] * y[2 * half];
        }
        fft_exec(plan, modes); /* CPP:344, unnormalised e^{+i} */
        double wr = w[2 * p], wi = conj_w ? -w[2 * p + 1] : w[2 * p + 1];
        for (size_t m = arg->m0; m < arg->m1; ++m) {
            double dr = oracle_bistatic_range(tx, rx, info->scat_posns + 3 * m) - r_srp;
            if (!isfinite(dr)) continue;
            double hr, hi;
            if (arg->interp == ORACLE_LINEAR) {
                double f;
                long n0 = oracle_bin(dr, by_dx, nw, ORACLE_LINEAR, 1, &f);
                long n1 = (long)(((size_t)n0 + 1) % nw);
                hr = (1.0 - f) * modes[2 * n0] + f * modes[2 * n1];
                hi = (1.0 - f) * modes[2 * n0 + 1] + f * modes[2 * n1 + 1];
            } else {
                long n = oracle_bin(dr, by_dx, nw, arg->interp, 1, NULL);
                hr = modes[2 * n];
                hi = modes[2 * n + 1];
            }
            double th = dr * kc;
            double er = cos(th), ei = -sin(th);
            /* h[n] * exp(-i dR k_c) * w_p (CPP:348-349) */
            double t_r = hr * er - hi * ei, t_i = hr * ei + hi * er;
            out[2 * m] += t_r * wr - t_i * wi;
            out[2 * m + 1] += t_r * wi + t_i * wr;
        }
    }
    free(modes);
    fft_plan_free(plan);
    arg->status = 0;
    return NULL;
}

static int check_info(const SarCalculationInfo *info, int interp) {
    if (!info || interp < 0 || interp > 2) return -1;
    if (!info->transmit_posns || !info->receive_posns || !info->stab_ref_posns ||
        !info->scat_posns || !info->phase_history || !info->scattering_amplitudes ||
        !info->waveform_fft || !info->slow_time_weighting)
        return -1;
    if (info->num_fast_times == 0) return -1;
    size_t nw = oracle_working_len(info->num_fast_times, info->upsample_ratio);
    if (nw < info->num_fast_times) return -1;
    return 0;
}

/* Scene-block decomposition of the reference's multi_*_evaluate
 * (forward.py:78-93, adjoint.py:76-91; block rule base/hypothesis.py:7-31):
 * T blocks of length M // T, the last taking the remainder; each block runs
 * the whole per-pulse pipeline (its own FFTs) on its own thread. */
static int run_blocks(SarCalculationInfo *info, int interp, int num_threads, int forward) {
    int rc = check_info(info, interp);
    if (rc) return rc;
    size_t M = info->num_scatterers, P = info->num_slow_times, N = info->num_fast_times;
    int T = num_threads < 1 ? 1 : num_threads;
    if ((size_t)T > M && M > 0) T = (int)M;
    if (M == 0) T = 1;
    worker_arg_t *args = (worker_arg_t *)calloc((size_t)T, sizeof(worker_arg_t));
    pthread_t *th = (pthread_t *)calloc((size_t)T, sizeof(pthread_t));
    if (!args || !th) { free(args); free(th); return -4; }
    size_t blen = M / (size_t)T;
    for (int t = 0; t < T; ++t) {
        args[t].info = info;
        args[t].interp = interp;
        args[t].m0 = blen * (size_t)t;
        args[t].m1 = (t == T - 1) ? M : blen * (size_t)(t + 1);
        args[t].status = -4;
        if (forward) {
            if (T == 1) {
                args[t].ph_out = info->phase_history;
            } else {
                args[t].ph_out = (double *)calloc(2 * N * P, sizeof(double));
                if (!args[t].ph_out) { rc = -4; }
            }
        }
    }
    if (rc == 0) {
        for (int t = 0; t < T; ++t)
            pthread_create(&th[t], NULL, forward ? forward_worker : adjoint_worker, &args[t]);
        for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
        for (int t = 0; t < T; ++t)
            if (args[t].status) rc = args[t].status;
    }
    if (forward && T > 1) {
        if (rc == 0) {
            /* sum of scene-block partials in block order (forward.py:93) */
            memset(info->phase_history, 0, sizeof(double) * 2 * N * P);
            for (int t = 0; t < T; ++t)
                for (size_t i = 0; i < 2 * N * P; ++i) info->phase_history[i] += args[t].ph_out[i];
        }
        for (int t = 0; t < T; ++t) free(args[t].ph_out);
    }
    free(args);
    free(th);
    return rc;
}

int oracle_forward(SarCalculationInfo *info, int interp, int num_threads) {
    return run_blocks(info, interp, num_threads, 1);
}

int oracle_adjoint(SarCalculationInfo *info, int interp, int num_threads) {
    return run_blocks(info, interp, num_threads, 0);
}

/* Per pixel-pulse differential ranges dR[p*M + m] (CPP:283,287 / 330,346). */
int oracle_differential_ranges(const SarCalculationInfo *info, double *dr_out) {
    if (!info || !dr_out) return -1;
    size_t M = info->num_scatterers, P = info->num_slow_times;
    for (size_t p = 0; p < P; ++p) {
        const double *tx = info->transmit_posns + 3 * p, *rx = info->receive_posns + 3 * p;
        double r_srp = oracle_bistatic_range(tx, rx, info->stab_ref_posns + 3 * p);
        for (size_t m = 0; m < M; ++m)
            dr_out[p * M + m] = oracle_bistatic_range(tx, rx, info->scat_posns + 3 * m) - r_srp;
    }
    return 0;
}
