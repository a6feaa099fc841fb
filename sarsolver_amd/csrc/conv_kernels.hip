// conv_kernels.hip -- 3x3 convolutions for the learned CG preconditioner (gfx950).
//
// Activations are NHWC bfloat16 with 32 channels (layers with fewer channels
// are zero-padded), weights bfloat16, accumulation FP32.
//
// conv3x3_fwd_kernel: implicit GEMM on the matrix cores.  M = output pixels,
// N = output channels, K = 9 taps x 32 input channels.  One workgroup owns a
// 16 x 16 output tile; its 18 x 18 input halo is staged in LDS; wave w
// computes rows 4w..4w+3 of the tile, each row one 16-pixel M block, for one
// or two 16-channel N blocks, with v_mfma_f32_16x16x32_bf16 (K = 32 = one tap).
// Operand layouts (checked on MI355X): lane l holds A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16], i < 8, and D[4(l/16)+i][l%16], i < 4.  Weights come
// pre-arranged in that B-fragment order ([tap][nblock][lane][8]) so each lane
// reads 16 contiguous bytes.  Epilogue: + bias, optional ReLU, bf16 or FP32 out.
//
// The data gradient of a layer is the same kernel run on the output gradient
// with weights flipped in space and transposed in channels.
//
// conv3x3_wgrad_kernel: dW[tap][ci][co] = sum_p x[p + off(tap)][ci] dy[p][co]
// and db[co] = sum_p dy[p][co], per 16 x 16 pixel tile in FP32 on the vector
// ALUs (training only), written as per-tile partials and summed in a fixed
// order by conv_wgrad_reduce_kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "sar_launch.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kC = 32;        // channels of the activation tensors
constexpr int kTile = 16;     // output tile side
constexpr int kHalo = kTile + 2;
constexpr int kPixStride = 40;  // bf16 per pixel in LDS (32 + 8 padding against bank conflicts)

__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
    return __uint_as_float((uint32_t)b << 16);
}

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}

template <int NB, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void conv3x3_fwd_kernel(const uint16_t *__restrict__ in,
                                                          const uint16_t *__restrict__ wfrag,
                                                          const float *__restrict__ bias, int H,
                                                          int W, void *__restrict__ out,
                                                          int out_c) {
    __shared__ __attribute__((aligned(16))) uint16_t tile[kHalo * kHalo * kPixStride];
    const int tx0 = blockIdx.x * kTile, ty0 = blockIdx.y * kTile, b = blockIdx.z;
    const uint16_t *img = in + (size_t)b * H * W * kC;
    // stage the 18 x 18 x 32 halo (16-byte chunks of 8 channels), zero outside the image
    for (int e = threadIdx.x; e < kHalo * kHalo * 4; e += 256) {
        const int pix = e >> 2, q = e & 3;
        const int hy = pix / kHalo, hx = pix - hy * kHalo;
        const int gy = ty0 + hy - 1, gx = tx0 + hx - 1;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (gy >= 0 && gy < H && gx >= 0 && gx < W)
            v = *(const uint4 *)(img + ((size_t)gy * W + gx) * kC + q * 8);
        *(uint4 *)(tile + pix * kPixStride + q * 8) = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;  // A row (pixel in block), K chunk
    f32x4 acc[4][NB];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[r][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int kh = tap / 3, kw = tap - 3 * (tap / 3);
        bf16x8 bw[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
            bw[nb] = *(const bf16x8 *)(wfrag + (((size_t)tap * NB + nb) * 64 + lane) * 8);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int hy = wave * 4 + r + kh, hx = m + kw;
            const bf16x8 a = *(const bf16x8 *)(tile + (hy * kHalo + hx) * kPixStride + kq * 8);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[nb], acc[r][nb], 0, 0, 0);
        }
    }
    // epilogue: lane holds D[px = 4*kq + i][ch = m] of each (row, nblock)
    if constexpr (NB == 2 && !OUT_F32) {
        if (out_c == kC) {
            // 32 bf16 channels: staged in LDS as [16 rows][16 px][32 ch] (the
            // halo is no longer read), then written as 16-byte chunks, each
            // output pixel one contiguous 64-byte run (2-byte stores of four
            // pixels 256 bytes apart took 6 of 15.6 us at 512 x 512)
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    const int ch = nb * 16 + m;
                    const float bb = bias[ch];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        float v = acc[r][nb][i] + bb;
                        if (RELU) v = fmaxf(v, 0.f);
                        tile[((wave * 4 + r) * kTile + 4 * kq + i) * kC + ch] = f32_to_bf16(v);
                    }
                }
            __syncthreads();
            for (int e = threadIdx.x; e < kTile * kTile * 4; e += 256) {
                const int pix = e >> 2, q = e & 3;
                const int gy = ty0 + (pix >> 4), gx = tx0 + (pix & 15);
                if (gy < H && gx < W)
                    *(uint4 *)((uint16_t *)out + ((size_t)b * H * W + (size_t)gy * W + gx) * kC + q * 8) =
                        *(const uint4 *)(tile + pix * kC + q * 8);
            }
            return;
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int gy = ty0 + wave * 4 + r;
        if (gy >= H) continue;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int ch = nb * 16 + m;
            const float bb = ch < out_c ? bias[ch] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int gx = tx0 + 4 * kq + i;
                if (gx >= W || ch >= out_c) continue;
                float v = acc[r][nb][i] + bb;
                if (RELU) v = fmaxf(v, 0.f);
                const size_t o = ((size_t)b * H * W + (size_t)gy * W + gx) * out_c + ch;
                if (OUT_F32) ((float *)out)[o] = v;
                else ((uint16_t *)out)[o] = f32_to_bf16(v);
            }
        }
    }
}

// Per 16x16 tile partial of dW[tap][ci][co] (9*32*CO floats) and db[co].
// x: [B,H,W,32] bf16 (layer input), dy: [B,H,W,CO] float (gradient w.r.t. the
// pre-activation output).  Thread t owns entries e = t + 256 j of the 9*32*CO
// weight gradient; each entry sums over the tile's 256 pixels from LDS.
template <int CO>
__global__ __launch_bounds__(256) void conv3x3_wgrad_kernel(const uint16_t *__restrict__ x,
                                                            const float *__restrict__ dy, int H,
                                                            int W, float *__restrict__ partial) {
    __shared__ uint16_t xs[kHalo * kHalo * kC];  // 20.7 KB (bf16)
    __shared__ float ds[kTile * kTile * CO];      // 32 KB for CO = 32
    const int tx0 = blockIdx.x * kTile, ty0 = blockIdx.y * kTile, b = blockIdx.z;
    const uint16_t *img = x + (size_t)b * H * W * kC;
    for (int e = threadIdx.x; e < kHalo * kHalo * kC; e += 256) {
        const int pix = e / kC, c = e - pix * kC;
        const int hy = pix / kHalo, hx = pix - hy * kHalo;
        const int gy = ty0 + hy - 1, gx = tx0 + hx - 1;
        xs[e] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? img[((size_t)gy * W + gx) * kC + c]
                                                          : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < kTile * kTile * CO; e += 256) {
        const int pix = e / CO, c = e - pix * CO;
        const int py = pix / kTile, px = pix - py * kTile;
        const int gy = ty0 + py, gx = tx0 + px;
        ds[e] = (gy < H && gx < W) ? dy[((size_t)b * H * W + (size_t)gy * W + gx) * CO + c] : 0.f;
    }
    __syncthreads();
    const size_t nblk = (size_t)gridDim.x * gridDim.y;
    float *outp = partial + ((size_t)b * nblk + (size_t)blockIdx.y * gridDim.x + blockIdx.x) *
                                (9 * kC * CO + CO);
    for (int e = threadIdx.x; e < 9 * kC * CO + CO; e += 256) {
        float s = 0.f;
        if (e < 9 * kC * CO) {
            const int tap = e / (kC * CO), rem = e - tap * kC * CO;
            const int ci = rem / CO, co = rem - ci * CO;
            const int kh = tap / 3, kw = tap - 3 * kh;
            for (int py = 0; py < kTile; ++py)
                for (int px = 0; px < kTile; ++px)
                    s = fmaf(bf16_to_f32(xs[((py + kh) * kHalo + px + kw) * kC + ci]),
                             ds[(py * kTile + px) * CO + co], s);
        } else {
            const int co = e - 9 * kC * CO;
            for (int p = 0; p < kTile * kTile; ++p) s += ds[p * CO + co];
        }
        outp[e] = s;
    }
}

// out[e] = sum over tiles t of partial[t][e], tiles in fixed order.
__global__ __launch_bounds__(256) void conv_wgrad_reduce_kernel(const float *__restrict__ partial,
                                                                size_t ntiles, int nelem,
                                                                float *__restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nelem) return;
    double s = 0.0;
    for (size_t t = 0; t < ntiles; ++t) s += partial[t * nelem + e];
    out[e] = (float)s;
}

// ---- preconditioner network inference in one call --------------------------
// Per image b of complex z [B][n]: s_b = max(sqrt(mean |z_b|^2), 1e-30) from
// FP64 partial sums in a fixed order; the network input is bf16(z / s) in
// channels 0 and 1 (others zero); the output is (z + s net(z / s)) times an
// optional per-pixel phasor.
constexpr int kRmsChunk = 4096;  // elements per partial sum

__global__ __launch_bounds__(256) void rms_partial_kernel(const float2 *__restrict__ z, size_t n,
                                                          int nchunk, double *__restrict__ part) {
    __shared__ double red[256];
    const int b = blockIdx.y, t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * kRmsChunk;
    const float2 *zb = z + (size_t)b * n;
    double acc = 0.0;
    for (int i = t; i < kRmsChunk; i += 256) {
        const size_t e = base + i;
        if (e < n) {
            const float2 v = zb[e];
            acc += (double)v.x * v.x + (double)v.y * v.y;
        }
    }
    red[t] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) part[(size_t)b * nchunk + blockIdx.x] = red[0];
}

__device__ __forceinline__ float image_rms(const double *part, int b, int nchunk, size_t n) {
    double acc = 0.0;
    for (int c = 0; c < nchunk; ++c) acc += part[(size_t)b * nchunk + c];
    return fmaxf((float)sqrt(acc / (double)n), 1e-30f);
}

// x [B][n][32] bf16: channels 0, 1 = z / s, the rest zero; s_out[b] = s.
__global__ __launch_bounds__(256) void precond_pack_kernel(const float2 *__restrict__ z, size_t n,
                                                           int nchunk, const double *__restrict__ part,
                                                           float *__restrict__ s_out,
                                                           uint16_t *__restrict__ x) {
    __shared__ float s_sh;
    const int b = blockIdx.y;
    if (threadIdx.x == 0) {
        s_sh = image_rms(part, b, nchunk, n);
        if (blockIdx.x == 0) s_out[b] = s_sh;
    }
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float s = s_sh;
    const float2 v = z[(size_t)b * n + i];
    const uint32_t w0 = (uint32_t)f32_to_bf16(v.x / s) | ((uint32_t)f32_to_bf16(v.y / s) << 16);
    uint4 *dst = (uint4 *)(x + ((size_t)b * n + i) * kC);
    dst[0] = make_uint4(w0, 0u, 0u, 0u);
    dst[1] = dst[2] = dst[3] = make_uint4(0u, 0u, 0u, 0u);
}

// out[b][i] = (z + s_b (y0 + i y1)) * demod[i] (demod optional), y [B][n][2] f32.
__global__ __launch_bounds__(256) void precond_unpack_kernel(const float2 *__restrict__ z,
                                                             const float *__restrict__ y,
                                                             const float *__restrict__ s_in,
                                                             const float2 *__restrict__ demod,
                                                             size_t n, float2 *__restrict__ out) {
    const int b = blockIdx.y;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t e = (size_t)b * n + i;
    const float s = s_in[b];
    const float2 v = z[e];
    const float2 yy = *(const float2 *)(y + 2 * e);
    float2 r = make_float2(fmaf(s, yy.x, v.x), fmaf(s, yy.y, v.y));
    if (demod) {
        const float2 d = demod[i];
        r = make_float2(fmaf(r.x, d.x, -r.y * d.y), fmaf(r.x, d.y, r.y * d.x));
    }
    out[e] = r;
}

struct NetWork {
    size_t act, y, part, s, total;  // byte offsets
};
NetWork net_workspace(int B, int H, int W) {
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t n = (size_t)H * W;
    const size_t nchunk = (n + kRmsChunk - 1) / kRmsChunk;
    NetWork w;
    w.act = 0;
    w.y = up(2 * (size_t)B * n * kC * sizeof(uint16_t));
    w.part = w.y + up((size_t)B * n * 2 * sizeof(float));
    w.s = w.part + up((size_t)B * nchunk * sizeof(double));
    w.total = w.s + up((size_t)B * sizeof(float));
    return w;
}

thread_local std::string g_conv_err;

// Static workspace of sarhip_precond_net_noalloc: net_workspace(1, 512, 512)
// = 35 652 352 bytes (two 32-channel bf16 activation buffers, the FP32 output
// pair, partial sums and scale), rounded up to 34 MiB + 4 KiB.  Each device
// that loads this library holds one copy for the life of the process; calls
// on one device share it.
constexpr size_t kNetScratchBytes = ((size_t)34 << 20) + ((size_t)4 << 10);
__device__ __align__(256) unsigned char scratch[kNetScratchBytes];

// Device address of `scratch` on the current device, looked up once per
// device with hipGetSymbolAddress (host code must use this address; the
// array's name in host code is only a stand-in for the symbol).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char *g_scratch_ptr[kMaxDevices] = {};

int net_scratch_ptr(char **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        g_conv_err = std::string("sarhip_precond_net_noalloc: no usable current device: ") +
                     (e != hipSuccess ? hipGetErrorString(e) : "index out of range");
        return -2;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_conv_err = std::string("sarhip_precond_net_noalloc: hipGetSymbolAddress: ") +
                         hipGetErrorString(e);
            return -2;
        }
        g_scratch_ptr[dev] = (char *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

const char *sarhip_conv_last_error(void) { return g_conv_err.c_str(); }

// y = act(conv3x3(x, W) + b); x [B,H,W,32] bf16, wfrag = B-fragment-ordered
// bf16 weights [9][NB][64][8] (NB = ceil(out_c / 16)), bias [out_c] f32,
// y [B,H,W,out_c] (bf16 or f32).  out_c <= 32.
int sarhip_conv3x3_bf16(const void *x, const void *wfrag, const float *bias, int B, int H, int W,
                        int out_c, int relu, int out_f32, void *y, void *stream) {
    if (!x || !wfrag || !bias || !y || B <= 0 || H <= 0 || W <= 0 || out_c <= 0 || out_c > 32) {
        g_conv_err = "sarhip_conv3x3_bf16: bad argument";
        return -1;
    }
    dim3 grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile, B);
    hipStream_t s = (hipStream_t)stream;
    const uint16_t *xi = (const uint16_t *)x, *wf = (const uint16_t *)wfrag;
#define SARHIP_CONV(NB, RL, OF) \
    conv3x3_fwd_kernel<NB, RL, OF><<<grid, 256, 0, s>>>(xi, wf, bias, H, W, y, out_c)
    const int nb = out_c > 16 ? 2 : 1;
    if (nb == 2) {
        if (relu) { if (out_f32) SARHIP_CONV(2, true, true); else SARHIP_CONV(2, true, false); }
        else { if (out_f32) SARHIP_CONV(2, false, true); else SARHIP_CONV(2, false, false); }
    } else {
        if (relu) { if (out_f32) SARHIP_CONV(1, true, true); else SARHIP_CONV(1, true, false); }
        else { if (out_f32) SARHIP_CONV(1, false, true); else SARHIP_CONV(1, false, false); }
    }
#undef SARHIP_CONV
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_conv_err = hipGetErrorString(e);
        return -2;
    }
    return 0;
}

// Workspace bytes for sarhip_conv3x3_wgrad.
size_t sarhip_conv3x3_wgrad_workspace(int B, int H, int W, int out_c) {
    const size_t tiles = (size_t)B * ((H + kTile - 1) / kTile) * ((W + kTile - 1) / kTile);
    const int co = out_c > 16 ? 32 : 16;
    return tiles * (9 * kC * co + co) * sizeof(float);
}

// dW [9][32][co] and db [co] (co = 16 or 32, the padded output width) of a
// layer with input x [B,H,W,32] bf16 and pre-activation gradient
// dy [B,H,W,co] f32; written to dwdb = [dW (9*32*co), db (co)].
int sarhip_conv3x3_wgrad(const void *x, const float *dy, int B, int H, int W, int out_c,
                         float *dwdb, void *workspace, size_t workspace_bytes, void *stream) {
    const int co = out_c > 16 ? 32 : 16;
    if (!x || !dy || !dwdb || !workspace || B <= 0 || H <= 0 || W <= 0 ||
        workspace_bytes < sarhip_conv3x3_wgrad_workspace(B, H, W, out_c)) {
        g_conv_err = "sarhip_conv3x3_wgrad: bad argument or workspace too small";
        return -1;
    }
    dim3 grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile, B);
    hipStream_t s = (hipStream_t)stream;
    if (co == 32)
        conv3x3_wgrad_kernel<32><<<grid, 256, 0, s>>>((const uint16_t *)x, dy, H, W,
                                                      (float *)workspace);
    else
        conv3x3_wgrad_kernel<16><<<grid, 256, 0, s>>>((const uint16_t *)x, dy, H, W,
                                                      (float *)workspace);
    const int nelem = 9 * kC * co + co;
    const size_t tiles = (size_t)grid.x * grid.y * grid.z;
    conv_wgrad_reduce_kernel<<<(nelem + 255) / 256, 256, 0, s>>>((const float *)workspace, tiles,
                                                                nelem, dwdb);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_conv_err = hipGetErrorString(e);
        return -2;
    }
    return 0;
}

size_t sarhip_precond_net_workspace(int B, int H, int W) {
    if (B <= 0 || H <= 0 || W <= 0) return 0;
    return net_workspace(B, H, W).total;
}

// Checks shared by the precond_net entry points; 0 or -1 with the message set.
static int precond_net_check(const void *z, int B, int H, int W, int layers,
                             const void *const *wfrags, const float *const *biases,
                             const int *out_c, const void *out, const char *name) {
    if (!z || !out || !wfrags || !biases || !out_c || B <= 0 || H <= 0 || W <= 0 ||
        layers < 1 || (size_t)B * H * W > ((size_t)1 << 31)) {
        g_conv_err = std::string(name) + ": bad argument";
        return -1;
    }
    for (int i = 0; i < layers; ++i) {
        if (!wfrags[i] || !biases[i] || out_c[i] != (i == layers - 1 ? 2 : kC)) {
            g_conv_err = std::string(name) + ": layer widths must be 32, ..., 32, 2";
            return -1;
        }
    }
    return 0;
}

// The network on device workspace `wk` laid out as net_workspace(B, H, W):
// every byte the kernels read there is written earlier in the same call, so
// the workspace needs no clearing between calls.
static int precond_net_run(const void *z, const void *demod, int B, int H, int W, int layers,
                           const void *const *wfrags, const float *const *biases,
                           const int *out_c, char *wk, void *out, void *stream) {
    const NetWork nw = net_workspace(B, H, W);
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)H * W;
    const int nchunk = (int)((n + kRmsChunk - 1) / kRmsChunk);
    uint16_t *act[2] = {(uint16_t *)(wk + nw.act), (uint16_t *)(wk + nw.act) + (size_t)B * n * kC};
    float *y = (float *)(wk + nw.y);
    double *part = (double *)(wk + nw.part);
    float *sv = (float *)(wk + nw.s);
    const float2 *zz = (const float2 *)z;
    rms_partial_kernel<<<dim3(nchunk, B), 256, 0, st>>>(zz, n, nchunk, part);
    const dim3 pgrid((unsigned)((n + 255) / 256), B);
    precond_pack_kernel<<<pgrid, 256, 0, st>>>(zz, n, nchunk, part, sv, act[0]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_conv_err = hipGetErrorString(e);
        return -2;
    }
    for (int i = 0; i < layers; ++i) {
        const bool last = i == layers - 1;
        void *dst = last ? (void *)y : (void *)act[(i + 1) & 1];
        const int rc = sarhip_conv3x3_bf16(act[i & 1], wfrags[i], biases[i], B, H, W, out_c[i],
                                           last ? 0 : 1, last ? 1 : 0, dst, stream);
        if (rc != 0) return rc;
    }
    precond_unpack_kernel<<<pgrid, 256, 0, st>>>(zz, y, sv, (const float2 *)demod, n,
                                                 (float2 *)out);
    e = hipGetLastError();
    if (e != hipSuccess) {
        g_conv_err = hipGetErrorString(e);
        return -2;
    }
    return 0;
}

size_t sarhip_spectral_filter_workspace(int B, int n, int m) {
    if (B <= 0 || n <= 0 || m <= 0) return 0;
    return sizeof(float2) * (size_t)B * n * m;
}

int sarhip_spectral_filter(const void *in, const void *demod_in, int B, int n, int m,
                           const float *filter, void *work, size_t work_bytes, void *out,
                           const void *demod_out, void *stream) {
    if (!in || !filter || !work || !out || B <= 0 || !sarhip::spectral_filter_supported(n, m) ||
        (size_t)B * n * m > ((size_t)1 << 31)) {
        g_conv_err = "sarhip_spectral_filter: bad argument (need n <= m, m a power of two in "
                     "[64, 1024])";
        return -1;
    }
    if (work_bytes < sarhip_spectral_filter_workspace(B, n, m)) {
        g_conv_err = "sarhip_spectral_filter: workspace too small";
        return -1;
    }
    const hipError_t e = sarhip::launch_spectral_filter(
        (const float2 *)in, (const float2 *)demod_in, B, n, m, filter, (float2 *)work,
        (float2 *)out, (const float2 *)demod_out, (hipStream_t)stream);
    if (e != hipSuccess) {
        g_conv_err = std::string("sarhip_spectral_filter: ") + hipGetErrorString(e);
        return -2;
    }
    return 0;
}

int sarhip_precond_net(const void *z, const void *demod, int B, int H, int W, int layers,
                       const void *const *wfrags, const float *const *biases, const int *out_c,
                       void *work, size_t work_bytes, void *out, void *stream) {
    int rc = precond_net_check(z, B, H, W, layers, wfrags, biases, out_c, out,
                               "sarhip_precond_net");
    if (rc) return rc;
    if (!work || work_bytes < net_workspace(B, H, W).total) {
        g_conv_err = "sarhip_precond_net: workspace missing or too small";
        return -1;
    }
    return precond_net_run(z, demod, B, H, W, layers, wfrags, biases, out_c, (char *)work, out,
                           stream);
}

int sarhip_precond_net_noalloc(const void *z, const void *demod, int B, int H, int W, int layers,
                               const void *const *wfrags, const float *const *biases,
                               const int *out_c, void *out, void *stream) {
    int rc = precond_net_check(z, B, H, W, layers, wfrags, biases, out_c, out,
                               "sarhip_precond_net_noalloc");
    if (rc) return rc;
    const size_t need = net_workspace(B, H, W).total;
    if (need > kNetScratchBytes) {
        g_conv_err = "sarhip_precond_net_noalloc: B x H x W needs " + std::to_string(need) +
                     " bytes of workspace, above the static scratch of " +
                     std::to_string(kNetScratchBytes) +
                     " bytes (one 512 x 512 image); use sarhip_precond_net with a workspace";
        return -1;
    }
    char *wk = nullptr;
    if ((rc = net_scratch_ptr(&wk))) return rc;
    return precond_net_run(z, demod, B, H, W, layers, wfrags, biases, out_c, wk, out, stream);
}

}  // extern "C"
