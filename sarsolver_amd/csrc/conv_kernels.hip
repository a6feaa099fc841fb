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

#include <string>

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

thread_local std::string g_conv_err;

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

}  // extern "C"
