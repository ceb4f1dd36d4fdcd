// Denoiser kernels for gfx950: 3x3 convolutions on MFMA (bf16 inputs, fp32
// accumulation), their gradients, and the U-Net's glue (packing, pooling,
// upsampling, loss, Adam).
//
// No reference counterpart: the reference uses NVIDIA NRD (RELAX, closed) in its
// DirectX backend (DirectX/Denoiser.cpp:133).  This denoiser is new: a U-Net over
// noisy radiance plus the albedo / normal / depth guides of temporal.hip.
//
// Layouts: activations NHWC bf16 with the channel count padded to a multiple of
// 32 (channel stride S); convolution weights bf16 [tap][co][ci] (tap = 3*ky+kx,
// co padded to a multiple of 32, ci to the input stride); fp32 master weights
// and Adam moments in the same order.
//
// MFMA v_mfma_f32_16x16x32_bf16 computes D[16x16] += A[16x32] B[32x16]: lane l
// supplies A row (l % 16), B column (l % 16), each with the 8 consecutive k of
// group (l / 16); D lands as rows 4*(l/16)+i, column (l % 16), i = 0..3.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "denoise.h"

namespace aph {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 toBf(float f) { return (__bf16)f; }
__device__ __forceinline__ float fromBf(__bf16 b) { return (float)b; }

__device__ __forceinline__ bf16x8 load8(const __bf16* p)
{
    return *reinterpret_cast<const bf16x8*>(p);
}

// ---- convolution (forward, and data gradient with flipped weights) ----
// One wave: 16 consecutive pixels x up to 8 output tiles of 16 channels.
// Block: 4 waves (64 pixels); blockIdx.y selects a group of 128 output channels.
template <bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void kConv3x3(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (p0 >= npix) return;
    const int coBase = blockIdx.y * 128;
    const int nTiles = min(8, (a.coutP - coBase) / 16);
    const int row = lane & 15, kg = lane >> 4;
    // the pixel this lane feeds into A
    const int64_t pa = p0 + row;
    const bool paValid = pa < npix;
    int n = 0, y = 0, x = 0;
    if (paValid) {
        n = (int)(pa / ((int64_t)a.H * a.W));
        int r = (int)(pa - (int64_t)n * a.H * a.W);
        y = r / a.W;
        x = r - y * a.W;
    }
    f32x4 acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int t = 0; t < 9; t++) {
        const int sy = y + t / 3 - 1, sx = x + t % 3 - 1;
        const bool ok = paValid && sy >= 0 && sy < a.H && sx >= 0 && sx < a.W;
        const __bf16* src = a.in + (((int64_t)n * a.H + sy) * a.W + sx) * a.inStride + kg * 8;
        const __bf16* wt = a.w + ((int64_t)t * a.coutP + coBase + row) * a.cinP + kg * 8;
        for (int c0 = 0; c0 < a.cinP; c0 += 32) {
            bf16x8 av = ok ? load8(src + c0) : bf16x8{};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (j < nTiles) {
                    bf16x8 bv = load8(wt + (int64_t)j * 16 * a.cinP + c0);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[j], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
        if (j >= nTiles) continue;
        const int co = coBase + j * 16 + row;
        const float b = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t q = p0 + kg * 4 + i;
            if (q >= npix || co >= a.coutStore) continue;
            float v = acc[j][i] + b;
            if (RELU) v = fmaxf(v, 0.0f);
            if (OUT_F32) a.outF[q * a.outStride + a.outOffset + co] = v;
            else a.out[q * a.outStride + a.outOffset + co] = toBf(v);
        }
    }
}

// ---- weight gradient: dW[t][co][ci] += sum_p dY[p][co] X[p + off(t)][ci] ----
// One wave per (16 co x 16 ci tile, tap, pixel chunk); K = pixels, 32 per MFMA,
// fragments gathered (8 pixels per lane); partial sums added atomically.
__global__ __launch_bounds__(64) void kConvWGrad(WGradArgs a)
{
    const int lane = threadIdx.x, row = lane & 15, kg = lane >> 4;
    const int t = blockIdx.y;
    const int tilesCi = a.cinP / 16;
    const int coT = blockIdx.z / tilesCi, ciT = blockIdx.z % tilesCi;
    const int co = coT * 16 + row, ci = ciT * 16 + row;
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    const int64_t c0 = (int64_t)blockIdx.x * a.chunk, c1 = min(npix, c0 + a.chunk);
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int64_t p = c0; p < c1; p += 32) {
        bf16x8 av, bv;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int64_t q = p + kg * 8 + e;
            float gy = 0.0f, xv = 0.0f;
            if (q < c1) {
                gy = fromBf(a.dy[q * a.dyStride + co]);
                int n = (int)(q / ((int64_t)a.H * a.W));
                int r = (int)(q - (int64_t)n * a.H * a.W);
                int y = r / a.W, x = r - (r / a.W) * a.W;
                int sy = y + dy, sx = x + dx;
                if (sy >= 0 && sy < a.H && sx >= 0 && sx < a.W)
                    xv = fromBf(a.x[(((int64_t)n * a.H + sy) * a.W + sx) * a.xStride + ci]);
            }
            av[e] = toBf(gy);
            bv[e] = toBf(xv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    // D rows are co (A rows), columns ci (B columns)
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int dco = coT * 16 + kg * 4 + i, dci = ciT * 16 + row;
        atomicAdd(&a.dw[((int64_t)t * a.coutP + dco) * a.cinP + dci], acc[i]);
    }
}

__global__ void kBiasGrad(const __bf16* dy, int64_t npix, int stride, int coutP, float* db)
{
    const int co = blockIdx.y * blockDim.x + threadIdx.x;
    if (co >= coutP) return;
    float s = 0.0f;
    for (int64_t p = blockIdx.x; p < npix; p += gridDim.x) s += fromBf(dy[p * stride + co]);
    atomicAdd(&db[co], s);
}

// dY *= (Y > 0) for a ReLU output Y: channels [0, C) of two strided buffers
__global__ void kReluMask(__bf16* dy, int dyStride, const __bf16* y, int yStride, int64_t npix, int C)
{
    const int64_t n = npix * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int c = (int)(i % C);
        int64_t p = i / C;
        if (!(fromBf(y[p * yStride + c]) > 0.0f)) dy[p * dyStride + c] = toBf(0.0f);
    }
}

__global__ void kBf16ToF32(const __bf16* in, float* out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = fromBf(in[i]);
}

// ---- pooling and upsampling (2x2, NHWC; channel slices through offsets) ----
__global__ void kAvgPool2(const __bf16* in, int inStride, int inOff, __bf16* out, int outStride, int outOff, int N,
                          int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = (int64_t)N * Ho * Wo * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int c = (int)(i % C);
        int64_t p = i / C;
        int xo = (int)(p % Wo), yo = (int)((p / Wo) % Ho), b = (int)(p / ((int64_t)Wo * Ho));
        float s = 0.0f;
        for (int k = 0; k < 4; k++) {
            int yy = 2 * yo + k / 2, xx = 2 * xo + k % 2;
            s += fromBf(in[(((int64_t)b * H + yy) * W + xx) * inStride + inOff + c]);
        }
        out[p * outStride + outOff + c] = toBf(0.25f * s);
    }
}

__global__ void kAvgPool2Back(const __bf16* dout, int doStride, __bf16* din, int diStride, int diOff, int N, int H,
                              int W, int C, int accumulate)
{
    const int64_t n = (int64_t)N * H * W * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int c = (int)(i % C);
        int64_t p = i / C;
        int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
        float g = 0.25f * fromBf(dout[(((int64_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * doStride + c]);
        __bf16* d = &din[p * diStride + diOff + c];
        *d = toBf(accumulate ? fromBf(*d) + g : g);
    }
}

__global__ void kUpsample2(const __bf16* in, int inStride, __bf16* out, int outStride, int outOff, int N, int Ho,
                           int Wo, int C)
{
    const int64_t n = (int64_t)N * Ho * Wo * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int c = (int)(i % C);
        int64_t p = i / C;
        int x = (int)(p % Wo), y = (int)((p / Wo) % Ho), b = (int)(p / ((int64_t)Wo * Ho));
        out[p * outStride + outOff + c] = in[(((int64_t)b * (Ho / 2) + y / 2) * (Wo / 2) + x / 2) * inStride + c];
    }
}

__global__ void kUpsample2Back(const __bf16* dout, int doStride, int doOff, __bf16* din, int diStride, int N, int Hi,
                               int Wi, int C)
{
    const int64_t n = (int64_t)N * Hi * Wi * C;
    const int Ho = 2 * Hi, Wo = 2 * Wi;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int c = (int)(i % C);
        int64_t p = i / C;
        int x = (int)(p % Wi), y = (int)((p / Wi) % Hi), b = (int)(p / ((int64_t)Wi * Hi));
        float s = 0.0f;
        for (int k = 0; k < 4; k++) {
            int yy = 2 * y + k / 2, xx = 2 * x + k % 2;
            s += fromBf(dout[(((int64_t)b * Ho + yy) * Wo + xx) * doStride + doOff + c]);
        }
        din[p * diStride + c] = toBf(s);
    }
}

// Source pixel of destination (x, y) under a training transform: bit 2
// transposes (square images), bit 0 mirrors x, bit 1 mirrors y.
__device__ __forceinline__ void sourcePixel(int x, int y, int W, int H, uint32_t xf, int& sx, int& sy)
{
    if (xf & 4u) {
        int t = x;
        x = y;
        y = t;
    }
    sx = (xf & 1u) ? W - 1 - x : x;
    sy = (xf & 2u) ? H - 1 - y : y;
}

// ---- input packing: log1p radiance, albedo, metalness, normal, roughness, depth ----
__global__ void kPack(const float4* noisy, const float* depth, const float4* normalRough, const float4* albedo,
                      int W, int H, int Wp, int Hp, int img, uint32_t xf, __bf16* out, int stride)
{
    const int64_t n = (int64_t)Wp * Hp;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(i % Wp), y = (int)(i / Wp);
        __bf16* o = out + ((int64_t)img * n + i) * stride;
        float f[kDnInChannels] = {};
        if (x < W && y < H) {
            int sx, sy;
            sourcePixel(x, y, W, H, xf, sx, sy);
            int64_t s = (int64_t)sy * W + sx;
            float4 c = noisy[s], nr = normalRough[s], al = albedo[s];
            f[0] = log1pf(fmaxf(c.x, 0.0f));
            f[1] = log1pf(fmaxf(c.y, 0.0f));
            f[2] = log1pf(fmaxf(c.z, 0.0f));
            f[3] = al.x;
            f[4] = al.y;
            f[5] = al.z;
            f[6] = al.w;
            f[7] = nr.x;
            f[8] = nr.y;
            f[9] = nr.z;
            f[10] = nr.w;
            f[11] = log1pf(fabsf(depth[s])) * 0.2f;
#pragma unroll
            for (int k = 0; k < kDnInChannels; k++)
                if (!(f[k] == f[k]) || fabsf(f[k]) > 1e30f) f[k] = 0.0f;
        }
        for (int c = 0; c < stride; c++) o[c] = toBf(c < kDnInChannels ? f[c] : 0.0f);
    }
}

// ---- output: expm1 of the predicted log radiance, alpha from the noisy input ----
__global__ void kUnpack(const float* pred, int W, int H, int Wp, int Hp, int img, const float4* noisy, float4* out)
{
    const int64_t n = (int64_t)W * H;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(i % W), y = (int)(i / W);
        const float* p = pred + ((int64_t)img * Wp * Hp + (int64_t)y * Wp + x) * kDnPredStride;
        out[i] = make_float4(expm1f(fmaxf(p[0], 0.0f)), expm1f(fmaxf(p[1], 0.0f)), expm1f(fmaxf(p[2], 0.0f)),
                             noisy[i].w);
    }
}

// ---- loss against the reference and its gradient, mean over the unpadded
// pixels, 3 channels and nImg images.  With p the predicted log1p radiance and
// q = log1p(reference): L1 of p - q, or L2 of the tonemapped values
// x / (1 + x) = 1 - exp(-p) ----
__global__ void kLoss(const float* pred, const float4* ref, int W, int H, int Wp, int Hp, int img, int nImg,
                      uint32_t xf, int kind, __bf16* dpred, int dStride, float* loss)
{
    const int64_t n = (int64_t)Wp * Hp;
    const float scale = 1.0f / (3.0f * (float)W * (float)H * (float)nImg);
    float part = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(i % Wp), y = (int)(i / Wp);
        const int64_t q = (int64_t)img * n + i;
        float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const bool inside = x < W && y < H;
        if (inside) {
            int sx, sy;
            sourcePixel(x, y, W, H, xf, sx, sy);
            r = ref[(int64_t)sy * W + sx];
        }
        for (int c = 0; c < dStride; c++) {
            float g = 0.0f;
            if (c < 3 && inside) {
                float rv = log1pf(fmaxf(c == 0 ? r.x : (c == 1 ? r.y : r.z), 0.0f));
                float p = pred[q * kDnPredStride + c];
                if (kind == kDnLossL2Tonemapped) {
                    float ep = expf(-p);
                    float d = expf(-rv) - ep;  // (1 - e^-p) - (1 - e^-q)
                    part += d * d * scale;
                    g = 2.0f * d * ep * scale;
                } else {
                    float d = p - rv;
                    part += fabsf(d) * scale;
                    g = (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f)) * scale;
                }
            }
            dpred[q * dStride + c] = toBf(g);
        }
    }
    // wave sum, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, part);
}

// ---- Adam on the fp32 master weights; refreshes the bf16 forward weights and
// the flipped, transposed copy the data gradient convolves with ----
__global__ void kAdam(float* w, float* m, float* v, const float* g, int64_t n, float lr, float b1, float b2, float eps,
                      float c1, float c2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float gi = g[i];
        float mi = b1 * m[i] + (1.0f - b1) * gi;
        float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void kConvWeights(const float* w, int coutP, int cinP, __bf16* wf, __bf16* wt)
{
    const int64_t n = (int64_t)9 * coutP * cinP;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int ci = (int)(i % cinP);
        int co = (int)((i / cinP) % coutP);
        int t = (int)(i / ((int64_t)cinP * coutP));
        __bf16 b = toBf(w[i]);
        wf[i] = b;
        // data gradient: tap 8 - t, input channels = co, output channels = ci
        wt[((int64_t)(8 - t) * cinP + ci) * coutP + co] = b;
    }
}

unsigned gridFor(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 4096); }

}  // namespace

hipError_t launchConv(const ConvArgs& a, bool relu, bool outF32, hipStream_t st)
{
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    if (npix == 0) return hipSuccess;
    dim3 grid((unsigned)((npix + 63) / 64), (unsigned)((a.coutP + 127) / 128));
    if (outF32) {
        if (relu) kConv3x3<true, true><<<grid, 256, 0, st>>>(a);
        else kConv3x3<false, true><<<grid, 256, 0, st>>>(a);
    } else {
        if (relu) kConv3x3<true, false><<<grid, 256, 0, st>>>(a);
        else kConv3x3<false, false><<<grid, 256, 0, st>>>(a);
    }
    return hipGetLastError();
}

hipError_t launchWGrad(const WGradArgs& a, hipStream_t st)
{
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    dim3 grid((unsigned)((npix + a.chunk - 1) / a.chunk), 9, (unsigned)((a.coutP / 16) * (a.cinP / 16)));
    kConvWGrad<<<grid, 64, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launchBiasGrad(const __bf16* dy, int64_t npix, int stride, int coutP, float* db, hipStream_t st)
{
    dim3 grid((unsigned)std::min<int64_t>(npix, 1024), (unsigned)((coutP + 63) / 64));
    kBiasGrad<<<grid, 64, 0, st>>>(dy, npix, stride, coutP, db);
    return hipGetLastError();
}

hipError_t launchReluMask(__bf16* dy, int dyStride, const __bf16* y, int yStride, int64_t npix, int C, hipStream_t st)
{
    kReluMask<<<gridFor(npix * C), 256, 0, st>>>(dy, dyStride, y, yStride, npix, C);
    return hipGetLastError();
}

hipError_t launchBf16ToF32(const __bf16* in, float* out, int64_t n, hipStream_t st)
{
    kBf16ToF32<<<gridFor(n), 256, 0, st>>>(in, out, n);
    return hipGetLastError();
}

hipError_t launchAvgPool(const __bf16* in, int inStride, int inOff, __bf16* out, int outStride, int outOff, int N,
                         int H, int W, int C, hipStream_t st)
{
    kAvgPool2<<<gridFor((int64_t)N * H * W * C / 4), 256, 0, st>>>(in, inStride, inOff, out, outStride, outOff, N, H,
                                                                   W, C);
    return hipGetLastError();
}

hipError_t launchAvgPoolBack(const __bf16* dout, int doStride, __bf16* din, int diStride, int diOff, int N, int H,
                             int W, int C, bool accumulate, hipStream_t st)
{
    kAvgPool2Back<<<gridFor((int64_t)N * H * W * C), 256, 0, st>>>(dout, doStride, din, diStride, diOff, N, H, W, C,
                                                                   accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t launchUpsample(const __bf16* in, int inStride, __bf16* out, int outStride, int outOff, int N, int Ho,
                          int Wo, int C, hipStream_t st)
{
    kUpsample2<<<gridFor((int64_t)N * Ho * Wo * C), 256, 0, st>>>(in, inStride, out, outStride, outOff, N, Ho, Wo, C);
    return hipGetLastError();
}

hipError_t launchUpsampleBack(const __bf16* dout, int doStride, int doOff, __bf16* din, int diStride, int N, int Hi,
                              int Wi, int C, hipStream_t st)
{
    kUpsample2Back<<<gridFor((int64_t)N * Hi * Wi * C), 256, 0, st>>>(dout, doStride, doOff, din, diStride, N, Hi, Wi,
                                                                      C);
    return hipGetLastError();
}

hipError_t launchPack(const float4* noisy, const float* depth, const float4* normalRough, const float4* albedo, int W,
                      int H, int Wp, int Hp, int img, uint32_t transform, __bf16* out, int stride, hipStream_t st)
{
    kPack<<<gridFor((int64_t)Wp * Hp), 256, 0, st>>>(noisy, depth, normalRough, albedo, W, H, Wp, Hp, img, transform,
                                                     out, stride);
    return hipGetLastError();
}

hipError_t launchUnpack(const float* pred, int W, int H, int Wp, int Hp, int img, const float4* noisy, float4* out,
                        hipStream_t st)
{
    kUnpack<<<gridFor((int64_t)W * H), 256, 0, st>>>(pred, W, H, Wp, Hp, img, noisy, out);
    return hipGetLastError();
}

hipError_t launchLoss(const float* pred, const float4* ref, int W, int H, int Wp, int Hp, int img, int nImg,
                      uint32_t transform, int kind, __bf16* dpred, int dStride, float* loss, hipStream_t st)
{
    kLoss<<<gridFor((int64_t)Wp * Hp), 256, 0, st>>>(pred, ref, W, H, Wp, Hp, img, nImg, transform, kind, dpred,
                                                     dStride, loss);
    return hipGetLastError();
}

hipError_t launchAdam(float* w, float* m, float* v, const float* g, int64_t n, float lr, int step, hipStream_t st)
{
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    kAdam<<<gridFor(n), 256, 0, st>>>(w, m, v, g, n, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t launchConvWeights(const float* w, int coutP, int cinP, __bf16* wf, __bf16* wt, hipStream_t st)
{
    kConvWeights<<<gridFor((int64_t)9 * coutP * cinP), 256, 0, st>>>(w, coutP, cinP, wf, wt);
    return hipGetLastError();
}

}  // namespace aph
