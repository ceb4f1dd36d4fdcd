// Denoiser kernels (denoise.hip) and the U-Net built from them (denoiser.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace aph {

constexpr int kDnInChannels = 12;  // log radiance 3, albedo 3, metalness, normal 3, roughness, depth
constexpr int kDnAlign = 16;       // image sides are padded to this (four 2x2 poolings)
constexpr int kDnPredStride = 4;   // fp32 prediction: log radiance 3 + pad
constexpr int kDnLossL1Log = 0;        // L1 of log1p radiance
constexpr int kDnLossL2Tonemapped = 1; // L2 of x / (1 + x)

// 3x3 convolution, stride 1, zero padding 1, NHWC.  Input channels [0, cinP)
// of a buffer with stride inStride (cinP a multiple of 32); weights
// [9][coutP][cinP] bf16 (coutP a multiple of 16); output channels [0,
// coutStore) written at outOffset of a buffer with stride outStride.  With
// pool set (bf16 output, even H and W), the 2x2 average pooling of the stored
// output is also written at poolOffset of a buffer of (H/2) x (W/2) pixels with
// stride poolStride, exactly as launchAvgPool computes it from the output.
struct ConvArgs {
    const __bf16* in;
    const __bf16* w;
    const float* bias;  // coutP entries or null
    __bf16* out;        // bf16 output, or
    float* outF;        // fp32 output
    int N, H, W;
    int inStride, cinP, coutP, coutStore, outStride, outOffset;
    __bf16* pool;       // optional pooled output
    int poolStride, poolOffset;
    // Optional nearest 2x upsampling folded into the input: with in2 set, input
    // channels [in2From, cinP) (in2From a multiple of 32) are read from in2, a
    // buffer of (H/2) x (W/2) pixels with stride in2Stride, at pixel (y/2, x/2).
    const __bf16* in2;
    int in2Stride, in2From;
};

// dW[9][coutP][cinP] += sum over pixels of dY (x) shifted X; db[coutP] +=
// sum over pixels of dY (when db is not null).
struct WGradArgs {
    const __bf16* dy;
    const __bf16* x;
    float* dw;
    float* db;
    int N, H, W;
    int dyStride, xStride, coutP, cinP;
    int chunk;  // pixels per block (a multiple of 128)
};

hipError_t launchConv(const ConvArgs& a, bool relu, bool outF32, hipStream_t st);
hipError_t launchWGrad(const WGradArgs& a, hipStream_t st);
hipError_t launchReluMask(__bf16* dy, int dyStride, const __bf16* y, int yStride, int64_t npix, int C, hipStream_t st);
hipError_t launchAvgPool(const __bf16* in, int inStride, int inOff, __bf16* out, int outStride, int outOff, int N,
                         int H, int W, int C, hipStream_t st);
hipError_t launchAvgPoolBack(const __bf16* dout, int doStride, __bf16* din, int diStride, int diOff, int N, int H,
                             int W, int C, bool accumulate, hipStream_t st);
hipError_t launchUpsample(const __bf16* in, int inStride, __bf16* out, int outStride, int outOff, int N, int Ho,
                          int Wo, int C, hipStream_t st);
hipError_t launchUpsampleBack(const __bf16* dout, int doStride, int doOff, __bf16* din, int diStride, int N, int Hi,
                              int Wi, int C, hipStream_t st);
// transform: destination (x, y) reads source (sx, sy), (u, v) = bit 2 ? (y, x) : (x, y),
// sx = bit 0 ? W-1-u : u, sy = bit 1 ? H-1-v : v (transposition needs a square image)
hipError_t launchPack(const float4* noisy, const float* depth, const float4* normalRough, const float4* albedo, int W,
                      int H, int Wp, int Hp, int img, uint32_t transform, __bf16* out, int stride, hipStream_t st);
// k: index of this pass in an ensemble over transforms (running mean into out)
hipError_t launchUnpack(const float* pred, int W, int H, int Wp, int Hp, int img, uint32_t transform, int k,
                        const float4* noisy, float4* out, hipStream_t st);
hipError_t launchLoss(const float* pred, const float4* ref, int W, int H, int Wp, int Hp, int img, int nImg,
                      uint32_t transform, int kind, __bf16* dpred, int dStride, float* loss, hipStream_t st);
hipError_t launchAdam(float* w, float* m, float* v, const float* g, int64_t n, float lr, int step, hipStream_t st);
hipError_t launchConvWeights(const float* w, int coutP, int cinP, __bf16* wf, __bf16* wt, hipStream_t st);
hipError_t launchBf16ToF32(const __bf16* in, float* out, int64_t n, hipStream_t st);

}  // namespace aph
