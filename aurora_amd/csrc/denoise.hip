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

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Output channels co..co+3 of pixel q (bias b, optional ReLU): one 8-byte bf16
// or 16-byte fp32 store when all four are stored (channel offsets and strides
// are multiples of 4), single stores for a partial group.
template <bool RELU, bool OUT_F32>
__device__ __forceinline__ void storeQuad(const ConvArgs& a, int64_t q, int co, f32x4 acc, const float* b)
{
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        v[e] = acc[e] + b[e];
        if (RELU) v[e] = fmaxf(v[e], 0.0f);
    }
    const int64_t at = q * a.outStride + a.outOffset + co;
    if (co + 4 <= a.coutStore) {
        if (OUT_F32) *reinterpret_cast<float4*>(a.outF + at) = make_float4(v[0], v[1], v[2], v[3]);
        else *reinterpret_cast<bf16x4*>(a.out + at) = bf16x4{toBf(v[0]), toBf(v[1]), toBf(v[2]), toBf(v[3])};
    } else {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (co + e >= a.coutStore) continue;
            if (OUT_F32) a.outF[at + e] = v[e];
            else a.out[at + e] = toBf(v[e]);
        }
    }
}

// ---- convolution (forward, and data gradient with flipped weights) ----
// One wave: P subtiles of 16 consecutive pixels x T subtiles of 16 output
// channels (P * T accumulators); each A fragment (8 input channels of one
// pixel) feeds T MFMAs and each B fragment (8 weights) feeds P.  Block: 4
// waves; blockIdx.y selects 16 * T output channels.
template <bool RELU, bool OUT_F32, int P, int T>
__global__ __launch_bounds__(256) void kConv3x3(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * (16 * P);
    if (p0 >= npix) return;
    const int coBase = blockIdx.y * (16 * T);
    const int nTiles = min(T, (a.coutP - coBase) / 16);
    const int row = lane & 15, kg = lane >> 4;
    const int HW = a.H * a.W;
    // pixel fed into A by this lane in each pixel subtile
    int n[P], y[P], x[P];
#pragma unroll
    for (int i = 0; i < P; i++) {
        const int64_t pa = p0 + i * 16 + row;
        if (pa < npix) {
            n[i] = (int)(pa / HW);
            const int r = (int)(pa - (int64_t)n[i] * HW);
            y[i] = r / a.W;
            x[i] = r - y[i] * a.W;
        } else {
            n[i] = 0;
            y[i] = -4;  // every tap out of the image
            x[i] = 0;
        }
    }
    f32x4 acc[P][T];
#pragma unroll
    for (int i = 0; i < P; i++)
#pragma unroll
        for (int j = 0; j < T; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int t = 0; t < 9; t++) {
        const int ty = t / 3 - 1, tx = t % 3 - 1;
        const __bf16* src[P];
        const __bf16* src2[P];
        bool ok[P];
#pragma unroll
        for (int i = 0; i < P; i++) {
            const int sy = y[i] + ty, sx = x[i] + tx;
            ok[i] = sy >= 0 && sy < a.H && sx >= 0 && sx < a.W;
            const int cy = ok[i] ? sy : 0, cx = ok[i] ? sx : 0;
            src[i] = a.in + (((int64_t)n[i] * a.H + cy) * a.W + cx) * a.inStride + kg * 8;
            // the upsampled channels: pixel (y/2, x/2) of the half-resolution buffer
            src2[i] = a.in2 ? a.in2 + (((int64_t)n[i] * (a.H / 2) + cy / 2) * (a.W / 2) + cx / 2) * a.in2Stride + kg * 8
                            : src[i];
        }
        const int split = a.in2 ? a.in2From : a.cinP;
        const __bf16* wt = a.w + ((int64_t)t * a.coutP + coBase + row) * a.cinP + kg * 8;
        for (int c0 = 0; c0 < a.cinP; c0 += 32) {
            bf16x8 av[P], bv[T];
#pragma unroll
            for (int i = 0; i < P; i++) av[i] = ok[i] ? load8(c0 < split ? src[i] + c0 : src2[i] + (c0 - split)) : bf16x8{};
#pragma unroll
            for (int j = 0; j < T; j++) bv[j] = j < nTiles ? load8(wt + (int64_t)j * 16 * a.cinP + c0) : bf16x8{};
#pragma unroll
            for (int j = 0; j < T; j++) {
                if (j < nTiles) {
#pragma unroll
                    for (int i = 0; i < P; i++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[j], av[i], acc[i][j], 0, 0, 0);
                }
            }
        }
    }
    // D of (j, i), operands swapped: rows = 16 output channels, columns = the
    // 16 pixels of subtile i; lane (row, kg) holds channels 4kg..4kg+3 of pixel row.
#pragma unroll
    for (int j = 0; j < T; j++) {
        if (j >= nTiles) continue;
        const int co = coBase + j * 16 + kg * 4;
        if (co >= a.coutStore) continue;
        float b[4];
#pragma unroll
        for (int e = 0; e < 4; e++) b[e] = a.bias ? a.bias[co + e] : 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) {
            const int64_t q = p0 + i * 16 + row;
            if (q < npix) storeQuad<RELU, OUT_F32>(a, q, co, acc[i][j], b);
        }
    }
}

// Fused 2x2 average pooling of a convolution's stored output.  Rows y0 + i and
// y0 + i + 1 (i even) are two accumulators of the same lane and columns x, x + 1
// are neighbouring lanes (swapped with a DPP quad permutation), so each even lane
// forms one pooled pixel from the bf16-rounded outputs, summed in the order of
// kAvgPool2 ((y, x), (y, x+1), (y+1, x), (y+1, x+1)) and rounded once.
template <bool RELU, int P, int T>
__device__ __forceinline__ void poolQuads(const ConvArgs& a, const f32x4 (&acc)[P][T], int nTiles, int coBase, int n,
                                          int y0, int x, int lane, int kg)
{
    static_assert(P % 2 == 0, "row pairs within a wave");
    const int Ho = a.H / 2, Wo = a.W / 2;
#pragma unroll
    for (int j = 0; j < T; j++) {
        if (j >= nTiles) continue;
        const int co = coBase + j * 16 + kg * 4;
        float b[4];
#pragma unroll
        for (int e = 0; e < 4; e++) b[e] = a.bias ? a.bias[co + e] : 0.0f;
#pragma unroll
        for (int i = 0; i < P; i += 2) {
            float s[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                float v0 = acc[i][j][e] + b[e], v1 = acc[i + 1][j][e] + b[e];
                if (RELU) {
                    v0 = fmaxf(v0, 0.0f);
                    v1 = fmaxf(v1, 0.0f);
                }
                v0 = fromBf(toBf(v0));
                v1 = fromBf(toBf(v1));
                // quad_perm [1, 0, 3, 2]: the value of lane l ^ 1
                const float w0 = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v0), __float_as_int(v0),
                                                                            0xB1, 0xF, 0xF, false));
                const float w1 = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v1), __float_as_int(v1),
                                                                            0xB1, 0xF, 0xF, false));
                float t = 0.0f;
                t += v0;
                t += w0;
                t += v1;
                t += w1;
                s[e] = t;
            }
            const int y = y0 + i;
            if ((lane & 1) == 0 && y < a.H && x < a.W && co < a.coutStore) {
                __bf16* q = a.pool + ((int64_t)(n * Ho + y / 2) * Wo + x / 2) * a.poolStride + a.poolOffset + co;
                if (co + 4 <= a.coutStore) {
                    *reinterpret_cast<bf16x4*>(q) =
                        bf16x4{toBf(0.25f * s[0]), toBf(0.25f * s[1]), toBf(0.25f * s[2]), toBf(0.25f * s[3])};
                } else {
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (co + e < a.coutStore) q[e] = toBf(0.25f * s[e]);
                }
            }
        }
    }
}

// ---- convolution through LDS (32 or 64 output channels) ----
// Block: 4 waves over a tile of 4P rows x 16 columns of one image and 16*T
// output channels.  For each chunk of 32 input channels the block stages the
// tile's input with a one-pixel halo and the chunk's weights of all 9 taps in
// LDS; every wave then takes its A fragments (P rows of 16 pixels) and B
// fragments from there.  Pixels are 40 bf16 apart in LDS (80 B: the 16 lanes
// of a fragment read 16 disjoint 4-bank groups).
constexpr int kLdsPix = 40;

template <bool RELU, bool OUT_F32, int P, int T>
__global__ __launch_bounds__(256) void kConvLds(ConvArgs a)
{
    constexpr int TR = 4 * P, HR = TR + 2, HC = 18;
    __shared__ __attribute__((aligned(16))) __bf16 tile[HR * HC * kLdsPix];
    __shared__ __attribute__((aligned(16))) __bf16 wts[9 * 16 * T * kLdsPix];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, row = lane & 15, kg = lane >> 4;
    const int tilesX = (a.W + 15) / 16, tilesY = (a.H + TR - 1) / TR;
    const int tx0 = (int)(blockIdx.x % tilesX) * 16;
    const int rest = (int)(blockIdx.x / tilesX);
    const int ty0 = (rest % tilesY) * TR, n = rest / tilesY;
    const int coBase = blockIdx.y * (16 * T);
    const int nTiles = min(T, (a.coutP - coBase) / 16);
    f32x4 acc[P][T];
#pragma unroll
    for (int i = 0; i < P; i++)
#pragma unroll
        for (int j = 0; j < T; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // Staging of one 32-channel chunk: every thread's segments of the haloed
    // input tile (4 x 8 channels per pixel) and of the 9-tap weights are loaded
    // into registers in one batch, then written to LDS, so a block pays one
    // global-memory latency per chunk; the next chunk's loads are issued before
    // the current chunk's MFMAs.
    constexpr int NSEG = HR * HC * 4, NIT = (NSEG + 255) / 256;
    constexpr int NWSEG = 9 * 16 * T * 4, NWIT = (NWSEG + 255) / 256;
    bf16x8 sv[NIT], sw[NWIT];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int k = 0; k < NIT; k++) {
            const int sgi = tid + k * 256;
            const int px = sgi >> 2, part = sgi & 3;
            const int hy = px / HC, hx = px - hy * HC;
            const int gy = ty0 + hy - 1, gx = tx0 + hx - 1;
            sv[k] = bf16x8{};
            if (sgi < NSEG && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                if (a.in2 && c0 >= a.in2From)
                    sv[k] = load8(a.in2 + (((int64_t)n * (a.H / 2) + gy / 2) * (a.W / 2) + gx / 2) * a.in2Stride +
                                  (c0 - a.in2From) + part * 8);
                else
                    sv[k] = load8(a.in + (((int64_t)n * a.H + gy) * a.W + gx) * a.inStride + c0 + part * 8);
            }
        }
#pragma unroll
        for (int k = 0; k < NWIT; k++) {
            const int sgi = tid + k * 256;
            const int r = sgi >> 2, part = sgi & 3;
            const int t = r / (16 * T), co = r - t * (16 * T);
            sw[k] = bf16x8{};
            if (sgi < NWSEG && co < nTiles * 16)
                sw[k] = load8(a.w + ((int64_t)t * a.coutP + coBase + co) * a.cinP + c0 + part * 8);
        }
    };
    fetch(0);
    for (int c0 = 0; c0 < a.cinP; c0 += 32) {
#pragma unroll
        for (int k = 0; k < NIT; k++) {
            const int sgi = tid + k * 256;
            if (sgi < NSEG) *reinterpret_cast<bf16x8*>(&tile[(sgi >> 2) * kLdsPix + (sgi & 3) * 8]) = sv[k];
        }
#pragma unroll
        for (int k = 0; k < NWIT; k++) {
            const int sgi = tid + k * 256;
            if (sgi < NWSEG) *reinterpret_cast<bf16x8*>(&wts[(sgi >> 2) * kLdsPix + (sgi & 3) * 8]) = sw[k];
        }
        __syncthreads();
        if (c0 + 32 < a.cinP) fetch(c0 + 32);
#pragma unroll 1
        for (int t = 0; t < 9; t++) {
            const int dy = t / 3, dx = t % 3;  // halo offsets of the tap
            bf16x8 av[P], bv[T];
#pragma unroll
            for (int i = 0; i < P; i++)
                av[i] = *reinterpret_cast<const bf16x8*>(
                    &tile[((wave * P + i + dy) * HC + row + dx) * kLdsPix + kg * 8]);
#pragma unroll
            for (int j = 0; j < T; j++)
                bv[j] = *reinterpret_cast<const bf16x8*>(&wts[(t * 16 * T + j * 16 + row) * kLdsPix + kg * 8]);
#pragma unroll
            for (int j = 0; j < T; j++) {
                if (j < nTiles) {
#pragma unroll
                    for (int i = 0; i < P; i++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[j], av[i], acc[i][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D of (j, i), operands swapped: rows = 16 output channels, columns = the 16
    // pixels of tile row wave*P + i; lane (row, kg) holds channels 4kg..4kg+3 of
    // pixel tx0 + row.
    const int x = tx0 + row;
    if (!OUT_F32 && a.pool) poolQuads<RELU, P, T>(a, acc, nTiles, coBase, n, ty0 + wave * P, x, lane, kg);
#pragma unroll
    for (int j = 0; j < T; j++) {
        if (j >= nTiles) continue;
        const int co = coBase + j * 16 + kg * 4;
        if (co >= a.coutStore || x >= a.W) continue;
        float b[4];
#pragma unroll
        for (int e = 0; e < 4; e++) b[e] = a.bias ? a.bias[co + e] : 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) {
            const int y = ty0 + wave * P + i;
            if (y < a.H) storeQuad<RELU, OUT_F32>(a, ((int64_t)n * a.H + y) * a.W + x, co, acc[i][j], b);
        }
    }
}

// ---- weight gradient: dW[t][co][ci] += sum_p dY[p][co] X[p + off(t)][ci] ----
// Block: one tap, 16*TC output x 16*TI input channels, one pixel chunk; its 4
// waves take interleaved steps of 32 pixels (K = pixels, 8 per lane,
// gathered), each wave adding its TC x TI tile of partial sums atomically.
// The centre tap's blocks of the first input-channel group also sum dY into
// the bias gradient db[co].
template <int TC, int TI>
__global__ __launch_bounds__(256) void kConvWGrad(WGradArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, row = lane & 15, kg = lane >> 4;
    const int t = blockIdx.y;
    const int ciGroups = (a.cinP + 16 * TI - 1) / (16 * TI);
    const int coBase = (blockIdx.z / ciGroups) * 16 * TC, ciBase = (blockIdx.z % ciGroups) * 16 * TI;
    const int nTc = min(TC, (a.coutP - coBase) / 16), nTi = min(TI, (a.cinP - ciBase) / 16);
    const int HW = a.H * a.W;
    const int64_t npix = (int64_t)a.N * HW;
    const int64_t c0 = (int64_t)blockIdx.x * a.chunk, c1 = min(npix, c0 + a.chunk);
    const int ty = t / 3 - 1, tx = t % 3 - 1;
    const int64_t shift = (int64_t)ty * a.W + tx;
    f32x4 acc[TC][TI];
#pragma unroll
    for (int i = 0; i < TC; i++)
#pragma unroll
        for (int j = 0; j < TI; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const __bf16* dyBase = a.dy + coBase + row;
    const __bf16* xBase = a.x + ciBase + row;
    const bool biasBlock = a.db && t == 4 && ciBase == 0;
    float bsum[TC] = {};
    for (int64_t p = c0 + wave * 32; p < c1; p += 128) {
        const int64_t q0 = p + kg * 8;
        int n = (int)(q0 / HW);
        int r = (int)(q0 - (int64_t)n * HW);
        int y = r / a.W, x = r - (r / a.W) * a.W;
        bf16x8 av[TC], bv[TI];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int64_t q = q0 + e;
            const bool valid = q < c1;
            const int sy = y + ty, sx = x + tx;
            const bool okx = valid && sy >= 0 && sy < a.H && sx >= 0 && sx < a.W;
#pragma unroll
            for (int i = 0; i < TC; i++)
                av[i][e] = (valid && i < nTc) ? dyBase[q * a.dyStride + i * 16] : toBf(0.0f);
#pragma unroll
            for (int j = 0; j < TI; j++)
                bv[j][e] = (okx && j < nTi) ? xBase[(q + shift) * a.xStride + j * 16] : toBf(0.0f);
            if (++x == a.W) {
                x = 0;
                if (++y == a.H) y = 0;
            }
        }
#pragma unroll
        for (int i = 0; i < TC; i++)
#pragma unroll
            for (int j = 0; j < TI; j++)
                if (i < nTc && j < nTi)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        if (biasBlock) {
#pragma unroll
            for (int i = 0; i < TC; i++)
#pragma unroll
                for (int e = 0; e < 8; e++) bsum[i] += fromBf(av[i][e]);
        }
    }
    if (biasBlock) {
        // lanes row, row+16, row+32, row+48 hold the same channel
#pragma unroll
        for (int i = 0; i < TC; i++) {
            float v = bsum[i];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (kg == 0 && i < nTc) atomicAdd(&a.db[coBase + i * 16 + row], v);
        }
    }
    // D rows are co (A rows), columns ci (B columns)
#pragma unroll
    for (int i = 0; i < TC; i++)
#pragma unroll
        for (int j = 0; j < TI; j++) {
            if (i >= nTc || j >= nTi) continue;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int dco = coBase + i * 16 + kg * 4 + e, dci = ciBase + j * 16 + row;
                atomicAdd(&a.dw[((int64_t)t * a.coutP + dco) * a.cinP + dci], acc[i][j][e]);
            }
        }
}

__device__ __forceinline__ void store8(__bf16* p, bf16x8 v) { *reinterpret_cast<bf16x8*>(p) = v; }

// The elementwise kernels below take 8 channels (16 bytes) per thread; channel
// counts, strides and offsets are multiples of 8.

// dY *= (Y > 0) for a ReLU output Y: channels [0, C) of two strided buffers
__global__ void kReluMask(__bf16* dy, int dyStride, const __bf16* y, int yStride, int64_t npix, int C)
{
    const int G = C / 8;
    const int64_t n = npix * G;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = i / G;
        const int c = (int)(i - p * G) * 8;
        bf16x8 g = load8(dy + p * dyStride + c), v = load8(y + p * yStride + c);
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (!(fromBf(v[k]) > 0.0f)) g[k] = toBf(0.0f);
        store8(dy + p * dyStride + c, g);
    }
}

__global__ void kBf16ToF32(const __bf16* in, float* out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = fromBf(in[i]);
}

// ---- pooling and upsampling (2x2, NHWC; channel slices through offsets) ----
// The glue kernels run on a grid of (runs of pixels of one row, rows, images) with
// 32-bit index arithmetic: one thread per 8 channels of one pixel.
__global__ void kAvgPool2(const __bf16* in, int inStride, int inOff, __bf16* out, int outStride, int outOff, int N,
                          int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2, G = C / 8;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int xo = t / G, c = (t - xo * G) * 8;
    if (xo >= Wo) return;
    const int yo = blockIdx.y, b = blockIdx.z;
    float s[8] = {};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bf16x8 v = load8(in + ((int64_t)(b * H + 2 * yo + k / 2) * W + 2 * xo + k % 2) * inStride + inOff + c);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += fromBf(v[e]);
    }
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = toBf(0.25f * s[e]);
    store8(out + ((int64_t)(b * Ho + yo) * Wo + xo) * outStride + outOff + c, o);
}

__global__ void kAvgPool2Back(const __bf16* dout, int doStride, __bf16* din, int diStride, int diOff, int N, int H,
                              int W, int C, int accumulate)
{
    const int G = C / 8;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = t / G, c = (t - x * G) * 8;
    if (x >= W) return;
    const int y = blockIdx.y, b = blockIdx.z;
    const bf16x8 g = load8(dout + ((int64_t)(b * (H / 2) + y / 2) * (W / 2) + x / 2) * doStride + c);
    __bf16* d = din + ((int64_t)(b * H + y) * W + x) * diStride + diOff + c;
    bf16x8 o = accumulate ? load8(d) : bf16x8{};
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = toBf((accumulate ? fromBf(o[e]) : 0.0f) + 0.25f * fromBf(g[e]));
    store8(d, o);
}

__global__ void kUpsample2(const __bf16* in, int inStride, __bf16* out, int outStride, int outOff, int N, int Ho,
                           int Wo, int C)
{
    const int G = C / 8;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = t / G, c = (t - x * G) * 8;
    if (x >= Wo) return;
    const int y = blockIdx.y, b = blockIdx.z;
    store8(out + ((int64_t)(b * Ho + y) * Wo + x) * outStride + outOff + c,
           load8(in + ((int64_t)(b * (Ho / 2) + y / 2) * (Wo / 2) + x / 2) * inStride + c));
}

__global__ void kUpsample2Back(const __bf16* dout, int doStride, int doOff, __bf16* din, int diStride, int N, int Hi,
                               int Wi, int C)
{
    const int G = C / 8;
    const int Ho = 2 * Hi, Wo = 2 * Wi;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = t / G, c = (t - x * G) * 8;
    if (x >= Wi) return;
    const int y = blockIdx.y, b = blockIdx.z;
    float s[8] = {};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bf16x8 v = load8(dout + ((int64_t)(b * Ho + 2 * y + k / 2) * Wo + 2 * x + k % 2) * doStride + doOff + c);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += fromBf(v[e]);
    }
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = toBf(s[e]);
    store8(din + ((int64_t)(b * Hi + y) * Wi + x) * diStride + c, o);
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
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < Wp) {
        __bf16* o = out + (((int64_t)img * Hp + y) * Wp + x) * stride;
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
        // 16-byte stores of 8 channels: the 12 features, then zeros to the stride
        static_assert(kDnInChannels > 8 && kDnInChannels <= 16, "features fill two 8-channel groups");
        bf16x8 v0, v1;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v0[k] = toBf(f[k]);
            v1[k] = toBf(8 + k < kDnInChannels ? f[(8 + k) % kDnInChannels] : 0.0f);
        }
        store8(o, v0);
        store8(o + 8, v1);
        for (int c0 = 16; c0 < stride; c0 += 8) store8(o + c0, bf16x8{});
    }
}

// ---- output: expm1 of the predicted log radiance, alpha from the noisy input.
// The prediction of batch image img was made on the input under transform xf:
// its pixel (x, y) belongs to source pixel sourcePixel(x, y).  Pass k of an
// ensemble folds into the running mean of the k passes before it ----
__global__ void kUnpack(const float* pred, int W, int H, int Wp, int Hp, int img, uint32_t xf, int k,
                        const float4* noisy, float4* out)
{
    const float wNew = 1.0f / (float)(k + 1);
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < W) {
        const float* p = pred + ((int64_t)img * Wp * Hp + (int64_t)y * Wp + x) * kDnPredStride;
        int sx, sy;
        sourcePixel(x, y, W, H, xf, sx, sy);
        const int64_t s = (int64_t)sy * W + sx;
        float3 v = make_float3(expm1f(fmaxf(p[0], 0.0f)), expm1f(fmaxf(p[1], 0.0f)), expm1f(fmaxf(p[2], 0.0f)));
        if (k > 0) {
            const float4 o = out[s];
            v = make_float3(o.x + (v.x - o.x) * wNew, o.y + (v.y - o.y) * wNew, o.z + (v.z - o.z) * wNew);
        }
        out[s] = make_float4(v.x, v.y, v.z, noisy[s].w);
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

// Grid of the glue kernels: 256-thread blocks over one row of w pixels x c/8
// channel groups, one block row per image row (h), one layer per image (n).
dim3 rowGrid(int w, int c, int h, int n)
{
    return dim3((unsigned)((w * (c / 8) + 255) / 256), (unsigned)h, (unsigned)n);
}

}  // namespace

template <int P, int T>
void convLaunch(const ConvArgs& a, bool relu, bool outF32, hipStream_t st)
{
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    dim3 grid((unsigned)((npix + 64 * P - 1) / (64 * P)), (unsigned)((a.coutP + 16 * T - 1) / (16 * T)));
    if (outF32) {
        if (relu) kConv3x3<true, true, P, T><<<grid, 256, 0, st>>>(a);
        else kConv3x3<false, true, P, T><<<grid, 256, 0, st>>>(a);
    } else {
        if (relu) kConv3x3<true, false, P, T><<<grid, 256, 0, st>>>(a);
        else kConv3x3<false, false, P, T><<<grid, 256, 0, st>>>(a);
    }
}

template <int P, int T>
void convLdsLaunch(const ConvArgs& a, bool relu, bool outF32, hipStream_t st)
{
    const unsigned tiles = (unsigned)((a.W + 15) / 16) * (unsigned)((a.H + 4 * P - 1) / (4 * P)) * (unsigned)a.N;
    dim3 grid(tiles, (unsigned)((a.coutP + 16 * T - 1) / (16 * T)));
    if (outF32) {
        if (relu) kConvLds<true, true, P, T><<<grid, 256, 0, st>>>(a);
        else kConvLds<false, true, P, T><<<grid, 256, 0, st>>>(a);
    } else {
        if (relu) kConvLds<true, false, P, T><<<grid, 256, 0, st>>>(a);
        else kConvLds<false, false, P, T><<<grid, 256, 0, st>>>(a);
    }
}

hipError_t launchConv(const ConvArgs& a, bool relu, bool outF32, hipStream_t st)
{
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    if (npix == 0) return hipSuccess;
    // 32 or 64 output channels (the full- and half-resolution layers): LDS
    // tiles of 16 x 16 pixels (49 KB or 72 KB of LDS: three or two blocks per CU);
    // wider layers: 2 pixel x 4 channel subtiles per wave from global memory.
    // Measured on a 1080p frame against 32 x 16 tiles for 32 channels and 2 x 8
    // subtiles: 5 % faster full-resolution layers, 10 % faster wide ones
    if (a.pool && (outF32 || (a.H & 1) || (a.W & 1))) return hipErrorInvalidValue;
    if (a.coutP <= 32) convLdsLaunch<4, 2>(a, relu, outF32, st);
    else if (a.coutP <= 64) convLdsLaunch<4, 4>(a, relu, outF32, st);
    else {
        // pooling as its own pass after the global-memory kernel
        convLaunch<2, 4>(a, relu, outF32, st);
        if (a.pool) {
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            return launchAvgPool(a.out, a.outStride, a.outOffset, a.pool, a.poolStride, a.poolOffset, a.N, a.H, a.W,
                                 a.coutStore, st);
        }
    }
    return hipGetLastError();
}

hipError_t launchWGrad(const WGradArgs& a, hipStream_t st)
{
    constexpr int TC = 4, TI = 4;
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    const unsigned groups = (unsigned)(((a.coutP + 16 * TC - 1) / (16 * TC)) * ((a.cinP + 16 * TI - 1) / (16 * TI)));
    dim3 grid((unsigned)((npix + a.chunk - 1) / a.chunk), 9, groups);
    kConvWGrad<TC, TI><<<grid, 256, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launchReluMask(__bf16* dy, int dyStride, const __bf16* y, int yStride, int64_t npix, int C, hipStream_t st)
{
    kReluMask<<<gridFor(npix * C / 8), 256, 0, st>>>(dy, dyStride, y, yStride, npix, C);
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
    kAvgPool2<<<rowGrid(W / 2, C, H / 2, N), 256, 0, st>>>(in, inStride, inOff, out, outStride, outOff, N, H, W, C);
    return hipGetLastError();
}

hipError_t launchAvgPoolBack(const __bf16* dout, int doStride, __bf16* din, int diStride, int diOff, int N, int H,
                             int W, int C, bool accumulate, hipStream_t st)
{
    kAvgPool2Back<<<rowGrid(W, C, H, N), 256, 0, st>>>(dout, doStride, din, diStride, diOff, N, H, W, C,
                                                       accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t launchUpsample(const __bf16* in, int inStride, __bf16* out, int outStride, int outOff, int N, int Ho,
                          int Wo, int C, hipStream_t st)
{
    kUpsample2<<<rowGrid(Wo, C, Ho, N), 256, 0, st>>>(in, inStride, out, outStride, outOff, N, Ho, Wo, C);
    return hipGetLastError();
}

hipError_t launchUpsampleBack(const __bf16* dout, int doStride, int doOff, __bf16* din, int diStride, int N, int Hi,
                              int Wi, int C, hipStream_t st)
{
    kUpsample2Back<<<rowGrid(Wi, C, Hi, N), 256, 0, st>>>(dout, doStride, doOff, din, diStride, N, Hi, Wi, C);
    return hipGetLastError();
}

hipError_t launchPack(const float4* noisy, const float* depth, const float4* normalRough, const float4* albedo, int W,
                      int H, int Wp, int Hp, int img, uint32_t transform, __bf16* out, int stride, hipStream_t st)
{
    kPack<<<rowGrid(Wp, 8, Hp, 1), 256, 0, st>>>(noisy, depth, normalRough, albedo, W, H, Wp, Hp, img, transform, out,
                                                  stride);
    return hipGetLastError();
}

hipError_t launchUnpack(const float* pred, int W, int H, int Wp, int Hp, int img, uint32_t transform, int k,
                        const float4* noisy, float4* out, hipStream_t st)
{
    kUnpack<<<rowGrid(W, 8, H, 1), 256, 0, st>>>(pred, W, H, Wp, Hp, img, transform, k, noisy, out);
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
