// Learned denoiser: a U-Net over noisy radiance and the guide images, trained
// and run with the MFMA kernels of denoise.hip, and its C-ABI.
//
// No reference counterpart for the network: the reference's DirectX backend
// calls NVIDIA NRD RELAX (DirectX/Denoiser.cpp:133, Denoiser.h:70
// denoise(DenoiserState)), which is closed.  aph_denoise takes the place of
// that call in the frame loop (PTRenderer.cpp:1052 submitDenoising) and reads
// the guides the reference hands NRD (Denoiser.h:24-35: view depth, normal +
// roughness, motion), plus base colour and metalness.
//
// Network: four 2x2 average-pool levels, two 3x3 convolutions + ReLU per level
// (encoder widths 32, 64, 64, 96; 128 at the bottom), nearest upsampling and
// channel concatenation with the encoder output of the same level, a 3x3
// output convolution to log radiance.  Activations are bf16 NHWC, weights bf16
// for the kernels with fp32 master copies for Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"
#include "denoise.h"

static_assert(APH_DN_LOSS_L1_LOG == aph::kDnLossL1Log && APH_DN_LOSS_L2_TONEMAPPED == aph::kDnLossL2Tonemapped,
              "loss kinds of the C-ABI and the kernels agree");

using namespace aph;

namespace aph {
namespace {

enum BufId {
    B_IN, B_E0A, B_CAT0, B_D0A, B_D0B,
    B_P1, B_E1A, B_CAT1, B_D1A, B_D1B,
    B_P2, B_E2A, B_CAT2, B_D2A, B_D2B,
    B_P3, B_E3A, B_CAT3, B_D3A, B_D3B,
    B_P4, B_MA, B_MB,
    B_OUT,  // gradient of the prediction only (the prediction itself is fp32)
    kNumBufs
};

const aph_dn_buffer kBufs[kNumBufs] = {
    {0, 32}, {0, 32}, {0, 96}, {0, 32}, {0, 32},
    {1, 32}, {1, 64}, {1, 128}, {1, 64}, {1, 64},
    {2, 64}, {2, 64}, {2, 160}, {2, 64}, {2, 64},
    {3, 64}, {3, 96}, {3, 224}, {3, 96}, {3, 96},
    {4, 96}, {4, 128}, {4, 128},
    {0, 32},
};

struct LayerDef {
    const char* name;
    int cin, cinP, cout, coutP;
    int in, out, outOff;
    bool relu;
};

const LayerDef kLayers[] = {
    {"enc0a", kDnInChannels, 32, 32, 32, B_IN, B_E0A, 0, true},
    {"enc0b", 32, 32, 32, 32, B_E0A, B_CAT0, 0, true},
    {"enc1a", 32, 32, 64, 64, B_P1, B_E1A, 0, true},
    {"enc1b", 64, 64, 64, 64, B_E1A, B_CAT1, 0, true},
    {"enc2a", 64, 64, 64, 64, B_P2, B_E2A, 0, true},
    {"enc2b", 64, 64, 64, 64, B_E2A, B_CAT2, 0, true},
    {"enc3a", 64, 64, 96, 96, B_P3, B_E3A, 0, true},
    {"enc3b", 96, 96, 96, 96, B_E3A, B_CAT3, 0, true},
    {"mid_a", 96, 96, 128, 128, B_P4, B_MA, 0, true},
    {"mid_b", 128, 128, 128, 128, B_MA, B_MB, 0, true},
    {"dec3a", 224, 224, 96, 96, B_CAT3, B_D3A, 0, true},
    {"dec3b", 96, 96, 96, 96, B_D3A, B_D3B, 0, true},
    {"dec2a", 160, 160, 64, 64, B_CAT2, B_D2A, 0, true},
    {"dec2b", 64, 64, 64, 64, B_D2A, B_D2B, 0, true},
    {"dec1a", 128, 128, 64, 64, B_CAT1, B_D1A, 0, true},
    {"dec1b", 64, 64, 64, 64, B_D1A, B_D1B, 0, true},
    {"dec0a", 96, 96, 32, 32, B_CAT0, B_D0A, 0, true},
    {"dec0b", 32, 32, 32, 32, B_D0A, B_D0B, 0, true},
    {"output", 32, 32, 3, 32, B_D0B, B_OUT, 0, false},
};
constexpr int kNumLayers = (int)(sizeof(kLayers) / sizeof(kLayers[0]));

// Forward schedule; the backward pass walks it in reverse.  Every consumer of
// a buffer slice comes after its producer, and each encoder output (the low
// channels of a concatenation buffer) is read by a pooling op before the
// decoder convolution that reads the whole buffer.
const aph_dn_op kOps[] = {
    {APH_DN_OP_CONV, 0, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 1, 0, 0, 0, 0, 0},
    {APH_DN_OP_POOL, -1, B_CAT0, 0, 32, B_P1, 0},
    {APH_DN_OP_CONV, 2, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 3, 0, 0, 0, 0, 0},
    {APH_DN_OP_POOL, -1, B_CAT1, 0, 64, B_P2, 0},
    {APH_DN_OP_CONV, 4, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 5, 0, 0, 0, 0, 0},
    {APH_DN_OP_POOL, -1, B_CAT2, 0, 64, B_P3, 0},
    {APH_DN_OP_CONV, 6, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 7, 0, 0, 0, 0, 0},
    {APH_DN_OP_POOL, -1, B_CAT3, 0, 96, B_P4, 0},
    {APH_DN_OP_CONV, 8, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 9, 0, 0, 0, 0, 0},
    {APH_DN_OP_UPSAMPLE, -1, B_MB, 0, 128, B_CAT3, 96},
    {APH_DN_OP_CONV, 10, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 11, 0, 0, 0, 0, 0},
    {APH_DN_OP_UPSAMPLE, -1, B_D3B, 0, 96, B_CAT2, 64},
    {APH_DN_OP_CONV, 12, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 13, 0, 0, 0, 0, 0},
    {APH_DN_OP_UPSAMPLE, -1, B_D2B, 0, 64, B_CAT1, 64},
    {APH_DN_OP_CONV, 14, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 15, 0, 0, 0, 0, 0},
    {APH_DN_OP_UPSAMPLE, -1, B_D1B, 0, 64, B_CAT0, 32},
    {APH_DN_OP_CONV, 16, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 17, 0, 0, 0, 0, 0},
    {APH_DN_OP_CONV, 18, 0, 0, 0, 0, 0},
};
constexpr int kNumOps = (int)(sizeof(kOps) / sizeof(kOps[0]));

struct Offsets {
    std::vector<size_t> w, b, wb;  // fp32 weight / bias offsets, bf16 weight offsets
    size_t nParams = 0, nWeightsBf = 0;
    Offsets()
    {
        for (const LayerDef& L : kLayers) {
            w.push_back(nParams);
            nParams += (size_t)9 * L.coutP * L.cinP;
            b.push_back(nParams);
            nParams += (size_t)L.coutP;
            wb.push_back(nWeightsBf);
            nWeightsBf += (size_t)9 * L.coutP * L.cinP;
        }
    }
};
const Offsets& offsets()
{
    static const Offsets o;
    return o;
}

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace
}  // namespace aph

struct aph_denoiser {
    aph_ctx* ctx = nullptr;
    DevBuf params, grads, adamM, adamV, wf, wt;
    DevBuf acts, actGrads, pred, loss, tmp;
    size_t actOff[kNumBufs] = {}, gradOff[kNumBufs] = {};
    size_t actBytes = 0, gradBytes = 0;
    int N = 0, Hp = 0, Wp = 0;  // shape of the last forward pass
    bool haveGrads = false;     // actGrads hold the last training step's gradients
    int step = 0;
    int lossKind = APH_DN_LOSS_L1_LOG;

    int levelH(int l) const { return Hp >> l; }
    int levelW(int l) const { return Wp >> l; }
    int64_t levelPix(int l) const { return (int64_t)N * levelH(l) * levelW(l); }
    __bf16* act(int b) const { return (__bf16*)((char*)acts.p + actOff[b]); }
    __bf16* grad(int b) const { return (__bf16*)((char*)actGrads.p + gradOff[b]); }
};

void aph::DenoiserDeleter::operator()(aph_denoiser* d) const { delete d; }

namespace {

aph_denoiser* findDenoiser(aph_ctx* c, uint64_t h)
{
    auto it = c->denoisers.find(h);
    return it == c->denoisers.end() ? nullptr : it->second.get();
}

// Lays out the activation (and, for training, gradient) buffers of N images of
// Hp x Wp; grows the allocations when needed.  The shape is recorded only once
// every buffer is in place (a failed allocation leaves no shape behind).
hipError_t reserveShape(aph_denoiser* d, int N, int Hp, int Wp, bool training)
{
    auto pix = [&](int level) { return (int64_t)N * (Hp >> level) * (Wp >> level); };
    size_t actOff[kNumBufs], gradOff[kNumBufs];
    size_t a = 0, g = 0;
    for (int b = 0; b < kNumBufs; b++) {
        const size_t bytes = ((size_t)pix(kBufs[b].level) * kBufs[b].channels * 2 + 255) & ~(size_t)255;
        actOff[b] = a;
        gradOff[b] = g;
        if (b != B_OUT) a += bytes;
        g += bytes;
    }
    d->N = 0;
    d->haveGrads = false;
    hipError_t e = d->acts.reserve(a);
    if (e == hipSuccess) e = d->pred.reserve((size_t)pix(0) * kDnPredStride * 4);
    if (e == hipSuccess && training) e = d->actGrads.reserve(g);
    if (e != hipSuccess) return e;
    std::memcpy(d->actOff, actOff, sizeof(actOff));
    d->actBytes = a;
    if (training) {
        std::memcpy(d->gradOff, gradOff, sizeof(gradOff));
        d->gradBytes = g;
    }
    d->N = N;
    d->Hp = Hp;
    d->Wp = Wp;
    return hipSuccess;
}

hipError_t refreshWeights(aph_denoiser* d, hipStream_t st)
{
    const Offsets& o = offsets();
    for (int l = 0; l < kNumLayers; l++) {
        const LayerDef& L = kLayers[l];
        hipError_t e = launchConvWeights(d->params.as<float>() + o.w[l], L.coutP, L.cinP, d->wf.as<__bf16>() + o.wb[l],
                                         d->wt.as<__bf16>() + o.wb[l], st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Inference folds each upsampling into the decoder convolution that reads it
// (the upsampled channels of its input are read from the lower level directly and
// never stored); training stores them, as its gradients read them.
hipError_t forward(aph_denoiser* d, hipStream_t st, bool inference)
{
    const Offsets& o = offsets();
    const aph_dn_op* up = nullptr;  // an upsampling folded into the next convolution
    for (int i = 0; i < kNumOps; i++) {
        const aph_dn_op& op = kOps[i];
        hipError_t e = hipSuccess;
        if (op.kind == APH_DN_OP_UPSAMPLE && inference && i + 1 < kNumOps && kOps[i + 1].kind == APH_DN_OP_CONV) {
            const LayerDef& Ln = kLayers[kOps[i + 1].layer];
            if (Ln.in == op.dst && op.dst_offset % 32 == 0 && op.dst_offset + op.channels == Ln.cinP &&
                op.src_offset == 0) {
                up = &op;
                continue;
            }
        }
        if (op.kind == APH_DN_OP_CONV) {
            const LayerDef& L = kLayers[op.layer];
            const int lv = kBufs[L.in].level;
            ConvArgs a{};
            if (up) {
                a.in2 = d->act(up->src);
                a.in2Stride = kBufs[up->src].channels;
                a.in2From = up->dst_offset;
                up = nullptr;
            }
            a.in = d->act(L.in);
            a.w = d->wf.as<__bf16>() + o.wb[op.layer];
            a.bias = d->params.as<float>() + o.b[op.layer];
            a.N = d->N;
            a.H = d->levelH(lv);
            a.W = d->levelW(lv);
            a.inStride = kBufs[L.in].channels;
            a.cinP = L.cinP;
            a.coutP = L.coutP;
            a.coutStore = L.cout;
            const bool isPred = L.out == B_OUT;
            if (isPred) {
                a.outF = d->pred.as<float>();
                a.outStride = kDnPredStride;
            } else {
                a.out = d->act(L.out);
                a.outStride = kBufs[L.out].channels;
                a.outOffset = L.outOff;
            }
            // A pooling of exactly this output right after it is fused into the
            // convolution's epilogue (the same bf16 result as the separate pass).
            if (!isPred && i + 1 < kNumOps) {
                const aph_dn_op& nx = kOps[i + 1];
                if (nx.kind == APH_DN_OP_POOL && nx.src == L.out && nx.src_offset == L.outOff &&
                    nx.channels == L.cout) {
                    a.pool = d->act(nx.dst);
                    a.poolStride = kBufs[nx.dst].channels;
                    a.poolOffset = nx.dst_offset;
                    i++;
                }
            }
            e = launchConv(a, L.relu, isPred, st);
        } else if (op.kind == APH_DN_OP_POOL) {
            const int lv = kBufs[op.src].level;
            e = launchAvgPool(d->act(op.src), kBufs[op.src].channels, op.src_offset, d->act(op.dst),
                              kBufs[op.dst].channels, op.dst_offset, d->N, d->levelH(lv), d->levelW(lv), op.channels,
                              st);
        } else {
            const int lv = kBufs[op.dst].level;
            e = launchUpsample(d->act(op.src), kBufs[op.src].channels, d->act(op.dst), kBufs[op.dst].channels,
                               op.dst_offset, d->N, d->levelH(lv), d->levelW(lv), op.channels, st);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Gradients of every parameter (accumulated into d->grads, cleared by the
// caller) and of every activation, from the prediction gradient in grad(B_OUT).
hipError_t backward(aph_denoiser* d, hipStream_t st)
{
    const Offsets& o = offsets();
    for (int i = kNumOps - 1; i >= 0; i--) {
        const aph_dn_op& op = kOps[i];
        hipError_t e = hipSuccess;
        if (op.kind == APH_DN_OP_CONV) {
            const int l = op.layer;
            const LayerDef& L = kLayers[l];
            const int lv = kBufs[L.in].level;
            const int64_t npix = d->levelPix(lv);
            const int dyStride = kBufs[L.out].channels;
            __bf16* dy = d->grad(L.out) + L.outOff;
            // every reader of this output has added its share by now
            if (L.relu) {
                e = launchReluMask(dy, dyStride, d->act(L.out) + L.outOff, dyStride, npix, L.coutP, st);
                if (e != hipSuccess) return e;
            }
            WGradArgs g{};
            g.dy = dy;
            g.x = d->act(L.in);
            g.dw = d->grads.as<float>() + o.w[l];
            g.db = d->grads.as<float>() + o.b[l];
            g.N = d->N;
            g.H = d->levelH(lv);
            g.W = d->levelW(lv);
            g.dyStride = dyStride;
            g.xStride = kBufs[L.in].channels;
            g.coutP = L.coutP;
            g.cinP = L.cinP;
            // about 2048 blocks of (tap, 64 x 64 channel group, pixel chunk)
            const int64_t groups = 9ll * ((L.coutP + 63) / 64) * ((L.cinP + 63) / 64);
            const int64_t chunks = std::max<int64_t>(1, 2048 / groups);
            int64_t chunk = (npix + chunks - 1) / chunks;
            chunk = std::max<int64_t>(512, (chunk + 127) / 128 * 128);
            g.chunk = (int)std::min<int64_t>(chunk, 1 << 30);
            e = launchWGrad(g, st);
            if (e != hipSuccess) return e;
            if (l == 0) continue;  // no gradient of the packed input
            ConvArgs a{};
            a.in = dy;
            a.w = d->wt.as<__bf16>() + o.wb[l];
            a.bias = nullptr;
            a.out = d->grad(L.in);
            a.N = d->N;
            a.H = d->levelH(lv);
            a.W = d->levelW(lv);
            a.inStride = dyStride;
            a.cinP = L.coutP;
            a.coutP = L.cinP;
            a.coutStore = L.cinP;
            a.outStride = kBufs[L.in].channels;
            a.outOffset = 0;
            e = launchConv(a, false, false, st);
        } else if (op.kind == APH_DN_OP_POOL) {
            // adds to the decoder convolution's gradient of the same slice
            const int lv = kBufs[op.src].level;
            e = launchAvgPoolBack(d->grad(op.dst), kBufs[op.dst].channels, d->grad(op.src), kBufs[op.src].channels,
                                  op.src_offset, d->N, d->levelH(lv), d->levelW(lv), op.channels, true, st);
        } else {
            const int lv = kBufs[op.src].level;
            e = launchUpsampleBack(d->grad(op.dst), kBufs[op.dst].channels, op.dst_offset, d->grad(op.src),
                                   kBufs[op.src].channels, d->N, d->levelH(lv), d->levelW(lv), op.channels, st);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int padded(uint32_t v) { return (int)((v + kDnAlign - 1) / kDnAlign * kDnAlign); }

// Resolves the noisy target and guides of n images of one size.
int resolveInputs(aph_ctx* c, const uint64_t* noisy, const uint64_t* aovs, uint32_t n, std::vector<Target*>& tg,
                  std::vector<AovSet*>& av, uint32_t& W, uint32_t& H)
{
    tg.clear();
    av.clear();
    for (uint32_t i = 0; i < n; i++) {
        Target* t = lookup(c->targets, noisy[i]);
        AovSet* a = lookup(c->aovs, aovs[i]);
        if (!t) return fail(c, APH_E_INVALID, "item %u: unknown noisy target", i);
        if (!a) return fail(c, APH_E_INVALID, "item %u: unknown AOV set", i);
        if (a->w != t->w || a->h != t->h) return fail(c, APH_E_INVALID, "item %u: AOV size differs from the target", i);
        if (i == 0) {
            W = t->w;
            H = t->h;
        } else if (t->w != W || t->h != H) {
            return fail(c, APH_E_INVALID, "item %u: all images of one call must have one size", i);
        }
        tg.push_back(t);
        av.push_back(a);
    }
    return APH_OK;
}

hipError_t packInputs(aph_denoiser* d, const std::vector<Target*>& tg, const std::vector<AovSet*>& av, uint32_t W,
                      uint32_t H, const std::vector<uint32_t>& xf, hipStream_t st)
{
    for (size_t i = 0; i < tg.size(); i++) {
        hipError_t e = launchPack(tg[i]->accum.as<float4>(), av[i]->depth.as<float>(),
                                  av[i]->normalRoughness.as<float4>(), av[i]->baseColorMetalness.as<float4>(), (int)W,
                                  (int)H, d->Wp, d->Hp, (int)i, xf.empty() ? 0u : xf[i], d->act(B_IN),
                                  kBufs[B_IN].channels, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

int aph_denoiser_arch(aph_dn_layer* layers, uint32_t* n_layers, aph_dn_buffer* buffers, uint32_t* n_buffers,
                      aph_dn_op* ops, uint32_t* n_ops)
{
    const Offsets& o = offsets();
    if (n_layers) {
        if (layers && *n_layers < (uint32_t)kNumLayers) return APH_E_INVALID;
        if (layers)
            for (int l = 0; l < kNumLayers; l++) {
                const LayerDef& L = kLayers[l];
                aph_dn_layer& x = layers[l];
                std::memset(&x, 0, sizeof(x));
                std::strncpy(x.name, L.name, sizeof(x.name) - 1);
                x.cin = (uint32_t)L.cin;
                x.cin_padded = (uint32_t)L.cinP;
                x.cout = (uint32_t)L.cout;
                x.cout_padded = (uint32_t)L.coutP;
                x.in_buffer = (uint32_t)L.in;
                x.out_buffer = (uint32_t)L.out;
                x.out_offset = (uint32_t)L.outOff;
                x.relu = L.relu ? 1u : 0u;
                x.weight_offset = o.w[l];
                x.bias_offset = o.b[l];
            }
        *n_layers = (uint32_t)kNumLayers;
    }
    if (n_buffers) {
        if (buffers && *n_buffers < (uint32_t)kNumBufs) return APH_E_INVALID;
        if (buffers) std::memcpy(buffers, kBufs, sizeof(kBufs));
        *n_buffers = (uint32_t)kNumBufs;
    }
    if (n_ops) {
        if (ops && *n_ops < (uint32_t)kNumOps) return APH_E_INVALID;
        if (ops) std::memcpy(ops, kOps, sizeof(kOps));
        *n_ops = (uint32_t)kNumOps;
    }
    return APH_OK;
}

uint64_t aph_denoiser_param_count(void) { return offsets().nParams; }

int aph_denoiser_create(aph_ctx_t c, uint64_t seed, uint64_t* out)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!out) return fail(c, APH_E_INVALID, "null denoiser output");
    *out = 0;
    const Offsets& o = offsets();
    std::unique_ptr<aph_denoiser, DenoiserDeleter> d(new aph_denoiser());
    d->ctx = c;
    HIPCALL(c, d->params.alloc(o.nParams * 4), "alloc denoiser parameters");
    HIPCALL(c, d->grads.alloc(o.nParams * 4), "alloc denoiser gradients");
    HIPCALL(c, d->adamM.alloc(o.nParams * 4), "alloc denoiser moments");
    HIPCALL(c, d->adamV.alloc(o.nParams * 4), "alloc denoiser moments");
    HIPCALL(c, d->wf.alloc(o.nWeightsBf * 2), "alloc denoiser weights");
    HIPCALL(c, d->wt.alloc(o.nWeightsBf * 2), "alloc denoiser weights");
    HIPCALL(c, d->loss.alloc(256), "alloc denoiser loss");
    // He-normal weights over the real (unpadded) fan-in, zero biases and padding
    std::vector<float> p(o.nParams, 0.0f);
    uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1;
    for (int l = 0; l < kNumLayers; l++) {
        const LayerDef& L = kLayers[l];
        const double sd = std::sqrt(2.0 / (9.0 * L.cin));
        for (int t = 0; t < 9; t++)
            for (int co = 0; co < L.coutP; co++)
                for (int ci = 0; ci < L.cinP; ci++) {
                    double u1 = ((splitmix(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
                    double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
                    double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                    if (co < L.cout && ci < L.cin)
                        p[o.w[l] + ((size_t)t * L.coutP + co) * L.cinP + ci] = (float)(g * sd);
                }
    }
    HIPCALL(c, hipMemcpy(d->params.p, p.data(), o.nParams * 4, hipMemcpyHostToDevice), "upload denoiser parameters");
    HIPCALL(c, hipMemsetAsync(d->adamM.p, 0, o.nParams * 4, c->stream), "clear moments");
    HIPCALL(c, hipMemsetAsync(d->adamV.p, 0, o.nParams * 4, c->stream), "clear moments");
    HIPCALL(c, hipMemsetAsync(d->grads.p, 0, o.nParams * 4, c->stream), "clear gradients");
    HIPCALL(c, hipMemsetAsync(d->wt.p, 0, o.nWeightsBf * 2, c->stream), "clear weights");
    HIPCALL(c, refreshWeights(d.get(), c->stream), "denoiser weights");
    HIPCALL(c, hipStreamSynchronize(c->stream), "denoiser create");
    uint64_t h = c->nextHandle++;
    c->denoisers[h] = std::move(d);
    *out = h;
    return APH_OK;
}

int aph_denoiser_destroy(aph_ctx_t c, uint64_t dn)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!findDenoiser(c, dn)) return fail(c, APH_E_INVALID, "unknown denoiser %llu", (unsigned long long)dn);
    HIPCALL(c, hipStreamSynchronize(c->stream), "denoiser destroy");
    c->denoisers.erase(dn);
    return APH_OK;
}

int aph_denoiser_set_params(aph_ctx_t c, uint64_t dn, const float* params, uint64_t count)
{
    int rc = enter(c);
    if (rc) return rc;
    aph_denoiser* d = findDenoiser(c, dn);
    if (!d) return fail(c, APH_E_INVALID, "unknown denoiser %llu", (unsigned long long)dn);
    const Offsets& o = offsets();
    if (!params || count != o.nParams)
        return fail(c, APH_E_INVALID, "expected %llu parameters, got %llu", (unsigned long long)o.nParams,
                    (unsigned long long)count);
    HIPCALL(c, hipStreamSynchronize(c->stream), "set parameters");
    HIPCALL(c, hipMemcpy(d->params.p, params, count * 4, hipMemcpyHostToDevice), "upload parameters");
    HIPCALL(c, hipMemsetAsync(d->adamM.p, 0, o.nParams * 4, c->stream), "clear moments");
    HIPCALL(c, hipMemsetAsync(d->adamV.p, 0, o.nParams * 4, c->stream), "clear moments");
    d->step = 0;
    HIPCALL(c, refreshWeights(d, c->stream), "denoiser weights");
    return APH_OK;
}

int aph_denoiser_set_loss(aph_ctx_t c, uint64_t dn, int kind)
{
    int rc = enter(c);
    if (rc) return rc;
    aph_denoiser* d = findDenoiser(c, dn);
    if (!d) return fail(c, APH_E_INVALID, "unknown denoiser %llu", (unsigned long long)dn);
    if (kind != APH_DN_LOSS_L1_LOG && kind != APH_DN_LOSS_L2_TONEMAPPED)
        return fail(c, APH_E_INVALID, "unknown loss %d", kind);
    d->lossKind = kind;
    return APH_OK;
}

int aph_denoiser_tensor_size(aph_ctx_t c, uint64_t dn, int kind, uint32_t index, uint64_t* count)
{
    int rc = enter(c);
    if (rc) return rc;
    aph_denoiser* d = findDenoiser(c, dn);
    if (!d || !count) return fail(c, APH_E_INVALID, "unknown denoiser or null output");
    switch (kind) {
    case APH_DN_PARAMS:
    case APH_DN_PARAM_GRADS: *count = offsets().nParams; return APH_OK;
    case APH_DN_PREDICTION: *count = (uint64_t)d->levelPix(0) * kDnPredStride; return APH_OK;
    case APH_DN_ACTIVATIONS:
    case APH_DN_ACTIVATION_GRADS:
        if (index >= (uint32_t)kNumBufs) return fail(c, APH_E_INVALID, "buffer %u out of range", index);
        *count = (uint64_t)d->levelPix(kBufs[index].level) * kBufs[index].channels;
        return APH_OK;
    default: return fail(c, APH_E_INVALID, "unknown tensor kind %d", kind);
    }
}

int aph_denoiser_tensor(aph_ctx_t c, uint64_t dn, int kind, uint32_t index, float* out, uint64_t count)
{
    uint64_t need = 0;
    int rc = aph_denoiser_tensor_size(c, dn, kind, index, &need);
    if (rc) return rc;
    aph_denoiser* d = findDenoiser(c, dn);
    if (!out || count != need)
        return fail(c, APH_E_INVALID, "tensor holds %llu values, asked for %llu", (unsigned long long)need,
                    (unsigned long long)count);
    if (d->N == 0 && (kind == APH_DN_PREDICTION || kind == APH_DN_ACTIVATIONS || kind == APH_DN_ACTIVATION_GRADS))
        return fail(c, APH_E_STATE, "no forward pass has run");
    if (kind == APH_DN_ACTIVATIONS && index == B_OUT)
        return fail(c, APH_E_INVALID, "the prediction is APH_DN_PREDICTION");
    if (kind == APH_DN_ACTIVATION_GRADS && !d->haveGrads)
        return fail(c, APH_E_STATE, "activation gradients exist only after a training step");
    HIPCALL(c, hipStreamSynchronize(c->stream), "tensor readback");
    const void* src = nullptr;
    if (kind == APH_DN_PARAMS) src = d->params.p;
    else if (kind == APH_DN_PARAM_GRADS) src = d->grads.p;
    else if (kind == APH_DN_PREDICTION) src = d->pred.p;
    if (!src) {
        const __bf16* b = kind == APH_DN_ACTIVATIONS ? d->act((int)index) : d->grad((int)index);
        HIPCALL(c, d->tmp.reserve(need * 4), "alloc readback");
        HIPCALL(c, launchBf16ToF32(b, d->tmp.as<float>(), (int64_t)need, c->stream), "convert tensor");
        HIPCALL(c, hipStreamSynchronize(c->stream), "tensor readback");
        src = d->tmp.p;
    }
    HIPCALL(c, hipMemcpy(out, src, need * 4, hipMemcpyDeviceToHost), "tensor readback");
    return APH_OK;
}

int aph_denoiser_train_step(aph_ctx_t c, uint64_t dn, const aph_denoise_example* ex, uint32_t n, float lr,
                            float* loss)
{
    int rc = enter(c);
    if (rc) return rc;
    aph_denoiser* d = findDenoiser(c, dn);
    if (!d) return fail(c, APH_E_INVALID, "unknown denoiser %llu", (unsigned long long)dn);
    if (!ex || n == 0 || n > 256) return fail(c, APH_E_INVALID, "need 1..256 training examples");
    if (!(lr >= 0.0f) || !std::isfinite(lr)) return fail(c, APH_E_INVALID, "bad learning rate");
    std::vector<uint64_t> noisy(n), aov(n);
    std::vector<uint32_t> xf(n);
    for (uint32_t i = 0; i < n; i++) {
        noisy[i] = ex[i].noisy;
        aov[i] = ex[i].aov;
        xf[i] = ex[i].transform;
    }
    std::vector<Target*> tg;
    std::vector<AovSet*> av;
    uint32_t W = 0, H = 0;
    rc = resolveInputs(c, noisy.data(), aov.data(), n, tg, av, W, H);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; i++) {
        if (xf[i] > 7u) return fail(c, APH_E_INVALID, "example %u: transform %u out of range", i, xf[i]);
        if ((xf[i] & 4u) && W != H) return fail(c, APH_E_INVALID, "example %u: transposing needs a square image", i);
    }
    std::vector<Target*> refs(n);
    for (uint32_t i = 0; i < n; i++) {
        refs[i] = lookup(c->targets, ex[i].reference);
        if (!refs[i]) return fail(c, APH_E_INVALID, "example %u: unknown reference target", i);
        if (refs[i]->w != W || refs[i]->h != H) return fail(c, APH_E_INVALID, "example %u: reference size differs", i);
    }
    if ((uint64_t)n * padded(W) * padded(H) > (1ull << 28)) return fail(c, APH_E_INVALID, "training batch too large");
    const Offsets& o = offsets();
    hipStream_t st = c->stream;
    TimedScope ts(c, "denoiser_train_step");
    HIPCALL(c, reserveShape(d, (int)n, padded(H), padded(W), true), "alloc denoiser activations");
    HIPCALL(c, packInputs(d, tg, av, W, H, xf, st), "denoiser pack");
    HIPCALL(c, forward(d, st, false), "denoiser forward");
    HIPCALL(c, hipMemsetAsync(d->loss.p, 0, 4, st), "clear loss");
    HIPCALL(c, hipMemsetAsync(d->grads.p, 0, o.nParams * 4, st), "clear gradients");
    for (uint32_t i = 0; i < n; i++)
        HIPCALL(c, launchLoss(d->pred.as<float>(), refs[i]->accum.as<float4>(), (int)W, (int)H, d->Wp, d->Hp, (int)i,
                              (int)n, xf[i], d->lossKind, d->grad(B_OUT), kBufs[B_OUT].channels,
                              d->loss.as<float>(), st),
                "denoiser loss");
    HIPCALL(c, backward(d, st), "denoiser backward");
    d->haveGrads = true;
    d->step++;
    HIPCALL(c, launchAdam(d->params.as<float>(), d->adamM.as<float>(), d->adamV.as<float>(), d->grads.as<float>(),
                          (int64_t)o.nParams, lr, d->step, st),
            "denoiser Adam");
    HIPCALL(c, refreshWeights(d, st), "denoiser weights");
    float l = 0.0f;
    HIPCALL(c, hipMemcpyAsync(&l, d->loss.p, 4, hipMemcpyDeviceToHost, st), "loss readback");
    HIPCALL(c, hipStreamSynchronize(st), "denoiser train step");
    if (loss) *loss = l;
    return APH_OK;
}

int aph_denoise(aph_ctx_t c, uint64_t dn, const aph_denoise_item* items, uint32_t n)
{
    int rc = enter(c);
    if (rc) return rc;
    aph_denoiser* d = findDenoiser(c, dn);
    if (!d) return fail(c, APH_E_INVALID, "unknown denoiser %llu", (unsigned long long)dn);
    if (n == 0) return APH_OK;
    if (!items || n > 256) return fail(c, APH_E_INVALID, "need 1..256 images");
    std::vector<uint64_t> noisy(n), aov(n);
    for (uint32_t i = 0; i < n; i++) {
        noisy[i] = items[i].noisy;
        aov[i] = items[i].aov;
    }
    std::vector<Target*> tg;
    std::vector<AovSet*> av;
    uint32_t W = 0, H = 0;
    rc = resolveInputs(c, noisy.data(), aov.data(), n, tg, av, W, H);
    if (rc) return rc;
    std::vector<Target*> outs(n);
    for (uint32_t i = 0; i < n; i++) {
        outs[i] = lookup(c->targets, items[i].output);
        if (!outs[i]) return fail(c, APH_E_INVALID, "item %u: unknown output target", i);
        if (outs[i]->w != W || outs[i]->h != H) return fail(c, APH_E_INVALID, "item %u: output size differs", i);
        for (uint32_t j = 0; j < n; j++)
            if (j != i && (outs[i] == tg[j] || (j < i && outs[i] == outs[j])))
                return fail(c, APH_E_INVALID, "item %u: output aliases another item's image", i);
    }
    // batch images: one per item, or one per transform of an ensemble item
    std::vector<Target*> btg;
    std::vector<AovSet*> bav;
    std::vector<uint32_t> bxf;
    std::vector<uint32_t> bitem;
    for (uint32_t i = 0; i < n; i++) {
        if (items[i].flags & ~(uint32_t)APH_DN_ENSEMBLE) return fail(c, APH_E_INVALID, "item %u: unknown flags", i);
        const uint32_t K = (items[i].flags & APH_DN_ENSEMBLE) ? (W == H ? 8u : 4u) : 1u;
        for (uint32_t k = 0; k < K; k++) {
            btg.push_back(tg[i]);
            bav.push_back(av[i]);
            bxf.push_back(k);
            bitem.push_back(i);
        }
    }
    const uint32_t nb = (uint32_t)btg.size();
    if ((uint64_t)nb * padded(W) * padded(H) > (1ull << 28)) return fail(c, APH_E_INVALID, "denoise batch too large");
    hipStream_t st = c->stream;
    TimedScope ts(c, "denoise");
    HIPCALL(c, reserveShape(d, (int)nb, padded(H), padded(W), false), "alloc denoiser activations");
    HIPCALL(c, packInputs(d, btg, bav, W, H, bxf, st), "denoiser pack");
    HIPCALL(c, forward(d, st, true), "denoiser forward");
    for (uint32_t b = 0, k = 0; b < nb; b++) {
        const uint32_t i = bitem[b];
        k = (b > 0 && bitem[b - 1] == i) ? k + 1 : 0;
        HIPCALL(c, launchUnpack(d->pred.as<float>(), (int)W, (int)H, d->Wp, d->Hp, (int)b, bxf[b], (int)k,
                                tg[i]->accum.as<float4>(), outs[i]->accum.as<float4>(), st),
                "denoiser output");
    }
    return APH_OK;
}

}  // extern "C"
