"""Learned denoiser over the C-ABI (aph_denoiser_* / aph_denoise).

A U-Net over the noisy radiance of a target and the guides of an AOV set,
trained and run by the bf16 MFMA kernels of libaph.so.  The reference hands
the same guides to NVIDIA NRD (DirectX/Denoiser.h:24-35); this network is
its own (no reference counterpart); train_denoiser.py trains it on the path
tracer's own low-spp / 4096-spp renders.
"""
import ctypes as C
import os

import numpy as np

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "denoiser_unet.npy")
ALIGN = 16


def arch():
    """(layers, buffers, ops) of the network as the library defines them."""
    lib = capi.load()
    nl, nb, no = C.c_uint32(), C.c_uint32(), C.c_uint32()
    assert lib.aph_denoiser_arch(None, C.byref(nl), None, C.byref(nb), None, C.byref(no)) == capi.APH_OK
    layers = (capi.DnLayer * nl.value)()
    bufs = (capi.DnBuffer * nb.value)()
    ops = (capi.DnOp * no.value)()
    assert lib.aph_denoiser_arch(C.cast(layers, C.c_void_p), C.byref(nl), C.cast(bufs, C.c_void_p), C.byref(nb),
                                 C.cast(ops, C.c_void_p), C.byref(no)) == capi.APH_OK
    return list(layers), list(bufs), list(ops)


def param_count():
    return int(capi.load().aph_denoiser_param_count())


def padded(v):
    return (v + ALIGN - 1) // ALIGN * ALIGN


class Denoiser:
    """One network on a Renderer's context (aph_denoiser_create / _destroy)."""

    def __init__(self, renderer, seed=1, params=None):
        self.r = renderer
        self.lib = renderer.lib
        h = C.c_uint64()
        self.r._check(self.lib.aph_denoiser_create(self.r.ctx, seed, C.byref(h)))
        self.h = h.value
        self.layers, self.buffers, self.ops = arch()
        if params is not None:
            self.set_params(params)

    def close(self):
        if self.h and self.r.ctx:
            self.lib.aph_denoiser_destroy(self.r.ctx, self.h)
        self.h = 0

    # ---- parameters
    def params(self):
        return self.tensor(capi.APH_DN_PARAMS)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
        self.r._check(self.lib.aph_denoiser_set_params(self.r.ctx, self.h, capi.ptr(p), p.size))

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.save(path, self.params())

    def load(self, path=DEFAULT_WEIGHTS):
        self.set_params(np.load(path, allow_pickle=False))

    # ---- training and inference
    def set_loss(self, kind):
        self.r._check(self.lib.aph_denoiser_set_loss(self.r.ctx, self.h, kind))

    def train_step(self, examples, lr):
        """examples: [(noisy_target, aov, reference_target[, transform])] of one size; returns the loss."""
        arr = (capi.DenoiseExample * len(examples))(*[capi.DenoiseExample(*e) for e in examples])
        loss = C.c_float()
        self.r._check(self.lib.aph_denoiser_train_step(self.r.ctx, self.h, C.cast(arr, C.c_void_p), len(examples),
                                                       float(lr), C.byref(loss)))
        return float(loss.value)

    def denoise(self, items):
        """items: [(noisy_target, aov, output_target[, flags])] of one size (asynchronous);
        flags capi.APH_DN_ENSEMBLE averages over mirrored / transposed passes."""
        arr = (capi.DenoiseItem * len(items))(*[capi.DenoiseItem(*i) for i in items])
        self.r._check(self.lib.aph_denoise(self.r.ctx, self.h, C.cast(arr, C.c_void_p), len(items)))

    # ---- inspection of the last pass
    def tensor(self, kind, index=0):
        n = C.c_uint64()
        self.r._check(self.lib.aph_denoiser_tensor_size(self.r.ctx, self.h, kind, index, C.byref(n)))
        out = np.zeros(n.value, dtype=np.float32)
        self.r._check(self.lib.aph_denoiser_tensor(self.r.ctx, self.h, kind, index, capi.ptr(out), n.value))
        return out

    def activation(self, buf, n, hp, wp, grad=False):
        """Buffer `buf` of the last pass as [n, Hp >> level, Wp >> level, channels]."""
        b = self.buffers[buf]
        kind = capi.APH_DN_ACTIVATION_GRADS if grad else capi.APH_DN_ACTIVATIONS
        return self.tensor(kind, buf).reshape(n, hp >> b.level, wp >> b.level, b.channels)

    def layer_params(self, p, i):
        """(weights [9][cout_p][cin_p], bias [cout_p]) of layer i in parameter vector p."""
        L = self.layers[i]
        nw = 9 * L.cout_padded * L.cin_padded
        w = p[L.weight_offset:L.weight_offset + nw].reshape(9, L.cout_padded, L.cin_padded)
        return w, p[L.bias_offset:L.bias_offset + L.cout_padded]


def main():
    """Times aph_denoise on a 4-spp 1080p frame of config2 (the committed weights):
    python -m aurora_amd.denoiser [--runs N] [--ensemble]."""
    import argparse
    import time

    from . import scenes
    from .session import Renderer

    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=20)
    ap.add_argument("--ensemble", action="store_true")
    args = ap.parse_args()
    r = Renderer(0)
    sc = scenes.config2()
    w, h = sc.width, sc.height
    tlas = r.load_scene(sc, capi.APH_BUILDER_SAH)
    r.set_launch_mode(capi.APH_LAUNCH_EAGER)
    noisy, out, aov = r.target(w, h), r.target(w, h), r.aov(w, h)
    r.render([r.request(noisy, tlas, sc.frame, 0, 4)], capi.APH_MODE_WAVEFRONT)
    vp = scenes.view_proj(sc.frame)
    r.render_aovs(aov, tlas, sc.frame, vp, vp, np.array(sc.frame.inv_view, dtype=np.float32))
    dn = Denoiser(r)
    dn.load()
    item = [(noisy, aov, out, capi.APH_DN_ENSEMBLE if args.ensemble else 0)]
    dn.denoise(item)
    r.wait()
    t = time.perf_counter()
    for _ in range(args.runs):
        dn.denoise(item)
    r.wait()
    ms = (time.perf_counter() - t) / args.runs * 1e3
    print(f"denoise {w}x{h}{' ensemble' if args.ensemble else ''}: {ms:.3f} ms per call over {args.runs} calls",
          flush=True)
    dn.close()
    r.close()


if __name__ == "__main__":
    main()
