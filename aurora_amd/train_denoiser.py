"""Trains the denoiser on the path tracer's own low-spp / 4096-spp pairs.

Data: `--scenes` seeded scenes, alternately of the config-1 family (ground +
16 icospheres, sky gradient) and the config-2 family (displaced icospheres of
subdivision 3 or 4 with transmission, coat and a decal layer under a sun sky),
each seen by one random camera (eye, target and field of view drawn from the
scene's seed; config-2 seeds start at 1000, so the benchmark's seed-2 scene and
its camera are never trained on).  Per scene
the wavefront path tracer renders a 4096-spp reference and noisy inputs (two
4-spp and one 1-spp realisation with seeds disjoint from the reference's), and
the guide images of the same camera.  Everything stays resident on the GPU.

Training: Adam (aph_denoiser_train_step) on the L1 loss of log1p radiance or
the L2 loss of tonemapped radiance x / (1 + x), random batches of one image
size, each pair randomly mirrored / transposed, linear warm-up then cosine
decay.  The last
`--holdout` scenes are never trained on; their denoised error is reported.

    python -m aurora_amd.train_denoiser --out aurora_amd/weights/denoiser_unet.npy
"""
import argparse
import json
import math
import os
import time

import numpy as np

from . import capi, scenes
from .denoiser import DEFAULT_WEIGHTS, Denoiser
from .session import Renderer


LOSSES = {"l1log": capi.APH_DN_LOSS_L1_LOG, "l2tm": capi.APH_DN_LOSS_L2_TONEMAPPED}


def scene_for(i, size, seed0):
    """Scene i of the training set and its camera (every fourth of the config-1 family)."""
    seed = seed0 + i
    rng = np.random.Generator(np.random.PCG64(seed * 7919 + 13))
    if i % 4 == 0:
        sc = scenes.config1(seed=seed, width=size, height=size, trace_depth=5)
        radius, target, yr = rng.uniform(8.0, 16.0), (0.0, 0.6, 0.0), (2.0, 9.0)
        fov = float(rng.uniform(25.0, 60.0))
    else:
        sc = scenes.config2(seed=seed, subdiv=3 + i % 2, width=size, height=size, env_w=1024, env_h=512,
                            instances=int(rng.integers(60, 200)), trace_depth=5)
        radius, target, yr = rng.uniform(16.0, 40.0), (0.0, 1.0, 0.0), (4.0, 18.0)
        fov = float(rng.uniform(6.0, 26.0))
    th = rng.uniform(0.0, 2.0 * math.pi)
    eye = (radius * math.cos(th), float(rng.uniform(*yr)), radius * math.sin(th))
    tgt = tuple(float(t + rng.uniform(-2.0, 2.0)) for t in target)
    cam = scenes.look_at_frame(eye, tgt, (0, 1, 0), fov, 1.0, 5)
    return sc, cam


def render_pairs(r, n_scenes, size, ref_spp, seed0, log):
    """[(noisy, aov, reference)] per scene (lists of examples per scene)."""
    per_scene = []
    t0 = time.perf_counter()
    for i in range(n_scenes):
        sc, cam = scene_for(i, size, seed0)
        tlas = r.load_scene(sc)
        ref = r.target(size, size)
        r.render([r.request(ref, tlas, cam, 0, ref_spp, 1_000_003)], capi.APH_MODE_WAVEFRONT)
        aov = r.aov(size, size)
        vp = scenes.view_proj(cam)
        r.render_aovs(aov, tlas, cam, vp, vp, np.array(cam.inv_view, dtype=np.float32))
        ex = []
        for k, spp in enumerate((4, 4, 1)):
            t = r.target(size, size)
            r.render([r.request(t, tlas, cam, 0, spp, 17 + 101 * k)], capi.APH_MODE_WAVEFRONT)
            ex.append((t, aov, ref))
        per_scene.append(ex)
        if (i + 1) % 16 == 0 or i + 1 == n_scenes:
            r.wait()
            log(f"rendered {i + 1}/{n_scenes} scenes ({time.perf_counter() - t0:.1f} s)")
    r.wait()
    return per_scene


def evaluate(r, dn, examples, size):
    """Means over examples of the noisy and denoised errors: L1 of log1p radiance,
    RMSE of radiance and RMSE of x / (1 + x)."""
    out = r.target(size, size)
    names = ["l1_log", "rmse", "rmse_tonemapped"]
    acc = np.zeros((2, 3))
    tm = lambda x: np.maximum(x, 0) / (1.0 + np.maximum(x, 0))  # noqa: E731
    for noisy, aov, ref in examples:
        dn.denoise([(noisy, aov, out)])
        g = np.maximum(r.readback(ref, size, size)[..., :3].astype(np.float64), 0)
        for k, t in enumerate((noisy, out)):
            x = np.maximum(r.readback(t, size, size)[..., :3].astype(np.float64), 0)
            acc[k] += [np.abs(np.log1p(x) - np.log1p(g)).mean(), np.sqrt(np.mean((x - g) ** 2)),
                       np.sqrt(np.mean((tm(x) - tm(g)) ** 2))]
    acc /= max(1, len(examples))
    return {f"{w}_{n}": float(acc[k, j]) for k, w in enumerate(("noisy", "denoised")) for j, n in enumerate(names)}


def train(r, args, train_set, held, loss_name, out_path, log):
    dn = Denoiser(r, seed=args.init_seed)
    if args.resume:
        dn.load(args.resume)  # fine-tuning: start from saved weights (fresh Adam moments)
    dn.set_loss(LOSSES[loss_name])
    rng = np.random.Generator(np.random.PCG64(args.seed))
    warm = max(1, args.steps // 40)
    losses = []
    t0 = time.perf_counter()
    for step in range(args.steps):
        if step < warm:
            lr = args.lr * (step + 1) / warm
        else:
            q = (step - warm) / max(1, args.steps - warm)
            lr = 2e-5 + 0.5 * (args.lr - 2e-5) * (1.0 + math.cos(math.pi * q))
        pick = rng.choice(len(train_set), args.batch, replace=False)
        xf = rng.integers(0, 8 if args.augment else 1, args.batch)
        batch = [train_set[j] + (int(t),) for j, t in zip(pick, xf)]
        losses.append(dn.train_step(batch, lr))
        if (step + 1) % 500 == 0:
            log(f"[{loss_name}] step {step + 1}/{args.steps} lr {lr:.2e} loss {np.mean(losses[-500:]):.6f} "
                f"({time.perf_counter() - t0:.1f} s)")
    t_train = time.perf_counter() - t0
    resumed = None
    if args.resume:
        prior = os.path.splitext(args.resume)[0] + ".json"
        resumed = {"weights": os.path.basename(args.resume)}
        if os.path.exists(prior):
            with open(prior) as f:
                p = json.load(f)
            resumed.update({k: p[k] for k in ("steps", "batch", "lr", "resumed_from") if k in p})
    ev = evaluate(r, dn, held, args.size)
    log(f"[{loss_name}] held-out: " + json.dumps({k: round(v, 5) for k, v in ev.items()}))
    dn.save(out_path)
    summary = dict(scenes=args.scenes, holdout=args.holdout, size=args.size, ref_spp=args.ref_spp,
                   noisy_spp=[4, 4, 1], loss=loss_name, augment=bool(args.augment), steps=args.steps,
                   batch=args.batch, lr=args.lr, seed=args.seed, init_seed=args.init_seed,
                   resumed_from=resumed,
                   train_seconds=round(t_train, 1),
                   loss_curve=[float(np.mean(losses[i:i + 500])) for i in range(0, len(losses), 500)],
                   held_out={k: round(v, 6) for k, v in ev.items()})
    with open(os.path.splitext(out_path)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(f"saved {out_path}")
    dn.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--holdout", type=int, default=16)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=44000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--loss", default="l1log",
                    help="l1log, l2tm, or a comma-separated list: each trains its own network on the same data "
                         "and saves <out>_<loss>.npy")
    ap.add_argument("--no-augment", dest="augment", action="store_false",
                    help="no random mirroring / transposition of the training pairs")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--resume", default=None,
                    help="weights (.npy) to start from instead of a random initialisation; the schedule "
                         "(warm-up, cosine decay from --lr) runs again over --steps")
    args = ap.parse_args()

    def log(msg):
        print(f"[train_denoiser] {msg}", flush=True)

    losses = [x for x in args.loss.split(",") if x]
    for x in losses:
        if x not in LOSSES:
            raise SystemExit(f"unknown loss {x}")
    r = Renderer(0)
    per_scene = render_pairs(r, args.scenes, args.size, args.ref_spp, args.seed, log)
    train_set = [e for ex in per_scene[:args.scenes - args.holdout] for e in ex]
    held = [e for ex in per_scene[args.scenes - args.holdout:] for e in ex]
    log(f"{len(train_set)} training pairs, {len(held)} held-out pairs of {args.size}x{args.size}")
    for x in losses:
        out = args.out if len(losses) == 1 else f"{os.path.splitext(args.out)[0]}_{x}.npy"
        train(r, args, train_set, held, x, out, log)
    r.close()


if __name__ == "__main__":
    main()
