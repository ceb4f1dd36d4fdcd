"""Benchmark of the gfx950 path tracer on the 1M-triangle 1080p workload.

One step = one full frame: 1920x1080 pixels x 64 samples x up to 5 bounces of
the path tracer (raygen, BVH traversal, Standard Surface shading, environment
MIS with shadow rays, accumulation), inputs resident in HBM.

Prints one JSON line.  value = Mrays/s over the timed steps, where a ray is a
BVH query actually executed (extension + shadow rays, counted by the kernels).
The roofline entry prices the dominant kernel by its algorithmic bytes (BVH
nodes, triangles, instance records, ray and hit records, accumulation) per
launch over its HIP-event launch time.  The CPU baseline is the oracle's
restatement of the reference path tracer on the host cores, on a bounded
pixel sample of the same frame.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from aurora_amd import capi, scenes  # noqa: E402
from aurora_amd.session import Renderer  # noqa: E402

HBM_PEAK_GBS = 8000.0
NODE_B, TRI_B, INST_B, RAY_B, HIT_B, ACC_B = 64, 48, 64, 32, 16, 32


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_sync():
    hip = C.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


def cpu_baseline(sc, budget_s=12.0):
    """Oracle path tracer on the host cores over a bounded sample of the same frame:
    random pixels, successive sample indices, until about budget_s of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    threads = max(1, min(16, os.cpu_count() or 1))
    ora = pyoracle.OracleScene(sc)
    ora.env()  # alias-table build outside the timed sample
    w, h = sc.width, sc.height
    npix = 4096
    rng = np.random.default_rng(0)
    total_rays = total_pix = 0
    total_time = 0.0
    sample = 0
    while total_time < budget_s and sample < sc.spp:
        pix = np.sort(rng.choice(w * h, size=min(npix, w * h), replace=False)).astype(np.uint32)
        t = time.perf_counter()
        _, cnt = ora.render(sc.frame, w, h, sample, 1, pixels=pix, threads=threads)
        dt = time.perf_counter() - t
        total_rays += cnt["extension_rays"] + cnt["shadow_rays"]
        total_time += dt
        total_pix += len(pix)
        sample += 1
        npix = int(min(w * h, max(npix, npix * (budget_s - total_time) / max(dt, 1e-3) / 2)))
    return {
        "value": round(total_rays / total_time / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{total_pix} random pixel-samples of the 1920x1080 frame ({sample} sample indices), 5 bounces: "
                  f"{total_rays} rays in {total_time:.1f} s on {threads} host threads",
        "samples_per_s": round(total_pix / total_time, 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--subdiv", type=int, default=4, help="4: 1.02M instanced tris; 5: 4.1M")
    ap.add_argument("--mode", choices=["megakernel", "wavefront"], default="wavefront")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        hip = C.CDLL("libamdhip64.so")
        n = 1 << 30
        a, b = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
        hip.hipMemset(a, 1, C.c_size_t(n))
        for _ in range(3):
            assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the path shards by sample range across GPUs, see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    sc = scenes.config2(subdiv=args.subdiv, spp=args.spp)
    mode = capi.APH_MODE_MEGAKERNEL if args.mode == "megakernel" else capi.APH_MODE_WAVEFRONT
    kname = "render_megakernel" if args.mode == "megakernel" else "wf_extend"
    log(f"scene: {sc.instanced_triangles} instanced / {sc.unique_triangles} unique triangles, "
        f"{len(sc.instances)} instances, {sc.width}x{sc.height} x {sc.spp} spp, mode {args.mode}")
    r = Renderer(0)
    tlas = r.load_scene(sc)
    tgt = r.target(sc.width, sc.height)
    req = [r.request(tgt, tlas, sc.frame, 0, sc.spp)]

    # Algorithmic bytes of one frame from the kernels' visit counters (deterministic).
    r.set_stats_level(1)
    r.reset_stats()
    r.render(req, mode)
    r.wait()
    st = r.stats()
    r.set_stats_level(0)
    queries = st["extension_rays"] + st["shadow_rays"]
    frame_bytes = (st["node_visits"] * NODE_B + st["tri_tests"] * TRI_B + st["inst_visits"] * INST_B
                   + queries * RAY_B + st["extension_rays"] * HIT_B + sc.width * sc.height * ACC_B)
    log(f"stats: {st}")

    for i in range(args.warmup):
        r.render(req, mode)
        r.wait()
        log(f"warmup step {i + 1}/{args.warmup} done")
    r.reset_stats()
    r.reset_timing()
    r.enable_timing(True)
    device_sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render(req, mode)
        r.wait()
        log(f"step {i + 1}/{args.steps} done")
    device_sync()
    wall = time.perf_counter() - t0
    r.enable_timing(False)
    tst = r.stats()
    rays = tst["extension_rays"] + tst["shadow_rays"]
    kms, klaunch = r.timing(kname)
    avg_launch_ms = kms / max(klaunch, 1)
    launches_per_frame = klaunch / max(args.steps, 1)
    bytes_per_launch = frame_bytes / max(launches_per_frame, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    mrays = rays / wall / 1e6
    out = {
        "metric": "Mrays/s (1M tri, 1080p)",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 shading, f64 ray/triangle",
        "data": "synthetic (seeded scene generator)",
        "config": {
            "workload": f"config2: {sc.width}x{sc.height}, {sc.spp} spp, 5 bounces, 200 instances of 5 displaced "
                        f"icospheres (subdiv {args.subdiv}), {sc.instanced_triangles} instanced triangles, "
                        "2048x1024 HDR environment with 1e4 sun",
            "mode": args.mode,
            "samples_per_s": round(sc.width * sc.height * sc.spp * args.steps / wall, 1),
            "extension_rays_per_frame": tst["extension_rays"] // args.steps,
            "shadow_rays_per_frame": tst["shadow_rays"] // args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_launch": int(bytes_per_launch),
            "avg_launch_ms": round(avg_launch_ms, 3),
        },
    }
    if args.mode == "wavefront":
        # the plain one-request-at-a-time form on the same frame, for the batched-vs-plain ratio
        r.reset_stats()
        t1 = time.perf_counter()
        r.render(req, capi.APH_MODE_MEGAKERNEL)
        r.wait()
        t_plain = time.perf_counter() - t1
        pst = r.stats()
        out["config"]["plain_megakernel_mrays_s"] = round(
            (pst["extension_rays"] + pst["shadow_rays"]) / t_plain / 1e6, 2)
        log(f"plain megakernel frame: {t_plain * 1e3:.1f} ms")
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)...")
        out["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
