"""Benchmark of the gfx950 path tracer on the 1M-triangle 1080p workload.

One step = one full frame: 1920x1080 pixels x 64 samples x up to 5 bounces of
the path tracer (raygen, BVH traversal, Standard Surface shading, environment
MIS with shadow rays, accumulation), inputs resident in HBM.

Prints one JSON line.  value = Mrays/s over the timed steps, where a ray is a
BVH query actually executed (extension + shadow rays, counted by the kernels).
The roofline entry prices the dominant kernel by its algorithmic bytes (BVH
nodes, triangles, instance records, ray and hit records, accumulation) per
launch over its HIP-event launch time.  Its HBM traffic comes from two
rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE) over one frame, run as
child processes before this process touches the GPU; each counter is scaled by
what it reads for 1 GiB device-to-device copies made in the same pass.  The
CPU baseline is the oracle's restatement of the reference path tracer on the
host cores, on a bounded pixel sample of the same frame.
"""
import argparse
import ctypes as C
import glob
import json
import os
import shutil
import sqlite3
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from aurora_amd import capi, scenes  # noqa: E402
from aurora_amd import denoiser as dnmod  # noqa: E402
from aurora_amd.session import Renderer  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per unit of work: a four-child BVH node (four child boxes +
# refs, one 128 B line), a triangle record, an instance traversal record, a ray
# (origin + request tag, direction + path id), a hit record (wavefront:
# t/inst/prim/u + v; megakernel: kept in registers, the 16 B of the RGBA32F sample
# is charged to accumulation) and one RGBA32F accumulation read+write per pixel.
NODE_B, TRI_B, INST_B, RAY_B, HIT_B, ACC_B = 128, 48, 64, 32, 20, 32


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_sync():
    hip = C.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


KERNEL_SYMBOL = {"render_megakernel": "kRenderMega<false>", "wf_extend": "kExtend<false>"}
COPY_BYTES, COPIES = 1 << 30, 3


def copy_calibration():
    """Three 1 GiB device-to-device copies: the plain copy kernel whose counter
    readings calibrate FETCH_SIZE and WRITE_SIZE."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(COPY_BYTES)) == 0
    for _ in range(COPIES):
        assert hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def pmc_dispatches(db_dir, counter):
    """(kernel name, counter value in KiB) per dispatch from a rocprofv3 database."""
    rows = []
    for db in glob.glob(os.path.join(db_dir, "**", "*.db"), recursive=True):
        con = sqlite3.connect(db)
        rows += con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                            (counter,)).fetchall()
        con.close()
    return rows


def measure_traffic(args, kname):
    """HBM bytes per launch of the dominant kernel: FETCH_SIZE and WRITE_SIZE per
    dispatch (two rocprofv3 passes over one frame), each scaled by actual bytes /
    counter reading of the calibration copies in the same pass."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    sym = KERNEL_SYMBOL[kname]
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="aph_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "240", prof, "--pmc", counter, "-d", d, "-o", "pmc", "--",
               sys.executable, "-u", os.path.abspath(__file__), "--pmc-child", "--spp", str(args.spp),
               "--subdiv", str(args.subdiv), "--mode", args.mode, "--builder", args.builder,
               "--tuning", args.tuning]
        log(f"pmc pass {counter}: {' '.join(cmd)}")
        env = dict(os.environ, TMPDIR="/tmp")
        p = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        rows = pmc_dispatches(d, counter)
        shutil.rmtree(d, ignore_errors=True)
        if p.returncode != 0 or not rows:
            log(p.stdout[-2000:])
            return None, f"{counter} pass failed (rc {p.returncode})"
        kern = [v for n, v in rows if sym in n]
        copies = sorted((v for n, v in rows if "copyBuffer" in n), reverse=True)[:COPIES]
        if not kern or len(copies) < COPIES:
            return None, f"{counter} pass: {len(kern)} {sym} dispatches, {len(copies)} copies"
        factor = COPY_BYTES * COPIES / (sum(copies) * 1024.0)
        res[counter] = (sum(kern) / len(kern), factor, len(kern))
        log(f"{counter}: {res[counter][0]:.1f} KiB per {sym} dispatch over {len(kern)}, copy factor {factor:.4f}")
    f_kib, f_fac, nd = res["FETCH_SIZE"]
    w_kib, w_fac, _ = res["WRITE_SIZE"]
    detail = {"fetch_kib_per_dispatch": round(f_kib, 1), "fetch_factor": round(f_fac, 4),
              "write_kib_per_dispatch": round(w_kib, 1), "write_factor": round(w_fac, 4), "dispatches": nd}
    return int((f_kib * f_fac + w_kib * w_fac) * 1024), detail


def apply_tuning(r, tuning):
    if tuning:
        r.set_tuning(*(int(v) for v in tuning.split(",")))


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


def autotune_child(args):
    """aph_wavefront_autotune on the headline frame in a child process (before this
    process touches the GPU): the launch configuration the measured line uses."""
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--autotune-child", "--spp", str(args.spp),
           "--subdiv", str(args.subdiv), "--builder", args.builder]
    log("autotune (child process)...")
    try:
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    except subprocess.TimeoutExpired:
        return None
    if p.returncode != 0:
        log(p.stderr[-3000:])
        return None
    return json.loads(p.stdout.strip().splitlines()[-1])


def cpu_baseline_child(args):
    """Runs cpu_baseline in a child process that never touches the GPU, so that a
    failure there costs the baseline and not the measured line."""
    cmd = [sys.executable, "-X", "faulthandler", "-u", os.path.abspath(__file__), "--cpu-baseline-child",
           "--spp", str(args.spp), "--subdiv", str(args.subdiv)]
    try:
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    except subprocess.TimeoutExpired:
        return {"value": None, "error": "CPU baseline child timed out"}
    if p.returncode != 0:
        log(p.stderr[-3000:])
        return {"value": None, "error": f"CPU baseline child failed (rc {p.returncode})"}
    return json.loads(p.stdout.strip().splitlines()[-1])


RESOLUTIONS = scenes.C3_RESOLUTIONS


def rays_of(st):
    return st["extension_rays"] + st["shadow_rays"]


def literal_config2(r, builder, spp):
    """One frame of config2 as specified (icospheres subdiv 5: 4,096,002 instanced triangles)."""
    sc = scenes.config2(subdiv=5, spp=spp)
    tlas = r.load_scene(sc, builder)
    tgt = r.target(sc.width, sc.height)
    req = [r.request(tgt, tlas, sc.frame, 0, sc.spp)]
    r.render(req, capi.APH_MODE_WAVEFRONT)
    r.wait()
    r.reset_stats()
    t = time.perf_counter()
    r.render(req, capi.APH_MODE_WAVEFRONT)
    r.wait()
    dt = time.perf_counter() - t
    return {"instanced_triangles": sc.instanced_triangles, "unique_triangles": sc.unique_triangles,
            "mrays_s": round(rays_of(r.stats()) / dt / 1e6, 2), "ms_per_frame": round(dt * 1e3, 1)}


def c3_cameras(r, builder, seed=3, autotune=True):
    """64 simultaneous cameras on the config1 scene, resolutions from RESOLUTIONS, spp U{1..64},
    5 bounces: all in one batched wavefront call vs 64 sequential plain megakernel launches.
    The batched call runs with the autotuner's choice for this request set (tuned on the
    same 64 cameras at 1/8 of their spp); the context's previous tuning is restored after."""
    sc = scenes.config1()
    tlas = r.load_scene(sc, builder)
    cams = scenes.c3_cameras(seed)
    tg = [r.target(w, h) for w, h, _, _ in cams]
    reqs = [r.request(t, tlas, fr, 0, spp) for t, (w, h, spp, fr) in zip(tg, cams)]
    keep = r.tuning()
    tuned = None
    if autotune:
        small = [r.request(t, tlas, fr, 0, max(1, spp // 8)) for t, (w, h, spp, fr) in zip(tg, cams)]
        best = r.autotune(small)
        tuned = [best.chunk_paths, best.trace_blocks_per_cu, best.shade_blocks_per_cu, best.trace_refill_lanes,
                 best.trace_leaf_wait_lanes]
        log(f"c3 autotune: {tuned}")
    r.render(reqs[:4], capi.APH_MODE_WAVEFRONT)  # warm
    r.wait()
    r.reset_stats()
    t = time.perf_counter()
    r.render(reqs, capi.APH_MODE_WAVEFRONT)
    r.wait()
    dt_b = time.perf_counter() - t
    r.set_tuning(keep.chunk_paths, keep.trace_blocks_per_cu, keep.shade_blocks_per_cu, keep.trace_refill_lanes,
                 keep.trace_leaf_wait_lanes)
    rays_b = rays_of(r.stats())
    samples = sum(w * h * spp for w, h, spp, _ in cams)
    r.render(reqs[:1], capi.APH_MODE_MEGAKERNEL)  # warm
    r.wait()
    r.reset_stats()
    t = time.perf_counter()
    for q in reqs:
        r.render([q], capi.APH_MODE_MEGAKERNEL)
        r.wait()
    dt_p = time.perf_counter() - t
    rays_p = rays_of(r.stats())
    b, p = rays_b / dt_b / 1e6, rays_p / dt_p / 1e6
    return {"cameras": 64, "samples": samples, "batched_mrays_s": round(b, 2), "plain_mrays_s": round(p, 2),
            "batched_vs_plain": round(b / p, 2), "batched_s": round(dt_b, 3), "plain_s": round(dt_p, 3),
            "batched_samples_per_s": round(samples / dt_b, 1), "plain_samples_per_s": round(samples / dt_p, 1),
            "plain": "64 sequential megakernel launches, one per camera",
            "launch_config": tuned or [keep.chunk_paths, keep.trace_blocks_per_cu, keep.shade_blocks_per_cu,
                                       keep.trace_refill_lanes, keep.trace_leaf_wait_lanes],
            "launch_config_source": "autotuner on these 64 cameras at 1/8 spp" if tuned else "context"}


def serving(r, builder, jobs=24, rates=(4.0, 16.0), max_spp=32, seed=5, budget_gib=24.0):
    """Poisson arrivals of render jobs over the four resident serving scenes (about 10k, 100k,
    1M and 5M triangles); resolution from RESOLUTIONS, spp U{1..max_spp}; HBM-budget admission.
    Batched (continuous batching in shared wavefronts) vs plain (one job at a time, megakernel)."""
    from aurora_amd.session import Server
    scs = scenes.serving_scenes()
    tl = r.load_scenes(scs, builder)
    rng = np.random.default_rng(seed)
    out = {"scenes_triangles": [s.instanced_triangles for s in scs], "jobs": jobs, "max_spp": max_spp,
           "hbm_budget_gib": budget_gib, "rates": {}}
    js = []
    for _ in range(jobs):
        k = int(rng.integers(0, 4))
        w, h = RESOLUTIONS[int(rng.integers(0, 4))]
        js.append(Server.job(tl[k], scenes.orbit_frame(rng, w, h, radius=30.0 if k else 14.0), w, h,
                             int(rng.integers(1, max_spp + 1))))
    for rate in rates:
        arr = np.cumsum(rng.exponential(1.0 / rate, jobs))
        res = {}
        for name, mode in (("batched", capi.APH_MODE_WAVEFRONT), ("plain", capi.APH_MODE_MEGAKERNEL)):
            srv = Server(r, hbm_budget=int(budget_gib * (1 << 30)), mode=mode, keep_results=False)
            r.reset_stats()
            t = time.perf_counter()
            tickets = srv.replay(js, arr)
            dt = time.perf_counter() - t
            lat = np.array([srv.status(tk).t_done - srv.status(tk).t_arrival for tk in tickets])
            resident = srv.usage()[0]
            srv.close()
            res[name] = {"p50_s": round(float(np.percentile(lat, 50)), 3),
                         "p99_s": round(float(np.percentile(lat, 99)), 3),
                         "jobs_per_s": round(jobs / dt, 2), "mrays_s": round(rays_of(r.stats()) / dt / 1e6, 2),
                         "wall_s": round(dt, 2)}
            log(f"serving rate {rate}/s {name}: {res[name]}")
        out["rates"][str(rate)] = res
    out["resident_scene_bytes"] = int(resident)
    return out


def equal_time_rmse(r, builder, width=480, height=270, ref_spp=4096, plain_spp=64):
    """Equal-time RMSE vs a 4096-spp reference (config2 at reduced resolution): the plain
    megakernel's 64-spp image against the wavefront image with as many samples as it renders
    in the same time (same sampler, so images differ only by sample count)."""
    sc = scenes.config2(width=width, height=height)
    tlas = r.load_scene(sc, builder)
    ref = r.target(width, height)
    r.render([r.request(ref, tlas, sc.frame, 0, ref_spp)], capi.APH_MODE_WAVEFRONT)
    r.wait()
    ref_img = r.readback(ref, width, height)[..., :3].astype(np.float64)
    tp = r.target(width, height)
    t = time.perf_counter()
    r.render([r.request(tp, tlas, sc.frame, 0, plain_spp)], capi.APH_MODE_MEGAKERNEL)
    r.wait()
    t_plain = time.perf_counter() - t
    tw = r.target(width, height)
    t = time.perf_counter()
    r.render([r.request(tw, tlas, sc.frame, 0, plain_spp)], capi.APH_MODE_WAVEFRONT)
    r.wait()
    t_wf = time.perf_counter() - t
    spp_eq = max(plain_spp, int(plain_spp * t_plain / t_wf))
    if spp_eq > plain_spp:
        r.render([r.request(tw, tlas, sc.frame, plain_spp, spp_eq - plain_spp)], capi.APH_MODE_WAVEFRONT)
        r.wait()

    def rmse(tgt):
        return float(np.sqrt(np.mean((r.readback(tgt, width, height)[..., :3] - ref_img) ** 2)))
    a, b = rmse(tp), rmse(tw)
    return {"resolution": f"{width}x{height}", "reference_spp": ref_spp, "time_s": round(t_plain, 3),
            "plain_spp": plain_spp, "plain_rmse": round(a, 5), "wavefront_spp": spp_eq, "wavefront_rmse": round(b, 5),
            "rmse_ratio": round(a / b, 3)}


def autotune_section(r, sc, tlas, spp=32):
    """aph_wavefront_autotune on the headline frame at reduced spp: the chosen launch
    configuration and its frame time against the default one (the context keeps
    the chosen one)."""
    tgt = r.target(sc.width, sc.height)
    req = [r.request(tgt, tlas, sc.frame, 0, spp)]
    default = r.tuning()

    def frame_ms():
        r.render(req, capi.APH_MODE_WAVEFRONT)
        r.wait()
        t = time.perf_counter()
        r.render(req, capi.APH_MODE_WAVEFRONT)
        r.wait()
        return (time.perf_counter() - t) * 1e3

    base = frame_ms()
    best = r.autotune(req)
    tuned = frame_ms()
    out = {"spp": spp, "default": [default.chunk_paths, default.trace_blocks_per_cu, default.shade_blocks_per_cu,
                                   default.trace_refill_lanes, default.trace_leaf_wait_lanes],
           "default_ms": round(base, 2),
           "chosen": [best.chunk_paths, best.trace_blocks_per_cu, best.shade_blocks_per_cu, best.trace_refill_lanes,
                      best.trace_leaf_wait_lanes],
           "chosen_ms": round(tuned, 2), "candidates": 57}
    return out


def config5_sequence(r, builder, frames=120, spp=4, width=1920, height=1080):
    """Config 5 (24.7 M instanced triangles) along its 120-frame camera spline at 1080p:
    per frame a 4-spp wavefront render, the guide images, the denoiser (committed
    weights) and the temporal resolve, each frame synchronized; latency p50/p99."""
    t = time.perf_counter()
    sc = scenes.config5(width=width, height=height)
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    tlas = r.load_scene(sc, builder)
    r.wait()
    t_load = time.perf_counter() - t
    cams = scenes.camera_spline(sc.seed, frames, width, height)
    dn = dnmod.Denoiser(r)
    dn.load()
    noisy, den, aov = r.target(width, height), r.target(width, height), r.aov(width, height)
    hist = [r.target(width, height), r.target(width, height)]
    lat = []
    prev = cams[0]
    r.reset_stats()
    for i, fr in enumerate(cams):
        t = time.perf_counter()
        r.render([r.request(noisy, tlas, fr, 0, spp, i)], capi.APH_MODE_WAVEFRONT)
        r.render_aovs(aov, tlas, fr, scenes.view_proj(fr), scenes.view_proj(prev),
                      np.array(prev.inv_view, dtype=np.float32))
        dn.denoise([(noisy, aov, den)])
        r.temporal_resolve(den, aov, hist[i & 1], hist[(i + 1) & 1], is_reset=(i == 0))
        r.wait()
        lat.append(time.perf_counter() - t)
        prev = fr
    dn.close()
    lat = np.array(lat[1:]) * 1e3  # frame 0 allocates
    return {"scene": f"config5: {len(sc.instances) - 1} rocks of 2048 triangles on a {len(sc.meshes[0][2])}-triangle "
                     f"terrain, {sc.instanced_triangles} instanced triangles",
            "resolution": f"{width}x{height}", "spp": spp, "frames": frames,
            "pipeline": "wavefront render + guide images + U-Net + temporal resolve",
            "generate_s": round(t_gen, 2), "load_build_s": round(t_load, 3),
            "frame_ms_p50": round(float(np.percentile(lat, 50)), 2),
            "frame_ms_p99": round(float(np.percentile(lat, 99)), 2),
            "fps": round(1e3 / float(np.mean(lat)), 1)}


def config5_quality(r, builder, frames=120, spp=4, width=1920, height=1080, score=(29, 59, 89, 119), ref_spp=1024):
    """Image quality of the config-5 sequence: the same per-frame pipeline as
    config5_sequence, and at the frames in `score` the 4-spp render, the denoised
    frame and the temporally resolved output are compared with a `ref_spp` render of
    the same camera (disjoint seeds).  RMSE of linear radiance and of x / (1 + x)."""
    sc = scenes.config5(width=width, height=height)
    tlas = r.load_scene(sc, builder)
    cams = scenes.camera_spline(sc.seed, frames, width, height)
    dn = dnmod.Denoiser(r)
    dn.load()
    noisy, den, aov = r.target(width, height), r.target(width, height), r.aov(width, height)
    hist = [r.target(width, height), r.target(width, height)]
    kept = {}
    prev = cams[0]
    for i, fr in enumerate(cams):
        r.render([r.request(noisy, tlas, fr, 0, spp, i)], capi.APH_MODE_WAVEFRONT)
        r.render_aovs(aov, tlas, fr, scenes.view_proj(fr), scenes.view_proj(prev),
                      np.array(prev.inv_view, dtype=np.float32))
        dn.denoise([(noisy, aov, den)])
        r.temporal_resolve(den, aov, hist[i & 1], hist[(i + 1) & 1], is_reset=(i == 0))
        if i in score:
            r.wait()
            kept[i] = [r.readback(t, width, height)[..., :3].astype(np.float64) for t in (noisy, den, hist[(i + 1) & 1])]
        prev = fr
    dn.close()
    tm = lambda x: np.maximum(x, 0) / (1.0 + np.maximum(x, 0))  # noqa: E731
    ref = r.target(width, height)
    per = {}
    for i, imgs in kept.items():
        r.render([r.request(ref, tlas, cams[i], 0, ref_spp, 1_000_003 + i)], capi.APH_MODE_WAVEFRONT)
        r.wait()
        g = r.readback(ref, width, height)[..., :3].astype(np.float64)
        per[i] = {name: [round(float(np.sqrt(np.mean((x - g) ** 2))), 5),
                         round(float(np.sqrt(np.mean((tm(x) - tm(g)) ** 2))), 5)]
                  for name, x in zip(("noisy", "denoised", "resolved"), imgs)}
        log(f"config5 quality frame {i}: {per[i]}")
    mean = {name: [round(float(np.mean([per[i][name][k] for i in per])), 5) for k in (0, 1)]
            for name in ("noisy", "denoised", "resolved")}
    return {"frames_scored": list(score), "reference_spp": ref_spp, "spp": spp, "resolution": f"{width}x{height}",
            "rmse": "[linear, tonemapped x/(1+x)]", "mean": mean, "per_frame": {str(i): v for i, v in per.items()}}


def denoiser_flops(width, height):
    """Algorithmic work of one denoise: 2 * 9 * cin * cout multiply-adds per output
    pixel of every convolution, over the real (unpadded) channels."""
    layers, bufs, _ = dnmod.arch()
    hp, wp = dnmod.padded(height), dnmod.padded(width)
    return sum(2 * 9 * L.cin * L.cout * (hp >> bufs[L.in_buffer].level) * (wp >> bufs[L.in_buffer].level)
               for L in layers)


def denoise_vs_plain(r, builder, width=1920, height=1080, ref_spp=4096, plain_spp=256, weights=None):
    """Denoised low-spp frames against the plain path tracer at 256 spp (config2, 1080p,
    its own camera, which no training image uses): RMSE against a 4096-spp reference
    rendered with disjoint seeds, and wall time with the device synchronized.  The
    denoised path renders 4 (or 1) spp with the wavefront pipeline, the guide images,
    and the U-Net with the committed weights; the plain path is one megakernel launch.
    Further weight files, if given, are scored on the same images under "variants"."""
    weights = list(weights or [dnmod.DEFAULT_WEIGHTS])
    sc = scenes.config2(width=width, height=height)
    tlas = r.load_scene(sc, builder)
    dn = dnmod.Denoiser(r)
    w, h = width, height
    ref = r.target(w, h)
    step = 512
    for s0 in range(0, ref_spp, step):
        r.render([r.request(ref, tlas, sc.frame, s0, min(step, ref_spp - s0), 1_000_003)], capi.APH_MODE_WAVEFRONT)
        r.wait()
        log(f"denoise: reference {s0 + step}/{ref_spp} spp")
    g = r.readback(ref, w, h)[..., :3].astype(np.float64)

    def errors(tgt):
        img = r.readback(tgt, w, h)[..., :3].astype(np.float64)
        tm = lambda x: np.maximum(x, 0) / (1.0 + np.maximum(x, 0))  # noqa: E731
        return (round(float(np.sqrt(np.mean((img - g) ** 2))), 5),
                round(float(np.sqrt(np.mean((tm(img) - tm(g)) ** 2))), 5))

    vp = scenes.view_proj(sc.frame)
    ivp = np.array(sc.frame.inv_view, dtype=np.float32)
    aov, noisy, out = r.aov(w, h), r.target(w, h), r.target(w, h)

    def denoised(spp, flags=0):
        r.render([r.request(noisy, tlas, sc.frame, 0, spp, 0)], capi.APH_MODE_WAVEFRONT)
        r.render_aovs(aov, tlas, sc.frame, vp, vp, ivp)
        dn.denoise([(noisy, aov, out, flags)])
        r.wait()

    res = {"resolution": f"{w}x{h}", "reference_spp": ref_spp, "scene": "config2 (1M instanced triangles)",
           "rmse_domain": "linear radiance; tonemapped = x/(1+x)",
           "denoised_path": "wavefront render + guide images + U-Net (ensemble: mean over the 4 mirrored inputs)"}
    dn.load(weights[0])
    denoised(4)  # warm-up: activation buffers, graphs
    tp = r.target(w, h)
    t = time.perf_counter()
    r.render([r.request(tp, tlas, sc.frame, 0, plain_spp, 0)], capi.APH_MODE_MEGAKERNEL)
    r.wait()
    res["plain_spp"] = plain_spp
    res["plain_s"] = round(time.perf_counter() - t, 4)
    res["plain_rmse"], res["plain_rmse_tonemapped"] = errors(tp)

    def score(into):
        for spp, flags, key in ((4, 0, "denoised4"), (1, 0, "denoised1"),
                                (4, capi.APH_DN_ENSEMBLE, "denoised4_ensemble")):
            denoised(spp, flags)  # first use of a batch shape allocates
            t = time.perf_counter()
            denoised(spp, flags)
            into[f"{key}_s"] = round(time.perf_counter() - t, 4)
            into[f"{key}_rmse"], into[f"{key}_rmse_tonemapped"] = errors(out)
            into[f"noisy{spp}_rmse"], into[f"noisy{spp}_rmse_tonemapped"] = errors(noisy)
        for key in ("denoised4", "denoised4_ensemble"):
            tag = key.replace("denoised4", "4spp")
            into[f"time_ratio_{tag}_vs_plain"] = round(into[f"{key}_s"] / res["plain_s"], 4)
            into[f"rmse_ratio_{tag}_vs_plain"] = round(into[f"{key}_rmse"] / res["plain_rmse"], 3)
            into[f"rmse_ratio_{tag}_vs_plain_tonemapped"] = round(
                into[f"{key}_rmse_tonemapped"] / res["plain_rmse_tonemapped"], 3)
        # plain samples that would match the denoised error (error ~ 1/sqrt(spp))
        into["plain_spp_equal_rmse"] = int(plain_spp * (res["plain_rmse"] / into["denoised4_rmse"]) ** 2)
        into["plain_spp_equal_rmse_tonemapped"] = int(
            plain_spp * (res["plain_rmse_tonemapped"] / into["denoised4_rmse_tonemapped"]) ** 2)

    score(res)
    res["weights"] = os.path.relpath(weights[0], ROOT)
    if len(weights) > 1:
        res["variants"] = {}
        for wpath in weights[1:]:
            dn.load(wpath)
            v = {}
            score(v)
            res["variants"][os.path.basename(wpath)] = v
            log(f"denoise: {os.path.basename(wpath)} {v}")
        dn.load(weights[0])
    # the U-Net alone: HIP events around aph_denoise, MFMA work over real channels
    r.reset_timing()
    r.enable_timing(True)
    for _ in range(5):
        dn.denoise([(noisy, aov, out)])
    r.wait()
    r.enable_timing(False)
    ms, n = r.timing("denoise")
    flops = denoiser_flops(w, h)
    res["unet_ms"] = round(ms / max(n, 1), 3)
    res["unet_tflops"] = round(flops / (ms / max(n, 1) * 1e-3) / 1e12, 1)
    res["unet_mfma_frac_of_bf16_peak"] = round(res["unet_tflops"] / 2500.0, 4)
    dn.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--subdiv", type=int, default=4, help="4: 1.02M instanced tris; 5: 4.1M")
    ap.add_argument("--mode", choices=["megakernel", "wavefront"], default="wavefront")
    ap.add_argument("--builder", choices=["sah", "lbvh"], default="sah", help="BLAS builder")
    ap.add_argument("--sections", default="literal,c3,serving,rmse,denoise,config5",
                    help="extra measurements after the headline line's frame: literal (config2 at subdiv 5), "
                         "c3 (64 cameras), serving (Poisson replay), rmse (equal-time RMSE), denoise (denoised "
                         "4 spp vs plain 256 spp at 1080p), config5 (24.7M-triangle 120-frame sequence with "
                         "denoiser and temporal resolve); config5q (not by default: that sequence's frames "
                         "against 1024-spp references); '' for none")
    ap.add_argument("--tuning", default="",
                    help="wavefront launch configuration 'chunk_paths,trace_blocks_per_cu,shade_blocks_per_cu"
                         "[,trace_refill_lanes[,trace_leaf_wait_lanes]]'; "
                         "default: the autotuner's choice on the headline frame (32 spp, child process)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default launch configuration")
    ap.add_argument("--autotune-child", action="store_true",
                    help="internal: run the autotuner on the headline frame and print its choice")
    ap.add_argument("--denoiser-weights", default="",
                    help="comma-separated weight files for the denoise section (default: the committed weights); "
                         "the first is the reported one, the rest are scored under denoise.variants")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--cpu-baseline-child", action="store_true",
                    help="internal: the oracle's CPU baseline alone (never touches the GPU)")
    ap.add_argument("--pmc-child", action="store_true",
                    help="internal: calibration copies + one frame, run under rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard by sample range across GPUs (per-pixel sums reduced once per frame).",
              file=sys.stderr)
        sys.exit(2)

    sc = scenes.config2(subdiv=args.subdiv, spp=args.spp)
    mode = capi.APH_MODE_MEGAKERNEL if args.mode == "megakernel" else capi.APH_MODE_WAVEFRONT
    kname = "render_megakernel" if args.mode == "megakernel" else "wf_extend"
    builder = capi.APH_BUILDER_SAH if args.builder == "sah" else capi.APH_BUILDER_LBVH
    if args.cpu_baseline_child:
        print(json.dumps(cpu_baseline(sc)), flush=True)
        return
    if args.autotune_child:
        r = Renderer(0)
        print(json.dumps(autotune_section(r, sc, r.load_scene(sc, builder))), flush=True)
        r.close()
        return
    if args.pmc_child:
        copy_calibration()
        r = Renderer(0)
        apply_tuning(r, args.tuning)
        tgt = r.target(sc.width, sc.height)
        r.render([r.request(tgt, r.load_scene(sc, builder), sc.frame, 0, sc.spp)], mode)
        r.wait()
        r.close()
        print("pmc child done", flush=True)
        return
    tune_info = None
    if not args.tuning and not args.no_autotune and args.mode == "wavefront":
        tune_info = autotune_child(args)
        if tune_info:
            args.tuning = ",".join(str(v) for v in tune_info["chosen"])
            log(f"autotune: {tune_info}")
    traffic, traffic_detail = (None, "skipped (--no-traffic)") if args.no_traffic else measure_traffic(args, kname)
    log(f"scene: {sc.instanced_triangles} instanced / {sc.unique_triangles} unique triangles, "
        f"{len(sc.instances)} instances, {sc.width}x{sc.height} x {sc.spp} spp, mode {args.mode}")
    r = Renderer(0)
    apply_tuning(r, args.tuning)
    tlas = r.load_scene(sc, builder)
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
    if args.mode == "wavefront":
        # kExtend alone: closest-hit traversal of the extension queue
        frame_bytes = ((st["node_visits"] - st["shadow_node_visits"]) * NODE_B
                       + (st["tri_tests"] - st["shadow_tri_tests"]) * TRI_B
                       + (st["inst_visits"] - st["shadow_inst_visits"]) * INST_B
                       + st["extension_rays"] * (RAY_B + HIT_B))
    else:
        # the megakernel does the whole frame: every query, and accumulation
        frame_bytes = (st["node_visits"] * NODE_B + st["tri_tests"] * TRI_B + st["inst_visits"] * INST_B
                       + queries * RAY_B + sc.width * sc.height * ACC_B)
    log(f"stats: {st}")

    for i in range(args.warmup):
        r.render(req, mode)
        r.wait()
        log(f"warmup step {i + 1}/{args.warmup} done")
    # Timed steps: the production launch path (one captured graph per wavefront chunk).
    r.reset_stats()
    device_sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render(req, mode)
        r.wait()
        log(f"step {i + 1}/{args.steps} done")
    device_sync()
    wall = time.perf_counter() - t0
    tst = r.stats()
    rays = tst["extension_rays"] + tst["shadow_rays"]
    # Per-kernel HIP events need eager launches: one more frame, timed per stage.
    r.reset_timing()
    r.enable_timing(True)
    r.render(req, mode)
    r.wait()
    r.enable_timing(False)
    kms, klaunch = r.timing(kname)
    avg_launch_ms = kms / max(klaunch, 1)
    launches_per_frame = klaunch
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
            "blas_builder": args.builder,
            "launch_config": {"chunk_paths": r.tuning().chunk_paths,
                              "trace_blocks_per_cu": r.tuning().trace_blocks_per_cu,
                              "shade_blocks_per_cu": r.tuning().shade_blocks_per_cu,
                              "trace_refill_lanes": r.tuning().trace_refill_lanes,
                              "trace_leaf_wait_lanes": r.tuning().trace_leaf_wait_lanes,
                              "source": "autotuner" if tune_info else ("--tuning" if args.tuning else "default")},
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
            "traffic": traffic,
            "traffic_detail": traffic_detail,
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
    sections = [s for s in args.sections.split(",") if s]
    if tune_info:
        out["autotune"] = tune_info
    if "literal" in sections:
        out["config"]["config2_subdiv5"] = literal_config2(r, builder, sc.spp)
        log(f"config2 subdiv 5: {out['config']['config2_subdiv5']}")
    if "c3" in sections:
        out["c3_cameras"] = c3_cameras(r, builder)
        log(f"c3: {out['c3_cameras']}")
    if "serving" in sections:
        out["serving"] = serving(r, builder)
    if "rmse" in sections:
        out["equal_time_rmse"] = equal_time_rmse(r, builder)
        log(f"rmse: {out['equal_time_rmse']}")
    if "config5" in sections:
        out["config5_sequence"] = config5_sequence(r, builder)
        log(f"config5: {out['config5_sequence']}")
    if "config5q" in sections:
        out["config5_quality"] = config5_quality(r, builder)
        log(f"config5 quality: {out['config5_quality']['mean']}")
    if "denoise" in sections:
        out["denoise"] = denoise_vs_plain(r, builder,
                                          weights=[x for x in args.denoiser_weights.split(",") if x] or None)
        log(f"denoise: {out['denoise']}")
    r.close()
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores, child process)...")
        out["cpu_baseline"] = cpu_baseline_child(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
