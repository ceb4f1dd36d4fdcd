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
NODE_B, TRI_B, INST_B, RAY_B, HIT_B, ACC_B, VIS_B = 128, 48, 64, 32, 20, 32, 1
# kShade per queue entry: ray (32) + hit (20) + path state read (pixel, request, flags, contribution,
# rng: 36) and written (flags, contribution, rng, pending emission, continue and shadow-mask bytes: 46);
# per hit: shading data (indices + 3 x position, normal, uv: 108), instance shading record (96),
# material (192), environment (2 alias entries + 2 texels: 64), written next ray (32), pending
# contribution (16) and two shadow rays with their colours (2 x 48); per miss: one environment texel.
SHADE_ENTRY_B = 32 + 20 + 36 + 46
SHADE_HIT_B = 108 + 96 + 192 + 64 + 32 + 16 + 2 * 48
SHADE_MISS_B = 16


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_sync():
    hip = C.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


# Kernels priced by the measurement, by their timing name: the symbol substring that
# picks their dispatches out of a rocprofv3 database.
KERNEL_SYMBOL = {"render_megakernel": "kRenderMega<false>", "wf_extend": "kExtend<false>",
                 "wf_shadow": "kShadow<false>", "wf_shade": "kShade("}
COPY_BYTES, COPIES = 1 << 30, 3
# Issue-side counters of one pass (at most 8 SQ and 2 GRBM counters per pass).
ISSUE_COUNTERS = ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES",
                  "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"]
# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz; a wave64 VALU instruction occupies a 16-lane SIMD for 4
# cycles, so the chip issues at most 256 * 4 * 2.4e9 / 4 wave-level VALU instructions per second.
PEAK_VALU_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


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
    """(kernel name, counter value) per dispatch from a rocprofv3 database."""
    rows = []
    for db in glob.glob(os.path.join(db_dir, "**", "*.db"), recursive=True):
        con = sqlite3.connect(db)
        rows += con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                            (counter,)).fetchall()
        con.close()
    return rows


def pmc_pass(args, counters):
    """One rocprofv3 --pmc pass over the pmc child (calibration copies + one frame):
    {counter: [(kernel name, value) per dispatch]}, or None."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    d = tempfile.mkdtemp(prefix="aph_pmc_", dir="/tmp")
    cmd = ["timeout", "-s", "KILL", "240", prof, "--pmc", *counters, "-d", d, "-o", "pmc", "--",
           sys.executable, "-u", os.path.abspath(__file__), "--pmc-child", "--spp", str(args.spp),
           "--subdiv", str(args.subdiv), "--mode", args.mode, "--builder", args.builder, "--tuning", args.tuning]
    log(f"pmc pass {' '.join(counters)}: {' '.join(cmd)}")
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = {c: pmc_dispatches(d, c) for c in counters}
    shutil.rmtree(d, ignore_errors=True)
    if p.returncode != 0 or not all(out.values()):
        log(p.stdout[-2000:])
        return None
    return out


def measure_counters(args, kernels):
    """Per kernel (timing name): HBM bytes per launch from FETCH_SIZE and WRITE_SIZE (two
    passes, each scaled by actual bytes / counter reading of 1 GiB calibration copies made in
    the same pass) and the VALU issue counters per launch (a third pass).  Returns
    ({kernel: {...}}, detail) or (None, reason)."""
    res = {k: {} for k in kernels}
    factors = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = pmc_pass(args, [counter])
        if rows is None:
            return None, f"{counter} pass failed"
        rows = rows[counter]
        copies = sorted((v for n, v in rows if "copyBuffer" in n), reverse=True)[:COPIES]
        if len(copies) < COPIES:
            return None, f"{counter} pass: {len(copies)} calibration copies"
        factors[counter] = COPY_BYTES * COPIES / (sum(copies) * 1024.0)
        for k in kernels:
            vals = [v for n, v in rows if KERNEL_SYMBOL[k] in n]
            if vals:
                res[k][counter] = sum(vals) / len(vals)
                res[k]["dispatches"] = len(vals)
        log(f"{counter}: copy factor {factors[counter]:.4f}; " +
            ", ".join(f"{k} {res[k].get(counter, 0):.1f} KiB" for k in kernels))
    rows = pmc_pass(args, ISSUE_COUNTERS)
    for k in kernels:
        r = res[k]
        if "FETCH_SIZE" in r and "WRITE_SIZE" in r:
            r["traffic_bytes"] = int((r["FETCH_SIZE"] * factors["FETCH_SIZE"] + r["WRITE_SIZE"] * factors["WRITE_SIZE"])
                                     * 1024)
        if rows is not None:
            for c in ISSUE_COUNTERS:
                vals = [v for n, v in rows[c] if KERNEL_SYMBOL[k] in n]
                if vals:
                    r[c] = sum(vals) / len(vals)
    detail = {"fetch_factor": round(factors["FETCH_SIZE"], 4), "write_factor": round(factors["WRITE_SIZE"], 4),
              "issue_pass": rows is not None}
    return res, detail


def apply_tuning(r, tuning):
    if tuning:
        r.set_tuning(*(int(v) for v in tuning.split(",")))


def host_cpu():
    """(CPU model, logical CPUs of the host, CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1, len(os.sched_getaffinity(0))


# A one-GPU job on the GPU host is allotted 16 host cores; worker pools are sized to that share.
HOST_SHARE = 16


def cpu_baseline(sc, budget_s=12.0):
    """Oracle path tracer on the host cores over a bounded sample of the same frame:
    random pixels, successive sample indices, until about budget_s of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    model, nproc, allowed = host_cpu()
    threads = max(1, min(HOST_SHARE, allowed))
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
        "cpu_model": model,
        "host_logical_cpus": nproc,
        "cpus_allowed": allowed,
        "cores_note": f"all {threads} cores of the one-GPU job's host share ({HOST_SHARE}); the host has {nproc}",
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


def serving(r, builder, jobs=32, rates=(10.0, 40.0), max_spp=64, seed=5, budget_gib=48.0, width=1920, height=1080,
            plain_jobs=None):
    """Config 4 (by default at reduced length): Poisson arrivals of 1080p jobs with spp
    U{1..max_spp} over the four resident serving scenes (about 10k, 100k, 1M and 5M
    triangles), HBM-budget admission (scenes, job buffers and the wavefront working set).
    Batched (continuous batching in shared wavefronts) vs plain (one job at a time,
    megakernel).  Both default rates saturate the plain form; the batched jobs/s at the
    higher rate is its saturation throughput.  plain_jobs: replay only the first that many
    jobs of each trace in the plain form (its rate is then measured on that prefix)."""
    from aurora_amd.session import Server
    scs = scenes.serving_scenes()
    tl = r.load_scenes(scs, builder)
    rng = np.random.default_rng(seed)
    out = {"scenes_triangles": [s.instanced_triangles for s in scs], "jobs": jobs, "resolution": f"{width}x{height}",
           "spp": f"U{{1..{max_spp}}}", "hbm_budget_gib": budget_gib, "rates": {}}
    if jobs != 1024:
        out["differs_from_config4"] = (f"{jobs} jobs per rate (config 4: 1,024) at rates {list(rates)} /s (config 4: "
                                       "10-500 /s); every rate above about 2 jobs/s saturates one GPU, so longer "
                                       "traces only lengthen the queue")
    if plain_jobs is not None:
        out["plain_jobs"] = plain_jobs
    js = []
    for _ in range(jobs):
        k = int(rng.integers(0, 4))
        js.append(Server.job(tl[k], scenes.orbit_frame(rng, width, height, radius=30.0 if k else 14.0), width,
                             height, int(rng.integers(1, max_spp + 1))))
    for rate in rates:
        arr = np.cumsum(rng.exponential(1.0 / rate, jobs))
        res = {}
        for name, mode in (("batched", capi.APH_MODE_WAVEFRONT), ("plain", capi.APH_MODE_MEGAKERNEL)):
            n = jobs if name == "batched" or plain_jobs is None else min(plain_jobs, jobs)
            srv = Server(r, hbm_budget=int(budget_gib * (1 << 30)), mode=mode, keep_results=False)
            r.reset_stats()
            t = time.perf_counter()
            tickets = srv.replay(js[:n], arr[:n])
            dt = time.perf_counter() - t
            lat = np.array([srv.status(tk).t_done - srv.status(tk).t_arrival for tk in tickets])
            resident = srv.usage()[0]
            wf_bytes, wf_paths = srv.working_set()
            srv.close()
            ns = sum(j.width * j.height * j.spp for j in js[:n])
            res[name] = {"jobs": n, "p50_s": round(float(np.percentile(lat, 50)), 3),
                         "p99_s": round(float(np.percentile(lat, 99)), 3),
                         "jobs_per_s": round(n / dt, 2), "mrays_s": round(rays_of(r.stats()) / dt / 1e6, 2),
                         "samples_per_s": round(ns / dt, 1), "wall_s": round(dt, 2),
                         "last_arrival_s": round(float(arr[n - 1]), 2)}
            if mode == capi.APH_MODE_WAVEFRONT:
                res[name]["working_set_gib"] = round(wf_bytes / (1 << 30), 2)
                res[name]["paths_in_flight"] = wf_paths
            log(f"serving rate {rate}/s {name}: {res[name]}")
        res["batched_vs_plain_throughput"] = round(res["batched"]["samples_per_s"] / res["plain"]["samples_per_s"], 2)
        out["rates"][str(rate)] = res
    out["resident_scene_bytes"] = int(resident)
    return out


def config2_reference(r, builder, width=1920, height=1080, ref_spp=4096):
    """The 1080p config2 frame (the headline scene and camera) and its 4096-spp wavefront render
    with seeds disjoint from every measured render (seed offset 1,000,003); rendered once per
    renderer and shared by the sections that score against it."""
    key = (builder, width, height, ref_spp)
    cache = r.__dict__.setdefault("_reference_cache", {})
    sc = scenes.config2(width=width, height=height)
    tlas = r.load_scene(sc, builder)  # loaded again each time: other sections replace the material table
    if key not in cache:
        ref = r.target(width, height)
        step = 512
        for s0 in range(0, ref_spp, step):
            r.render([r.request(ref, tlas, sc.frame, s0, min(step, ref_spp - s0), 1_000_003)],
                     capi.APH_MODE_WAVEFRONT)
            r.wait()
            log(f"reference: {s0 + step}/{ref_spp} spp")
        cache[key] = r.readback(ref, width, height)[..., :3].astype(np.float64)
        r.destroy_target(ref)
    return sc, tlas, cache[key]


def equal_time_rmse(r, builder, plain_spp=64):
    """Equal-time RMSE on the 1080p headline frame against its 4096-spp reference: the plain
    megakernel's 64-spp image against the wavefront image with as many samples as the
    wavefront renders in the same wall time (same sampler, so the images differ only by
    sample count).  Linear radiance and tonemapped x / (1 + x)."""
    sc, tlas, ref_img = config2_reference(r, builder)
    width, height = sc.width, sc.height
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
    t = time.perf_counter()
    if spp_eq > plain_spp:
        r.render([r.request(tw, tlas, sc.frame, plain_spp, spp_eq - plain_spp)], capi.APH_MODE_WAVEFRONT)
        r.wait()
    t_wf += time.perf_counter() - t
    tm = lambda x: np.maximum(x, 0) / (1.0 + np.maximum(x, 0))  # noqa: E731

    def rmse(tgt):
        img = r.readback(tgt, width, height)[..., :3].astype(np.float64)
        return (float(np.sqrt(np.mean((img - ref_img) ** 2))), float(np.sqrt(np.mean((tm(img) - tm(ref_img)) ** 2))))
    a, b = rmse(tp), rmse(tw)
    return {"resolution": f"{width}x{height}", "reference_spp": 4096, "plain_s": round(t_plain, 3),
            "wavefront_s": round(t_wf, 3), "plain_spp": plain_spp, "wavefront_spp": spp_eq,
            "plain_rmse": round(a[0], 5), "wavefront_rmse": round(b[0], 5), "rmse_ratio": round(a[0] / b[0], 3),
            "plain_rmse_tonemapped": round(a[1], 5), "wavefront_rmse_tonemapped": round(b[1], 5),
            "rmse_ratio_tonemapped": round(a[1] / b[1], 3)}


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
    sc, tlas, g = config2_reference(r, builder, width, height, ref_spp)
    dn = dnmod.Denoiser(r)
    w, h = width, height

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
                         "denoiser and temporal resolve); not by default: config5q (that sequence's frames "
                         "against 1024-spp references), c4full (config 4's 1,024 jobs at 10 and 500 /s, the "
                         "plain form on 64 of them); '' for none")
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
    priced = [kname] if args.mode == "megakernel" else ["wf_extend", "wf_shadow", "wf_shade"]
    counters, counter_detail = (None, "skipped (--no-traffic)") if args.no_traffic else measure_counters(args, priced)
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
        ext, hits = st["extension_rays"], st["hits"]
        frame_bytes = {
            # kExtend: closest-hit traversal of the extension queue
            "wf_extend": ((st["node_visits"] - st["shadow_node_visits"]) * NODE_B
                          + (st["tri_tests"] - st["shadow_tri_tests"]) * TRI_B
                          + (st["inst_visits"] - st["shadow_inst_visits"]) * INST_B + ext * (RAY_B + HIT_B)),
            # kShadow: any-hit traversal of the shadow queue, one visibility byte per query
            "wf_shadow": (st["shadow_node_visits"] * NODE_B + st["shadow_tri_tests"] * TRI_B
                          + st["shadow_inst_visits"] * INST_B + st["shadow_rays"] * (RAY_B + VIS_B)),
            # kShade: every queue entry's ray, hit and path state; per hit the shading data,
            # instance record, material, environment lookups and the rays and shadow records written;
            # per miss one environment texel
            "wf_shade": ext * SHADE_ENTRY_B + hits * SHADE_HIT_B + (ext - hits) * SHADE_MISS_B,
        }
    else:
        # the megakernel does the whole frame: every query, and accumulation
        frame_bytes = {kname: st["node_visits"] * NODE_B + st["tri_tests"] * TRI_B + st["inst_visits"] * INST_B
                       + queries * RAY_B + sc.width * sc.height * ACC_B}
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
    kernels = {}
    for k in priced:
        kms, klaunch = r.timing(k)
        avg_ms = kms / max(klaunch, 1)
        per_launch = frame_bytes[k] / max(klaunch, 1)
        e = {"launches_per_frame": klaunch, "avg_launch_ms": round(avg_ms, 3), "ms_per_frame": round(kms, 2),
             "algorithmic_bytes_per_launch": int(per_launch),
             "algorithmic_gbs": round(per_launch / (avg_ms * 1e-3) / 1e9, 1)}
        c = (counters or {}).get(k, {})
        if "traffic_bytes" in c:
            e["hbm_bytes_per_launch"] = c["traffic_bytes"]
            e["hbm_gbs"] = round(c["traffic_bytes"] / (avg_ms * 1e-3) / 1e9, 1)
            e["hbm_frac"] = round(e["hbm_gbs"] / HBM_PEAK_GBS, 4)
            e["algorithmic_over_hbm_bytes"] = round(per_launch / max(c["traffic_bytes"], 1), 2)
        if "SQ_INSTS_VALU" in c and "SQ_ACTIVE_INST_VALU" in c:
            e["valu_wave_insts_per_launch"] = int(c["SQ_INSTS_VALU"])
            e["valu_issue_frac"] = round(c["SQ_INSTS_VALU"] / (avg_ms * 1e-3) / PEAK_VALU_WAVE_INSTS, 4)
            if c.get("SQ_THREAD_CYCLES_VALU") and c["SQ_ACTIVE_INST_VALU"]:
                e["valu_lanes_active"] = round(c["SQ_THREAD_CYCLES_VALU"] / c["SQ_ACTIVE_INST_VALU"], 1)
            e["counters_per_launch"] = {n: int(c[n]) for n in ISSUE_COUNTERS if n in c}
        if "hbm_frac" in e and "valu_issue_frac" in e:
            # a resource more than 60 % busy is the limiter; below that on both, the
            # waves mostly wait on memory latency
            if e["valu_issue_frac"] >= 0.6 and e["valu_issue_frac"] >= e["hbm_frac"]:
                e["limiter"] = "VALU issue (divergent waves)"
            elif e["hbm_frac"] >= 0.6:
                e["limiter"] = "HBM bandwidth"
            else:
                e["limiter"] = "memory latency (VALU and HBM both under 60 % busy)"
        kernels[k] = e
    dom = kernels[kname]
    avg_launch_ms = dom["avg_launch_ms"]
    bytes_per_launch = dom["algorithmic_bytes_per_launch"]
    achieved = dom["algorithmic_gbs"]
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
            "traffic": dom.get("hbm_bytes_per_launch"),
            "traffic_detail": counter_detail,
            "bytes_per_launch": int(bytes_per_launch),
            "avg_launch_ms": round(avg_launch_ms, 3),
            "achieved_meaning": "algorithmic bytes (BVH nodes, triangles, instance records, rays, hits) per launch "
                                "over the launch time: effective bandwidth on L2/MALL-resident data",
            "frac_of_hbm_by_counters": dom.get("hbm_frac"),
            "limiter": dom.get("limiter"),
            "kernels": kernels,
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
    if "c4full" in sections:
        # config 4 at its length: 1,024 jobs per rate; the plain form on each trace's first 64 jobs
        out["serving_config4_full"] = serving(r, builder, jobs=1024, rates=(10.0, 500.0), plain_jobs=64)
        log(f"config4 full: {out['serving_config4_full']}")
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
