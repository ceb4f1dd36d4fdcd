"""Python host wrapper over the C-ABI: one backend context per object.

This is plumbing for tests and benchmarks; every operation is a call into
libaph.so (hand-written HIP for gfx950).  There is no CPU fallback: creating a
Renderer without the built library or without a GPU raises.
"""
import ctypes as C

import numpy as np

from . import capi
from . import scenes


class AphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"aph error {code}: {msg}")
        self.code = code


class Renderer:
    """A backend context (aph_create / aph_destroy): createRenderer(Backend::HIP)."""

    def __init__(self, device=0):
        self.lib = capi.load()
        self.ctx = C.c_void_p()
        rc = self.lib.aph_create(device, C.byref(self.ctx))
        if rc != capi.APH_OK:
            raise AphError(rc, "aph_create failed (no usable HIP device?)")
        self._keep = []
        self._mesh_tris = {}

    def close(self):
        if self.ctx:
            self.lib.aph_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != capi.APH_OK:
            raise AphError(rc, self.lib.aph_last_error(self.ctx).decode())

    # ---- scene objects
    def mesh(self, pos, idx, nrm=None, uv=None):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        idx = None if idx is None else np.ascontiguousarray(idx, dtype=np.uint32)
        nrm = None if nrm is None else np.ascontiguousarray(nrm, dtype=np.float32)
        uv = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32)
        h = C.c_uint64()
        nt = (len(idx.reshape(-1)) // 3) if idx is not None else len(pos.reshape(-1, 3)) // 3
        self._check(self.lib.aph_mesh_upload(self.ctx, capi.ptr(pos), len(pos.reshape(-1, 3)), capi.ptr(idx), nt,
                                             capi.ptr(nrm), capi.ptr(uv), C.byref(h)))
        self._mesh_tris[h.value] = nt
        return h.value

    def build_blas(self, mesh, builder=capi.APH_BUILDER_LBVH):
        self._check(self.lib.aph_blas_build(self.ctx, mesh, builder))

    # ---- launch control
    def set_launch_mode(self, mode):
        self._check(self.lib.aph_set_launch_mode(self.ctx, mode))

    def tuning(self):
        t = capi.Tuning()
        self._check(self.lib.aph_get_tuning(self.ctx, C.byref(t)))
        return t

    def set_tuning(self, chunk_paths, trace_blocks_per_cu, shade_blocks_per_cu, trace_refill_lanes=32,
                   trace_leaf_wait_lanes=16):
        t = capi.Tuning(chunk_paths, trace_blocks_per_cu, shade_blocks_per_cu, trace_refill_lanes,
                        trace_leaf_wait_lanes, 0.0)
        self._check(self.lib.aph_set_tuning(self.ctx, C.byref(t)))

    def autotune(self, requests):
        arr = (capi.Request * len(requests))(*requests)
        t = capi.Tuning()
        self._check(self.lib.aph_wavefront_autotune(self.ctx, C.cast(arr, C.c_void_p), len(requests), C.byref(t)))
        return t

    # ---- guide images and temporal resolve
    def aov(self, w, h):
        a = C.c_uint64()
        self._check(self.lib.aph_aov_create(self.ctx, w, h, C.byref(a)))
        return a.value

    def render_aovs(self, aov, tlas, frame, view_proj, view_proj_prev, inv_view_prev, sky_view_z=-2000.0,
                    sample_index=0, seed_offset=0):
        r = capi.AovRequest()
        r.aov, r.tlas, r.frame = aov, tlas, frame
        for name, m in (("view_proj", view_proj), ("view_proj_prev", view_proj_prev),
                        ("inv_view_prev", inv_view_prev)):
            getattr(r, name)[:] = [float(v) for v in np.asarray(m, dtype=np.float32).reshape(16)]
        r.sky_view_z, r.sample_index, r.seed_offset = sky_view_z, sample_index, seed_offset
        self._check(self.lib.aph_render_aovs(self.ctx, C.byref(r)))

    def aov_readback(self, aov, w, h):
        d = np.zeros((h, w), dtype=np.float32)
        n, b, m = (np.zeros((h, w, 4), dtype=np.float32) for _ in range(3))
        self._check(self.lib.aph_aov_readback(self.ctx, aov, d.ctypes.data, n.ctypes.data, b.ctypes.data,
                                              m.ctypes.data))
        return d, n, b, m

    def temporal_resolve(self, current, aov, history_in, history_out, is_reset=False, alpha_min=1.0 / 16.0,
                         alpha_max=0.6, speed_reference=6.0):
        r = capi.TaaRequest(current, aov, history_in, history_out, int(is_reset), alpha_min, alpha_max,
                            speed_reference)
        self._check(self.lib.aph_temporal_resolve(self.ctx, C.byref(r)))

    def target_upload(self, target, rgba):
        """Writes an RGBA32F image into a target (test and tool helper, through the device pointer)."""
        p = C.c_void_p()
        self._check(self.lib.aph_target_device_ptr(self.ctx, target, C.byref(p)))
        hip = C.CDLL("libamdhip64.so")
        a = np.ascontiguousarray(rgba, dtype=np.float32)
        assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device

    def blas_export(self, mesh):
        """(nodes float32[N, 16], tris float32[T, 12], depth) of a built BLAS."""
        cnt, depth = C.c_uint32(), C.c_uint32()
        self._check(self.lib.aph_blas_export(self.ctx, mesh, None, 0, None, 0, C.byref(cnt), C.byref(depth)))
        nodes = np.zeros((max(cnt.value, 1), 16), dtype=np.float32)
        nt = self._mesh_tris[mesh]
        tris = np.zeros((max(nt, 1), 12), dtype=np.float32)
        self._check(self.lib.aph_blas_export(self.ctx, mesh, nodes.ctypes.data, nodes.nbytes, tris.ctypes.data,
                                             tris.nbytes, C.byref(cnt), C.byref(depth)))
        return nodes[:cnt.value], tris[:nt], depth.value

    def tlas(self, instances):
        inst = np.ascontiguousarray(instances, dtype=capi.INSTANCE_DTYPE)
        h = C.c_uint64()
        self._check(self.lib.aph_tlas_build(self.ctx, capi.ptr(inst), len(inst), C.byref(h)))
        return h.value

    def update_tlas(self, tlas, instances, mode=capi.APH_TLAS_REBUILD):
        """aph_tlas_update: the handle takes the new instances (BLASes kept)."""
        inst = np.ascontiguousarray(instances, dtype=capi.INSTANCE_DTYPE)
        self._check(self.lib.aph_tlas_update(self.ctx, tlas, capi.ptr(inst), len(inst), mode))

    def update_materials(self, first, mats):
        """aph_materials_update: patches materials first .. first + len(mats) - 1 in place."""
        arr = (capi.Material * len(mats))(*mats)
        self._check(self.lib.aph_materials_update(self.ctx, first, C.cast(arr, C.c_void_p), len(mats)))

    def set_layers(self, tlas, layers, n):
        """Per-instance material layer lists, bottom to top (aph_tlas_set_layers)."""
        off = np.zeros(n + 1, dtype=np.uint32)
        flat = []
        for i in range(n):
            ls = layers[i] if i < len(layers) and layers[i] else []
            flat.extend(ls)
            off[i + 1] = len(flat)
        mats = np.ascontiguousarray(flat if flat else [0], dtype=np.uint32)
        self._check(self.lib.aph_tlas_set_layers(self.ctx, tlas, off.ctypes.data, mats.ctypes.data))

    def set_layer_geometry(self, tlas, geometry, n):
        """Per instance, per layer (as set_layers): the mesh handle whose texture
        coordinates the layer reads, or 0 for the instance's (aph_tlas_set_layer_geometry)."""
        flat = []
        for i in range(n):
            flat.extend(geometry[i] if i < len(geometry) and geometry[i] else [])
        arr = np.ascontiguousarray(flat if flat else [0], dtype=np.uint64)
        self._check(self.lib.aph_tlas_set_layer_geometry(self.ctx, tlas, arr.ctypes.data))

    def _scene_layer_geometry(self, sc, tlas, handles):
        lg = getattr(sc, "layer_geometry", None)
        if lg is not None:
            per = [[0 if k is None else handles[k] for k in (g or [])] for g in lg]
            self.set_layer_geometry(tlas, per, len(sc.instances))

    def materials(self, mats):
        arr = (capi.Material * len(mats))(*mats)
        self._check(self.lib.aph_materials_upload(self.ctx, C.cast(arr, C.c_void_p), len(mats)))

    def image(self, pixels, linearize=True):
        """aph_image_create from an (h, w, 4) uint8 (RGBA8), (h, w, 4) float32 (RGBA32F)
        or (h, w, 3) float32 (RGB32F) array."""
        a = np.ascontiguousarray(pixels)
        if a.dtype == np.uint8:
            fmt = capi.APH_IMAGE_RGBA8
        else:
            a = np.ascontiguousarray(a, dtype=np.float32)
            fmt = capi.APH_IMAGE_RGBA32F if a.shape[2] == 4 else capi.APH_IMAGE_RGB32F
        d = capi.ImageDesc(a.shape[1], a.shape[0], fmt, int(linearize), a.ctypes.data)
        h = C.c_uint64()
        self._check(self.lib.aph_image_create(self.ctx, C.byref(d), C.byref(h)))
        return h.value

    def destroy_image(self, image):
        self._check(self.lib.aph_image_destroy(self.ctx, image))

    def bind_textures(self, nmat, slots):
        """slots: {(material, slot): (image handle, address_u, address_v)} for a table of nmat."""
        arr = (capi.TextureBinding * (nmat * capi.APH_TEX_SLOTS))()
        for (mi, k), (img, au, av) in slots.items():
            arr[mi * capi.APH_TEX_SLOTS + k] = capi.TextureBinding(img, au, av)
        self._check(self.lib.aph_materials_bind_textures(self.ctx, C.cast(arr, C.c_void_p), nmat))

    def _scene_textures(self, sc, mat_base=0, nmat=None):
        """Images and texture bindings of a scene (Scene.images / Scene.textures)."""
        textures = getattr(sc, "textures", None)
        if not textures:
            return {}
        handles = [self.image(im["pixels"], im.get("linearize", True)) for im in sc.images]
        return {(mi + mat_base, k): (handles[ii], au, av) for (mi, k), (ii, au, av) in textures.items()}

    def environment(self, env):
        """Sets the context's environment (scenes without one of their own)."""
        e, keep = scenes.environment_struct(env)
        self._check(self.lib.aph_environment_upload(self.ctx, C.byref(e)))

    def create_environment(self, env):
        """A resident environment object; returns its handle."""
        e, keep = scenes.environment_struct(env)
        h = C.c_uint64()
        self._check(self.lib.aph_environment_create(self.ctx, C.byref(e), C.byref(h)))
        return h.value

    def destroy_environment(self, handle):
        self._check(self.lib.aph_environment_destroy(self.ctx, handle))

    def set_environment(self, tlas, handle):
        """Renders of `tlas` use environment `handle` (0: the context's)."""
        self._check(self.lib.aph_tlas_set_environment(self.ctx, tlas, handle))

    def set_ground_plane(self, tlas, props):
        """Renders of `tlas` see a ground plane with these properties (a dict over the
        library's defaults); None removes it."""
        g = None if props is None else C.byref(capi.ground_plane(props))
        self._check(self.lib.aph_tlas_set_ground_plane(self.ctx, tlas, g))

    def load_scenes(self, scs, builder=capi.APH_BUILDER_LBVH):
        """Uploads several scenes into this context, resident together: one material
        table (each scene's instance material indices offset into it) and one
        environment object per scene, set on its TLAS.  Returns one TLAS handle per scene."""
        mats, tl, slots = [], [], {}
        for sc in scs:
            handles = []
            for k, (pos, nrm, idx) in enumerate(sc.meshes):
                m = self.mesh(pos, idx, nrm, sc.uvs.get(k))
                self.build_blas(m, builder)
                handles.append(m)
            inst = sc.instances.copy()
            inst["mesh"] = np.array([handles[i] for i in sc.instances["mesh"]], dtype=np.uint32)
            inst["material"] = inst["material"] + len(mats)
            tl.append(self.tlas(inst))
            if getattr(sc, "ground", None):
                self.set_ground_plane(tl[-1], sc.ground)
            if getattr(sc, "layers", None) is not None:
                self.set_layers(tl[-1], [[m + len(mats) for m in (ls or [])] for ls in sc.layers], len(inst))
                self._scene_layer_geometry(sc, tl[-1], handles)
            slots.update(self._scene_textures(sc, len(mats)))
            mats.extend(sc.materials)
        self.materials(mats)
        if slots:
            self.bind_textures(len(mats), slots)
        for sc, t in zip(scs, tl):
            self.set_environment(t, self.create_environment(sc.env or {}))
        return tl

    def load_scene(self, sc, builder=capi.APH_BUILDER_LBVH):
        """Uploads a scenes.Scene; returns the TLAS handle."""
        handles = []
        for k, (pos, nrm, idx) in enumerate(sc.meshes):
            m = self.mesh(pos, idx, nrm, sc.uvs.get(k))
            self.build_blas(m, builder)
            handles.append(m)
        inst = sc.instances.copy()
        inst["mesh"] = np.array([handles[i] for i in sc.instances["mesh"]], dtype=np.uint32)
        t = self.tlas(inst)
        self.mesh_handles = handles  # the scene's mesh index -> mesh handle, for updates of its instances
        if getattr(sc, "ground", None):
            self.set_ground_plane(t, sc.ground)
        if getattr(sc, "layers", None) is not None:
            self.set_layers(t, sc.layers, len(inst))
            self._scene_layer_geometry(sc, t, handles)
        self.materials(sc.materials)
        slots = self._scene_textures(sc)
        if slots:
            self.bind_textures(len(sc.materials), slots)
        self.environment(sc.env or {})
        return t

    # ---- ray queries
    def trace_closest(self, tlas, rays):
        rays = np.ascontiguousarray(rays, dtype=capi.RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=capi.HIT_DTYPE)
        self._check(self.lib.aph_trace_closest(self.ctx, tlas, capi.ptr(rays), len(rays), capi.ptr(hits)))
        return hits

    def trace_any(self, tlas, rays):
        rays = np.ascontiguousarray(rays, dtype=capi.RAY_DTYPE)
        occ = np.zeros(len(rays), dtype=np.uint8)
        self._check(self.lib.aph_trace_any(self.ctx, tlas, capi.ptr(rays), len(rays), capi.ptr(occ)))
        return occ

    # ---- rendering
    def target(self, w, h):
        t = C.c_uint64()
        self._check(self.lib.aph_target_create(self.ctx, w, h, C.byref(t)))
        return t.value

    def destroy_target(self, target):
        self._check(self.lib.aph_target_destroy(self.ctx, target))

    def render(self, requests, mode=capi.APH_MODE_MEGAKERNEL):
        arr = (capi.Request * len(requests))(*requests)
        self._check(self.lib.aph_render(self.ctx, C.cast(arr, C.c_void_p), len(requests), mode))

    def request(self, target, tlas, frame, sample_start, sample_count, seed_offset=0):
        r = capi.Request()
        r.target = target
        r.tlas = tlas
        r.frame = frame
        r.sample_start = sample_start
        r.sample_count = sample_count
        r.seed_offset = seed_offset
        return r

    def wait(self):
        self._check(self.lib.aph_wait(self.ctx))

    def readback(self, target, w, h):
        out = np.zeros((h, w, 4), dtype=np.float32)
        self._check(self.lib.aph_readback(self.ctx, target, capi.ptr(out), out.nbytes))
        return out

    def post(self, target, w, h, brightness=(1.0, 1.0, 1.0), tone_map=False, gamma=True):
        out = np.zeros((h, w, 4), dtype=np.float32)
        b = (C.c_float * 3)(*brightness)
        self._check(self.lib.aph_post(self.ctx, target, C.cast(b, C.c_void_p), int(tone_map), int(gamma),
                                      capi.ptr(out), out.nbytes))
        return out

    # ---- instrumentation
    def set_stats_level(self, level):
        self._check(self.lib.aph_set_stats_level(self.ctx, level))

    def stats(self):
        s = capi.Stats()
        self._check(self.lib.aph_stats_get(self.ctx, C.byref(s)))
        return {k: getattr(s, k) for k, _ in capi.Stats._fields_}

    def reset_stats(self):
        self._check(self.lib.aph_stats_reset(self.ctx))

    def enable_timing(self, on=True):
        self._check(self.lib.aph_timing_enable(self.ctx, int(on)))

    def timing(self, kernel):
        ms = C.c_double()
        n = C.c_uint64()
        self._check(self.lib.aph_timing_get(self.ctx, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_timing(self):
        self._check(self.lib.aph_timing_reset(self.ctx))


class Server:
    """Serving loop over a Renderer's resident scenes (aph_server_*): continuous
    batching of render jobs with HBM-budget admission."""

    def __init__(self, renderer, hbm_budget=0, max_active=0, step_samples=0, mode=capi.APH_MODE_WAVEFRONT,
                 keep_results=True):
        self.r = renderer
        self.lib = renderer.lib
        cfg = capi.ServerConfig(hbm_budget, max_active, step_samples, mode, 1 if keep_results else 0)
        self.h = C.c_void_p()
        renderer._check(self.lib.aph_server_create(renderer.ctx, C.byref(cfg), C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.aph_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def job(tlas, frame, width, height, spp, seed_offset=0):
        return capi.Job(tlas, frame, width, height, spp, seed_offset)

    def clock(self):
        return self.lib.aph_server_clock(self.h)

    def submit(self, job, t_arrival=-1.0):
        t = C.c_uint64()
        self.r._check(self.lib.aph_server_submit(self.h, C.byref(job), t_arrival, C.byref(t)))
        return t.value

    def step(self):
        a, p = C.c_uint32(), C.c_uint32()
        self.r._check(self.lib.aph_server_step(self.h, C.byref(a), C.byref(p)))
        return a.value, p.value

    def status(self, ticket):
        st = capi.JobStatus()
        self.r._check(self.lib.aph_server_status(self.h, ticket, C.byref(st)))
        return st

    def result(self, ticket, width, height):
        out = np.zeros((height, width, 4), dtype=np.float32)
        self.r._check(self.lib.aph_server_result(self.h, ticket, out.ctypes.data, out.nbytes))
        return out

    def usage(self):
        """(resident scene bytes, admitted job bytes, running, waiting)."""
        r, j, a, p = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        self.r._check(self.lib.aph_server_usage(self.h, C.byref(r), C.byref(j), C.byref(a), C.byref(p)))
        return r.value, j.value, a.value, p.value

    def working_set(self):
        """(wavefront working-set bytes, paths in flight) of the last batched step."""
        b, n = C.c_uint64(), C.c_uint32()
        self.r._check(self.lib.aph_server_working_set(self.h, C.byref(b), C.byref(n)))
        return b.value, n.value

    def release(self, ticket):
        self.r._check(self.lib.aph_server_release(self.h, ticket))

    def run(self):
        """Steps until no job is running or waiting."""
        while True:
            a, p = self.step()
            if a == 0 and p == 0:
                return

    def replay(self, jobs, arrivals):
        """Real-time replay of an arrival trace; returns the tickets."""
        n = len(jobs)
        arr = (capi.Job * n)(*jobs)
        t = np.ascontiguousarray(arrivals, dtype=np.float64)
        tickets = np.zeros(n, dtype=np.uint64)
        self.r._check(self.lib.aph_server_replay(self.h, C.cast(arr, C.c_void_p), t.ctypes.data, n,
                                                 tickets.ctypes.data))
        return [int(x) for x in tickets]
