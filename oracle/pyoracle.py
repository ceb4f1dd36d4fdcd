"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("inst", "<u4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])
ALIAS_DTYPE = np.dtype([("alias", "<u4"), ("prob", "<f4"), ("pdf", "<f4"), ("pad", "<f4")])


class Mesh(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("nrm", C.c_void_p), ("uv", C.c_void_p), ("idx", C.c_void_p),
                ("nv", C.c_uint32), ("nt", C.c_uint32)]


class Texture(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("w", C.c_uint32), ("h", C.c_uint32), ("addr_u", C.c_int32),
                ("addr_v", C.c_int32)]


def texture_texels(pixels, linearize=True):
    """Linear RGBA float texels of an image as the oracle reads it: RGBA8 as unorm, with
    linearize the sRGB transfer function undone per colour channel (ora_srgb_to_linear)."""
    a = np.asarray(pixels)
    if a.dtype == np.uint8:
        lib = load()
        unorm = np.arange(256, dtype=np.float32) / np.float32(255.0)
        lin = np.array([lib.ora_srgb_to_linear(float(v)) for v in unorm], dtype=np.float32)
        out = unorm[a].astype(np.float32)
        if linearize:
            out[..., :3] = lin[a[..., :3]]
        return np.ascontiguousarray(out)
    a = a.astype(np.float32)
    if a.shape[2] == 3:
        a = np.concatenate([a, np.ones(a.shape[:2] + (1,), np.float32)], axis=2)
    return np.ascontiguousarray(a)


class Env(C.Structure):
    _fields_ = [("light_top", C.c_float * 3), ("light_bottom", C.c_float * 3), ("bg_top", C.c_float * 3),
                ("bg_bottom", C.c_float * 3), ("light_xform", C.c_float * 16), ("light_xform_inv", C.c_float * 16),
                ("bg_xform", C.c_float * 16), ("bg_use_screen", C.c_int32), ("filter", C.c_int32),
                ("light_rgb", C.c_void_p), ("light_w", C.c_uint32), ("light_h", C.c_uint32),
                ("alias", C.c_void_p), ("light_integral", C.c_float), ("bg_rgb", C.c_void_p),
                ("bg_w", C.c_uint32), ("bg_h", C.c_uint32)]


class GroundPlane(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("position", C.c_float * 3), ("normal", C.c_float * 3),
                ("shadow_opacity", C.c_float), ("shadow_color", C.c_float * 3), ("reflection_opacity", C.c_float),
                ("reflection_color", C.c_float * 3), ("reflection_roughness", C.c_float)]


# The ground plane's properties and their defaults (PTGroundPlane.cpp propertySet)
GROUND_DEFAULTS = dict(enabled=1, position=(0.0, 0.0, 0.0), normal=(0.0, 1.0, 0.0), shadow_opacity=1.0,
                       shadow_color=(0.0, 0.0, 0.0), reflection_opacity=0.5, reflection_color=(0.5, 0.5, 0.5),
                       reflection_roughness=0.1)


def ground_plane_struct(d):
    g = GroundPlane()
    for k, dv in GROUND_DEFAULTS.items():
        v = d.get(k, dv)
        if isinstance(dv, tuple):
            getattr(g, k)[:] = [float(x) for x in v]
        else:
            setattr(g, k, type(getattr(g, k))(v))
    return g


class Frame3(C.Structure):
    _fields_ = [("N", C.c_float * 3), ("X", C.c_float * 3), ("Y", C.c_float * 3)]


class Counters(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
    lib = C.CDLL(LIB_PATH)
    P, u32, i = C.c_void_p, C.c_uint32, C.c_int
    sig = {
        "ora_pcg2d_init": (None, [u32, u32, u32, u32, P]),
        "ora_pcg2d": (None, [P]),
        "ora_random2d": (None, [P, P]),
        "ora_random_stream": (None, [u32, u32, u32, P]),
        "ora_inverse3x4": (i, [P, P]),
        "ora_scene_create": (P, [P, u32, P, u32]),
        "ora_scene_destroy": (None, [P]),
        "ora_scene_set_layers": (None, [P, P, P]),
        "ora_trace_closest": (None, [P, P, u32, P, i]),
        "ora_trace_closest_brute": (None, [P, P, u32, P]),
        "ora_trace_any": (None, [P, P, u32, P, i]),
        "ora_trace_any_brute": (None, [P, P, u32, P]),
        "ora_sah_sweep_cost": (C.c_double, [P, P, u32, u32, C.c_float, C.c_float]),
        "ora_intersect_triangle": (i, [P, P, C.c_float, C.c_float, P, P, P, P, P, P]),
        "ora_material_defaults": (None, [P]),
        "ora_alias_build": (None, [P, u32, u32, P, P]),
        "ora_env_evaluate": (None, [P, P, i, u32, u32, u32, u32, P]),
        "ora_env_sample": (None, [P, P, P, P, P]),
        "ora_dir_to_latlong": (None, [P, P]),
        "ora_latlong_to_dir": (None, [P, P]),
        "ora_bsdf_evaluate": (None, [P, P, P, P, P]),
        "ora_bsdf_evaluate_pdf": (None, [P, P, P, P, P, P, P, P]),
        "ora_bsdf_sample": (None, [P, P, P, P, P, P, P, P]),
        "ora_sheen_albedo": (C.c_float, [C.c_float, C.c_float]),
        "ora_render": (None, [P, P, u32, P, P, u32, u32, u32, u32, P, u32, i, P, P]),
        "ora_pixel_sample": (None, [P, P, u32, P, P, u32, u32, u32, u32, u32, P, P]),
        "ora_post": (None, [P, u32, P, i, i, P]),
        "ora_camera_ray": (None, [P, u32, u32, u32, u32, u32, P, P]),
        "ora_triangle_hit_owned": (i, [P, P, P, P, P]),
        "ora_temporal_resolve": (None, [P, P, P, P, P, u32, u32, i, C.c_float, C.c_float, C.c_float]),
        "ora_view_z": (C.c_float, [P, P]),
        "ora_screen_motion": (None, [P, P, P, u32, u32, P]),
        "ora_halton_jitter": (None, [u32, u32, P]),
        "ora_aov_pixel": (None, [P, P, u32, P, u32, u32, u32, u32, u32, P, P, P, C.c_float, P]),
        "ora_srgb_to_linear": (C.c_float, [C.c_float]),
        "ora_uv_transform": (None, [P, P, P]),
        "ora_texture_sample": (None, [P, P, P]),
        "ora_scene_set_textures": (None, [P, P, u32, P, u32]),
        "ora_normal_from_map": (None, [P, P, P, P]),
        "ora_hit_uv": (None, [P, u32, u32, C.c_float, C.c_float, P]),
        "ora_scene_set_ground_plane": (None, [P, P]),
        "ora_scene_set_layer_uvs": (None, [P, P]),
        "ora_ground_plane_hit": (i, [P, P, P, P, P]),
        "ora_ggx_smith_sample": (None, [P, P, C.c_float, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f32(*vals):
    return np.ascontiguousarray(np.array(vals, dtype=np.float32).reshape(-1))


class OracleScene:
    """Meshes + instances of a scenes.Scene with the oracle's own BVH."""

    def __init__(self, sc):
        lib = load()
        self._keep = []
        meshes = (Mesh * len(sc.meshes))()
        uvs = getattr(sc, "uvs", {}) or {}
        for k, (pos, nrm, idx) in enumerate(sc.meshes):
            pos = np.ascontiguousarray(pos, dtype=np.float32)
            idx = np.ascontiguousarray(idx, dtype=np.uint32)
            nrm = None if nrm is None else np.ascontiguousarray(nrm, dtype=np.float32)
            uv = uvs.get(k)
            uv = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32)
            self._keep += [pos, idx, nrm, uv]
            meshes[k].pos = _p(pos)
            meshes[k].nrm = _p(nrm)
            meshes[k].uv = _p(uv)
            meshes[k].idx = _p(idx)
            meshes[k].nv = len(pos.reshape(-1, 3))
            meshes[k].nt = len(idx.reshape(-1)) // 3
        inst = np.ascontiguousarray(sc.instances)
        self._keep += [meshes, inst]
        self.ptr = lib.ora_scene_create(C.cast(meshes, C.c_void_p), len(sc.meshes), _p(inst), len(inst))
        self.scene = sc
        layers = getattr(sc, "layers", None)
        if layers is not None:
            off, mats = layer_csr(layers, len(inst))
            lib.ora_scene_set_layers(self.ptr, _p(off), _p(mats))
            geom = getattr(sc, "layer_geometry", None)
            if geom is not None:
                # per layer: the texture coordinates of its mesh (given, or u = v = i / nv)
                ptrs = (C.c_void_p * max(1, int(off[-1])))()
                for i in range(len(inst)):
                    g = geom[i] if i < len(geom) and geom[i] else []
                    for j, k in enumerate(g):
                        if k is None:
                            continue
                        uv = uvs.get(k)
                        if uv is None:
                            nv = len(np.asarray(sc.meshes[k][0]).reshape(-1, 3))
                            uv = np.repeat((np.arange(nv, dtype=np.float32) / np.float32(nv))[:, None], 2, axis=1)
                        uv = np.ascontiguousarray(uv, dtype=np.float32)
                        self._keep.append(uv)
                        ptrs[int(off[i]) + j] = uv.ctypes.data
                lib.ora_scene_set_layer_uvs(self.ptr, C.cast(ptrs, C.c_void_p))
        textures = getattr(sc, "textures", None)
        if textures:
            # one oracle texture per (material, slot) binding, as the library's table
            nmat = len(sc.materials)
            slot_tex = np.full(nmat * 5, -1, dtype=np.int32)
            tex = (Texture * len(textures))()
            texels = []
            for n, ((mi, k), (ii, au, av)) in enumerate(sorted(textures.items())):
                im = sc.images[ii]
                t = texture_texels(im["pixels"], im.get("linearize", True))
                texels.append(t)
                tex[n] = Texture(_p(t), t.shape[1], t.shape[0], au, av)
                slot_tex[mi * 5 + k] = n
            lib.ora_scene_set_textures(self.ptr, C.cast(tex, C.c_void_p), len(textures), _p(slot_tex), nmat)
        ground = getattr(sc, "ground", None)
        if ground:
            g = ground_plane_struct(ground)
            lib.ora_scene_set_ground_plane(self.ptr, C.cast(C.pointer(g), C.c_void_p))
        self._env = None

    def __del__(self):
        try:
            load().ora_scene_destroy(self.ptr)
        except Exception:
            pass

    def trace_closest(self, rays, threads=8, brute=False):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        if brute:
            load().ora_trace_closest_brute(self.ptr, _p(rays), len(rays), _p(hits))
        else:
            load().ora_trace_closest(self.ptr, _p(rays), len(rays), _p(hits), threads)
        return hits

    def trace_any(self, rays, threads=8, brute=False):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.zeros(len(rays), dtype=np.uint8)
        if brute:
            load().ora_trace_any_brute(self.ptr, _p(rays), len(rays), _p(occ))
        else:
            load().ora_trace_any(self.ptr, _p(rays), len(rays), _p(occ), threads)
        return occ

    def env(self):
        if self._env is None:
            self._env = make_env(self.scene.env or {})
        return self._env

    def aovs(self, frame, width, height, sample_index, view_proj, view_proj_prev, inv_view_prev, sky_view_z):
        """Guide images of one camera ray per pixel: depth [H,W], normal_roughness,
        base_color_metalness, motion [H,W,4]."""
        lib = load()
        mats = (type(self.scene.materials[0]) * len(self.scene.materials))(*self.scene.materials)
        vp, vpp, ivp = (np.ascontiguousarray(m, dtype=np.float32).reshape(16)
                        for m in (view_proj, view_proj_prev, inv_view_prev))
        out = np.zeros((height, width, 13), dtype=np.float32)
        for y in range(height):
            for x in range(width):
                lib.ora_aov_pixel(self.ptr, C.cast(mats, C.c_void_p), len(self.scene.materials), C.byref(frame),
                                  width, height, x, y, sample_index, _p(vp), _p(vpp), _p(ivp), sky_view_z,
                                  out[y, x].ctypes.data_as(C.c_void_p))
        return out[..., 0], out[..., 1:5], out[..., 5:9], out[..., 9:13]

    def render(self, frame, width, height, sample_start, sample_count, pixels=None, threads=8, accum=None):
        lib = load()
        mats = (type(self.scene.materials[0]) * len(self.scene.materials))(*self.scene.materials)
        if accum is None:
            accum = np.zeros((height, width, 4), dtype=np.float32)
        pix = None if pixels is None else np.ascontiguousarray(pixels, dtype=np.uint32)
        cnt = Counters()
        env, _keep = self.env()
        lib.ora_render(self.ptr, C.cast(mats, C.c_void_p), len(self.scene.materials), C.byref(env),
                       C.byref(frame), width, height, sample_start, sample_count, _p(pix),
                       0 if pix is None else len(pix), threads, _p(accum), C.byref(cnt))
        return accum, {"extension_rays": cnt.extension_rays, "shadow_rays": cnt.shadow_rays}


def layer_csr(layers, n):
    """(offsets[n+1], materials) of per-instance layer lists (bottom to top)."""
    off = np.zeros(n + 1, dtype=np.uint32)
    flat = []
    for i in range(n):
        ls = layers[i] if i < len(layers) and layers[i] else []
        flat.extend(ls)
        off[i + 1] = len(flat)
    return off, np.ascontiguousarray(flat if flat else [0], dtype=np.uint32)


def temporal_resolve(accum, depth, motion, hist_in, is_reset=0, alpha_min=1.0 / 16.0, alpha_max=0.6,
                     speed_reference=6.0):
    """TemporalResolve.hlsl over float32 images: accum/motion/hist [H,W,4], depth [H,W]."""
    h, w = depth.shape
    args = [np.ascontiguousarray(a, dtype=np.float32) for a in (accum, depth, motion, hist_in)]
    out = np.zeros((h, w, 4), dtype=np.float32)
    load().ora_temporal_resolve(*[_p(a) for a in args], _p(out), w, h, int(is_reset), alpha_min, alpha_max,
                                speed_reference)
    return out


def alias_build(rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    out = np.zeros(w * h, dtype=ALIAS_DTYPE)
    integral = C.c_float()
    load().ora_alias_build(_p(rgb), w, h, _p(out), C.byref(integral))
    return out, integral.value


def _inv4(m):
    return np.linalg.inv(np.asarray(m, dtype=np.float64).reshape(4, 4)).astype(np.float32).reshape(16)


def make_env(env):
    """Oracle environment from the same dict scenes.environment_struct consumes."""
    e = Env()
    defaults = dict(light_top=(1, 1, 1), light_bottom=(0, 0, 0), background_top=(0.02, 0.25, 0.60),
                    background_bottom=(0.32, 0.79, 1.0))
    names = dict(light_top="light_top", light_bottom="light_bottom", background_top="bg_top",
                 background_bottom="bg_bottom")
    for k, val in defaults.items():
        val = env.get(k, val)
        arr = getattr(e, names[k])
        for i in range(3):
            arr[i] = val[i]
    ident = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]
    lt = env.get("light_transform", ident)
    bt = env.get("background_transform", ident)
    lti = _inv4(lt)
    for i in range(16):
        e.light_xform[i] = lt[i]
        e.light_xform_inv[i] = lti[i]
        e.bg_xform[i] = bt[i]
    e.bg_use_screen = int(env.get("background_use_screen", 0))
    e.filter = int(env.get("filter", 0))
    keep = []
    if env.get("light_rgb") is not None:
        a = np.ascontiguousarray(env["light_rgb"], dtype=np.float32)
        al, integral = alias_build(a)
        keep += [a, al]
        e.light_rgb = _p(a)
        e.light_h, e.light_w = a.shape[0], a.shape[1]
        e.alias = _p(al)
        e.light_integral = integral
    if env.get("background_rgb") is not None:
        a = np.ascontiguousarray(env["background_rgb"], dtype=np.float32)
        keep.append(a)
        e.bg_rgb = _p(a)
        e.bg_h, e.bg_w = a.shape[0], a.shape[1]
    return e, keep
