"""ctypes mirror of include/aph.h, the C-ABI of libaph.so.

The structures below have exactly the C layouts of aph.h; numpy structured
dtypes with the same layouts are provided for arrays of rays, hits and
instances.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaph.so")

APH_OK = 0
APH_E_INVALID = -1
APH_E_DEVICE = -2
APH_E_MEMORY = -3
APH_E_STATE = -4
APH_E_UNSUPPORTED = -5
APH_E_OVERFLOW = -6

APH_BUILDER_LBVH = 0
APH_BUILDER_SAH = 1
APH_MODE_MEGAKERNEL = 0
APH_TLAS_REBUILD, APH_TLAS_REFIT = 0, 1
APH_MODE_WAVEFRONT = 1

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("inst", "<u4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])
INSTANCE_DTYPE = np.dtype([("o2w", "<f4", 12), ("w2o", "<f4", 12), ("mesh", "<u4"), ("material", "<u4"),
                           ("mask", "<u4"), ("flags", "<u4")])
ALIAS_DTYPE = np.dtype([("alias", "<u4"), ("prob", "<f4"), ("pdf", "<f4"), ("pad", "<f4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 20 and INSTANCE_DTYPE.itemsize == 112


class Material(C.Structure):
    """aph_material: MaterialConstants of DefaultMaterialUniformBuffer.slang:3-72 (400 B)."""
    _fields_ = [
        ("base", C.c_float), ("base_color", C.c_float * 3), ("diffuse_roughness", C.c_float),
        ("metalness", C.c_float), ("specular", C.c_float), ("_pad0", C.c_int32),
        ("specular_color", C.c_float * 3), ("specular_roughness", C.c_float), ("specular_ior", C.c_float),
        ("specular_anisotropy", C.c_float), ("specular_rotation", C.c_float), ("transmission", C.c_float),
        ("transmission_color", C.c_float * 3), ("subsurface", C.c_float), ("subsurface_color", C.c_float * 3),
        ("_pad1", C.c_int32), ("subsurface_radius", C.c_float * 3), ("subsurface_scale", C.c_float),
        ("subsurface_anisotropy", C.c_float), ("sheen", C.c_float), ("_pad2", C.c_int32), ("_pad3", C.c_int32),
        ("sheen_color", C.c_float * 3), ("sheen_roughness", C.c_float), ("coat", C.c_float),
        ("coat_color", C.c_float * 3), ("coat_roughness", C.c_float), ("coat_anisotropy", C.c_float),
        ("coat_rotation", C.c_float), ("coat_ior", C.c_float), ("coat_affect_color", C.c_float),
        ("coat_affect_roughness", C.c_float), ("emission", C.c_float), ("_pad4", C.c_int32),
        ("emission_color", C.c_float * 3), ("_pad5", C.c_int32), ("opacity", C.c_float * 3),
        ("thin_walled", C.c_int32), ("texture_block", C.c_int32 * 44),
    ]


assert C.sizeof(Material) == 400


class Environment(C.Structure):
    _fields_ = [
        ("light_top", C.c_float * 3), ("light_bottom", C.c_float * 3),
        ("background_top", C.c_float * 3), ("background_bottom", C.c_float * 3),
        ("light_transform", C.c_float * 16), ("background_transform", C.c_float * 16),
        ("background_use_screen", C.c_int32), ("filter", C.c_int32),
        ("light_rgb", C.POINTER(C.c_float)), ("light_width", C.c_uint32), ("light_height", C.c_uint32),
        ("background_rgb", C.POINTER(C.c_float)), ("background_width", C.c_uint32),
        ("background_height", C.c_uint32),
    ]


class DistantLight(C.Structure):
    _fields_ = [("color_intensity", C.c_float * 4), ("direction", C.c_float * 3), ("cos_radius", C.c_float)]


class Frame(C.Structure):
    _fields_ = [
        ("inv_view", C.c_float * 16), ("view_size", C.c_float * 2), ("jitter", C.c_float * 2),
        ("is_ortho", C.c_int32), ("focal_distance", C.c_float), ("lens_radius", C.c_float),
        ("trace_depth", C.c_int32), ("max_luminance", C.c_float), ("rr_enabled", C.c_int32),
        ("rr_start_depth", C.c_int32), ("force_opaque_shadows", C.c_int32), ("display_errors", C.c_int32),
        ("temporal_jitter", C.c_int32), ("light_count", C.c_int32), ("lights", DistantLight * 4),
    ]


class Request(C.Structure):
    _fields_ = [("target", C.c_uint64), ("tlas", C.c_uint64), ("frame", Frame), ("sample_start", C.c_uint32),
                ("sample_count", C.c_uint32), ("seed_offset", C.c_uint32), ("_pad", C.c_uint32)]


class GroundPlane(C.Structure):
    """aph_ground_plane: the properties of PTGroundPlane.cpp."""
    _fields_ = [("enabled", C.c_int32), ("position", C.c_float * 3), ("normal", C.c_float * 3),
                ("shadow_opacity", C.c_float), ("shadow_color", C.c_float * 3), ("reflection_opacity", C.c_float),
                ("reflection_color", C.c_float * 3), ("reflection_roughness", C.c_float)]


def ground_plane(props=None):
    """GroundPlane with the library's defaults, overridden by the dict `props`
    (keys: the field names)."""
    g = GroundPlane()
    load().aph_ground_plane_defaults(C.byref(g))
    for k, v in (props or {}).items():
        if isinstance(v, (tuple, list, np.ndarray)):
            getattr(g, k)[:] = [float(x) for x in v]
        else:
            setattr(g, k, type(getattr(g, k))(v))
    return g


class ObjShape(C.Structure):
    _fields_ = [("nv", C.c_uint32), ("nt", C.c_uint32), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("material", C.c_int32)]


class ObjFile(C.Structure):
    _fields_ = [("nshape", C.c_uint32), ("nmat", C.c_uint32), ("shapes", C.POINTER(ObjShape)),
                ("materials", C.POINTER(Material)), ("material_names", C.POINTER(C.c_char_p)),
                ("textures", C.POINTER(C.c_char_p))]


class GltfMesh(C.Structure):
    _fields_ = [("nv", C.c_uint32), ("nt", C.c_uint32), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32))]


class GltfInstance(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("material", C.c_int32), ("o2w", C.c_float * 12), ("node", C.c_int32),
                ("materialx", C.c_char_p)]


class GltfCamera(C.Structure):
    _fields_ = [("type", C.c_int32), ("node", C.c_int32), ("view", C.c_float * 16), ("yfov", C.c_float),
                ("aspect", C.c_float), ("znear", C.c_float), ("zfar", C.c_float), ("xmag", C.c_float),
                ("ymag", C.c_float)]


class GltfFile(C.Structure):
    _fields_ = [("nmesh", C.c_uint32), ("ninst", C.c_uint32), ("nmat", C.c_uint32), ("ncam", C.c_uint32),
                ("meshes", C.POINTER(GltfMesh)), ("instances", C.POINTER(GltfInstance)),
                ("materials", C.POINTER(Material)), ("cameras", C.POINTER(GltfCamera))]


class Stats(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("node_visits", C.c_uint64),
                ("tri_tests", C.c_uint64), ("inst_visits", C.c_uint64), ("paths", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("shadow_inst_visits", C.c_uint64), ("hits", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("chunk_paths", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("trace_refill_lanes", C.c_uint32),
                ("trace_leaf_wait_lanes", C.c_uint32), ("ms", C.c_float)]


APH_LAUNCH_GRAPH, APH_LAUNCH_EAGER = 0, 1


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_active", C.c_uint32), ("step_samples", C.c_uint32),
                ("mode", C.c_int), ("keep_results", C.c_int)]


class Job(C.Structure):
    _fields_ = [("tlas", C.c_uint64), ("frame", Frame), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("seed_offset", C.c_uint32)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int), ("samples_done", C.c_uint32), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


APH_JOB_PENDING, APH_JOB_ACTIVE, APH_JOB_DONE, APH_JOB_FAILED = 0, 1, 2, 3


class AovRequest(C.Structure):
    _fields_ = [("aov", C.c_uint64), ("tlas", C.c_uint64), ("frame", Frame), ("view_proj", C.c_float * 16),
                ("view_proj_prev", C.c_float * 16), ("inv_view_prev", C.c_float * 16), ("sky_view_z", C.c_float),
                ("sample_index", C.c_uint32), ("seed_offset", C.c_uint32), ("_pad", C.c_uint32)]


class TaaRequest(C.Structure):
    _fields_ = [("current", C.c_uint64), ("aov", C.c_uint64), ("history_in", C.c_uint64),
                ("history_out", C.c_uint64), ("is_reset", C.c_int32), ("blend_alpha_min", C.c_float),
                ("blend_alpha_max", C.c_float), ("speed_reference", C.c_float)]

class ImageDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_int32), ("linearize", C.c_int32),
                ("pixels", C.c_void_p)]


class TextureBinding(C.Structure):
    _fields_ = [("image", C.c_uint64), ("address_u", C.c_int32), ("address_v", C.c_int32)]


APH_IMAGE_RGBA8, APH_IMAGE_RGBA32F, APH_IMAGE_RGB32F = 0, 1, 2
APH_ADDRESS_WRAP, APH_ADDRESS_MIRROR, APH_ADDRESS_CLAMP, APH_ADDRESS_MIRROR_ONCE, APH_ADDRESS_BORDER = 0, 1, 2, 3, 4
APH_TEX_BASE_COLOR, APH_TEX_SPECULAR_ROUGHNESS, APH_TEX_EMISSION_COLOR, APH_TEX_OPACITY, APH_TEX_NORMAL = 0, 1, 2, 3, 4
APH_TEX_SLOTS = 5
# word offsets (from byte 224, Material.texture_block) of each slot's flag, offset, scale,
# pivot and rotation (DefaultMaterialUniformBuffer.slang:43-71)
TEX_WORDS = [(0, 1, 4, 6, 8), (9, 10, 12, 14, 16), (17, 18, 20, 22, 24), (25, 26, 28, 30, 32), (33, 34, 36, 38, 40)]


def set_texture_slot(m, slot, offset=(0.0, 0.0), scale=(1.0, 1.0), pivot=(0.0, 0.0), rotation=0.0, enabled=True):
    """Sets a material's has_*_image flag and UV transform (the 400-byte layout)."""
    fl, off, sc, pv, rot = TEX_WORDS[slot]
    words = np.zeros(44, dtype=np.int32)
    words[:] = list(m.texture_block)
    f = words.view(np.float32)
    words[fl] = 1 if enabled else 0
    f[off:off + 2] = offset
    f[sc:sc + 2] = scale
    f[pv:pv + 2] = pivot
    f[rot] = rotation
    for i in range(44):
        m.texture_block[i] = int(words[i])


class DnBuffer(C.Structure):
    _fields_ = [("level", C.c_uint32), ("channels", C.c_uint32)]


class DnOp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("layer", C.c_int32), ("src", C.c_int32), ("src_offset", C.c_int32),
                ("channels", C.c_int32), ("dst", C.c_int32), ("dst_offset", C.c_int32)]


class DnLayer(C.Structure):
    _fields_ = [("name", C.c_char * 16), ("cin", C.c_uint32), ("cin_padded", C.c_uint32), ("cout", C.c_uint32),
                ("cout_padded", C.c_uint32), ("in_buffer", C.c_uint32), ("out_buffer", C.c_uint32),
                ("out_offset", C.c_uint32), ("relu", C.c_uint32), ("weight_offset", C.c_uint64),
                ("bias_offset", C.c_uint64)]


class DenoiseExample(C.Structure):
    _fields_ = [("noisy", C.c_uint64), ("aov", C.c_uint64), ("reference", C.c_uint64), ("transform", C.c_uint32),
                ("_pad", C.c_uint32)]


class DenoiseItem(C.Structure):
    _fields_ = [("noisy", C.c_uint64), ("aov", C.c_uint64), ("output", C.c_uint64), ("flags", C.c_uint32),
                ("_pad", C.c_uint32)]


APH_DN_OP_CONV, APH_DN_OP_POOL, APH_DN_OP_UPSAMPLE = 0, 1, 2
APH_DN_LOSS_L1_LOG, APH_DN_LOSS_L2_TONEMAPPED = 0, 1
APH_DN_ENSEMBLE = 1
APH_DN_PARAMS, APH_DN_PARAM_GRADS, APH_DN_ACTIVATIONS, APH_DN_ACTIVATION_GRADS, APH_DN_PREDICTION = 0, 1, 2, 3, 4

_P = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_i = C.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "aph_version": (_i, []),
    "aph_create": (_i, [_i, C.POINTER(_P)]),
    "aph_destroy": (None, [_P]),
    "aph_last_error": (C.c_char_p, [_P]),
    "aph_get_stream": (_i, [_P, C.POINTER(_P)]),
    "aph_mesh_upload": (_i, [_P, _P, _u32, _P, _u32, _P, _P, C.POINTER(_u64)]),
    "aph_blas_build": (_i, [_P, _u64, _i]),
    "aph_blas_export": (_i, [_P, _u64, _P, _u64, _P, _u64, C.POINTER(_u32), C.POINTER(_u32)]),
    "aph_tlas_build": (_i, [_P, _P, _u32, C.POINTER(_u64)]),
    "aph_tlas_set_layers": (_i, [_P, _u64, _P, _P]),
    "aph_tlas_set_layer_geometry": (_i, [_P, C.c_uint64, _P]),
    "aph_inverse3x4": (_i, [_P, _P]),
    "aph_materials_upload": (_i, [_P, _P, _u32]),
    "aph_material_defaults": (None, [C.POINTER(Material)]),
    "aph_environment_upload": (_i, [_P, C.POINTER(Environment)]),
    "aph_environment_create": (_i, [_P, C.POINTER(Environment), C.POINTER(C.c_uint64)]),
    "aph_environment_destroy": (_i, [_P, C.c_uint64]),
    "aph_tlas_set_environment": (_i, [_P, C.c_uint64, C.c_uint64]),
    "aph_ground_plane_defaults": (None, [C.POINTER(GroundPlane)]),
    "aph_tlas_set_ground_plane": (_i, [_P, C.c_uint64, C.POINTER(GroundPlane)]),
    "aph_alias_build": (_i, [_P, _u32, _u32, _P, C.POINTER(C.c_float)]),
    "aph_trace_closest": (_i, [_P, _u64, _P, _u32, _P]),
    "aph_trace_any": (_i, [_P, _u64, _P, _u32, _P]),
    "aph_trace_closest_device": (_i, [_P, _u64, _P, _u32, _P]),
    "aph_trace_any_device": (_i, [_P, _u64, _P, _u32, _P]),
    "aph_target_create": (_i, [_P, _u32, _u32, C.POINTER(_u64)]),
    "aph_target_destroy": (_i, [_P, _u64]),
    "aph_target_device_ptr": (_i, [_P, _u64, C.POINTER(_P)]),
    "aph_render": (_i, [_P, _P, _u32, _i]),
    "aph_wait": (_i, [_P]),
    "aph_readback": (_i, [_P, _u64, _P, C.c_size_t]),
    "aph_post": (_i, [_P, _u64, _P, _i, _i, _P, C.c_size_t]),
    "aph_set_stats_level": (_i, [_P, _i]),
    "aph_stats_get": (_i, [_P, C.POINTER(Stats)]),
    "aph_stats_reset": (_i, [_P]),
    "aph_timing_enable": (_i, [_P, _i]),
    "aph_timing_get": (_i, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(_u64)]),
    "aph_timing_reset": (_i, [_P]),
    "aph_set_launch_mode": (_i, [_P, _i]),
    "aph_get_tuning": (_i, [_P, C.POINTER(Tuning)]),
    "aph_set_tuning": (_i, [_P, C.POINTER(Tuning)]),
    "aph_wavefront_autotune": (_i, [_P, _P, _u32, C.POINTER(Tuning)]),
    "aph_aov_create": (_i, [_P, _u32, _u32, C.POINTER(_u64)]),
    "aph_aov_destroy": (_i, [_P, _u64]),
    "aph_aov_readback": (_i, [_P, _u64, _P, _P, _P, _P]),
    "aph_render_aovs": (_i, [_P, C.POINTER(AovRequest)]),
    "aph_temporal_resolve": (_i, [_P, C.POINTER(TaaRequest)]),
    "aph_halton_jitter": (None, [_u32, _u32, _P]),
    "aph_server_create": (_i, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "aph_server_destroy": (None, [_P]),
    "aph_server_clock": (C.c_double, [_P]),
    "aph_server_submit": (_i, [_P, C.POINTER(Job), C.c_double, C.POINTER(_u64)]),
    "aph_server_step": (_i, [_P, C.POINTER(_u32), C.POINTER(_u32)]),
    "aph_server_status": (_i, [_P, _u64, C.POINTER(JobStatus)]),
    "aph_server_result": (_i, [_P, _u64, _P, C.c_size_t]),
    "aph_server_release": (_i, [_P, _u64]),
    "aph_server_usage": (_i, [_P, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u32), C.POINTER(_u32)]),
    "aph_server_working_set": (_i, [_P, C.POINTER(_u64), C.POINTER(_u32)]),
    "aph_render_noalloc": (_i, [_P, _P, _u32]),
    "aph_tlas_update": (_i, [_P, _u64, _P, _u32, _i]),
    "aph_obj_load": (_i, [C.c_char_p, C.POINTER(C.POINTER(ObjFile))]),
    "aph_obj_free": (None, [C.POINTER(ObjFile)]),
    "aph_gltf_load": (_i, [C.c_char_p, C.POINTER(C.POINTER(GltfFile))]),
    "aph_gltf_free": (None, [C.POINTER(GltfFile)]),
    "aph_hdr_load": (_i, [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(_u32), C.POINTER(_u32)]),
    "aph_host_free": (None, [_P]),
    "aph_materials_update": (_i, [_P, _u32, _P, _u32]),
    "aph_server_replay": (_i, [_P, _P, _P, _u32, _P]),
    "aph_image_create": (_i, [_P, C.POINTER(ImageDesc), C.POINTER(_u64)]),
    "aph_image_destroy": (_i, [_P, _u64]),
    "aph_materials_bind_textures": (_i, [_P, _P, _u32]),
    "aph_denoiser_arch": (_i, [_P, C.POINTER(_u32), _P, C.POINTER(_u32), _P, C.POINTER(_u32)]),
    "aph_denoiser_param_count": (_u64, []),
    "aph_denoiser_create": (_i, [_P, _u64, C.POINTER(_u64)]),
    "aph_denoiser_destroy": (_i, [_P, _u64]),
    "aph_denoiser_set_params": (_i, [_P, _u64, _P, _u64]),
    "aph_denoiser_set_loss": (_i, [_P, _u64, _i]),
    "aph_denoiser_tensor_size": (_i, [_P, _u64, _i, _u32, C.POINTER(_u64)]),
    "aph_denoiser_tensor": (_i, [_P, _u64, _i, _u32, _P, _u64]),
    "aph_denoiser_train_step": (_i, [_P, _u64, _P, _u32, C.c_float, C.POINTER(C.c_float)]),
    "aph_denoise": (_i, [_P, _u64, _P, _u32]),
}

_lib = None


def load():
    """Loads libaph.so; raises when the HIP backend has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP backend library missing: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def default_material():
    m = Material()
    load().aph_material_defaults(C.byref(m))
    return m


TEXTURE_SLOT_NAMES = ("base_color", "specular_roughness", "emission_color", "opacity", "normal")


def load_obj(path):
    """Parses a Wavefront OBJ file (and its MTL libraries) with aph_obj_load.

    Returns a dict: ``shapes`` - list of dicts with float32 ``positions`` (nv, 3),
    ``normals`` (nv, 3), ``uvs`` (nv, 2) or None, uint32 ``indices`` (nt, 3) and the
    ``material`` index (-1: default material); ``materials`` - list of Material;
    ``material_names``; ``textures`` - per material a dict slot name -> file path."""
    lib = load()
    f = C.POINTER(ObjFile)()
    rc = lib.aph_obj_load(os.fsencode(path), C.byref(f))
    if rc != APH_OK:
        raise RuntimeError(f"aph_obj_load({path!r}) failed: {rc}")
    try:
        o = f.contents
        shapes = []
        for k in range(o.nshape):
            s = o.shapes[k]

            def arr(p, n, c, dt):
                return np.ctypeslib.as_array(p, shape=(n, c)).astype(dt, copy=True) if n else np.zeros((0, c), dt)

            shapes.append(dict(positions=arr(s.positions, s.nv, 3, np.float32),
                               normals=arr(s.normals, s.nv, 3, np.float32),
                               uvs=arr(s.uvs, s.nv, 2, np.float32) if s.uvs else None,
                               indices=arr(s.indices, s.nt, 3, np.uint32), material=int(s.material)))
        mats, names, tex = [], [], []
        for i in range(o.nmat):
            m = Material()
            C.memmove(C.byref(m), C.byref(o.materials[i]), C.sizeof(Material))
            mats.append(m)
            names.append(o.material_names[i].decode())
            tex.append({TEXTURE_SLOT_NAMES[s]: o.textures[5 * i + s].decode()
                        for s in range(5) if o.textures[5 * i + s]})
        return dict(shapes=shapes, materials=mats, material_names=names, textures=tex)
    finally:
        lib.aph_obj_free(f)


def load_gltf(path):
    """Parses a glTF 2.0 file (.gltf or .glb) with aph_gltf_load.

    Returns a dict: ``meshes`` - list of dicts with float32 ``positions`` (nv, 3),
    ``normals`` (nv, 3), ``uvs`` (nv, 2) or None and uint32 ``indices`` (nt, 3);
    ``instances`` - list of dicts with ``mesh``, ``material`` (-1: default), float32
    ``o2w`` (12,), ``node`` and ``materialx`` (path or None); ``materials`` - list of
    Material; ``cameras`` - list of dicts with ``type`` (0 perspective, 1 orthographic),
    ``node``, float32 ``view`` (4, 4) row-major world-to-view and the projection fields."""
    lib = load()
    f = C.POINTER(GltfFile)()
    rc = lib.aph_gltf_load(os.fsencode(path), C.byref(f))
    if rc != APH_OK:
        raise RuntimeError(f"aph_gltf_load({path!r}) failed: {rc}")
    try:
        g = f.contents

        def arr(p, n, c, dt):
            return np.ctypeslib.as_array(p, shape=(n, c)).astype(dt, copy=True) if n else np.zeros((0, c), dt)

        meshes = []
        for k in range(g.nmesh):
            m = g.meshes[k]
            meshes.append(dict(positions=arr(m.positions, m.nv, 3, np.float32),
                               normals=arr(m.normals, m.nv, 3, np.float32),
                               uvs=arr(m.uvs, m.nv, 2, np.float32) if m.uvs else None,
                               indices=arr(m.indices, m.nt, 3, np.uint32)))
        insts = []
        for k in range(g.ninst):
            i = g.instances[k]
            insts.append(dict(mesh=int(i.mesh), material=int(i.material), node=int(i.node),
                              o2w=np.array(i.o2w[:], dtype=np.float32),
                              materialx=i.materialx.decode() if i.materialx else None))
        mats = []
        for k in range(g.nmat):
            m = Material()
            C.memmove(C.byref(m), C.byref(g.materials[k]), C.sizeof(Material))
            mats.append(m)
        cams = []
        for k in range(g.ncam):
            c = g.cameras[k]
            cams.append(dict(type=int(c.type), node=int(c.node),
                             view=np.array(c.view[:], dtype=np.float32).reshape(4, 4), yfov=c.yfov,
                             aspect=c.aspect, znear=c.znear, zfar=c.zfar, xmag=c.xmag, ymag=c.ymag))
        return dict(meshes=meshes, instances=insts, materials=mats, cameras=cams)
    finally:
        lib.aph_gltf_free(f)


def load_hdr(path):
    """Reads a Radiance .hdr image with aph_hdr_load: float32 (height, width, 3), top row first."""
    lib = load()
    p = C.POINTER(C.c_float)()
    w, h = C.c_uint32(), C.c_uint32()
    rc = lib.aph_hdr_load(os.fsencode(path), C.byref(p), C.byref(w), C.byref(h))
    if rc != APH_OK:
        raise RuntimeError(f"aph_hdr_load({path!r}) failed: {rc}")
    try:
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 3)).copy()
    finally:
        lib.aph_host_free(p)
