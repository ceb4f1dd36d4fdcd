"""Seeded synthetic scenes, environments and cameras for the path tracer.

All generators are deterministic for a given seed (numpy PCG64) so that the
GPU backend, the CPU oracle and the benchmark see identical inputs.

Workloads:
  * ``config1`` - plain path tracer, CPU-runnable size: ground quad + 16
    icospheres (subdivision 3, 1,280 triangles each), 256x128 sky-gradient
    lat-long environment, 512x512, 16 spp, 3 bounces.
  * ``config2`` - the 1M-triangle, 1080p workload: 200 instances of 5
    noise-displaced icospheres (subdivision 4 gives 1.02 M instanced
    triangles; subdivision 5 gives 4.1 M), random rotations, materials with
    transmission and coat, 2048x1024 HDR environment with a 1e4 sun,
    1920x1080, 64 spp, 5 bounces.
  * ``config5`` - the denoiser / temporal-resolve sequence: 10,000 instances
    of five 2,048-triangle rocks on a 4.19 M-triangle terrain (24.7 M
    instanced triangles), a sun sky, and a seeded 120-frame camera spline
    (``camera_spline``).
"""
import math
import os

import numpy as np

from . import capi

# ---------------------------------------------------------------- meshes


def icosphere(subdiv):
    """Unit icosphere: 10*4^s+2 vertices, 20*4^s triangles, outward winding."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(e, axis=1)
        uniq, inv = np.unique(key[:, 0] * (1 << 32) + key[:, 1], return_inverse=True)
        a = uniq >> 32
        b = uniq & 0xFFFFFFFF
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v.astype(np.float32), f.astype(np.uint32)


def vertex_normals(pos, idx):
    p = pos.astype(np.float64)
    fn = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n.astype(np.float32)


def displaced_icosphere(subdiv, rng, amplitude=0.15, octaves=6):
    """Icosphere displaced along the normal by a seeded smooth noise (sum of sinusoids)."""
    v, f = icosphere(subdiv)
    p = v.astype(np.float64)
    disp = np.zeros(len(p))
    for o in range(octaves):
        freq = 1.5 * (1.9 ** o)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        disp += (amplitude / (1.6 ** o)) * np.sin(freq * (p @ d) * math.pi + rng.uniform(0, 2 * math.pi))
    p *= (1.0 + disp)[:, None]
    pos = p.astype(np.float32)
    return pos, vertex_normals(pos, f), f


def rock(rng, rings=33, segments=32, amplitude=0.25, octaves=5):
    """A UV sphere of 2 * segments * (rings - 1) triangles (2,048 by default) displaced
    along the normal by seeded smooth noise."""
    th = np.linspace(0.0, math.pi, rings + 1)[1:-1]
    ph = np.arange(segments) * (2 * math.pi / segments)
    ring = np.stack([np.sin(th)[:, None] * np.cos(ph)[None, :], np.cos(th)[:, None] * np.ones_like(ph)[None, :],
                     np.sin(th)[:, None] * np.sin(ph)[None, :]], -1).reshape(-1, 3)
    p = np.concatenate([[[0.0, 1.0, 0.0]], ring, [[0.0, -1.0, 0.0]]])
    disp = np.zeros(len(p))
    for o in range(octaves):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        disp += (amplitude / (1.7 ** o)) * np.sin(1.3 * (1.9 ** o) * (p @ d) * math.pi + rng.uniform(0, 2 * math.pi))
    p *= (1.0 + disp)[:, None]
    nr = rings - 1  # latitude rows of ring vertices
    f = []
    j = np.arange(segments)
    jn = (j + 1) % segments
    f.append(np.stack([np.zeros(segments, dtype=np.int64), 1 + jn, 1 + j], 1))  # top cap
    for r in range(nr - 1):
        a, b = 1 + r * segments, 1 + (r + 1) * segments
        f.append(np.stack([a + j, a + jn, b + jn], 1))
        f.append(np.stack([a + j, b + jn, b + j], 1))
    last, bottom = 1 + (nr - 1) * segments, len(p) - 1
    f.append(np.stack([np.full(segments, bottom), last + j, last + jn], 1))  # bottom cap
    idx = np.concatenate(f).astype(np.uint32)
    pos = p.astype(np.float32)
    return pos, vertex_normals(pos, idx), idx


def terrain_height(x, z, seed, amplitude=6.0):
    """Seeded smooth height field (sum of rotated sinusoids)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = np.zeros(np.broadcast(x, z).shape)
    for o in range(6):
        a = rng.uniform(0, 2 * math.pi)
        k = 0.012 * (2.0 ** o)
        h += (amplitude / (1.8 ** o)) * np.sin(k * (math.cos(a) * x + math.sin(a) * z) + rng.uniform(0, 2 * math.pi))
    return h


def terrain(n, half, seed):
    """n x n quads (2 n^2 triangles) over [-half, half]^2 with analytic-gradient normals."""
    g = np.linspace(-half, half, n + 1)
    x, z = np.meshgrid(g, g, indexing="xy")
    y = terrain_height(x, z, seed)
    e = 1e-3 * half
    dx = (terrain_height(x + e, z, seed) - terrain_height(x - e, z, seed)) / (2 * e)
    dz = (terrain_height(x, z + e, seed) - terrain_height(x, z - e, seed)) / (2 * e)
    nrm = np.stack([-dx, np.ones_like(dx), -dz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    pos = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
    i = np.arange(n)
    a = (i[:, None] * (n + 1) + i[None, :]).reshape(-1)  # row-major quad corners
    b, c, d = a + 1, a + n + 1, a + n + 2
    idx = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)]).astype(np.uint32)
    return pos, nrm.reshape(-1, 3).astype(np.float32), idx


def ground_quad(half=10.0):
    pos = np.array([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], dtype=np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], dtype=np.float32), (4, 1))
    idx = np.array([[0, 2, 1], [0, 3, 2]], dtype=np.uint32)
    return pos, nrm, idx


def load_teapot(path):
    d = np.load(path, allow_pickle=False)
    return d["positions"], d["normals"], d["indices"].astype(np.uint32)


# ---------------------------------------------------------------- transforms


def affine(rot=None, scale=1.0, trans=(0.0, 0.0, 0.0)):
    """Row-major 3x4 object-to-world matrix T * R * S."""
    r = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64)
    m = np.zeros((3, 4))
    m[:, :3] = r * scale
    m[:, 3] = trans
    return m.astype(np.float32).reshape(12)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def inverse3x4(m12):
    """Same definition as aph_inverse3x4: double cofactors, rounded to float."""
    m = np.asarray(m12, dtype=np.float64)
    a, b, c, d, e, f, g, h, k = m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]
    A = e * k - f * h
    B = f * g - d * k
    Cc = d * h - e * g
    det = a * A + b * B + c * Cc
    r = np.array([A / det, (c * h - b * k) / det, (b * f - c * e) / det,
                  B / det, (a * k - c * g) / det, (c * d - a * f) / det,
                  Cc / det, (b * g - a * h) / det, (a * e - b * d) / det])
    out = np.zeros(12, dtype=np.float32)
    t = (m[3], m[7], m[11])
    for i in range(3):
        out[4 * i:4 * i + 3] = r[3 * i:3 * i + 3].astype(np.float32)
        out[4 * i + 3] = np.float32(-((r[3 * i] * t[0] + r[3 * i + 1] * t[1]) + r[3 * i + 2] * t[2]))
    return out


# ---------------------------------------------------------------- camera


def look_at_frame(eye, target, up, fov_y_deg, aspect, trace_depth=5, max_luminance=1000.0):
    """aph_frame for a glm::lookAt / glm::perspective camera (RendererBase.cpp:237-246)."""
    eye = np.asarray(eye, dtype=np.float64)
    fwd = np.asarray(target, dtype=np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    s = np.cross(fwd, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, fwd)
    fr = capi.Frame()
    inv = [s[0], s[1], s[2], 0.0, u[0], u[1], u[2], 0.0, -fwd[0], -fwd[1], -fwd[2], 0.0,
           eye[0], eye[1], eye[2], 1.0]
    for i, val in enumerate(inv):
        fr.inv_view[i] = float(np.float32(val))
    f = 1.0 / math.tan(math.radians(fov_y_deg) / 2.0)
    p00, p11 = np.float32(f / aspect), np.float32(f)
    fr.view_size[0] = float(np.float32(2.0) / p00)
    fr.view_size[1] = float(np.float32(2.0) / p11)
    fr.trace_depth = trace_depth
    fr.max_luminance = max_luminance
    fr.rr_enabled = 1
    fr.rr_start_depth = 3
    fr.force_opaque_shadows = 1
    return fr


def view_proj(frame, near=0.1, far=1000.0):
    """Row-major view-projection (clip = M (p, 1)) of an aph_frame's perspective camera:
    the glm::perspective * glm::lookAt product FrameData.cameraViewProj holds
    (RendererBase.cpp:237)."""
    inv = np.array(frame.inv_view, dtype=np.float64).reshape(4, 4)  # rows right, up, back, eye
    c2w = np.eye(4)
    c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = inv[0, :3], inv[1, :3], inv[2, :3], inv[3, :3]
    view = np.linalg.inv(c2w)
    p = np.zeros((4, 4))
    p[0, 0] = 2.0 / frame.view_size[0]
    p[1, 1] = 2.0 / frame.view_size[1]
    p[2, 2] = -(far + near) / (far - near)
    p[2, 3] = -2.0 * far * near / (far - near)
    p[3, 2] = -1.0
    return (p @ view).astype(np.float32)


# ---------------------------------------------------------------- environments


def sky_gradient(w, h, rng):
    """Seeded RGB32F sky: horizon/zenith colors blended by latitude plus a soft glow."""
    zen = rng.uniform(0.1, 0.6, 3) * np.array([0.5, 0.7, 1.0])
    hor = rng.uniform(0.6, 1.2, 3)
    gnd = rng.uniform(0.05, 0.3, 3)
    v = (np.arange(h) + 0.5) / h
    lat = (0.5 - v) * math.pi
    up = np.clip(np.sin(lat), 0, 1)[:, None]
    down = np.clip(-np.sin(lat), 0, 1)[:, None]
    row = hor[None, :] * (1 - up) + zen[None, :] * up
    row = row * (1 - down) + gnd[None, :] * down
    img = np.repeat(row[:, None, :], w, axis=1)
    u = (np.arange(w) + 0.5) / w
    glow = 0.5 * np.exp(-((u - rng.uniform(0.2, 0.8)) ** 2) / 0.01)[None, :, None] * up[:, :, None]
    img = img + glow
    return np.ascontiguousarray(img.astype(np.float32))


def sun_sky(w, h, rng, sun_radiance=1.0e4, sun_radius_deg=0.5):
    img = sky_gradient(w, h, rng)
    u = (np.arange(w) + 0.5) / w
    v = (np.arange(h) + 0.5) / h
    phi = (u[None, :] + 0.25) * 2 * math.pi
    theta = v[:, None] * math.pi
    dirs = np.stack([np.cos(phi) * np.sin(theta), np.cos(theta) * np.ones_like(phi), np.sin(phi) * np.sin(theta)], -1)
    el = math.radians(rng.uniform(25, 60))
    az = rng.uniform(0, 2 * math.pi)
    sd = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    cosang = dirs @ sd
    mask = cosang > math.cos(math.radians(sun_radius_deg))
    if not mask.any():
        i = np.unravel_index(np.argmax(cosang), cosang.shape)
        mask[i] = True
    img[mask] = np.array([1.0, 0.95, 0.85], dtype=np.float32) * sun_radiance
    return np.ascontiguousarray(img)


def identity4():
    return [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]


# ---------------------------------------------------------------- scene container


class Scene:
    """Meshes, instances, materials, environment and camera of one workload."""

    def __init__(self):
        self.meshes = []        # list of (pos, nrm or None, idx)
        self.instances = None   # INSTANCE_DTYPE array; mesh field = 0-based mesh index
        self.materials = []     # list of capi.Material
        self.env = None         # dict
        self.layers = None      # per instance: material indices of its layers, bottom to top
        self.layer_geometry = None  # per instance, per layer: mesh index whose UVs the layer reads, or None
        self.ground = None      # ground plane properties (dict over aph_ground_plane_defaults) or None
        self.uvs = {}           # mesh index -> (nv, 2) texture coordinates (others: u = v = i / nv)
        self.images = []        # dicts: pixels (h, w, 4) uint8 / (h, w, 4|3) float32, linearize
        self.textures = {}      # (material, slot) -> (image index, address_u, address_v)
        self.frame = None       # capi.Frame
        self.width = 0
        self.height = 0
        self.spp = 1
        self.name = ""

    @property
    def instanced_triangles(self):
        return int(sum(len(self.meshes[i][2]) for i in self.instances["mesh"]))

    @property
    def unique_triangles(self):
        return int(sum(len(m[2]) for m in self.meshes))


def _instances(list_of):
    arr = np.zeros(len(list_of), dtype=capi.INSTANCE_DTYPE)
    for i, (mesh, mat, o2w) in enumerate(list_of):
        arr[i]["o2w"] = o2w
        arr[i]["w2o"] = inverse3x4(o2w)
        arr[i]["mesh"] = mesh
        arr[i]["material"] = mat
        arr[i]["mask"] = 0xFF
    return arr


def _material(base_color=(0.8, 0.8, 0.8), metalness=0.0, roughness=0.2, **kw):
    m = capi.default_material()
    for i in range(3):
        m.base_color[i] = base_color[i]
    m.metalness = metalness
    m.specular_roughness = roughness
    for k, val in kw.items():
        if isinstance(val, (tuple, list, np.ndarray)):
            arr = getattr(m, k)
            for i in range(3):
                arr[i] = val[i]
        else:
            setattr(m, k, val)
    return m


def config1(seed=1, width=512, height=512, spp=16, trace_depth=3):
    """Ground quad + 16 icospheres (subdiv 3), 256x128 sky, 512^2 / 16 spp / 3 bounces."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    sc.name = "config1"
    gp, gn, gi = ground_quad(10.0)
    sp, sf = icosphere(3)
    sn = vertex_normals(sp, sf)
    sc.meshes = [(gp, gn, gi), (sp, sn, sf)]
    inst = [(0, 0, affine())]
    sc.materials = [_material((0.6, 0.6, 0.6), 0.0, 0.5)]
    for i in range(16):
        c = (rng.uniform(-5, 5), rng.uniform(0.2, 2.0), rng.uniform(-5, 5))
        r = rng.uniform(0.2, 1.0)
        inst.append((1, i + 1, affine(scale=r, trans=c)))
        sc.materials.append(_material(tuple(rng.uniform(0, 1, 3)), float(rng.uniform() < 0.3),
                                      float(rng.uniform(0.05, 0.8))))
    sc.instances = _instances(inst)
    sky = sky_gradient(256, 128, rng)
    sc.env = dict(light_rgb=sky, background_rgb=sky)
    sc.width, sc.height, sc.spp = width, height, spp
    sc.frame = look_at_frame((0.0, 6.0, 14.0), (0.0, 0.5, 0.0), (0, 1, 0), 45.0, width / height, trace_depth)
    return sc


def decal_images(size=64):
    """A decal's colour pattern (RGBA32F) and its round opacity mask (clear at the border)."""
    y, x = np.mgrid[0:size, 0:size]
    t = (np.arange(size) + 0.5) / size
    pattern = np.stack([np.broadcast_to(t[None, :], (size, size)), np.broadcast_to(t[:, None], (size, size)),
                        0.5 + 0.5 * np.sin(x * 0.3) * np.cos(y * 0.2), np.ones((size, size))], -1)
    r = np.hypot(x + 0.5 - size / 2, y + 0.5 - size / 2)
    mask = np.zeros((size, size, 4), np.float32)
    mask[..., :3] = (r < 0.4 * size)[..., None]
    mask[..., 3] = 1.0
    return [dict(pixels=pattern.astype(np.float32)), dict(pixels=mask)]


def config2(seed=2, subdiv=4, width=1920, height=1080, spp=64, trace_depth=5, env_w=2048, env_h=1024,
            instances=200):
    """200 instances of 5 displaced icospheres, transmission/coat materials, a textured decal
    layer on every fourth sphere, 2048x1024 sun sky."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    sc.name = "config2"
    gp, gn, gi = ground_quad(40.0)
    sc.meshes = [(gp, gn, gi)]
    for _ in range(5):
        p, n, f = displaced_icosphere(subdiv, rng)
        sc.meshes.append((p, n, f))
    inst = [(0, 0, affine())]
    sc.materials = [_material((0.5, 0.5, 0.5), 0.0, 0.6)]
    for i in range(instances):
        mesh = 1 + int(rng.integers(0, 5))
        s = rng.uniform(0.4, 1.2)
        c = (rng.uniform(-18, 18), rng.uniform(0.5, 3.0), rng.uniform(-18, 18))
        inst.append((mesh, i + 1, affine(random_rotation(rng), s, c)))
        sc.materials.append(_material(tuple(rng.uniform(0, 1, 3)), float(rng.uniform() < 0.3),
                                      float(rng.uniform(0.05, 0.8)), transmission=float(rng.uniform(0, 1)),
                                      coat=float(rng.uniform(0, 1)), coat_roughness=float(rng.uniform(0.02, 0.3))))
    sc.instances = _instances(inst)
    # one decal layer (DirectX layer semantics, with layer geometry): a textured, round-masked
    # decal over every fourth sphere instance, placed by its own geometry - a copy of the
    # sphere mesh carrying a spherical projection as texture coordinates
    decal_mat = _material((1.0, 1.0, 1.0), 0.0, 0.3)
    capi.set_texture_slot(decal_mat, capi.APH_TEX_BASE_COLOR)
    capi.set_texture_slot(decal_mat, capi.APH_TEX_OPACITY)
    sc.materials.append(decal_mat)
    decal = len(sc.materials) - 1
    for k in range(1, 6):
        p, n, f = sc.meshes[k]
        sc.meshes.append((p.copy(), n.copy(), f.copy()))
        sc.uvs[len(sc.meshes) - 1] = np.ascontiguousarray(sphere_uvs(p) * 3.0 - 1.0, dtype=np.float32)
    sc.images, sc.textures = decal_images(), {
        (decal, capi.APH_TEX_BASE_COLOR): (0, capi.APH_ADDRESS_CLAMP, capi.APH_ADDRESS_CLAMP),
        (decal, capi.APH_TEX_OPACITY): (1, capi.APH_ADDRESS_CLAMP, capi.APH_ADDRESS_CLAMP)}
    decorated = [k > 0 and k % 4 == 0 for k in range(len(inst))]
    sc.layers = [[decal] if d else [] for d in decorated]
    sc.layer_geometry = [[inst[k][0] + 5] if d else [] for k, d in enumerate(decorated)]
    sky = sun_sky(env_w, env_h, rng)
    sc.env = dict(light_rgb=sky, background_rgb=sky)
    sc.width, sc.height, sc.spp = width, height, spp
    sc.frame = look_at_frame((0.0, 14.0, 34.0), (0.0, 1.0, 0.0), (0, 1, 0), 50.0, width / height, trace_depth)
    return sc


def config5(seed=5, rocks=10000, terrain_quads=1448, half=200.0, width=1920, height=1080, spp=4, trace_depth=5,
            env_w=2048, env_h=1024):
    """10,000 instances of five 2,048-triangle rocks scattered on a 1448^2-quad
    (4.19 M-triangle) terrain, sun sky; the camera is frame 0 of camera_spline."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    sc.name = "config5"
    tp, tn, ti = terrain(terrain_quads, half, seed)
    sc.meshes = [(tp, tn, ti)] + [rock(rng) for _ in range(5)]
    sc.materials = [_material((0.35, 0.3, 0.25), 0.0, 0.8)]
    for _ in range(31):
        sc.materials.append(_material(tuple(rng.uniform(0.1, 0.9, 3)), float(rng.uniform() < 0.2),
                                      float(rng.uniform(0.1, 0.9)), coat=float(rng.uniform() < 0.3)))
    inst = [(0, 0, affine())]
    xz = rng.uniform(-0.95 * half, 0.95 * half, (rocks, 2))
    ys = terrain_height(xz[:, 0], xz[:, 1], seed)
    for k in range(rocks):
        s = float(rng.uniform(0.3, 2.5))
        inst.append((1 + int(rng.integers(0, 5)), 1 + int(rng.integers(0, 31)),
                     affine(random_rotation(rng), s, (xz[k, 0], ys[k] + 0.3 * s, xz[k, 1]))))
    sc.instances = _instances(inst)
    sky = sun_sky(env_w, env_h, rng)
    sc.env = dict(light_rgb=sky, background_rgb=sky)
    sc.width, sc.height, sc.spp = width, height, spp
    sc.seed = seed
    sc.frame = camera_spline(seed, 1, width, height, trace_depth, half)[0]
    return sc


def camera_spline(seed, frames, width, height, trace_depth=5, half=200.0, fov_y_deg=50.0):
    """A seeded closed Catmull-Rom loop of 6 eye points above the terrain of config5,
    each frame looking at a point further along the loop."""
    rng = np.random.Generator(np.random.PCG64(seed * 31 + 7))
    ang = np.sort(rng.uniform(0, 2 * math.pi, 6))
    rad = rng.uniform(0.2, 0.6, 6) * half
    ctrl = np.stack([rad * np.cos(ang), np.zeros(6), rad * np.sin(ang)], -1)
    ctrl[:, 1] = terrain_height(ctrl[:, 0], ctrl[:, 2], seed) + rng.uniform(8.0, 25.0, 6)

    def at(u):
        u = u % 6.0
        i = int(u)
        t = u - i
        p0, p1, p2, p3 = (ctrl[(i + k) % 6] for k in (-1, 0, 1, 2))
        return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t
                      + (-p0 + 3 * p1 - 3 * p2 + p3) * t * t * t)

    out = []
    for f in range(frames):
        u = 6.0 * f / max(frames, 1) * 0.25  # a quarter of the loop over the sequence
        eye = at(u)
        look = at(u + 0.35)
        look[1] = terrain_height(look[0], look[2], seed) + 1.0
        out.append(look_at_frame(tuple(eye), tuple(look), (0, 1, 0), fov_y_deg, width / height, trace_depth))
    return out


def sphere_uvs(pos):
    """Spherical texture coordinates of points on a unit sphere."""
    p = np.asarray(pos, dtype=np.float64)
    u = 0.5 + np.arctan2(p[:, 2], p[:, 0]) / (2 * math.pi)
    v = np.arccos(np.clip(p[:, 1] / np.linalg.norm(p, axis=1), -1, 1)) / math.pi
    return np.stack([u, v], 1).astype(np.float32)


def texture_images(rng):
    """Five test images: an sRGB RGBA8 checker, an RGBA32F roughness ramp, an RGB32F
    emission pattern, an RGBA8 opacity stripe image and an RGBA8 bump normal map."""
    y, x = np.mgrid[0:64, 0:64]
    a, b = rng.integers(0, 256, 4), rng.integers(0, 256, 4)
    checker = np.where((((x // 8) + (y // 8)) % 2 == 0)[..., None], a, b).astype(np.uint8)
    checker[..., 3] = 255
    y32, x32 = np.mgrid[0:32, 0:32]
    rough = np.stack([(x32 + 0.5) / 32.0, (y32 + 0.5) / 32.0, np.zeros((32, 32)), np.ones((32, 32))], -1)
    y16, x16 = np.mgrid[0:16, 0:16]
    emis = np.stack([4.0 * ((x16 + y16) % 4 == 0), 2.0 * (x16 % 3 == 0), rng.uniform(0, 1, (16, 16))], -1)
    stripes = np.zeros((32, 32, 4), np.uint8)
    stripes[..., :3] = np.where(((x32 // 4) % 2 == 0)[..., None], 255, 64)
    stripes[..., 3] = 255
    nx = 0.4 * np.sin(x * 0.4) * np.cos(y * 0.3)
    ny = 0.4 * np.cos(x * 0.25) * np.sin(y * 0.5)
    n = np.stack([nx, ny, np.ones_like(nx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    nmap = np.concatenate([np.round((n * 0.5 + 0.5) * 255), np.full((64, 64, 1), 255)], -1).astype(np.uint8)
    return [dict(pixels=checker, linearize=True), dict(pixels=rough.astype(np.float32)),
            dict(pixels=emis.astype(np.float32)), dict(pixels=stripes, linearize=False),
            dict(pixels=nmap, linearize=False)]


def textured_scene(seed=11, width=96, height=64, spp=4, trace_depth=3):
    """config1 with material textures in every slot, UV transform, address mode and
    image format: a textured ground quad (UVs 0..1, base colour checker repeated and
    rotated, mirrored normal map) and 16 spheres, half with spherical UVs and half
    with the generated u = v = i / nv, cycling through base colour (clamp),
    roughness (mirror once / wrap), emission (border), opacity (wrap / clamp) and
    normal map + base colour (wrap); sphere 13 flags an emission image without
    binding one (reads 0)."""
    sc = config1(seed=seed, width=width, height=height, spp=spp, trace_depth=trace_depth)
    rng = np.random.Generator(np.random.PCG64(seed + 100))
    sc.name = "textured"
    gp, gn, gi = sc.meshes[0]
    sp, sn, sf = sc.meshes[1]
    sc.meshes = [(gp, gn, gi), (sp, sn, sf), (sp.copy(), sn.copy(), sf.copy())]
    sc.uvs = {0: np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), 1: sphere_uvs(sp)}
    for k in range(9, 17):
        sc.instances[k]["mesh"] = 2
    sc.images = texture_images(rng)
    A = capi
    g = sc.materials[0]
    A.set_texture_slot(g, A.APH_TEX_BASE_COLOR, offset=(0.1, 0.2), scale=(0.25, 0.25), pivot=(0.5, 0.5),
                       rotation=30.0)
    A.set_texture_slot(g, A.APH_TEX_NORMAL, scale=(0.5, 0.5))
    sc.textures = {(0, A.APH_TEX_BASE_COLOR): (0, A.APH_ADDRESS_WRAP, A.APH_ADDRESS_WRAP),
                   (0, A.APH_TEX_NORMAL): (4, A.APH_ADDRESS_MIRROR, A.APH_ADDRESS_MIRROR)}
    for i in range(1, 17):
        m = sc.materials[i]
        kind = (i - 1) % 5
        if kind == 0:
            A.set_texture_slot(m, A.APH_TEX_BASE_COLOR, scale=(0.5, 0.5))
            sc.textures[(i, A.APH_TEX_BASE_COLOR)] = (0, A.APH_ADDRESS_CLAMP, A.APH_ADDRESS_CLAMP)
        elif kind == 1:
            A.set_texture_slot(m, A.APH_TEX_SPECULAR_ROUGHNESS, rotation=45.0, pivot=(0.5, 0.5))
            sc.textures[(i, A.APH_TEX_SPECULAR_ROUGHNESS)] = (1, A.APH_ADDRESS_MIRROR_ONCE, A.APH_ADDRESS_WRAP)
        elif kind == 2:
            m.emission = 0.5
            A.set_texture_slot(m, A.APH_TEX_EMISSION_COLOR, scale=(0.7, 0.7), offset=(-0.1, 0.05))
            if i != 13:
                sc.textures[(i, A.APH_TEX_EMISSION_COLOR)] = (2, A.APH_ADDRESS_BORDER, A.APH_ADDRESS_BORDER)
        elif kind == 3:
            A.set_texture_slot(m, A.APH_TEX_OPACITY, scale=(2.0, 1.0))
            sc.textures[(i, A.APH_TEX_OPACITY)] = (3, A.APH_ADDRESS_WRAP, A.APH_ADDRESS_CLAMP)
        else:
            A.set_texture_slot(m, A.APH_TEX_NORMAL)
            A.set_texture_slot(m, A.APH_TEX_BASE_COLOR, rotation=-20.0)
            sc.textures[(i, A.APH_TEX_NORMAL)] = (4, A.APH_ADDRESS_WRAP, A.APH_ADDRESS_WRAP)
            sc.textures[(i, A.APH_TEX_BASE_COLOR)] = (0, A.APH_ADDRESS_WRAP, A.APH_ADDRESS_WRAP)
    return sc


def serving_scenes(seed=7, env_w=1024, env_h=512):
    """The four resident scenes of the serving workload, about 10k, 100k, 1M and 5M
    instanced triangles: config1 (20,482), 80 x subdiv-3 (102,402), 200 x subdiv-4
    (1,024,002) and 250 x subdiv-5 (5,120,002) displaced icospheres."""
    return [config1(seed=seed), config2(seed=seed + 1, subdiv=3, instances=80, env_w=env_w, env_h=env_h),
            config2(seed=seed + 2, subdiv=4, env_w=env_w, env_h=env_h),
            config2(seed=seed + 3, subdiv=5, instances=250, env_w=env_w, env_h=env_h)]


def orbit_frame(rng, width, height, radius=14.0, target=(0.0, 0.5, 0.0), trace_depth=5):
    """A camera on a random orbit around the scene, looking at its centre."""
    th = rng.uniform(0.0, 2.0 * math.pi)
    eye = (radius * math.cos(th), float(rng.uniform(3.0, 9.0)), radius * math.sin(th))
    return look_at_frame(eye, target, (0, 1, 0), 45.0, width / height, trace_depth)


C3_RESOLUTIONS = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]


def c3_cameras(seed=3, count=64):
    """The many-camera workload on the config1 scene: `count` cameras on random orbits,
    resolutions drawn from C3_RESOLUTIONS, spp U{1..64}, 5 bounces.  Returns
    [(width, height, spp, frame)]."""
    rng = np.random.default_rng(seed)
    cams = []
    for _ in range(count):
        w, h = C3_RESOLUTIONS[int(rng.integers(0, 4))]
        cams.append((w, h, int(rng.integers(1, 65)), orbit_frame(rng, w, h)))
    return cams


def teapot_scene(path, width=128, height=128, spp=4, trace_depth=5):
    """The reference fixture: teapot, 45 degree camera at (0,0,-5) (AuroraTestHelpers.cpp:617-622)."""
    sc = Scene()
    sc.name = "teapot"
    p, n, f = load_teapot(path)
    sc.meshes = [(p, n, f)]
    sc.materials = [_material((0.8, 0.8, 0.8), 0.0, 0.2)]
    sc.instances = _instances([(0, 0, affine())])
    sc.env = dict()  # default gradient environment (EnvironmentBase.cpp:33-45)
    sc.width, sc.height, sc.spp = width, height, spp
    sc.frame = look_at_frame((0.0, 0.0, -5.0), (0.0, 0.0, 0.0), (0, 1, 0), 45.0, width / height, trace_depth)
    return sc


def obj_scene(obj_path, hdr_path=None, width=128, height=128, spp=4, trace_depth=5, eye=(0.0, 0.0, -5.0),
              target=(0.0, 0.0, 0.0), fov_y_deg=45.0):
    """A scene read from disk as Plasma reads one (Applications/Plasma/OBJLoader.cpp): one
    mesh and one identity instance per OBJ shape, the MTL materials translated to Standard
    Surface (shapes without a material get the default one, appended last), and, given
    ``hdr_path``, that Radiance image as the lat-long light and background (Plasma.cpp:1341).
    Texture maps named by the MTL file are listed in ``sc.texture_paths`` but not decoded."""
    obj = capi.load_obj(obj_path)
    sc = Scene()
    sc.name = os.path.basename(obj_path)
    sc.materials = list(obj["materials"])
    default = None
    inst = []
    for k, sh in enumerate(obj["shapes"]):
        sc.meshes.append((sh["positions"], sh["normals"], sh["indices"]))
        if sh["uvs"] is not None:
            sc.uvs[k] = sh["uvs"]
        mat = sh["material"]
        if mat < 0:
            if default is None:
                default = len(sc.materials)
                sc.materials.append(capi.default_material())
            mat = default
        inst.append((k, mat, affine()))
    if not inst:
        raise ValueError(f"{obj_path}: no faces")
    sc.instances = _instances(inst)
    sc.texture_paths = obj["textures"]
    sc.env = dict()
    if hdr_path is not None:
        rgb = capi.load_hdr(hdr_path)
        sc.env = dict(light_rgb=rgb, background_rgb=rgb)
    sc.width, sc.height, sc.spp = width, height, spp
    sc.frame = look_at_frame(eye, target, (0, 1, 0), fov_y_deg, width / height, trace_depth)
    return sc


def camera_frame(cam, width, height, trace_depth=5):
    """aph_frame of a glTF camera (aph_gltf_camera): the inverse of its view matrix as the
    camera-to-world transform; a perspective yfov (its aspect ratio, or the image's when
    the file gives none), or an orthographic xmag / ymag."""
    fr = look_at_frame((0, 0, 1), (0, 0, 0), (0, 1, 0), 45.0, width / height, trace_depth)
    c2w = np.linalg.inv(np.asarray(cam["view"], dtype=np.float64))
    for c in range(4):
        for r in range(4):
            fr.inv_view[4 * c + r] = float(np.float32(c2w[r, c]))
    if cam["type"] == 1:
        fr.is_ortho = 1
        fr.view_size[0] = float(np.float32(2.0 * cam["xmag"]))
        fr.view_size[1] = float(np.float32(2.0 * cam["ymag"]))
    else:
        aspect = cam["aspect"] if cam["aspect"] > 0 else width / height
        f = 1.0 / math.tan(cam["yfov"] / 2.0)
        p00, p11 = np.float32(f / aspect), np.float32(f)
        fr.view_size[0] = float(np.float32(2.0) / p00)
        fr.view_size[1] = float(np.float32(2.0) / p11)
    return fr


def gltf_scene(gltf_path, hdr_path=None, width=128, height=128, spp=4, trace_depth=5, camera=None,
               eye=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), fov_y_deg=45.0):
    """A glTF scene as Plasma loads one (Applications/Plasma/glTFLoader.cpp): one mesh per
    (mesh, triangle primitive), one instance per node using it at the node's world
    transform, the glTF materials translated to Standard Surface (primitives without one
    get the default material, appended last); ``camera`` picks one of the file's cameras,
    otherwise the look-at camera; ``hdr_path`` as in obj_scene."""
    g = capi.load_gltf(gltf_path)
    sc = Scene()
    sc.name = os.path.basename(gltf_path)
    sc.materials = list(g["materials"])
    for k, m in enumerate(g["meshes"]):
        sc.meshes.append((m["positions"], m["normals"], m["indices"]))
        if m["uvs"] is not None:
            sc.uvs[k] = m["uvs"]
    default = None
    inst = []
    for i in g["instances"]:
        mat = i["material"]
        if mat < 0:
            if default is None:
                default = len(sc.materials)
                sc.materials.append(capi.default_material())
            mat = default
        inst.append((i["mesh"], mat, i["o2w"]))
    if not inst:
        raise ValueError(f"{gltf_path}: no triangles")
    sc.instances = _instances(inst)
    sc.cameras = g["cameras"]
    sc.env = dict()
    if hdr_path is not None:
        rgb = capi.load_hdr(hdr_path)
        sc.env = dict(light_rgb=rgb, background_rgb=rgb)
    sc.width, sc.height, sc.spp = width, height, spp
    if camera is not None:
        sc.frame = camera_frame(g["cameras"][camera], width, height, trace_depth)
    else:
        sc.frame = look_at_frame(eye, target, (0, 1, 0), fov_y_deg, width / height, trace_depth)
    return sc


def decal_scene(path, width=96, height=96, spp=4, trace_depth=3, seed=17):
    """A textured decal (the DirectX layer loop with layer geometry, HitShaderEntryPoints.slang:84-113,
    GlobalBufferAccessors.slang:116-127): the fixture teapot, its base material checkered
    through the generated u = v = i / nv, under a decal layer whose own geometry (a copy of
    the teapot carrying a planar projection as texture coordinates) places a colour pattern
    and a round opacity mask on the teapot's side; outside the mask the layer is clear."""
    sc = teapot_scene(path, width, height, spp, trace_depth)
    sc.name = "decal"
    rng = np.random.default_rng(seed)
    p, n, f = sc.meshes[0]
    # planar projection onto the xy plane over the teapot's front: u from x, v from y
    lo, hi = np.array([-0.9, 0.1]), np.array([0.9, 1.5])
    proj = np.ascontiguousarray(((p[:, :2] - lo) / (hi - lo)).astype(np.float32))
    sc.meshes = [(p, n, f), (p.copy(), n.copy(), f.copy())]
    sc.uvs = {1: proj}
    y, x = np.mgrid[0:64, 0:64]
    pattern = np.stack([(x / 63.0), (y / 63.0), 0.5 + 0.5 * np.sin(x * 0.3) * np.cos(y * 0.2),
                        np.ones_like(x, dtype=np.float64)], -1).astype(np.float32)
    r = np.hypot(x - 31.5, y - 31.5)
    mask = np.zeros((64, 64, 4), np.float32)
    mask[..., :3] = (r < 24)[..., None].astype(np.float32)
    mask[..., 3] = 1.0
    a, b = rng.integers(0, 256, 4), rng.integers(0, 256, 4)
    checker = np.where((((x // 8) + (y // 8)) % 2 == 0)[..., None], a, b).astype(np.uint8)
    checker[..., 3] = 255
    sc.images = [dict(pixels=checker, linearize=True), dict(pixels=pattern), dict(pixels=mask)]
    base = _material((0.8, 0.8, 0.8), 0.0, 0.3)
    capi.set_texture_slot(base, capi.APH_TEX_BASE_COLOR, scale=(0.05, 0.05))
    decal = _material((1.0, 1.0, 1.0), 0.0, 0.4)
    capi.set_texture_slot(decal, capi.APH_TEX_BASE_COLOR)
    capi.set_texture_slot(decal, capi.APH_TEX_OPACITY)
    sc.materials = [base, decal]
    sc.textures = {(0, capi.APH_TEX_BASE_COLOR): (0, capi.APH_ADDRESS_WRAP, capi.APH_ADDRESS_WRAP),
                   (1, capi.APH_TEX_BASE_COLOR): (1, capi.APH_ADDRESS_CLAMP, capi.APH_ADDRESS_CLAMP),
                   (1, capi.APH_TEX_OPACITY): (2, capi.APH_ADDRESS_CLAMP, capi.APH_ADDRESS_CLAMP)}
    sc.layers = [[1]]
    sc.layer_geometry = [[1]]
    return sc


GROUND_KINDS = ("reference", "mixed")


def ground_plane_scene(path, kind="reference", width=96, height=64, spp=4, trace_depth=3, seed=13):
    """A teapot standing on the ground plane.
      reference  the reference's TestRendererGroundPlane (Tests/Aurora/Tests/TestRenderer.cpp:476-511):
                 a red teapot, the plane at the origin facing +Y with shadow and reflection
                 opacity 1 (other properties at their defaults), camera at (0, 1, -5)
      mixed      reflection opacity 0.5 at roughness 0.3 in a warm grey, a blue-black shadow at
                 opacity 0.8, a second (glass) teapot, a seeded sun sky image and a distant
                 light: both plane branches, the matte shadow against an importance-sampled sky."""
    if kind not in GROUND_KINDS:
        raise ValueError(kind)
    sc = teapot_scene(path, width, height, spp, trace_depth)
    sc.name = f"ground_{kind}"
    if kind == "reference":
        sc.materials = [_material((1.0, 0.0, 0.0))]
        sc.ground = dict(enabled=1, position=(0.0, 0.0, 0.0), normal=(0.0, 1.0, 0.0), shadow_opacity=1.0,
                         reflection_opacity=1.0)
        sc.frame = look_at_frame((0.0, 1.0, -5.0), (0.0, 0.0, 0.0), (0, 1, 0), 45.0, width / height, trace_depth)
        return sc
    rng = np.random.default_rng(seed)
    sc.materials = [_material((0.8, 0.3, 0.1), 0.0, 0.3),
                    _material((1.0, 1.0, 1.0), transmission=0.9, specular_roughness=0.05)]
    sc.instances = _instances([(0, 0, affine(trans=(-1.2, 0.0, 0.0))),
                               (0, 1, affine(rot=None, scale=0.7, trans=(1.6, 0.0, 0.8)))])
    sky = sun_sky(128, 64, rng, sun_radiance=200.0, sun_radius_deg=3.0)
    sc.env = dict(light_rgb=sky, background_rgb=sky)
    sc.ground = dict(enabled=1, position=(0.0, 0.0, 0.0), normal=(0.0, 1.0, 0.0), shadow_opacity=0.8,
                     shadow_color=(0.05, 0.05, 0.1), reflection_opacity=0.5, reflection_color=(0.8, 0.7, 0.6),
                     reflection_roughness=0.3)
    sc.frame = look_at_frame((1.0, 2.5, -6.0), (0.2, 0.3, 0.0), (0, 1, 0), 50.0, width / height, trace_depth)
    sc.frame.light_count = 1
    d = np.array([-0.4, 0.8, -0.3]) / np.linalg.norm([-0.4, 0.8, -0.3])
    sc.frame.lights[0].direction[:] = d
    sc.frame.lights[0].color_intensity[:] = [1.0, 0.95, 0.9, 2.0]
    sc.frame.lights[0].cos_radius = math.cos(math.radians(1.0))
    return sc


SWEEP_KINDS = ("subsurface", "sheen", "coat", "coat_affect", "transmission", "anisotropy", "thin_walled")


def material_sweep(path, kind, width=96, height=96, spp=4, trace_depth=5):
    """A 5 x 5 grid of teapots, one material each, varied along u = column / 4 and
    v = row / 4: the grid and sweeps of the reference's TestMaterialAdvancedMaterialProperties
    (Tests/Aurora/Tests/TestMaterial.cpp:613-830; grid at :271-306: gap 3.1, centred at
    z = 17) seen by the fixture camera (AuroraTestHelpers.cpp:617-622).
      subsurface   grey base, subsurface 0.75, colour bc0->bc1 (u), radius bc2->bc3 (v)
      sheen        grey base, sheen 1, colour bc0->bc1 (u), roughness v
      coat         coat 1 at IOR 0.5 over nothing (base 0, specular 0), colour (u), roughness (v)
      coat_affect  coat 1, colour bc0, coat_affect_color u, coat_affect_roughness v - over a
                   grey base with specular 1 (the reference sweeps it over nothing, where
                   neither parameter shows)
      transmission base 0.75, specular u, transmission v
    and two sweeps of parameters that test does not vary:
      anisotropy   metal, roughness 0.4, specular_anisotropy u, specular_rotation v / 2, with a
                   clear coat of anisotropy u and rotation v
      thin_walled  transmission 1, thin_walled, transmission colour bc0->bc1 (u), IOR 1 + v"""
    if kind not in SWEEP_KINDS:
        raise ValueError(kind)
    sc = teapot_scene(path, width, height, spp, trace_depth)
    sc.name = f"sweep_{kind}"
    bc0, bc1 = np.array([1, 0.4, 0.25]), np.array([0.35, 1, 0.3])
    bc2, bc3 = np.array([0.3, 0.2, 0.95]), np.array([0.7, 0.8, 0.05])
    grey = (0.25, 0.25, 0.25)
    mats, inst = [], []
    for i in range(5):
        v = i / 4.0
        for j in range(5):
            u = j / 4.0
            mix = tuple(bc0 + (bc1 - bc0) * u)
            if kind == "subsurface":
                m = _material(grey, subsurface=0.75, subsurface_color=mix,
                              subsurface_radius=tuple(bc2 + (bc3 - bc2) * v))
            elif kind == "sheen":
                m = _material(grey, sheen=1.0, sheen_color=mix, sheen_roughness=v)
            elif kind == "coat":
                m = _material(coat=1.0, coat_ior=0.5, base=0.0, specular=0.0, coat_color=mix, coat_roughness=v)
            elif kind == "coat_affect":
                m = _material(grey, coat=1.0, coat_ior=0.5, base=0.8, specular=1.0, coat_color=tuple(bc0),
                              coat_affect_color=u, coat_affect_roughness=v)
            elif kind == "transmission":
                m = _material(base=0.75, specular=u, transmission=v)
            elif kind == "anisotropy":
                m = _material((0.9, 0.7, 0.5), 1.0, 0.4, specular_anisotropy=u, specular_rotation=v * 0.5,
                              coat=0.5, coat_roughness=0.3, coat_anisotropy=u, coat_rotation=v)
            else:
                m = _material(transmission=1.0, thin_walled=1, transmission_color=mix, specular_ior=1.0 + v)
            mats.append(m)
            pos = np.array([j * -3.1, i * 3.1, 0.0]) + np.array([0.0, 0.0, 17.0]) - np.array([-3.1, 3.1, 0.0]) * 2
            inst.append((0, len(mats) - 1, affine(trans=pos)))
    sc.materials = mats
    sc.instances = _instances(inst)
    return sc


def environment_struct(env):
    """capi.Environment for an env dict (keeps numpy arrays alive via the returned tuple)."""
    e = capi.Environment()
    defaults = dict(light_top=(1, 1, 1), light_bottom=(0, 0, 0), background_top=(0.02, 0.25, 0.60),
                    background_bottom=(0.32, 0.79, 1.0))
    for k, val in defaults.items():
        val = env.get(k, val)
        arr = getattr(e, k)
        for i in range(3):
            arr[i] = val[i]
    lt = env.get("light_transform", identity4())
    bt = env.get("background_transform", identity4())
    for i in range(16):
        e.light_transform[i] = lt[i]
        e.background_transform[i] = bt[i]
    e.background_use_screen = int(env.get("background_use_screen", 0))
    e.filter = int(env.get("filter", 0))
    keep = []
    import ctypes as C
    if env.get("light_rgb") is not None:
        a = np.ascontiguousarray(env["light_rgb"], dtype=np.float32)
        keep.append(a)
        e.light_rgb = a.ctypes.data_as(C.POINTER(C.c_float))
        e.light_height, e.light_width = a.shape[0], a.shape[1]
    if env.get("background_rgb") is not None:
        a = np.ascontiguousarray(env["background_rgb"], dtype=np.float32)
        keep.append(a)
        e.background_rgb = a.ctypes.data_as(C.POINTER(C.c_float))
        e.background_height, e.background_width = a.shape[0], a.shape[1]
    return e, keep
