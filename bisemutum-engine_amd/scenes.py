"""Seeded procedural scenes and cameras for the wavefront path-tracing pass.

Every random draw comes from one PCG32 stream (O'Neill 2014) of the scene seed, in
the order documented per generator, so the GPU path and the CPU restatement receive
identical inputs.  Geometry is computed in float64 and stored as float32.

Scene layout follows the inputs of PathTracingPass (src/renderer/pass/path_tracing.hpp:13-23):
  meshes     object-space positions/normals (float32 xyz) + uint32 triangle indices
  instances  3x4 row-major transforms (AccelerationStructureInstanceDesc.transform,
             include/bisemutum/rhi/accel.hpp:48-55), mesh id, material id
  materials  bi_pt_material records: base_color, roughness, f0, anisotropy, f90, ior,
             two_sided -- f0 = lerp(0.04, base_color, metallic), f90 = 1, ior = 1.5
             (src/scene_basic/menu_actions/import_model.cpp:208-230)
  lights     DirLightData records (src/renderer/context/lights.hpp:12-22)
  sky        RGBA32F equirect, colour multiplier, 3x3 rotation
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MASK64 = (1 << 64) - 1


class Pcg32:
    """PCG32 (pcg32_random_r / pcg32_srandom_r), Python integers."""

    MULT = 6364136223846793005

    def __init__(self, initstate: int, initseq: int = 0):
        self.state = 0
        self.inc = ((initseq << 1) | 1) & MASK64
        self.next_u32()
        self.state = (self.state + initstate) & MASK64
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * self.MULT + self.inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self) -> float:
        return (self.next_u32() >> 8) * (1.0 / (1 << 24))

    def range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()


@dataclasses.dataclass
class Scene:
    name: str
    meshes: list          # list of (positions f32 [V,3], normals f32 [V,3], indices u32 [T,3])
    instances: np.ndarray  # structured: xform f32[12], mesh u32, material u32
    materials: np.ndarray  # f32 [M,16]
    dir_lights: np.ndarray  # f32 [L,16] (sm_index stored as int32 bits)
    sky_rgba: np.ndarray   # f32 [H,W,4]
    sky_color: np.ndarray  # f32 [3]
    sky_rot: np.ndarray    # f32 [9] row-major
    notes: dict = dataclasses.field(default_factory=dict)
    mesh_uv: list = None    # per mesh: f32 [V, 2] texture coordinates, or None (all meshes without)
    mesh_tan: list = None   # per mesh: f32 [V, 4] tangents (xyz, bitangent sign), or None
    textures: list = dataclasses.field(default_factory=list)   # Texture records
    sky_layout: int = 0     # SKY_EQUIRECT: sky_rgba [H, W, 4]; SKY_CUBE: [6, S, S, 4] (+X, -X, +Y, -Y, +Z, -Z)
    point_lights: np.ndarray = None   # f32 [P, 16] (bi_pt_point_light records), after the directional lights
    rect_lights: np.ndarray = None    # f32 [R, 28] (bi_pt_rect_light records)
    ltc: tuple = None                 # LTC tables (matrix rows 0, 1, 2: [4096, 4]; norm: [4096, 2])
    rect_textures: list = dataclasses.field(default_factory=list)   # RectTexture records

    @property
    def num_tris(self) -> int:
        return int(sum(len(self.meshes[m][2]) for m in self.instances["mesh"]))


@dataclasses.dataclass
class Texture:
    """A material texture (bi_pt_texture): level 0, texels [H, W, 4] from v = 0 (uint8
    RGBA or float32 RGBA), format TEXEL_*, filter 0 nearest / 1 linear, addressing 0
    repeat / 1 mirrored repeat / 2 clamp to edge."""
    texels: np.ndarray
    format: int = 0          # TEXEL_RGBA8_UNORM
    filter: int = 1
    address_u: int = 0
    address_v: int = 0


INSTANCE_DTYPE = np.dtype([("xform", np.float32, (12,)), ("mesh", np.uint32), ("material", np.uint32)])


MATERIAL_FLOATS = 32   # bi_pt_material: 128 bytes
# float/word index of the bi_pt_material fields
M_BLEND, M_OPACITY, M_OPACITY_TEX, M_OPACITY_CHANNEL, M_OPACITY_STEP = 15, 16, 17, 18, 19
M_STEP_AT, M_STEP_BELOW, M_STEP_ABOVE = 20, 21, 22
M_BASE_TEX, M_F0_TEX, M_NORMAL_TEX, M_ROUGH_TEX, M_ROUGH_CHANNEL = 23, 24, 25, 26, 27
BLEND_OPAQUE, BLEND_ALPHA_TEST, BLEND_TRANSLUCENT = 0, 1, 2
TEXEL_RGBA8_UNORM, TEXEL_RGBA8_SRGB, TEXEL_RGBA32F = 0, 1, 2
ADDRESS_REPEAT, ADDRESS_MIRRORED_REPEAT, ADDRESS_CLAMP = 0, 1, 2


def material(base_color, roughness, metallic, two_sided=True, anisotropy=0.0):
    """bi_pt_material record (32 floats) with the importer's metallic mapping: opaque,
    opacity 1, no textures."""
    base = np.asarray(base_color, np.float32)
    m = np.float32(metallic)
    f0 = (np.float32(0.04) + (base - np.float32(0.04)) * m).astype(np.float32)  # lerp(0.04, base, metallic)
    rec = np.zeros(MATERIAL_FLOATS, np.float32)
    rec[M_OPACITY] = 1.0
    rec[0:3] = base
    rec[3] = roughness
    rec[4:7] = f0
    rec[7] = anisotropy
    rec[8:11] = 1.0
    rec[11] = 1.5
    rec.view(np.uint32)[12] = 1 if two_sided else 0
    return rec


def dir_light(emission, direction, cast_shadow=True, shadow_strength=1.0):
    """DirLightData record: emission, sm_index, direction (to the light), shadow_strength, ..."""
    rec = np.zeros(16, np.float32)
    rec[0:3] = emission
    d = np.asarray(direction, np.float64)
    rec[4:7] = d / np.linalg.norm(d)
    rec[7] = shadow_strength
    rec.view(np.int32)[3] = 0 if cast_shadow else -1
    return rec


def point_light(emission, position, radius=20.0, spot_axis=None, inner_deg=0.0, outer_deg=0.0, cast_shadow=True,
                shadow_strength=1.0):
    """A bi_pt_point_light record (PointLightData layout): range_sqr_inv = 1 / radius^2;
    a spot when spot_axis (pointing back to the light) and inner < outer angles are given
    (cos_inner > cos_outer), as LightsContext fills it (lights.cpp:130-142)."""
    r = np.zeros(16, np.float32)
    r[0:3] = emission
    r[3] = 1.0 / (radius * radius)
    r[4:7] = position
    if spot_axis is not None:
        a = np.asarray(spot_axis, np.float64)
        r[8:11] = a / np.linalg.norm(a)
        r[7] = math.cos(math.radians(inner_deg))
        r[11] = math.cos(math.radians(outer_deg))
    r[12:13].view(np.int32)[0] = 0 if cast_shadow else -1
    r[13] = shadow_strength
    return r


def rect_light(emission, center, width, height, x_axis=(1, 0, 0), y_axis=(0, 1, 0), two_sided=False):
    """A bi_pt_rect_light record as LightsContext fills it (lights.cpp:211-227): corners
    centre + (w, h), (-w, h), (-w, -h), (w, -h) along the light's x and y axes (half sizes),
    normal x cross y, untextured."""
    x = np.asarray(x_axis, np.float64); x = x / np.linalg.norm(x)
    y = np.asarray(y_axis, np.float64); y = y / np.linalg.norm(y)
    c = np.asarray(center, np.float64)
    w, h = 0.5 * width, 0.5 * height
    r = np.zeros(28, np.float32)
    r[0:3] = emission
    r[3:4].view(np.int32)[0] = -1
    r[4:7] = c
    r[7:8].view(np.uint32)[0] = 1 if two_sided else 0
    for k, (sx, sy) in enumerate(((w, h), (-w, h), (-w, -h), (w, -h))):
        r[8 + 4 * k:11 + 4 * k] = c + sx * x + sy * y
    r[11] = 1.0 / (width * width)
    r[15] = 1.0 / (height * height)
    r[24:27] = np.cross(x, y)
    return r


@dataclasses.dataclass
class RectTexture:
    """A rect light's texture (bi_pt_rect_texture): levels, a list of [h_l, w_l, 4] arrays
    (uint8 for the RGBA8 formats, float32 for RGBA32F), level l of max(1, w >> l) x
    max(1, h >> l) texels, and the sampler."""
    levels: list
    format: int = TEXEL_RGBA32F
    mag_filter: int = 1
    min_filter: int = 1
    mipmap_mode: int = 1
    address_u: int = ADDRESS_CLAMP
    address_v: int = ADDRESS_CLAMP
    lod_bias: float = 0.0
    lod_min: float = 0.0
    lod_max: float = 1000.0

    def chain(self):
        """The levels one after the other, as 32-bit words."""
        return np.concatenate([np.ascontiguousarray(l).view(np.uint32).reshape(-1) for l in self.levels])


def textured_rect_light(record, texture_index, tex_width, tex_height, width, height):
    """A rect light record given a texture: texture_index and inv_texel_size =
    max(tex_width / width, tex_height / height) (lights.cpp:229-238)."""
    r = np.array(record, np.float32, copy=True)
    r[3:4].view(np.int32)[0] = int(texture_index)
    r[19] = max(tex_width / width, tex_height / height)
    return r


def identity_ltc(norm=(1.0, 0.0)):
    """LTC tables whose transform is the identity everywhere (the specular integral is then the
    cosine integral) and whose (norm, Fresnel) pair is constant."""
    rows = [np.zeros((4096, 4), np.float32) for _ in range(3)]
    for i in range(3):
        rows[i][:, i] = 1.0
    return (rows[0], rows[1], rows[2], np.tile(np.asarray(norm, np.float32), (4096, 1)))


def smooth_ltc(seed=0):
    """Synthetic LTC tables: transforms near diag(a, a, 1) varying smoothly over the table
    (invertible everywhere) and a (norm, Fresnel) pair in (0, 1) (a table for parity tests,
    not a fit)."""
    rng = np.random.default_rng(seed)
    i = np.arange(4096)
    x, y, z = i % 8, (i // 8) % 8, i // 64
    a = 0.3 + 0.09 * x + 0.04 * y + 0.005 * z
    rows = [np.zeros((4096, 4), np.float32) for _ in range(3)]
    rows[0][:, 0] = a
    rows[1][:, 1] = a * (0.8 + 0.02 * y)
    rows[2][:, 2] = 1.0
    for r in rows:
        r[:, :3] += (0.02 * rng.standard_normal((4096, 3))).astype(np.float32)
    norm = np.stack([0.9 - 0.05 * x, 0.02 + 0.01 * y], axis=1).astype(np.float32)
    return (rows[0], rows[1], rows[2], norm)


def random_material(rng: Pcg32, two_sided=True):
    """albedo U[0.1,0.9]^3, roughness U[0.05,1], metallic Bernoulli(0.3); draw order as written."""
    albedo = [rng.range(0.1, 0.9) for _ in range(3)]
    rough = rng.range(0.05, 1.0)
    metal = 1.0 if rng.uniform() < 0.3 else 0.0
    return material(albedo, rough, metal, two_sided)


def translate_scale(t, s=(1.0, 1.0, 1.0), rot_y=0.0):
    c, si = math.cos(rot_y), math.sin(rot_y)
    r = np.array([[c, 0, si], [0, 1, 0], [-si, 0, c]], np.float64)
    m = r @ np.diag(s)
    x = np.zeros((3, 4), np.float64)
    x[:, :3] = m
    x[:, 3] = t
    return x.reshape(-1).astype(np.float32)


# ---------------------------------------------------------------- meshes

def icosphere(subdiv: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                v = verts[a] + verts[b]
                verts.append(v / np.linalg.norm(v))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    p = np.array(verts, np.float64)
    return p.astype(np.float32), p.astype(np.float32), np.array(faces, np.uint32)


def grid_quad(origin, du, dv, nu, nv, normal):
    """nu x nv quads spanning origin + [0,1]du + [0,1]dv, flat normal; 2*nu*nv triangles."""
    origin, du, dv = (np.asarray(a, np.float64) for a in (origin, du, dv))
    iu, iv = np.meshgrid(np.arange(nu + 1), np.arange(nv + 1), indexing="ij")
    p = origin + (iu[..., None] / nu) * du + (iv[..., None] / nv) * dv
    p = p.reshape(-1, 3)
    n = np.tile(np.asarray(normal, np.float64), (len(p), 1))
    idx = []
    for a in range(nu):
        for b in range(nv):
            v00 = a * (nv + 1) + b
            v10 = (a + 1) * (nv + 1) + b
            v01 = a * (nv + 1) + b + 1
            v11 = (a + 1) * (nv + 1) + b + 1
            idx += [(v00, v10, v11), (v00, v11, v01)]
    return p.astype(np.float32), n.astype(np.float32), np.array(idx, np.uint32)


def merge(parts):
    ps, ns, ids = [], [], []
    base = 0
    for p, n, i in parts:
        ps.append(p)
        ns.append(n)
        ids.append(i + base)
        base += len(p)
    return np.concatenate(ps), np.concatenate(ns), np.concatenate(ids).astype(np.uint32)


def lathe_column(segments=64, rings=28, radius=0.25, height=3.6, flutes=16):
    """Fluted column: a lathe of `rings` x `segments` quads (2*rings*segments triangles)."""
    ys = np.linspace(0.0, height, rings + 1)
    th = np.linspace(0.0, 2 * math.pi, segments + 1)

    def prof(y):
        base = 0.35 * np.exp(-((y - 0.0) / 0.25) ** 2)      # base flare
        cap = 0.30 * np.exp(-((y - height) / 0.25) ** 2)   # capital
        return radius * (1.0 + base + cap)

    Y, T = np.meshgrid(ys, th, indexing="ij")
    flute = 1.0 + 0.06 * np.cos(flutes * T) * (1.0 - np.exp(-((Y - 0) / 0.3) ** 2) - np.exp(-((Y - height) / 0.3) ** 2))
    R = prof(Y) * flute
    P = np.stack([R * np.sin(T), Y, R * np.cos(T)], -1)
    # analytic-ish normals by finite differences on the parametric surface
    dY = np.gradient(P, ys, axis=0)
    dT = np.gradient(P, th, axis=1)
    N = np.cross(dT, dY)
    N /= np.maximum(np.linalg.norm(N, axis=-1, keepdims=True), 1e-12)
    idx = []
    for r in range(rings):
        for s in range(segments):
            a = r * (segments + 1) + s
            b = a + 1
            c = a + (segments + 1)
            d = c + 1
            idx += [(a, b, d), (a, d, c)]
    return P.reshape(-1, 3).astype(np.float32), N.reshape(-1, 3).astype(np.float32), np.array(idx, np.uint32)


def unit_arch(arc_segments=14, tube_segments=14, tube_radius=0.12, rise=0.5):
    """Semicircular tube from (0,0,0) to (1,0,0) peaking at y=rise: 2*arc*tube triangles."""
    u = np.linspace(0.0, 1.0, arc_segments + 1)
    a = np.linspace(0.0, 2 * math.pi, tube_segments + 1)
    ang = u * math.pi
    cx = 0.5 - 0.5 * np.cos(ang)
    cy = rise * np.sin(ang)
    tx = 0.5 * np.sin(ang)
    ty = rise * np.cos(ang)
    tl = np.sqrt(tx * tx + ty * ty)
    tx, ty = tx / tl, ty / tl
    nx, ny = -ty, tx           # in-plane normal of the centre curve
    U, A = np.meshgrid(np.arange(arc_segments + 1), a, indexing="ij")
    ca, sa = np.cos(A), np.sin(A)
    Nv = np.stack([nx[U] * ca, ny[U] * ca, sa], -1)
    P = np.stack([cx[U], cy[U], np.zeros_like(ca)], -1) + tube_radius * Nv
    idx = []
    for i in range(arc_segments):
        for j in range(tube_segments):
            p00 = i * (tube_segments + 1) + j
            p01 = p00 + 1
            p10 = p00 + tube_segments + 1
            p11 = p10 + 1
            idx += [(p00, p10, p11), (p00, p11, p01)]
    return P.reshape(-1, 3).astype(np.float32), Nv.reshape(-1, 3).astype(np.float32), np.array(idx, np.uint32)


# ---------------------------------------------------------------- sky

SKY_EQUIRECT, SKY_CUBE = 0, 1


def cube_face_directions(size):
    """Unit directions of the texel centres of a cube map's six faces, f32 [6, S, S, 3]:
    face f (+X, -X, +Y, -Y, +Z, -Z), row from v, column from u, with the orientation of
    the engine's layered-cube mapping (shaders/core/utils/cubemap.hlsl:3-21)."""
    c = (np.arange(size, dtype=np.float64) + 0.5) / size * 2.0 - 1.0
    b, a = np.meshgrid(c, c, indexing="ij")
    one = np.ones_like(a)
    faces = [(one, -b, -a), (-one, -b, a), (a, one, b), (a, -one, -b), (a, -b, one), (-a, -b, -one)]
    d = np.stack([np.stack(f, -1) for f in faces])
    return (d / np.linalg.norm(d, axis=-1, keepdims=True)).astype(np.float32)


def equirect_to_cube(rgba, size):
    """A cube map made from an equirect as the engine imports a skybox
    (shaders/core/equitangular_to_cubemap.hlsl:20-34): each face texel takes the
    equirect's bilinear value (wrap in u, clamp in v) in its centre direction."""
    d = cube_face_directions(size).astype(np.float64)
    H, W = rgba.shape[:2]
    th = np.arccos(np.clip(d[..., 1], -1.0, 1.0))
    ph = np.where(d[..., 2] == 0.0, 0.0, np.arctan2(d[..., 0], d[..., 2]))
    ph = np.where(ph < 0.0, ph + 2 * math.pi, ph)
    x = ph / (2 * math.pi) * W - 0.5
    y = th / math.pi * H - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = (x - x0)[..., None], (y - y0)[..., None]
    i0, j0 = y0.astype(np.int64), x0.astype(np.int64)
    ja, jb = j0 % W, (j0 + 1) % W
    ia, ib = np.clip(i0, 0, H - 1), np.clip(i0 + 1, 0, H - 1)
    src = rgba.astype(np.float64)
    top = src[ia, ja] + (src[ia, jb] - src[ia, ja]) * fx
    bot = src[ib, ja] + (src[ib, jb] - src[ib, ja]) * fx
    return (top + (bot - top) * fy).astype(np.float32)

def constant_sky(color=(0.5, 0.6, 0.8), w=64, h=32):
    rgba = np.ones((h, w, 4), np.float32)
    return rgba, np.asarray(color, np.float32)


def preetham_like_sky(sun_dir, w=512, h=256):
    """Zenith/horizon gradient above, dim ground below, plus a sun lobe (FP32 RGBA)."""
    th = (np.arange(h) + 0.5) / h * math.pi
    ph = (np.arange(w) + 0.5) / w * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.sin(P), np.cos(T), np.sin(T) * np.cos(P)], -1)
    s = np.asarray(sun_dir, np.float64)
    s = s / np.linalg.norm(s)
    zen = np.array([0.22, 0.42, 0.95])
    hor = np.array([0.85, 0.88, 0.95])
    gnd = np.array([0.12, 0.11, 0.10])
    y = d[..., 1:2]
    t = np.clip(y, 0.0, 1.0) ** 0.4
    up = hor * (1 - t) + zen * t
    col = np.where(y > 0, up, gnd + (hor - gnd) * np.exp(8.0 * np.minimum(y, 0.0)))
    cg = np.clip((d * s).sum(-1, keepdims=True), -1.0, 1.0)
    col = col + np.array([30.0, 27.0, 22.0]) * np.exp(2000.0 * (cg - 1.0)) \
        + np.array([0.6, 0.5, 0.35]) * np.exp(30.0 * (cg - 1.0))
    rgba = np.concatenate([col, np.ones_like(y)], -1).astype(np.float32)
    return rgba, np.ones(3, np.float32)


# ---------------------------------------------------------------- configs

def room_c0(half=6.0, height=4.0):
    """Open-top box room: floor 2x2 quads + 4 walls of 1x2 / 2x1 quads (24 triangles)."""
    return merge([
        grid_quad((-half, 0, -half), (2 * half, 0, 0), (0, 0, 2 * half), 2, 2, (0, 1, 0)),
        grid_quad((-half, 0, -half), (0, height, 0), (2 * half, 0, 0), 1, 2, (0, 0, 1)),
        grid_quad((-half, 0, half), (2 * half, 0, 0), (0, height, 0), 2, 1, (0, 0, -1)),
        grid_quad((-half, 0, -half), (0, 0, 2 * half), (0, height, 0), 2, 1, (1, 0, 0)),
        grid_quad((half, 0, -half), (0, height, 0), (0, 0, 2 * half), 1, 2, (-1, 0, 0)),
    ])


def scene_c0(seed: int = 0xB15E0000):
    """Box room (12 quads: floor 2x2 + 4 walls x 2, open top) + 64 icospheres (subdiv 2).

    Draw order: for each sphere: centre x, y, z, radius; then materials (room first,
    then spheres): albedo r, g, b, roughness, metallic."""
    rng = Pcg32(seed, 0)
    half, height = 6.0, 4.0
    room = room_c0(half, height)
    sphere = icosphere(2)
    inst = []
    centres = []
    for _ in range(64):
        c = (rng.range(-half + 0.6, half - 0.6), rng.range(0.6, height - 0.6), rng.range(-half + 0.6, half - 0.6))
        r = rng.range(0.2, 0.6)
        centres.append((c, r))
    mats = [random_material(rng)]
    inst.append((translate_scale((0, 0, 0)), 0, 0))
    for i, (c, r) in enumerate(centres):
        mats.append(random_material(rng))
        inst.append((translate_scale(c, (r, r, r)), 1, i + 1))
    instances = np.array(inst, INSTANCE_DTYPE)
    sky, col = constant_sky()
    sun = dir_light((2.6, 2.5, 2.3), (0.35, 0.85, 0.4))
    return Scene("c0_box_room", [room, sphere], instances, np.stack(mats), sun[None], sky, col,
                 np.eye(3, dtype=np.float32).reshape(-1),
                 notes={"room_tris": len(room[2]), "sphere_tris": len(sphere[2])})


_ICO_CACHE: dict = {}


def icosphere_cached(subdiv: int):
    if subdiv not in _ICO_CACHE:
        _ICO_CACHE[subdiv] = icosphere(subdiv)
    return _ICO_CACHE[subdiv]


def sized_room_plan(target_tris: int):
    """(icosphere subdivision, sphere count) for a c0-style room of about target_tris:
    the finest subdivision that still leaves at least 64 spheres (subdivision 0 below
    64 x 20 triangles), then as many spheres as fill the target."""
    body = max(target_tris - 24, 20)
    s = 0
    while s < 6 and 64 * 20 * 4 ** (s + 1) <= body:
        s += 1
    return s, max(1, int(round(body / (20 * 4 ** s))))


def scene_random_triangles(n: int, seed: int):
    """n triangles of edge ~1 scattered in a 10 m cube (one instance, c0's first
    material, sky and sun): small trees whose treelets reach 6 and 7 leaves."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5.0, 5.0, (n, 1, 3)).astype(np.float32)
    v = (c + rng.uniform(-0.5, 0.5, (n, 3, 3))).astype(np.float32).reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    nrm = np.tile(np.array([0.0, 1.0, 0.0], np.float32), (3 * n, 1))
    base = scene_c0()
    inst = np.zeros(1, INSTANCE_DTYPE)
    inst["xform"][0] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    return dataclasses.replace(base, name=f"random_triangles_{n}", meshes=[(v, nrm, idx)], instances=inst,
                               materials=base.materials[:1])


def scene_sized(target_tris: int, seed: int):
    """c0-style room (24 triangles) with icospheres adding up to about target_tris.

    Spheres of one subdivision (sized_room_plan); draw order as scene_c0: for each sphere
    centre x, y, z, radius; then materials (room first, then spheres)."""
    rng = Pcg32(seed, 0)
    half, height = 6.0, 4.0
    room = room_c0(half, height)
    subdiv, count = sized_room_plan(target_tris)
    sphere = icosphere_cached(subdiv)
    centres = []
    for _ in range(count):
        c = (rng.range(-half + 0.6, half - 0.6), rng.range(0.6, height - 0.6), rng.range(-half + 0.6, half - 0.6))
        centres.append((c, rng.range(0.2, 0.6)))
    mats = [random_material(rng)]
    inst = [(translate_scale((0, 0, 0)), 0, 0)]
    for i, (c, r) in enumerate(centres):
        mats.append(random_material(rng))
        inst.append((translate_scale(c, (r, r, r)), 1, i + 1))
    sky, col = constant_sky()
    sun = dir_light((2.6, 2.5, 2.3), (0.35, 0.85, 0.4))
    return Scene(f"room_{count}x{20 * 4 ** subdiv}", [room, sphere], np.array(inst, INSTANCE_DTYPE), np.stack(mats),
                 sun[None], sky, col, np.eye(3, dtype=np.float32).reshape(-1),
                 notes={"spheres": count, "sphere_tris": 20 * 4 ** subdiv})


def log_uniform_tris(rng: Pcg32, lo: float, hi: float) -> int:
    return int(round(math.exp(rng.range(math.log(lo), math.log(hi)))))


def camera_orbit(rng: Pcg32, width: int, height: int, radius=12.0, centre=(0.0, 2.0, 0.0)):
    """Camera on a sphere around a c0-style room, looking at its centre over the walls:
    azimuth U[0, 2pi), elevation U[30, 75] degrees (draw order as written)."""
    az = rng.range(0.0, 2.0 * math.pi)
    el = math.radians(rng.range(30.0, 75.0))
    eye = (centre[0] + radius * math.cos(el) * math.sin(az), centre[1] + radius * math.sin(el),
           centre[2] + radius * math.cos(el) * math.cos(az))
    return make_camera(eye, centre, width, height, 55.0)


C2_RESOLUTIONS = [(256, 256), (640, 480), (1280, 720)]
C3_RESOLUTIONS = [(256, 256), (512, 512), (1280, 720)]
C3_SPP = [1, 4, 16]


def batch_c2(count: int = 64, seed: int = 0xB15E0002):
    """Multi-scene batch: `count` c0-style scenes with triangle counts log-uniform in
    [1k, 500k], each with one job: resolution uniform over C2_RESOLUTIONS, spp U{1..8},
    depth 6 (max_bounces 7), an orbit camera.  Draws per scene: triangle count, then
    resolution, spp, camera; each scene's own geometry uses seed + 1 + index."""
    rng = Pcg32(seed, 1)
    out = []
    for i in range(count):
        tris = log_uniform_tris(rng, 1e3, 5e5)
        w, h = C2_RESOLUTIONS[min(int(rng.uniform() * 3), 2)]
        spp = 1 + min(int(rng.uniform() * 8), 7)
        cam = camera_orbit(rng, w, h)
        out.append((scene_sized(tris, seed + 1 + i), cam, dict(spp=spp, max_bounces=7, seed=0x5EED0002 + i)))
    return out


def templates_c3(count: int = 8, seed: int = 0xB15E0003):
    """Serving templates: `count` c0-style scenes, triangle counts log-uniform in [10k, 1M]."""
    rng = Pcg32(seed, 1)
    return [scene_sized(log_uniform_tris(rng, 1e4, 1e6), seed + 1 + i) for i in range(count)]


def arrivals_c3(seconds: float = 20.0, rate: float = 200.0, templates: int = 8, seed: int = 0xA11C0000):
    """Serving workload: Poisson arrivals at `rate` jobs/s over [0, seconds).  Per job
    (draw order as written): exponential inter-arrival time, template index, resolution
    (C3_RESOLUTIONS), spp (C3_SPP), orbit camera.  Returns (arrival_s, template, camera, spp)."""
    rng = Pcg32(seed, 0)
    jobs = []
    t = 0.0
    while True:
        t += -math.log(1.0 - rng.uniform()) / rate
        if t >= seconds:
            return jobs
        k = min(int(rng.uniform() * templates), templates - 1)
        w, h = C3_RESOLUTIONS[min(int(rng.uniform() * 3), 2)]
        spp = C3_SPP[min(int(rng.uniform() * 3), 2)]
        jobs.append((t, k, camera_orbit(rng, w, h), spp))


C3_MAX_BOUNCES = 7     # depth 6
C3_SEED = 0x5EED0003


def scene_c1(seed: int = 0xB15E0001):
    """Atrium of exactly 262,144 triangles.

    Split: 64 fluted lathe columns x 3,584 = 229,376; 64 arches x 392 = 25,088; ground
    floor 2,048; gallery floor 1,536; 4 walls x 1,024 = 4,096.  Open top (sky).
    Draw order: materials per object in instance order (columns floor 0 then 1, arches
    floor 0 then 1, ground, gallery, 4 walls): albedo r, g, b, roughness, metallic."""
    rng = Pcg32(seed, 0)
    outer, court, storey = 12.0, 6.0, 4.0
    column = lathe_column()
    arch = unit_arch()
    ground = grid_quad((-outer, 0, -outer), (2 * outer, 0, 0), (0, 0, 2 * outer), 32, 32, (0, 1, 0))
    # gallery: 32x32 grid minus the central 16x16 court
    gp, gn, gi = grid_quad((-outer, storey, -outer), (2 * outer, 0, 0), (0, 0, 2 * outer), 32, 32, (0, 1, 0))
    keep = []
    for a in range(32):
        for b in range(32):
            if not (8 <= a < 24 and 8 <= b < 24):
                keep += [2 * (a * 32 + b), 2 * (a * 32 + b) + 1]
    gallery = (gp, gn, gi[keep])
    wall = grid_quad((0, 0, 0), (1, 0, 0), (0, 1, 0), 32, 16, (0, 0, 1))  # unit wall, normal +z
    meshes = [column, arch, ground, gallery, wall]
    # 32 column positions on the court perimeter (8 per side)
    ring = []
    for side in range(4):
        for k in range(8):
            t = -court + 1.5 * k
            ring.append([(t, -court), (court, t), (-t, court), (-court, -t)][side])
    inst = []
    mats = []

    def add(x, mesh):
        mats.append(random_material(rng))
        inst.append((x, mesh, len(mats) - 1))

    for fl in range(2):
        for (x, z) in ring:
            add(translate_scale((x, fl * storey, z)), 0)
    col_h = 3.6
    for fl in range(2):
        for k in range(32):
            (x0, z0), (x1, z1) = ring[k], ring[(k + 1) % 32]
            span = math.hypot(x1 - x0, z1 - z0)
            ang = math.atan2(-(z1 - z0), x1 - x0)
            add(translate_scale((x0, fl * storey + col_h, z0), (span, 0.8, 1.0), ang), 1)
    add(translate_scale((0, 0, 0)), 2)
    add(translate_scale((0, 0, 0)), 3)
    wall_h = 2 * storey
    for k, (pos, ang) in enumerate([((-outer, 0, -outer), 0.0), ((outer, 0, outer), math.pi),
                                     ((-outer, 0, outer), math.pi / 2), ((outer, 0, -outer), -math.pi / 2)]):
        add(translate_scale(pos, (2 * outer, wall_h, 1.0), ang), 4)
    instances = np.array(inst, INSTANCE_DTYPE)
    sun_dir = (0.45, 0.8, 0.35)
    sky, col = preetham_like_sky(sun_dir)
    sun = dir_light((3.2, 3.0, 2.7), sun_dir)
    sc = Scene("c1_atrium", meshes, instances, np.stack(mats), sun[None], sky, col,
               np.eye(3, dtype=np.float32).reshape(-1),
               notes={"columns": 64 * len(column[2]), "arches": 64 * len(arch[2]), "ground": len(ground[2]),
                      "gallery": len(gallery[2]), "walls": 4 * len(wall[2])})
    assert sc.num_tris == 262144, sc.num_tris
    return sc


# ---------------------------------------------------------------- camera

def look_at_rh(eye, center, up):
    eye, center, up = (np.asarray(a, np.float64) for a in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    v = np.eye(4)
    v[0, :3], v[1, :3], v[2, :3] = s, u, -f
    v[0, 3], v[1, 3], v[2, 3] = -s @ eye, -u @ eye, f @ eye
    return v


def perspective_reverse_z(fovy, aspect, near, far):
    """glm::perspective (zero-to-one, RH) with the reverse-Z patch of src/math/math.cpp:5-11."""
    t = math.tan(fovy / 2)
    p = np.zeros((4, 4))
    p[0, 0] = 1.0 / (aspect * t)
    p[1, 1] = 1.0 / t
    p[2, 2] = near / (far - near)
    p[2, 3] = near * far / (far - near)
    p[3, 2] = -1.0
    return p


@dataclasses.dataclass
class Camera:
    inv_view: np.ndarray
    inv_proj: np.ndarray
    proj_view: np.ndarray
    width: int
    height: int


def make_camera(eye, center, width, height, yfov_deg=60.0, up=(0, 1, 0), near=0.05, far=500.0):
    view = look_at_rh(eye, center, up)
    proj = perspective_reverse_z(math.radians(yfov_deg), width / height, near, far)
    return Camera(np.linalg.inv(view).astype(np.float32), np.linalg.inv(proj).astype(np.float32),
                  (proj @ view).astype(np.float32), width, height)


def camera_c0(width=320, height=240):
    return make_camera((0.0, 7.5, 10.5), (0.0, 0.8, 0.0), width, height, 55.0)


def camera_c1(width=1920, height=1080):
    return make_camera((0.0, 1.7, 11.0), (0.0, 3.2, 0.0), width, height, 65.0)


CONFIGS = {
    "c0": dict(scene=scene_c0, camera=camera_c0, width=320, height=240, spp=1, max_bounces=5,
               seed=0x5EED0000),
    "c1": dict(scene=scene_c1, camera=camera_c1, width=1920, height=1080, spp=16, max_bounces=9,
               seed=0x5EED0001),
}
