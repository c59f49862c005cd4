"""Scene ingestion from the engine's own formats: StaticMesh assets and scene projects.

StaticMesh asset (`.static_mesh.biasset`, src/scene_basic/static_mesh.cpp:10-43):
  u32 magic 0x0B1A55E7 (include/bisemutum/runtime/asset.hpp:15), string "StaticMesh",
  u32 version; version 1 = MeshData follows, version 2 = u64 uncompressed length,
  u64 compressed length, zlib stream of MeshData (ReadByteStream::read_compressed_part,
  src/prelude/byte_stream.cpp:28-41; miniz's mz_uncompress reads the zlib format).
  Strings and vectors are a u64 element count followed by the raw elements
  (include/bisemutum/prelude/byte_stream.hpp: read(std::string&), read(std::vector<T>&)),
  all little-endian.  MeshData (src/graphics/mesh.cpp:137-146) in order: positions float3,
  normals float3, tangents float4, colors float3, texcoords float2, texcoords2 float2,
  indices u32, submeshes SubmeshDesc {u32 base_vertex, u32 index_offset, u32 num_indices,
  u8 topology, 3 padding bytes} (include/bisemutum/graphics/mesh.hpp:16-22,75-84).

Scene project (examples/scene_basic): project.toml names the asset metadata and the scene
file under the virtual root "/project"; scene.toml lists objects (optional "parent"
index, src/runtime/scene.cpp:103-135) with components.  Used here:
  Transform        translation, scaling, rotation = Euler degrees, R = Rz(z) Rx(x) Ry(y)
                   (Transform::from_value, src/math/transform.cpp:119-124: glm::eulerAngleZXY);
                   world = parent_world * local as matrices (transform.cpp:89-91,
                   src/runtime/scene_object.cpp:31-52)
  CameraComponent  position = T(0), front = T_dir(0,0,-1), up = T_dir(0,1,0)
                   (src/scene_basic/camera_system.cpp:19-21); perspective yfov / near / far
                   over render_target_size (src/graphics/camera.cpp:83-98)
  DirectionalLightComponent  emission = color * strength (zero emission: no light), direction
                   = R (0,1,0) without scaling, shadow when cast_shadow, shadow_strength > 0 and
                   shadow_range > 0 (src/renderer/context/lights.cpp:55-94)
  StaticMeshComponent + MeshRendererComponent  one drawable per submesh from
                   submesh_start_index, material i for drawable i
                   (src/scene_basic/static_mesh_render_system.cpp:58-83)
  BasicRendererOverrideVolume.settings.path_tracing  max_bounces, ray_length, accumulate
                   (include/bisemutum/renderer/basic.hpp:76-81)
A project has no skybox here, so the sky is black, as the engine's default skybox.

Materials (`materials/*.toml`): blend_mode, surface_model, params and a material_function
of HLSL statements run on surface_data_default() (shaders/core/material/utils.hlsl:18-31).
The pass takes parameter materials, so a function is read when every statement assigns a
surface field from a parameter or a literal, or selects between two of them by the world
checkerboard cell (int(floor(x)) ^ int(floor(z))) & 1 of the hit point, as the project's
checkerboard material does (the pass's pattern materials).  Texture sampling, other
expressions over vertex data and non-opaque blending with opacity below one are reported
as unsupported.
"""
from __future__ import annotations

import dataclasses
import math
import os
import re
import struct
import zlib

import numpy as np

import scenes

ASSET_MAGIC = 0x0B1A55E7
STATIC_MESH_TYPE = "StaticMesh"
TOPOLOGY_TRIANGLE_LIST = 3  # rhi::PrimitiveTopology (include/bisemutum/rhi/pipeline.hpp:34-43)

SUBMESH_DTYPE = np.dtype([("base_vertex", "<u4"), ("index_offset", "<u4"), ("num_indices", "<u4"),
                          ("topology", "u1"), ("_pad", "u1", (3,))])
assert SUBMESH_DTYPE.itemsize == 16

_MESH_FIELDS = (("positions", 3), ("normals", 3), ("tangents", 4), ("colors", 3),
                ("texcoords", 2), ("texcoords2", 2))


class AssetError(ValueError):
    """A file that is not a valid asset of the expected type."""


class UnsupportedMaterial(ValueError):
    """A material whose function the pass cannot represent as parameters."""


@dataclasses.dataclass
class MeshData:
    positions: np.ndarray   # f32 [V,3]
    normals: np.ndarray     # f32 [V,3] (may be empty)
    tangents: np.ndarray    # f32 [V,4]
    colors: np.ndarray      # f32 [V,3]
    texcoords: np.ndarray   # f32 [V,2]
    texcoords2: np.ndarray  # f32 [V,2]
    indices: np.ndarray     # u32 [I]
    submeshes: np.ndarray   # SUBMESH_DTYPE [S]


class _Reader:
    def __init__(self, data: bytes, what: str):
        self.data = memoryview(data)
        self.off = 0
        self.what = what

    def take(self, n: int) -> memoryview:
        if n < 0 or self.off + n > len(self.data):
            raise AssetError(f"{self.what}: truncated at byte {self.off} (need {n} more)")
        v = self.data[self.off:self.off + n]
        self.off += n
        return v

    def u32(self) -> int:
        return struct.unpack("<I", self.take(4))[0]

    def u64(self) -> int:
        return struct.unpack("<Q", self.take(8))[0]

    def string(self) -> str:
        return bytes(self.take(self.u64())).decode("utf-8", "replace")

    def vector(self, dtype, width: int = 1) -> np.ndarray:
        dt = np.dtype(dtype)
        count = self.u64()
        if count > (len(self.data) - self.off) // (dt.itemsize * width):
            raise AssetError(f"{self.what}: vector of {count} elements exceeds the file")
        a = np.frombuffer(self.take(count * dt.itemsize * width), dt).copy()
        return a.reshape(count, width) if width > 1 else a


def _read_mesh_data(r: _Reader) -> MeshData:
    fields = {name: r.vector("<f4", w) for name, w in _MESH_FIELDS}
    indices = r.vector("<u4")
    submeshes = r.vector(SUBMESH_DTYPE)
    return MeshData(indices=indices, submeshes=submeshes, **fields)


def read_static_mesh(data: bytes, what: str = "static mesh") -> MeshData:
    """Parse a StaticMesh asset (versions 1 and 2)."""
    r = _Reader(data, what)
    magic = r.u32()
    type_name = r.string()
    version = r.u32()
    if magic != ASSET_MAGIC or type_name != STATIC_MESH_TYPE:
        raise AssetError(f"{what}: not a StaticMesh asset (magic {magic:#010x}, type {type_name!r})")
    if version == 1:
        return _read_mesh_data(r)
    if version == 2:
        length = r.u64()
        compressed = r.u64()
        raw = zlib.decompress(bytes(r.take(compressed)))
        if len(raw) != length:
            raise AssetError(f"{what}: compressed part holds {len(raw)} bytes, header says {length}")
        return _read_mesh_data(_Reader(raw, what))
    raise AssetError(f"{what}: unknown StaticMesh version {version}")


def encode_static_mesh(mesh: MeshData, version: int = 2) -> bytes:
    """Write a StaticMesh asset in the layout read_static_mesh reads (StaticMesh::save writes version 2)."""
    body = bytearray()

    def vec(a, dtype):
        a = np.ascontiguousarray(a, dtype)
        body.extend(struct.pack("<Q", len(a)))
        body.extend(a.tobytes())

    for name, _ in _MESH_FIELDS:
        vec(getattr(mesh, name), "<f4")
    vec(mesh.indices, "<u4")
    vec(mesh.submeshes, SUBMESH_DTYPE)
    name = STATIC_MESH_TYPE.encode()
    head = struct.pack("<IQ", ASSET_MAGIC, len(name)) + name + struct.pack("<I", version)
    if version == 1:
        return head + bytes(body)
    if version == 2:
        z = zlib.compress(bytes(body))
        return head + struct.pack("<QQ", len(body), len(z)) + z
    raise ValueError(f"unknown StaticMesh version {version}")


def submesh_triangles(mesh: MeshData, s: int):
    """(positions, normals, indices [T,3]) of submesh s as the pass's mesh input.

    Indexed: indices[index_offset : index_offset+n] + base_vertex with n = num_indices, where
    0xFFFFFFFF means every index from index_offset on (mesh.cpp:84-87), clamped to the index
    data and cut to whole triangles as the bottom-level build counts them
    (graphics_manager.cpp:634).  Without index data num_indices counts vertices from
    base_vertex (mesh.hpp:19)."""
    sd = mesh.submeshes[s]
    if int(sd["topology"]) != TOPOLOGY_TRIANGLE_LIST:
        raise AssetError(f"submesh {s}: topology {int(sd['topology'])} is not a triangle list")
    n = int(sd["num_indices"])
    base = int(sd["base_vertex"])
    if len(mesh.indices):
        off = int(sd["index_offset"])
        if off > len(mesh.indices):
            raise AssetError(f"submesh {s}: index offset {off} beyond {len(mesh.indices)} indices")
        n = min(n, len(mesh.indices) - off) // 3 * 3
        idx = mesh.indices[off:off + n].astype(np.int64) + base
    else:
        if n == 0xFFFFFFFF:
            n = len(mesh.positions) - base
        n = n // 3 * 3
        idx = np.arange(base, base + n, dtype=np.int64)
    nv = len(mesh.positions)
    if n and (idx.min() < 0 or idx.max() >= nv):
        raise AssetError(f"submesh {s}: vertex index beyond {nv} vertices")
    if len(mesh.normals) != nv:
        raise AssetError(f"submesh {s}: the pass needs one normal per vertex ({len(mesh.normals)} for {nv})")
    return (np.ascontiguousarray(mesh.positions, np.float32), np.ascontiguousarray(mesh.normals, np.float32),
            idx.astype(np.uint32).reshape(-1, 3))


def submesh_attributes(mesh: MeshData):
    """(texcoords [V,2] or None, tangents [V,4] or None): the asset's per-vertex texture
    coordinates and tangents when it has one per vertex (texcoords_buffer, tangents_buffer)."""
    nv = len(mesh.positions)
    uv = np.ascontiguousarray(mesh.texcoords, np.float32) if len(mesh.texcoords) == nv and nv else None
    tan = np.ascontiguousarray(mesh.tangents, np.float32) if len(mesh.tangents) == nv and nv else None
    return uv, tan


# ---------------------------------------------------------------- textures

TEXTURE_TYPE = "Texture"
# rhi::ResourceFormat values of the formats the pass samples (include/bisemutum/rhi/defines.hpp:44-)
FORMAT_RGBA8_UNORM, FORMAT_RGBA8_SRGB, FORMAT_RGBA32_SFLOAT = 37, 43, 109
FORMAT_RG32_SFLOAT = 103
_DIM_2D, _DIM_3D = 1, 2
_FILTER_LINEAR = 1
_ADDRESS = {0: 0, 1: 1, 2: 2}   # repeat, mirrored repeat, clamp to edge (clamp to border / mirror clamp: no)


@dataclasses.dataclass
class TextureData:
    width: int
    height: int
    levels: int
    format: int          # rhi::ResourceFormat
    sampler: dict        # mag/min filter, mipmap mode, address modes, lod range
    texels: np.ndarray   # level 0: [H, W, 4] uint8 or float32


def read_texture(data: bytes, what: str = "texture") -> TextureData:
    """Parse a Texture asset (TextureAsset::load, src/scene_basic/texture.cpp:92-136):
    magic, type name, version, rhi::SamplerDesc (28 bytes: 9 one-byte enums and flags,
    then anisotropy, lod bias, lod min, lod max), rhi::TextureDesc (20 bytes: extent
    width/height/layers, levels, one-byte format, dimension, usages), then level 0's texel
    bytes as a vector, zlib-compressed (version 2) or raw (version 1, storage type 0;
    storage type 1, PNG layers, is not read)."""
    r = _Reader(data, what)
    magic = r.u32()
    type_name = r.string()
    version = r.u32()
    if magic != ASSET_MAGIC or type_name != TEXTURE_TYPE:
        raise AssetError(f"{what}: not a Texture asset (magic {magic:#010x}, type {type_name!r})")
    sb = bytes(r.take(28))
    sampler = {"mag_filter": sb[0], "min_filter": sb[1], "mipmap_mode": sb[2], "address_u": sb[3],
               "address_v": sb[4], "address_w": sb[5], "border_color": sb[6], "compare_enabled": sb[7],
               "compare_op": sb[8]}
    sampler.update(zip(("anisotropy", "lod_bias", "lod_min", "lod_max"), struct.unpack("<4f", sb[12:28])))
    width, height, layers, levels = struct.unpack("<4I", bytes(r.take(16)))
    fmt, dim, _usages, _pad = bytes(r.take(4))
    if version == 1:
        storage = r.u32()
        if storage != 0:
            raise AssetError(f"{what}: texture storage type {storage} (PNG layers) is not read")
        body = r
    elif version == 2:
        length = r.u64()
        compressed = r.u64()
        raw = zlib.decompress(bytes(r.take(compressed)))
        if len(raw) != length:
            raise AssetError(f"{what}: compressed part holds {len(raw)} bytes, header says {length}")
        body = _Reader(raw, what)
    else:
        raise AssetError(f"{what}: unknown Texture version {version}")
    if dim != 1 or layers != 1:
        raise AssetError(f"{what}: only 2D textures of one layer are read (dimension {dim}, {layers} layers)")
    if width == 0 or height == 0 or width > 16384 or height > 16384:
        raise AssetError(f"{what}: texture size {width}x{height}")
    if fmt in (FORMAT_RGBA8_UNORM, FORMAT_RGBA8_SRGB):
        texels = body.vector("u1")
        if len(texels) < 4 * width * height:
            raise AssetError(f"{what}: {len(texels)} texel bytes for {width}x{height} RGBA8")
        texels = texels[:4 * width * height].reshape(height, width, 4)
    elif fmt == FORMAT_RGBA32_SFLOAT:
        raw = body.vector("u1")
        if len(raw) < 16 * width * height:
            raise AssetError(f"{what}: {len(raw)} texel bytes for {width}x{height} RGBA32F")
        texels = np.frombuffer(raw[:16 * width * height].tobytes(), "<f4").reshape(height, width, 4).copy()
    else:
        raise AssetError(f"{what}: texture format {fmt} is not RGBA8 or RGBA32F")
    return TextureData(width, height, levels, fmt, sampler, texels)


def read_texture3d(data: bytes, what: str = "texture") -> np.ndarray:
    """Level 0 of a 3D Texture asset (the engine's LTC tables, lights.cpp:26-37): float32
    texels [depth, height, width, channels] of an RGBA32F or RG32F texture (the header as
    read_texture; extent.depth_or_layers is the depth)."""
    r = _Reader(data, what)
    magic = r.u32()
    type_name = r.string()
    version = r.u32()
    if magic != ASSET_MAGIC or type_name != TEXTURE_TYPE:
        raise AssetError(f"{what}: not a Texture asset (magic {magic:#010x}, type {type_name!r})")
    r.take(28)
    width, height, depth, _levels = struct.unpack("<4I", bytes(r.take(16)))
    fmt, dim, _usages, _pad = bytes(r.take(4))
    if dim != _DIM_3D:
        raise AssetError(f"{what}: not a 3D texture (dimension {dim})")
    if version == 1:
        if r.u32() != 0:
            raise AssetError(f"{what}: texture storage type is not raw")
        body = r
    elif version == 2:
        length = r.u64()
        compressed = r.u64()
        raw = zlib.decompress(bytes(r.take(compressed)))
        if len(raw) != length:
            raise AssetError(f"{what}: compressed part holds {len(raw)} bytes, header says {length}")
        body = _Reader(raw, what)
    else:
        raise AssetError(f"{what}: unknown Texture version {version}")
    channels = {FORMAT_RGBA32_SFLOAT: 4, FORMAT_RG32_SFLOAT: 2}.get(fmt)
    if channels is None:
        raise AssetError(f"{what}: 3D texture format {fmt} is not RGBA32F or RG32F")
    n = width * height * depth * channels
    raw = body.vector("u1")
    if len(raw) < 4 * n or n == 0:
        raise AssetError(f"{what}: {len(raw)} texel bytes for {width}x{height}x{depth}x{channels} float32")
    return np.frombuffer(raw[:4 * n].tobytes(), "<f4").reshape(depth, height, width, channels).copy()


def read_ltc_tables(asset_dir: str):
    """The engine's LTC tables (ltc_matrix_lut0..2 and ltc_norm_lut texture assets in
    asset_dir), as bi_pt_set_ltc_luts takes them: three [4096, 4] and one [4096, 2]."""
    out = []
    for name in ("ltc_matrix_lut0", "ltc_matrix_lut1", "ltc_matrix_lut2", "ltc_norm_lut"):
        with open(os.path.join(asset_dir, name + ".texture.biasset"), "rb") as f:
            t = read_texture3d(f.read(), name)
        if t.shape[:3] != (64, 8, 8):
            raise AssetError(f"{name}: LTC tables are 8 x 8 x 64 texels, not {t.shape[2]}x{t.shape[1]}x{t.shape[0]}")
        out.append(t.reshape(4096, t.shape[3]))
    return tuple(out)


def encode_texture(t: TextureData) -> bytes:
    """A Texture asset (version 2, as TextureAsset::save writes) of level 0 of t."""
    name = TEXTURE_TYPE.encode()
    sm = t.sampler
    sb = bytes([sm["mag_filter"], sm["min_filter"], sm["mipmap_mode"], sm["address_u"], sm["address_v"],
                sm["address_w"], sm["border_color"], sm["compare_enabled"], sm["compare_op"], 0, 0, 0])
    sb += struct.pack("<4f", sm["anisotropy"], sm["lod_bias"], sm["lod_min"], sm["lod_max"])
    td = struct.pack("<4I", t.width, t.height, 1, t.levels) + bytes([t.format, 1, 7, 0])
    raw = np.ascontiguousarray(t.texels).tobytes()
    body = struct.pack("<Q", len(raw)) + raw
    z = zlib.compress(body)
    return struct.pack("<IQ", ASSET_MAGIC, len(name)) + name + struct.pack("<I", 2) + sb + td + \
        struct.pack("<QQ", len(body), len(z)) + z


def texture_record(t: TextureData) -> "scenes.Texture":
    """The pass's texture of an asset: level 0 (the reference's ray-tracing materials sample
    level 0), the sampler's magnification filter and u / v addressing."""
    if t.sampler["address_u"] not in _ADDRESS or t.sampler["address_v"] not in _ADDRESS:
        raise AssetError("texture addressing clamp-to-border / mirror-clamp-to-edge is not supported")
    fmt = {FORMAT_RGBA8_UNORM: scenes.TEXEL_RGBA8_UNORM, FORMAT_RGBA8_SRGB: scenes.TEXEL_RGBA8_SRGB,
           FORMAT_RGBA32_SFLOAT: scenes.TEXEL_RGBA32F}[t.format]
    return scenes.Texture(np.ascontiguousarray(t.texels), fmt, 1 if t.sampler["mag_filter"] == _FILTER_LINEAR else 0,
                          _ADDRESS[t.sampler["address_u"]], _ADDRESS[t.sampler["address_v"]])


# ---------------------------------------------------------------- transforms

def euler_zxy_degrees(rot_deg) -> np.ndarray:
    """R = Rz(z) Rx(x) Ry(y) from Euler angles in degrees, right-handed rotations."""
    x, y, z = (math.radians(float(a)) for a in rot_deg)
    cx, sx, cy, sy, cz, sz = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ rx @ ry


def transform_matrix(value: dict) -> np.ndarray:
    """4x4 float64 matrix [R S | t] of a Transform component value (missing keys: identity)."""
    t = np.asarray(value.get("translation", (0.0, 0.0, 0.0)), np.float64)
    s = np.asarray(value.get("scaling", (1.0, 1.0, 1.0)), np.float64)
    m = np.eye(4)
    m[:3, :3] = euler_zxy_degrees(value.get("rotation", (0.0, 0.0, 0.0))) * s[None, :]
    m[:3, 3] = t
    return m


def _rotation_of(m: np.ndarray) -> np.ndarray:
    """Rotation part of a TRS matrix: columns divided by their lengths (Transform::from_matrix)."""
    r = m[:3, :3].copy()
    for c in range(3):
        n = np.linalg.norm(r[:, c])
        if n != 0.0:
            r[:, c] /= n
    return r


# ---------------------------------------------------------------- materials

SURFACE_FIELDS = {"base_color": 3, "f0_color": 3, "f90_color": 3, "roughness": 1, "anisotropy": 1,
                  "ior": 1, "opacity": 1, "two_sided": 1, "emission": 3, "normal_map_value": 3}


def surface_default() -> dict:
    """surface_data_default() (shaders/core/material/utils.hlsl:18-31)."""
    return {"emission": (0.0, 0.0, 0.0), "base_color": (0.5, 0.5, 0.5), "f0_color": (0.04, 0.04, 0.04),
            "f90_color": (1.0, 1.0, 1.0), "normal_map_value": (0.5, 0.5, 1.0), "roughness": 0.5,
            "anisotropy": 0.0, "ior": 1.5, "opacity": 1.0, "two_sided": False}


_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?f?"
_STMT = re.compile(r"^surface\.(\w+)\s*=\s*(.+?)\s*$")
_VEC = re.compile(r"^float3\s*\(\s*(" + _NUM + r")\s*,\s*(" + _NUM + r")\s*,\s*(" + _NUM + r")\s*\)$")


_SAMPLE = r"PARAM_(\w+)\s*\.\s*Sample\s*\(\s*PARAM_\w+\s*,\s*vertex\.texcoord\s*\)"
_SWZ = r"(xyz|rgb|[xyzwrgba])"
_TEX_DECL = re.compile(r"^float4\s+(\w+)\s*=\s*" + _SAMPLE + r"$")
_TEX_USE = re.compile(r"^(?:" + _SAMPLE + r"|(\w+))\s*\.\s*" + _SWZ + r"$")
_TEX_STEP = re.compile(r"^(?:" + _SAMPLE + r"|(\w+))\s*\.\s*" + _SWZ + r"\s*<\s*(" + r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?f?" +
                       r")\s*\?\s*(.+?)\s*:\s*(.+)$")
_CHANNEL = {"x": 0, "r": 0, "y": 1, "g": 1, "z": 2, "b": 2, "w": 3, "a": 3}
_CELL = r"int\s*\(\s*floor\s*\(\s*vertex\.position_world\.([xz])\s*\)\s*\)"
_DECL = re.compile(r"^int\s+(\w+)\s*=\s*" + _CELL + r"\s*\^\s*" + _CELL + r"$")
_SEL = re.compile(r"^\(\s*(\w+)\s*&\s*1\s*\)\s*==\s*([01])\s*\?\s*(.+?)\s*:\s*(.+)$")


def _operand(rhs: str):
    if re.fullmatch(r"PARAM_\w+", rhs):
        return ("param", rhs[len("PARAM_"):])
    if rhs in ("true", "false"):
        return ("const", rhs == "true")
    if re.fullmatch(_NUM, rhs):
        return ("const", float(rhs.rstrip("f")))
    m = _VEC.match(rhs)
    if m:
        return ("const", tuple(float(g.rstrip("f")) for g in m.groups()))
    return None


def parse_material_function(text: str) -> list:
    """Statements of a parameter material: [(field, rhs), ...].

    rhs is ("param", name) or ("const", value) for `surface.<field> = PARAM_<name>;`, a
    number, `true`/`false` or `float3(a, b, c)`.  The world checkerboard of the example
    project is read too: `int g = int(floor(vertex.position_world.x)) ^
    int(floor(vertex.position_world.z));` gives ("$g", ("checker_xz",)), and
    `surface.<field> = (g & 1) == 1 ? A : B;` gives (field, ("checker", "g", A, B)) with
    A for odd cells ("== 0" swaps them).  Textures sampled at vertex.texcoord:
    `surface.<field> = PARAM_<t>.Sample(PARAM_<t>_sampler, vertex.texcoord).<swizzle>;`, or
    through `float4 v = PARAM_<t>.Sample(...);` then `surface.<field> = v.<swizzle>;`, give
    (field, ("texture", t, swizzle)); `surface.<field> = v.<c> < a ? b : c;` with constants
    gives (field, ("step", t, c, a, b, c)).  Anything else raises UnsupportedMaterial
    naming the reason."""
    out = []
    cells = set()
    samples = {}   # float4 variable -> texture parameter
    for raw in text.split(";"):
        stmt = " ".join(raw.split())
        if not stmt:
            continue
        td = _TEX_DECL.match(stmt)
        if td:
            samples[td.group(1)] = td.group(2)
            out.append(("$" + td.group(1), ("sample", td.group(2))))
            continue
        d = _DECL.match(stmt)
        if d and {d.group(2), d.group(3)} == {"x", "z"}:
            cells.add(d.group(1))
            out.append(("$" + d.group(1), ("checker_xz",)))
            continue
        m = _STMT.match(stmt)
        if not m:
            raise UnsupportedMaterial(f"statement outside the parameter subset: {stmt!r}")
        field, rhs = m.group(1), m.group(2)
        if field not in SURFACE_FIELDS:
            raise UnsupportedMaterial(f"unknown surface field {field!r}")
        op = _operand(rhs)
        if op is not None:
            out.append((field, op))
            continue
        tu = _TEX_USE.match(rhs)
        if tu and (tu.group(1) or tu.group(2) in samples):
            out.append((field, ("texture", tu.group(1) or samples[tu.group(2)], tu.group(3))))
            continue
        ts = _TEX_STEP.match(rhs)
        if ts and (ts.group(1) or ts.group(2) in samples) and len(ts.group(3)) == 1:
            below, above = _operand(ts.group(5)), _operand(ts.group(6))
            if below is not None and above is not None and below[0] == above[0] == "const":
                out.append((field, ("step", ts.group(1) or samples[ts.group(2)], ts.group(3),
                                    float(ts.group(4).rstrip("f")), float(below[1]), float(above[1]))))
                continue
        if ".Sample" in stmt or "Texture" in stmt:
            raise UnsupportedMaterial(f"texture sampling outside the supported forms: {stmt!r}")
        sel = _SEL.match(rhs)
        if sel and sel.group(1) in cells:
            a, b = _operand(sel.group(3)), _operand(sel.group(4))
            if a is not None and b is not None:
                odd, even = (a, b) if sel.group(2) == "1" else (b, a)
                out.append((field, ("checker", sel.group(1), odd, even)))
                continue
        raise UnsupportedMaterial(f"expression outside the parameter subset: {stmt!r}")
    return out


def _as_field(field: str, v):
    w = SURFACE_FIELDS[field]
    if field == "two_sided":
        if not isinstance(v, bool):
            raise UnsupportedMaterial(f"two_sided needs a bool, got {v!r}")
        return v
    if isinstance(v, dict):
        raise UnsupportedMaterial(f"{field} from a texture parameter")
    if w == 1:
        if isinstance(v, (list, tuple)):
            raise UnsupportedMaterial(f"{field} needs a scalar, got {v!r}")
        return float(v)
    if isinstance(v, (int, float)):
        return (float(v),) * 3          # HLSL broadcasts a scalar to float3
    if len(v) != 3:
        raise UnsupportedMaterial(f"{field} needs 3 components, got {v!r}")
    return tuple(float(c) for c in v)


def evaluate_material(desc: dict) -> dict:
    """Surface of a parameter material: the defaults with the function's assignments applied.

    desc: the material's TOML table (blend_mode, surface_model, params, material_function),
    or the same with a parsed `statements` list in place of material_function."""
    if desc.get("surface_model", "lit") != "lit":
        raise UnsupportedMaterial(f"surface model {desc.get('surface_model')!r}")
    if "outside_subset" in desc:
        raise UnsupportedMaterial(desc["outside_subset"])
    stmts = desc["statements"] if "statements" in desc else parse_material_function(desc.get("material_function", ""))
    params = {p["name"]: p.get("value") for p in desc.get("params", [])}

    def value(field, rhs):
        kind, v = rhs[0], rhs[1]
        if kind == "param":
            if v not in params:
                raise UnsupportedMaterial(f"parameter {v!r} is not declared")
            v = params[v]
        return _as_field(field, tuple(v) if isinstance(v, list) else v)

    odd, even = surface_default(), surface_default()
    cells, pattern = set(), False
    textured = {}

    def texture_asset(name):
        v = params.get(name)
        if not isinstance(v, dict) or "asset_id" not in v:
            raise UnsupportedMaterial(f"parameter {name!r} is not a texture")
        return int(v["asset_id"])

    for field, rhs in stmts:
        if field.startswith("$"):
            if rhs[0] == "checker_xz":
                cells.add(field[1:])
        elif rhs[0] == "texture":
            tex, swz = texture_asset(rhs[1]), rhs[2]
            if field in ("base_color", "f0_color", "normal_map_value") and swz in ("xyz", "rgb"):
                textured[field] = (tex,)
            elif field in ("roughness", "opacity") and len(swz) == 1:
                textured[field] = (tex, _CHANNEL[swz])
            else:
                raise UnsupportedMaterial(f"{field} from texture channels {swz!r}")
        elif rhs[0] == "step":
            if field != "opacity":
                raise UnsupportedMaterial(f"a stepped texture channel for {field}")
            textured[field] = (texture_asset(rhs[1]), _CHANNEL[rhs[2]], rhs[3], rhs[4], rhs[5])
        elif rhs[0] == "checker":
            if rhs[1] not in cells:
                raise UnsupportedMaterial(f"checker variable {rhs[1]!r} is not declared")
            odd[field], even[field] = value(field, rhs[2]), value(field, rhs[3])
            pattern = True
        else:
            odd[field] = even[field] = value(field, rhs)
    blend = desc.get("blend_mode", "opaque")
    if blend not in ("opaque", "alpha_test", "translucent", "additive", "modulate"):
        raise UnsupportedMaterial(f"blend mode {blend!r}")
    if pattern and textured:
        raise UnsupportedMaterial("a checkerboard of textured fields")
    for surf in (odd, even):
        if tuple(surf["normal_map_value"]) != (0.5, 0.5, 1.0) and "normal_map_value" not in textured:
            raise UnsupportedMaterial("constant normal map value other than (0.5, 0.5, 1)")
        # the ray-tracing any-hit: alpha test for alpha_test, stochastic opacity for the
        # other non-opaque modes (hits/rt_gbuffer_hit.hlsl:20-33)
        surf["blend"] = {"opaque": 0, "alpha_test": 1}.get(blend, 2)
        surf["textures"] = dict(textured)
    odd["pattern_even"] = even if pattern else None
    return odd


PATTERN_CHECKER_XZ = 1   # bi_pt_material.pattern: odd cells this record, even cells pattern_material


def material_record(surf: dict, tex_index=None) -> np.ndarray:
    """bi_pt_material record (32 floats) of an evaluated surface; emission is not used by the pass.
    A checkerboard surface becomes two records: this one (odd cells, pattern set) followed by
    the even cells' record.  tex_index maps a texture asset id to its index in the scene's
    textures (references are stored as index + 1)."""
    rec = np.zeros(scenes.MATERIAL_FLOATS, np.float32)
    w = rec.view(np.uint32)
    rec[0:3] = surf["base_color"]
    rec[3] = surf["roughness"]
    rec[4:7] = surf["f0_color"]
    rec[7] = surf["anisotropy"]
    rec[8:11] = surf["f90_color"]
    rec[11] = surf["ior"]
    w[12] = 1 if surf["two_sided"] else 0
    w[scenes.M_BLEND] = surf.get("blend", 0)
    rec[scenes.M_OPACITY] = surf["opacity"]
    for field, t in surf.get("textures", {}).items():
        ref = tex_index[t[0]] + 1
        if field == "base_color":
            w[scenes.M_BASE_TEX] = ref
        elif field == "f0_color":
            w[scenes.M_F0_TEX] = ref
        elif field == "normal_map_value":
            w[scenes.M_NORMAL_TEX] = ref
        elif field == "roughness":
            w[scenes.M_ROUGH_TEX], w[scenes.M_ROUGH_CHANNEL] = ref, t[1]
        elif field == "opacity":
            w[scenes.M_OPACITY_TEX], w[scenes.M_OPACITY_CHANNEL] = ref, t[1]
            if len(t) == 5:
                w[scenes.M_OPACITY_STEP] = 1
                rec[scenes.M_STEP_AT], rec[scenes.M_STEP_BELOW], rec[scenes.M_STEP_ABOVE] = t[2], t[3], t[4]
    return rec


# ---------------------------------------------------------------- scenes

@dataclasses.dataclass
class ProjectScene:
    scene: scenes.Scene
    camera: scenes.Camera
    settings: dict                 # max_bounces, ray_length, accumulate of the path-tracing settings
    unsupported: dict              # material asset id -> reason (materials replaced by the default surface)
    objects: list                  # (object name, mesh asset id, submesh, material asset id) per instance


def _components(obj: dict) -> dict:
    out = {}
    for c in obj.get("components", []):
        out.setdefault(c["type"], c.get("value", {}))
    return out


def scene_from_description(scene_desc: dict, fetch, name="project", width=None, height=None,
                           on_unsupported="raise") -> ProjectScene:
    """Build the pass's scene and camera from a parsed scene file.

    fetch(asset_id) returns a StaticMesh's or Texture's bytes or a material's table.  on_unsupported:
    "raise", or "default" to replace such a material by surface_data_default() and list it."""
    if on_unsupported not in ("raise", "default"):
        raise ValueError("on_unsupported must be 'raise' or 'default'")
    objects = scene_desc.get("objects", [])
    comps = [_components(o) for o in objects]
    parents = [int(o.get("parent", -1)) for o in objects]
    world: list = [None] * len(objects)

    def world_of(i, depth=0):
        if depth > len(objects):
            raise AssetError("cycle in the object hierarchy")
        if world[i] is None:
            local = transform_matrix(comps[i].get("Transform", {}))
            p = parents[i]
            world[i] = world_of(p, depth + 1) @ local if 0 <= p < len(objects) else local
        return world[i]

    meshes, mats, inst, placed = [], [], [], []
    mesh_uv, mesh_tan = [], []
    textures, tex_index = [], {}   # pass textures; texture asset id -> index
    parsed = {}        # mesh asset id -> MeshData
    mesh_index = {}    # (mesh asset id, submesh) -> index into meshes
    mat_index = {}     # material asset id -> index into mats
    unsupported: dict = {}
    lights = []
    camera = None
    settings = {"max_bounces": 3, "ray_length": 100.0, "accumulate": True}
    for i, (obj, c) in enumerate(zip(objects, comps)):
        if "CameraComponent" in c and camera is None:
            cv = c["CameraComponent"]
            if cv.get("projection_type", "perspective") != "perspective":
                raise AssetError("only perspective cameras are supported")
            w = world_of(i)
            eye = w[:3, 3]
            front = w[:3, :3] @ np.array([0.0, 0.0, -1.0])
            up = w[:3, :3] @ np.array([0.0, 1.0, 0.0])
            size = cv.get("render_target_size", (1, 1))
            cw, ch = int(width or size[0]), int(height or size[1])
            camera = scenes.make_camera(eye, eye + front, cw, ch, float(cv.get("yfov", 45.0)), up=up,
                                        near=float(cv.get("near_z", 0.1)), far=float(cv.get("far_z", 1000.0)))
        if "DirectionalLightComponent" in c:
            lv = c["DirectionalLightComponent"]
            emission = np.asarray(lv.get("color", (1.0, 1.0, 1.0)), np.float64) * float(lv.get("strength", 1.0))
            if np.any(emission != 0.0):
                d = _rotation_of(world_of(i)) @ np.array([0.0, 1.0, 0.0])
                shadow_strength = float(lv.get("shadow_strength", 1.0))
                shadow = bool(lv.get("cast_shadow", True)) and shadow_strength > 0.0 and \
                    float(lv.get("shadow_range", 100.0)) > 0.0
                lights.append(scenes.dir_light(emission, d, cast_shadow=shadow,
                                               shadow_strength=min(shadow_strength, 1.0)))
        if "BasicRendererOverrideVolume" in c:
            pt = c["BasicRendererOverrideVolume"].get("settings", {}).get("path_tracing", {})
            for k in settings:
                if k in pt:
                    settings[k] = type(settings[k])(pt[k])
        if "StaticMeshComponent" in c and "MeshRendererComponent" in c:
            mesh_id = int(c["StaticMeshComponent"]["static_mesh"]["asset_id"])
            r = c["MeshRendererComponent"]
            start = int(r.get("submesh_start_index", 0))
            if mesh_id not in parsed:
                parsed[mesh_id] = read_static_mesh(fetch(mesh_id), f"asset {mesh_id}")
            md = parsed[mesh_id]
            count = len(md.submeshes) - start
            mat_ids = [int(m["asset_id"]) for m in r.get("materials", [])]
            if count > len(mat_ids):
                raise AssetError(f"object {obj.get('name')!r}: {count} submeshes, {len(mat_ids)} materials")
            xform = world_of(i)[:3, :].reshape(-1).astype(np.float32)
            for k in range(count):
                sub = start + k
                key = (mesh_id, sub)
                if key not in mesh_index:
                    mesh_index[key] = len(meshes)
                    meshes.append(submesh_triangles(md, sub))
                    uv, tan = submesh_attributes(md)
                    mesh_uv.append(uv)
                    mesh_tan.append(tan)
                mid = mat_ids[k]
                if mid not in mat_index:
                    try:
                        surf = evaluate_material(fetch(mid))
                    except UnsupportedMaterial as e:
                        if on_unsupported == "raise":
                            raise UnsupportedMaterial(f"material asset {mid}: {e}") from None
                        unsupported[mid] = str(e)
                        surf = surface_default()
                    for t in surf.get("textures", {}).values():
                        if t[0] not in tex_index:
                            tex_index[t[0]] = len(textures)
                            textures.append(texture_record(read_texture(fetch(t[0]), f"asset {t[0]}")))
                    mat_index[mid] = len(mats)
                    rec = material_record(surf, tex_index)
                    if surf.get("pattern_even") is not None:
                        rec.view(np.uint32)[13] = PATTERN_CHECKER_XZ
                        rec.view(np.uint32)[14] = len(mats) + 1
                        mats.append(rec)
                        rec = material_record(surf["pattern_even"])
                    mats.append(rec)
                inst.append((xform, mesh_index[key], mat_index[mid]))
                placed.append((obj.get("name", f"object {i}"), mesh_id, sub, mid))
    if camera is None:
        raise AssetError("the scene has no camera")
    if not inst:
        raise AssetError("the scene has no static meshes")
    sky = np.zeros((1, 1, 4), np.float32)
    any_uv = textures and any(u is not None for u in mesh_uv)
    any_tan = textures and any(t is not None for t in mesh_tan)
    sc = scenes.Scene(name, meshes, np.array(inst, scenes.INSTANCE_DTYPE), np.stack(mats),
                      np.stack(lights) if lights else np.zeros((0, 16), np.float32), sky,
                      np.zeros(3, np.float32), np.eye(3, dtype=np.float32).reshape(-1),
                      mesh_uv=mesh_uv if any_uv else None, mesh_tan=mesh_tan if any_tan else None,
                      textures=textures)
    return ProjectScene(sc, camera, settings, unsupported, placed)


def _project_path(root: str, virtual: str) -> str:
    if virtual.startswith("/project/"):
        virtual = virtual[len("/project/"):]
    return os.path.join(root, virtual.lstrip("/"))


def load_project(root: str, width=None, height=None, on_unsupported="raise") -> ProjectScene:
    """Load a scene project directory (project.toml, asset metadata, scene, assets)."""
    import tomli

    def toml(path):
        with open(path, "rb") as f:
            return tomli.load(f)

    project = toml(os.path.join(root, "project.toml"))
    meta = toml(_project_path(root, project["asset_metadata_file"]))
    by_id = {int(a["id"]): a for a in meta.get("assets", [])}
    scene_desc = toml(_project_path(root, project["scene_file"]))

    def fetch(asset_id):
        a = by_id.get(asset_id)
        if a is None:
            raise AssetError(f"asset {asset_id} is not in the asset metadata")
        path = _project_path(root, a["path"])
        if a["type"] == "StaticMesh":
            with open(path, "rb") as f:
                return f.read()
        if a["type"] == "Material":
            return toml(path)
        if a["type"] == "Texture":
            with open(path, "rb") as f:
                return f.read()
        raise AssetError(f"asset {asset_id} has type {a['type']!r}, not a mesh, material or texture")

    return scene_from_description(scene_desc, fetch, project.get("name", "project"), width, height, on_unsupported)
