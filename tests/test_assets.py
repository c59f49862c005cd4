"""Scene ingestion: StaticMesh assets, transforms, parameter materials and the example
project (fixtures under tests/golden/scene_basic, written by make_scene_basic.py)."""
import json
import math
import os

import numpy as np
import pytest

import assets

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scene_basic")
REF_PROJECT = "/root/reference/examples/scene_basic"


def mesh_bytes(name):
    with open(os.path.join(GOLD, "meshes", f"{name}.static_mesh.biasset"), "rb") as f:
        return f.read()


def fixture_scene(width=None, height=None, on_unsupported="default"):
    with open(os.path.join(GOLD, "scene.json")) as f:
        scene = json.load(f)
    with open(os.path.join(GOLD, "assets.json")) as f:
        table = {int(k): v for k, v in json.load(f).items()}
    with open(os.path.join(GOLD, "materials.json")) as f:
        mats = {int(k): v for k, v in json.load(f).items()}

    def fetch(i):
        if table[i]["type"] in ("StaticMesh", "Texture"):
            with open(os.path.join(GOLD, table[i]["file"]), "rb") as f:
                return f.read()
        return mats[i]
    return assets.scene_from_description(scene, fetch, "scene_basic", width, height, on_unsupported)


# ---------------------------------------------------------------- StaticMesh

def winding_agrees(p, n, idx):
    a, b, c = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
    g = np.cross(b - a, c - a)
    return np.all((g * (n[idx[:, 0]] + n[idx[:, 1]] + n[idx[:, 2]])).sum(1) > 0)


def test_cube_asset_is_the_unit_cube():
    m = assets.read_static_mesh(mesh_bytes("cube"))
    assert len(m.positions) == 24 and len(m.indices) == 36 and len(m.submeshes) == 1
    assert len(m.normals) == 24 and len(m.tangents) == 24 and len(m.texcoords) == 24
    assert len(m.colors) == 0 and len(m.texcoords2) == 0
    assert int(m.submeshes[0]["num_indices"]) == 0xFFFFFFFF and int(m.submeshes[0]["topology"]) == 3
    p, n, idx = assets.submesh_triangles(m, 0)
    assert idx.shape == (12, 3)
    assert np.array_equal(np.abs(p), np.ones_like(p))               # corners of [-1,1]^3
    assert np.array_equal(np.abs(n).sum(1), np.ones(24, np.float32))  # axis-aligned face normals
    assert np.array_equal((n * p).sum(1), np.ones(24, np.float32))    # each vertex on its face plane
    assert winding_agrees(p, n, idx)
    # every face is a quad of 4 vertices with one normal, 6 faces
    assert len({tuple(v) for v in n}) == 6


def test_plane_and_sphere_assets():
    pl = assets.read_static_mesh(mesh_bytes("plane"))
    p, n, idx = assets.submesh_triangles(pl, 0)
    assert idx.shape == (2, 3) and np.all(p[:, 1] == 0) and np.all(n == [0, 1, 0])
    assert np.array_equal(np.sort(np.abs(p[:, [0, 2]]).ravel()), np.ones(8, np.float32))
    assert winding_agrees(p, n, idx)
    sp = assets.read_static_mesh(mesh_bytes("sphere"))
    p, n, idx = assets.submesh_triangles(sp, 0)
    assert idx.shape == (960, 3) and len(p) == 559
    r = np.linalg.norm(p.astype(np.float64), axis=1)
    assert np.all(np.abs(r - 1.0) < 1e-5)                          # unit sphere
    assert np.all(np.abs(np.linalg.norm(n.astype(np.float64), axis=1) - 1.0) < 1e-3)
    assert np.all((n * p).sum(1) > 0.99)                           # smooth normals point outwards
    assert winding_agrees(p, n, idx)
    # closed mesh: every edge is shared by exactly two triangles (up to welded seams)
    assert np.all(np.unique(idx) == np.arange(len(p))) or len(np.unique(idx)) <= len(p)


@pytest.mark.parametrize("version", [1, 2])
def test_static_mesh_round_trip(version):
    m = assets.read_static_mesh(mesh_bytes("sphere"))
    data = assets.encode_static_mesh(m, version)
    m2 = assets.read_static_mesh(data)
    for f in ("positions", "normals", "tangents", "colors", "texcoords", "texcoords2", "indices"):
        assert np.array_equal(getattr(m, f), getattr(m2, f)), f
    assert m.submeshes.tobytes() == m2.submeshes.tobytes()
    if version == 2:   # same layout as the asset written by the engine, up to the zlib stream
        assert data[:26] == mesh_bytes("sphere")[:26]


def test_static_mesh_errors():
    good = mesh_bytes("cube")
    with pytest.raises(assets.AssetError, match="not a StaticMesh"):
        assets.read_static_mesh(b"\0\0\0\0" + good[4:])
    with pytest.raises(assets.AssetError, match="truncated"):
        assets.read_static_mesh(good[:20])
    bad_version = bytearray(good)
    bad_version[22:26] = (7).to_bytes(4, "little")
    with pytest.raises(assets.AssetError, match="version 7"):
        assets.read_static_mesh(bytes(bad_version))
    m = assets.read_static_mesh(good)
    m.submeshes = m.submeshes.copy()
    m.submeshes["topology"] = 4   # triangle strip
    with pytest.raises(assets.AssetError, match="triangle list"):
        assets.submesh_triangles(m, 0)


def test_submesh_ranges():
    m = assets.read_static_mesh(mesh_bytes("cube"))
    sub = np.zeros(2, assets.SUBMESH_DTYPE)
    sub[0] = (0, 0, 18, 3, 0)
    sub[1] = (0, 18, 0xFFFFFFFF, 3, 0)
    m.submeshes = sub
    _, _, a = assets.submesh_triangles(m, 0)
    _, _, b = assets.submesh_triangles(m, 1)
    assert np.array_equal(np.concatenate([a, b]).ravel(), m.indices)
    m.indices = np.zeros(0, np.uint32)           # no index data: vertices in order
    sub[0] = (3, 0, 6, 3, 0)
    _, _, c = assets.submesh_triangles(m, 0)
    assert np.array_equal(c.ravel(), np.arange(3, 9))


# ---------------------------------------------------------------- transforms

def test_euler_zxy_single_axes_and_order():
    r = assets.euler_zxy_degrees
    assert np.allclose(r((90, 0, 0)) @ [0, 1, 0], [0, 0, 1])      # about x: y -> z
    assert np.allclose(r((0, 90, 0)) @ [0, 0, 1], [1, 0, 0])      # about y: z -> x
    assert np.allclose(r((0, 0, 90)) @ [1, 0, 0], [0, 1, 0])      # about z: x -> y
    # glm::eulerAngleZXY(t1 = z, t2 = x, t3 = y) in closed form, column by column
    x, y, z = 0.3, -1.1, 0.7
    c1, s1, c2, s2, c3, s3 = math.cos(z), math.sin(z), math.cos(x), math.sin(x), math.cos(y), math.sin(y)
    cols = np.array([[c1 * c3 - s1 * s2 * s3, c3 * s1 + c1 * s2 * s3, -c2 * s3],
                     [-c2 * s1, c1 * c2, s2],
                     [c1 * s3 + c3 * s1 * s2, s1 * s3 - c1 * c3 * s2, c2 * c3]])
    got = r(tuple(math.degrees(a) for a in (x, y, z)))
    assert np.allclose(got, cols.T, atol=1e-12)


def test_transform_matrix_is_trs():
    v = {"translation": [1.0, 2.0, 3.0], "scaling": [2.0, 1.0, 0.5], "rotation": [0.0, 90.0, 0.0]}
    m = assets.transform_matrix(v)
    # scale, then rotate z -> x, then translate
    assert np.allclose(m @ [0, 0, 1, 1], [1.5, 2.0, 3.0, 1.0])
    assert np.allclose(m @ [1, 0, 0, 1], [1.0, 2.0, 1.0, 1.0])
    assert np.allclose(assets.transform_matrix({}), np.eye(4))


# ---------------------------------------------------------------- materials

def test_parse_material_function_subset():
    s = assets.parse_material_function(
        "surface.base_color = PARAM_albedo;\n surface.roughness = 0.25f;\n"
        "surface.two_sided = true;\nsurface.f0_color = float3(0.1, 0.2, 0.3);\n")
    assert s == [("base_color", ("param", "albedo")), ("roughness", ("const", 0.25)),
                 ("two_sided", ("const", True)), ("f0_color", ("const", (0.1, 0.2, 0.3)))]
    for bad, why in (("surface.base_color = PARAM_t.Sample(PARAM_s, vertex.texcoord2).xyz;", "texture"),
                     ("surface.base_color = PARAM_t.Sample(PARAM_s, vertex.texcoord).xyz * 2.0;", "texture"),
                     ("surface.roughness = vertex.position_world.x > 0 ? 1.0 : 0.5;", "expression"),
                     ("float a = 1.0;", "statement"),
                     ("surface.sheen = 1.0;", "unknown surface field")):
        with pytest.raises(assets.UnsupportedMaterial, match=why):
            assets.parse_material_function(bad)


def test_parse_checker_select():
    text = ("int cell = int(floor(vertex.position_world.z)) ^ int(floor(vertex.position_world.x));\n"
            "surface.base_color = (cell & 1) == 0 ? PARAM_even : float3(1, 0, 0);\n"
            "surface.roughness = (cell & 1) == 1 ? 0.25 : PARAM_r;")
    s = assets.parse_material_function(text)
    assert s == [("$cell", ("checker_xz",)),
                 ("base_color", ("checker", "cell", ("const", (1.0, 0.0, 0.0)), ("param", "even"))),
                 ("roughness", ("checker", "cell", ("const", 0.25), ("param", "r")))]
    surf = assets.evaluate_material({"material_function": text, "params": [
        {"name": "even", "value": [0.0, 0.0, 1.0]}, {"name": "r", "value": 0.75}]})
    assert surf["base_color"] == (1.0, 0.0, 0.0) and surf["roughness"] == 0.25
    assert surf["pattern_even"]["base_color"] == (0.0, 0.0, 1.0) and surf["pattern_even"]["roughness"] == 0.75
    for bad in ("int c = int(floor(vertex.position_world.x)) ^ int(floor(vertex.position_world.y));",
                "surface.roughness = (other & 1) == 1 ? 0.5 : 1.0;"):
        with pytest.raises(assets.UnsupportedMaterial):
            assets.parse_material_function(bad)


def test_evaluate_material_defaults_and_params():
    surf = assets.evaluate_material({"blend_mode": "opaque", "material_function": "surface.roughness = PARAM_r;",
                                     "params": [{"name": "r", "value": 0.3}]})
    assert surf["roughness"] == 0.3 and surf["base_color"] == (0.5, 0.5, 0.5) and surf["f0_color"] == (0.04,) * 3
    rec = assets.material_record(surf)
    assert np.allclose(rec[0:12], [0.5, 0.5, 0.5, 0.3, 0.04, 0.04, 0.04, 0.0, 1, 1, 1, 1.5])
    assert rec.view(np.uint32)[12] == 0
    scalar = assets.evaluate_material({"material_function": "surface.base_color = PARAM_c;",
                                       "params": [{"name": "c", "value": 0.25}]})
    assert scalar["base_color"] == (0.25, 0.25, 0.25)                   # HLSL scalar broadcast
    tr = assets.evaluate_material({"blend_mode": "translucent", "material_function": "surface.opacity = 0.5;"})
    rec = assets.material_record(tr)
    assert rec.view(np.uint32)[15] == 2 and rec[16] == np.float32(0.5)          # stochastic opacity
    for mode, word in (("alpha_test", 1), ("additive", 2), ("modulate", 2), ("opaque", 0)):
        assert assets.material_record(assets.evaluate_material({"blend_mode": mode})).view(np.uint32)[15] == word
    with pytest.raises(assets.UnsupportedMaterial, match="not declared"):
        assets.evaluate_material({"material_function": "surface.roughness = PARAM_missing;"})
    with pytest.raises(assets.UnsupportedMaterial, match="texture parameter"):
        assets.evaluate_material({"material_function": "surface.base_color = PARAM_t;",
                                  "params": [{"name": "t", "value": {"asset_id": 6}}]})


def test_textured_material_statements_and_record():
    text = ("float4 value = PARAM_cage_tex.Sample(PARAM_cage_tex_sampler, vertex.texcoord);\n"
            "surface.base_color = value.xyz;\nsurface.f0_color = value.rgb;\n"
            "surface.opacity = value.w < 0.5 ? 0.0 : 1.0;\n"
            "surface.normal_map_value = PARAM_n.Sample(PARAM_n_sampler, vertex.texcoord).xyz;\n"
            "surface.roughness = PARAM_r.Sample(PARAM_r_sampler, vertex.texcoord).g;\nsurface.two_sided = true;")
    s = assets.parse_material_function(text)
    assert s == [("$value", ("sample", "cage_tex")), ("base_color", ("texture", "cage_tex", "xyz")),
                 ("f0_color", ("texture", "cage_tex", "rgb")), ("opacity", ("step", "cage_tex", "w", 0.5, 0.0, 1.0)),
                 ("normal_map_value", ("texture", "n", "xyz")), ("roughness", ("texture", "r", "g")),
                 ("two_sided", ("const", True))]
    surf = assets.evaluate_material({"blend_mode": "alpha_test", "material_function": text, "params": [
        {"name": "cage_tex", "value": {"asset_id": 10}}, {"name": "n", "value": {"asset_id": 7}},
        {"name": "r", "value": {"asset_id": 3}}]})
    rec = assets.material_record(surf, {10: 0, 7: 1, 3: 2})
    w = rec.view(np.uint32)
    assert w[15] == 1 and w[23] == 1 and w[24] == 1 and w[25] == 2 and w[26] == 3 and w[27] == 1
    assert w[17] == 1 and w[18] == 3 and w[19] == 1 and (rec[20], rec[21], rec[22]) == (0.5, 0.0, 1.0)
    with pytest.raises(assets.UnsupportedMaterial, match="texture channels"):
        assets.evaluate_material({"material_function": "surface.f90_color = PARAM_t.Sample(PARAM_s, vertex.texcoord).xyz;",
                                  "params": [{"name": "t", "value": {"asset_id": 1}}]})


def test_texture_asset_round_trip_and_layout():
    rng = np.random.default_rng(3)
    texels = rng.integers(0, 256, size=(5, 7, 4), dtype=np.uint8)
    sampler = {"mag_filter": 1, "min_filter": 0, "mipmap_mode": 1, "address_u": 2, "address_v": 1, "address_w": 0,
               "border_color": 0, "compare_enabled": 0, "compare_op": 0, "anisotropy": 0.0, "lod_bias": 0.0,
               "lod_min": 0.0, "lod_max": 1000.0}
    t = assets.TextureData(7, 5, 3, assets.FORMAT_RGBA8_SRGB, sampler, texels)
    data = assets.encode_texture(t)
    # layout: magic, type name, version 2, 28-byte SamplerDesc, 20-byte TextureDesc
    assert data[:4] == (0x0B1A55E7).to_bytes(4, "little") and data[12:19] == b"Texture"
    off = 4 + 8 + 7 + 4
    assert list(data[off:off + 9]) == [1, 0, 1, 2, 1, 0, 0, 0, 0]
    back = assets.read_texture(data)
    assert (back.width, back.height, back.format) == (7, 5, assets.FORMAT_RGBA8_SRGB)
    assert np.array_equal(back.texels, texels) and back.sampler == sampler
    rec = assets.texture_record(back)
    assert (rec.format, rec.filter, rec.address_u, rec.address_v) == (1, 1, 2, 1)
    with pytest.raises(assets.AssetError):
        assets.read_texture(data[:40])


@pytest.mark.skipif(not os.path.isdir(REF_PROJECT), reason="the engine's example project is not present")
def test_reference_texture_assets():
    """The example project's three textures: level 0 of the sizes and formats their
    descriptions give, linear filtering with repeat addressing, and the fixtures are every
    8th texel of them."""
    want = {"earth_diffuse": (2048, 1024, assets.FORMAT_RGBA8_SRGB), "earth_normal": (2048, 1024, assets.FORMAT_RGBA8_UNORM),
            "cage": (1299, 1300, assets.FORMAT_RGBA8_UNORM)}
    for name, (w, h, fmt) in want.items():
        with open(os.path.join(REF_PROJECT, "textures", f"{name}.texture.biasset"), "rb") as f:
            t = assets.read_texture(f.read(), name)
        assert (t.width, t.height, t.format) == (w, h, fmt) and t.texels.shape == (h, w, 4)
        assert t.sampler["mag_filter"] == 1 and t.sampler["address_u"] == 0 and t.sampler["address_v"] == 0
        with open(os.path.join(GOLD, "textures", f"{name}.texture.biasset"), "rb") as f:
            small = assets.read_texture(f.read())
        assert np.array_equal(small.texels, t.texels[::8, ::8])
    # the normal map's texels encode tangent-space normals (the shader normalises them):
    # |2 t - 1| near 1 for most texels, all pointing away from the surface
    with open(os.path.join(REF_PROJECT, "textures", "earth_normal.texture.biasset"), "rb") as f:
        n = assets.read_texture(f.read()).texels[::37, ::37, :3].astype(np.float64) / 255.0 * 2.0 - 1.0
    assert np.mean(np.abs(np.linalg.norm(n, axis=-1) - 1.0) < 0.1) > 0.9 and np.all(n[..., 2] > 0.5)
    # the cage's alpha is a cut-out mask: almost all texels fully opaque or fully clear
    with open(os.path.join(REF_PROJECT, "textures", "cage.texture.biasset"), "rb") as f:
        a = assets.read_texture(f.read()).texels[..., 3]
    assert np.mean((a == 0) | (a == 255)) > 0.9 and 0.05 < np.mean(a >= 128) < 0.95


# ---------------------------------------------------------------- the example project

def test_scene_basic_from_fixtures():
    ps = fixture_scene()
    sc = ps.scene
    assert [o[0] for o in ps.objects] == ["Sphere", "Cube", "Transparent Cube", "Cage Cube", "Plane"]
    assert len(sc.meshes) == 3 and sc.num_tris == 960 + 3 * 12 + 2
    assert ps.unsupported == {}
    assert [t.texels.shape for t in sc.textures] == [(128, 256, 4), (128, 256, 4), (163, 163, 4)]
    assert [t.format for t in sc.textures] == [1, 0, 0]          # earth colour is sRGB
    m_earth, m_tr, m_cage = (sc.materials[sc.instances["material"][k]] for k in (0, 2, 3))
    assert m_earth.view(np.uint32)[23] == 1 and m_earth.view(np.uint32)[25] == 2   # base colour, normal map
    assert m_tr.view(np.uint32)[15] == 2 and m_tr[16] == np.float32(0.5)         # translucent, opacity 0.5
    assert m_cage.view(np.uint32)[15] == 1 and m_cage.view(np.uint32)[17] == 3   # alpha test from the cage's w
    assert sc.mesh_uv is not None and all(u is not None for u in sc.mesh_uv)
    assert sc.mesh_tan is not None
    assert ps.settings == {"max_bounces": 3, "ray_length": 100.0, "accumulate": True}
    white = sc.materials[sc.instances["material"][1]]
    assert np.array_equal(white[:3], [1, 1, 1]) and white[3] == np.float32(0.5)
    # checkerboard floor: odd cells 0.8 grey / roughness 1, even cells 0.1 / 0.1
    k = int(sc.instances["material"][4])
    odd, even = sc.materials[k], sc.materials[sc.materials[k].view(np.uint32)[14]]
    assert odd.view(np.uint32)[13] == assets.PATTERN_CHECKER_XZ and even.view(np.uint32)[13] == 0
    assert np.allclose(odd[:4], [0.8, 0.8, 0.8, 1.0]) and np.allclose(even[:4], [0.1, 0.1, 0.1, 0.1])
    assert np.allclose(odd[4:7], 0.04) and np.allclose(even[4:7], 0.04)
    # plane: scaled 5 x 1 x 5 and lowered by one
    plane = sc.instances[4]["xform"].reshape(3, 4)
    assert np.allclose(plane, [[5, 0, 0, 0], [0, 1, 0, -1], [0, 0, 5, 0]])
    # light: colour * strength, direction R(10, 0, -30) (0, 1, 0)
    li = sc.dir_lights[0]
    assert np.allclose(li[0:3], [4.0, 3.6, 3.2])
    d = assets.euler_zxy_degrees((10.000005722045898, 0.0, -30.000001907348633)) @ [0, 1, 0]
    assert np.allclose(li[4:7], d, atol=1e-6) and li.view(np.int32)[3] == 0 and li[7] == 1.0
    # camera at its translation, looking down its -z
    cam = ps.camera
    assert (cam.width, cam.height) == (1298, 635)
    assert np.allclose(cam.inv_view[:3, 3], [10.5, 4.5, 6.0], atol=1e-5)
    front = assets.euler_zxy_degrees((-9.3324127197265625, 63.346549987792969, 17.90446662902832)) @ [0, 0, -1]
    assert np.allclose(cam.inv_view[:3, :3] @ [0, 0, -1], front, atol=1e-5)
    fixture_scene(on_unsupported="raise")   # every material of the project is in the subset


@pytest.mark.skipif(not os.path.isdir(REF_PROJECT), reason="the engine's example project is not present")
def test_load_project_matches_fixtures():
    a = assets.load_project(REF_PROJECT, 130, 64, on_unsupported="default")
    b = fixture_scene(130, 64)
    assert a.objects == b.objects and a.settings == b.settings and set(a.unsupported) == set(b.unsupported)
    assert a.scene.instances.tobytes() == b.scene.instances.tobytes()
    assert a.scene.materials.tobytes() == b.scene.materials.tobytes()
    assert a.scene.dir_lights.tobytes() == b.scene.dir_lights.tobytes()
    for (p, n, i), (q, m, j) in zip(a.scene.meshes, b.scene.meshes):
        assert p.tobytes() == q.tobytes() and n.tobytes() == m.tobytes() and i.tobytes() == j.tobytes()
    assert a.camera.inv_view.tobytes() == b.camera.inv_view.tobytes()


# ---------------------------------------------------------------- the C++ loader (bi_assets.h)

ASSETS_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "bisemutum-engine_amd", "libbi_assets.so")


class _PtMesh(__import__("ctypes").Structure):
    import ctypes as C
    _fields_ = [("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("num_vertices", C.c_uint32), ("num_indices", C.c_uint32),
                ("texcoords", C.POINTER(C.c_float)), ("tangents", C.POINTER(C.c_float))]


def native_lib():
    import ctypes as C
    if not os.path.exists(ASSETS_LIB):
        pytest.skip("libbi_assets.so not built")
    lib = C.CDLL(ASSETS_LIB)
    lib.bi_static_mesh_load.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
    lib.bi_static_mesh_destroy.argtypes = [C.c_void_p]
    lib.bi_static_mesh_counts.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    lib.bi_static_mesh_submesh.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(_PtMesh), C.c_char_p, C.c_size_t]
    return lib


def native_load(lib, data):
    import ctypes as C
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    buf = C.create_string_buffer(data, len(data))
    rc = lib.bi_static_mesh_load(buf, len(data), C.byref(h), err, 256)
    return rc, h, err.value.decode()


@pytest.mark.parametrize("name", ["cube", "plane", "sphere"])
@pytest.mark.parametrize("version", [1, 2])
def test_native_loader_matches(name, version):
    import ctypes as C
    lib = native_lib()
    m = assets.read_static_mesh(mesh_bytes(name))
    data = mesh_bytes(name) if version == 2 else assets.encode_static_mesh(m, 1)
    rc, h, err = native_load(lib, data)
    assert rc == 0, err
    counts = (C.c_uint64 * 8)()
    lib.bi_static_mesh_counts(h, counts)
    assert list(counts) == [len(m.positions), len(m.normals), len(m.tangents), len(m.colors), len(m.texcoords),
                            len(m.texcoords2), len(m.indices), len(m.submeshes)]
    out = _PtMesh()
    assert lib.bi_static_mesh_submesh(h, 0, C.byref(out), None, 0) == 0
    p, n, idx = assets.submesh_triangles(m, 0)
    got_p = np.ctypeslib.as_array(out.positions, (out.num_vertices * 3,)).reshape(-1, 3)
    got_n = np.ctypeslib.as_array(out.normals, (out.num_vertices * 3,)).reshape(-1, 3)
    got_i = np.ctypeslib.as_array(out.indices, (out.num_indices,)).reshape(-1, 3)
    assert np.array_equal(got_p.view(np.uint32), p.view(np.uint32))
    assert np.array_equal(got_n.view(np.uint32), n.view(np.uint32))
    assert np.array_equal(got_i, idx)
    uv, tan = assets.submesh_attributes(m)
    for got, want, w in ((out.texcoords, uv, 2), (out.tangents, tan, 4)):
        if want is None:
            assert not got
        else:
            arr = np.ctypeslib.as_array(got, (out.num_vertices * w,)).reshape(-1, w)
            assert np.array_equal(arr.view(np.uint32), want.view(np.uint32))
    lib.bi_static_mesh_destroy(h)


def test_native_loader_errors():
    import ctypes as C
    lib = native_lib()
    good = mesh_bytes("cube")
    bad_version = bytearray(good)
    bad_version[22:26] = (7).to_bytes(4, "little")
    bad_z = bytearray(good)
    bad_z[50] ^= 0xFF
    for data, why in ((b"\0\0\0\0" + good[4:], "not a StaticMesh"), (good[:20], "truncated"),
                      (bytes(bad_version), "version 7"), (bytes(bad_z), "inflate")):
        rc, h, err = native_load(lib, data)
        assert rc != 0 and why in err and not h.value, (why, err)
    m = assets.read_static_mesh(good)
    m.submeshes = m.submeshes.copy()
    m.submeshes["topology"] = 4
    rc, h, err = native_load(lib, assets.encode_static_mesh(m, 2))
    assert rc == 0
    msg = C.create_string_buffer(128)
    out = _PtMesh()
    assert lib.bi_static_mesh_submesh(h, 0, C.byref(out), msg, 128) != 0 and b"triangle list" in msg.value
    assert lib.bi_static_mesh_submesh(h, 5, C.byref(out), msg, 128) != 0 and b"out of range" in msg.value
    lib.bi_static_mesh_destroy(h)


def test_parent_transforms_compose():
    """An object under a parent is placed by parent_world @ local (matrix product of TRS
    transforms); the camera and light follow their parents too."""
    def obj(name, comps, parent=None):
        o = {"name": name, "components": [{"type": t, "value": v} for t, v in comps]}
        if parent is not None:
            o["parent"] = parent
        return o
    root = {"translation": [1.0, 0.0, 0.0], "rotation": [0.0, 90.0, 0.0], "scaling": [2.0, 2.0, 2.0]}
    child = {"translation": [0.0, 0.0, 1.0]}
    mesh = [("StaticMeshComponent", {"static_mesh": {"asset_id": 1}}),
            ("MeshRendererComponent", {"materials": [{"asset_id": 5}]})]
    desc = {"objects": [
        obj("root", [("Transform", root)]),
        obj("cube", [("Transform", child)] + mesh, parent=0),
        obj("cam", [("Transform", {"translation": [0.0, 0.0, 5.0]}),
                    ("CameraComponent", {"yfov": 40.0, "render_target_size": [64, 32]})], parent=0),
        obj("sun", [("Transform", {"rotation": [90.0, 0.0, 0.0]}), ("DirectionalLightComponent", {"strength": 2.0})],
            parent=0),
    ]}
    white = {"blend_mode": "opaque", "statements": [], "params": []}
    ps = assets.scene_from_description(desc, lambda i: mesh_bytes("cube") if i == 1 else white)
    x = ps.scene.instances[0]["xform"].reshape(3, 4).astype(np.float64)
    # child origin: parent maps (0,0,1) -> R_y(90) (0,0,2) + (1,0,0) = (3,0,0); scale 2 carried
    assert np.allclose(x[:, 3], [3.0, 0.0, 0.0], atol=1e-6)
    assert np.allclose(x[:, :3], assets.euler_zxy_degrees((0, 90, 0)) * 2.0, atol=1e-6)
    assert np.allclose(ps.camera.inv_view[:3, 3], [11.0, 0.0, 0.0], atol=1e-5)
    # light: R_parent R_local (0,1,0) with the scale removed; R_x(90) takes y to z, R_y(90) z to x
    assert np.allclose(ps.scene.dir_lights[0][4:7], [1.0, 0.0, 0.0], atol=1e-6)
    assert np.allclose(ps.scene.dir_lights[0][0:3], [2.0, 2.0, 2.0])
