"""Scene ingestion from the engine's GPU-scene buffers (GpuSceneData, gpu_scene_data.hpp:
10-24) and drawable shader records (DrawableSbtData, graphics_manager.cpp:1314-1330):
a scene is laid out in flat buffers the way the engine's mesh allocator lays it out
(meshes one after the other, submesh offsets into them, glm column-major matrices), and
bi_pt_drawables_layout (host only) must give back the scene's own meshes and 3x4
instance transforms (accel.cpp:119-131 transposes the matrix); bad records are refused.
The GPU half renders the result bit for bit like bi_pt_set_scene."""
import numpy as np
import pytest

import bi_pt
import scenes


def gpu_scene_of(scene, with_tangents=False, with_uv=False, split_first=False):
    """Flat buffers of the scene's meshes (each mesh's vertices after the previous mesh's),
    one drawable per instance.  split_first: the first mesh's triangles go to two submeshes
    with a base vertex of their own, as a multi-submesh StaticMesh is drawn."""
    P, N, T, UV, I = [], [], [], [], []
    subs = []   # per mesh: list of (position_offset, index_offset, index_count)
    voff = ioff = 0
    for m, (p, n, idx) in enumerate(scene.meshes):
        p = np.asarray(p, np.float32).reshape(-1, 3)
        n = np.asarray(n, np.float32).reshape(-1, 3)
        idx = np.asarray(idx, np.uint32).reshape(-1)
        P.append(p); N.append(n)
        tan = (scene.mesh_tan[m] if scene.mesh_tan is not None and scene.mesh_tan[m] is not None
               else np.tile(np.array([1, 0, 0, 1], np.float32), (len(p), 1)))
        uv = (scene.mesh_uv[m] if scene.mesh_uv is not None and scene.mesh_uv[m] is not None
              else np.zeros((len(p), 2), np.float32))
        T.append(np.asarray(tan, np.float32)); UV.append(np.asarray(uv, np.float32))
        if split_first and m == 0 and len(idx) >= 6:
            half = (len(idx) // 6) * 3
            # second submesh: its own base vertex (the smallest index it uses)
            base = int(idx[half:].min())
            I.append(idx[:half]); I.append(idx[half:] - base)
            subs.append([(voff, ioff, half), (voff + base, ioff + half, len(idx) - half)])
        else:
            I.append(idx)
            subs.append([(voff, ioff, len(idx))])
        voff += len(p)
        ioff += len(idx)
    drawables = []
    for k, rec in enumerate(scene.instances):
        M = np.eye(4, dtype=np.float32)
        M[:3, :4] = np.asarray(rec["xform"], np.float32).reshape(3, 4)
        for (vo, io, ic) in subs[int(rec["mesh"])]:
            attrs = bi_pt.VA_POSITION | bi_pt.VA_NORMAL | (bi_pt.VA_TANGENT if with_tangents else 0) | \
                (bi_pt.VA_TEXCOORD if with_uv else 0)
            drawables.append(dict(drawable_index=len(drawables), position_offset=3 * vo, normal_offset=3 * vo,
                                  tangent_offset=4 * vo, color_offset=3 * vo, texcoord_offset=2 * vo,
                                  texcoord2_offset=2 * vo, index_offset=io, material_offset=64 * int(rec["material"]),
                                  index_count=ic, attributes=attrs, material=int(rec["material"]),
                                  matrix=M.T.reshape(-1)))   # glm: column-major storage
    return bi_pt.GpuScene(np.concatenate(P).reshape(-1), np.concatenate(N).reshape(-1),
                          np.concatenate(T).reshape(-1), np.concatenate(UV).reshape(-1), np.concatenate(I), drawables)


def test_layout_gives_back_meshes_and_transposed_transforms(c0_scene):
    gs = gpu_scene_of(c0_scene)
    ranges, xf, mesh, mat = gs.layout()
    assert len(ranges) == len(c0_scene.meshes)
    np.testing.assert_array_equal(xf, np.asarray(c0_scene.instances["xform"], np.float32).reshape(-1, 12))
    np.testing.assert_array_equal(mat, c0_scene.instances["material"])
    voff = ioff = 0
    for m, (p, n, idx) in enumerate(c0_scene.meshes):
        r = ranges[m]
        assert r["position_offset"] == 3 * voff and r["index_offset"] == ioff
        assert r["index_count"] == np.asarray(idx).size and r["num_vertices"] == int(np.asarray(idx).max()) + 1
        assert r["tangent_offset"] == 2 ** 64 - 1 and r["texcoord_offset"] == 2 ** 64 - 1   # not named
        voff += len(p)
        ioff += np.asarray(idx).size
    # instances of the same mesh share its range
    for k, rec in enumerate(c0_scene.instances):
        assert mesh[k] == int(rec["mesh"])


def test_layout_transposes_an_oblique_matrix():
    """A non-symmetric matrix: glm stores column c at [4c, 4c + 4); the instance row r is
    (m[0][r], m[1][r], m[2][r], m[3][r])."""
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (3, 1))
    glm = np.arange(16, dtype=np.float32) + 1   # columns (1,2,3,4), (5,6,7,8), (9,10,11,12), (13,14,15,16)
    gs = bi_pt.GpuScene(p.reshape(-1), n.reshape(-1), None, None, np.array([0, 1, 2], np.uint32),
                        [dict(position_offset=0, normal_offset=0, index_offset=0, index_count=3,
                              attributes=bi_pt.VA_POSITION | bi_pt.VA_NORMAL, material=0, matrix=glm)])
    _, xf, _, _ = gs.layout()
    np.testing.assert_array_equal(xf[0], np.array([1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15], np.float32))


def test_layout_splits_submeshes_and_checks_bounds(c0_scene):
    gs = gpu_scene_of(c0_scene, with_tangents=True, with_uv=True, split_first=True)
    ranges, _, mesh, _ = gs.layout()
    assert len(ranges) == len(c0_scene.meshes) + 1
    assert all(r["tangent_offset"] != 2 ** 64 - 1 and r["texcoord_offset"] != 2 ** 64 - 1 for r in ranges)
    bad = [("index_count", 4), ("index_offset", 10 ** 8), ("position_offset", 10 ** 8), ("attributes", 0)]
    for key, value in bad:
        d = dict(position_offset=0, normal_offset=0, index_offset=0, index_count=3,
                 attributes=bi_pt.VA_POSITION | bi_pt.VA_NORMAL, material=0, matrix=np.eye(4).reshape(-1))
        d[key] = value
        g = bi_pt.GpuScene(gs.arrays[0], gs.arrays[1], None, None, gs.arrays[4], [d])
        with pytest.raises(bi_pt.PtError):
            g.layout()
    # an attribute named without its buffer
    d = dict(position_offset=0, normal_offset=0, index_offset=0, index_count=3,
             attributes=bi_pt.VA_POSITION | bi_pt.VA_TANGENT, material=0, matrix=np.eye(4).reshape(-1))
    with pytest.raises(bi_pt.PtError):
        bi_pt.GpuScene(gs.arrays[0], gs.arrays[1], None, None, gs.arrays[4], [d]).layout()


@pytest.mark.gpu
def test_drawables_render_like_set_scene(oracle_lib, c0_scene):
    print("case: c0 through GPU-scene buffers and drawables, 96x72 2spp", flush=True)
    cam = scenes.camera_c0(96, 72)
    a = bi_pt.PathTracer()
    a.set_scene(c0_scene)
    a.render(cam, spp=2, max_bounces=5, seed=0x6D7A)
    want = a.sums(cam.width * cam.height)
    b = bi_pt.PathTracer()
    gs = gpu_scene_of(c0_scene, split_first=True)
    b.set_scene_drawables(gs, c0_scene.materials, c0_scene.dir_lights, c0_scene.sky_rgba, c0_scene.sky_color,
                          c0_scene.sky_rot)
    b.render(cam, spp=2, max_bounces=5, seed=0x6D7A)
    got = b.sums(cam.width * cam.height)
    # same triangles in another order of meshes (a split submesh): same closest hits, same image
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    a.close()
    b.close()


@pytest.mark.gpu
def test_drawables_example_project(oracle_lib):
    """The example project (textures, normal maps, tangents and texture coordinates from
    the flat buffers) through the drawables equals the restatement."""
    print("case: scene_basic through GPU-scene buffers, 130x64 1spp", flush=True)
    from test_assets import fixture_scene
    from test_gpu_parity import radiance_check
    ps = fixture_scene(130, 64, on_unsupported="raise")
    sc = ps.scene
    gs = gpu_scene_of(sc, with_tangents=True, with_uv=True)
    g = bi_pt.PathTracer()
    g.set_scene_drawables(gs, sc.materials, sc.dir_lights, sc.sky_rgba, sc.sky_color, sc.sky_rot, sc.textures)
    st = ps.settings
    g.render(ps.camera, spp=1, max_bounces=st["max_bounces"], seed=0x5EED5B, ray_length=st["ray_length"])
    os_ = oracle_lib.OracleScene(sc).render32(ps.camera, st["max_bounces"], 0x5EED5B, spp=1,
                                              ray_length=st["ray_length"])["sums"]
    assert radiance_check(g.sums(130 * 64), os_)
    g.close()
