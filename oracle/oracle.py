"""ctypes binding of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker, never the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")   # ORACLE_LIB: e.g. the sanitized build

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)

GBUFFER_QUANTIZE = 1 << 30   # render32 flag: drift mode (the reference's GBuffer round trip per surface)


class TextureDesc(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("filter", C.c_uint32), ("address_u", C.c_uint32), ("address_v", C.c_uint32)]


class RectTextureDesc(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("num_levels", C.c_uint32),
                ("format", C.c_uint32), ("mag_filter", C.c_uint32), ("min_filter", C.c_uint32),
                ("mipmap_mode", C.c_uint32), ("address_u", C.c_uint32), ("address_v", C.c_uint32),
                ("lod_bias", C.c_float), ("lod_min", C.c_float), ("lod_max", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_meshes", C.c_uint32), ("positions", _f32p), ("normals", _f32p), ("indices", _u32p),
        ("vtx_off", _u32p), ("idx_off", _u32p), ("idx_cnt", _u32p),
        ("num_instances", C.c_uint32), ("inst_xform", _f32p), ("inst_mesh", _u32p), ("inst_mat", _u32p),
        ("num_materials", C.c_uint32), ("materials", _f32p),
        ("num_dir_lights", C.c_uint32), ("dir_lights", _f32p),
        ("sky_w", C.c_uint32), ("sky_h", C.c_uint32), ("sky_rgba", _f32p),
        ("sky_color", C.c_float * 3), ("sky_rot", C.c_float * 9),
        ("texcoords", _f32p), ("tangents", _f32p),
        ("num_textures", C.c_uint32), ("textures", C.POINTER(TextureDesc)),
        ("sky_layout", C.c_uint32),
        ("num_point_lights", C.c_uint32), ("point_lights", _f32p),
        ("num_rect_lights", C.c_uint32), ("rect_lights", _f32p), ("ltc_matrix", _f32p * 3), ("ltc_norm", _f32p),
        ("num_rect_textures", C.c_uint32), ("rect_textures", C.POINTER(RectTextureDesc)),
    ]


class Camera(C.Structure):
    _fields_ = [("inv_view", C.c_float * 16), ("inv_proj", C.c_float * 16), ("proj_view", C.c_float * 16),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class Settings(C.Structure):
    _fields_ = [("ray_length", C.c_float), ("max_bounces", C.c_uint32), ("seed", C.c_uint32),
                ("flags", C.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_create.restype = C.c_void_p
        L.ora_create.argtypes = [C.POINTER(SceneDesc)]
        L.ora_destroy.argtypes = [C.c_void_p]
        L.ora_set_treelet_rounds.argtypes = [C.c_void_p, C.c_uint32]
        L.ora_num_tris.restype = C.c_uint32
        L.ora_num_tris.argtypes = [C.c_void_p]
        L.ora_get_world.argtypes = [C.c_void_p, _f32p, _f32p, _u32p]
        L.ora_get_bvh.argtypes = [C.c_void_p, _u32p, _u32p, _i32p, _f32p, _i32p, _f32p]
        L.ora_get_bvh4.argtypes = [C.c_void_p, _i32p, _f32p]
        L.ora_trace.argtypes = [C.c_void_p, C.c_uint32, _f32p, C.c_int, _f32p, _u32p, _f32p, _f32p]
        L.ora_trace_brute64.argtypes = [C.c_void_p, C.c_uint32, _f32p, _f64p, _u32p]
        L.ora_render32.restype = C.c_int
        L.ora_render32.argtypes = [C.c_void_p, C.POINTER(Camera), C.POINTER(Settings), C.c_uint32, C.c_uint32,
                                   _f32p, _u8p, _u8p, _u64p, C.c_int]
        L.ora_render32_pixels.restype = C.c_int
        L.ora_render32_pixels.argtypes = [C.c_void_p, C.POINTER(Camera), C.POINTER(Settings), C.c_uint32, C.c_uint32,
                                          C.c_uint32, _u32p, _f32p, C.c_int]
        L.ora_render32_cached.restype = C.c_int
        L.ora_render32_cached.argtypes = [C.c_void_p, C.POINTER(Camera), C.POINTER(Settings), C.c_uint32, _f32p,
                                          C.c_uint32, C.c_uint32, _f32p, C.c_int]
        L.ora_train_samples.restype = C.c_int
        L.ora_train_samples.argtypes = [C.c_void_p, C.POINTER(Camera), C.POINTER(Settings), C.c_uint32,
                                        C.c_uint32, C.c_uint32, _f32p, _f32p]
        L.ora_pixel_of.restype = C.c_uint32
        L.ora_pixel_of.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.ora_render64.restype = C.c_int
        L.ora_render64.argtypes = [C.c_void_p, C.POINTER(Camera), C.POINTER(Settings), C.c_uint32, _u32p,
                                   C.c_uint32, C.c_uint32, _f64p, _f64p, C.c_int]
        L.ora_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _u32p]
        L.ora_tex_sample32.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, _f32p, _f32p]
        L.ora_pcg32_keyed.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        L.ora_morton3.restype = C.c_uint32
        L.ora_morton3.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.ora_math32.argtypes = [C.c_int, C.c_uint32, _f32p, _f32p, _f32p]
        L.ora_bsdf32.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p]
        L.ora_bsdf64.argtypes = [_f32p, _f64p, _f64p, _f64p, _f64p]
        L.ora_vndf32.argtypes = [_f32p, C.c_float, C.c_float, C.c_float, _f32p, _f32p]
        L.ora_vndf_pdf64.restype = C.c_double
        L.ora_vndf_pdf64.argtypes = [_f64p, _f64p, C.c_double]
        L.ora_sky_eval32.argtypes = [C.c_void_p, _f32p, _f32p]
        L.ora_gbuffer32.restype = C.c_int
        L.ora_gbuffer32.argtypes = [C.c_void_p, C.POINTER(Camera), C.POINTER(Settings), C.c_uint32, C.c_uint32, _u32p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ora_half32.restype = C.c_uint16
        L.ora_half32.argtypes = [C.c_float]
        L.ora_rect_light32.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float, _f32p, _f32p, _f32p,
                                       _f32p, _f32p]
        L.ora_ltc_fetch32.argtypes = [_f32p, C.c_float, C.c_float, C.c_float, C.c_float, _f32p, _f32p]
        L.ora_sky_sample32.argtypes = [C.c_void_p, C.c_float, C.c_float, _f32p, _f32p]
        L.ora_sky_sample_le32.argtypes = [C.c_void_p, C.c_float, C.c_float, _f32p, _f32p, _f32p]
        L.ora_rect_light_scene32.argtypes = [C.c_void_p, C.c_uint32, _f32p, _f32p, _f32p, C.c_float, C.c_float,
                                             _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.ora_rect_tex_lookup32.argtypes = [C.c_void_p, C.c_uint32, _f32p, _f32p, C.c_float, _f32p, _f32p]
        L.ora_rect_tex_sample32.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_float, C.c_float, _f32p]
        L.ora_log2_32.argtypes = [C.c_uint32, _f32p, _f32p]
        L.ora_sky_pdf32.restype = C.c_float
        L.ora_sky_pdf32.argtypes = [C.c_void_p, _f32p]
        L.ora_camera_ray32.argtypes = [C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                       _f32p, _f32p]
        L.ora_bsdf64_batch.argtypes = [_f32p, _f64p, _f64p, C.c_uint32, _f64p, _f64p]
        L.ora_vndf_pdf64_batch.argtypes = [_f64p, C.c_uint32, _f64p, C.c_double, _f64p]
        L.ora_sky_batch32.argtypes = [C.c_void_p, C.c_uint32, _f32p, _f32p, _f32p]
        _lib = L
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def camera_struct(cam) -> Camera:
    c = Camera()
    C.memmove(C.addressof(c.inv_view), np.ascontiguousarray(cam.inv_view, np.float32).ctypes.data, 64)
    C.memmove(C.addressof(c.inv_proj), np.ascontiguousarray(cam.inv_proj, np.float32).ctypes.data, 64)
    C.memmove(C.addressof(c.proj_view), np.ascontiguousarray(cam.proj_view, np.float32).ctypes.data, 64)
    c.width, c.height = cam.width, cam.height
    return c


def settings_struct(max_bounces, seed, ray_length=1000.0, flags=0) -> Settings:
    return Settings(ray_length, max_bounces, seed, flags)


class OracleScene:
    """A scenes.Scene loaded into the CPU restatement."""

    def __init__(self, scene):
        self.scene = scene
        pos = [np.ascontiguousarray(m[0], np.float32) for m in scene.meshes]
        nrm = [np.ascontiguousarray(m[1], np.float32) for m in scene.meshes]
        idx = [np.ascontiguousarray(m[2], np.uint32).reshape(-1) for m in scene.meshes]
        self._keep = dict(
            positions=np.concatenate(pos), normals=np.concatenate(nrm), indices=np.concatenate(idx),
            vtx_off=np.array(np.cumsum([0] + [len(p) for p in pos[:-1]]), np.uint32),
            idx_off=np.array(np.cumsum([0] + [len(i) for i in idx[:-1]]), np.uint32),
            idx_cnt=np.array([len(i) for i in idx], np.uint32),
            inst_xform=np.ascontiguousarray(scene.instances["xform"], np.float32),
            inst_mesh=np.ascontiguousarray(scene.instances["mesh"], np.uint32),
            inst_mat=np.ascontiguousarray(scene.instances["material"], np.uint32),
            materials=np.ascontiguousarray(scene.materials, np.float32),
            dir_lights=np.ascontiguousarray(scene.dir_lights, np.float32),
            point_lights=np.ascontiguousarray(getattr(scene, "point_lights", None) if getattr(scene, "point_lights", None)
                                              is not None else np.zeros((0, 16), np.float32), np.float32).reshape(-1, 16),
            sky=np.ascontiguousarray(scene.sky_rgba, np.float32),
        )
        k = self._keep
        d = SceneDesc()
        d.num_meshes = len(pos)
        d.positions, d.normals, d.indices = ptr(k["positions"], _f32p), ptr(k["normals"], _f32p), ptr(k["indices"], _u32p)
        d.vtx_off, d.idx_off, d.idx_cnt = ptr(k["vtx_off"], _u32p), ptr(k["idx_off"], _u32p), ptr(k["idx_cnt"], _u32p)
        d.num_instances = len(scene.instances)
        d.inst_xform, d.inst_mesh, d.inst_mat = ptr(k["inst_xform"], _f32p), ptr(k["inst_mesh"], _u32p), ptr(k["inst_mat"], _u32p)
        d.num_materials = len(scene.materials)
        d.materials = ptr(k["materials"], _f32p)
        d.num_dir_lights = len(scene.dir_lights)
        d.dir_lights = ptr(k["dir_lights"], _f32p)
        d.num_point_lights = len(k["point_lights"])
        d.point_lights = ptr(k["point_lights"], _f32p) if len(k["point_lights"]) else None
        rects = getattr(scene, "rect_lights", None)
        if rects is not None and len(rects):
            k["rect_lights"] = np.ascontiguousarray(rects, np.float32).reshape(-1, 28)
            luts = [np.ascontiguousarray(t, np.float32).reshape(-1) for t in scene.ltc]
            k["ltc"] = luts
            d.num_rect_lights = len(k["rect_lights"])
            d.rect_lights = ptr(k["rect_lights"], _f32p)
            for i in range(3):
                d.ltc_matrix[i] = ptr(luts[i], _f32p)
            d.ltc_norm = ptr(luts[3], _f32p)
        rtex = getattr(scene, "rect_textures", None) or []
        if rtex:
            arr = (RectTextureDesc * len(rtex))()
            chains = []
            for i, t in enumerate(rtex):
                a = t.chain()
                chains.append(a)
                h, w = t.levels[0].shape[:2]
                arr[i] = RectTextureDesc(a.ctypes.data, w, h, len(t.levels), t.format, t.mag_filter, t.min_filter,
                                         t.mipmap_mode, t.address_u, t.address_v, t.lod_bias, t.lod_min, t.lod_max)
            k["rect_tex"] = (arr, chains)
            d.num_rect_textures = len(rtex)
            d.rect_textures = arr
        layout = getattr(scene, "sky_layout", 0)
        if layout == 1:   # cube faces [6, S, S, 4]: rows face-major
            if k["sky"].size and (k["sky"].ndim != 4 or k["sky"].shape[0] != 6 or k["sky"].shape[1] != k["sky"].shape[2]):
                raise ValueError(f"cube sky must be [6, S, S, 4], got {k['sky'].shape}")
            d.sky_w = k["sky"].shape[1] if k["sky"].size else 0
            d.sky_h = 6 * d.sky_w
            d.sky_layout = 1 if d.sky_w else 0
        else:
            d.sky_h, d.sky_w = scene.sky_rgba.shape[0], scene.sky_rgba.shape[1]
        d.sky_rgba = ptr(k["sky"], _f32p) if k["sky"].size else None
        for i in range(3):
            d.sky_color[i] = float(scene.sky_color[i])
        for i in range(9):
            d.sky_rot[i] = float(scene.sky_rot[i])
        nv = [len(p) for p in pos]
        if scene.mesh_uv is not None:
            k["uv"] = np.ascontiguousarray(np.concatenate(
                [np.zeros((n, 2), np.float32) if u is None else np.asarray(u, np.float32).reshape(n, 2)
                 for n, u in zip(nv, scene.mesh_uv)]), np.float32)
            d.texcoords = ptr(k["uv"], _f32p)
        if scene.mesh_tan is not None:
            k["tan"] = np.ascontiguousarray(np.concatenate(
                [np.tile(np.array([1, 0, 0, 1], np.float32), (n, 1)) if t is None else
                 np.asarray(t, np.float32).reshape(n, 4) for n, t in zip(nv, scene.mesh_tan)]), np.float32)
            d.tangents = ptr(k["tan"], _f32p)
        if scene.textures:
            k["tex_data"] = [np.ascontiguousarray(t.texels) for t in scene.textures]
            k["tex"] = (TextureDesc * len(scene.textures))()
            for i, (t, a) in enumerate(zip(scene.textures, k["tex_data"])):
                k["tex"][i] = TextureDesc(a.ctypes.data, a.shape[1], a.shape[0], t.format, t.filter,
                                          t.address_u, t.address_v)
            d.num_textures = len(scene.textures)
            d.textures = k["tex"]
        self.h = lib().ora_create(C.byref(d))
        self.n = lib().ora_num_tris(self.h)

    def set_treelet_rounds(self, rounds):
        """Rebuild the BVH with `rounds` rounds of treelet restructuring (0: the LBVH as built)."""
        lib().ora_set_treelet_rounds(self.h, rounds)

    def close(self):
        if self.h:
            lib().ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def world(self):
        v = np.zeros((self.n, 9), np.float32)
        nn = np.zeros((self.n, 9), np.float32)
        m = np.zeros(self.n, np.uint32)
        lib().ora_get_world(self.h, ptr(v, _f32p), ptr(nn, _f32p), ptr(m, _u32p))
        return v, nn, m

    def bvh(self):
        n = self.n
        ni = max(n - 1, 1)
        morton = np.zeros(n, np.uint32)
        order = np.zeros(n, np.uint32)
        children = np.zeros((ni, 2), np.int32)
        bounds = np.zeros((ni, 12), np.float32)
        parents = np.zeros(ni + n, np.int32)
        cb = np.zeros(6, np.float32)
        lib().ora_get_bvh(self.h, ptr(morton, _u32p), ptr(order, _u32p), ptr(children, _i32p),
                          ptr(bounds, _f32p), ptr(parents, _i32p), ptr(cb, _f32p))
        return dict(morton=morton, order=order, children=children, bounds=bounds, parents=parents,
                    centroid_bounds=cb)

    def bvh4(self):
        ni = max(self.n - 1, 1)
        child4 = np.zeros((ni, 4), np.int32)
        box4 = np.zeros((ni, 24), np.float32)
        lib().ora_get_bvh4(self.h, ptr(child4, _i32p), ptr(box4, _f32p))
        return child4, box4

    def trace(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        t = np.zeros(n, np.float32)
        leaf = np.zeros(n, np.uint32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        lib().ora_trace(self.h, n, ptr(rays, _f32p), int(any_hit), ptr(t, _f32p), ptr(leaf, _u32p),
                        ptr(u, _f32p), ptr(v, _f32p))
        return (leaf,) if any_hit else (t, leaf, u, v)

    def trace_brute64(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        t = np.zeros(n, np.float64)
        tri = np.zeros(n, np.uint32)
        lib().ora_trace_brute64(self.h, n, ptr(rays, _f32p), ptr(t, _f64p), ptr(tri, _u32p))
        return t, tri

    def render32(self, cam, max_bounces, seed, spp, sample_base=0, sums=None, ray_length=1000.0, flags=0,
                 record=False, threads=0):
        npix = cam.width * cam.height
        if sums is None:
            sums = np.zeros((npix, 3), np.float32)
        segs = np.zeros(spp * npix, np.uint8) if record else None
        bits = np.zeros((spp * npix, 16), np.uint8) if record else None
        counts = np.zeros(4, np.uint64)
        c = camera_struct(cam)
        s = settings_struct(max_bounces, seed, ray_length, flags)
        rc = lib().ora_render32(self.h, C.byref(c), C.byref(s), sample_base, spp, ptr(sums, _f32p),
                                ptr(segs, _u8p), ptr(bits, _u8p), ptr(counts, _u64p), threads)
        assert rc == 0
        out = dict(sums=sums, counts=counts)
        if record:
            out["segments"] = segs
            out["shadow_bits"] = bits
        return out

    def render32_pixels(self, cam, max_bounces, seed, spp, pixels, sample_base=0, ray_length=1000.0, flags=0,
                        threads=0):
        """Running sums [len(pixels), 3] of the listed pixels (row-major image indices) over
        samples [sample_base, sample_base + spp), as render32 sums them."""
        pixels = np.ascontiguousarray(pixels, np.uint32)
        sums = np.zeros((len(pixels), 3), np.float32)
        rc = lib().ora_render32_pixels(self.h, C.byref(camera_struct(cam)),
                                       C.byref(settings_struct(max_bounces, seed, ray_length, flags)), sample_base,
                                       spp, len(pixels), ptr(pixels, _u32p), ptr(sums, _f32p), threads)
        assert rc == 0
        return sums

    def render32_cached(self, cam, max_bounces, seed, spp, cache_bounce, cache_rgb, flags=0, ray_length=1000.0,
                        threads=0):
        """Running sums of paths that end at segment cache_bounce in a constant cache."""
        sums = np.zeros((cam.width * cam.height, 3), np.float32)
        rgb = np.ascontiguousarray(cache_rgb, np.float32)
        rc = lib().ora_render32_cached(self.h, C.byref(camera_struct(cam)),
                                       C.byref(settings_struct(max_bounces, seed, ray_length, flags)), cache_bounce,
                                       ptr(rgb, _f32p), 0, spp, ptr(sums, _f32p), threads)
        assert rc == 0
        return sums

    def train_samples(self, cam, max_bounces, seed, cache_bounce, n, suffixes=1, flags=0, ray_length=1000.0):
        """Records [n, 16] and targets [n, 3] of the cache's training slots 0..n-1 (targets
        averaged over `suffixes` continuations from each recorded vertex)."""
        rec = np.zeros((n, 16), np.float32)
        tgt = np.zeros((n, 3), np.float32)
        rc = lib().ora_train_samples(self.h, C.byref(camera_struct(cam)),
                                     C.byref(settings_struct(max_bounces, seed, ray_length, flags)), cache_bounce, n,
                                     suffixes, ptr(rec, _f32p), ptr(tgt, _f32p))
        assert rc == 0
        return rec, tgt

    def render64(self, cam, max_bounces, pixels, spp, seed64=0x64, ray_length=1000.0, flags=0, threads=0):
        pixels = np.ascontiguousarray(pixels, np.uint32)
        mean = np.zeros((len(pixels), 3), np.float64)
        var = np.zeros((len(pixels), 3), np.float64)
        c = camera_struct(cam)
        s = settings_struct(max_bounces, 0, ray_length, flags)
        rc = lib().ora_render64(self.h, C.byref(c), C.byref(s), len(pixels), ptr(pixels, _u32p), spp, seed64,
                                ptr(mean, _f64p), ptr(var, _f64p), threads)
        assert rc == 0
        return mean, var

    def tex_sample(self, tex, uv):
        """Level-0 samples [n, 4] of texture `tex` at uv [n, 2]."""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((len(uv), 4), np.float32)
        lib().ora_tex_sample32(self.h, tex, len(uv), ptr(uv, _f32p), ptr(out, _f32p))
        return out

    def gbuffer32(self, cam, pixels, sample=0, seed=0, flags=0, ray_length=1000.0, max_bounces=5):
        """Primary-hit GBuffer of `sample` at the listed pixels: dict of [n, 4] arrays (uint16
        half bits base_color / normal_roughness, uint16 fresnel, uint8 material_0)."""
        pixels = np.ascontiguousarray(pixels, np.uint32)
        n = len(pixels)
        out = dict(base_color=np.zeros((n, 4), np.uint16), normal_roughness=np.zeros((n, 4), np.uint16),
                   fresnel=np.zeros((n, 4), np.uint16), material_0=np.zeros((n, 4), np.uint8))
        rc = lib().ora_gbuffer32(self.h, C.byref(camera_struct(cam)),
                                 C.byref(settings_struct(max_bounces, seed, ray_length, flags)), sample, n,
                                 ptr(pixels, _u32p), out["base_color"].ctypes.data, out["normal_roughness"].ctypes.data,
                                 out["fresnel"].ctypes.data, out["material_0"].ctypes.data)
        if rc != 0:
            raise RuntimeError(f"ora_gbuffer32 failed ({rc})")
        return out

    def sky_eval(self, d):
        d = np.ascontiguousarray(d, np.float32)
        out = np.zeros(3, np.float32)
        lib().ora_sky_eval32(self.h, ptr(d, _f32p), ptr(out, _f32p))
        return out

    def sky_sample(self, u0, u1):
        d = np.zeros(3, np.float32)
        p = np.zeros(1, np.float32)
        lib().ora_sky_sample32(self.h, u0, u1, ptr(d, _f32p), ptr(p, _f32p))
        return d, float(p[0])

    def sky_sample_le(self, u0, u1):
        """(direction, pdf, radiance at the sample) of the sky sampler."""
        d = np.zeros(3, np.float32)
        p = np.zeros(1, np.float32)
        le = np.zeros(3, np.float32)
        lib().ora_sky_sample_le32(self.h, u0, u1, ptr(d, _f32p), ptr(p, _f32p), ptr(le, _f32p))
        return d, float(p[0]), le

    def rect_light_scene32(self, light, P, N, V, rx, ry, base, f0, f90):
        """(rgb, terms, dirs [2, 3]) of rect light `light` of the scene at a lit surface, with its
        texture: dirs are the diffuse and specular lookup directions (zero when unused)."""
        a = [np.ascontiguousarray(x, np.float32) for x in (P, N, V, base, f0, f90)]
        rgb, terms, dirs = np.zeros(3, np.float32), np.zeros(4, np.float32), np.zeros(6, np.float32)
        lib().ora_rect_light_scene32(self.h, light, *[ptr(x, _f32p) for x in a[:3]], float(rx), float(ry),
                                     *[ptr(x, _f32p) for x in a[3:]], ptr(rgb, _f32p), ptr(terms, _f32p),
                                     ptr(dirs, _f32p))
        return rgb, terms, dirs.reshape(2, 3)

    def rect_tex_lookup32(self, light, d, P, roughness):
        """(rgb, (u, v, level)) of rect light `light`'s texture seen from P along d."""
        d = np.ascontiguousarray(d, np.float32)
        P = np.ascontiguousarray(P, np.float32)
        rgb, uvl = np.zeros(3, np.float32), np.zeros(3, np.float32)
        lib().ora_rect_tex_lookup32(self.h, light, ptr(d, _f32p), ptr(P, _f32p), float(roughness), ptr(rgb, _f32p),
                                    ptr(uvl, _f32p))
        return rgb, uvl

    def rect_tex_sample32(self, tex, u, v, level):
        out = np.zeros(4, np.float32)
        lib().ora_rect_tex_sample32(self.h, tex, float(u), float(v), float(level), ptr(out, _f32p))
        return out

    def sky_pdf(self, d):
        d = np.ascontiguousarray(d, np.float32)
        return float(lib().ora_sky_pdf32(self.h, ptr(d, _f32p)))


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().ora_pcg32(initstate, initseq, n, ptr(out, _u32p))
    return out


def pcg32_keyed(seed, pixel, sample, bounce, n):
    out = np.zeros(n, np.float32)
    lib().ora_pcg32_keyed(seed, pixel, sample, bounce, n, ptr(out, _f32p))
    return out


def morton3(x, y, z):
    return lib().ora_morton3(x, y, z)


def math32(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float32)
    out = np.zeros_like(x)
    lib().ora_math32(fn, len(x), ptr(x, _f32p), ptr(y, _f32p), ptr(out, _f32p))
    return out


def bsdf32(mat, n, v, l):
    out = np.zeros(3, np.float32)
    args = [np.ascontiguousarray(a, np.float32) for a in (mat, n, v, l)]
    lib().ora_bsdf32(*[ptr(a, _f32p) for a in args], ptr(out, _f32p))
    return out


def bsdf64(mat, n, v, l):
    out = np.zeros(3, np.float64)
    m = np.ascontiguousarray(mat, np.float32)
    args = [np.ascontiguousarray(a, np.float64) for a in (n, v, l)]
    lib().ora_bsdf64(ptr(m, _f32p), *[ptr(a, _f64p) for a in args], ptr(out, _f64p))
    return out


def vndf32(v_local, alpha, u0, u1):
    v = np.ascontiguousarray(v_local, np.float32)
    h = np.zeros(3, np.float32)
    p = np.zeros(1, np.float32)
    lib().ora_vndf32(ptr(v, _f32p), alpha, u0, u1, ptr(h, _f32p), ptr(p, _f32p))
    return h, float(p[0])


def vndf_pdf64(v_local, l_local, alpha):
    v = np.ascontiguousarray(v_local, np.float64)
    l = np.ascontiguousarray(l_local, np.float64)
    return float(lib().ora_vndf_pdf64(ptr(v, _f64p), ptr(l, _f64p), alpha))


def bsdf64_batch(mat, n, v, ls):
    m = np.ascontiguousarray(mat, np.float32)
    n = np.ascontiguousarray(n, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    ls = np.ascontiguousarray(ls, np.float64)
    out = np.zeros((len(ls), 3), np.float64)
    lib().ora_bsdf64_batch(ptr(m, _f32p), ptr(n, _f64p), ptr(v, _f64p), len(ls), ptr(ls, _f64p), ptr(out, _f64p))
    return out


def vndf_pdf64_batch(v_local, ls, alpha):
    v = np.ascontiguousarray(v_local, np.float64)
    ls = np.ascontiguousarray(ls, np.float64)
    out = np.zeros(len(ls), np.float64)
    lib().ora_vndf_pdf64_batch(ptr(v, _f64p), len(ls), ptr(ls, _f64p), alpha, ptr(out, _f64p))
    return out


def ltc_tables(m0, m1, m2, norm):
    """The four LTC tables in the kernels' layout: [4 * 4096 * 4] float32 (matrix rows, then
    the norm pair padded to four floats)."""
    t = np.zeros((4, 4096, 4), np.float32)
    for i, m in enumerate((m0, m1, m2)):
        t[i] = np.asarray(m, np.float32).reshape(4096, 4)
    t[3, :, :2] = np.asarray(norm, np.float32).reshape(4096, 2)
    return t.reshape(-1)


def rect_light32(rect, tables, P, N, V, rx, ry, base, f0, f90):
    """(rgb, (diffuse integral, specular integral, norm, Fresnel term)) of one rect light at
    a lit surface; tables from ltc_tables."""
    a = [np.ascontiguousarray(x, np.float32) for x in (rect, tables, P, N, V, base, f0, f90)]
    rgb = np.zeros(3, np.float32)
    terms = np.zeros(4, np.float32)
    lib().ora_rect_light32(ptr(a[0], _f32p), ptr(a[1], _f32p), ptr(a[2], _f32p), ptr(a[3], _f32p), ptr(a[4], _f32p),
                           float(rx), float(ry), ptr(a[5], _f32p), ptr(a[6], _f32p), ptr(a[7], _f32p),
                           ptr(rgb, _f32p), ptr(terms, _f32p))
    return rgb, terms


def ltc_fetch32(tables, u3, u2, u1, u0):
    """(transform [3, 3] rows, (norm, Fresnel term)) of the LTC tables at (u3, u2, u1, u0)."""
    t = np.ascontiguousarray(tables, np.float32)
    m = np.zeros(9, np.float32)
    b = np.zeros(2, np.float32)
    lib().ora_ltc_fetch32(ptr(t, _f32p), float(u3), float(u2), float(u1), float(u0), ptr(m, _f32p), ptr(b, _f32p))
    return m.reshape(3, 3), b


def log2_32(x):
    """The kernels' log2 of x >= 1 (float32 array)."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    out = np.zeros_like(x)
    lib().ora_log2_32(len(x), ptr(x, _f32p), ptr(out, _f32p))
    return out


def half32(x):
    """float -> IEEE half bits (round to nearest even), the restatement's conversion."""
    return int(lib().ora_half32(float(np.float32(x))))


def sky_batch32(scene, dirs):
    d = np.ascontiguousarray(dirs, np.float32)
    rgb = np.zeros((len(d), 3), np.float32)
    pdf = np.zeros(len(d), np.float32)
    lib().ora_sky_batch32(scene.h, len(d), ptr(d, _f32p), ptr(rgb, _f32p), ptr(pdf, _f32p))
    return rgb, pdf


def camera_ray32(cam, px, py, jx=0.5, jy=0.5):
    c = camera_struct(cam)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().ora_camera_ray32(C.byref(c), px, py, jx, jy, ptr(o, _f32p), ptr(d, _f32p))
    return o, d


def pixel_order(width, height):
    """Pixel index of each path-order position (q -> pixel)."""
    L = lib()
    return np.array([L.ora_pixel_of(q, width, height) for q in range(width * height)], np.uint32)
