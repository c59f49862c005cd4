"""Python binding (ctypes) of the C ABI in include/bi_pt.h.

The pass runs in libbi_pt.so (hand-written HIP kernels for gfx950).  Output
buffers live in device memory allocated through torch, which is only plumbing
here; there is no CPU fallback: a missing library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# BI_PT_LIBRARY names another build of the library (for example one made with
# different compile-time options); by default the one beside this file is used.
LIB_PATH = os.environ.get("BI_PT_LIBRARY") or os.path.join(HERE, "libbi_pt.so")

FLAG_JITTER = 1 << 0
FLAG_NO_SKY_NEE = 1 << 1
FLAG_OUTPUT_RGBA16F = 1 << 2
FLAG_RECORD_QUEUES = 1 << 8
FLAG_TIMING = 1 << 9
FLAG_REUSE_BVH = 1 << 10

STATUS = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "HIP", 4: "NO_SCENE", 5: "DEVICE_CHECK",
          6: "UNSUPPORTED"}

# every entry point declared in include/bi_pt.h
EXPORTS = ["bi_pt_create", "bi_pt_destroy", "bi_pt_last_error", "bi_pt_set_scene", "bi_pt_update_transforms",
           "bi_pt_set_lights", "bi_pt_set_point_lights", "bi_pt_set_rect_lights", "bi_pt_set_rect_light_textures", "bi_pt_rect_texture_mips", "bi_pt_set_ltc_luts", "bi_pt_set_sky", "bi_pt_set_sky_cube", "bi_pt_render", "bi_pt_render_gbuffer", "bi_pt_drawables_layout", "bi_pt_set_scene_drawables", "bi_pt_reset_accumulation", "bi_pt_get_stats",
           "bi_pt_debug_get_bvh", "bi_pt_debug_get_bvh4", "bi_pt_debug_num_tris", "bi_pt_debug_queue", "bi_pt_debug_get_sums",
           "bi_pt_debug_trace", "bi_pt_trace_closest", "bi_pt_trace_occlusion", "bi_pt_batch_create", "bi_pt_batch_destroy", "bi_pt_batch_last_error",
           "bi_pt_batch_set_tuning", "bi_pt_batch_get_tuning", "bi_pt_render_batch", "bi_pt_batch_get_stats",
           "bi_pt_batch_autotune", "bi_pt_create_view", "bi_pt_server_create", "bi_pt_server_destroy",
           "bi_pt_server_last_error", "bi_pt_server_clock", "bi_pt_server_add_scene", "bi_pt_server_submit",
           "bi_pt_server_step", "bi_pt_server_poll", "bi_pt_server_take_result", "bi_pt_server_get_stats",
           "bi_pt_nrc_create", "bi_pt_nrc_destroy", "bi_pt_nrc_last_error", "bi_pt_render_nrc",
           "bi_pt_nrc_get_stats", "bi_pt_nrc_get_params", "bi_pt_nrc_set_params", "bi_pt_nrc_set_bounds",
           "bi_pt_nrc_infer", "bi_pt_nrc_train", "bi_pt_nrc_debug_gradient", "bi_pt_nrc_debug_samples",
           "bi_pt_set_build_options", "bi_pt_debug_inject_error", "bi_pt_server_debug_inject_error",
           "bi_pt_trace_closest_noalloc", "bi_pt_set_textures"]
NOALLOC_MAX_RAYS = 1 << 18
NRC_PARAMS = 17408
NRC_FREEZE = 1


class PtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"bi_pt {STATUS.get(status, status)}: {msg}")
        self.status = status


class DeviceDesc(C.Structure):
    _fields_ = [("device", C.c_int32), ("_pad", C.c_uint32), ("hbm_budget_bytes", C.c_uint64)]


class Mesh(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("num_vertices", C.c_uint32), ("num_indices", C.c_uint32),
                ("texcoords", C.POINTER(C.c_float)), ("tangents", C.POINTER(C.c_float))]


VA_POSITION, VA_NORMAL, VA_TANGENT, VA_TEXCOORD = 0x1, 0x4, 0x8, 0x20


class GpuSceneC(C.Structure):
    """bi_pt_gpu_scene: the flat vertex/index buffers (host) and their element counts."""
    _fields_ = [("positions", C.c_void_p), ("num_positions", C.c_uint64), ("normals", C.c_void_p),
                ("num_normals", C.c_uint64), ("tangents", C.c_void_p), ("num_tangents", C.c_uint64),
                ("texcoords", C.c_void_p), ("num_texcoords", C.c_uint64), ("indices", C.c_void_p),
                ("num_indices", C.c_uint64)]


class DrawableC(C.Structure):
    """bi_pt_drawable: DrawableSbtData + index count, attributes, material, glm matrix."""
    _fields_ = [(k, C.c_uint32) for k in ("drawable_index", "position_offset", "normal_offset", "tangent_offset",
                                          "color_offset", "texcoord_offset", "texcoord2_offset", "index_offset",
                                          "material_offset", "index_count", "attributes", "material")] + \
               [("matrix", C.c_float * 16)]


class SubmeshRangeC(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("position_offset", "normal_offset", "tangent_offset", "texcoord_offset",
                                          "index_offset")] + [("index_count", C.c_uint32), ("num_vertices", C.c_uint32)]


class GpuScene:
    """The engine's GPU-scene view of a scene: flat buffers (numpy, kept alive here) and
    drawable records, for bi_pt_drawables_layout / bi_pt_set_scene_drawables."""

    def __init__(self, positions, normals, tangents, texcoords, indices, drawables):
        self.arrays = [np.ascontiguousarray(a, t) if a is not None else None
                       for a, t in ((positions, np.float32), (normals, np.float32), (tangents, np.float32),
                                    (texcoords, np.float32), (indices, np.uint32))]
        p, n, t, uv, i = self.arrays
        ptr = lambda a: a.ctypes.data if a is not None else None
        size = lambda a: a.size if a is not None else 0
        self.c = GpuSceneC(ptr(p), size(p), ptr(n), size(n), ptr(t), size(t), ptr(uv), size(uv), ptr(i), size(i))
        self.drawables = (DrawableC * max(len(drawables), 1))()
        for k, dr in enumerate(drawables):
            for key, v in dr.items():
                if key == "matrix":
                    self.drawables[k].matrix[:] = [float(x) for x in np.asarray(v, np.float32).reshape(-1)]
                else:
                    setattr(self.drawables[k], key, int(v))
        self.n = len(drawables)

    def layout(self):
        """bi_pt_drawables_layout (host only): ([range dicts], instance transforms [n, 12],
        instance mesh indices, instance materials)."""
        ranges = (SubmeshRangeC * max(self.n, 1))()
        inst = (Instance * max(self.n, 1))()
        nr = C.c_uint32()
        st = lib().bi_pt_drawables_layout(C.byref(self.c), self.drawables, self.n, ranges, C.byref(nr), inst)
        if st != 0:
            raise PtError(st, "drawable records do not fit the GPU-scene buffers")
        out = [{k: getattr(ranges[m], k) for k, _ in SubmeshRangeC._fields_} for m in range(nr.value)]
        xf = np.array([list(inst[k].transform) for k in range(self.n)], np.float32).reshape(self.n, 12)
        return out, xf, np.array([inst[k].mesh for k in range(self.n)]), np.array([inst[k].material for k in range(self.n)])


class TextureC(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("filter", C.c_uint32), ("address_u", C.c_uint32), ("address_v", C.c_uint32)]


class RectTextureC(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("num_levels", C.c_uint32),
                ("format", C.c_uint32), ("mag_filter", C.c_uint32), ("min_filter", C.c_uint32),
                ("mipmap_mode", C.c_uint32), ("address_u", C.c_uint32), ("address_v", C.c_uint32),
                ("lod_bias", C.c_float), ("lod_min", C.c_float), ("lod_max", C.c_float)]


class Instance(C.Structure):
    _fields_ = [("transform", C.c_float * 12), ("mesh", C.c_uint32), ("material", C.c_uint32)]


class CameraC(C.Structure):
    _fields_ = [("matrix_inv_view", C.c_float * 16), ("matrix_inv_proj", C.c_float * 16),
                ("matrix_proj_view", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32)]


class GBufferC(C.Structure):
    """bi_pt_gbuffer: device pointers of the four GBuffer images (nullable)."""
    _fields_ = [("base_color", C.c_void_p), ("normal_roughness", C.c_void_p), ("fresnel", C.c_void_p),
                ("material_0", C.c_void_p)]


class Settings(C.Structure):
    _fields_ = [("ray_length", C.c_float), ("max_bounces", C.c_uint32), ("accumulate", C.c_uint32),
                ("spp", C.c_uint32), ("seed", C.c_uint32), ("flags", C.c_uint32), ("wave_paths", C.c_uint32),
                ("stack_lds", C.c_uint32), ("refill_lanes", C.c_uint32), ("work_parts", C.c_uint32),
                ("block_threads", C.c_uint32)]


class JobC(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("camera", C.POINTER(CameraC)), ("settings", C.POINTER(Settings)),
                ("out_rgba32f", C.c_void_p), ("out_depth", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("wave_paths", C.c_uint32), ("stack_lds", C.c_uint32), ("refill_lanes", C.c_uint32),
                ("work_parts", C.c_uint32), ("use_graph", C.c_uint32), ("block_threads", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerDesc(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_jobs_per_round", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("spp_per_round", C.c_uint32), ("wave_paths", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("scene_id", C.c_uint32), ("_pad", C.c_uint32), ("arrival_s", C.c_double), ("camera", CameraC),
                ("settings", Settings)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_uint32), ("spp_done", C.c_uint32), ("submit_s", C.c_double),
                ("admit_s", C.c_double), ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("rounds", C.c_uint64), ("waiting", C.c_uint32), ("running", C.c_uint32),
                ("hbm_used", C.c_uint64)]


class NrcDesc(C.Structure):
    _fields_ = [("train_samples", C.c_uint32), ("train_steps", C.c_uint32), ("cache_bounce", C.c_uint32),
                ("learning_rate", C.c_float), ("use_after_steps", C.c_uint32), ("seed", C.c_uint32),
                ("flags", C.c_uint32), ("target_clamp", C.c_float)]


class NrcStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("steps", C.c_uint64), ("last_loss", C.c_float), ("_pad", C.c_uint32),
                ("last_valid_samples", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("rays_camera", C.c_uint64), ("rays_extension", C.c_uint64), ("rays_light_shadow", C.c_uint64),
                ("rays_sky_shadow", C.c_uint64), ("frames", C.c_uint64), ("samples_accumulated", C.c_uint64),
                ("ms_build", C.c_float), ("ms_trace_closest", C.c_float), ("ms_shade", C.c_float),
                ("ms_trace_shadow", C.c_float), ("ms_resolve", C.c_float), ("ms_total", C.c_float),
                ("launches_trace_closest", C.c_uint32), ("launches_trace_shadow", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libbi_pt.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run __graft_entry__.build() (make -C bisemutum-engine_amd)")
        L = C.CDLL(LIB_PATH)
        vp, u32, i32, f32p, u32p, i32p = C.c_void_p, C.c_uint32, C.c_int32, C.POINTER(C.c_float), \
            C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
        L.bi_pt_create.argtypes = [C.POINTER(DeviceDesc), C.POINTER(vp)]
        L.bi_pt_destroy.argtypes = [vp]
        L.bi_pt_last_error.argtypes = [vp]
        L.bi_pt_last_error.restype = C.c_char_p
        L.bi_pt_set_scene.argtypes = [vp, C.POINTER(Mesh), u32, C.POINTER(Instance), u32, f32p, u32]
        L.bi_pt_update_transforms.argtypes = [vp, f32p, u32]
        L.bi_pt_set_lights.argtypes = [vp, f32p, u32]
        L.bi_pt_set_sky.argtypes = [vp, f32p, u32, u32, f32p, f32p]
        L.bi_pt_set_point_lights.argtypes = [vp, f32p, u32]
        L.bi_pt_set_rect_lights.argtypes = [vp, f32p, u32]
        L.bi_pt_set_rect_light_textures.argtypes = [vp, C.POINTER(RectTextureC), u32]
        L.bi_pt_rect_texture_mips.argtypes = [u32, u32, u32, u32, vp, vp]
        L.bi_pt_set_ltc_luts.argtypes = [vp, f32p, f32p, f32p, f32p]
        L.bi_pt_set_sky_cube.argtypes = [vp, f32p, u32, f32p, f32p]
        L.bi_pt_render.argtypes = [vp, C.POINTER(CameraC), C.POINTER(Settings), vp, vp, vp]
        L.bi_pt_render_gbuffer.argtypes = [vp, C.POINTER(CameraC), C.POINTER(Settings), vp, vp, C.POINTER(GBufferC), vp]
        L.bi_pt_reset_accumulation.argtypes = [vp]
        L.bi_pt_set_build_options.argtypes = [vp, C.POINTER(C.c_uint32 * 4)]
        L.bi_pt_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.bi_pt_debug_get_bvh.argtypes = [vp, u32p, u32p, i32p, f32p, i32p, f32p]
        L.bi_pt_debug_get_bvh4.argtypes = [vp, i32p, f32p]
        L.bi_pt_debug_num_tris.argtypes = [vp]
        L.bi_pt_debug_num_tris.restype = u32
        L.bi_pt_debug_queue.argtypes = [vp, u32, u32, u32, u32p, u32, u32p]
        L.bi_pt_debug_get_sums.argtypes = [vp, f32p]
        L.bi_pt_debug_trace.argtypes = [vp, u32, f32p, C.c_int, f32p, u32p, f32p, f32p]
        L.bi_pt_batch_create.argtypes = [C.POINTER(DeviceDesc), C.POINTER(vp)]
        L.bi_pt_batch_destroy.argtypes = [vp]
        L.bi_pt_batch_last_error.argtypes = [vp]
        L.bi_pt_batch_last_error.restype = C.c_char_p
        L.bi_pt_batch_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.bi_pt_batch_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.bi_pt_render_batch.argtypes = [vp, C.POINTER(JobC), u32, vp]
        L.bi_pt_batch_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.bi_pt_batch_autotune.argtypes = [vp, C.POINTER(JobC), u32, u32, vp, C.POINTER(Tuning)]
        L.bi_pt_create_view.argtypes = [vp, C.POINTER(vp)]
        L.bi_pt_server_create.argtypes = [C.POINTER(ServerDesc), C.POINTER(vp)]
        L.bi_pt_server_destroy.argtypes = [vp]
        L.bi_pt_server_last_error.argtypes = [vp]
        L.bi_pt_server_last_error.restype = C.c_char_p
        L.bi_pt_server_clock.argtypes = [vp]
        L.bi_pt_server_clock.restype = C.c_double
        L.bi_pt_server_add_scene.argtypes = [vp, vp, u32p]
        L.bi_pt_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.bi_pt_server_step.argtypes = [vp, vp, u32p]
        L.bi_pt_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(JobStatus)]
        L.bi_pt_server_take_result.argtypes = [vp, C.c_uint64, f32p, C.c_uint64]
        L.bi_pt_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.bi_pt_nrc_create.argtypes = [vp, C.POINTER(NrcDesc), C.POINTER(vp)]
        L.bi_pt_nrc_destroy.argtypes = [vp]
        L.bi_pt_nrc_last_error.argtypes = [vp]
        L.bi_pt_nrc_last_error.restype = C.c_char_p
        L.bi_pt_render_nrc.argtypes = [vp, C.POINTER(CameraC), C.POINTER(Settings), vp, vp, vp]
        L.bi_pt_nrc_get_stats.argtypes = [vp, C.POINTER(NrcStats)]
        L.bi_pt_nrc_get_params.argtypes = [vp, f32p, u32]
        L.bi_pt_nrc_set_params.argtypes = [vp, f32p, u32]
        L.bi_pt_nrc_set_bounds.argtypes = [vp, f32p]
        L.bi_pt_nrc_infer.argtypes = [vp, vp, u32, vp, vp]
        L.bi_pt_nrc_train.argtypes = [vp, vp, vp, u32, vp]
        L.bi_pt_nrc_debug_gradient.argtypes = [vp, f32p, u32]
        L.bi_pt_nrc_debug_samples.argtypes = [vp, f32p, f32p, u32]
        L.bi_pt_trace_closest.argtypes = [vp, u32, vp, C.c_float, vp, vp, vp, vp, vp]
        L.bi_pt_trace_occlusion.argtypes = [vp, u32, vp, C.c_float, vp, vp]
        L.bi_pt_set_textures.argtypes = [vp, vp, u32]
        L.bi_pt_drawables_layout.argtypes = [C.POINTER(GpuSceneC), C.POINTER(DrawableC), u32, C.POINTER(SubmeshRangeC),
                                             C.POINTER(C.c_uint32), vp]
        L.bi_pt_set_scene_drawables.argtypes = [vp, C.POINTER(GpuSceneC), C.POINTER(DrawableC), u32, vp, u32]
        L.bi_pt_trace_closest_noalloc.argtypes = [vp, u32, vp, C.c_float, vp, vp, vp, vp, vp]
        L.bi_pt_debug_inject_error.argtypes = [vp, u32]
        L.bi_pt_server_debug_inject_error.argtypes = [vp, u32]
        for name in EXPORTS:
            if name not in ("bi_pt_last_error", "bi_pt_debug_num_tris", "bi_pt_batch_last_error",
                            "bi_pt_server_last_error", "bi_pt_server_clock", "bi_pt_nrc_last_error"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


class Rays(C.Structure):
    """bi_pt_rays: six device arrays (origin xyz, direction xyz)."""
    _fields_ = [("ox", C.c_void_p), ("oy", C.c_void_p), ("oz", C.c_void_p),
                ("dx", C.c_void_p), ("dy", C.c_void_p), ("dz", C.c_void_p)]


def _p(a, t=C.c_float):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def camera_struct(cam) -> CameraC:
    c = CameraC()
    C.memmove(C.addressof(c.matrix_inv_view), np.ascontiguousarray(cam.inv_view, np.float32).ctypes.data, 64)
    C.memmove(C.addressof(c.matrix_inv_proj), np.ascontiguousarray(cam.inv_proj, np.float32).ctypes.data, 64)
    C.memmove(C.addressof(c.matrix_proj_view), np.ascontiguousarray(cam.proj_view, np.float32).ctypes.data, 64)
    c.width, c.height = cam.width, cam.height
    return c


class PathTracer:
    """One bi_pt context on one GPU."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, view_of=None):
        """A context on `device`; with view_of=<PathTracer>, a view of that tracer's scene."""
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("bi_pt needs a ROCm GPU; none is visible")
        self.torch = torch
        self.device = device
        h = C.c_void_p()
        if view_of is not None:
            self.device = view_of.device
            st = lib().bi_pt_create_view(view_of.h, C.byref(h))
            if st != 0:
                raise PtError(st, lib().bi_pt_last_error(view_of.h).decode())
            self._scene_owner = view_of   # the scene context outlives the view
        else:
            st = lib().bi_pt_create(C.byref(DeviceDesc(device, 0, hbm_budget_bytes)), C.byref(h))
            if st != 0:
                raise PtError(st, "bi_pt_create failed")
        self.h = h
        self._out = None
        self._depth = None
        self._n_point_lights = 0

    def _check(self, st):
        if st != 0:
            raise PtError(st, lib().bi_pt_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().bi_pt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- inputs
    def set_scene(self, scene):
        keep = []
        meshes = (Mesh * len(scene.meshes))()
        for i, (p, n, idx) in enumerate(scene.meshes):
            p = np.ascontiguousarray(p, np.float32)
            n = np.ascontiguousarray(n, np.float32)
            idx = np.ascontiguousarray(idx, np.uint32).reshape(-1)
            uv = tan = None
            if scene.mesh_uv is not None and scene.mesh_uv[i] is not None:
                uv = np.ascontiguousarray(scene.mesh_uv[i], np.float32)
            if scene.mesh_tan is not None and scene.mesh_tan[i] is not None:
                tan = np.ascontiguousarray(scene.mesh_tan[i], np.float32)
            keep += [p, n, idx, uv, tan]
            meshes[i] = Mesh(_p(p), _p(n), _p(idx, C.c_uint32), len(p), len(idx),
                             _p(uv) if uv is not None else None, _p(tan) if tan is not None else None)
        self.set_textures(scene.textures)
        inst = (Instance * len(scene.instances))()
        for i, rec in enumerate(scene.instances):
            x = np.ascontiguousarray(rec["xform"], np.float32)
            inst[i].transform[:] = list(map(float, x))
            inst[i].mesh = int(rec["mesh"])
            inst[i].material = int(rec["material"])
        mats = np.ascontiguousarray(scene.materials, np.float32)
        self._check(lib().bi_pt_set_scene(self.h, meshes, len(scene.meshes), inst, len(scene.instances), _p(mats),
                                          len(mats)))
        if len(scene.dir_lights) + self._n_point_lights > 4:
            self.set_point_lights(None)   # the light list holds four lights, directional and point together
        self.set_lights(scene.dir_lights)
        self.set_point_lights(getattr(scene, "point_lights", None))
        if getattr(scene, "ltc", None) is not None:
            self.set_ltc_luts(*scene.ltc)
        self.set_rect_light_textures(getattr(scene, "rect_textures", None) or [])
        self.set_rect_lights(getattr(scene, "rect_lights", None))
        if getattr(scene, "sky_layout", 0) == 1:
            self.set_sky_cube(scene.sky_rgba, scene.sky_color, scene.sky_rot)
        else:
            self.set_sky(scene.sky_rgba, scene.sky_color, scene.sky_rot)

    def set_scene_drawables(self, gpu_scene, materials, lights, sky_rgba, sky_color, sky_rot, textures=()):
        """bi_pt_set_scene_drawables from a GpuScene, then lights, sky and textures."""
        self.set_textures(list(textures))
        mats = np.ascontiguousarray(materials, np.float32)
        self._check(lib().bi_pt_set_scene_drawables(self.h, C.byref(gpu_scene.c), gpu_scene.drawables, gpu_scene.n,
                                                    _p(mats), len(mats)))
        self.set_lights(lights)
        self.set_sky(sky_rgba, sky_color, sky_rot)

    def set_textures(self, textures):
        """bi_pt_set_textures: scenes.Texture records (texels [H, W, 4] uint8 or float32)."""
        arr = (TextureC * max(len(textures), 1))()
        keep = []
        for i, t in enumerate(textures):
            a = np.ascontiguousarray(t.texels)
            keep.append(a)
            arr[i] = TextureC(a.ctypes.data, a.shape[1], a.shape[0], t.format, t.filter, t.address_u, t.address_v)
        self._check(lib().bi_pt_set_textures(self.h, arr, len(textures)))

    def update_transforms(self, xforms):
        x = np.ascontiguousarray(xforms, np.float32)
        self._check(lib().bi_pt_update_transforms(self.h, _p(x), len(x)))

    def set_point_lights(self, lights):
        """bi_pt_set_point_lights: records [P, 16] (scenes.point_light), or None."""
        lights = np.ascontiguousarray(lights if lights is not None else np.zeros((0, 16)), np.float32).reshape(-1, 16)
        self._check(lib().bi_pt_set_point_lights(self.h, _p(lights) if len(lights) else None, len(lights)))
        self._n_point_lights = len(lights)

    def set_rect_lights(self, lights):
        """bi_pt_set_rect_lights: records [R, 28] (scenes.rect_light), or None."""
        lights = np.ascontiguousarray(lights if lights is not None else np.zeros((0, 28)), np.float32).reshape(-1, 28)
        self._check(lib().bi_pt_set_rect_lights(self.h, _p(lights) if len(lights) else None, len(lights)))

    def set_rect_light_textures(self, textures):
        """bi_pt_set_rect_light_textures: scenes.RectTexture records."""
        arr = (RectTextureC * max(len(textures), 1))()
        keep = []
        for i, t in enumerate(textures):
            a = t.chain()
            keep.append(a)
            h, w = t.levels[0].shape[:2]
            arr[i] = RectTextureC(a.ctypes.data, w, h, len(t.levels), t.format, t.mag_filter, t.min_filter,
                                  t.mipmap_mode, t.address_u, t.address_v, t.lod_bias, t.lod_min, t.lod_max)
        self._check(lib().bi_pt_set_rect_light_textures(self.h, arr, len(textures)))

    def set_ltc_luts(self, m0, m1, m2, norm):
        """bi_pt_set_ltc_luts: three [4096, 4] matrix tables and the [4096, 2] norm table."""
        t = [np.ascontiguousarray(x, np.float32) for x in (m0, m1, m2, norm)]
        if any(x.size != 4096 * 4 for x in t[:3]) or t[3].size != 4096 * 2:
            raise ValueError("LTC tables are 8 x 8 x 64 texels")
        self._check(lib().bi_pt_set_ltc_luts(self.h, *[_p(x) for x in t]))

    def set_lights(self, lights):
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 16)
        self._check(lib().bi_pt_set_lights(self.h, _p(lights), len(lights)))

    def set_sky(self, rgba, color, rot):
        rgba = np.ascontiguousarray(rgba, np.float32)
        col = np.ascontiguousarray(color, np.float32)
        rot = np.ascontiguousarray(rot, np.float32)
        h, w = (rgba.shape[0], rgba.shape[1]) if rgba.size else (0, 0)
        self._check(lib().bi_pt_set_sky(self.h, _p(rgba) if rgba.size else None, w, h, _p(col), _p(rot)))

    def set_sky_cube(self, faces, color, rot):
        """bi_pt_set_sky_cube: faces f32 [6, S, S, 4] in the order +X, -X, +Y, -Y, +Z, -Z."""
        faces = np.ascontiguousarray(faces, np.float32)
        if faces.size and (faces.ndim != 4 or faces.shape[0] != 6 or faces.shape[1] != faces.shape[2]
                           or faces.shape[3] != 4):
            raise ValueError(f"cube faces must be [6, S, S, 4], got {faces.shape}")
        col = np.ascontiguousarray(color, np.float32)
        rot = np.ascontiguousarray(rot, np.float32)
        s = faces.shape[1] if faces.size else 0
        self._check(lib().bi_pt_set_sky_cube(self.h, _p(faces) if faces.size else None, s, _p(col), _p(rot)))

    # ---------------------------------------------------------------- render
    def render(self, cam, spp=1, max_bounces=5, seed=0, flags=0, ray_length=1000.0, accumulate=False,
               wave_paths=0, stack_lds=0, refill_lanes=0, work_parts=0, block_threads=0, depth=False, stream=None):
        torch = self.torch
        npix = cam.width * cam.height
        dev = torch.device("cuda", self.device)
        dtype = torch.float16 if flags & FLAG_OUTPUT_RGBA16F else torch.float32
        if self._out is None or self._out.numel() != npix * 4 or self._out.dtype != dtype:
            self._out = torch.empty(npix * 4, dtype=dtype, device=dev)
        dptr = None
        if depth:
            if self._depth is None or self._depth.numel() != npix:
                self._depth = torch.empty(npix, dtype=torch.float32, device=dev)
            dptr = C.c_void_p(self._depth.data_ptr())
        st = Settings(ray_length, max_bounces, 1 if accumulate else 0, spp, seed, flags, wave_paths, stack_lds,
                      refill_lanes, work_parts, block_threads)
        s = C.c_void_p(stream) if stream else None
        self._check(lib().bi_pt_render(self.h, C.byref(camera_struct(cam)), C.byref(st),
                                       C.c_void_p(self._out.data_ptr()), dptr, s))
        return self._out

    def render_gbuffer(self, cam, spp=1, max_bounces=5, seed=0, flags=0, ray_length=1000.0, accumulate=False,
                       parts=("base_color", "normal_roughness", "fresnel", "material_0"), stream=None):
        """bi_pt_render_gbuffer: render as render() and write the primary-hit GBuffer of the
        call's first sample.  Returns {part: numpy array [npix, 4]} (uint16 half bits for
        base_color / normal_roughness, uint16 unorm for fresnel, uint8 for material_0)."""
        torch = self.torch
        npix = cam.width * cam.height
        dev = torch.device("cuda", self.device)
        dtype = torch.float16 if flags & FLAG_OUTPUT_RGBA16F else torch.float32
        if self._out is None or self._out.numel() != npix * 4 or self._out.dtype != dtype:
            self._out = torch.empty(npix * 4, dtype=dtype, device=dev)
        bufs = {}
        for name in parts:
            bufs[name] = torch.full((npix * 4,), 0x5A, dtype=torch.uint8 if name == "material_0" else torch.int16,
                                    device=dev)
        g = GBufferC(*[C.c_void_p(bufs[k].data_ptr()) if k in bufs else None
                       for k in ("base_color", "normal_roughness", "fresnel", "material_0")])
        st = Settings(ray_length, max_bounces, 1 if accumulate else 0, spp, seed, flags, 0, 0, 0, 0, 0)
        s = C.c_void_p(stream) if stream else None
        self._check(lib().bi_pt_render_gbuffer(self.h, C.byref(camera_struct(cam)), C.byref(st),
                                               C.c_void_p(self._out.data_ptr()), None, C.byref(g), s))
        torch.cuda.synchronize()
        out = {}
        for k, t in bufs.items():
            a = t.cpu().numpy().reshape(npix, 4)
            out[k] = a if k == "material_0" else a.view(np.uint16)
        return out

    def image(self):
        self.torch.cuda.synchronize()
        return self._out.cpu().numpy().reshape(-1, 4)

    def depth_image(self):
        self.torch.cuda.synchronize()
        return self._depth.cpu().numpy()

    def reset(self):
        self._check(lib().bi_pt_reset_accumulation(self.h))

    def inject_error(self, flags):
        """Raise device error flags from the host (reported by the next stats())."""
        self._check(lib().bi_pt_debug_inject_error(self.h, flags))

    def set_build_options(self, treelet_rounds=3):
        """Rounds of treelet restructuring after the LBVH (0: the Morton/Karras tree)."""
        opt = (C.c_uint32 * 4)(treelet_rounds, 0, 0, 0)
        self._check(lib().bi_pt_set_build_options(self.h, C.byref(opt)))

    def stats(self):
        s = Stats()
        self._check(lib().bi_pt_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    # ---------------------------------------------------------------- inspection
    def num_tris(self):
        return lib().bi_pt_debug_num_tris(self.h)

    def bvh(self):
        n = self.num_tris()
        ni = max(n - 1, 1)
        morton = np.zeros(n, np.uint32)
        order = np.zeros(n, np.uint32)
        children = np.zeros((ni, 2), np.int32)
        bounds = np.zeros((ni, 12), np.float32)
        parents = np.zeros(ni + n, np.int32)
        cb = np.zeros(6, np.float32)
        self._check(lib().bi_pt_debug_get_bvh(self.h, _p(morton, C.c_uint32), _p(order, C.c_uint32),
                                              _p(children, C.c_int32), _p(bounds), _p(parents, C.c_int32), _p(cb)))
        return dict(morton=morton, order=order, children=children, bounds=bounds, parents=parents,
                    centroid_bounds=cb)

    def bvh4(self):
        ni = max(self.num_tris() - 1, 1)
        child4 = np.zeros((ni, 4), np.int32)
        box4 = np.zeros((ni, 24), np.float32)
        self._check(lib().bi_pt_debug_get_bvh4(self.h, _p(child4, C.c_int32), _p(box4)))
        return child4, box4

    def queue(self, wave, bounce, kind):
        cnt = C.c_uint32()
        self._check(lib().bi_pt_debug_queue(self.h, wave, bounce, kind, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, np.uint32)
        self._check(lib().bi_pt_debug_queue(self.h, wave, bounce, kind, _p(out, C.c_uint32), cnt.value,
                                            C.byref(cnt)))
        return out

    def sums(self, npix):
        out = np.zeros((npix, 3), np.float32)
        self._check(lib().bi_pt_debug_get_sums(self.h, _p(out)))
        return out

    def _query_rays(self, origins, dirs):
        torch = self.torch
        dev = torch.device("cuda", self.device)
        o = torch.as_tensor(np.ascontiguousarray(origins, np.float32)).to(dev)
        d = torch.as_tensor(np.ascontiguousarray(dirs, np.float32)).to(dev)
        cols = [o[:, k].contiguous() for k in range(3)] + [d[:, k].contiguous() for k in range(3)]
        r = Rays(*[C.c_void_p(c.data_ptr()) for c in cols])
        return cols, r, dev

    def trace_closest(self, origins, dirs, tmax=1000.0, stream=None):
        """Closest hits of n rays (origins, dirs: (n, 3)) through bi_pt_trace_closest on
        device arrays: (t, triangle, u, v) as numpy arrays (t = inf, triangle = 0xffffffff
        on a miss)."""
        torch = self.torch
        cols, r, dev = self._query_rays(origins, dirs)
        n = cols[0].numel()
        t = torch.empty(n, dtype=torch.float32, device=dev)
        tri = torch.empty(n, dtype=torch.int32, device=dev)
        u = torch.empty(n, dtype=torch.float32, device=dev)
        v = torch.empty(n, dtype=torch.float32, device=dev)
        s = C.c_void_p(stream) if stream else None
        self._check(lib().bi_pt_trace_closest(self.h, n, C.byref(r), tmax, C.c_void_p(t.data_ptr()),
                                              C.c_void_p(tri.data_ptr()), C.c_void_p(u.data_ptr()),
                                              C.c_void_p(v.data_ptr()), s))
        torch.cuda.synchronize()
        return t.cpu().numpy(), tri.cpu().numpy().view(np.uint32), u.cpu().numpy(), v.cpu().numpy()

    def trace_closest_noalloc(self, origins, dirs, tmax=1000.0, stream=None):
        """trace_closest through bi_pt_trace_closest_noalloc (the library's static scratch;
        the scene's LBVH must be current)."""
        torch = self.torch
        cols, r, dev = self._query_rays(origins, dirs)
        n = cols[0].numel()
        t = torch.empty(n, dtype=torch.float32, device=dev)
        tri = torch.empty(n, dtype=torch.int32, device=dev)
        u = torch.empty(n, dtype=torch.float32, device=dev)
        v = torch.empty(n, dtype=torch.float32, device=dev)
        s = C.c_void_p(stream) if stream else None
        self._check(lib().bi_pt_trace_closest_noalloc(self.h, n, C.byref(r), tmax, C.c_void_p(t.data_ptr()),
                                                      C.c_void_p(tri.data_ptr()), C.c_void_p(u.data_ptr()),
                                                      C.c_void_p(v.data_ptr()), s))
        torch.cuda.synchronize()
        return t.cpu().numpy(), tri.cpu().numpy().view(np.uint32), u.cpu().numpy(), v.cpu().numpy()

    def trace_occlusion(self, origins, dirs, tmax=1000.0, stream=None):
        """1 where a ray hits any triangle in [0.001, tmax], else 0 (bi_pt_trace_occlusion)."""
        torch = self.torch
        cols, r, dev = self._query_rays(origins, dirs)
        n = cols[0].numel()
        occ = torch.empty(n, dtype=torch.int32, device=dev)
        s = C.c_void_p(stream) if stream else None
        self._check(lib().bi_pt_trace_occlusion(self.h, n, C.byref(r), tmax, C.c_void_p(occ.data_ptr()), s))
        torch.cuda.synchronize()
        return occ.cpu().numpy().view(np.uint32)

    def trace(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        t = np.zeros(n, np.float32)
        leaf = np.zeros(n, np.uint32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        self._check(lib().bi_pt_debug_trace(self.h, n, _p(rays), int(any_hit), _p(t), _p(leaf, C.c_uint32),
                                            _p(u), _p(v)))
        return (leaf,) if any_hit else (t, leaf, u, v)


class Job:
    """One render job of a batch: a PathTracer (its scene, lights, sky and running
    image), a camera and the settings of bi_pt_render."""

    def __init__(self, tracer, cam, spp=1, max_bounces=5, seed=0, flags=0, ray_length=1000.0, accumulate=False,
                 depth=False):
        self.tracer = tracer
        self.cam = cam
        self.settings = Settings(ray_length, max_bounces, 1 if accumulate else 0, spp, seed, flags, 0, 0, 0, 0)
        self.depth = depth

    def _buffers(self):
        t = self.tracer
        torch = t.torch
        npix = self.cam.width * self.cam.height
        dev = torch.device("cuda", t.device)
        dtype = torch.float16 if self.settings.flags & FLAG_OUTPUT_RGBA16F else torch.float32
        if t._out is None or t._out.numel() != npix * 4 or t._out.dtype != dtype:
            t._out = torch.empty(npix * 4, dtype=dtype, device=dev)
        if self.depth and (t._depth is None or t._depth.numel() != npix):
            t._depth = torch.empty(npix, dtype=torch.float32, device=dev)
        return t._out, (t._depth if self.depth else None)


class Batch:
    """bi_pt_batch: renders the jobs of several PathTracers in shared waves."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("bi_pt needs a ROCm GPU; none is visible")
        self.device = device
        h = C.c_void_p()
        st = lib().bi_pt_batch_create(C.byref(DeviceDesc(device, 0, hbm_budget_bytes)), C.byref(h))
        if st != 0:
            raise PtError(st, "bi_pt_batch_create failed")
        self.h = h

    def _check(self, st):
        if st != 0:
            raise PtError(st, lib().bi_pt_batch_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().bi_pt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _jobs(self, jobs):
        arr = (JobC * len(jobs))()
        keep = []
        for i, j in enumerate(jobs):
            out, depth = j._buffers()
            cam = camera_struct(j.cam)
            keep.append(cam)
            arr[i] = JobC(j.tracer.h.value, C.pointer(cam), C.pointer(j.settings), out.data_ptr(),
                          depth.data_ptr() if depth is not None else None)
        return arr, keep

    def render(self, jobs, stream=None):
        arr, keep = self._jobs(jobs)
        self._check(lib().bi_pt_render_batch(self.h, arr, len(jobs), C.c_void_p(stream) if stream else None))

    def set_tuning(self, wave_paths=0, stack_lds=0, refill_lanes=0, work_parts=0, use_graph=False, block_threads=0):
        t = Tuning(wave_paths, stack_lds, refill_lanes, work_parts, 1 if use_graph else 0, block_threads)
        self._check(lib().bi_pt_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = Tuning()
        self._check(lib().bi_pt_batch_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def autotune(self, jobs, rounds=1, stream=None):
        arr, keep = self._jobs(jobs)
        t = Tuning()
        self._check(lib().bi_pt_batch_autotune(self.h, arr, len(jobs), rounds,
                                               C.c_void_p(stream) if stream else None, C.byref(t)))
        return t.as_dict()

    def stats(self):
        s = Stats()
        self._check(lib().bi_pt_batch_get_stats(self.h, C.byref(s)))
        return s.as_dict()


class Server:
    """bi_pt_server: the serving loop (admission by HBM budget, continuous batching)."""

    def __init__(self, device=0, hbm_budget_bytes=0, max_jobs_per_round=0, spp_per_round=0, wave_paths=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("bi_pt needs a ROCm GPU; none is visible")
        h = C.c_void_p()
        st = lib().bi_pt_server_create(C.byref(ServerDesc(device, max_jobs_per_round, hbm_budget_bytes,
                                                          spp_per_round, wave_paths)), C.byref(h))
        if st != 0:
            raise PtError(st, "bi_pt_server_create failed")
        self.h = h
        self.scenes = []

    def _check(self, st):
        if st != 0:
            raise PtError(st, lib().bi_pt_server_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().bi_pt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self):
        return lib().bi_pt_server_clock(self.h)

    def add_scene(self, tracer):
        sid = C.c_uint32()
        self._check(lib().bi_pt_server_add_scene(self.h, tracer.h, C.byref(sid)))
        self.scenes.append(tracer)
        return sid.value

    def submit(self, scene_id, cam, spp=1, max_bounces=5, seed=0, flags=0, ray_length=1000.0, arrival_s=-1.0):
        r = Request(scene_id, 0, arrival_s, camera_struct(cam),
                    Settings(ray_length, max_bounces, 0, spp, seed, flags, 0, 0, 0, 0))
        jid = C.c_uint64()
        self._check(lib().bi_pt_server_submit(self.h, C.byref(r), C.byref(jid)))
        return jid.value

    def step(self, stream=None):
        act = C.c_uint32()
        self._check(lib().bi_pt_server_step(self.h, C.c_void_p(stream) if stream else None, C.byref(act)))
        return act.value

    def poll(self, job_id):
        s = JobStatus()
        self._check(lib().bi_pt_server_poll(self.h, job_id, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def take_result(self, job_id, npix):
        """The finished job's RGBA32F image ([npix, 4]); npix None discards it."""
        if npix is None:
            self._check(lib().bi_pt_server_take_result(self.h, job_id, None, 0))
            return None
        out = np.zeros((npix, 4), np.float32)
        self._check(lib().bi_pt_server_take_result(self.h, job_id, _p(out), out.size))
        return out

    def inject_error(self, flags):
        """Raise device error flags from the host: the next round fails its jobs."""
        self._check(lib().bi_pt_server_debug_inject_error(self.h, flags))

    def stats(self):
        s = ServerStats()
        self._check(lib().bi_pt_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}


class Nrc:
    """bi_pt_nrc: the neural radiance cache of one PathTracer's scene."""

    def __init__(self, tracer, train_samples=0, train_steps=0, cache_bounce=0, learning_rate=0.0,
                 use_after_steps=0, seed=0, flags=0, target_clamp=0.0):
        self.tracer = tracer
        self.torch = tracer.torch
        h = C.c_void_p()
        st = lib().bi_pt_nrc_create(tracer.h, C.byref(NrcDesc(train_samples, train_steps, cache_bounce,
                                                                learning_rate, use_after_steps, seed, flags,
                                                                target_clamp)),
                                    C.byref(h))
        if st != 0:
            raise PtError(st, "bi_pt_nrc_create failed")
        self.h = h

    def _check(self, st):
        if st != 0:
            raise PtError(st, lib().bi_pt_nrc_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().bi_pt_nrc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, cam, spp=1, max_bounces=5, seed=0, flags=0, ray_length=1000.0, accumulate=False,
               wave_paths=0, stack_lds=0, refill_lanes=0, work_parts=0, block_threads=0, stream=None):
        """bi_pt_render_nrc into the tracer's output buffer (tracer.image() reads it)."""
        t = self.tracer
        torch = self.torch
        npix = cam.width * cam.height
        if t._out is None or t._out.numel() != npix * 4:
            t._out = torch.empty(npix * 4, dtype=torch.float32, device=torch.device("cuda", t.device))
        st = Settings(ray_length, max_bounces, 1 if accumulate else 0, spp, seed, flags, wave_paths, stack_lds,
                      refill_lanes, work_parts, block_threads)
        self._check(lib().bi_pt_render_nrc(self.h, C.byref(camera_struct(cam)), C.byref(st),
                                           C.c_void_p(t._out.data_ptr()), None,
                                           C.c_void_p(stream) if stream else None))
        return t._out

    def stats(self):
        s = NrcStats()
        self._check(lib().bi_pt_nrc_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_ if k != "_pad"}

    def params(self):
        out = np.zeros(NRC_PARAMS, np.float32)
        self._check(lib().bi_pt_nrc_get_params(self.h, _p(out), NRC_PARAMS))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(lib().bi_pt_nrc_set_params(self.h, _p(w), w.size))

    def set_bounds(self, lo, hi):
        b = np.ascontiguousarray(list(lo) + list(hi), np.float32)
        self._check(lib().bi_pt_nrc_set_bounds(self.h, _p(b)))

    def infer(self, records):
        """Raw outputs [n, 3] for device records (torch float32 [n, 16])."""
        torch = self.torch
        out = torch.empty((records.shape[0], 3), dtype=torch.float32, device=records.device)
        self._check(lib().bi_pt_nrc_infer(self.h, C.c_void_p(records.data_ptr()), records.shape[0],
                                          C.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        return out

    def train(self, records, targets):
        """One Adam step on device records [n, 16] and targets [n, 3] (torch float32)."""
        self._check(lib().bi_pt_nrc_train(self.h, C.c_void_p(records.data_ptr()), C.c_void_p(targets.data_ptr()),
                                          records.shape[0], None))

    def gradient(self):
        out = np.zeros(NRC_PARAMS, np.float32)
        self._check(lib().bi_pt_nrc_debug_gradient(self.h, _p(out), NRC_PARAMS))
        return out

    def samples(self, n):
        """The last frame's first n training records [n, 16] and targets [n, 3]."""
        rec = np.zeros((n, 16), np.float32)
        tgt = np.zeros((n, 3), np.float32)
        self._check(lib().bi_pt_nrc_debug_samples(self.h, _p(rec), _p(tgt), n))
        return rec, tgt


def nrc_train_seed(seed, frame):
    """Seed of the training paths of frame `frame` of bi_pt_render_nrc with render seed `seed`."""
    return ((seed * 0x9E3779B1) & 0xFFFFFFFF) ^ ((frame * 0x85EBCA77 + 0x7A1B) & 0xFFFFFFFF)


def rect_texture_mips(level0, num_levels, fmt):
    """bi_pt_rect_texture_mips (host only): the engine's mip chain of level0 ([h, w, 4], uint8
    for the RGBA8 formats, float32 for RGBA32F) as a list of num_levels [h_l, w_l, 4] arrays."""
    a = np.ascontiguousarray(level0)
    h, w = a.shape[:2]
    sizes, lw, lh = [], w, h
    for _ in range(num_levels):
        sizes.append((lh, lw))
        lw, lh = max(lw // 2, 1), max(lh // 2, 1)
    out = np.zeros(sum(y * x * 4 for y, x in sizes), a.dtype)
    st = lib().bi_pt_rect_texture_mips(fmt, w, h, num_levels, a.ctypes.data, out.ctypes.data)
    if st != 0:
        raise PtError(st, "bi_pt_rect_texture_mips: bad format, size or level count")
    levels, o = [], 0
    for y, x in sizes:
        levels.append(out[o:o + y * x * 4].reshape(y, x, 4))
        o += y * x * 4
    return levels
