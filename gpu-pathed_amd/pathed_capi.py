"""ctypes binding of libpathed.so (include/pathed.h).

Loads the in-tree library next to this file (or the one PATHED_LIB names) and fails loudly
when it is missing: there is no CPU fallback for any GPU entry point.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PATHED_LIB") or os.path.join(HERE, "libpathed.so")

PT_OK = 0
ERR_INVALID, ERR_DEVICE, ERR_MEMORY, ERR_STATE, ERR_NUMERIC, ERR_IO = -1, -2, -3, -4, -5, -6
LAMBERTIAN, MIRROR, GLASS, MICROFACET = 0, 1, 2, 3
RENDERER_CUDA, RENDERER_OPTIX, RENDERER_NORMALS, RENDERER_GUIDED = 0, 1, 2, 3
OPT_POOL_PATHS, OPT_TRACE_BLOCKS_PER_CU, OPT_SHADE_BLOCKS_PER_CU, OPT_PROFILE_KERNELS = 0, 1, 2, 3
OPT_CONSISTENT_MIS, OPT_GUIDE_TRAIN_STEPS, OPT_GUIDE_BATCH, OPT_GUIDE_LR_MICRO, OPT_GUIDE_RESET = 4, 5, 6, 7, 8
OPT_TRACE_LDS_STACK, OPT_BATCH_GRAPH, OPT_GUIDE_MAX_FRACTION_MILLI, OPT_TRACE_REFILL_LANES, OPT_GUIDE_BOUNCES = 9, 10, 11, 12, 13
OPT_BVH_BUILDER, OPT_BATCH_WAVEFRONT, OPT_TRACE_STATS = 14, 15, 16
BVH_LBVH, BVH_PLOC = 1, 2
MISS = np.iinfo(np.int32).min

# every symbol declared in include/pathed.h
EXPORTS = [
    "pt_last_error", "pt_abi_version", "pt_device_count",
    "pt_scene_create", "pt_scene_destroy", "pt_scene_add_material", "pt_scene_add_triangles", "pt_scene_add_spheres",
    "pt_scene_set_environment", "pt_scene_set_camera", "pt_scene_preset", "pt_scene_get_info", "pt_scene_export",
    "pt_renderer_create", "pt_renderer_destroy", "pt_renderer_set_scene", "pt_renderer_set_camera", "pt_renderer_set_material",
    "pt_renderer_set_max_depth", "pt_renderer_set_nee", "pt_renderer_set_type", "pt_renderer_set_seed", "pt_renderer_set_option",
    "pt_renderer_get_option", "pt_renderer_render_async", "pt_renderer_poll", "pt_renderer_synchronize", "pt_renderer_reset",
    "pt_renderer_get_spp", "pt_renderer_get_radiance", "pt_renderer_get_radiance_sum", "pt_renderer_get_rgba8",
    "pt_renderer_hit_test", "pt_renderer_get_stats", "pt_renderer_reset_stats", "pt_renderer_get_kernel_stats",
    "pt_renderer_get_path_radiance", "pt_renderer_intersect",
    "pt_unit_triangle_test", "pt_unit_bsdf_eval", "pt_unit_bsdf_sample", "pt_unit_env_tables", "pt_unit_env_sample",
    "pt_unit_path_uniforms", "pt_unit_guide_forward", "pt_unit_guide_records", "pt_unit_guide_gradients", "pt_unit_guide_adam", "pt_unit_mfma_tile",
    "pt_unit_vmf_sample", "pt_unit_tonemap", "pt_pool_layout",
    "pt_scene_load_obj", "pt_scene_load_ply", "pt_scene_load_exr_environment", "pt_scene_reference", "pt_exr_info",
    "pt_exr_read_rgba", "pt_exr_write_rgb",
    "pt_renderer_autotune", "pt_batch_create", "pt_batch_destroy", "pt_batch_add", "pt_batch_set_option", "pt_batch_get_option",
    "pt_batch_autotune", "pt_batch_run", "pt_batch_get_radiance", "pt_batch_get_render_stats", "pt_batch_get_stats",
    "pt_env_tables_noalloc", "pt_renderer_get_stream", "pt_unit_mfma_tile16",
]
GUIDE_PARAMS = 20800


class PtError(RuntimeError):
    pass


class SceneInfo(C.Structure):
    _fields_ = [
        ("n_triangles", C.c_int64), ("n_spheres", C.c_int64), ("n_materials", C.c_int64), ("n_emitters", C.c_int64),
        ("env_width", C.c_int), ("env_height", C.c_int), ("has_normals", C.c_int),
        ("camera_origin", C.c_float * 3), ("camera_target", C.c_float * 3), ("camera_up", C.c_float * 3),
        ("camera_vfov", C.c_float), ("camera_flip", C.c_int),
        ("bounds_lo", C.c_float * 3), ("bounds_hi", C.c_float * 3),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("rays_total", C.c_int64), ("extension_rays", C.c_int64), ("shadow_rays", C.c_int64), ("paths", C.c_int64),
        ("passes", C.c_int64), ("gpu_ms", C.c_double), ("bvh_depth", C.c_int), ("bvh_nodes", C.c_int64),
        ("stack_overflows", C.c_int), ("guide_steps", C.c_int64), ("guide_loss", C.c_double),
        ("bounce_rays", C.c_int64 * 16), ("bounce_shadow", C.c_int64 * 16),
        ("trace_rays", C.c_int64 * 2), ("trace_nodes", C.c_int64 * 2), ("trace_prims", C.c_int64 * 2),
    ]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 16), ("launches", C.c_int64), ("units", C.c_int64), ("ms", C.c_double), ("bytes", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("pool_paths", C.c_int64), ("trace_blocks_per_cu", C.c_int), ("shade_blocks_per_cu", C.c_int),
                ("lds_stack", C.c_int), ("mrays_per_s", C.c_double), ("probes", C.c_int)]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchRenderStats(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("triangles", C.c_int64), ("rays", C.c_int64),
                ("latency_ms", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("renders", C.c_int), ("chunks", C.c_int), ("graph_nodes", C.c_int64), ("rays", C.c_int64), ("wall_ms", C.c_double),
                ("stack_overflows", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PtError(f"libpathed.so not built at {LIB_PATH}; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        L = _lib
        vp = C.c_void_p
        L.pt_last_error.restype = C.c_char_p
        for name in EXPORTS:
            if name != "pt_last_error":
                getattr(L, name).restype = C.c_int
        L.pt_scene_create.argtypes = [C.POINTER(vp)]
        L.pt_scene_destroy.argtypes = [vp]
        L.pt_scene_add_material.argtypes = [vp, C.c_int, vp, vp, C.c_float, C.c_float, C.POINTER(C.c_int)]
        L.pt_scene_add_triangles.argtypes = [vp, vp, vp, vp, C.c_int64]
        L.pt_scene_add_spheres.argtypes = [vp, vp, vp, vp, C.c_int64]
        L.pt_scene_set_environment.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int]
        L.pt_scene_set_camera.argtypes = [vp, vp, vp, vp, C.c_float, C.c_int]
        L.pt_scene_preset.argtypes = [vp, C.c_int, C.c_uint64, C.c_int64]
        L.pt_scene_get_info.argtypes = [vp, C.POINTER(SceneInfo)]
        L.pt_scene_export.argtypes = [vp] + [vp] * 11
        L.pt_renderer_create.argtypes = [C.c_int, C.c_int, C.POINTER(vp)]
        L.pt_renderer_destroy.argtypes = [vp]
        L.pt_renderer_set_scene.argtypes = [vp, vp]
        L.pt_renderer_set_camera.argtypes = [vp, vp, vp, vp, C.c_float, C.c_int]
        L.pt_renderer_set_material.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_float, C.c_float]
        L.pt_renderer_set_max_depth.argtypes = [vp, C.c_int]
        L.pt_renderer_set_nee.argtypes = [vp, C.c_int]
        L.pt_renderer_set_type.argtypes = [vp, C.c_int]
        L.pt_renderer_set_seed.argtypes = [vp, C.c_uint64]
        L.pt_renderer_set_option.argtypes = [vp, C.c_int, C.c_int64]
        L.pt_renderer_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.pt_renderer_render_async.argtypes = [vp, C.c_int]
        L.pt_renderer_poll.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.pt_renderer_synchronize.argtypes = [vp]
        L.pt_renderer_reset.argtypes = [vp]
        L.pt_renderer_get_spp.argtypes = [vp, C.POINTER(C.c_int64)]
        L.pt_renderer_get_radiance.argtypes = [vp, vp]
        L.pt_renderer_get_radiance_sum.argtypes = [vp, vp]
        L.pt_renderer_get_rgba8.argtypes = [vp, vp]
        L.pt_renderer_hit_test.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pt_renderer_get_stats.argtypes = [vp, C.POINTER(RenderStats)]
        L.pt_renderer_reset_stats.argtypes = [vp]
        L.pt_renderer_get_kernel_stats.argtypes = [vp, C.POINTER(KernelStat), C.c_int, C.POINTER(C.c_int)]
        L.pt_renderer_get_path_radiance.argtypes = [vp, vp, C.c_int64, C.POINTER(C.c_int64)]
        L.pt_renderer_intersect.argtypes = [vp, vp, C.c_int64, C.c_int, vp]
        L.pt_renderer_autotune.argtypes = [vp, C.c_int, C.POINTER(TuneResult)]
        L.pt_scene_load_obj.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(C.c_int)]
        L.pt_scene_load_ply.argtypes = [vp, C.c_char_p, C.c_int]
        L.pt_scene_load_exr_environment.argtypes = [vp, C.c_char_p, C.c_int]
        L.pt_scene_reference.argtypes = [vp, C.c_int, C.c_char_p]
        L.pt_exr_info.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pt_exr_read_rgba.argtypes = [C.c_char_p, vp, C.c_int64]
        L.pt_exr_write_rgb.argtypes = [C.c_char_p, vp, C.c_int, C.c_int, C.c_int]
        L.pt_batch_create.argtypes = [C.POINTER(vp)]
        L.pt_batch_destroy.argtypes = [vp]
        L.pt_batch_add.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.pt_batch_set_option.argtypes = [vp, C.c_int, C.c_int64]
        L.pt_batch_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.pt_batch_autotune.argtypes = [vp, C.c_int, C.POINTER(TuneResult)]
        L.pt_batch_run.argtypes = [vp]
        L.pt_batch_get_radiance.argtypes = [vp, C.c_int, vp]
        L.pt_batch_get_render_stats.argtypes = [vp, C.c_int, C.POINTER(BatchRenderStats)]
        L.pt_batch_get_stats.argtypes = [vp, C.POINTER(BatchStats)]
        L.pt_unit_triangle_test.argtypes = [vp, vp, C.c_int64, C.c_int, vp]
        L.pt_unit_bsdf_eval.argtypes = [vp, vp, vp, C.c_int64, vp, vp]
        L.pt_unit_bsdf_sample.argtypes = [vp, vp, vp, C.c_int64, vp]
        L.pt_unit_env_tables.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.pt_env_tables_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
        L.pt_unit_env_sample.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int64, vp, vp, vp, vp, C.c_int64, vp]
        L.pt_unit_path_uniforms.argtypes = [C.c_uint64, vp, vp, C.c_int64, C.c_int, vp]
        L.pt_device_count.argtypes = [C.POINTER(C.c_int)]
        L.pt_unit_guide_forward.argtypes = [vp, vp, C.c_int64, vp]
        L.pt_unit_guide_records.argtypes = [vp, vp, C.c_int64, vp]
        L.pt_unit_guide_gradients.argtypes = [vp, vp, C.c_int64, vp, vp]
        L.pt_unit_guide_adam.argtypes = [vp, vp, vp, vp, C.c_int, C.c_float]
        L.pt_unit_mfma_tile.argtypes = [vp, vp, vp]
        L.pt_unit_mfma_tile16.argtypes = [vp, vp, vp]
        L.pt_unit_vmf_sample.argtypes = [vp, vp, vp, C.c_int64, vp, vp]
        L.pt_unit_tonemap.argtypes = [vp, C.c_int64, vp]
        L.pt_pool_layout.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]
    return _lib


def check(status):
    if status != PT_OK:
        raise PtError(f"libpathed error {status}: {lib().pt_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def device_count():
    n = C.c_int(0)
    check(lib().pt_device_count(C.byref(n)))
    return n.value


class Scene:
    """Host scene description (pt_scene)."""

    def __init__(self):
        self.h = C.c_void_p()
        check(lib().pt_scene_create(C.byref(self.h)))

    def __del__(self):
        try:
            if self.h:
                lib().pt_scene_destroy(self.h)
        except Exception:
            pass

    @classmethod
    def preset(cls, preset, seed=0, param=0):
        s = cls()
        check(lib().pt_scene_preset(s.h, preset, seed, param))
        return s

    def add_material(self, mtype, albedo=(0, 0, 0), emit=(0, 0, 0), ior=1.5, alpha=0.1):
        out = C.c_int()
        a, e = _f32(albedo), _f32(emit)
        check(lib().pt_scene_add_material(self.h, mtype, _p(a), _p(e), ior, alpha, C.byref(out)))
        return out.value

    def add_triangles(self, positions, material_ids, normals=None):
        p = _f32(positions).reshape(-1)
        m = np.ascontiguousarray(material_ids, dtype=np.int32)
        n = None if normals is None else _f32(normals).reshape(-1)
        check(lib().pt_scene_add_triangles(self.h, _p(p), _p(n), _p(m), m.size))

    def add_spheres(self, centers, radii, material_ids):
        c = _f32(centers).reshape(-1)
        r = _f32(radii).reshape(-1)
        m = np.ascontiguousarray(material_ids, dtype=np.int32)
        check(lib().pt_scene_add_spheres(self.h, _p(c), _p(r), _p(m), m.size))

    @classmethod
    def reference(cls, index, scenes_dir):
        """The reference's built-in scene `index` (0..3) read from its scenes/ folder layout."""
        s = cls()
        check(lib().pt_scene_reference(s.h, index, os.fsencode(scenes_dir)))
        return s

    def load_obj(self, path, default_material=0):
        first = C.c_int()
        check(lib().pt_scene_load_obj(self.h, os.fsencode(path), default_material, C.byref(first)))
        return first.value

    def load_ply(self, path, material):
        check(lib().pt_scene_load_ply(self.h, os.fsencode(path), material))

    def load_exr_environment(self, path, sin_theta_weight=False):
        check(lib().pt_scene_load_exr_environment(self.h, os.fsencode(path), int(sin_theta_weight)))

    def set_environment(self, rgba, sin_theta_weight=False):
        e = _f32(rgba)
        check(lib().pt_scene_set_environment(self.h, _p(e), e.shape[1], e.shape[0], int(sin_theta_weight)))

    def set_camera(self, origin, target, up, vfov, flip=False):
        o, t, u = _f32(origin), _f32(target), _f32(up)
        check(lib().pt_scene_set_camera(self.h, _p(o), _p(t), _p(u), vfov, int(flip)))

    def info(self):
        i = SceneInfo()
        check(lib().pt_scene_get_info(self.h, C.byref(i)))
        return i

    def export(self):
        """Flat arrays in the layout the oracle takes."""
        i = self.info()
        T, S, M = i.n_triangles, i.n_spheres, i.n_materials
        pos = np.zeros((T, 9), np.float32)
        nrm = np.zeros((T, 9), np.float32)
        tm = np.zeros(T, np.int32)
        sph = np.zeros((S, 4), np.float32)
        sm = np.zeros(S, np.int32)
        mt = np.zeros(M, np.int32)
        ma = np.zeros((M, 3), np.float32)
        me = np.zeros((M, 3), np.float32)
        mi = np.zeros(M, np.float32)
        mal = np.zeros(M, np.float32)
        env = np.zeros((i.env_height, i.env_width, 4), np.float32) if i.env_width else None
        check(lib().pt_scene_export(self.h, _p(pos), _p(nrm), _p(tm), _p(sph), _p(sm), _p(mt), _p(ma), _p(me), _p(mi), _p(mal), _p(env)))
        return dict(
            tri_pos=pos, tri_nrm=nrm if i.has_normals else None, tri_mat=tm, sph=sph, sph_mat=sm,
            mat_type=mt, mat_albedo=ma, mat_emit=me, mat_ior=mi, mat_alpha=mal, env_rgba=env,
            camera=dict(origin=tuple(i.camera_origin), target=tuple(i.camera_target), up=tuple(i.camera_up),
                        vfov=i.camera_vfov, flip=i.camera_flip),
        )


class Renderer:
    """Device renderer (pt_renderer)."""

    def __init__(self, width, height):
        self.W, self.H = width, height
        self.h = C.c_void_p()
        check(lib().pt_renderer_create(width, height, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().pt_renderer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scene(self, scene):
        check(lib().pt_renderer_set_scene(self.h, scene.h))

    def set_camera(self, origin, target, up, vfov, flip=False):
        o, t, u = _f32(origin), _f32(target), _f32(up)
        check(lib().pt_renderer_set_camera(self.h, _p(o), _p(t), _p(u), vfov, int(flip)))

    def set_material(self, mid, mtype, albedo=(0, 0, 0), emit=(0, 0, 0), ior=1.5, alpha=0.1):
        a, e = _f32(albedo), _f32(emit)
        check(lib().pt_renderer_set_material(self.h, mid, mtype, _p(a), _p(e), ior, alpha))

    def set_max_depth(self, d):
        check(lib().pt_renderer_set_max_depth(self.h, d))

    def set_nee(self, on):
        check(lib().pt_renderer_set_nee(self.h, int(on)))

    def set_type(self, t):
        check(lib().pt_renderer_set_type(self.h, t))

    def set_seed(self, s):
        check(lib().pt_renderer_set_seed(self.h, s))

    def set_option(self, opt, v):
        check(lib().pt_renderer_set_option(self.h, opt, v))

    def get_option(self, opt):
        v = C.c_int64()
        check(lib().pt_renderer_get_option(self.h, opt, C.byref(v)))
        return v.value

    def render_async(self, spp):
        check(lib().pt_renderer_render_async(self.h, spp))

    def poll(self):
        f = C.c_int()
        ms = C.c_float()
        check(lib().pt_renderer_poll(self.h, C.byref(f), C.byref(ms)))
        return bool(f.value), ms.value

    def synchronize(self):
        check(lib().pt_renderer_synchronize(self.h))

    def render(self, spp):
        self.render_async(spp)
        self.synchronize()

    def reset(self):
        check(lib().pt_renderer_reset(self.h))

    def spp(self):
        v = C.c_int64()
        check(lib().pt_renderer_get_spp(self.h, C.byref(v)))
        return v.value

    def radiance(self):
        out = np.zeros((self.H, self.W, 3), np.float32)
        check(lib().pt_renderer_get_radiance(self.h, _p(out)))
        return out

    def radiance_sum(self):
        out = np.zeros((self.H, self.W, 3), np.float64)
        check(lib().pt_renderer_get_radiance_sum(self.h, _p(out)))
        return out

    def rgba8(self):
        out = np.zeros((self.H, self.W, 4), np.uint8)
        check(lib().pt_renderer_get_rgba8(self.h, _p(out)))
        return out

    def hit_test(self, x, y):
        m = C.c_int()
        h = C.c_int()
        check(lib().pt_renderer_hit_test(self.h, x, y, C.byref(m), C.byref(h)))
        return m.value, bool(h.value)

    def stats(self):
        s = RenderStats()
        check(lib().pt_renderer_get_stats(self.h, C.byref(s)))
        return s

    def reset_stats(self):
        check(lib().pt_renderer_reset_stats(self.h))

    def kernel_stats(self):
        arr = (KernelStat * 16)()
        n = C.c_int()
        check(lib().pt_renderer_get_kernel_stats(self.h, arr, 16, C.byref(n)))
        return {arr[i].name.decode(): dict(launches=arr[i].launches, units=arr[i].units, ms=arr[i].ms, bytes=arr[i].bytes) for i in range(n.value)}

    def path_radiance(self, n):
        out = np.zeros((n, 3), np.float32)
        cnt = C.c_int64()
        check(lib().pt_renderer_get_path_radiance(self.h, _p(out), n, C.byref(cnt)))
        return out[: cnt.value]

    def intersect(self, rays, any_hit=False):
        r = _f32(rays).reshape(-1, 8)
        out = np.zeros((r.shape[0], 4), np.float32)
        check(lib().pt_renderer_intersect(self.h, _p(r), r.shape[0], int(any_hit), _p(out)))
        prim = out[:, 3].copy().view(np.int32)
        return out[:, 0], out[:, 1], out[:, 2], prim

    def autotune(self, spp):
        t = TuneResult()
        check(lib().pt_renderer_autotune(self.h, spp, C.byref(t)))
        return t.asdict()


class Batch:
    """Independent renders sharing one path pool, run as one captured graph (pt_batch)."""

    def __init__(self):
        self.h = C.c_void_p()
        check(lib().pt_batch_create(C.byref(self.h)))
        self.sizes = []

    def close(self):
        if self.h:
            lib().pt_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, scene, width, height, spp, max_depth=8):
        i = C.c_int()
        check(lib().pt_batch_add(self.h, scene.h, width, height, spp, max_depth, C.byref(i)))
        self.sizes.append((width, height))
        return i.value

    def set_option(self, opt, v):
        check(lib().pt_batch_set_option(self.h, opt, v))

    def get_option(self, opt):
        v = C.c_int64()
        check(lib().pt_batch_get_option(self.h, opt, C.byref(v)))
        return v.value

    def autotune(self, probe_divisor=8):
        t = TuneResult()
        check(lib().pt_batch_autotune(self.h, probe_divisor, C.byref(t)))
        return t.asdict()

    def run(self):
        check(lib().pt_batch_run(self.h))

    def radiance(self, i):
        w, h = self.sizes[i]
        out = np.zeros((h, w, 3), np.float32)
        check(lib().pt_batch_get_radiance(self.h, i, _p(out)))
        return out

    def render_stats(self, i):
        s = BatchRenderStats()
        check(lib().pt_batch_get_render_stats(self.h, i, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def stats(self):
        s = BatchStats()
        check(lib().pt_batch_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}


def triangle_test(tris, rays, mode):
    t = _f32(tris).reshape(-1, 9)
    r = _f32(rays).reshape(-1, 8)
    out = np.zeros((t.shape[0], 4), np.float32)
    check(lib().pt_unit_triangle_test(_p(t), _p(r), t.shape[0], mode, _p(out)))
    return out


def bsdf_eval(mat, wo, wi):
    m = _f32(mat).reshape(-1, 6)
    o = _f32(wo).reshape(-1, 3)
    i = _f32(wi).reshape(-1, 3)
    f = np.zeros((m.shape[0], 3), np.float32)
    pdf = np.zeros(m.shape[0], np.float32)
    check(lib().pt_unit_bsdf_eval(_p(m), _p(o), _p(i), m.shape[0], _p(f), _p(pdf)))
    return f, pdf


def bsdf_sample(mat, wo, xi):
    m = _f32(mat).reshape(-1, 6)
    o = _f32(wo).reshape(-1, 3)
    x = _f32(xi).reshape(-1, 2)
    out = np.zeros((m.shape[0], 8), np.float32)
    check(lib().pt_unit_bsdf_sample(_p(m), _p(o), _p(x), m.shape[0], _p(out)))
    return out


def env_tables(rgba, sin_theta_weight=False):
    e = _f32(rgba)
    h, w = e.shape[:2]
    rp = np.zeros((h, w), np.float64)
    ra = np.zeros((h, w), np.int32)
    mp = np.zeros(h, np.float64)
    ma = np.zeros(h, np.int32)
    tp = np.zeros((h, w), np.float32)
    check(lib().pt_unit_env_tables(_p(e), w, h, int(sin_theta_weight), _p(rp), _p(ra), _p(mp), _p(ma), _p(tp)))
    return rp, ra, mp, ma, tp


def env_sample(rgba, xi, pdf_dirs=None, sin_theta_weight=False):
    e = _f32(rgba)
    h, w = e.shape[:2]
    x = _f32(xi).reshape(-1, 2)
    n = x.shape[0]
    pd = _f32(pdf_dirs if pdf_dirs is not None else np.zeros((0, 3))).reshape(-1, 3)
    m = pd.shape[0]
    d = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    em = np.zeros((n, 3), np.float32)
    po = np.zeros(m, np.float32)
    check(lib().pt_unit_env_sample(_p(e), w, h, int(sin_theta_weight), _p(x), n, _p(d), _p(pdf), _p(em), _p(pd), m, _p(po)))
    return d, pdf, em, po


def path_uniforms(seed, pixels, samples, k):
    p = np.ascontiguousarray(pixels, dtype=np.uint32)
    s = np.ascontiguousarray(samples, dtype=np.uint32)
    out = np.zeros((p.size, k), np.float32)
    check(lib().pt_unit_path_uniforms(seed, _p(p), _p(s), p.size, k, _p(out)))
    return out


def guide_forward(params, feats):
    p = _f32(params)
    f = _f32(feats).reshape(-1, 8)
    out = np.zeros((f.shape[0], 48), np.float32)
    check(lib().pt_unit_guide_forward(_p(p), _p(f), f.shape[0], _p(out)))
    return out


def guide_records(params, feats):
    """The inference kernel's 64-byte mixture records of every feature row, (n, 16) uint32."""
    p = _f32(params)
    f = _f32(feats).reshape(-1, 8)
    out = np.zeros((f.shape[0], 16), np.uint32)
    check(lib().pt_unit_guide_records(_p(p), _p(f), f.shape[0], _p(out)))
    return out


def guide_gradients(params, records):
    p = _f32(params)
    r = _f32(records).reshape(-1, 16)
    g = np.zeros(GUIDE_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    check(lib().pt_unit_guide_gradients(_p(p), _p(r), r.shape[0], _p(g), _p(loss)))
    return float(loss[0]), g


def guide_adam(params, m, v, grad, step, lr):
    p = _f32(params).copy()
    mm = _f32(m).copy()
    vv = _f32(v).copy()
    g = _f32(grad)
    check(lib().pt_unit_guide_adam(_p(p), _p(mm), _p(vv), _p(g), step, lr))
    return p, mm, vv


def mfma_tile(a, b):
    a = _f32(a).reshape(32, 16)
    b = _f32(b).reshape(16, 32)
    c = np.zeros((32, 32), np.float32)
    check(lib().pt_unit_mfma_tile(_p(a), _p(b), _p(c)))
    return c


def mfma_tile16(a, b):
    a = _f32(a).reshape(16, 32)
    b = _f32(b).reshape(32, 16)
    c = np.zeros((16, 16), np.float32)
    check(lib().pt_unit_mfma_tile16(_p(a), _p(b), _p(c)))
    return c


def vmf_sample(mu, kappa, xi):
    """Directions drawn from vMF lobes on the GPU and the lobe density at them."""
    m = _f32(mu).reshape(-1, 3)
    k = _f32(kappa).reshape(-1)
    x = _f32(xi).reshape(-1, 2)
    d = np.zeros((k.size, 3), np.float32)
    p = np.zeros(k.size, np.float32)
    check(lib().pt_unit_vmf_sample(_p(m), _p(k), _p(x), k.size, _p(d), _p(p)))
    return d, p


def tonemap(radiance):
    """framebuffer bytes (n, 4) of radiance triples (n, 3) through the renderer's GPU tonemap."""
    r = _f32(radiance).reshape(-1, 3)
    out = np.zeros((r.shape[0], 4), np.uint8)
    check(lib().pt_unit_tonemap(_p(r), r.shape[0], _p(out)))
    return out


def pool_layout(width, height, spp, pool_option=0):
    """(capacity, chunks, max_slot) of a pass's path pool; host only."""
    cap, chunks, hi = C.c_int64(), C.c_int64(), C.c_int64()
    check(lib().pt_pool_layout(width, height, spp, pool_option, C.byref(cap), C.byref(chunks), C.byref(hi)))
    return cap.value, chunks.value, hi.value


def exr_read(path):
    """RGBA float image (H, W, 4) of an OpenEXR file, scanline order (row 0 = top)."""
    w, h = C.c_int(), C.c_int()
    check(lib().pt_exr_info(os.fsencode(path), C.byref(w), C.byref(h)))
    out = np.zeros((h.value, w.value, 4), np.float32)
    check(lib().pt_exr_read_rgba(os.fsencode(path), _p(out), out.size))
    return out


def exr_write(path, radiance, half=True):
    """Writes an (H, W, 3) radiance buffer (row 0 = bottom) as an OpenEXR file."""
    r = _f32(radiance)
    h, w = r.shape[0], r.shape[1]
    check(lib().pt_exr_write_rgb(os.fsencode(path), _p(r), w, h, int(half)))
