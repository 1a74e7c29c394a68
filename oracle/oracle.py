"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP product is compared against.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force=False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-B" if force else "all", "-s"], cwd=_HERE)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_tris", C.c_int64), ("tri_pos", C.c_void_p), ("tri_nrm", C.c_void_p), ("tri_mat", C.c_void_p),
        ("n_spheres", C.c_int64), ("sph", C.c_void_p), ("sph_mat", C.c_void_p),
        ("n_mats", C.c_int64), ("mat_type", C.c_void_p), ("mat_albedo", C.c_void_p), ("mat_emit", C.c_void_p),
        ("mat_ior", C.c_void_p), ("mat_alpha", C.c_void_p),
        ("env_w", C.c_int), ("env_h", C.c_int), ("env_rgba", C.c_void_p), ("env_sin_weight", C.c_int),
        ("cam_origin", C.c_float * 3), ("cam_target", C.c_float * 3), ("cam_up", C.c_float * 3),
        ("cam_vfov", C.c_float), ("cam_flip", C.c_int),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("max_depth", C.c_int), ("nee", C.c_int), ("mode", C.c_int), ("tri_test", C.c_int), ("ref_env", C.c_int),
        ("use_bvh", C.c_int), ("precision", C.c_int), ("threads", C.c_int),
        ("seed", C.c_uint64), ("sample_start", C.c_int64), ("spp", C.c_int),
        ("row_begin", C.c_int), ("row_end", C.c_int), ("pixel_stride", C.c_int),
        ("mis_sample", C.c_int),
    ]


def _setup(L):
    vp = C.c_void_p
    L.or_scene_create.restype = vp
    L.or_scene_create.argtypes = [C.POINTER(SceneDesc), C.c_int, C.c_int, C.c_int]
    L.or_scene_destroy.argtypes = [vp]
    L.or_render.argtypes = [vp, C.POINTER(RenderParams), vp, vp, vp]
    L.or_trace_paths.argtypes = [vp, C.POINTER(RenderParams), vp, vp, C.c_int64, vp]
    L.or_render_pixels.argtypes = [vp, C.POINTER(RenderParams), vp, C.c_int64, vp, vp, vp]
    L.or_intersect.argtypes = [vp, vp, C.c_int64, C.c_int, C.c_int, vp]
    L.or_tri_test.argtypes = [vp, vp, C.c_int64, C.c_int, vp]
    L.or_bsdf_eval.argtypes = [vp, vp, vp, C.c_int64, C.c_int, vp, vp]
    L.or_bsdf_sample.argtypes = [vp, vp, vp, C.c_int64, C.c_int, vp]
    L.or_fresnel.restype = C.c_double
    L.or_fresnel.argtypes = [C.c_double, C.c_double, C.c_double]
    L.or_dist1d.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, vp, vp]
    L.or_dist2d.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_int, vp, vp]
    L.or_phitheta_pdf_canonical.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
    L.or_env_pdf.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int64, C.c_int, vp]
    L.or_env_sample.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int64, C.c_int, vp, vp, vp]
    L.or_env_alias.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.or_alias_build.argtypes = [vp, C.c_int, vp, vp, vp]
    L.or_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int, vp]
    L.or_path_uniforms.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, vp]
    L.or_camera_rays.argtypes = [vp, vp, vp, C.c_float, C.c_int, C.c_int, C.c_int, vp, C.c_int64, vp]
    L.or_framebuffer_update.argtypes = [vp, vp, C.c_int64, C.c_int, C.c_int, vp]


def _p(a):
    return None if a is None else a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Scene:
    """A scene given as flat arrays (the layout the product's scene export produces)."""

    def __init__(self, arrays, width, height, bvh=False):
        a = arrays
        self._keep = {}
        d = SceneDesc()

        def keep(name, arr):
            self._keep[name] = arr
            return _p(arr)

        tri = _f32(a["tri_pos"]).reshape(-1)
        d.n_tris = tri.size // 9
        d.tri_pos = keep("tri_pos", tri)
        d.tri_nrm = keep("tri_nrm", _f32(a["tri_nrm"]).reshape(-1)) if a.get("tri_nrm") is not None else None
        d.tri_mat = keep("tri_mat", _i32(a["tri_mat"]))
        sph = _f32(a.get("sph", np.zeros((0, 4)))).reshape(-1)
        d.n_spheres = sph.size // 4
        d.sph = keep("sph", sph if sph.size else np.zeros(4, np.float32))
        d.sph_mat = keep("sph_mat", _i32(a.get("sph_mat", np.zeros(1, np.int32))))
        mt = _i32(a["mat_type"])
        d.n_mats = mt.size
        d.mat_type = keep("mat_type", mt)
        d.mat_albedo = keep("mat_albedo", _f32(a["mat_albedo"]).reshape(-1))
        d.mat_emit = keep("mat_emit", _f32(a["mat_emit"]).reshape(-1))
        d.mat_ior = keep("mat_ior", _f32(a["mat_ior"]))
        d.mat_alpha = keep("mat_alpha", _f32(a["mat_alpha"]))
        env = a.get("env_rgba")
        if env is not None and env.size:
            env = _f32(env)
            d.env_h, d.env_w = env.shape[0], env.shape[1]
            d.env_rgba = keep("env", env.reshape(-1))
        else:
            d.env_w = d.env_h = 0
            d.env_rgba = None
        d.env_sin_weight = int(a.get("env_sin_weight", 0))
        cam = a["camera"]
        for i in range(3):
            d.cam_origin[i] = cam["origin"][i]
            d.cam_target[i] = cam["target"][i]
            d.cam_up[i] = cam["up"][i]
        d.cam_vfov = cam["vfov"]
        d.cam_flip = int(cam.get("flip", 0))
        self.width, self.height = width, height
        self.handle = lib().or_scene_create(C.byref(d), width, height, 1 if bvh else 0)
        self.bvh = bvh

    def __del__(self):
        try:
            if self.handle:
                lib().or_scene_destroy(self.handle)
        except Exception:
            pass

    def params(self, spp=1, max_depth=6, nee=1, mode=0, tri_test=0, ref_env=1, precision=32,
               threads=0, seed=0, sample_start=0, row_begin=0, row_end=0, pixel_stride=1, mis_sample=0):
        p = RenderParams()
        p.width, p.height = self.width, self.height
        p.max_depth, p.nee, p.mode, p.tri_test, p.ref_env = max_depth, nee, mode, tri_test, ref_env
        p.use_bvh = 1 if self.bvh else 0
        p.precision, p.threads, p.seed = precision, threads, seed
        p.sample_start, p.spp = sample_start, spp
        p.row_begin, p.row_end, p.pixel_stride = row_begin, row_end, pixel_stride
        p.mis_sample = mis_sample
        return p

    def render(self, want_sq=False, **kw):
        """Per-pixel FP64 sums of the radiance of `spp` samples; returns (sum, sumsq, rays)."""
        p = self.params(**kw)
        n = self.width * self.height * 3
        s = np.zeros(n, np.float64)
        sq = np.zeros(n, np.float64) if want_sq else None
        rays = np.zeros(1, np.int64)
        lib().or_render(self.handle, C.byref(p), _p(s), _p(sq), _p(rays))
        return s.reshape(self.height, self.width, 3), (None if sq is None else sq.reshape(self.height, self.width, 3)), int(rays[0])

    def render_pixels(self, pixels, **kw):
        """FP64 sums and sums of squares (n, 3) of `spp` samples at each listed pixel index
        (row * width + col); returns (sum, sumsq, rays)."""
        p = self.params(**kw)
        pix = np.ascontiguousarray(pixels, dtype=np.int32)
        s = np.zeros((pix.size, 3), np.float64)
        sq = np.zeros((pix.size, 3), np.float64)
        rays = np.zeros(1, np.int64)
        lib().or_render_pixels(self.handle, C.byref(p), _p(pix), pix.size, _p(s), _p(sq), _p(rays))
        return s, sq, int(rays[0])

    def trace_paths(self, pixels, samples, **kw):
        p = self.params(**kw)
        pix = _i32(pixels)
        smp = _i32(samples)
        out = np.zeros((pix.size, 3), np.float64)
        lib().or_trace_paths(self.handle, C.byref(p), _p(pix), _p(smp), pix.size, _p(out))
        return out

    def intersect(self, rays, tri_test=1, any_hit=False):
        r = _f32(rays).reshape(-1, 8)
        out = np.zeros((r.shape[0], 4), np.float32)
        lib().or_intersect(self.handle, _p(r), r.shape[0], tri_test, 1 if any_hit else 0, _p(out))
        prim = out[:, 3].copy().view(np.int32)
        return out[:, 0], out[:, 1], out[:, 2], prim


def tri_test(tris, rays, mode):
    t = _f32(tris).reshape(-1, 9)
    r = _f32(rays).reshape(-1, 8)
    out = np.zeros((t.shape[0], 4), np.float32)
    lib().or_tri_test(_p(t), _p(r), t.shape[0], mode, _p(out))
    return out


def bsdf_eval(mat, wo, wi, precision=32, sample_pdf=False):
    """f and pdf() at (wo, wi); sample_pdf=True returns the density sample() draws with instead."""
    if sample_pdf:
        precision += 1
    m = _f32(mat).reshape(-1, 6)
    o = _f32(wo).reshape(-1, 3)
    i = _f32(wi).reshape(-1, 3)
    f = np.zeros((m.shape[0], 3), np.float64)
    pdf = np.zeros(m.shape[0], np.float64)
    lib().or_bsdf_eval(_p(m), _p(o), _p(i), m.shape[0], precision, _p(f), _p(pdf))
    return f, pdf


def bsdf_sample(mat, wo, xi, precision=32):
    m = _f32(mat).reshape(-1, 6)
    o = _f32(wo).reshape(-1, 3)
    x = _f32(xi).reshape(-1, 2)
    out = np.zeros((m.shape[0], 8), np.float64)
    lib().or_bsdf_sample(_p(m), _p(o), _p(x), m.shape[0], precision, _p(out))
    return out


def fresnel(cos_i, eta_i, eta_t):
    return lib().or_fresnel(cos_i, eta_i, eta_t)


def dist1d(values, xis=()):
    v = _f32(values)
    x = _f32(np.asarray(xis, dtype=np.float32))
    pmf = np.zeros(v.size, np.float32)
    tot = np.zeros(1, np.float32)
    idx = np.zeros(max(x.size, 1), np.int32)
    xp = np.zeros(max(x.size, 1), np.float32)
    lib().or_dist1d(_p(v), v.size, _p(pmf), _p(tot), _p(x) if x.size else None, x.size, _p(idx), _p(xp))
    return pmf, float(tot[0]), idx[: x.size], xp[: x.size]


def dist2d(values, w, h, xis=()):
    v = _f32(values)
    x = _f32(np.asarray(xis, dtype=np.float32)).reshape(-1)
    pmf = np.zeros(w * h, np.float32)
    n = x.size // 2
    xy = np.zeros(max(2 * n, 2), np.int32)
    xp = np.zeros(max(n, 1), np.float32)
    lib().or_dist2d(_p(v), w, h, _p(pmf), _p(x) if n else None, n, _p(xy), _p(xp))
    return pmf.reshape(h, w), xy[: 2 * n].reshape(-1, 2), xp[:n]


def phitheta_pdf_canonical(values, w, h, nx, ny):
    v = _f32(values)
    out = np.zeros(nx * ny, np.float32)
    lib().or_phitheta_pdf_canonical(_p(v), w, h, nx, ny, _p(out))
    return out.reshape(ny, nx)


def env_pdf(rgba, dirs, ref=True, sin_weight=False):
    e = _f32(rgba)
    h, w = e.shape[:2]
    d = _f32(dirs).reshape(-1, 3)
    out = np.zeros(d.shape[0], np.float32)
    lib().or_env_pdf(_p(e), w, h, int(sin_weight), _p(d), d.shape[0], int(ref), _p(out))
    return out


def env_sample(rgba, xi, ref=False, sin_weight=False):
    e = _f32(rgba)
    h, w = e.shape[:2]
    x = _f32(xi).reshape(-1, 2)
    n = x.shape[0]
    d = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    em = np.zeros((n, 3), np.float32)
    lib().or_env_sample(_p(e), w, h, int(sin_weight), _p(x), n, int(ref), _p(d), _p(pdf), _p(em))
    return d, pdf, em


def env_alias(rgba, sin_weight=False):
    e = _f32(rgba)
    h, w = e.shape[:2]
    rp = np.zeros(w * h, np.float64)
    ra = np.zeros(w * h, np.int32)
    mp = np.zeros(h, np.float64)
    ma = np.zeros(h, np.int32)
    tp = np.zeros(w * h, np.float32)
    lib().or_env_alias(_p(e), w, h, int(sin_weight), _p(rp), _p(ra), _p(mp), _p(ma), _p(tp))
    return rp.reshape(h, w), ra.reshape(h, w), mp, ma, tp.reshape(h, w)


def alias_build(weights):
    w = np.ascontiguousarray(weights, dtype=np.float64)
    prob = np.zeros(w.size, np.float64)
    alias = np.zeros(w.size, np.int32)
    s = np.zeros(1, np.float64)
    lib().or_alias_build(_p(w), w.size, _p(prob), _p(alias), _p(s))
    return prob, alias, float(s[0])


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().or_pcg32(initstate, initseq, n, _p(out))
    return out


def path_uniforms(seed, pixel, sample, n):
    out = np.zeros(n, np.float32)
    lib().or_path_uniforms(seed, pixel, sample, n, _p(out))
    return out


def camera_rays(cam, W, H, queries):
    o = _f32(cam["origin"]); t = _f32(cam["target"]); u = _f32(cam["up"])
    q = _f32(queries).reshape(-1, 4)
    out = np.zeros((q.shape[0], 6), np.float32)
    lib().or_camera_rays(_p(o), _p(t), _p(u), cam["vfov"], int(cam.get("flip", 0)), W, H, _p(q), q.shape[0], _p(out))
    return out


def framebuffer_update(pass_rad, mean, pass_samples, previous_samples):
    p = _f32(pass_rad).reshape(-1)
    m = np.ascontiguousarray(mean, dtype=np.float32).reshape(-1)
    npix = p.size // 3
    rgba = np.zeros(npix * 4, np.uint8)
    lib().or_framebuffer_update(_p(p), _p(m), npix, pass_samples, previous_samples, _p(rgba))
    return m, rgba
