"""Inputs shared by the oracle tests and the GPU parity tests (plain data builders)."""
import numpy as np

import pathed_capi as P


def srgb_sweep():
    """Radiance values that cover the byte range densely: a uniform grid, the float
    neighbours of every byte boundary of the transfer (found by bisection in FP64), the
    segment switch at 0.0031308, and the special values (0, 1, > 1, inf, nan)."""
    g = 1.0 / 2.2
    v = list(np.linspace(0, 1.2, 200001))
    for k in range(1, 256):
        y = k / 255.0
        lo = y / 12.92 if y <= 0.0031308 * 12.92 else ((y + 0.055) / 1.055) ** (1 / g)
        for d in range(-3, 4):
            v.append(np.nextafter(np.float32(lo), np.float32(np.inf if d > 0 else -np.inf)) if d else np.float32(lo))
    t = np.float32(0.0031308)
    v += [t, np.nextafter(t, np.float32(1)), np.nextafter(t, np.float32(0)), 0.0, 1.0, 2.0, 1e30, np.inf, np.nan, 1e-30]
    return np.array(v, np.float32)


def sphere_light_scene(target, rho=0.5, le=2.0, radius=1.0, height=1.5):
    """A Lambertian floor (y = 0, albedo rho, 100 x 100) under one emissive sphere (radiance
    le, `radius`, centre (0, height, 0)), seen through a 1e-4 rad pinhole aimed at `target`
    on the floor.  With height > radius the whole sphere is above every floor point's horizon,
    so the radiance leaving the floor towards the camera after one bounce is the closed form
    rho * le * radius^2 * height / d^3 (d = distance from the point to the centre): the
    irradiance of a uniform sphere source, pi le (r / d)^2 cos(theta), times rho / pi.
    Returns (scene, closed form)."""
    s = P.Scene()
    fl = s.add_material(P.LAMBERTIAN, (rho, rho, rho))
    li = s.add_material(P.LAMBERTIAN, (0, 0, 0), (le, le, le))
    h = 50.0
    quad = np.array([[-h, 0, -h, -h, 0, h, h, 0, h], [-h, 0, -h, h, 0, h, h, 0, -h]], np.float32)
    s.add_triangles(quad, [fl, fl])
    s.add_spheres([[0, height, 0]], [radius], [li])
    s.set_camera([0, 1, 3], target, [0, 1, 0], 1e-4)
    d = np.linalg.norm(np.array([0, height, 0], np.float64) - np.asarray(target, np.float64))
    return s, rho * le * radius**2 * height / d**3


def path_agreement(o, gpu, pixels, samples, tol=1e-3, **kw):
    """Per-path parity of GPU path radiances with the oracle's product-mode FP32 integrator on
    the same PCG32 streams: the fraction of paths within `tol` relative (denominator |L| + 1e-3),
    and, for scale, the same fraction between the oracle's own FP32 and FP64 runs of those
    paths.  A path through glass or a near-specular lobe can branch differently after a
    rounding difference anywhere on it; the second number is how often that happens to the
    oracle itself, the floor that FP32 arithmetic sets for any two implementations."""
    kw = dict(dict(nee=1, mode=1, tri_test=1, ref_env=0), **kw)
    c32 = o.trace_paths(pixels, samples, precision=32, **kw)
    c64 = o.trace_paths(pixels, samples, precision=64, **kw)

    def agree(a, b):
        return float((np.abs(a - b).max(1) / (np.abs(b).max(1) + 1e-3) < tol).mean())

    return agree(gpu, c32), agree(c32, c64), c32


def one_texel_plane_scene(scenes_dir, vfov_deg=10.0):
    """The reference's scene 2 (scene.cu:73-94: the plane of tests/plane.obj, albedo 1, lit
    only by the 1000 x 500 map 1_pixel_test.exr, whose single non-black texel is row 239,
    column 753 at radiance 1e4) seen from straight above through a narrow pinhole, so that every
    pixel sees the plane.  The plane's normal is +y and the texel spans polar angles
    [239, 240] pi / 500 from the +y pole and 2 pi / 1000 of azimuth; irradiance from a uniform
    patch is L * dphi * integral of cos(theta) sin(theta) dtheta = L * dphi *
    (sin^2(theta1) - sin^2(theta0)) / 2, and the Lambertian plane returns albedo / pi of it.
    Returns (scene, closed-form radiance)."""
    s = P.Scene.reference(2, scenes_dir)
    s.set_camera([0, 1, 0], [0, 0, 0], [0, 0, -1], np.radians(vfov_deg))
    th0, th1 = 239 * np.pi / 500, 240 * np.pi / 500
    irradiance = 1e4 * (2 * np.pi / 1000) * (np.sin(th1) ** 2 - np.sin(th0) ** 2) / 2
    return s, irradiance / np.pi
