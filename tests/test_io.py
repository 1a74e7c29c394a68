"""Scene and image files (host code of libpathed; no GPU needed).

Pins: the reference's PLY known answer (test/test_ply_parser.cpp:11-22), the reference
scenes' own OBJ/MTL/PLY/EXR files (fixtures under tests/golden/scenes, copied data), OpenEXR
round trips, IEEE half rounding against numpy's float16, and a ZIP/ZIPS-compressed file
written by an independent encoder in this test (zlib over the byte-interleaved,
delta-predicted block of the OpenEXR format description).
"""
import os
import struct
import zlib

import numpy as np
import pytest

import pathed_capi as P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SCENES = os.path.join(GOLD, "scenes")


def test_ply_known_answer():
    """8 vertices, 4 faces, first vertex and first face's v0 = (-10, -4.14615, -10)."""
    s = P.Scene()
    m = s.add_material(P.LAMBERTIAN, (0.5, 0.5, 0.5))
    s.load_ply(os.path.join(GOLD, "ply_example.ply"), m)
    a = s.export()
    assert a["tri_pos"].shape == (4, 9)
    assert np.array_equal(a["tri_pos"][0, :3], np.array([-10, -4.14615, -10], np.float32))
    assert a["tri_nrm"] is None  # flat normals from the winding
    assert (a["tri_mat"] == m).all()


def test_ply_rejects_bad_material_and_missing_file():
    s = P.Scene()
    with pytest.raises(P.PtError):
        s.load_ply(os.path.join(GOLD, "ply_example.ply"), 0)  # no material yet
    s.add_material(P.LAMBERTIAN)
    with pytest.raises(P.PtError):
        s.load_ply(os.path.join(GOLD, "does_not_exist.ply"), 0)


def _mtl_names(path):
    return sorted(line.split()[1] for line in open(path) if line.strip().startswith("newmtl"))


def test_obj_cornell_original():
    """18 quads -> 36 triangles; 8 MTL materials appended in name order after the default
    material; the light (Ke 17 12 4) is the only emitter (2 triangles)."""
    s = P.Scene()
    s.add_material(P.LAMBERTIAN, (1, 0, 0))
    first = s.load_obj(os.path.join(SCENES, "cornell-box", "CornellBox-Original.obj"), 0)
    a = s.export()
    names = _mtl_names(os.path.join(SCENES, "cornell-box", "CornellBox-Original.mtl"))
    assert first == 1 and len(names) == 8
    assert a["tri_pos"].shape[0] == 36
    assert a["mat_type"].size == 9 and (a["mat_type"] == P.LAMBERTIAN).all()
    light = 1 + names.index("light")
    assert np.array_equal(a["mat_emit"][light], np.array([17, 12, 4], np.float32))
    assert np.allclose(a["mat_albedo"][1 + names.index("leftWall")], [0.63, 0.065, 0.05])
    emissive = (a["mat_emit"][a["tri_mat"]] != 0).any(1)
    assert emissive.sum() == 2 and (a["tri_mat"][emissive] == light).all()
    # quads split (0,1,2) (0,2,3): the floor quad (first face) shares its corners 0 and 2
    t = a["tri_pos"].reshape(-1, 3, 3)
    assert np.array_equal(t[0, 0], t[1, 0]) and np.array_equal(t[0, 2], t[1, 1])
    # flat normals: face normal of the floor points up
    n = np.cross(t[0, 1] - t[0, 0], t[0, 2] - t[0, 0])
    assert n[1] > 0


def test_obj_glossy_box_with_normals_takes_default_material():
    """box.obj: v/t/n triangles (per-vertex normals), usemtl without mtllib -> the default
    material for every face."""
    s = P.Scene()
    s.add_material(P.LAMBERTIAN, (1, 0, 0))
    mirror = s.add_material(P.MIRROR)
    first = s.load_obj(os.path.join(SCENES, "cornell-glossy", "box.obj"), mirror)
    a = s.export()
    assert first == -1
    assert a["tri_pos"].shape[0] == 12 and (a["tri_mat"] == mirror).all()
    assert a["tri_nrm"] is not None
    # the file's normals as written (four decimals)
    assert np.allclose(np.linalg.norm(a["tri_nrm"].reshape(-1, 3), axis=1), 1.0, atol=1e-3)


def test_reference_scenes_0_and_1():
    s0 = P.Scene.reference(0, SCENES)
    i0 = s0.info()
    assert (i0.n_triangles, i0.n_spheres, i0.n_materials) == (36, 0, 9)
    assert np.allclose(list(i0.camera_origin), [0, 1, 6.8]) and abs(i0.camera_vfov - np.radians(19.5)) < 1e-6
    s1 = P.Scene.reference(1, SCENES)
    i1 = s1.info()
    nglossy = len(_mtl_names(os.path.join(SCENES, "cornell-glossy", "CornellBox-Glossy.mtl")))
    assert i1.n_spheres == 1 and i1.n_materials == 3 + nglossy
    a = s1.export()
    assert a["mat_type"][2] == P.GLASS and a["mat_ior"][2] == pytest.approx(1.4)
    assert a["sph"][0, 3] == pytest.approx(0.45)


def test_reference_scene_3_mis_plates():
    """scene.cu:95-167: the floor and four Beckmann plates (PLY, alpha 0.005 .. 0.1), five
    emissive spheres in the reference's order, the wells6 environment (1000 x 500) and the
    tilted, handedness-flipped camera."""
    s = P.Scene.reference(3, SCENES)
    i = s.info()
    assert (i.n_triangles, i.n_spheres, i.n_materials, i.n_emitters) == (12, 5, 10, 5)
    assert (i.env_width, i.env_height) == (1000, 500) and i.camera_flip == 1
    assert np.allclose(list(i.camera_origin), [0, 2, 15]) and i.camera_vfov == pytest.approx(np.radians(28.0000262073138))
    a = s.export()
    assert a["tri_mat"].tolist() == [0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4]
    assert a["mat_type"][:5].tolist() == [P.LAMBERTIAN] + [P.MICROFACET] * 4
    assert np.allclose(a["mat_alpha"][1:5], [0.005, 0.02, 0.05, 0.1])
    assert np.allclose(a["mat_emit"][5:, 0], [800, 100, 901.803, 11.1111, 1.234567])
    assert np.allclose(a["sph"][:, 3], [0.5, 0.1, 0.03333, 0.3, 0.9])
    assert np.allclose(a["sph"][:, :3], [[10, 10, 4], [-1.25, 0, 0], [-3.75, 0, 0], [1.25, 0, 0], [3.75, 0, 0]])
    assert (a["sph_mat"] == np.arange(5, 10)).all()
    env = a["env_rgba"]
    assert np.isfinite(env).all() and env[..., :3].min() >= 0 and env[..., :3].max() > 1


def test_exr_float_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.lognormal(0, 2, (7, 11, 3)).astype(np.float32)
    p = str(tmp_path / "f.exr")
    P.exr_write(p, img, half=False)
    back = P.exr_read(p)
    # radiance rows run bottom to top, EXR scanlines top to bottom
    assert np.array_equal(back[::-1, :, :3], img)
    assert (back[..., 3] == 1).all()


def test_exr_half_rounding_matches_ieee(tmp_path):
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.lognormal(0, 6, 3000), rng.uniform(0, 7e-5, 300), [0, 6.1e-5, 5.96e-8, 2.98e-8, 65504, 65519, 65520, 1e6]])
    vals = vals[: (vals.size // 3) * 3].astype(np.float32)
    img = vals.reshape(1, -1, 3)
    p = str(tmp_path / "h.exr")
    P.exr_write(p, img, half=True)
    back = P.exr_read(p)[0, :, :3]
    with np.errstate(over="ignore"):  # 65520 and above round to infinity
        want = img[0].astype(np.float16).astype(np.float32)
    assert np.array_equal(back, want)


def _zip_block(raw):
    """OpenEXR ZIP block encoder: even bytes then odd bytes, then differences + 128."""
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int16)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return zlib.compress(d.astype(np.uint8).tobytes())


def _write_zip_exr(path, rgba, lines_per_block, half):
    """Independent OpenEXR writer: channels A, B, G, R, ZIP (16 lines) or ZIPS (1 line)."""
    h, w, _ = rgba.shape
    pt = 1 if half else 2
    dt = np.float16 if half else np.float32

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(data)) + data

    ch = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", pt, 0, 1, 1) for n in "ABGR") + b"\0"
    hdr = struct.pack("<II", 20000630, 2)
    hdr += attr("channels", "chlist", ch)
    hdr += attr("compression", "compression", bytes([3 if lines_per_block == 16 else 2]))
    hdr += attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("lineOrder", "lineOrder", b"\0")
    hdr += attr("pixelAspectRatio", "float", struct.pack("<f", 1))
    hdr += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    hdr += attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0"
    nblk = (h + lines_per_block - 1) // lines_per_block
    blocks = []
    for b in range(nblk):
        rows = rgba[b * lines_per_block:(b + 1) * lines_per_block]
        raw = b"".join(rows[r, :, c].astype(dt).tobytes() for r in range(rows.shape[0]) for c in (3, 2, 1, 0))
        z = _zip_block(raw)
        if len(z) >= len(raw):
            z = raw
        blocks.append(struct.pack("<ii", b * lines_per_block, len(z)) + z)
    offs = []
    pos = len(hdr) + 8 * nblk
    for blk in blocks:
        offs.append(pos)
        pos += len(blk)
    with open(path, "wb") as f:
        f.write(hdr + b"".join(struct.pack("<Q", o) for o in offs) + b"".join(blocks))


@pytest.mark.parametrize("lines,half", [(16, False), (16, True), (1, True)])
def test_exr_zip_read(tmp_path, lines, half):
    rng = np.random.default_rng(2)
    h, w = 37, 23
    rgba = rng.lognormal(0, 1, (h, w, 4)).astype(np.float32)
    rgba[5:9] = 0.25  # runs that compress
    p = str(tmp_path / "z.exr")
    _write_zip_exr(p, rgba, lines, half)
    back = P.exr_read(p)
    want = rgba.astype(np.float16).astype(np.float32) if half else rgba
    assert np.array_equal(back, want)


def test_exr_environment_and_errors(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.lognormal(0, 1, (8, 16, 3)).astype(np.float32)
    p = str(tmp_path / "env.exr")
    P.exr_write(p, img, half=False)
    s = P.Scene()
    s.load_exr_environment(p)
    i = s.info()
    assert (i.env_width, i.env_height) == (16, 8)
    env = s.export()["env_rgba"]
    assert np.array_equal(env[..., :3], img[::-1])
    bad = tmp_path / "bad.exr"
    bad.write_bytes(b"not an exr at all")
    with pytest.raises(P.PtError):
        P.exr_read(str(bad))


def test_reference_environment_assets():
    """The reference's own environment maps: wells6 (ZIP, FLOAT A/B/G/R, 1000x500) loads;
    the 1-pixel test map uses PIZ, which this reader reports as unsupported."""
    img = P.exr_read(os.path.join(SCENES, "assets", "20060807_wells6_hd.exr"))
    assert img.shape == (500, 1000, 4)
    assert np.isfinite(img).all() and (img[..., :3] >= 0).all() and img[..., :3].max() > 1
    with pytest.raises(P.PtError, match="compression 4"):
        P.exr_read(os.path.join(SCENES, "assets", "1_pixel_test.exr"))


def test_exr_crafted_offsets_and_windows_are_errors(tmp_path):
    """Block offsets near 2^64, block sizes past the end, block line numbers outside the data
    window and data windows whose width overflows 32 bits are reported as errors, not read."""
    rgba = np.ones((4, 3, 4), np.float32)
    good = str(tmp_path / "g.exr")
    _write_zip_exr(good, rgba, 16, False)
    data = open(good, "rb").read()
    # locate the offset table: the header ends with a zero byte after screenWindowWidth's float
    marker = b"screenWindowWidth\0float\0" + struct.pack("<I", 4) + struct.pack("<f", 1) + b"\0"
    tpos = data.index(marker) + len(marker)
    (boff,) = struct.unpack("<Q", data[tpos:tpos + 8])
    cases = {
        "offset near 2^64": data[:tpos] + struct.pack("<Q", (1 << 64) - 4) + data[tpos + 8:],
        "offset past the end": data[:tpos] + struct.pack("<Q", len(data) - 4) + data[tpos + 8:],
        "size past the end": data[:boff + 4] + struct.pack("<i", 1 << 30) + data[boff + 8:],
        "line outside the window": data[:boff] + struct.pack("<i", -(1 << 31)) + data[boff + 4:],
    }
    win = struct.pack("<4i", 0, 0, 2, 3)
    wpos = data.index(b"dataWindow\0box2i\0") + len(b"dataWindow\0box2i\0") + 4
    assert data[wpos:wpos + 16] == win
    cases["window width overflows int32"] = data[:wpos] + struct.pack("<4i", -(1 << 31), 0, (1 << 31) - 1, 3) + data[wpos + 16:]
    for name, blob in cases.items():
        p = tmp_path / "bad.exr"
        p.write_bytes(blob)
        with pytest.raises(P.PtError):
            P.exr_read(str(p))
    assert np.array_equal(P.exr_read(good), rgba)
