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


def test_reference_scene_2_one_texel_environment():
    """scene.cu:73-94: the plane of tests/plane.obj (white Lambertian) lit only by the
    PIZ-compressed 1_pixel_test.exr map."""
    s = P.Scene.reference(2, SCENES)
    i = s.info()
    assert (i.n_triangles, i.n_spheres, i.n_materials, i.n_emitters) == (2, 0, 1, 0)
    assert (i.env_width, i.env_height) == (1000, 500)
    env = s.export()["env_rgba"]
    # the environment's rows run from the +y pole; the file's row 239 is the lit one
    assert np.argwhere(env[..., :3].max(2) > 0).tolist() == [[239, 753]]


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
    """The reference's own environment maps: wells6 (ZIP, FLOAT A/B/G/R, 1000x500) and the
    1-pixel test map (PIZ, FLOAT B/G/R, 1000x500, 16 PIZ blocks of 32 lines): the latter decodes
    to black everywhere except one texel of radiance 1e4 (row 239, column 753), which a wrong
    Huffman table, bit order or wavelet step would not produce."""
    img = P.exr_read(os.path.join(SCENES, "assets", "20060807_wells6_hd.exr"))
    assert img.shape == (500, 1000, 4)
    assert np.isfinite(img).all() and (img[..., :3] >= 0).all() and img[..., :3].max() > 1
    one = P.exr_read(os.path.join(SCENES, "assets", "1_pixel_test.exr"))
    assert one.shape == (500, 1000, 4) and (one[..., 3] == 1).all()
    lit = np.argwhere(one[..., :3].max(2) != 0)
    assert lit.tolist() == [[239, 753]]
    assert one[239, 753, :3].tolist() == [1e4, 1e4, 1e4]


# ------------------------------------------------------------------ PIZ: independent encoder
# Written from the OpenEXR PIZ description (bitmap of the values present and their compaction
# table, 2D lifting wavelet per channel plane, canonical Huffman codes with a run-length
# pseudo-symbol), to check the reader on data that the reference's one PIZ file does not
# cover: HALF channels, several blocks, ragged block heights, both wavelet widths and runs.

def _piz_wavelet(v, nx, ox, ny, oy, maxv):
    """Forward 2D wavelet in place on a uint16 numpy array (element (x, y) at x*ox + y*oy)."""
    w14 = maxv < (1 << 14)

    def lift(a, b):
        if w14:
            a16, b16 = int(np.int16(np.uint16(a).view(np.int16))), int(np.int16(np.uint16(b).view(np.int16)))
            m = (a16 + b16) >> 1
            d = a16 - b16
            return np.uint16(m & 0xffff), np.uint16(d & 0xffff)
        ao = (int(a) + 0x8000) & 0xffff
        m = (ao + int(b)) >> 1
        d = ao - int(b)
        if d < 0:
            m = (m + 0x8000) & 0xffff
        return np.uint16(m), np.uint16(d & 0xffff)

    n = min(nx, ny)
    p, p2 = 1, 2
    while p2 <= n:
        y = 0
        while y + p2 <= ny:
            x = 0
            while x + p2 <= nx:
                i00 = x * ox + y * oy
                i01, i10 = i00 + p * ox, i00 + p * oy
                i11 = i10 + p * ox
                a00, a01 = lift(v[i00], v[i01])
                a10, a11 = lift(v[i10], v[i11])
                v[i00], v[i10] = lift(a00, a10)
                v[i01], v[i11] = lift(a01, a11)
                x += p2
            if nx & p:
                i00 = x * ox + y * oy
                i10 = i00 + p * oy
                v[i00], v[i10] = lift(v[i00], v[i10])
            y += p2
        if ny & p:
            x = 0
            while x + p2 <= nx:
                i00 = x * ox + y * oy
                i01 = i00 + p * ox
                v[i00], v[i01] = lift(v[i00], v[i01])
                x += p2
        p, p2 = p2, p2 << 1


class _BitWriter:
    def __init__(self):
        self.bits = []

    def put(self, value, n):
        self.bits += [(value >> (n - 1 - k)) & 1 for k in range(n)]

    def bytes(self):
        b = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(int("".join(map(str, b[i:i + 8])), 2) for i in range(0, len(b), 8))


def _huffman_lengths(freq):
    import heapq
    heap = [(f, i, (s,)) for i, (s, f) in enumerate(sorted(freq.items()))]
    heapq.heapify(heap)
    lengths = {s: 0 for s in freq}
    if len(heap) == 1:
        lengths[heap[0][2][0]] = 1
        return lengths
    tie = len(heap)
    while len(heap) > 1:
        f1, _, a = heapq.heappop(heap)
        f2, _, b = heapq.heappop(heap)
        for s in a + b:
            lengths[s] += 1
        heapq.heappush(heap, (f1 + f2, tie, a + b))
        tie += 1
    return lengths


def _huffman_encode(words):
    """Huffman stream of uint16 words: header, run-length-coded code lengths, codes; runs of
    one word longer than 3 repeats use the pseudo-symbol (max word + 1) and an 8-bit count."""
    words = [int(x) for x in words]
    rlc = max(words) + 1
    tokens = []
    i = 0
    while i < len(words):
        j = i + 1
        while j < len(words) and words[j] == words[i] and j - i - 1 < 255:
            j += 1
        run = j - i - 1
        tokens.append(words[i])
        if run > 3:
            tokens.append(("run", run))
        else:
            tokens += [words[i]] * run
        i = j
    freq = {}
    for t in tokens:
        s_ = rlc if isinstance(t, tuple) else t
        freq[s_] = freq.get(s_, 0) + 1
    lengths = _huffman_lengths(freq)
    im, iM = min(freq), rlc
    # canonical codes: first code per length from the longest down, symbols in order
    count = [0] * 60
    for s_ in range(im, iM + 1):
        count[lengths.get(s_, 0)] += 1
    count[0] = 0
    first, c = [0] * 60, 0
    for l in range(58, 0, -1):
        first[l] = c
        c = (c + count[l]) >> 1
    nxt, codes = list(first), {}
    for s_ in range(im, iM + 1):
        l = lengths.get(s_, 0)
        if l:
            codes[s_] = (nxt[l], l)
            nxt[l] += 1
    table = _BitWriter()
    s_ = im
    while s_ <= iM:
        l = lengths.get(s_, 0)
        if l == 0:
            run = 1
            while s_ + run <= iM and lengths.get(s_ + run, 0) == 0 and run < 261:
                run += 1
            if run >= 6:
                table.put(63, 6)
                table.put(run - 6, 8)
            elif run >= 2:
                table.put(59 + run - 2, 6)
            else:
                table.put(0, 6)
            s_ += run
            continue
        table.put(l, 6)
        s_ += 1
    data = _BitWriter()
    for t in tokens:
        if isinstance(t, tuple):
            data.put(*codes[rlc])
            data.put(t[1], 8)
        else:
            data.put(codes[t][0], codes[t][1])
    tb = table.bytes()
    return struct.pack("<5I", im, iM, len(tb), len(data.bits), 0) + tb + data.bytes()


def _piz_block(planes):
    """planes: list of (uint16 array of one channel's words for the block's lines, nx, words
    per pixel, ny).  Returns the PIZ block bytes."""
    allw = np.concatenate([p for p, *_ in planes])
    present = np.zeros(1 << 16, bool)
    present[allw] = True
    present[0] = False
    bitmap = np.packbits(present.reshape(-1, 8)[:, ::-1], axis=1).reshape(-1)
    nz = np.nonzero(bitmap)[0]
    lo, hi = (int(nz[0]), int(nz[-1])) if nz.size else (8191, 0)
    values = np.concatenate([[0], np.nonzero(present)[0]])
    fwd = np.zeros(1 << 16, np.uint16)
    fwd[values] = np.arange(values.size, dtype=np.uint16)
    maxv = values.size - 1
    coded = []
    for words, nx, wd, ny in planes:
        v = fwd[words].copy()
        for j in range(wd):  # each interleaved word of a pixel is its own plane
            _piz_wavelet(v[j:], nx, wd, ny, nx * wd, maxv)
        coded.append(v)
    huf = _huffman_encode(np.concatenate(coded))
    out = struct.pack("<HH", lo, hi)
    if lo <= hi:
        out += bitmap[lo:hi + 1].tobytes()
    return out + struct.pack("<I", len(huf)) + huf, maxv


def _write_piz_exr(path, rgba, half):
    h, w, _ = rgba.shape
    pt = 1 if half else 2
    dt = np.float16 if half else np.float32

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(data)) + data

    names = "ABGR"
    ch = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", pt, 0, 1, 1) for n in names) + b"\0"
    hdr = struct.pack("<II", 20000630, 2)
    hdr += attr("channels", "chlist", ch)
    hdr += attr("compression", "compression", bytes([4]))
    hdr += attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("lineOrder", "lineOrder", b"\0")
    hdr += attr("pixelAspectRatio", "float", struct.pack("<f", 1))
    hdr += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    hdr += attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0"
    nblk = (h + 31) // 32
    blocks, kinds = [], []
    for b in range(nblk):
        rows = rgba[b * 32:(b + 1) * 32]
        ny = rows.shape[0]
        planes = []
        for c in (3, 2, 1, 0):
            words = np.frombuffer(rows[:, :, c].astype(dt).tobytes(), np.uint16).copy()
            planes.append((words, w, 1 if half else 2, ny))
        blk, maxv = _piz_block(planes)
        raw = b"".join(rows[r, :, c].astype(dt).tobytes() for r in range(ny) for c in (3, 2, 1, 0))
        # a block exactly as long as its raw data is read as raw data, so it is never kept;
        # PIZ blocks longer than raw are legal and are kept to exercise the codec
        assert len(blk) != len(raw)
        kinds.append(("piz", 14 if maxv < (1 << 14) else 16))
        blocks.append(struct.pack("<ii", b * 32, len(blk)) + blk)
    offs, pos = [], len(hdr) + 8 * nblk
    for blk in blocks:
        offs.append(pos)
        pos += len(blk)
    with open(path, "wb") as f:
        f.write(hdr + b"".join(struct.pack("<Q", o) for o in offs) + b"".join(blocks))
    return kinds


@pytest.mark.parametrize("half,h,w,kind,bits", [(True, 37, 23, "smooth", 14), (False, 70, 19, "smooth", 14), (True, 33, 8, "noise", 14),
                                                (False, 40, 256, "noise", 16), (True, 64, 16, "flat", 14)])
def test_exr_piz_read(tmp_path, half, h, w, kind, bits):
    """PIZ files from the independent encoder above decode to the exact pixels: HALF and FLOAT
    channels, several 32-line blocks with a ragged last one, smooth data (few distinct values:
    14-bit wavelet), noise (many values: 16-bit wavelet) and constant runs (run-length codes)."""
    rng = np.random.default_rng(h * 100 + w)
    y, x = np.mgrid[0:h, 0:w]
    if kind == "smooth":
        base = (np.sin(x / 5.0) + np.cos(y / 7.0) + 2.5).astype(np.float32)
        rgba = np.stack([base, base * 0.5, base * 0.25, np.ones_like(base)], -1)
        rgba = np.round(rgba * 8) / 8
    elif kind == "noise":
        rgba = rng.lognormal(0, 2, (h, w, 4)).astype(np.float32)
    else:
        rgba = np.full((h, w, 4), 0.75, np.float32)
        rgba[10:20, 3:9, 0] = 3.0
    rgba = rgba.astype(np.float32)
    p = str(tmp_path / "p.exr")
    kinds = _write_piz_exr(p, rgba, half)
    # the first block went through the codec, with the wavelet width the case is meant for
    assert kinds[0] == ("piz", bits), kinds
    back = P.exr_read(p)
    want = rgba.astype(np.float16).astype(np.float32) if half else rgba
    assert back.shape == (h, w, 4)
    assert np.array_equal(back, want)


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
