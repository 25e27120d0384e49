"""On-disk formats (host code in libhana_gpu.so, no GPU needed): the OBJ and
TGA loaders and the TGA writer against the reference's own Model and
TGAImage code compiled into oracle/_ref/libhanaref.so, on the reference's
assets and on synthetic files that exercise every branch (raw and RLE data,
1/3/4 bytes per texel, both row orders, mirrored files, quads, CRLF, normals
that are not fixed points of normalize()).  Full frames of the two assets
are then rendered by the oracle from the loaded data and compared with the
reference's own face loop.

Tests that need the reference tree or its compiled library skip without it."""
import os

import numpy as np
import pytest

from hana_amd import gpu
from hana_amd import scenegen as sg
from pyoracle import REF_SO, CPURenderer, HostTarget, write_obj

ASSETS = "/root/reference/Hana-SoftwareRenderer/assets"
NAMES = ("african_head", "diablo3_pose")

need_ref = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref/libhanaref.so not built")
need_assets = pytest.mark.skipif(not os.path.isdir(ASSETS), reason="reference assets not present")


@pytest.fixture(scope="module")
def ref():
    return CPURenderer("reference")


def corners(m):
    c = m.idx.reshape(-1)
    return m.pos[c], m.nrm[c], m.uv[c]


def bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


# ---------------------------------------------------------------- OBJ

@need_ref
@need_assets
@pytest.mark.parametrize("name", NAMES)
def test_obj_assets_match_model(ref, name):
    path = os.path.join(ASSETS, name, name + ".obj")
    m = gpu.load_obj(path)
    p, n, u = ref.model_corners(path)
    gp, gn, gu = corners(m)
    assert bits_equal(gp, p) and bits_equal(gn, n) and bits_equal(gu, u)
    assert m.nverts < 3 * m.nfaces          # corners sharing (v, vt, fetched vn) share a vertex


@need_ref
def test_obj_synthetic_edge_cases(ref, tmp_path):
    """Normals far from unit length (so every fetch renormalises), a quad
    (only its first three corners count), CRLF, comments, blank and unknown
    lines, extra spaces, exponents."""
    rng = np.random.default_rng(7)
    V = 40
    pos = rng.normal(size=(V, 3)).astype(np.float32)
    nrm = (rng.normal(size=(V, 3)) * rng.uniform(0.2, 5.0, size=(V, 1))).astype(np.float32)
    uv = rng.uniform(0, 1, size=(V, 2)).astype(np.float32)
    faces = rng.integers(0, V, size=(120, 4))
    path = tmp_path / "edge.obj"
    with open(path, "w", newline="") as f:
        f.write("# comment\r\n\r\no thing\r\ng group\r\n")
        for p in pos:
            f.write("v  %.9g %.9e   %.9g\r\n" % tuple(float(x) for x in p))
        for t in uv:
            f.write("vt %.9g %.9g 0.0\r\n" % tuple(float(x) for x in t))
        for n in nrm:
            f.write("vn %.9g %.9g %.9g\r\n" % tuple(float(x) for x in n))
        f.write("s off\r\nusemtl none\r\n")
        for i, q in enumerate(faces + 1):
            if i % 3 == 0:   # a quad
                f.write("f %d/%d/%d %d/%d/%d %d/%d/%d %d/%d/%d\r\n" % tuple(int(q[k // 3]) for k in range(12)))
            else:            # mixed v / vt / vn indices
                a, b, c, d = (int(x) for x in q)
                f.write(f"f {a}/{b}/{c} {b}/{c}/{d} {c}/{d}/{a}\r\n")
    m = gpu.load_obj(path)
    p, n, u = ref.model_corners(str(path))
    gp, gn, gu = corners(m)
    assert m.nfaces == len(faces)
    assert bits_equal(gp, p) and bits_equal(gu, u)
    assert bits_equal(gn, n)       # the fetch-by-fetch renormalisation sequence


def test_obj_roundtrip_of_generated_mesh(tmp_path):
    sc = sg.small_scene(sg.BLINN, seed=3)
    path = tmp_path / "m.obj"
    write_obj(sc.mesh, path)
    m = gpu.load_obj(path)
    a, b = corners(m), (sc.mesh.pos[sc.mesh.idx.reshape(-1)], sc.mesh.nrm[sc.mesh.idx.reshape(-1)],
                        sc.mesh.uv[sc.mesh.idx.reshape(-1)])
    assert all(bits_equal(x, y) for x, y in zip(a, b))
    assert m.nverts <= sc.mesh.nverts


def test_obj_rejects_bad_faces(tmp_path):
    for body in ("v 0 0 0\nvt 0 0\nvn 0 0 1\nf 1/1/1 1/1/1\n",             # two corners
                 "v 0 0 0\nvt 0 0\nvn 0 0 1\nf 1/1/1 2/1/1 1/1/1\n",       # past the vertex list
                 "v 0 0 0\nvt 0 0\nvn 0 0 1\nf 1//1 1//1 1//1\n"):         # no uv index
        path = tmp_path / "bad.obj"
        path.write_text(body)
        with pytest.raises(gpu.HanaError):
            gpu.load_obj(path)
    with pytest.raises(gpu.HanaError):
        gpu.load_obj(tmp_path / "missing.obj")


# ---------------------------------------------------------------- TGA

@need_ref
@need_assets
@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("suffix", ["_diffuse.tga", "_nm_tangent.tga", "_spec.tga", "_nm.tga"])
def test_tga_assets_match_tgaimage(ref, name, suffix):
    path = os.path.join(ASSETS, name, name + suffix)
    for flip in (False, True):
        assert np.array_equal(gpu.load_tga(path, flip), ref.tga_read(path, flip))


def synthetic_image(rng, h, w, bpp):
    """Runs, literal stretches, packets longer than 128 and single pixels."""
    shape = (h, w) if bpp == 1 else (h, w, bpp)
    img = rng.integers(0, 256, size=shape, dtype=np.uint8)
    flat = img.reshape(h * w, -1)
    flat[5:300] = flat[5]               # a run across rows, longer than 128
    flat[400:402] = flat[400]           # a pair of equal pixels
    flat[-130:] = flat[-130]            # a run up to the last pixel
    flat[1000:1129] = flat[999]         # a run of 130 straddling a literal
    return img


@need_ref
@pytest.mark.parametrize("bpp", [1, 3, 4])
@pytest.mark.parametrize("rle", [False, True])
def test_tga_writer_is_byte_identical(ref, tmp_path, bpp, rle):
    rng = np.random.default_rng(bpp * 10 + rle)
    img = synthetic_image(rng, 37, 61, bpp)
    ours, theirs = tmp_path / "ours.tga", tmp_path / "ref.tga"
    gpu.save_tga(ours, img, rle)
    assert ref.tga_write(str(theirs), img, rle)
    assert ours.read_bytes() == theirs.read_bytes()
    assert np.array_equal(gpu.load_tga(ours), img)
    assert np.array_equal(ref.tga_read(str(ours)), img)


def tga_bytes(img, datatype, descriptor, rle_packets=None):
    """A TGA file written by hand: 18-byte header, then raw pixels or the
    given RLE packet stream."""
    h, w = img.shape[:2]
    bpp = 1 if img.ndim == 2 else img.shape[2]
    hdr = np.zeros(18, np.uint8)
    hdr[2] = datatype
    hdr[12:14] = np.frombuffer(np.int16(w).tobytes(), np.uint8)
    hdr[14:16] = np.frombuffer(np.int16(h).tobytes(), np.uint8)
    hdr[16] = bpp * 8
    hdr[17] = descriptor
    body = rle_packets if rle_packets is not None else img.tobytes()
    return hdr.tobytes() + body


def rle_encode_simple(img):
    """Packets of one kind at a time (runs of 1-128, literals of 1-128) --
    a different but valid encoding from the writer's."""
    bpp = 1 if img.ndim == 2 else img.shape[2]
    px = img.reshape(-1, bpp)
    out = bytearray()
    i = 0
    while i < len(px):
        j = i
        while j + 1 < len(px) and j + 1 - i < 128 and np.array_equal(px[j + 1], px[i]):
            j += 1
        if j > i:
            out.append(127 + (j - i + 1))
            out += px[i].tobytes()
            i = j + 1
        else:
            k = min(len(px), i + 1 + (i % 7))
            out.append(k - i - 1)
            out += px[i:k].tobytes()
            i = k
    return bytes(out)


@need_ref
@pytest.mark.parametrize("bpp", [1, 3, 4])
@pytest.mark.parametrize("descriptor", [0x00, 0x20, 0x10, 0x30])
def test_tga_reader_orientations(ref, tmp_path, bpp, descriptor):
    rng = np.random.default_rng(bpp + descriptor)
    img = synthetic_image(rng, 29, 47, bpp)
    dt_raw, dt_rle = (3, 11) if bpp == 1 else (2, 10)
    for i, data in enumerate((tga_bytes(img, dt_raw, descriptor),
                              tga_bytes(img, dt_rle, descriptor, rle_encode_simple(img)))):
        path = tmp_path / f"o{i}.tga"
        path.write_bytes(data)
        for flip in (False, True):
            ours, theirs = gpu.load_tga(path, flip), ref.tga_read(str(path), flip)
            assert np.array_equal(ours, theirs), (i, flip)
    expect = img
    if not descriptor & 0x20:
        expect = expect[::-1]
    if descriptor & 0x10:
        expect = expect[:, ::-1]
    assert np.array_equal(gpu.load_tga(tmp_path / "o0.tga"), expect)


def test_tga_rejects_malformed(tmp_path):
    img = np.arange(64, dtype=np.uint8).reshape(8, 8)
    bad = {
        "runs past the image": tga_bytes(img, 11, 0x20, bytes([255, 7]) + bytes([127]) + bytes(128)),
        "truncated RLE": tga_bytes(img, 11, 0x20, bytes([10, 1, 2])),
        "short raw data": tga_bytes(img, 3, 0x20)[:-5],
        "unknown type": tga_bytes(img, 1, 0x20),
        "bad bpp": tga_bytes(img, 3, 0x20)[:16] + bytes([16, 0x20]) + img.tobytes(),
    }
    for what, data in bad.items():
        path = tmp_path / "bad.tga"
        path.write_bytes(data)
        with pytest.raises(gpu.HanaError):
            gpu.load_tga(path)
            raise AssertionError(what)


# ---------------------------------------------------------------- frames of the real assets

def asset_scene(name, shader, shadow, cam):
    d = gpu.load_tga(os.path.join(ASSETS, name, name + "_diffuse.tga"), True)       # Model::load_texture
    n = gpu.load_tga(os.path.join(ASSETS, name, name + "_nm_tangent.tga"), True)
    m = gpu.load_obj(os.path.join(ASSETS, name, name + ".obj"))
    return sg.Scene(m, 1000, 600, shader, shadow, cam, (2.0, 2.0, 2.0), d, n, name=name)


@need_ref
@need_assets
@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("shader,shadow", [(sg.NORMALMAP, True), (sg.BLINN, False), (sg.TEXTURE_LIGHT, False)])
def test_asset_frames_match_reference(ref, name, shader, shadow):
    """1000x600 frames of the reference's own assets, loaded by the library's
    loaders: the oracle equals the reference's face loop bit for bit."""
    o = CPURenderer("oracle")
    sc = asset_scene(name, shader, shadow, (0.0, 0.0, 1.5))
    a = o.render_passes(sc)
    b = ref.render_passes(sc)
    for x, y in ((a[0], b[0]), (a[1], b[1])):
        if x is None:
            continue
        assert np.array_equal(x.color[..., :3], y.color[..., :3]) and bits_equal(x.depth, y.depth)
    assert a[3]["clipped_faces"] > 0 and a[3]["fragments"] == b[3]["fragments"]


@need_ref
@need_assets
def test_asset_unmodified_entry(ref, tmp_path):
    """The reference's unmodified graphics_draw_triangle on its own Model
    (textures from its own TGAImage reader) against the oracle on the
    library-loaded data, for a camera with no clipped face."""
    name = "african_head"
    o = CPURenderer("oracle")
    sc = asset_scene(name, sg.BLINN, False, (0.0, 0.0, 3.0))
    u = o.build_uniforms(sc)
    t_ref, t_or = HostTarget(1000, 600), HostTarget(1000, 600)
    d = ref.tga_read(os.path.join(ASSETS, name, name + "_diffuse.tga"), True)
    n = ref.tga_read(os.path.join(ASSETS, name, name + "_nm_tangent.tga"), True)
    ref.draw_model(os.path.join(ASSETS, name, name + ".obj"), sg.BLINN, u, t_ref, d, n)
    st, _ = o.draw(sc.mesh, sg.BLINN, u, t_or, sc.diffuse, sc.normal)
    assert st["clipped_faces"] == 0 and st["fragments"] > 50_000
    assert np.array_equal(t_ref.color[..., :3], t_or.color[..., :3]) and bits_equal(t_ref.depth, t_or.depth)
