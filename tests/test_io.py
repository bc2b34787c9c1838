"""Input formats: Wavefront OBJ (OBJParser) and Radiance RGBE (HDRBitmapReader).

Both reference readers are in un-vendored jars, so parity with them is
unpinned; these tests pin the readers to the published formats: known
answers for records, index forms, fan triangulation and groups (OBJ), and
byte-level known answers plus encode -> decode round trips through all three
scanline encodings (RGBE)."""
import math

import numpy as np
import pytest

from mambatracer_amd import _lib, api

BOX = """# a closed box, one object per side pair, mixed index forms
o floor
v -1 -1 -1
v  1 -1 -1
v  1 -1  1
v -1 -1  1
vn 0 1 0
f 1//1 2//1 3//1 4//1
o ceiling
v -1 1 -1
v  1 1 -1
v  1 1  1
v -1 1  1
vt 0 0
vt 1 0
vt 1 1
vn 0 -1 0
s off
f 5/1/2 8/3/2 7/2/2
f -4/1 -2/3 -3/2
o empty_object
o back
usemtl wall
f 3 4 8 7
g
f 1 2 3
"""


def test_obj_box_known_answers():
    p, n, t, f, groups = api.OBJParser().readString(BOX)
    assert p.shape == (8, 4) and n.shape == (2, 4) and t.shape == (3, 2)
    np.testing.assert_array_equal(p[0], [-1, -1, -1, 0])
    np.testing.assert_array_equal(p[6], [1, 1, 1, 0])
    np.testing.assert_array_equal(n[1], [0, -1, 0, 0])
    np.testing.assert_array_equal(t[2], [1, 1])
    # quad 1 2 3 4 -> fan (1,2,3), (1,3,4), zero-based, normals 0, no uvs
    np.testing.assert_array_equal(f[0], [0, 1, 2, -1, -1, -1, 0, 0, 0, 0])
    np.testing.assert_array_equal(f[1], [0, 2, 3, -1, -1, -1, 0, 0, 0, 0])
    np.testing.assert_array_equal(f[2], [4, 7, 6, 0, 2, 1, 1, 1, 1, 1])
    # negative indices count back from the last vertex / uv so far
    np.testing.assert_array_equal(f[3], [4, 6, 5, 0, 2, 1, -1, -1, -1, 1])
    # `o empty_object` has no faces and gets no material; `o back` + `usemtl`
    # in a row start one group (2); a later `g` starts group 3
    np.testing.assert_array_equal(f[4], [2, 3, 7, -1, -1, -1, -1, -1, -1, 2])
    np.testing.assert_array_equal(f[5], [2, 7, 6, -1, -1, -1, -1, -1, -1, 2])
    np.testing.assert_array_equal(f[6], [0, 1, 2, -1, -1, -1, -1, -1, -1, 3])
    assert f.shape[0] == 7 and groups == 4


def test_obj_polygon_fan_and_scene_build():
    # a regular 12-gon: fan of 10 triangles, all with positive area, and the
    # parsed arrays build a valid BVH
    k = 12
    lines = [f"v {math.cos(2 * math.pi * i / k):.6f} {math.sin(2 * math.pi * i / k):.6f} 0" for i in range(k)]
    lines.append("f " + " ".join(str(i + 1) for i in range(k)))
    p, n, t, f, groups = api.OBJParser().readString("\n".join(lines))
    assert f.shape[0] == k - 2 and groups == 1
    assert (f[:, 0] == 0).all() and (f[:, 1] == np.arange(1, k - 1)).all() and (f[:, 2] == np.arange(2, k)).all()
    a = p[f[:, 1], :3] - p[f[:, 0], :3]
    b = p[f[:, 2], :3] - p[f[:, 0], :3]
    assert (np.cross(a, b)[:, 2] > 0).all()
    nodes, bounds, depth = api.CNormalBVH.build(p, f)
    assert nodes.shape[0] == 2 * f.shape[0] - 1


@pytest.mark.parametrize("bad", ["v 1 2\n", "f 1 2\n", "v 0 0 0\nf 1 2 3\n", "v 0 0 0\nf 1/x 1 1\n"])
def test_obj_errors(bad):
    with pytest.raises(_lib.MambartError):
        api.OBJParser().readString(bad)


# ---------------------------------------------------------------- RGBE
def to_rgbe(rgb):
    """Ward's float -> RGBE (frexp of the largest component)."""
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    v = rgb.max(-1)
    ok = v > 1e-32
    m, e = np.frexp(np.where(ok, v, 1.0))
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    out[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return out


def from_rgbe(c):
    c = c.astype(np.float64)
    f = np.where(c[..., 3] > 0, np.ldexp(1.0, (c[..., 3] - 136).astype(int)), 0.0)
    return np.concatenate([c[..., :3] * f[..., None], np.ones(c.shape[:-1] + (1,))], -1)


def _rle_channel(b):
    out = bytearray()
    i, n = 0, len(b)
    while i < n:
        r = 1
        while i + r < n and r < 127 and b[i + r] == b[i]:
            r += 1
        if r >= 4:
            out += bytes([128 + r, b[i]])
            i += r
            continue
        j = i
        while j < n and j - i < 128:
            if j + 3 < n and b[j] == b[j + 1] == b[j + 2] == b[j + 3]:
                break
            j += 1
        out += bytes([j - i]) + bytes(b[i:j])
        i = j
    return bytes(out)


def encode_hdr(px, mode):
    """px: uint8[h, w, 4] RGBE.  mode: 'rle' (adaptive), 'flat', 'old' (old runs)."""
    h, w = px.shape[:2]
    data = bytearray(b"#?RADIANCE\n# test\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n")
    data += f"-Y {h} +X {w}\n".encode()
    for y in range(h):
        row = px[y]
        if mode == "rle":
            data += bytes([2, 2, w >> 8, w & 255])
            for ch in range(4):
                data += _rle_channel(bytes(row[:, ch]))
        elif mode == "flat":
            data += row.tobytes()
        else:  # old-style: a pixel then (1,1,1,count) for repeats of it
            x = 0
            while x < w:
                data += row[x].tobytes()
                r = 0
                while x + 1 + r < w and (row[x + 1 + r] == row[x]).all() and r < 255:
                    r += 1
                if r:
                    data += bytes([1, 1, 1, r])
                x += 1 + r
    return bytes(data)


def test_rgbe_known_answers():
    px = np.array([[[128, 64, 0, 129], [255, 255, 255, 136], [7, 9, 11, 0], [1, 2, 3, 120]]], np.uint8)
    img = api.HDRBitmapReader().decode(encode_hdr(px, "flat"))
    assert img.shape == (1, 4, 4)
    np.testing.assert_array_equal(img[0, 0], [1.0, 0.5, 0.0, 1.0])          # 128 * 2^-7
    np.testing.assert_array_equal(img[0, 1], [255.0, 255.0, 255.0, 1.0])    # 2^0
    np.testing.assert_array_equal(img[0, 2], [0.0, 0.0, 0.0, 1.0])          # exponent 0: black
    np.testing.assert_array_equal(img[0, 3], [2.0 ** -16, 2.0 ** -15, 3 * 2.0 ** -16, 1.0])  # m * 2^(120-136)


@pytest.mark.parametrize("mode", ["rle", "flat", "old"])
def test_rgbe_round_trip(mode):
    rng = np.random.default_rng(3)
    h, w = 37, 301
    rgb = rng.lognormal(0.0, 2.0, size=(h, w, 3))
    rgb[5:9, 20:200] = 3.25            # long runs for the run encodings
    rgb[10, :] = 0.0
    px = to_rgbe(rgb)
    img = api.HDRBitmapReader().decode(encode_hdr(px, mode))
    np.testing.assert_array_equal(img, from_rgbe(px).astype(np.float32))
    # RGBE keeps ~8 bits of the largest component
    big = rgb.max(-1) > 1e-30
    rel = np.abs(img[..., :3][big] - rgb[big]) / rgb.max(-1)[big][:, None]
    assert rel.max() < 2.0 ** -7


def test_rgbe_environment_map_round_trip():
    env = __import__("mambatracer_amd.workloads", fromlist=["env_map"]).env_map(256, 128)
    px = to_rgbe(env[..., :3].astype(np.float64))
    img = api.HDRBitmapReader().decode(encode_hdr(px, "rle"))
    assert img.shape == (128, 256, 4)
    np.testing.assert_array_equal(img, from_rgbe(px).astype(np.float32))


@pytest.mark.parametrize("bad", [b"P6\n", b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x00\x00\x00\x00",
                                 b"#?RADIANCE\n\n+Y 1 +X 1\n\x00\x00\x00\x00", b"#?RADIANCE\n\n-Y 2 +X 1\n\x01\x02\x03\x80"])
def test_rgbe_errors(bad):
    with pytest.raises(_lib.MambartError):
        api.HDRBitmapReader().decode(bad)
