"""CPU: on-disk formats (OBJ/MTL load, PNG read/write, face-id text).

Two independent implementations -- the C++ one of the drop-in API
(ugu_amd/cpp/{mesh_io,image_io}.cc, driven through ugu_amd/cpp/ugu_io_tool)
and the Python one (ugu_amd/io.py) -- are compared with each other and
pinned by facts that come from neither: PIL's PNG codec, the OBJ files' own
text parsed with numpy, closed forms (unit normals, quad splits), and the
reference's documented quirks (src/mesh.cc:500-735, src/image_io.cc).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
from PIL import Image

from ugu_amd import io as uio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# UGU_IO_TOOL: another build of the same driver (tests/sanitize builds one
# under AddressSanitizer and UBSan)
TOOL = os.environ.get("UGU_IO_TOOL") or os.path.join(ROOT, "ugu_amd", "cpp", "ugu_io_tool")
DATA = os.path.join(ROOT, "tests", "golden", "data")

KINDS = {"1b": (np.uint8, 1), "3b": (np.uint8, 3), "4b": (np.uint8, 4), "1w": (np.uint16, 1)}


def tool(*args):
    assert os.path.exists(TOOL), "run make -C ugu_amd/cpp"
    return subprocess.run([TOOL, *map(str, args)], capture_output=True, text=True, timeout=60)


def cpp_read(path, kind, tmp):
    out = os.path.join(str(tmp), "read.raw")
    r = tool("pngread", path, out, kind)
    if r.returncode != 0:
        return None
    w, h = map(int, r.stdout.split())
    dt, ch = KINDS[kind]
    a = np.fromfile(out, dt)
    return a.reshape((h, w) if ch == 1 else (h, w, ch))


def cpp_load_obj(path, tmp):
    d = os.path.join(str(tmp), "obj")
    os.makedirs(d, exist_ok=True)
    r = tool("obj", path, d)
    assert r.returncode == 0, r.stderr
    meta = open(os.path.join(d, "meta.txt")).read().split("\n")
    head = list(map(int, meta[0].split()))
    rd = lambda n, dt, c: np.fromfile(os.path.join(d, n), dt).reshape(-1, c)
    arrays = {"vertices": rd("verts.f32", np.float32, 3), "vertex_indices": rd("faces.i32", np.int32, 3),
              "normals": rd("normals.f32", np.float32, 3), "normal_indices": rd("nidx.i32", np.int32, 3),
              "face_normals": rd("fnormals.f32", np.float32, 3), "uv": rd("uv.f32", np.float32, 2),
              "uv_indices": rd("uvidx.i32", np.int32, 3), "vertex_colors": rd("colors.f32", np.float32, 3),
              "material_ids": np.fromfile(os.path.join(d, "matids.i32"), np.int32)}
    mats = [ln.split() for ln in meta[1:1 + head[8]]]
    texs = []
    for i, m in enumerate(mats):
        w, h = int(m[2]), int(m[3])
        texs.append(np.fromfile(os.path.join(d, f"tex{i}.u8"), np.uint8).reshape(h, w, 3))
    return bool(head[0]), arrays, mats, texs


def assert_same_mesh(py, cpp_arrays, label):
    for k, v in cpp_arrays.items():
        got = getattr(py, k)
        assert got.shape == v.shape, f"{label} {k}: {got.shape} vs {v.shape}"
        assert np.array_equal(got.view(np.uint32) if got.dtype == np.float32 else got,
                              v.view(np.uint32) if v.dtype == np.float32 else v), f"{label} {k}"


# ---------------------------------------------------------------------------
# PNG


def test_png_decode_matches_pil(tmp_path):
    path = os.path.join(DATA, "spot_texture.png")
    ref = np.asarray(Image.open(path).convert("RGB"))
    assert np.array_equal(cpp_read(path, "3b", tmp_path), ref)
    assert np.array_equal(uio.imread(path), ref)


@pytest.mark.parametrize("kind", sorted(KINDS))
def test_png_write_read_roundtrip(tmp_path, kind):
    rng = np.random.default_rng(3)
    dt, ch = KINDS[kind]
    h, w = 37, 53
    shape = (h, w) if ch == 1 else (h, w, ch)
    img = rng.integers(0, np.iinfo(dt).max + 1, size=shape, dtype=np.int64).astype(dt)
    raw = os.path.join(str(tmp_path), "in.raw")
    img.tofile(raw)
    p_cpp, p_py = os.path.join(str(tmp_path), "c.png"), os.path.join(str(tmp_path), "p.png")
    assert tool("pngwrite", raw, w, h, kind, p_cpp).returncode == 0
    assert uio.imwrite(p_py, img)
    for p in (p_cpp, p_py):
        pil = np.asarray(Image.open(p))
        assert np.array_equal(pil.astype(dt), img), f"{kind} {p}: PIL decodes a different image"
        # read back by both implementations; an RGBA file needs
        # IMREAD_UNCHANGED (the tool uses it for Image4b): IMREAD_COLOR drops
        # the alpha channel
        assert np.array_equal(cpp_read(p, kind, tmp_path), img)
        if kind != "4b":
            assert np.array_equal(uio.imread(p), img)
        else:
            assert np.array_equal(uio.imread(p, uio.IMREAD_UNCHANGED), img)
            assert np.array_equal(uio.imread(p), img[..., :3])
            assert np.array_equal(cpp_read(p, "3b", tmp_path), img[..., :3])


def _pil_variants(tmp):
    """PNGs written by PIL in forms the writers here never produce, with the
    pixels the stb rules give (8-bit header -> 8 bit; other depths -> 16 bit,
    low depths scaled to 8 bits then x257)."""
    rng = np.random.default_rng(11)
    out = []
    g = rng.integers(0, 256, (19, 23), dtype=np.uint8)
    p = os.path.join(tmp, "gray_opt.png")
    Image.fromarray(g).save(p, optimize=True)          # adaptive filters
    out.append((p, "1b", g))
    rgb = rng.integers(0, 256, (17, 29, 3), dtype=np.uint8)
    p = os.path.join(tmp, "rgb_c9.png")
    Image.fromarray(rgb).save(p, compress_level=9)
    out.append((p, "3b", rgb))
    pal_idx = rng.integers(0, 7, (13, 11), dtype=np.uint8)
    palette = rng.integers(0, 256, (7, 3), dtype=np.uint8)
    im = Image.fromarray(pal_idx, "P")
    im.putpalette(palette.reshape(-1).tolist())
    p = os.path.join(tmp, "pal.png")
    im.save(p, bits=8)
    exp = palette[pal_idx]
    out.append((p, "3b" if Image.open(p).info.get("transparency") is None else "4b", exp))
    b = rng.integers(0, 2, (9, 21), dtype=np.uint8).astype(bool)
    p = os.path.join(tmp, "bits1.png")
    Image.fromarray(b).save(p)                          # 1-bit gray
    out.append((p, "1w", (b.astype(np.uint16) * 255) * 257))
    w16 = rng.integers(0, 65536, (8, 10), dtype=np.uint16)
    p = os.path.join(tmp, "gray16.png")
    Image.fromarray(w16.astype(np.int32), "I").convert("I;16").save(p)
    out.append((p, "1w", w16))
    rgba = rng.integers(0, 256, (6, 7, 4), dtype=np.uint8)
    p = os.path.join(tmp, "rgba.png")
    Image.fromarray(rgba, "RGBA").save(p)
    out.append((p, "3b", rgba[..., :3]))                # IMREAD_COLOR drops alpha
    return out


def test_png_decode_pil_variants(tmp_path):
    for p, kind, exp in _pil_variants(str(tmp_path)):
        got_c = cpp_read(p, kind, tmp_path)
        got_p = uio.imread(p)
        assert got_c is not None and np.array_equal(got_c, exp), f"C++ {os.path.basename(p)}"
        assert got_p is not None and np.array_equal(got_p, exp), f"Python {os.path.basename(p)}"


def test_png_rejects(tmp_path):
    bad = os.path.join(str(tmp_path), "bad.png")
    open(bad, "wb").write(b"not a png at all")
    assert cpp_read(bad, "3b", tmp_path) is None and uio.imread(bad) is None
    # channel mismatch: a gray PNG asked for as Image3b is an empty image
    g = os.path.join(str(tmp_path), "g.png")
    Image.fromarray(np.zeros((4, 4), np.uint8)).save(g)
    assert cpp_read(g, "3b", tmp_path) is None
    # the directory case used by LoadObj: opens, does not decode
    assert cpp_read(str(tmp_path) + "/", "3b", tmp_path) is None
    assert uio.imread(str(tmp_path) + "/") is None


def test_face_id_text(tmp_path):
    rng = np.random.default_rng(5)
    f = rng.integers(-1, 100000, (23, 31), dtype=np.int32)
    raw = os.path.join(str(tmp_path), "f.i32")
    f.tofile(raw)
    a, b = os.path.join(str(tmp_path), "a.txt"), os.path.join(str(tmp_path), "b.txt")
    assert tool("faceid", raw, 31, 23, a).returncode == 0
    uio.write_face_id_as_text(f, b)
    assert open(a).read() == open(b).read()
    assert np.array_equal(np.loadtxt(a, dtype=np.int64), f.reshape(-1))


# ---------------------------------------------------------------------------
# OBJ


def _text_arrays(path):
    v, vt, f = [], [], []
    for ln in open(path):
        t = ln.split()
        if not t:
            continue
        if t[0] == "v":
            v.append([float(x) for x in t[1:4]])
        elif t[0] == "vt":
            vt.append([float(x) for x in t[1:3]])
        elif t[0] == "f":
            f.append([int(x.split("/")[0]) - 1 for x in t[1:]])
    return np.array(v, np.float32), np.array(vt, np.float32), f


def test_load_obj_icosphere3(tmp_path):
    path = os.path.join(DATA, "icosphere3_smart_uv.obj")
    py, ok_py = uio.load_obj(path)
    ok_c, arrays, mats, _ = cpp_load_obj(path, tmp_path)
    assert_same_mesh(py, arrays, "icosphere3")
    # the MTL has no map_Kd: the texture path is the directory itself, which
    # opens and fails to decode, so LoadObj returns false with the mesh filled
    assert ok_c is False and ok_py is False
    assert mats[0][0] == "None" and py.obj_materials[0].name == "None"
    v, vt, f = _text_arrays(path)
    assert len(py.vertices) == 162 and len(py.vertex_indices) == 320
    assert np.array_equal(py.vertices, v)
    assert np.array_equal(py.vertex_indices, np.array(f, np.int32))
    assert np.array_equal(py.uv, vt)
    assert np.allclose(np.linalg.norm(py.normals, axis=1), 1.0, atol=1e-6)
    assert np.allclose(np.linalg.norm(py.face_normals, axis=1), 1.0, atol=1e-6)
    # per-face normals of the export point outwards like the face normals
    n_face = py.normals[py.normal_indices[:, 0]]
    assert (np.sum(n_face * py.face_normals, 1) > 0.9).all()


def test_load_obj_quads_with_texture(tmp_path):
    d = str(tmp_path)
    for n in ("spot.mtl", "spot_texture.png"):
        shutil.copy(os.path.join(DATA, n), d)
    src = open(os.path.join(DATA, "spot_quadrangulated.obj")).read()
    path = os.path.join(d, "spot_q.obj")
    open(path, "w").write("mtllib spot.mtl\nusemtl material000\n" + src)
    py, ok_py = uio.load_obj(path)
    ok_c, arrays, mats, texs = cpp_load_obj(path, tmp_path)
    assert ok_py and ok_c
    assert_same_mesh(py, arrays, "spot quads")
    ref_tex = np.asarray(Image.open(os.path.join(DATA, "spot_texture.png")).convert("RGB"))
    assert np.array_equal(texs[0], ref_tex) and np.array_equal(py.materials[0], ref_tex)
    v, vt, quads = _text_arrays(os.path.join(DATA, "spot_quadrangulated.obj"))
    assert len(py.vertex_indices) == 2 * len(quads)
    # each quad becomes two triangles on the shorter diagonal (tinyobjloader)
    for qi, q in enumerate(quads[:200]):
        t0, t1 = py.vertex_indices[2 * qi], py.vertex_indices[2 * qi + 1]
        P = v[q]
        d02 = np.float32(((P[2] - P[0]) ** 2).sum())
        d13 = np.float32(((P[3] - P[1]) ** 2).sum())
        if abs(d02 - d13) < 1e-6 * max(d02, d13):
            continue
        exp = ([q[0], q[1], q[2]], [q[0], q[2], q[3]]) if d02 < d13 else \
            ([q[0], q[1], q[3]], [q[1], q[2], q[3]])
        assert list(t0) == exp[0] and list(t1) == exp[1], f"quad {qi}"
    assert (py.material_ids == 0).all()
    assert len(py.normals) == len(py.vertices)      # CalcNormal: the file has no vn


def test_load_obj_missing_texture_file(tmp_path):
    path = os.path.join(DATA, "spot_remesh.obj")
    py, ok_py = uio.load_obj(path)
    ok_c, arrays, mats, _ = cpp_load_obj(path, tmp_path)
    assert ok_py and ok_c                              # LOGW only
    assert_same_mesh(py, arrays, "spot remesh")
    assert py.materials[0].size == 0


def test_load_obj_synthetic_rules(tmp_path):
    """Relative indices, vertex colours, an unreferenced vertex (removed, uv
    taken per vertex index as RemoveVerticesBase does, normals recomputed),
    an unknown usemtl (-1 kept, as in the reference), a missing mtllib."""
    d = str(tmp_path)
    obj = """# synthetic
mtllib nothere.mtl
v 0 0 0 1 0 0
v 1 0 0 0 1 0
v 0 1 0 0 0 1
v 5 5 5 0.5 0.5 0.5
v 1 1 0 0.2 0.4 0.6
vt 0 0
vt 1 0
vt 0 1
vt 0.5 0.5
vt 1 1
vn 0 0 2
usemtl nosuch
f 1/1/1 2/2/1 3/3/1
f -4/-4/-1 -1/-1/-1 -3/-3/-1
"""
    path = os.path.join(d, "s.obj")
    open(path, "w").write(obj)
    py, ok_py = uio.load_obj(path)
    ok_c, arrays, _, _ = cpp_load_obj(path, tmp_path)
    assert ok_py == ok_c
    assert_same_mesh(py, arrays, "synthetic")
    assert len(py.vertices) == 4                        # vertex 4 (5,5,5) removed
    assert np.array_equal(py.vertex_indices, np.array([[0, 1, 2], [1, 3, 2]], np.int32))
    assert np.allclose(py.vertex_colors[3], np.float32([0.2, 0.4, 0.6]) * np.float32(255))
    # uv re-indexed per vertex: uv[i] of each kept vertex i (0, 1, 2, 4)
    assert np.array_equal(py.uv, np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.float32))
    assert np.array_equal(py.uv_indices, py.vertex_indices)
    assert np.array_equal(py.normal_indices, py.vertex_indices)     # CalcNormal after removal


def test_load_obj_errors(tmp_path):
    d = str(tmp_path)
    p = os.path.join(d, "z.obj")
    open(p, "w").write("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n")   # index 0 is invalid
    assert uio.load_obj(p)[1] is False
    assert cpp_load_obj(p, tmp_path)[0] is False
    assert uio.load_obj(os.path.join(d, "none.obj"))[1] is False
