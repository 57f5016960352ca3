"""On-disk formats: OBJ/MTL loading, PNG read/write and face-id text.

Python counterparts of the C++ functions in ugu_amd/cpp/{mesh_io,image_io}.cc,
which follow UGU's Mesh::LoadObj (src/mesh.cc:500-735), imread / imwrite
(src/image_io.cc) and WriteFaceIdAsText (src/util/io_util.cc:22-33). The two
implementations are independent and compared by tests/test_io.py.

* ``load_obj(path, mtl_dir="") -> (Mesh, ok)``: triangles (quads split on the
  shorter diagonal, larger polygons as a fan), v/vt/vn with relative indices,
  vertex colours (x255, kept only when every vertex has them), per-face
  materials, diffuse textures via ``imread``; unreferenced vertices removed as
  Mesh::RemoveVertices does. ``ok`` is LoadObj's return value: as in the
  reference it is False when a material's texture path opens but does not
  decode (this includes a material without map_Kd, whose path is the
  directory itself).
* ``imread(path, flags=IMREAD_COLOR)``: PNG as stbi_load (8-bit header) or
  stbi_load_16 (any other depth), own channel count, IMREAD_COLOR drops alpha.
* ``imwrite(path, image)``: PNG (uint8 with 1/3/4 channels, uint16 with 1) or
  raw ``.bin``.
"""
import os
import struct
import sys
import zlib

import numpy as np

from .types import Mesh, _normalize_rows_f32

IMREAD_UNCHANGED = -1
IMREAD_GRAYSCALE = 0
IMREAD_COLOR = 1


def _log(msg):
    sys.stderr.write(msg if msg.endswith("\n") else msg + "\n")


# ---------------------------------------------------------------------------
# PNG


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def _unfilter(raw, h, row, bpp):
    out = np.zeros((h, row), np.uint8)
    prev = np.zeros(row, np.int32)
    for y in range(h):
        ft = raw[y * (row + 1)]
        src = np.frombuffer(raw, np.uint8, row, y * (row + 1) + 1).astype(np.int32)
        if ft == 0:
            cur = src
        elif ft == 2:
            cur = (src + prev) & 255
        else:
            cur = np.zeros(row, np.int32)
            for i in range(row):  # left-dependent filters run per byte
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if ft == 1:
                    v = src[i] + a
                elif ft == 3:
                    v = src[i] + ((a + b) >> 1)
                elif ft == 4:
                    v = src[i] + int(_paeth(a, b, c))
                else:
                    raise ValueError("unknown row filter")
                cur[i] = v & 255
        out[y] = cur
        prev = cur
    return out


def _decode_png(data, out8):
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, plte, trns, hdr = 8, b"", b"", b"", None
    while pos + 12 <= len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body[:13])
        elif typ == b"PLTE":
            plte = body
        elif typ == b"tRNS":
            trns = body
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
        pos += 12 + n
    if hdr is None:
        raise ValueError("missing IHDR")
    w, h, bd, ct, _, _, interlace = hdr
    if interlace:
        raise ValueError("interlaced PNG is not supported")
    n = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ct]
    row = (w * n * bd + 7) // 8
    bpp = max(1, n * bd // 8)
    raw = zlib.decompress(idat)
    img = _unfilter(raw, h, row, bpp)
    if bd < 8:
        bits = np.unpackbits(img, axis=1)[:, :w * bd].reshape(h, w, bd)
        s = (bits * (1 << np.arange(bd - 1, -1, -1))).sum(-1).astype(np.uint32)[..., None]
        if ct != 3:
            s = s * (255 // ((1 << bd) - 1))
    elif bd == 16:
        s = img.reshape(h, w * n, 2).astype(np.uint32)
        s = ((s[..., 0] << 8) | s[..., 1]).reshape(h, w, n)
    else:
        s = img.reshape(h, w, n).astype(np.uint32)
    if ct == 3:
        pal = np.frombuffer(plte, np.uint8).reshape(-1, 3).astype(np.uint32)
        idx = s[..., 0]
        if idx.max() >= len(pal):
            raise ValueError("palette index out of range")
        s = pal[idx]
        if trns:
            alpha = np.full(len(pal), 255, np.uint32)
            t = np.frombuffer(trns, np.uint8)[:len(pal)]
            alpha[:len(t)] = t
            s = np.concatenate([s, alpha[idx][..., None]], -1)
    elif trns and ct in (0, 2):
        key = np.array([(trns[2 * k] << 8) | trns[2 * k + 1] if 2 * k + 1 < len(trns) else 0
                        for k in range(n)], np.uint32)
        if bd < 8:
            key = key * (255 // ((1 << bd) - 1))
        full = 65535 if bd == 16 else 255
        alpha = np.where((s == key).all(-1), 0, full).astype(np.uint32)
        s = np.concatenate([s, alpha[..., None]], -1)
    if bd == 8 or out8:
        if bd == 16:
            s = s >> 8
        return s.astype(np.uint8)
    if bd < 8:
        s = s * 257
    return s.astype(np.uint16)


def imread(path, flags=IMREAD_COLOR):
    """Image array (H, W, C) or (H, W) for one channel; None when it cannot be
    read (as an empty ugu::Image)."""
    try:
        with open(path, "rb") as f:
            data = f.read()
        ext = os.path.splitext(path)[1]
        img = _decode_png(data, out8=ext not in (".png", ".PNG"))
    except (OSError, ValueError, KeyError, zlib.error) as e:
        _log(f"failed to load {path}: {e}")
        return None
    if flags == IMREAD_COLOR and img.shape[-1] == 4 and img.dtype == np.uint8:
        img = img[..., :3]
    img = np.ascontiguousarray(img)
    return img[..., 0] if img.shape[-1] == 1 else img


def imwrite(path, image):
    """imwrite (src/image_io.cc:182-206) for .png and .bin; returns bool."""
    a = np.ascontiguousarray(image)
    ext = os.path.splitext(path)[1]
    if ext in (".bin", ".BIN"):
        a.tofile(path)
        return True
    if ext not in (".png", ".PNG"):
        _log(f"acceptable extention is .png, .jpg or .jpeg. this extention is not supported: "
             f"{path}")
        return False
    ch = 1 if a.ndim == 2 else a.shape[2]
    if a.dtype == np.uint16 and ch != 1:
        _log("1 byte per channel is required to save by stb_image")
        return False
    if a.dtype not in (np.uint8, np.uint16) or ch not in (1, 3, 4) or a.size == 0:
        _log("unsupported image for PNG")
        return False
    h, w = a.shape[:2]
    body = a.astype(">u2").tobytes() if a.dtype == np.uint16 else a.tobytes()
    row = len(body) // h
    raw = b"".join(b"\x00" + body[y * row:(y + 1) * row] for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    ct = {1: 0, 3: 2, 4: 6}[ch]
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h,
                                                               8 * a.itemsize, ct, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)
    return True


def write_face_id_as_text(face_id, path):
    """WriteFaceIdAsText (src/util/io_util.cc:22-33): one id per line, row order."""
    with open(path, "w") as f:
        f.write("".join(f"{int(v)}\n" for v in np.asarray(face_id).reshape(-1)))


# ---------------------------------------------------------------------------
# OBJ / MTL


class ObjMaterial:
    """ObjMaterial (include/ugu/mesh.h:26-50) fields filled by LoadObj."""

    def __init__(self, name="default_mat"):
        self.name = name
        self.ambient = [0.117647, 0.117647, 0.117647]
        self.diffuse = [0.752941, 0.752941, 0.752941]
        self.specular = [0.752941, 0.752941, 0.752941]
        self.shininess = 8.0
        self.dissolve = 1.0
        self.illum = 1
        self.diffuse_texname = ""
        self.diffuse_texpath = ""
        self.diffuse_tex = np.zeros((0, 0, 3), np.uint8)


def _load_mtl(path):
    mats = []
    try:
        f = open(path)
    except OSError:
        return None
    cur = None
    with f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            k = t[0]
            if k == "newmtl":
                if cur is not None:
                    mats.append(cur)
                cur = ObjMaterial(t[1] if len(t) > 1 else "")
                cur.ambient, cur.diffuse, cur.specular = [0.0] * 3, [0.0] * 3, [0.0] * 3
                cur.shininess, cur.dissolve, cur.illum = 1.0, 1.0, 0
            elif cur is None:
                continue
            elif k in ("Ka", "Kd", "Ks") and len(t) >= 4:
                vals = [float(np.float32(float(x))) for x in t[1:4]]
                setattr(cur, {"Ka": "ambient", "Kd": "diffuse", "Ks": "specular"}[k], vals)
            elif k == "Ns" and len(t) >= 2:
                cur.shininess = float(np.float32(float(t[1])))
            elif k == "d" and len(t) >= 2:
                cur.dissolve = float(np.float32(float(t[1])))
            elif k == "Tr" and len(t) >= 2:
                cur.dissolve = float(np.float32(1.0) - np.float32(float(t[1])))
            elif k == "illum" and len(t) >= 2:
                cur.illum = int(t[1])
            elif k == "map_Kd" and len(t) >= 2:
                cur.diffuse_texname = t[-1]
    if cur is not None:
        mats.append(cur)
    return mats


def _fix(idx, n):
    if idx > 0:
        return idx - 1
    if idx < 0:
        return n + idx
    raise ValueError("zero index")


def _remove_vertices(m, valid):
    """Mesh::RemoveVertices (src/mesh.cc:1066-1164, RemoveVerticesBase :2124-2186)."""
    with_uv = len(m.uv) > 0 and len(m.uv_indices) > 0
    with_vc = len(m.vertex_colors) > 0
    table = -np.ones(len(m.vertices), np.int64)
    keep = np.nonzero(valid)[0]
    table[keep] = np.arange(len(keep))
    mapped = table[m.vertex_indices]
    ok = (mapped >= 0).all(1)
    faces = mapped[ok].astype(np.int32)
    face_table = -np.ones(len(m.vertex_indices), np.int64)
    face_table[ok] = np.arange(ok.sum())
    if with_uv:
        uv = np.zeros((len(keep), 2), np.float32)
        inb = keep < len(m.uv)
        uv[inb] = m.uv[keep[inb]]
        m.uv, m.uv_indices = uv, faces.copy()
    if with_vc:
        vc = np.zeros((len(keep), 3), np.float32)
        inb = keep < len(m.vertex_colors)
        vc[inb] = m.vertex_colors[keep[inb]]
        m.vertex_colors = vc
    m.vertices = m.vertices[keep]
    m.vertex_indices = faces
    m.CalcNormal()
    ids = np.zeros(len(faces), np.int32)
    sel = face_table[:len(m.material_ids)] >= 0
    ids[face_table[:len(m.material_ids)][sel]] = m.material_ids[sel]
    if len(ids):
        m.material_ids = ids
    return int(len(valid) - len(keep))


def load_obj(obj_path, mtl_dir=""):
    """Mesh::LoadObj; returns (Mesh, ok). Mesh.materials holds the diffuse
    textures (empty arrays where none loaded); Mesh.obj_materials the
    ObjMaterial records."""
    m = Mesh()
    try:
        lines = open(obj_path).read().splitlines()
    except OSError:
        _log(f"Cannot open file [{obj_path}]")
        return m, False
    d = mtl_dir if mtl_dir else (os.path.dirname(obj_path) + "/" if os.path.dirname(obj_path)
                                 else "./")
    v, vn, vt, vc = [], [], [], []
    all_colors = True
    tri, tri_mat = [], []
    mats, mat_ids, cur = [], {}, -1
    for ln, line in enumerate(lines, 1):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        k = t[0]
        if k == "v":
            x = [0.0, 0.0, 0.0, 1.0, 1.0, 1.0]
            for i, s in enumerate(t[1:7]):
                x[i] = float(s)
            v.append(x[:3])
            vc.append(x[3:6])
            all_colors = all_colors and len(t) >= 7
        elif k == "vn":
            vn.append([float(s) for s in (t[1:4] + ["0"] * 3)[:3]])
        elif k == "vt":
            vt.append([float(s) for s in (t[1:3] + ["0"] * 2)[:2]])
        elif k == "f":
            poly = []
            try:
                for s in t[1:]:
                    p = s.split("/")
                    iv = _fix(int(p[0]), len(v))
                    it = _fix(int(p[1]), len(vt)) if len(p) > 1 and p[1] else -1
                    inn = _fix(int(p[2]), len(vn)) if len(p) > 2 and p[2] else -1
                    poly.append((iv, it, inn))
            except ValueError:
                _log(f"Failed parse `f' line (e.g. zero value for face index. line {ln}.)")
                return m, False
            if len(poly) < 3:
                continue
            if len(poly) == 4:
                P = np.array([v[p[0]] if 0 <= p[0] < len(v) else [0, 0, 0] for p in poly],
                             np.float32)
                e02, e13 = P[2] - P[0], P[3] - P[1]
                s02 = np.float32(0)
                s13 = np.float32(0)
                for c in range(3):
                    s02 = np.float32(s02 + np.float32(e02[c] * e02[c]))
                    s13 = np.float32(s13 + np.float32(e13[c] * e13[c]))
                order = [(0, 1, 2), (0, 2, 3)] if s02 < s13 else [(0, 1, 3), (1, 2, 3)]
            else:
                order = [(0, i, i + 1) for i in range(1, len(poly) - 1)]
            for a, b, c in order:
                tri += [poly[a], poly[b], poly[c]]
                tri_mat.append(cur)
        elif k == "usemtl":
            cur = mat_ids.get(t[1] if len(t) > 1 else "", -1)
        elif k == "mtllib":
            found = None
            for name in t[1:]:
                found = _load_mtl(d + name)
                if found is not None:
                    break
            if found is None:
                _log(f"Material file [ {t[1] if len(t) > 1 else ''} ] not found.")
            else:
                for mt in found:
                    mat_ids[mt.name] = len(mats)
                    mats.append(mt)
    v = np.array(v, np.float32).reshape(-1, 3)
    vn = np.array(vn, np.float32).reshape(-1, 3)
    vt = np.array(vt, np.float32).reshape(-1, 2)
    vc = np.array(vc, np.float32).reshape(-1, 3)
    tri = np.array(tri, np.int64).reshape(-1, 3, 3)
    nf = len(tri)
    if nf:
        bad = ((tri[..., 0] < 0) | (tri[..., 0] >= len(v))).any()
        if len(vn):
            bad |= ((tri[..., 2] < 0) | (tri[..., 2] >= len(vn))).any()
        if len(vt):
            bad |= ((tri[..., 1] < 0) | (tri[..., 1] >= len(vt))).any()
        if bad:
            _log("face index out of range")
            return m, False
    m.vertices = v.copy() if nf == 0 else np.zeros_like(v)
    valid = np.zeros(len(v), bool)
    if nf:
        fv = tri[..., 0]
        m.vertices[fv.reshape(-1)] = v[fv.reshape(-1)]
        valid[fv.reshape(-1)] = True
        m.vertex_indices = fv.astype(np.int32)
        m.material_ids = np.array(tri_mat, np.int32)
    else:
        m.vertex_indices = np.zeros((0, 3), np.int32)
        m.material_ids = np.zeros(0, np.int32)
    if len(vn):
        n = np.zeros_like(vn)
        ni = tri[..., 2].reshape(-1) if nf else np.arange(min(len(vn), len(v)))
        n[ni] = _normalize_rows_f32(vn[ni])
        m.normals = n
        if nf:
            m.normal_indices = tri[..., 2].astype(np.int32)
    if len(vt):
        uv = np.zeros_like(vt)
        ti = tri[..., 1].reshape(-1) if nf else np.arange(min(len(vt), len(v)))
        uv[ti] = vt[ti]
        m.uv = uv
        if nf:
            m.uv_indices = tri[..., 1].astype(np.int32)
    if all_colors and len(v):
        c = np.zeros_like(vc)
        ci = tri[..., 0].reshape(-1) if nf else np.arange(len(v))
        c[ci] = vc[ci] * np.float32(255.0)
        m.vertex_colors = c
    m.CalcFaceNormal()
    if len(m.normals) == 0:
        m.CalcNormal()
    ok = True
    if not mats:
        mats = [ObjMaterial()]
        m.material_ids = np.zeros(nf, np.int32)
        _log("Default material was added because material did not find on input obj")
    else:
        for mt in mats:
            mt.diffuse_texpath = d + mt.diffuse_texname
            if os.path.exists(mt.diffuse_texpath):
                img = imread(mt.diffuse_texpath)
                if img is None or img.ndim != 3 or img.shape[2] != 3 or img.dtype != np.uint8:
                    img = np.zeros((0, 0, 3), np.uint8)
                mt.diffuse_tex = img
                ok = img.size > 0
            else:
                _log(f"diffuse texture doesn't exist {mt.diffuse_texpath}")
    m.obj_materials = mats
    m.materials = [mt.diffuse_tex for mt in mats]
    if nf and not valid.all():
        _remove_vertices(m, valid)
    return m, ok
