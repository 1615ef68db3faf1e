"""Test infrastructure only: a Python restatement of the reference's glTF
loader (src/scene/scene_load_gltf.cpp:61-529) and of the viewer's camera set
up (src/main.cpp:170-190, src/app/render_loop.cpp:86-117), used as the
checker for lx_scene_load_glb.  Nothing in the product imports this module.

Transforms are composed in float64 (numpy) where the reference composes
them in float32, so they are compared with a tolerance; the camera is
restated in float32 in the reference's operation order.  Integer data
(indices, material and texture indices, decoded 8/16-bit texels) compare
exactly.
PNG and JPEG texels come from PIL (PNG is lossless, so exact; JPEG decoders
differ in rounding, so JPEG is compared with a tolerance).  EXR: ZIP / ZIPS /
uncompressed scanline files are decoded here with zlib; PIZ has no decoder in
this image: PIZ files written by the tests' format-restated PIZ encoder
(tests/gltf_util.py) must decode back exactly, and the reference's own PIZ
HDRIs are checked by invariants.
"""
import io
import json
import math
import struct
import zlib

import numpy as np

try:
    from PIL import Image
except ImportError:  # pragma: no cover
    Image = None


def _u32(b, o):
    return struct.unpack_from("<I", b, o)[0]


def parse_glb(path):
    data = open(path, "rb").read()
    assert _u32(data, 0) == 0x46546C67 and _u32(data, 4) == 2
    i, js, bin_ = 12, None, None
    while i + 8 <= len(data):
        ln, ty = _u32(data, i), _u32(data, i + 4)
        chunk = data[i + 8:i + 8 + ln]
        if ty == 0x4E4F534A:
            js = json.loads(chunk.decode("utf-8"))
        elif ty == 0x004E4942 and bin_ is None:
            bin_ = chunk
        i += 8 + ((ln + 3) & ~3)
    return js, bin_


def _view(js, bin_, v):
    bv = js["bufferViews"][v]
    off = bv.get("byteOffset", 0)
    return bin_[off:off + bv["byteLength"]], bv.get("byteStride", 0)


def _accessor(js, bin_, a, ncomp, dtype):
    acc = js["accessors"][a]
    count = acc["count"]
    buf, stride = _view(js, bin_, acc["bufferView"])
    off = acc.get("byteOffset", 0)
    es = np.dtype(dtype).itemsize * ncomp
    stride = stride or es
    if count == 0:
        return np.zeros((0, ncomp), dtype)
    raw = np.frombuffer(buf, np.uint8, (count - 1) * stride + es, off)
    rows = np.lib.stride_tricks.as_strided(raw, (count, es), (stride, 1))
    return np.ascontiguousarray(rows).view(dtype).reshape(count, ncomp)


def _indices(js, bin_, a):
    acc = js["accessors"][a]
    dt = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32}[acc["componentType"]]
    return _accessor(js, bin_, a, 1, dt).reshape(-1).astype(np.uint32)


def _node_matrix(node):
    """glTF local matrix (column-major list or T * R * S), as a numpy 4x4 in
    row-major mathematical form."""
    if "matrix" in node:
        return np.array(node["matrix"], np.float64).reshape(4, 4).T
    M = np.eye(4)
    if "translation" in node:
        M[:3, 3] = node["translation"]
    if "rotation" in node:
        x, y, z, w = node["rotation"]
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        M[:3, :3] = M[:3, :3] @ R
    if "scale" in node:
        M[:3, :3] = M[:3, :3] @ np.diag(node["scale"])
    return M


def decode_image(raw):
    """Four channels, 8 or 16 bits, as tinygltf asks of its image decoder."""
    im = Image.open(io.BytesIO(raw))
    if im.format == "PNG" and im.mode in ("I;16", "I;16B", "I"):
        a = np.asarray(im).astype(np.uint16)
        out = np.empty(a.shape + (4,), np.uint16)
        out[..., 0] = out[..., 1] = out[..., 2] = a
        out[..., 3] = 65535
        return out, 16
    if im.format == "PNG" and im.mode in ("RGB", "RGBA") and im.info.get("bitdepth", 8) == 16:
        raise NotImplementedError("PIL cannot return 16-bit RGB(A)")
    return np.asarray(im.convert("RGBA")), 8


def load(path):
    """Returns dict(meshes=[(pos, nrm|None, uv|None, idx)], instances=[(mesh, material, 3x4)],
    materials=[dict], textures=[(array, bits, fmt)], camera=dict, trace_depth=8)."""
    js, bin_ = parse_glb(path)
    textures = []
    for t in js.get("textures", []):
        img = js["images"][t["source"]]
        raw, _ = _view(js, bin_, img["bufferView"])
        textures.append(decode_image(bytes(raw)) + (img.get("mimeType", ""),))

    def xf(info):
        tt = (info or {}).get("extensions", {}).get("KHR_texture_transform", {})
        return dict(offset=tuple(tt.get("offset", (0.0, 0.0))), scale=tuple(tt.get("scale", (1.0, 1.0))),
                    rotation=tt.get("rotation", 0.0))

    mats = [dict(base_color=(1.0, 0.0, 0.0), roughness=0.0, metallic=0.0, ior=0.0, transmission=0.0,
                 emission_color=(0.0, 0.0, 0.0), emission_strength=0.0, alpha=1.0, alpha_mode=0, alpha_cutoff=0.5,
                 albedo_tex=-1, metallic_roughness_tex=-1, normal_tex=-1, emissive_tex=-1, albedo_xf=xf(None),
                 normal_xf=xf(None), emissive_xf=xf(None), metallic_roughness_xf=xf(None), material_type=0)]

    def is_glass(m, base_alpha):
        """isGlass, include/scene/gltf_material_utils.h:9-47."""
        ext = m.get("extensions", {})
        if "KHR_materials_transmission" in ext or "KHR_materials_volume" in ext:
            return True
        if "ior" in ext.get("KHR_materials_ior", {}) and abs(float(ext["KHR_materials_ior"]["ior"]) - 1.0) > 0.01:
            return True
        if m.get("alphaMode", "OPAQUE") in ("BLEND", "MASK") and base_alpha < float(np.float32(0.99)):
            return True
        df = ext.get("KHR_materials_pbrSpecularGlossiness", {}).get("diffuseFactor")
        return isinstance(df, list) and len(df) == 4 and df[3] < 0.99

    def legacy_type(d, m, base_alpha):
        """Material class of the UBER_SHADER 0 loader branch
        (src/scene/scene_load_gltf.cpp:211-228): 3 emissive, 4 glass, 0 microfacet."""
        e = np.asarray(d["emission_color"], np.float32)
        ln = np.sqrt(np.float32(np.float32(e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]))
        if np.float32(np.float32(d["emission_strength"]) * ln) > np.float32(1e-5):
            return 3
        if is_glass(m, base_alpha) and np.float32(d["metallic"]) < np.float32(0.01) and \
                np.float32(d["roughness"]) < np.float32(0.01):
            return 4
        return 0
    for m in js.get("materials", []):
        pbr = m.get("pbrMetallicRoughness", {})
        bc = pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0])
        ext = m.get("extensions", {})
        mats.append(dict(
            base_color=tuple(bc[:3]), roughness=pbr.get("roughnessFactor", 1.0),
            metallic=pbr.get("metallicFactor", 1.0),
            ior=ext.get("KHR_materials_ior", {}).get("ior", 1.55),
            transmission=ext.get("KHR_materials_transmission", {}).get("transmissionFactor", 0.0),
            emission_color=tuple(m.get("emissiveFactor", [0.0, 0.0, 0.0])),
            emission_strength=ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0),
            alpha=bc[3], alpha_mode={"OPAQUE": 0, "MASK": 1, "BLEND": 2}[m.get("alphaMode", "OPAQUE")],
            alpha_cutoff=m.get("alphaCutoff", 0.5),
            albedo_tex=pbr.get("baseColorTexture", {}).get("index", -1),
            metallic_roughness_tex=pbr.get("metallicRoughnessTexture", {}).get("index", -1),
            normal_tex=m.get("normalTexture", {}).get("index", -1),
            emissive_tex=m.get("emissiveTexture", {}).get("index", -1),
            albedo_xf=xf(pbr.get("baseColorTexture")), normal_xf=xf(m.get("normalTexture")),
            emissive_xf=xf(m.get("emissiveTexture")),
            metallic_roughness_xf=xf(pbr.get("metallicRoughnessTexture"))))
        mats[-1]["material_type"] = legacy_type(mats[-1], m, bc[3])

    meshes, insts, mesh_of = [], [], {}
    cam = []

    def visit(ni, parent):
        node = js["nodes"][ni]
        G = parent @ _node_matrix(node)
        if "mesh" in node:
            mesh = js["meshes"][node["mesh"]]
            for pi, prim in enumerate(mesh["primitives"]):
                if prim.get("mode", 4) != 4 or "POSITION" not in prim["attributes"]:
                    continue
                key = (node["mesh"], pi)
                if key not in mesh_of:
                    at = prim["attributes"]
                    P = _accessor(js, bin_, at["POSITION"], 3, np.float32)
                    N = None
                    if "NORMAL" in at:
                        N = _accessor(js, bin_, at["NORMAL"], 3, np.float32).astype(np.float64)
                        N /= np.linalg.norm(N, axis=1, keepdims=True)
                    T = _accessor(js, bin_, at["TEXCOORD_0"], 2, np.float32) if "TEXCOORD_0" in at else None
                    I = _indices(js, bin_, prim["indices"]) if "indices" in prim else np.arange(len(P), dtype=np.uint32)
                    I = I[:len(I) // 3 * 3]
                    if len(I) == 0:
                        continue
                    mesh_of[key] = len(meshes)
                    meshes.append((P, N, T, I))
                pm = prim.get("material", -1)
                insts.append((mesh_of[key], pm + 1 if pm >= 0 else 0, G[:3, :4].copy()))
        if "camera" in node:
            cam.append((ni, G))
        for c in node.get("children", []):
            visit(c, G)

    for n in js["scenes"][js.get("scene", 0)]["nodes"]:
        visit(n, np.eye(4))
    assert len(cam) == 1, "exactly one camera"
    ni, G = cam[0]
    camera = _camera(js["cameras"][js["nodes"][ni]["camera"]]["perspective"], G)
    return dict(meshes=meshes, instances=insts, materials=mats, textures=textures, camera=camera, trace_depth=8)


def _camera(persp, G):
    """The camera in float32, in the reference's operation order (glm vec3
    float arithmetic and float libm calls): scene_load_gltf.cpp:457-495, then
    the spherical re-derivation with world up of main.cpp:178-189 and
    render_loop.cpp:91-116.  float32 matters: for a camera looking straight
    down, the sign of `right` follows float32 rounding of sin(pi)."""
    f = np.float32
    G = G.astype(np.float32)

    def norm(v):
        d = f(f(v[0] * v[0]) + f(v[1] * v[1])) + f(v[2] * v[2])
        return (v * f(f(1.0) / np.sqrt(f(d)))).astype(np.float32)

    def length(v):
        return np.sqrt(f(f(f(v[0] * v[0]) + f(v[1] * v[1])) + f(v[2] * v[2])))

    def cross(a, b):
        return np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]], np.float32)

    aspect = f(persp["aspectRatio"])
    if aspect >= 1.0:
        h = 1000
        w = int(f(aspect * f(h)))
    else:
        w = 1000
        h = int(f(f(w) / aspect))
    fovy = f(persp["yfov"] * (180.0 / math.pi))
    eye = G[:3, 3].copy()
    view = -norm(G[:3, 2])
    look = (eye + view).astype(np.float32)
    pi = f(math.pi)
    ys = np.tan(f(f(f(0.5) * fovy) * f(pi / f(180.0))))
    xs = f(f(ys * f(w)) / f(h))
    v = norm(view)
    phi = np.arctan2(v[0], -v[2])
    theta = np.arccos(np.clip(v[1], f(-1.0), f(1.0)))
    zoom = length(eye - look)
    sph = np.array([f(f(-zoom * np.sin(phi)) * np.sin(theta)), f(-zoom * np.cos(theta)),
                    f(f(zoom * np.cos(phi)) * np.sin(theta))], np.float32)
    v = -norm(sph)
    ref = np.array([0.0, 1.0, 0.0], np.float32)
    dv = f(f(f(ref[0] * v[0]) + f(ref[1] * v[1])) + f(ref[2] * v[2]))
    u = (ref - v * dv).astype(np.float32)
    if f(length(u)) ** 2 < 1e-6:
        u = ref.copy()
    u = norm(u)
    r = norm(cross(v, u))
    u = cross(r, v)
    return dict(width=w, height=h, position=(look - v * zoom).astype(np.float32), view=v, up=u, right=r,
                pixel_length=(f(f(f(2.0) * xs) / f(w)), f(f(f(2.0) * ys) / f(h))))


def _half(u16):
    return u16.view(np.float16).astype(np.float32)


def load_exr_zip(path):
    """Scanline EXR with NONE, ZIPS or ZIP compression, HALF / FLOAT channels,
    to RGBA32F with LoadEXR's channel rules."""
    b = open(path, "rb").read()
    assert _u32(b, 0) == 20000630
    i = 8
    chans, comp, dw = [], None, None
    while True:
        e = b.index(b"\0", i)
        name = b[i:e].decode()
        i = e + 1
        if not name:
            break
        e = b.index(b"\0", i)
        typ = b[i:e].decode()
        i = e + 1
        sz = _u32(b, i)
        i += 4
        val = b[i:i + sz]
        i += sz
        if name == "channels":
            k = 0
            while val[k] != 0:
                q = val.index(b"\0", k)
                chans.append((val[k:q].decode(), _u32(val, q + 1)))
                k = q + 17
        elif name == "compression":
            comp = val[0]
        elif name == "dataWindow":
            dw = struct.unpack("<4i", val)
    assert comp in (0, 2, 3), "this checker reads NONE / ZIPS / ZIP only"
    W, H = dw[2] - dw[0] + 1, dw[3] - dw[1] + 1
    lpb = 16 if comp == 3 else 1
    nb = (H + lpb - 1) // lpb
    offs = struct.unpack_from("<%dQ" % nb, b, i)
    sizes = [2 if t == 1 else 4 for _, t in chans]
    planes = {n: np.zeros((H, W), np.float32) for n, _ in chans}
    for off in offs:
        y, ln = struct.unpack_from("<iI", b, off)
        raw = b[off + 8:off + 8 + ln]
        lines = min(lpb, H - (y - dw[1]))
        expect = lines * W * sum(sizes)
        if ln != expect:
            t = np.frombuffer(zlib.decompress(raw), np.uint8).astype(np.int32)
            t = (np.cumsum(t - 128) + 128) % 256  # predictor: t[k] += t[k-1] - 128, with t[0] kept
            t[0] = np.frombuffer(zlib.decompress(raw), np.uint8)[0]
            t = t.astype(np.uint8)
            half = (expect + 1) // 2
            out = np.empty(expect, np.uint8)
            out[0::2] = t[:half][:len(out[0::2])]
            out[1::2] = t[half:]
            raw = out.tobytes()
        p = 0
        for l in range(lines):
            for (n, t), s in zip(chans, sizes):
                row = np.frombuffer(raw, np.uint16 if t == 1 else np.float32, W, p)
                planes[n][y - dw[1] + l] = _half(row) if t == 1 else row
                p += W * s
    rgba = np.empty((H, W, 4), np.float32)
    if len(chans) == 1:
        rgba[...] = planes[chans[0][0]][..., None]
    else:
        rgba[..., 0], rgba[..., 1], rgba[..., 2] = planes["R"], planes["G"], planes["B"]
        rgba[..., 3] = planes["A"] if "A" in planes else 1.0
    return rgba
