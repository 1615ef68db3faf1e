"""Writers for small synthetic glTF binaries and OpenEXR files used by the
loader tests (test data generation only)."""
import io
import json
import struct
import zlib

import numpy as np


def write_glb(path, images=(), mime="image/png", camera_rot=(0.0, 0.0, 0.0, 1.0), extra_materials=None,
              node_xf=None):
    """One textured quad (two triangles) per image, a camera at z = 3 looking
    down -z, and materials referencing the textures."""
    bin_ = bytearray()
    views, accessors = [], []

    def add(data, target=None):
        while len(bin_) % 4:
            bin_.append(0)
        off = len(bin_)
        bin_.extend(data)
        v = {"buffer": 0, "byteOffset": off, "byteLength": len(data)}
        if target:
            v["target"] = target
        views.append(v)
        return len(views) - 1

    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    nrm = np.array([[0, 0, 2]] * 4, np.float32)  # unnormalised on purpose
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    for arr, typ, ct in ((pos, "VEC3", 5126), (nrm, "VEC3", 5126), (uv, "VEC2", 5126), (idx, "SCALAR", 5123)):
        v = add(arr.tobytes())
        accessors.append({"bufferView": v, "componentType": ct, "count": len(arr), "type": typ})
    imgs, texs, mats = [], [], []
    for k, im in enumerate(images):
        v = add(im)
        imgs.append({"bufferView": v, "mimeType": mime})
        texs.append({"source": k})
        mats.append({"pbrMetallicRoughness": {"baseColorTexture": {
            "index": k, "extensions": {"KHR_texture_transform": {"offset": [0.25, 0.5], "scale": [2, 3],
                                                                  "rotation": 0.3}}}}})
    mats += list(extra_materials or [])
    meshes = [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 3}]}
              for k in range(max(1, len(mats)))]
    for k, m in enumerate(meshes):
        if mats:
            m["primitives"][0]["material"] = k
    nodes = [{"camera": 0, "translation": [0, 0, 3], "rotation": list(camera_rot)}]
    for k in range(len(meshes)):
        n = {"mesh": k, "translation": [3.0 * k, 0.0, 0.0]}
        if node_xf:
            n.update(node_xf)
        nodes.append(n)
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": list(range(len(nodes)))}],
           "nodes": nodes, "meshes": meshes, "materials": mats, "textures": texs, "images": imgs,
           "accessors": accessors, "bufferViews": views, "buffers": [{"byteLength": len(bin_)}],
           "cameras": [{"type": "perspective", "perspective": {"aspectRatio": 1.5, "yfov": 0.7, "znear": 0.1}}]}
    if not texs:
        del doc["textures"], doc["images"]
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    while len(bin_) % 4:
        bin_.append(0)
    total = 12 + 8 + len(js) + 8 + len(bin_)
    with open(path, "wb") as f:
        f.write(struct.pack("<III", 0x46546C67, 2, total))
        f.write(struct.pack("<II", len(js), 0x4E4F534A) + js)
        f.write(struct.pack("<II", len(bin_), 0x004E4942) + bytes(bin_))


def _attr(name, typ, val):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(val)) + val


def write_exr(path, planes, compression, half=True):
    """Scanline EXR; planes: dict name -> (H, W) float32.  compression 0
    (NONE), 1 (RLE), 2 (ZIPS) or 3 (ZIP)."""
    names = sorted(planes)
    H, W = planes[names[0]].shape
    chl = b""
    for n in names:
        chl += n.encode() + b"\0" + struct.pack("<iB3xii", 1 if half else 2, 0, 1, 1)
    chl += b"\0"
    hdr = struct.pack("<II", 20000630, 2)
    hdr += _attr("channels", "chlist", chl)
    hdr += _attr("compression", "compression", bytes([compression]))
    hdr += _attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1))
    hdr += _attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1))
    hdr += _attr("lineOrder", "lineOrder", b"\0")
    hdr += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    hdr += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    hdr += b"\0"
    lpb = 16 if compression == 3 else 1
    nb = (H + lpb - 1) // lpb
    blocks = []
    for b in range(nb):
        y0 = b * lpb
        raw = b""
        for y in range(y0, min(H, y0 + lpb)):
            for n in names:
                row = planes[n][y]
                raw += (row.astype(np.float16).tobytes() if half else row.astype(np.float32).tobytes())
        data = raw
        if compression in (1, 2, 3):
            t = np.frombuffer(raw, np.uint8)
            inter = np.concatenate([t[0::2], t[1::2]])
            d = inter.astype(np.int32)
            pred = d.copy()
            pred[1:] = (d[1:] - d[:-1] + 128) % 256
            pb = pred.astype(np.uint8).tobytes()
            if compression == 1:
                out = b""
                i = 0
                while i < len(pb):
                    j = i
                    while j + 1 < len(pb) and pb[j + 1] == pb[i] and j - i < 126:
                        j += 1
                    if j - i >= 2:
                        out += bytes([j - i]) + pb[i:i + 1]
                        i = j + 1
                    else:
                        k = i
                        while k < len(pb) and k - i < 127 and not (k + 2 < len(pb) and pb[k] == pb[k + 1] == pb[k + 2]):
                            k += 1
                        k = max(k, i + 1)
                        out += bytes([(256 - (k - i)) & 255]) + pb[i:k]
                        i = k
                comp = out
            else:
                comp = zlib.compress(pb)
            if len(comp) < len(raw):
                data = comp
        blocks.append(struct.pack("<iI", y0, len(data)) + data)
    off = len(hdr) + 8 * nb
    table = b""
    for blk in blocks:
        table += struct.pack("<Q", off)
        off += len(blk)
    with open(path, "wb") as f:
        f.write(hdr + table + b"".join(blocks))


def png_bytes(img, **kw):
    from PIL import Image
    buf = io.BytesIO()
    img.save(buf, format="PNG", **kw)
    return buf.getvalue()


def jpeg_bytes(img, **kw):
    buf = io.BytesIO()
    img.save(buf, format="JPEG", **kw)
    return buf.getvalue()
