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


def _wavelet_encode(P, wide):
    """Forward integer Haar wavelet of the OpenEXR PIZ format over the 2-D
    uint16 plane P (in place), finest level first: in each 2h x 2h cell,
    rows first, then columns; cells cut by the right / bottom edge get only
    their vertical / horizontal pairs.  wide: 16-bit modular lifting (else
    14-bit)."""
    def lift(a, b):
        if not wide:
            a16, b16 = a.astype(np.int16).astype(np.int32), b.astype(np.int16).astype(np.int32)
            return ((a16 + b16) >> 1).astype(np.uint16), (a16 - b16).astype(np.uint16)
        ao = (a.astype(np.int32) + 0x8000) & 0xffff
        m = (ao + b.astype(np.int32)) >> 1
        d = ao - b.astype(np.int32)
        m = np.where(d < 0, (m + 0x8000) & 0xffff, m)
        return m.astype(np.uint16), (d & 0xffff).astype(np.uint16)

    ny, nx = P.shape
    h = 1
    while 2 * h <= min(nx, ny):
        s = 2 * h
        cx, cy = nx // s * s, ny // s * s
        if cx and cy:
            p00, p01 = P[0:cy:s, 0:cx:s].copy(), P[0:cy:s, h:cx:s].copy()
            p10, p11 = P[h:cy:s, 0:cx:s].copy(), P[h:cy:s, h:cx:s].copy()
            i00, i01 = lift(p00, p01)
            i10, i11 = lift(p10, p11)
            P[0:cy:s, 0:cx:s], P[h:cy:s, 0:cx:s] = lift(i00, i10)
            P[0:cy:s, h:cx:s], P[h:cy:s, h:cx:s] = lift(i01, i11)
        if nx & h and cy:
            P[0:cy:s, cx], P[h:cy:s, cx] = lift(P[0:cy:s, cx].copy(), P[h:cy:s, cx].copy())
        if ny & h and cx:
            P[cy, 0:cx:s], P[cy, h:cx:s] = lift(P[cy, 0:cx:s].copy(), P[cy, h:cx:s].copy())
        h = s


def _huffman_lengths(freq):
    """Code length per symbol (0 = unused) from a Huffman tree over freq."""
    import heapq
    syms = [s for s in range(len(freq)) if freq[s]]
    lengths = np.zeros(len(freq), np.int64)
    if len(syms) == 1:
        lengths[syms[0]] = 1
        return lengths
    heap = [(int(freq[s]), k, [s]) for k, s in enumerate(syms)]
    heapq.heapify(heap)
    tick = len(heap)
    while len(heap) > 1:
        fa, _, a = heapq.heappop(heap)
        fb, _, b = heapq.heappop(heap)
        for s in a + b:
            lengths[s] += 1
        heapq.heappush(heap, (fa + fb, tick, a + b))
        tick += 1
    assert lengths.max() <= 58
    return lengths


def _canonical(lengths, lo, hi):
    """Code value per symbol under the format's canonical numbering: walking
    from length 58 down, each length's codes start where the (halved) run of
    the longer codes ended; symbols of one length count up in symbol order."""
    count = np.bincount(lengths[lo:hi + 1], minlength=59)
    start, c = [0] * 59, 0
    for L in range(58, 0, -1):
        start[L] = c
        c = (c + int(count[L])) >> 1
    codes = {}
    for s in range(lo, hi + 1):
        L = int(lengths[s])
        if L:
            codes[s] = (start[L], L)
            start[L] += 1
    return codes


class _BitWriter:
    def __init__(self):
        self.bits = []

    def put(self, value, n):
        self.bits.extend((value >> (n - 1 - k)) & 1 for k in range(n))

    def bytes(self):
        b = self.bits + [0] * (-len(self.bits) % 8)
        return np.packbits(np.array(b, np.uint8)).tobytes() if b else b""


def huffman_block(words, lengths_override=None):
    """OpenEXR Huffman block for the uint16 sequence `words`: header (lowest
    and highest symbol, table bytes, bit count, 4 spare bytes), packed code
    lengths (6 bits each; 59..62 = 2..5 zero lengths; 63 + 8 bits = 6..261
    zero lengths), then the codes, MSB first; runs of a repeated value use
    the escape symbol (highest symbol + 1) and an 8-bit repeat count.
    lengths_override: code lengths to write instead (malformed tables)."""
    words = np.asarray(words, np.uint16)
    freq = np.bincount(words, minlength=65537).astype(np.int64)
    lo, hi = int(words.min()), int(words.max()) + 1  # hi: the escape symbol
    freq[hi] = 1
    lengths = _huffman_lengths(freq) if lengths_override is None else np.asarray(lengths_override, np.int64)
    table = _BitWriter()
    s = lo
    while s <= hi:
        if lengths[s]:
            table.put(int(lengths[s]), 6)
            s += 1
            continue
        run = 1
        while s + run <= hi and lengths[s + run] == 0 and run < 261:
            run += 1
        if run >= 6:
            table.put(63, 6)
            table.put(run - 6, 8)
        elif run >= 2:
            table.put(59 + run - 2, 6)
        else:
            table.put(0, 6)
        s += run
    tb = table.bytes()
    codes = _canonical(lengths, lo, hi)
    out = _BitWriter()
    i, n = 0, len(words)
    esc = codes[hi]
    while i < n:
        v = int(words[i])
        j = i + 1
        while j < n and words[j] == v and j - i <= 255:
            j += 1
        rep = j - i - 1
        c = codes[v]
        out.put(*c)
        if rep >= 2:
            out.put(*esc)
            out.put(rep, 8)
        else:
            for _ in range(rep):
                out.put(*c)
        i = j
    nbits = len(out.bits)
    return struct.pack("<IIIII", lo, hi, len(tb), nbits, 0) + tb + out.bytes()


def piz_block(raw, chan_words, width, lines, lengths_override=None):
    """PIZ-compressed block of the interleaved scanline bytes `raw`
    (lines x channels x width samples, chan_words[c] uint16 words per sample
    of channel c): channel planes, the bitmap of present values and its dense
    index table, the wavelet per plane, then the Huffman block."""
    w16 = np.frombuffer(raw, np.uint16)
    planes, pos = [[] for _ in chan_words], 0
    for _ in range(lines):
        for c, k in enumerate(chan_words):
            planes[c].append(w16[pos:pos + width * k])
            pos += width * k
    planes = [np.concatenate(p).reshape(lines, width * k) for p, k in zip(planes, chan_words)]
    present = np.zeros(65536, bool)
    for p in planes:
        present[p.ravel()] = True
    present[0] = False
    bitmap = np.packbits(present.reshape(-1, 8)[:, ::-1], axis=1).ravel()  # bit v & 7 of byte v >> 3
    dense = np.cumsum(np.concatenate([[True], present[1:]])).astype(np.int64) - 1
    nz = np.nonzero(bitmap)[0]
    first, last = (int(nz[0]), int(nz[-1])) if len(nz) else (8191, 0)
    largest = int(np.count_nonzero(present))  # dense indices run 0 .. count of present non-zero values
    words = []
    for p, k in zip(planes, chan_words):
        q = dense[p].astype(np.uint16)
        for j in range(k):
            sub = q[:, j::k]
            _wavelet_encode(sub, largest >= (1 << 14))
            q[:, j::k] = sub
        words.append(q.ravel())
    words = np.concatenate(words)
    huf = huffman_block(words, lengths_override)
    head = struct.pack("<HH", first, last) + (bitmap[first:last + 1].tobytes() if first <= last else b"")
    return head + struct.pack("<I", len(huf)) + huf


def write_exr(path, planes, compression, half=True, piz_lengths=None, block_offset=None):
    """Scanline EXR; planes: dict name -> (H, W) float32.  compression 0
    (NONE), 1 (RLE), 2 (ZIPS), 3 (ZIP) or 4 (PIZ).  piz_lengths: code
    lengths written into every PIZ block instead of the real ones;
    block_offset: a value written over the first entry of the offset table
    (malformed files)."""
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
    lpb = {3: 16, 4: 32}.get(compression, 1)
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
        if compression == 4:
            comp = piz_block(raw, [1 if half else 2] * len(names), W, min(H, y0 + lpb) - y0, piz_lengths)
            if len(comp) < len(raw) or piz_lengths is not None:
                data = comp
        elif compression in (1, 2, 3):
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
    for k, blk in enumerate(blocks):
        table += struct.pack("<Q", block_offset if (k == 0 and block_offset is not None) else off)
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
