"""Deterministic procedural scenes for the benchmark configurations.

Every generator is a pure function of its arguments (numpy RNG streams from
one seed), so the GPU renderer and the CPU oracle receive identical arrays.

  cornell()          36-triangle Cornell box, one quad light (0.25 m^2, emission 15)
  icosphere_hdri()   subdivided, value-noise-displaced icosphere in 64 primitives
                     + ground plane + synthetic equirect HDRI with a sun disc
  instanced_interior() procedural meshes under many affine instances, emissive
                     meshes, glass, checker albedo + alpha-mask textures
  displaced_grid()   large displaced height-field grid (served-form scene)
"""
import ctypes as C
import math

import numpy as np

from . import _ffi as F


# ----------------------------------------------------------------- helpers
def look_at_camera(eye, target, up, fovy_deg, width, height):
    """Pinhole camera with the reference's Camera fields
    (src/scene/scene_load_gltf.cpp:465-493): view, right, up, pixelLength."""
    eye = np.asarray(eye, np.float64)
    view = np.asarray(target, np.float64) - eye
    view /= np.linalg.norm(view)
    right = np.cross(view, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    up2 = np.cross(right, view)
    yscaled = math.tan(0.5 * math.radians(fovy_deg))
    xscaled = yscaled * width / height
    return dict(width=int(width), height=int(height), position=eye.astype(np.float32),
                view=view.astype(np.float32), up=up2.astype(np.float32), right=right.astype(np.float32),
                pixel_length=np.array([2 * xscaled / width, 2 * yscaled / height], np.float32))


def material(color=(0.8, 0.8, 0.8), roughness=1.0, metallic=0.0, emission=(0.0, 0.0, 0.0), strength=0.0,
             ior=1.5, transmission=0.0, alpha=1.0, alpha_mode=F.LX_ALPHA_OPAQUE, alpha_cutoff=0.5,
             albedo_tex=-1, mr_tex=-1, normal_tex=-1, emissive_tex=-1, uv_scale=(1.0, 1.0), uv_rotation=0.0,
             uv_offset=(0.0, 0.0), material_type=F.LX_MATERIAL_MICROFACET):
    """material_type (LX_MATERIAL_*) is read only by LX_SHADING_LEGACY renders."""
    return dict(color=color, roughness=roughness, metallic=metallic, emission=emission, strength=strength, ior=ior,
                transmission=transmission, alpha=alpha, alpha_mode=alpha_mode, alpha_cutoff=alpha_cutoff,
                albedo_tex=albedo_tex, mr_tex=mr_tex, normal_tex=normal_tex, emissive_tex=emissive_tex,
                uv_scale=uv_scale, uv_rotation=uv_rotation, uv_offset=uv_offset, material_type=material_type)


def affine(translate=(0, 0, 0), rot_y=0.0, scale=(1, 1, 1), rot_x=0.0, rot_z=0.0):
    """Row-major 3x4 object-to-world matrix T * Rz * Ry * Rx * S (float32)."""
    cx, sx = math.cos(rot_x), math.sin(rot_x)
    cy, sy = math.cos(rot_y), math.sin(rot_y)
    cz, sz = math.cos(rot_z), math.sin(rot_z)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    M = Rz @ Ry @ Rx @ np.diag(np.asarray(scale, np.float64))
    out = np.zeros((3, 4))
    out[:, :3] = M
    out[:, 3] = translate
    return out.astype(np.float32).reshape(12)


IDENTITY = affine()


def _quad(p0, p1, p2, p3, normal):
    """Two triangles (0,1,2), (0,2,3) wound so the geometric normal
    cross(v1 - v0, v2 - v0) points along `normal`."""
    p = np.array([p0, p1, p2, p3], np.float32)
    n = np.cross(p[1] - p[0], p[2] - p[0])
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    if np.dot(n, normal) < 0:
        idx = idx[:, ::-1].copy()
    return p, idx


def _cube():
    """Unit cube [-0.5, 0.5]^3 with outward geometric normals (12 triangles)."""
    verts, tris = [], []
    faces = [((1, 0, 0), [(0.5, -0.5, -0.5), (0.5, 0.5, -0.5), (0.5, 0.5, 0.5), (0.5, -0.5, 0.5)]),
             ((-1, 0, 0), [(-0.5, -0.5, -0.5), (-0.5, -0.5, 0.5), (-0.5, 0.5, 0.5), (-0.5, 0.5, -0.5)]),
             ((0, 1, 0), [(-0.5, 0.5, -0.5), (-0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (0.5, 0.5, -0.5)]),
             ((0, -1, 0), [(-0.5, -0.5, -0.5), (0.5, -0.5, -0.5), (0.5, -0.5, 0.5), (-0.5, -0.5, 0.5)]),
             ((0, 0, 1), [(-0.5, -0.5, 0.5), (0.5, -0.5, 0.5), (0.5, 0.5, 0.5), (-0.5, 0.5, 0.5)]),
             ((0, 0, -1), [(-0.5, -0.5, -0.5), (-0.5, 0.5, -0.5), (0.5, 0.5, -0.5), (0.5, -0.5, -0.5)])]
    for n, q in faces:
        p, idx = _quad(*q, normal=np.array(n, np.float32))
        tris.append(idx + len(verts) * 4)
        verts.append(p)
    return np.concatenate(verts), np.concatenate(tris).astype(np.uint32)


class SceneData:
    """Host-side scene arrays; to_desc() builds the lx_scene_desc view."""

    def __init__(self):
        self.meshes = []      # dicts: pos (n,3) f32, idx (m,3) u32, nrm (n,3) or None, uv (n,2) or None
        self.instances = []   # (mesh, material, transform[12])
        self.materials = []
        self.textures = []    # (H, W, 4) uint8 / uint16 / float32 arrays
        self.env = None       # (H, W, 4) float32
        self.camera = None
        self.trace_depth = 8
        self.name = "scene"

    def add_mesh(self, pos, idx, nrm=None, uv=None):
        self.meshes.append(dict(pos=np.ascontiguousarray(pos, np.float32),
                                idx=np.ascontiguousarray(idx, np.uint32).reshape(-1, 3),
                                nrm=None if nrm is None else np.ascontiguousarray(nrm, np.float32),
                                uv=None if uv is None else np.ascontiguousarray(uv, np.float32)))
        return len(self.meshes) - 1

    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_instance(self, mesh, mat, transform=IDENTITY):
        self.instances.append((mesh, mat, np.asarray(transform, np.float32).reshape(12)))

    @property
    def num_tris(self):
        return sum(len(m["idx"]) for m in self.meshes)

    @property
    def num_instanced_tris(self):
        return sum(len(self.meshes[i[0]]["idx"]) for i in self.instances)

    def to_desc(self, camera=None):
        keep = []
        md = (F.lx_mesh_desc * len(self.meshes))()
        for k, m in enumerate(self.meshes):
            md[k].num_verts = len(m["pos"])
            md[k].positions = m["pos"].ctypes.data
            md[k].normals = m["nrm"].ctypes.data if m["nrm"] is not None else None
            md[k].uvs = m["uv"].ctypes.data if m["uv"] is not None else None
            md[k].num_tris = len(m["idx"])
            md[k].indices = m["idx"].ctypes.data
        inst = (F.lx_instance_desc * len(self.instances))()
        for k, (mesh, mat, xf) in enumerate(self.instances):
            inst[k].mesh = mesh
            inst[k].material = mat
            inst[k].transform = F.F12(*[float(v) for v in xf])
        mats = (F.lx_material_desc * len(self.materials))()
        for k, m in enumerate(self.materials):
            mats[k] = material_desc(m)
        texs = (F.lx_texture_desc * max(1, len(self.textures)))()
        for k, t in enumerate(self.textures):
            texs[k] = texture_desc(t)
            keep.append(t)
        env = None
        if self.env is not None:
            env = F.lx_texture_desc()
            env.width, env.height, env.format = self.env.shape[1], self.env.shape[0], F.LX_TEX_RGBA32F
            env.data = self.env.ctypes.data
        d = F.lx_scene_desc()
        d.meshes, d.num_meshes = md, len(self.meshes)
        d.instances, d.num_instances = inst, len(self.instances)
        d.materials, d.num_materials = mats, len(self.materials)
        d.textures, d.num_textures = texs, len(self.textures)
        d.envmap = C.pointer(env) if env is not None else None
        d.camera = camera_desc(camera or self.camera)
        d.trace_depth = self.trace_depth
        keep += [md, inst, mats, texs, env, self]
        return d, keep


def camera_desc(c):
    cd = F.lx_camera_desc()
    cd.width, cd.height = c["width"], c["height"]
    cd.position = F.F3(*[float(v) for v in c["position"]])
    cd.view = F.F3(*[float(v) for v in c["view"]])
    cd.up = F.F3(*[float(v) for v in c["up"]])
    cd.right = F.F3(*[float(v) for v in c["right"]])
    cd.pixel_length = F.F2(*[float(v) for v in c["pixel_length"]])
    return cd


def _xf(scale, rotation, offset):
    x = F.lx_tex_transform()
    x.offset = F.F2(*offset)
    x.scale = F.F2(*scale)
    x.rotation = rotation
    return x


def material_desc(m):
    if "_desc" in m:  # a loader material: every field and per-slot transforms given
        r = m["_desc"]
        d = F.lx_material_desc()
        for k, _ in F.lx_material_desc._fields_:
            v = r[k]
            if k.endswith("_xf"):
                v = _xf(v["scale"], v["rotation"], v["offset"])
            elif isinstance(v, tuple):
                v = F.F3(*v)
            setattr(d, k, v)
        return d
    d = F.lx_material_desc()
    d.base_color = F.F3(*m["color"])
    xf = _xf(m["uv_scale"], m["uv_rotation"], m["uv_offset"])
    d.albedo_tex, d.albedo_xf = m["albedo_tex"], xf
    d.roughness, d.metallic = m["roughness"], m["metallic"]
    d.metallic_roughness_tex, d.metallic_roughness_xf = m["mr_tex"], xf
    d.normal_tex, d.normal_xf = m["normal_tex"], xf
    d.emission_color = F.F3(*m["emission"])
    d.emission_strength = m["strength"]
    d.emissive_tex, d.emissive_xf = m["emissive_tex"], xf
    d.ior, d.transmission, d.alpha = m["ior"], m["transmission"], m["alpha"]
    d.alpha_mode, d.alpha_cutoff = m["alpha_mode"], m["alpha_cutoff"]
    d.material_type = m.get("material_type", F.LX_MATERIAL_MICROFACET)
    return d


def texture_desc(t):
    d = F.lx_texture_desc()
    d.height, d.width = t.shape[0], t.shape[1]
    d.format = {np.dtype(np.uint8): F.LX_TEX_RGBA8, np.dtype(np.uint16): F.LX_TEX_RGBA16,
                np.dtype(np.float32): F.LX_TEX_RGBA32F}[t.dtype]
    d.data = t.ctypes.data
    return d


def from_glb(path, envmap=None):
    """A SceneData holding exactly what lx_scene_load_glb builds from `path`
    (+ EXR), read back through the host-only parse hook; lets the oracle and
    lx_scene_create see the loader's scene."""
    from . import parse_glb
    g = parse_glb(path, envmap)
    env_len = 0 if g["env"] is None else g["env"][0] * g["env"][1] * 4
    if env_len:
        g = parse_glb(path, envmap, env_head=env_len)
    s = SceneData()
    s.name = path.rsplit("/", 1)[-1]
    for pos, nrm, uv, idx in g["meshes"]:
        s.add_mesh(pos, idx, nrm, uv)
    for m in g["materials"]:
        s.add_material({"_desc": m})
    for mesh, mat, xf in g["instances"]:
        s.add_instance(mesh, mat, np.asarray(xf, np.float32).reshape(12))
    s.textures = [t for t, _ in g["textures"]]
    if env_len:
        w, h, rgba = g["env"]
        s.env = np.ascontiguousarray(rgba.reshape(h, w, 4), np.float32)
    c = g["camera"]
    s.camera = dict(width=c["width"], height=c["height"], position=np.float32(c["position"]),
                    view=np.float32(c["view"]), up=np.float32(c["up"]), right=np.float32(c["right"]),
                    pixel_length=np.float32(c["pixel_length"]))
    s.trace_depth = g["trace_depth"]
    return s


# ----------------------------------------------------------------- cfg0
def cornell(width=256, height=256, trace_depth=8):
    """Procedural Cornell box: 5 walls (2 triangles each) + 2 boxes (12 each)
    + one 0.5 x 0.5 m quad light of emission 15 = 36 triangles.  Lambert-like
    uber materials (roughness 1, metallic 0): white 0.73, red (0.65, 0.05,
    0.05), green (0.12, 0.45, 0.15)."""
    s = SceneData()
    s.name = "cornell36"
    white = s.add_material(material((0.73, 0.73, 0.73)))
    red = s.add_material(material((0.65, 0.05, 0.05)))
    green = s.add_material(material((0.12, 0.45, 0.15)))
    light = s.add_material(material((0.73, 0.73, 0.73), emission=(1.0, 1.0, 1.0), strength=15.0))
    walls = [
        ((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1), (0, 1, 0), white),      # floor
        ((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0), white),     # ceiling
        ((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), (0, 0, 1), white),    # back
        ((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1), (1, 0, 0), red),      # left
        ((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0), green),       # right
    ]
    for p0, p1, p2, p3, n, m in walls:
        p, idx = _quad(p0, p1, p2, p3, np.array(n, np.float32))
        s.add_instance(s.add_mesh(p, idx), m)
    cp, ci = _cube()
    cube = s.add_mesh(cp, ci)
    s.add_instance(cube, white, affine((-0.33, 0.6, -0.3), math.radians(17.0), (0.6, 1.2, 0.6)))
    s.add_instance(cube, white, affine((0.35, 0.3, 0.3), math.radians(-18.0), (0.6, 0.6, 0.6)))
    lp, li = _quad((-0.25, 0, -0.25), (0.25, 0, -0.25), (0.25, 0, 0.25), (-0.25, 0, 0.25),
                   np.array((0, -1, 0), np.float32))
    # 1 mm under the ceiling: the light's back face is practically unreachable,
    # so the one-sided (-dmis) and two-sided (plain) emission rules agree
    s.add_instance(s.add_mesh(lp, li), light, affine((0.0, 1.999, 0.0)))
    s.camera = look_at_camera((0.0, 1.0, 3.4), (0.0, 1.0, 0.0), (0, 1, 0), 40.0, width, height)
    s.trace_depth = trace_depth
    return s


# ----------------------------------------------------------------- cfg1
def icosphere(level):
    """Unit icosphere by repeated 4-way subdivision; triangle i of level k
    becomes triangles 4i .. 4i+3 of level k+1 (spatially coherent order)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        nv = len(v)
        e = np.stack([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=1).reshape(-1, 2)
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // nv, uniq % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = nv + inv.reshape(-1, 3)
        v = np.concatenate([v, mid])
        v0, v1, v2 = f[:, 0], f[:, 1], f[:, 2]
        mab, mbc, mca = m[:, 0], m[:, 1], m[:, 2]
        f = np.stack([np.stack([v0, mab, mca], 1), np.stack([v1, mbc, mab], 1), np.stack([v2, mca, mbc], 1),
                      np.stack([mab, mbc, mca], 1)], axis=1).reshape(-1, 3)
    return v, f


def value_noise(p, seed, octaves=3, base_freq=3.0, lacunarity=2.0, gain=0.5):
    """Seeded 3-D lattice value noise in [-1, 1] (smoothstep trilinear),
    summed over octaves with the given lacunarity and gain."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256).astype(np.int64)
    vals = rng.uniform(-1.0, 1.0, 256)
    total = np.zeros(len(p))
    amp, freq = 1.0, base_freq
    for _ in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3.0 - 2.0 * fr)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    h = perm[(i0[:, 0] + dx) & 255]
                    h = perm[(h + i0[:, 1] + dy) & 255]
                    h = perm[(h + i0[:, 2] + dz) & 255]
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += vals[h] * wx * wy * wz
        total += amp * acc
        amp *= gain
        freq *= lacunarity
    return total


def vertex_normals(v, f):
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = np.zeros_like(v)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n


def synthetic_hdri(width=2048, height=1024, sun_theta=math.radians(50.0), sun_phi=0.6, sun_radiance=1.0e4,
                   sun_diameter_deg=0.5):
    """Equirect RGBA32F map in the reference's parameterisation (row 0 at
    theta = 0 = +y; phi = 2 pi u - pi): sky radiance 1.0 at the zenith down to
    0.1 at the nadir, plus a sun disc of the given angular diameter."""
    th = math.pi * (np.arange(height) + 0.5) / height
    ph = 2.0 * math.pi * (np.arange(width) + 0.5) / width - math.pi
    sky = (1.0 - 0.9 * th / math.pi).astype(np.float32)
    img = np.empty((height, width, 4), np.float32)
    img[..., 0:3] = sky[:, None, None]
    img[..., 3] = 1.0
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    d = np.stack([st * np.cos(ph)[None, :], np.broadcast_to(ct, (height, width)), st * np.sin(ph)[None, :]], -1)
    s = np.array([math.sin(sun_theta) * math.cos(sun_phi), math.cos(sun_theta),
                  math.sin(sun_theta) * math.sin(sun_phi)])
    ang = np.degrees(np.arccos(np.clip(d @ s, -1.0, 1.0)))
    img[ang <= 0.5 * sun_diameter_deg, 0:3] = sun_radiance
    return img


def icosphere_hdri(width=1920, height=1080, level=8, seed=0, pieces=64, env_size=(2048, 1024), trace_depth=12,
                   amp=0.05, normals=True):
    """Displaced icosphere (level 8: 1,310,720 triangles) split into `pieces`
    primitives with roughness ~ U[0.05, 0.8] and metallic ~ Bernoulli(0.5),
    on a 2-triangle ground plane, lit by a synthetic HDRI with a 0.5-degree
    sun disc of radiance 1e4."""
    s = SceneData()
    s.name = f"icosphere_l{level}_hdri"
    v, f = icosphere(level)
    disp = value_noise(v, seed + 1) * amp
    v = v * (1.0 + disp)[:, None]
    n = vertex_normals(v, f) if normals else None
    rng = np.random.default_rng(seed)
    nf = len(f)
    pieces = max(1, min(pieces, nf))
    bounds = np.linspace(0, nf, pieces + 1).astype(np.int64)
    for k in range(pieces):
        ff = f[bounds[k]:bounds[k + 1]]
        used, inv = np.unique(ff.reshape(-1), return_inverse=True)
        mesh = s.add_mesh(v[used], inv.reshape(-1, 3), None if n is None else n[used])
        col = tuple(float(c) for c in rng.uniform(0.2, 0.9, 3))
        mat = s.add_material(material(col, roughness=float(rng.uniform(0.05, 0.8)),
                                      metallic=float(rng.integers(0, 2))))
        s.add_instance(mesh, mat)
    gp, gi = _quad((-20, -1.1, -20), (-20, -1.1, 20), (20, -1.1, 20), (20, -1.1, -20), np.array((0, 1, 0)))
    ground = s.add_material(material((0.5, 0.5, 0.5), roughness=0.9))
    s.add_instance(s.add_mesh(gp, gi), ground)
    s.env = synthetic_hdri(env_size[0], env_size[1])
    s.camera = look_at_camera((0.0, 0.35, 3.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, width, height)
    s.trace_depth = trace_depth
    return s


# ----------------------------------------------------------------- cfg2
def checker_textures(size=1024, squares=16):
    """RGBA8 checker base colour and an alpha-mask texture (alpha 0 in a
    regular grid of holes)."""
    y, x = np.mgrid[0:size, 0:size]
    c = ((x * squares // size + y * squares // size) % 2).astype(np.uint8)
    base = np.empty((size, size, 4), np.uint8)
    base[..., 0] = np.where(c, 220, 40)
    base[..., 1] = np.where(c, 200, 60)
    base[..., 2] = np.where(c, 160, 90)
    base[..., 3] = 255
    mask = base.copy()
    holes = ((x * squares * 2 // size) % 2 == 0) & ((y * squares * 2 // size) % 2 == 0)
    mask[..., 3] = np.where(holes, 0, 255).astype(np.uint8)
    return base, mask


def sphere_uv(v):
    """Spherical uv of unit-sphere points: u from the azimuth, v from the polar angle."""
    r = np.linalg.norm(v, axis=1)
    return np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * math.pi),
                     np.arccos(np.clip(v[:, 1] / r, -1, 1)) / math.pi], 1)


def bump_normal_map(size=256, freq=8):
    """RGBA8 tangent-space normal map of a sinusoidal bump field."""
    y, x = np.mgrid[0:size, 0:size] / size
    dx = 0.35 * np.cos(2 * math.pi * freq * x)
    dy = 0.35 * np.cos(2 * math.pi * freq * y)
    n = np.stack([-dx, -dy, np.ones_like(dx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    out = np.empty((size, size, 4), np.uint8)
    out[..., 0:3] = np.round((n * 0.5 + 0.5) * 255).astype(np.uint8)
    out[..., 3] = 255
    return out


def material_zoo(width=320, height=180, level=3, env_size=(256, 128), trace_depth=8, flat=False):
    """One sphere per legacy material class (diffuse, mirror, glass, microfacet
    with metallic-roughness and normal maps, emissive) on a checker floor, an
    alpha-MASK and an alpha-BLEND panel, a quad light and a synthetic HDRI.
    Every material also renders under the uber shader.  flat: spheres
    without vertex normals (faceted)."""
    s = SceneData()
    s.name = "material_zoo"
    base, mask = checker_textures(128, 8)
    nm = bump_normal_map(128, 6)
    mr = np.empty((64, 64, 4), np.uint8)
    yy, xx = np.mgrid[0:64, 0:64]
    mr[..., 0] = 0
    mr[..., 1] = (40 + 3 * xx).astype(np.uint8)   # roughness scale
    mr[..., 2] = (255 - 3 * yy).astype(np.uint8)  # metallic scale
    mr[..., 3] = 255
    s.textures += [base, mask, nm, mr]
    T_BASE, T_MASK, T_NRM, T_MR = 0, 1, 2, 3
    m_floor = s.add_material(material((1, 1, 1), roughness=0.8, albedo_tex=T_BASE, uv_scale=(4.0, 4.0),
                                      material_type=F.LX_MATERIAL_DIFFUSE))
    m_diff = s.add_material(material((0.7, 0.3, 0.2), roughness=1.0, material_type=F.LX_MATERIAL_DIFFUSE))
    m_mirror = s.add_material(material((0.9, 0.9, 0.95), roughness=0.0, metallic=1.0,
                                       material_type=F.LX_MATERIAL_SPECULAR))
    m_glass = s.add_material(material((0.8, 1.0, 0.9), roughness=0.0, transmission=0.6, ior=1.45,
                                      material_type=F.LX_MATERIAL_GLASS))
    m_micro = s.add_material(material((0.9, 0.7, 0.3), roughness=0.6, metallic=0.8, mr_tex=T_MR, normal_tex=T_NRM,
                                      uv_scale=(2.0, 1.0), uv_rotation=0.3, material_type=F.LX_MATERIAL_MICROFACET))
    m_emit = s.add_material(material((0.2, 0.2, 0.2), emission=(1.0, 0.85, 0.6), strength=4.0,
                                     material_type=F.LX_MATERIAL_EMISSIVE))
    m_mask = s.add_material(material((1, 1, 1), roughness=0.5, albedo_tex=T_MASK, alpha_mode=F.LX_ALPHA_MASK,
                                     alpha_cutoff=0.5))
    m_blend = s.add_material(material((0.3, 0.5, 0.9), roughness=0.4, alpha=0.45, alpha_mode=F.LX_ALPHA_BLEND))
    m_light = s.add_material(material((1, 1, 1), emission=(1.0, 1.0, 1.0), strength=12.0,
                                      material_type=F.LX_MATERIAL_EMISSIVE))
    p, idx = _quad((-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4), np.array((0, 1, 0), np.float32))
    uv = (p[:, [0, 2]] + 4.0) / 8.0
    s.add_instance(s.add_mesh(p, idx, None, uv), m_floor)
    v, f = icosphere(level)
    sph = s.add_mesh(v, f, None if flat else v, sphere_uv(v))
    for k, m in enumerate((m_diff, m_mirror, m_glass, m_micro, m_emit)):
        s.add_instance(sph, m, affine((-2.4 + 1.2 * k, 0.5, 0.0), 0.4 * k, (0.5, 0.5, 0.5)))
    pp, pi = _quad((-0.6, 0, 0), (0.6, 0, 0), (0.6, 1.2, 0), (-0.6, 1.2, 0), np.array((0, 0, 1), np.float32))
    puv = np.stack([(pp[:, 0] + 0.6) / 1.2, pp[:, 1] / 1.2], 1)
    panel = s.add_mesh(pp, pi, None, puv)
    s.add_instance(panel, m_mask, affine((-1.0, 0.0, 1.3), 0.2))
    s.add_instance(panel, m_blend, affine((1.0, 0.0, 1.3), -0.2))
    lp, li = _quad((-0.5, 0, -0.5), (0.5, 0, -0.5), (0.5, 0, 0.5), (-0.5, 0, 0.5), np.array((0, -1, 0), np.float32))
    s.add_instance(s.add_mesh(lp, li), m_light, affine((0.0, 3.0, 0.5)))
    s.env = synthetic_hdri(env_size[0], env_size[1], sun_radiance=50.0, sun_diameter_deg=4.0)
    s.camera = look_at_camera((0.0, 1.6, 5.5), (0.0, 0.5, 0.0), (0, 1, 0), 45.0, width, height)
    s.trace_depth = trace_depth
    return s


def instanced_interior(width=1920, height=1080, n_meshes=64, n_instances=10000, tri_range=(10000, 500000),
                       n_emissive=8, seed=0, tex_size=1024, trace_depth=16):
    """Procedural instanced interior: displaced icosphere meshes with triangle
    counts drawn from tri_range under many random affine instances, emissive
    meshes for area-CDF NEE, glass (transmission 1, ior 1.5), checker albedo
    and alpha-mask (MASK) textures with uvs."""
    rng = np.random.default_rng(seed)
    s = SceneData()
    s.name = "instanced_interior"
    base, mask = checker_textures(tex_size)
    s.textures += [base, mask]
    levels = []
    for k in range(n_meshes):
        target = rng.uniform(*tri_range)
        lvl = int(round(math.log(target / 20.0, 4)))
        levels.append(max(1, min(lvl, 8)))
    cache = {}
    mesh_ids = []
    for k, lvl in enumerate(levels):
        if lvl not in cache:
            cache[lvl] = icosphere(lvl)
        v, f = cache[lvl]
        vv = v * (1.0 + 0.15 * value_noise(v, seed + 100 + k))[:, None]
        uv = np.stack([0.5 + np.arctan2(vv[:, 2], vv[:, 0]) / (2 * math.pi),
                       np.arccos(np.clip(vv[:, 1] / np.linalg.norm(vv, axis=1), -1, 1)) / math.pi], 1)
        mesh_ids.append(s.add_mesh(vv, f, vertex_normals(vv, f), uv))
    mats = []
    mats.append(s.add_material(material((1.0, 1.0, 1.0), roughness=0.02, transmission=1.0, ior=1.5)))
    mats.append(s.add_material(material((1, 1, 1), roughness=0.6, albedo_tex=0)))
    mats.append(s.add_material(material((1, 1, 1), roughness=0.5, albedo_tex=1, alpha_mode=F.LX_ALPHA_MASK,
                                        alpha_cutoff=0.5)))
    for k in range(5):
        mats.append(s.add_material(material(tuple(float(c) for c in rng.uniform(0.2, 0.9, 3)),
                                            roughness=float(rng.uniform(0.05, 0.9)),
                                            metallic=float(rng.integers(0, 2)))))
    emis = [s.add_material(material((1, 1, 1), emission=tuple(float(c) for c in rng.uniform(0.5, 1.0, 3)),
                                    strength=float(rng.uniform(5.0, 20.0)))) for _ in range(n_emissive)]
    extent = 30.0
    for k in range(n_instances):
        mesh = mesh_ids[int(rng.integers(0, n_meshes))]
        t = rng.uniform(-extent, extent, 3)
        t[1] = rng.uniform(-3.0, 12.0)
        sc = rng.uniform(0.2, 1.2, 3)
        xf = affine(t, rng.uniform(0, 2 * math.pi), sc, rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
        s.add_instance(mesh, mats[int(rng.integers(0, len(mats)))], xf)
    for k in range(n_emissive):
        t = rng.uniform(-extent, extent, 3)
        t[1] = rng.uniform(10.0, 14.0)
        s.add_instance(mesh_ids[k % n_meshes], emis[k], affine(t, 0.0, (1.5, 0.3, 1.5)))
    s.camera = look_at_camera((0.0, 5.0, 45.0), (0.0, 3.0, 0.0), (0, 1, 0), 50.0, width, height)
    s.trace_depth = trace_depth
    return s


def displaced_grid(quads=3873, seed=7, width=1920, height=1080, trace_depth=12):
    """Height-field grid of quads x quads cells (2 triangles each; 3873^2 x 2
    = 30,000,258 triangles) displaced by value noise."""
    s = SceneData()
    s.name = f"grid_{quads}"
    n = quads + 1
    xs = np.linspace(-10.0, 10.0, n, dtype=np.float64)
    X, Z = np.meshgrid(xs, xs)
    P = np.stack([X.reshape(-1), np.zeros(n * n), Z.reshape(-1)], 1)
    P[:, 1] = 0.8 * value_noise(P * 0.15, seed, octaves=3)
    i = np.arange(quads)
    a = (i[:, None] * n + i[None, :]).reshape(-1).astype(np.uint32)
    tri = np.empty((2 * len(a), 3), np.uint32)
    tri[0::2] = np.stack([a, a + n, a + n + 1], 1)
    tri[1::2] = np.stack([a, a + n + 1, a + 1], 1)
    m = s.add_material(material((0.6, 0.55, 0.5), roughness=0.7))
    s.add_instance(s.add_mesh(P.astype(np.float32), tri), m)
    s.env = synthetic_hdri(512, 256)
    s.camera = look_at_camera((0.0, 6.0, 14.0), (0.0, 0.0, 0.0), (0, 1, 0), 45.0, width, height)
    s.trace_depth = trace_depth
    return s
