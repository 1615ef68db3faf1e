"""The tail kernel's direct-light query against the closest-hit ray.

For -dmis jobs the tail kernel answers sampleDirectLight's shadow ray
(src/kernels/sample_direct_light.cu:9-66, a closest-hit ray in the
reference) with occlusion queries: any non-emitting candidate, the closest
emitting one before it over the emitter TLAS, and, when one comes first, a
second occlusion query up to it.  lx_trace_emitting exposes that query; it
must return lx_trace_closest's hit exactly when that hit's instance emits and
a miss otherwise.  Scenes with many emitting and non-emitting candidates
along each ray make every branch of the query occur."""
import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import _ffi as F
from luxion_amd import scenes

pytestmark = pytest.mark.gpu


def _emits(m):
    e = np.asarray(m["emission"], np.float32) * np.float32(m["strength"])
    return bool(np.any(e != 0) or np.any(np.isnan(e)))


def _confetti(seed=5, n_meshes=24, tris=150, n_emissive=9):
    """Random small triangles in the unit cube: meshes under the identity (one
    traversal group mixing emitting and non-emitting members) and under
    rotated / scaled transforms (groups of their own), a third of them
    emitting, one alpha-masked."""
    rng = np.random.default_rng(seed)
    s = scenes.SceneData()
    s.textures.append(scenes.checker_textures(64, 8)[1])
    mats = []
    for k in range(n_meshes):
        if k < n_emissive:
            mats.append(s.add_material(scenes.material((1, 1, 1), emission=tuple(rng.uniform(0.3, 1.0, 3)),
                                                       strength=float(rng.uniform(1, 5)))))
        elif k == n_emissive:
            mats.append(s.add_material(scenes.material((0.7, 0.7, 0.7), roughness=0.5, albedo_tex=0,
                                                       alpha_mode=F.LX_ALPHA_MASK, alpha_cutoff=0.5)))
        else:
            mats.append(s.add_material(scenes.material(tuple(rng.uniform(0.2, 0.9, 3)),
                                                       roughness=float(rng.uniform(0.1, 0.9)))))
    for k in range(n_meshes):
        c = rng.uniform(-0.9, 0.9, (tris, 1, 3))
        v = (c + rng.normal(0.0, 0.08, (tris, 3, 3))).astype(np.float32).reshape(-1, 3)
        idx = np.arange(3 * tris, dtype=np.uint32).reshape(-1, 3)
        uv = rng.uniform(0, 1, (3 * tris, 2)).astype(np.float32)
        m = s.add_mesh(v, idx, uv=uv)
        if k % 3 == 2:
            xf = scenes.affine(translate=tuple(rng.uniform(-0.2, 0.2, 3)), rot_y=float(rng.uniform(0, 6.28)),
                               scale=tuple(rng.uniform(0.7, 1.2, 3)), rot_x=float(rng.uniform(0, 1.0)))
            s.add_instance(m, mats[k], xf)
        else:
            s.add_instance(m, mats[k])
    s.camera = scenes.look_at_camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 40.0, 64, 64)
    return s


def _rays(n, seed, lo, hi, reach=1.0):
    """Origins uniform in the box [lo, hi], directions uniform; half the rays
    unbounded, half with tmax uniform in (0.05, 3) x reach."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmax = np.where(rng.uniform(size=n) < 0.5, 1e16, rng.uniform(0.05, 3.0, n) * reach)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 4:7], r[:, 7] = o, d, tmax
    return r


def _check(sc, data, rays, iteration):
    closest = lx.trace_closest(sc, rays, alpha=True, iteration=iteration)
    got = lx.trace_emitting(sc, rays, iteration=iteration)
    em = np.array([_emits(data.materials[data.instances[i][1]]) for i in range(len(data.instances))])
    hit = closest["instance"] != 0xFFFFFFFF
    keep = np.zeros(len(rays), bool)
    keep[hit] = em[closest["instance"][hit]]
    want = closest.copy()
    want["t"][~keep] = -1.0
    want["instance"][~keep] = 0xFFFFFFFF
    want["prim"][~keep] = 0xFFFFFFFF
    want["u"][~keep] = 0.0
    want["v"][~keep] = 0.0
    for f in ("t", "instance", "prim", "u", "v"):
        a, b = got[f], want[f]
        if a.dtype == np.float32:
            a, b = a.view(np.uint32), b.view(np.uint32)
        assert np.array_equal(a, b), (f, int(np.sum(a != b)))
    return int(hit.sum()), int(keep.sum())


@pytest.mark.parametrize("width", [4, 8])
def test_direct_light_query_equals_closest_hit_where_it_emits(ctx, width):
    data = _confetti()
    sc = lx.Scene(ctx, data, bvh_width=width)
    rays = _rays(60000, 3, -1.0, 1.0)
    hits, emitting = _check(sc, data, rays, iteration=7)
    # both outcomes occur in quantity: emitting closest hits and occluded ones
    assert emitting > 5000 and hits - emitting > 5000


def test_direct_light_query_on_the_instanced_interior(ctx):
    data = scenes.instanced_interior(96, 54, n_meshes=8, n_instances=200, tri_range=(80, 3000), n_emissive=3,
                                     tex_size=64, trace_depth=6)
    sc = lx.Scene(ctx, data)
    # origins inside the instances' world box (mesh boxes under their transforms)
    corners = []
    for mesh, _, xf in data.instances:
        p = data.meshes[mesh]["pos"]
        m = np.asarray(xf, np.float64).reshape(3, 4)
        corners.append(p.astype(np.float64) @ m[:, :3].T + m[:, 3])
    w = np.concatenate(corners)
    lo, hi = np.percentile(w, 5, axis=0), np.percentile(w, 95, axis=0)
    rays = _rays(60000, 11, lo, hi, reach=float(np.linalg.norm(hi - lo)) / 3.0)
    hits, emitting = _check(sc, data, rays, iteration=3)
    # a sparse scatter of instances: a few thousand hits, dozens of them on emitters
    assert hits > 2000 and emitting > 20


def test_direct_light_query_on_the_cornell_box(ctx):
    data = scenes.cornell(64, 64)
    sc = lx.Scene(ctx, data)
    rays = _rays(40000, 2, np.array([-0.95, -0.45, -0.95]), np.array([0.95, 1.95, 0.95]))
    hits, emitting = _check(sc, data, rays, iteration=1)
    assert emitting > 100 and hits - emitting > 10000
