"""Traversal groups: instances with one bitwise-identical transform whose
meshes have no other user share one BLAS whose triangles name their instance.
The renders stay the oracle's bits (the oracle traces every instance on its
own), and hits report the member instance and the triangle of its mesh."""
import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import _ffi as F
from luxion_amd import scenes

pytestmark = pytest.mark.gpu


def grouped_scene(width=128, height=96):
    """The material zoo (identity floor, a sphere mesh shared by five
    instances, a panel mesh shared by a MASK and a BLEND instance, a light
    with its own transform) plus three spheres of distinct meshes under one
    rotated, scaled transform (one of them MASK-textured, one glass) and two
    more identity-transform meshes that join the floor's group."""
    s = scenes.material_zoo(width, height, level=3, env_size=(128, 64))
    m_diff, m_mirror, m_glass, m_mask = 1, 2, 3, 6
    xf = scenes.affine((0.9, 1.0, -0.9), 0.7, (0.3, 0.3, 0.3))
    v, f = scenes.icosphere(2)
    for k, m in enumerate((m_mask, m_diff, m_glass)):
        off = np.array((2.2 * (k - 1), 0.0, 0.0), np.float32)
        s.add_instance(s.add_mesh(v + off, f, v, scenes.sphere_uv(v)), m, xf)
    for k, m in enumerate((m_mirror, m_mask)):
        p, idx = scenes._quad((-2.0 + 2.5 * k, 0.05, -2.2), (-1.2 + 2.5 * k, 0.05, -2.2), (-1.2 + 2.5 * k, 1.0, -2.4),
                              (-2.0 + 2.5 * k, 1.0, -2.4), np.array((0, 0, 1), np.float32))
        uv = np.stack([(p[:, 0] - p[:, 0].min()) / 0.8, p[:, 1]], 1)
        s.add_instance(s.add_mesh(p, idx, None, uv), m)
    return s


def test_grouped_scene_renders_oracle_bits(ctx, oracle):
    data = grouped_scene()
    sc = lx.Scene(ctx, data)
    ref = oracle.OracleScene(data)
    for kw in (dict(dmis=True), dict(emis=True), dict()):
        r = lx.Render(sc, **kw)
        r.iterate(1, 2)
        img, _ = r.accum()
        want, _ = ref.render(1, 2, threads=8, **kw)
        assert np.array_equal(img.reshape(-1, 3), want), kw


@pytest.mark.parametrize("width", [4, 8])
def test_grouped_scene_at_both_widths_and_builders(ctx, oracle, width):
    data = grouped_scene(96, 64)
    want, _ = oracle.OracleScene(data).render(1, 2, threads=8, emis=True)
    for builder in (0, 1):
        sc = lx.Scene(ctx, data, bvh_width=width, builder=builder)
        r = lx.Render(sc, emis=True)
        r.iterate(1, 2)
        assert np.array_equal(r.accum()[0].reshape(-1, 3), want), (width, builder)


def test_hits_name_the_member_instance_and_mesh_triangle(ctx):
    """Rays aimed at the centroid of chosen triangles of grouped instances
    hit that instance and triangle index of its own mesh."""
    data = grouped_scene()
    sc = lx.Scene(ctx, data)
    rng = np.random.default_rng(3)
    origins, dirs, want = [], [], []
    for inst, (mesh, mat, xf) in enumerate(data.instances):
        if inst < 9:
            continue  # the spheres and panels added to the zoo
        M = np.asarray(xf, np.float64).reshape(3, 4)
        m = data.meshes[mesh]
        for t in rng.choice(len(m["idx"]), min(3, len(m["idx"])), replace=False):
            c = m["pos"][m["idx"][t]].astype(np.float64).mean(0)
            wc = M[:, :3] @ c + M[:, 3]
            e = m["pos"][m["idx"][t]].astype(np.float64)
            n = np.cross(e[1] - e[0], e[2] - e[0])
            nw = np.linalg.inv(M[:, :3]).T @ n
            nw /= np.linalg.norm(nw)
            o = wc + 1e-3 * nw
            origins.append(o)
            dirs.append(-nw)
            want.append((inst, t))
    rays = np.zeros((len(origins), 8), np.float32)
    rays[:, 0:3] = origins
    rays[:, 4:7] = dirs
    rays[:, 3] = 0.0
    rays[:, 7] = 1.0
    hits = lx.trace_closest(sc, rays, alpha=False)
    got = [(int(h["instance"]), int(h["prim"])) for h in hits]
    assert got == want
    assert np.allclose(hits["t"], 1e-3, rtol=1e-2)


def test_trace_degenerate_rays_and_ties_vs_oracle(ctx, oracle):
    """Cornell box (five identity-transform walls form one tagged group, the
    two boxes share a mesh): axis-aligned rays (infinite reciprocal
    directions, signed zeros), rays lying in the floor plane, and rays aimed
    exactly at shared corners and quad diagonals, where several triangles of
    different instances of one group report the same t and the (instance,
    triangle) order decides -- bit-exact against the oracle's brute force."""
    data = scenes.cornell(64, 64)
    sc = lx.Scene(ctx, data)
    ref = oracle.OracleScene(data)
    rng = np.random.default_rng(11)
    parts = []

    def add(o, d, tmax=1e16):
        r = np.zeros((len(o), 8), np.float32)
        r[:, 0:3] = o
        r[:, 4:7] = d
        r[:, 7] = tmax
        parts.append(r)

    inside = ([-0.9, 0.1, -0.9], [0.9, 1.9, 0.9])
    for axis in range(3):
        for sgn in (1.0, -1.0):
            o = rng.uniform(*inside, (400, 3)).astype(np.float32)
            d = np.zeros((400, 3), np.float32)
            d[:, axis] = sgn
            d[:200, (axis + 1) % 3] = -0.0
            add(o, d)
    o = np.zeros((400, 3), np.float32)
    o[:, [0, 2]] = rng.uniform(-0.9, 0.9, (400, 2))
    d = np.zeros((400, 3), np.float32)
    d[:, [0, 2]] = rng.normal(size=(400, 2))
    add(o, d)
    targets = []
    for wall in data.meshes[:5]:
        p = wall["pos"]
        targets += [p[k] for k in range(len(p))]
        for a, b, c in wall["idx"]:
            targets += [0.5 * (p[a] + p[c]), 0.5 * (p[a] + p[b])]
    targets = np.array(targets, np.float32)
    o = rng.uniform(*inside, (len(targets) * 8, 3)).astype(np.float32)
    add(o, np.repeat(targets, 8, axis=0) - o)
    rays = np.concatenate(parts)
    g = lx.trace_closest(sc, rays)
    b = ref.trace(rays, mode=1)
    assert np.array_equal(g, b)
    assert (g["t"] > 0).mean() > 0.8  # +z rays leave through the open front
    assert np.array_equal(lx.trace_occluded(sc, rays), ref.trace(rays, mode=1, occluded=True))
