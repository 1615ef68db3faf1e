"""GPU BLAS builder (builder 0, linear BVH from Morton codes, collapsed on the
device): every width / leaf size gives the oracle's hits and the same
rendered bits as the host binned-SAH builder (the closest hit does not depend
on the tree), and its trees are well formed."""
import time

import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import scenes

pytestmark = pytest.mark.gpu


def _rays(n, lo, hi, seed, tmax=1e16):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    r[:, 4:7] = rng.normal(size=(n, 3))
    r[:, 7] = tmax
    return r


@pytest.mark.parametrize("width,leaf", [(4, 1), (4, 4), (8, 4), (8, 8), (4, 8)])
def test_gpu_built_trace_matches_oracle(ctx, oracle, width, leaf):
    data = scenes.icosphere_hdri(64, 36, level=6, pieces=16, env_size=(64, 32))
    sc = lx.Scene(ctx, data, builder=0, bvh_width=width, leaf_size=leaf)
    ref = oracle.OracleScene(data)
    r = _rays(40000, [-1.6, -1.2, -1.6], [1.6, 1.6, 1.6], 6)
    assert np.array_equal(lx.trace_closest(sc, r), ref.trace(r, mode=0))
    assert np.array_equal(lx.trace_occluded(sc, r), ref.trace(r, mode=0, occluded=True))


def test_gpu_built_instanced_alpha_and_cornell(ctx, oracle):
    data = scenes.instanced_interior(64, 36, n_meshes=6, n_instances=60, tri_range=(80, 1300), n_emissive=2,
                                     tex_size=64, trace_depth=6)
    sc = lx.Scene(ctx, data, builder=0)
    ref = oracle.OracleScene(data)
    r = _rays(30000, [-30, -3, -30], [30, 14, 30], 7)
    pix = np.arange(len(r), dtype=np.uint32)
    for alpha in (False, True):
        assert np.array_equal(lx.trace_closest(sc, r, alpha=alpha, iteration=3),
                              ref.trace(r, mode=1, alpha=alpha, iteration=3, pixel=pix)), alpha
    cd = scenes.cornell(64, 64)  # meshes of 2 and 12 triangles: single-leaf trees
    sc2 = lx.Scene(ctx, cd, builder=0)
    r2 = _rays(20000, [-0.99, 0.01, -0.99], [0.99, 1.99, 3.0], 5)
    assert np.array_equal(lx.trace_closest(sc2, r2), oracle.OracleScene(cd).trace(r2, mode=1))


def test_gpu_built_render_equals_host_built(ctx):
    data = scenes.icosphere_hdri(160, 90, level=6, pieces=8, env_size=(256, 128))
    imgs = []
    for builder in (0, 1):
        sc = lx.Scene(ctx, data, builder=builder)
        r = lx.Render(sc, emis=True)
        r.iterate(1, 2)
        imgs.append(r.accum()[0])
        info = sc.info()
        assert info["bvh_nodes"] > 0
    assert np.array_equal(imgs[0], imgs[1])


def test_gpu_builder_large_mesh_is_fast(ctx):
    data = scenes.icosphere_hdri(64, 36, level=8, pieces=1, env_size=(64, 32))  # one 1.31M-triangle mesh
    t0 = time.perf_counter()
    sc = lx.Scene(ctx, data, builder=0)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    sc1 = lx.Scene(ctx, data, builder=1)
    t_host = time.perf_counter() - t0
    print(f"1.31M-triangle mesh: GPU build {t_gpu * 1e3:.0f} ms, host build {t_host * 1e3:.0f} ms, "
          f"nodes {sc.info()['bvh_nodes']} vs {sc1.info()['bvh_nodes']}", flush=True)
    r = _rays(20000, [-1.6, -1.2, -1.6], [1.6, 1.6, 1.6], 9)
    assert np.array_equal(lx.trace_closest(sc, r), lx.trace_closest(sc1, r))
