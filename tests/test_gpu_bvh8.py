"""BVH width 8 (and other leaf sizes): the closest hit is the lexicographic
minimum of (t, instance, primitive) over the accepted candidates, which does
not depend on the tree, so every width and leaf size gives the oracle's hits
and the same rendered bits."""
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


@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_bvh8_trace_matches_oracle(ctx, oracle, leaf):
    data = scenes.icosphere_hdri(64, 36, level=6, pieces=16, env_size=(64, 32))
    sc = lx.Scene(ctx, data, bvh_width=8, leaf_size=leaf)
    assert sc.info()["bvh_nodes"] > 0
    ref = oracle.OracleScene(data)
    r = _rays(40000, [-1.6, -1.2, -1.6], [1.6, 1.6, 1.6], 6)
    assert np.array_equal(lx.trace_closest(sc, r), ref.trace(r, mode=0))
    assert np.array_equal(lx.trace_occluded(sc, r), ref.trace(r, mode=0, occluded=True))


def test_bvh8_instanced_alpha_matches_brute_force(ctx, oracle):
    data = scenes.instanced_interior(64, 36, n_meshes=6, n_instances=60, tri_range=(80, 1300), n_emissive=2,
                                     tex_size=64, trace_depth=6)
    sc = lx.Scene(ctx, data, bvh_width=8)
    ref = oracle.OracleScene(data)
    r = _rays(30000, [-30, -3, -30], [30, 14, 30], 7)
    pix = np.arange(len(r), dtype=np.uint32)
    for alpha in (False, True):
        assert np.array_equal(lx.trace_closest(sc, r, alpha=alpha, iteration=3),
                              ref.trace(r, mode=1, alpha=alpha, iteration=3, pixel=pix)), alpha
        assert np.array_equal(lx.trace_occluded(sc, r, alpha=alpha, iteration=3),
                              ref.trace(r, mode=1, alpha=alpha, iteration=3, pixel=pix, occluded=True)), alpha


def test_bvh8_render_equals_bvh4(ctx, oracle):
    data = scenes.icosphere_hdri(160, 90, level=5, pieces=16, env_size=(256, 128))
    a = lx.Scene(ctx, data, bvh_width=4)
    b = lx.Scene(ctx, data, bvh_width=8)
    ra = lx.Render(a, emis=True)
    ra.iterate(1, 2)
    rb = lx.Render(b, emis=True)
    rb.iterate(1, 2)
    ia, ib = ra.accum()[0], rb.accum()[0]
    assert np.array_equal(ia, ib)
    want, _ = oracle.OracleScene(data).render(1, 2, emis=True, threads=8)
    assert np.array_equal(ib, want.reshape(ib.shape))
    # wavefront kernels at width 8 too
    saved = ctx.tuning()
    try:
        ctx.set_tuning(wavefront_bounces=-1)
        rc = lx.Render(b, emis=True)
        rc.iterate(1, 2)
        assert np.array_equal(rc.accum()[0], ia)
    finally:
        ctx.set_tuning(**saved)


def test_batch_rejects_mixed_widths(ctx):
    data = scenes.cornell(16, 16)
    a = lx.Scene(ctx, data, bvh_width=4)
    b = lx.Scene(ctx, data, bvh_width=8)
    cam = data.camera
    with pytest.raises(lx.LuxionError, match="one BVH width"):
        lx.batch_run(ctx, [lx.make_job(a, cam, spp=1), lx.make_job(b, cam, spp=1)])
    res = lx.batch_run(ctx, [lx.make_job(b, cam, spp=2, dmis=True), lx.make_job(b, cam, spp=1, dmis=True, seed=4)])
    assert all(r["status"] == 0 for r in res)
