"""Queue sort stages (LX_SORT_RAYS, LX_SORT_MATERIALS): they only permute
the live-path queue, so every sort mode, with every per-bounce / tail split,
renders the oracle's bits with the same ray and bounce counts, alone, in a
multi-scene batch and through a graph captured before the tuning changed."""
import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import _ffi as F
from luxion_amd import scenes

pytestmark = pytest.mark.gpu

MODES = [F.LX_SORT_NONE, F.LX_SORT_RAYS, F.LX_SORT_MATERIALS, F.LX_SORT_RAYS | F.LX_SORT_MATERIALS]


@pytest.fixture
def tuned(ctx):
    saved = ctx.tuning()
    yield ctx
    ctx.set_tuning(**saved)


@pytest.mark.parametrize("which", ["cornell_dmis", "zoo_emis", "icosphere_emis"])
def test_sort_modes_render_the_oracle_bits(tuned, oracle, which):
    if which == "cornell_dmis":
        data, kw, spp = scenes.cornell(96, 96), dict(dmis=True), 3
    elif which == "zoo_emis":
        data, kw, spp = scenes.material_zoo(128, 72, level=3, env_size=(128, 64)), dict(emis=True), 2
    else:
        data, kw, spp = scenes.icosphere_hdri(128, 72, level=5, pieces=16, env_size=(256, 128)), dict(emis=True), 2
    sc = lx.Scene(tuned, data)
    want, _ = oracle.OracleScene(data).render(1, spp, threads=8, **kw)
    base_stats = None
    for wf in (-1, 0, 2):
        for mode in MODES:
            tuned.set_tuning(wavefront_bounces=wf, sort_mode=mode)
            r = lx.Render(sc, **kw)
            r.iterate(1, spp)
            img = r.accum()[0]
            st = r.stats()
            r.close()
            assert np.array_equal(img.reshape(-1, 3), want), (wf, mode)
            key = {k: st[k] for k in ("radiance_rays", "dl_shadow_rays", "env_shadow_rays", "path_bounces")}
            base_stats = base_stats or key
            assert key == base_stats, (wf, mode)


def test_sorted_batch_of_two_scenes_equals_unsorted(tuned):
    a = lx.Scene(tuned, scenes.cornell(64, 64))
    zd = scenes.material_zoo(96, 54, level=2, env_size=(64, 32))
    b = lx.Scene(tuned, zd)
    jobs = lambda: [lx.make_job(a, scenes.cornell(64, 64).camera, 3, dmis=True, seed=5),
                    lx.make_job(b, zd.camera, 2, emis=True, seed=9),
                    lx.make_job(a, scenes.cornell(48, 80).camera, 2, dmis=True, seed=11)]
    outs = []
    for wf, mode in ((-1, 0), (1, 3), (-1, 3), (0, 1)):
        tuned.set_tuning(wavefront_bounces=wf, sort_mode=mode)
        outs.append([o["image"] for o in lx.batch_run(tuned, jobs())])
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert np.array_equal(x, y)


def test_graph_recaptured_after_tuning_change(tuned):
    data = scenes.cornell(80, 80)
    sc = lx.Scene(tuned, data)
    tuned.set_tuning(wavefront_bounces=-1, sort_mode=0, use_graph=1, overlap_rounds=0)
    r = lx.Render(sc, dmis=True)
    r.iterate(1, 2)
    tuned.set_tuning(wavefront_bounces=1, sort_mode=3)
    r.iterate(3, 2)  # replays a graph captured with the new shapes
    img = r.accum()[0]
    tuned.set_tuning(wavefront_bounces=0, sort_mode=0)
    ref = lx.Render(sc, dmis=True)
    ref.iterate(1, 4)
    assert np.array_equal(img, ref.accum()[0])


def test_sort_mode_validated(tuned):
    with pytest.raises(lx.LuxionError, match="sort_mode"):
        tuned.set_tuning(sort_mode=4)


@pytest.mark.parametrize("tile,band", [(1, 0), (8, 0), (16, 0), (32, 0), (8, 3), (16, 8)])
def test_pixel_tile_orders_render_the_oracle_bits(tuned, oracle, tile, band):
    """Camera paths enter the queue square by square; odd image sizes leave
    partial squares, which the strips cover."""
    data = scenes.cornell(97, 61)
    sc = lx.Scene(tuned, data)
    want, _ = oracle.OracleScene(data).render(1, 2, dmis=True, threads=8)
    for wf in (0, -1):
        tuned.set_tuning(pixel_tile=tile, tile_band=band, wavefront_bounces=wf)
        r = lx.Render(sc, dmis=True)
        r.iterate(1, 2)
        assert np.array_equal(r.accum()[0].reshape(-1, 3), want), (tile, wf)
    jobs = [lx.make_job(sc, scenes.cornell(w, h).camera, 2, dmis=True) for w, h in ((97, 61), (8, 8), (40, 33))]
    for j, out in zip(jobs, lx.batch_run(tuned, jobs)):
        r = lx.Render(sc, dmis=True, camera=scenes.cornell(j.camera.width, j.camera.height).camera)
        tuned.set_tuning(pixel_tile=1)
        r.iterate(1, 2)
        assert np.array_equal(out["image"], r.accum()[0])
        tuned.set_tuning(pixel_tile=tile)


def test_pixel_tile_validated(tuned):
    with pytest.raises(lx.LuxionError, match="pixel_tile"):
        tuned.set_tuning(pixel_tile=4)
    with pytest.raises(lx.LuxionError, match="tile_band"):
        tuned.set_tuning(tile_band=65)
