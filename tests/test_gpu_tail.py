"""Tail kernel: bounces from `wavefront_bounces` on run in one persistent
launch where each lane carries a path to its end and then takes the next live
path.  A path's arithmetic is the same in both forms, so every split gives
the same bits as the all-wavefront round and as the oracle, the same ray and
bounce counts, and the same BVH visit totals."""
import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import scenes

pytestmark = pytest.mark.gpu

SPLITS = [-1, 0, 1, 3, 6]


@pytest.fixture
def tuned(ctx):
    saved = ctx.tuning()
    yield ctx
    ctx.set_tuning(**saved)


def _render(ctx, sc, wf, spp, timing=0, **kw):
    ctx.set_tuning(wavefront_bounces=wf)
    r = lx.Render(sc, **kw)
    if timing:
        r.set_timing(events=bool(timing & 1), count_work=bool(timing & 2))
    r.iterate(1, spp)
    img = r.accum()[0]
    st = r.stats()
    r.close()
    return img, st


def test_tail_splits_bitexact_cornell_dmis(tuned, oracle):
    data = scenes.cornell(96, 96)
    sc = lx.Scene(tuned, data)
    base, st0 = _render(tuned, sc, -1, 4, dmis=True)
    want, _ = oracle.OracleScene(data).render(1, 4, dmis=True, threads=8)
    assert np.array_equal(base, want.reshape(base.shape))
    for wf in SPLITS[1:]:
        img, st = _render(tuned, sc, wf, 4, dmis=True)
        print(f"cornell dmis wf={wf}: equal={np.array_equal(img, base)}", flush=True)
        assert np.array_equal(img, base)
        for k in ("radiance_rays", "dl_shadow_rays", "env_shadow_rays", "path_bounces"):
            assert st[k] == st0[k], k


def test_tail_splits_bitexact_icosphere_emis(tuned, oracle):
    data = scenes.icosphere_hdri(160, 90, level=5, pieces=16, env_size=(256, 128))
    sc = lx.Scene(tuned, data)
    base, st0 = _render(tuned, sc, -1, 2, emis=True)
    want, _ = oracle.OracleScene(data).render(1, 2, emis=True, threads=8)
    assert np.array_equal(base, want.reshape(base.shape))
    for wf in SPLITS[1:]:
        img, st = _render(tuned, sc, wf, 2, emis=True)
        print(f"icosphere emis wf={wf}: equal={np.array_equal(img, base)}", flush=True)
        assert np.array_equal(img, base)
        assert st["env_shadow_rays"] == st0["env_shadow_rays"]
        assert st["path_bounces"] == st0["path_bounces"]


def test_tail_counting_and_event_timing(tuned):
    data = scenes.icosphere_hdri(96, 64, level=4, pieces=8, env_size=(128, 64))
    sc = lx.Scene(tuned, data)
    base, s0 = _render(tuned, sc, -1, 2, timing=3, emis=True)
    img, s1 = _render(tuned, sc, 2, 2, timing=3, emis=True)
    assert np.array_equal(img, base)
    for k in ("closest_nodes", "closest_tris", "closest_insts", "any_nodes", "any_tris", "any_insts"):
        assert s0[k] == s1[k] and s0[k] > 0, k
    assert s0["tail_launches"] == 0 and s0["tail_closest_nodes"] == 0
    assert s1["tail_launches"] == 2 and s1["trace_launches"] == 4
    assert 0 < s1["tail_closest_nodes"] < s1["closest_nodes"]
    assert s1["tail_closest_rays"] > 0 and s1["tail_any_rays"] > 0
    assert s1["tail_ms"] > 0.0
    assert s0["tail_bounces"] == 0 and 0 < s1["tail_bounces"] < s1["path_bounces"]


def test_tail_in_batch_equals_alone(tuned):
    corn = lx.Scene(tuned, scenes.cornell(48, 48))
    cams = [scenes.look_at_camera((0.2 * k, 1.0, 3.4), (0, 1, 0), (0, 1, 0), 40, 24 + 8 * k, 32) for k in range(4)]
    spec = [dict(spp=2 + k, dmis=bool(k % 2), seed=k, first_iter=1 + k) for k in range(4)]
    tuned.set_tuning(wavefront_bounces=-1)
    want = []
    for cam, kw in zip(cams, spec):
        r = lx.Render(corn, dmis=kw["dmis"], seed=kw["seed"], camera=cam)
        r.iterate(kw["first_iter"], kw["spp"])
        want.append(r.accum()[0])
        r.close()
    tuned.set_tuning(wavefront_bounces=2)
    res = lx.batch_run(tuned, [lx.make_job(corn, cam, **kw) for cam, kw in zip(cams, spec)])
    for r, w in zip(res, want):
        assert r["status"] == 0
        assert np.array_equal(r["image"], w)


def test_autotune_picks_valid_tuning_same_image(tuned, tmp_path):
    data = scenes.icosphere_hdri(64, 48, level=4, pieces=8, env_size=(128, 64))
    sc = lx.Scene(tuned, data)
    cam = data.camera
    jobs = [lx.make_job(sc, cam, spp=2, emis=True), lx.make_job(sc, cam, spp=1, emis=True, seed=9)]
    before = lx.batch_run(tuned, jobs)
    path = str(tmp_path / "tuning.json")
    best = tuned.autotune(jobs, path)
    assert best["bvh_width"] in (4, 8) and best["block_size"] in (128, 256)
    assert tuned.tuning() == best
    after = lx.batch_run(tuned, jobs)
    for a, b in zip(before, after):
        assert np.array_equal(a["image"], b["image"])
    tuned.set_tuning(wavefront_bounces=-1, queue_chunk=32)
    tuned.load_tuning(path)
    assert tuned.tuning() == best


@pytest.mark.parametrize("overlap", [0, 1])
def test_event_timed_rounds_equal_graph_replay(tuned, overlap):
    """Event-timed eager rounds give the bits of the graph replay (overlap 0)
    and of the overlapped rounds (overlap 1)."""
    data = scenes.icosphere_hdri(96, 64, level=4, pieces=8, env_size=(128, 64))
    sc = lx.Scene(tuned, data)
    tuned.set_tuning(overlap_rounds=overlap, use_graph=1)
    a = lx.Render(sc, emis=True)
    a.iterate(1, 3)
    b = lx.Render(sc, emis=True)
    b.set_timing(events=True)
    b.iterate(1, 3)
    assert np.array_equal(a.accum()[0], b.accum()[0])
    s = b.stats()
    assert s["tail_launches"] == 3 and s["tail_ms"] > 0 and s["other_ms"] > 0
