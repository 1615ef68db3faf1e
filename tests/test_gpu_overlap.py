"""Overlapped rounds (tuning overlap_rounds = 1..3): consecutive rounds of one
render run on 1 + overlap_rounds streams, each round's colours in a buffer of
its own, added to the accumulation in round order.  The image bits, ray counts and
sample counts equal one-round-at-a-time rendering, for odd and even sample
counts, across calls, resets and mode switches, and against the oracle."""
import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuned(ctx):
    saved = ctx.tuning()
    yield ctx
    ctx.set_tuning(**saved)


def _render(ctx, sc, kw, calls, overlap, use_graph=1, wf=0):
    ctx.set_tuning(overlap_rounds=overlap, use_graph=use_graph, wavefront_bounces=wf)
    r = lx.Render(sc, **kw)
    for first, n in calls:
        r.iterate(first, n)
    img, spp = r.accum()
    st = r.stats()
    r.close()
    return img, spp, {k: st[k] for k in ("radiance_rays", "dl_shadow_rays", "env_shadow_rays", "path_bounces")}


@pytest.mark.parametrize("calls", [[(1, 2)], [(1, 5)], [(1, 3), (4, 4)], [(1, 1), (2, 2), (4, 3)], [(1, 7), (8, 2)]])
def test_overlapped_equals_sequential(tuned, calls):
    data = scenes.material_zoo(96, 64, level=3, env_size=(128, 64))
    sc = lx.Scene(tuned, data)
    for kw in (dict(emis=True), dict(dmis=True)):
        b = _render(tuned, sc, kw, calls, overlap=0)
        for depth in (1, 2, 3):
            a = _render(tuned, sc, kw, calls, overlap=depth)
            assert np.array_equal(a[0], b[0]), (calls, kw, depth)
            assert a[1] == b[1] == sum(n for _, n in calls)
            assert a[2] == b[2], (calls, kw, depth)


def test_changing_depth_between_calls(tuned):
    """A render continues one image across calls made with different numbers
    of rounds in flight (round sets are added as the depth grows)."""
    data = scenes.cornell(72, 64)
    sc = lx.Scene(tuned, data)
    r = lx.Render(sc, dmis=True)
    first = 1
    for depth, n in ((1, 3), (3, 5), (2, 2), (0, 1), (3, 4)):
        tuned.set_tuning(overlap_rounds=depth)
        r.iterate(first, n)
        first += n
    tuned.set_tuning(overlap_rounds=0)
    ref = lx.Render(sc, dmis=True)
    ref.iterate(1, first - 1)
    assert r.accum()[1] == first - 1
    assert np.array_equal(r.accum()[0], ref.accum()[0])
    assert r.stats()["radiance_rays"] == ref.stats()["radiance_rays"]


def test_overlapped_with_per_bounce_launches_and_eager(tuned):
    data = scenes.cornell(80, 72)
    sc = lx.Scene(tuned, data)
    want = _render(tuned, sc, dict(dmis=True), [(1, 6)], overlap=0, use_graph=0, wf=-1)
    for wf in (-1, 2, 0):
        got = _render(tuned, sc, dict(dmis=True), [(1, 6)], overlap=1, wf=wf)
        assert np.array_equal(got[0], want[0]), wf
        assert got[2] == want[2], wf


def test_overlapped_renders_oracle_bits(tuned, oracle):
    data = scenes.icosphere_hdri(128, 72, level=5, pieces=16, env_size=(256, 128))
    sc = lx.Scene(tuned, data)
    tuned.set_tuning(overlap_rounds=1)
    r = lx.Render(sc, emis=True)
    r.iterate(1, 3)
    img, spp = r.accum()
    want, _ = oracle.OracleScene(data).render(1, 3, emis=True, threads=8)
    assert spp == 3
    assert np.array_equal(img.reshape(-1, 3), want)


def test_reset_between_overlapped_calls_and_mode_switch(tuned):
    data = scenes.cornell(64, 64)
    sc = lx.Scene(tuned, data)
    tuned.set_tuning(overlap_rounds=1)
    r = lx.Render(sc, dmis=True)
    r.iterate(1, 4)
    r.reset()
    r.iterate(1, 3)
    tuned.set_tuning(overlap_rounds=0)
    r.iterate(4, 2)  # one round after another, continuing the same image
    tuned.set_tuning(overlap_rounds=1)
    r.iterate(6, 2)
    img, spp = r.accum()
    st = r.stats()
    ref = lx.Render(sc, dmis=True)
    tuned.set_tuning(overlap_rounds=0)
    ref.iterate(1, 7)
    assert spp == 7 and ref.accum()[1] == 7
    assert np.array_equal(img, ref.accum()[0])
    rs = ref.stats()
    assert st["radiance_rays"] == rs["radiance_rays"] and st["path_bounces"] == rs["path_bounces"]


def test_overlap_rounds_validated(tuned):
    with pytest.raises(lx.LuxionError, match="overlap_rounds"):
        tuned.set_tuning(overlap_rounds=4)


def _batch(ctx, jobs, **tuning):
    ctx.set_tuning(**tuning)
    return lx.batch_run(ctx, jobs())


def test_batch_overlapped_equals_sequential(tuned):
    """Two scenes (per-lane scene in the tail kernel), sample counts 1..5 and
    a later first iteration: images, sample counts and status agree."""
    a = lx.Scene(tuned, scenes.cornell(64, 64))
    zd = scenes.material_zoo(80, 48, level=2, env_size=(64, 32))
    b = lx.Scene(tuned, zd)
    jobs = lambda: [lx.make_job(a, scenes.cornell(64, 64).camera, 5, dmis=True, seed=5),
                    lx.make_job(b, zd.camera, 2, emis=True, seed=9),
                    lx.make_job(a, scenes.cornell(48, 80).camera, 1, dmis=True, seed=11),
                    lx.make_job(b, zd.camera, 4, emis=True, first_iter=7),
                    lx.make_job(a, scenes.cornell(33, 17).camera, 3, dmis=True)]
    seq = _batch(tuned, jobs, overlap_rounds=0)
    ovl = _batch(tuned, jobs, overlap_rounds=1)
    for x, y in zip(seq, ovl):
        assert x["status"] == y["status"] == 0
        assert x["spp"] == y["spp"]
        assert np.array_equal(x["image"], y["image"])
    one = _batch(tuned, lambda: jobs()[:1], overlap_rounds=1)  # a single job: scalar job and scene
    assert np.array_equal(one[0]["image"], seq[0]["image"])


def test_batch_with_limited_paths_runs_rounds_in_sequence(tuned):
    """With fewer path slots than the jobs' pixels some jobs wait for later
    rounds; the overlapped form then falls back to rounds one after another
    and gives the same images."""
    sc = lx.Scene(tuned, scenes.cornell(64, 64))
    jobs = lambda: [lx.make_job(sc, scenes.cornell(64, 64).camera, 3, dmis=True, seed=k) for k in range(4)]
    want = _batch(tuned, jobs, overlap_rounds=0)
    got = _batch(tuned, jobs, overlap_rounds=1, max_paths_per_round=2 * 64 * 64)
    for x, y in zip(want, got):
        assert x["spp"] == y["spp"] == 3
        assert np.array_equal(x["image"], y["image"])


def test_overlap_falls_back_when_the_budget_is_tight():
    """In a context whose HBM budget holds the render but not a second set of
    round buffers, overlapped iterate runs its rounds one after another."""
    data = scenes.cornell(128, 128)
    want = None
    with_room = lx.Context(0)
    try:
        sc = lx.Scene(with_room, data)
        r = lx.Render(sc, dmis=True)
        with_room.set_tuning(overlap_rounds=0)
        r.iterate(1, 3)
        want = r.accum()[0]
        r.close()
    finally:
        with_room.close()
    for mib in range(4, 400, 4):
        c = lx.Context(0, hbm_budget_bytes=mib << 20)
        try:
            sc = lx.Scene(c, data)
            r = lx.Render(sc, dmis=True)
        except lx.LuxionError:
            c.close()
            continue
        try:
            # the budget is tight: a second render (about the size of a second
            # set of round buffers) does not fit
            with pytest.raises(lx.LuxionError, match="budget"):
                lx.Render(sc, dmis=True)
            r.iterate(1, 3)  # overlap_rounds is on by default
            img, spp = r.accum()
            assert spp == 3 and np.array_equal(img, want), mib
            r.close()
        finally:
            c.close()
        return
    pytest.fail("no budget fitted the render")
