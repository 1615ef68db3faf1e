"""Wide rounds (tuning round_frames = R > 1): a round renders up to R
consecutive samples of a render as R copies of its job, each into a colour
region of its own, added to the accumulation in sample order; with
overlap_rounds, wide rounds overlap on separate streams.  The image bits, ray
counts and sample counts equal one-round-at-a-time rendering for any number
of samples, across calls, resets, changes of R and mode switches, and
against the oracle."""
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


KEYS = ("radiance_rays", "dl_shadow_rays", "env_shadow_rays", "path_bounces")


def _render(ctx, sc, kw, calls, frames, overlap):
    ctx.set_tuning(round_frames=frames, overlap_rounds=overlap)
    r = lx.Render(sc, **kw)
    for first, n in calls:
        r.iterate(first, n)
    img, spp = r.accum()
    st = r.stats()
    r.close()
    return img, spp, {k: st[k] for k in KEYS}


@pytest.mark.parametrize("calls", [[(1, 2)], [(1, 5)], [(1, 3), (4, 4)], [(1, 1), (2, 2), (4, 3)], [(1, 9), (10, 2)]])
def test_wide_equals_sequential(tuned, calls):
    data = scenes.material_zoo(96, 64, level=3, env_size=(128, 64))
    sc = lx.Scene(tuned, data)
    for kw in (dict(emis=True), dict(dmis=True)):
        b = _render(tuned, sc, kw, calls, frames=1, overlap=0)
        for frames, overlap in ((2, 0), (3, 1), (4, 3), (8, 1)):
            a = _render(tuned, sc, kw, calls, frames=frames, overlap=overlap)
            assert np.array_equal(a[0], b[0]), (calls, kw, frames, overlap)
            assert a[1] == b[1] == sum(n for _, n in calls)
            assert a[2] == b[2], (calls, kw, frames, overlap)


def test_wide_renders_oracle_bits(tuned, oracle):
    data = scenes.icosphere_hdri(128, 72, level=5, pieces=16, env_size=(256, 128))
    sc = lx.Scene(tuned, data)
    tuned.set_tuning(round_frames=4, overlap_rounds=1)
    r = lx.Render(sc, emis=True)
    r.iterate(1, 6)
    img, spp = r.accum()
    want, _ = oracle.OracleScene(data).render(1, 6, emis=True, threads=8)
    assert spp == 6
    assert np.array_equal(img.reshape(-1, 3).view(np.uint32), want.view(np.uint32))


def test_changing_width_reset_and_mode_switch(tuned):
    """One image continued across calls with different widths (the round
    buffers are remade), sequential calls and a reset in between."""
    data = scenes.cornell(72, 64)
    sc = lx.Scene(tuned, data)
    r = lx.Render(sc, dmis=True)
    tuned.set_tuning(round_frames=4, overlap_rounds=1)
    r.iterate(1, 3)
    r.reset()
    first = 1
    for frames, overlap, n in ((4, 1, 5), (2, 0, 3), (1, 1, 2), (1, 0, 1), (8, 3, 9), (3, 2, 4)):
        tuned.set_tuning(round_frames=frames, overlap_rounds=overlap)
        r.iterate(first, n)
        first += n
    img, spp = r.accum()
    st = r.stats()
    tuned.set_tuning(round_frames=1, overlap_rounds=0)
    ref = lx.Render(sc, dmis=True)
    ref.iterate(1, first - 1)
    assert spp == ref.accum()[1] == first - 1
    assert np.array_equal(img, ref.accum()[0])
    rs = ref.stats()
    assert all(st[k] == rs[k] for k in KEYS)


def test_wide_rounds_with_event_timing(tuned):
    """Event timing brackets each wide round's kernels on its own stream: the
    bits are unchanged and one tail launch is counted per round."""
    data = scenes.cornell(64, 48)
    sc = lx.Scene(tuned, data)
    want = _render(tuned, sc, dict(dmis=True), [(1, 7)], frames=1, overlap=0)
    tuned.set_tuning(round_frames=3, overlap_rounds=1)
    r = lx.Render(sc, dmis=True)
    r.set_timing(events=True)
    r.iterate(1, 7)
    img, spp = r.accum()
    st = r.stats()
    r.close()
    assert spp == 7 and np.array_equal(img, want[0])
    assert st["tail_launches"] == 3 and st["tail_ms"] > 0
    assert all(st[k] == want[2][k] for k in KEYS)


def test_round_frames_validated(tuned):
    for bad in (0, 33):
        with pytest.raises(lx.LuxionError, match="round_frames"):
            tuned.set_tuning(round_frames=bad)


def test_wide_falls_back_when_the_budget_is_tight():
    """In a context whose HBM budget holds the render but not R = 32 copies of
    its path buffers, wide iterate renders as without wide rounds."""
    data = scenes.cornell(128, 128)
    with_room = lx.Context(0)
    try:
        sc = lx.Scene(with_room, data)
        with_room.set_tuning(round_frames=1, overlap_rounds=0)
        r = lx.Render(sc, dmis=True)
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
            with pytest.raises(lx.LuxionError, match="budget"):
                lx.Render(sc, dmis=True)  # a second render does not fit, so 32 copies do not
            c.set_tuning(round_frames=32)
            r.iterate(1, 3)
            img, spp = r.accum()
            assert spp == 3 and np.array_equal(img, want), mib
            r.close()
        finally:
            c.close()
        return
    pytest.fail("no budget fitted the render")


def test_batch_wide_rounds_equal_one_sample_per_round(tuned):
    """The batched form with round_frames R takes up to R samples of every job
    per round (job copies, colour regions added in sample order): the same
    images, sample counts and status as one sample per round, for two scenes,
    spp 1..9 and later first iterations; with fewer path slots than R copies
    need it renders one sample per round."""
    a = lx.Scene(tuned, scenes.cornell(64, 64))
    zd = scenes.material_zoo(80, 48, level=2, env_size=(64, 32))
    b = lx.Scene(tuned, zd)
    jobs = lambda: [lx.make_job(a, scenes.cornell(64, 64).camera, 5, dmis=True, seed=5),
                    lx.make_job(b, zd.camera, 2, emis=True, seed=9),
                    lx.make_job(a, scenes.cornell(48, 80).camera, 1, dmis=True, seed=11),
                    lx.make_job(b, zd.camera, 9, emis=True, first_iter=7),
                    lx.make_job(a, scenes.cornell(33, 17).camera, 3, dmis=True)]
    tuned.set_tuning(round_frames=1, overlap_rounds=0)
    want = lx.batch_run(tuned, jobs())
    for frames, paths in ((2, 0), (3, 0), (4, 0), (8, 0), (4, 3 * 64 * 64)):
        tuned.set_tuning(round_frames=frames, overlap_rounds=1, max_paths_per_round=paths)
        got = lx.batch_run(tuned, jobs())
        for x, y in zip(want, got):
            assert x["status"] == y["status"] == 0, frames
            assert x["spp"] == y["spp"], frames
            assert np.array_equal(x["image"], y["image"]), (frames, paths)
