"""Material debug views and the legacy (UBER_SHADER 0) shading model on the
GPU, through the C ABI, against the oracle on the same seeded inputs.

Bars: debug views bit-exact; legacy renders bit-exact per pixel for the
plain, -dmis and -emis branches; the tail kernel and the per-bounce kernels give the same
bits; a batch of debug / legacy jobs gives each job's single-render bits;
the LDR read-back of a debug view is untonemapped.
"""
import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import _ffi as F
from luxion_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zoo(ctx, oracle):
    data = scenes.material_zoo(160, 90, level=3, env_size=(128, 64))
    return data, lx.Scene(ctx, data), oracle.OracleScene(data)


def _close(gpu, ref, what):
    gpu = np.asarray(gpu, np.float32).reshape(-1)
    ref = np.asarray(ref, np.float32).reshape(-1)
    exact = float(np.mean(gpu == ref))
    print(f"{what}: exact-match fraction {exact:.6f}, max |diff| {np.abs(gpu - ref).max():.3e}", flush=True)
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), what
    return exact


@pytest.mark.parametrize("flags", [dict(), dict(dmis=True), dict(emis=True)], ids=["plain", "dmis", "emis"])
def test_legacy_render_matches_oracle(ctx, zoo, flags):
    data, sc, ref = zoo
    r = lx.Render(sc, shading_model=F.LX_SHADING_LEGACY, **flags)
    r.iterate(1, 2)
    img, spp = r.accum()
    want, _ = ref.render(1, 2, shading_model=F.LX_SHADING_LEGACY, threads=8, **flags)
    assert spp == 2
    assert _close(img, want, f"legacy {flags}") == 1.0
    # the uber shader renders the same scene differently
    u = lx.Render(sc, **flags)
    u.iterate(1, 2)
    assert not np.array_equal(u.accum()[0], img)


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6])
def test_debug_views_match_oracle_bitexact(ctx, zoo, mode):
    data, sc, ref = zoo
    r = lx.Render(sc, material_debug_mode=mode, emis=True)
    r.iterate(1, 1)
    img, _ = r.accum()
    want, st = ref.render(1, 1, debug_mode=mode, emis=True, threads=8)
    assert np.array_equal(img.reshape(-1, 3), want), mode
    gs = r.stats()
    assert gs["path_bounces"] == 160 * 90 == st["path_bounces"]  # one bounce
    if mode == F.LX_DEBUG_NORMAL_MAP:
        assert len(np.unique(img.reshape(-1, 3), axis=0)) > 50  # the bump map shows


def test_legacy_debug_albedo_of_emissive_is_emission_colour(ctx, zoo):
    data, sc, ref = zoo
    for model in (F.LX_SHADING_UBER, F.LX_SHADING_LEGACY):
        r = lx.Render(sc, material_debug_mode=F.LX_DEBUG_ALBEDO, shading_model=model)
        r.iterate(1, 1)
        want, _ = ref.render(1, 1, debug_mode=F.LX_DEBUG_ALBEDO, shading_model=model, threads=8)
        got = r.accum()[0].reshape(-1, 3)
        assert np.array_equal(got, want)
        emit = np.all(np.isclose(got, (1.0, 0.85, 0.6)), axis=1)
        assert emit.any() == (model == F.LX_SHADING_LEGACY)


def test_tail_and_wavefront_splits_agree_for_legacy(ctx, zoo):
    data, sc, _ = zoo
    saved = ctx.tuning()
    imgs = []
    try:
        for wb in (0, 2, -1):
            ctx.set_tuning(wavefront_bounces=wb)
            r = lx.Render(sc, shading_model=F.LX_SHADING_LEGACY, emis=True)
            r.iterate(3, 2)
            imgs.append(r.accum()[0])
    finally:
        ctx.set_tuning(**saved)
    assert np.array_equal(imgs[0], imgs[1]) and np.array_equal(imgs[0], imgs[2])


def test_debug_ldr_is_untonemapped(ctx, zoo):
    data, sc, _ = zoo
    r = lx.Render(sc, material_debug_mode=F.LX_DEBUG_ROUGHNESS)
    r.iterate(1, 3)
    img, spp = r.accum()
    ldr = r.ldr(F.LX_TONEMAP_ACES)
    pix = img * np.float32(1.0 / spp)
    want = np.clip((pix.astype(np.float64) * 255.0).astype(np.int64), 0, 255)
    assert np.array_equal(ldr[..., :3].astype(np.int64), want)


def test_batch_jobs_with_views_and_models_equal_single_renders(ctx, zoo):
    data, sc, _ = zoo
    cam = data.camera
    specs = [dict(material_debug_mode=F.LX_DEBUG_ALBEDO), dict(shading_model=F.LX_SHADING_LEGACY, dmis=True),
             dict(material_debug_mode=F.LX_DEBUG_METALLIC, shading_model=F.LX_SHADING_LEGACY), dict(emis=True)]
    jobs = [lx.make_job(sc, cam, 2, seed=0, **s) for s in specs]
    res = lx.batch_run(ctx, jobs)
    for s, out in zip(specs, res):
        r = lx.Render(sc, **s)
        r.iterate(1, 2)
        assert out["status"] == 0 and out["spp"] == 2
        assert np.array_equal(out["image"], r.accum()[0]), s


def test_views_and_models_are_validated(ctx, zoo):
    _, sc, _ = zoo
    with pytest.raises(lx.LuxionError, match="material_debug_mode"):
        lx.Render(sc, material_debug_mode=7)
    with pytest.raises(lx.LuxionError, match="shading_model"):
        lx.Render(sc, shading_model=2)
