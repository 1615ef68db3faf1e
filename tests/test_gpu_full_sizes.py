"""The benchmark configurations at their stated sizes on the GPU, and the
display path's bytes:

  * the instanced interior at full size (64 meshes of 10k-500k triangles,
    10,008 instances, 8 emissive meshes, glass, 1024^2 RGBA8 checker albedo and
    alpha-mask textures, 1920x1080, depth 16, -dmis): 1 spp bit-exact against
    the oracle on 2,000 random pixels;
  * the batched 256-job mix on the 1,310,722-triangle HDRI scene (sides
    U{64..512}, spp capped at 4 for time, seeded orbit cameras): every job
    finishes, a sample of jobs equals the same jobs rendered alone bit for bit
    and the oracle on a pixel subset;
  * lx_render_read_ldr (sendImageToPBO, src/kernels/image_kernels.cu:6-46, with
    the tonemaps of src/tonemapping.cu:3-30): Reinhard, AgX and ACES bytes
    against a float32 numpy restatement, within 1 level;
  * a compiled C++ caller of include/luxion_mi355x.h (tests/capi_consumer.cpp)
    against the oracle.
"""
import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import _ffi as F
from luxion_amd import scenes, workloads

pytestmark = pytest.mark.gpu


def test_interior_full_size_1spp_matches_oracle(ctx, oracle):
    data = scenes.instanced_interior(1920, 1080)
    assert len(data.instances) == 10008 and data.num_tris > 15_000_000
    assert data.trace_depth == 16
    sc = lx.Scene(ctx, data)
    print(f"interior: {data.num_tris} unique triangles, {len(data.instances)} instances, "
          f"{sc.info()['device_bytes'] / 2**20:.0f} MiB on device", flush=True)
    r = lx.Render(sc, dmis=True)
    r.iterate(1, 1)
    img, spp = r.accum()
    st = r.stats()
    r.close()
    assert spp == 1 and st["dl_shadow_rays"] > 0
    ref = oracle.OracleScene(data)
    pix = np.sort(np.random.default_rng(7).choice(1920 * 1080, 2000, replace=False)).astype(np.uint32)
    want, _ = ref.render(1, 1, pixels=pix, dmis=True, threads=16)
    got = img.reshape(-1, 3)[pix]
    print(f"interior 1 spp: exact-match fraction {np.mean(got == want):.6f}, "
          f"lit pixels {np.mean(want.max(1) > 0):.3f}", flush=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (want.max(1) > 0).sum() >= 10  # a dark interior: about 1 % of pixels are lit at 1 spp


def test_batched_256_job_mix_on_hdri_scene(ctx, oracle):
    data = scenes.icosphere_hdri(64, 64)
    sc = lx.Scene(ctx, data)
    spec = workloads.mixed_jobs([(sc, dict(emis=True), (0.0, 0.0, 0.0), 3.0)], 256, seed=3, spp=(1, 4))
    assert len(spec) == 256
    sides = [cam["width"] for _, cam, _ in spec]
    assert min(sides) >= 64 and max(sides) <= 512 and len(set(sides)) > 100
    jobs = [lx.make_job(s_, cam, **kw) for s_, cam, kw in spec]
    res = lx.batch_run(ctx, jobs)
    assert all(r_["status"] == 0 and r_["spp"] == kw["spp"] for r_, (_, _, kw) in zip(res, spec))
    ref = oracle.OracleScene(data)
    rng = np.random.default_rng(11)
    for k in rng.choice(256, 6, replace=False):
        s_, cam, kw = spec[k]
        alone = lx.Render(s_, camera=cam, seed=kw["seed"], emis=True)
        alone.iterate(1, kw["spp"])
        a, _ = alone.accum()
        alone.close()
        assert np.array_equal(a.view(np.uint32), res[k]["image"].view(np.uint32)), k
        n = cam["width"] * cam["height"]
        pix = np.sort(rng.choice(n, min(n, 300), replace=False)).astype(np.uint32)
        want, _ = ref.render(1, kw["spp"], pixels=pix, emis=True, camera=cam, seed=kw["seed"], threads=16)
        got = res[k]["image"].reshape(-1, 3)[pix]
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k


def _tonemap_reference(sum_rgb, spp, mode):
    """sendImageToPBO in float32: average (1 / spp), tonemap, pow 0.45,
    (int)(v * 255.0) clamped to [0, 255]."""
    f = np.float32
    p = sum_rgb.astype(f) * (f(1.0) / f(spp))
    if mode == F.LX_TONEMAP_REINHARD:
        p = p / (p + f(1.0))
    elif mode == F.LX_TONEMAP_AGX:
        v = np.maximum(p, f(1e-6))
        v = (np.log2(v).astype(f) + f(10.0)) / f(16.0)
        v = np.clip(v, f(0), f(1))
        x2 = v * v
        x4 = x2 * x2
        v = f(15.5) * x4 * v - f(40.14) * x4 + f(31.96) * x2 * v - f(6.86) * x2 + f(0.429) * v + f(0.0322)
        r, g, b = v[..., 0], v[..., 1], v[..., 2]
        p = np.stack([r * f(1.1968790) - g * f(0.0980208) - b * f(0.0990297),
                      -r * f(0.0528968) + g * f(1.1519031) - b * f(0.0989611),
                      -r * f(0.0529716) - g * f(0.0980434) + b * f(1.1510736)], -1)
        p = np.clip(p, f(0), f(1))
    else:
        a, b, c, d, e = f(2.51), f(0.03), f(2.43), f(0.59), f(0.14)
        p = np.clip((p * (a * p + b)) / (p * (c * p + d) + e), f(0), f(1))
    p = np.power(p.astype(np.float64), 0.45)
    return np.clip((p * 255.0).astype(np.int64), 0, 255)


def test_ldr_tonemaps_match_numpy_restatement(ctx):
    data = scenes.cornell(96, 96)
    sc = lx.Scene(ctx, data)
    r = lx.Render(sc, dmis=True)
    r.iterate(1, 8)
    acc, spp = r.accum()
    assert acc.max() / spp > 4.0 and (acc == 0).any()  # the light is in view; unlit pixels exist
    for mode in (F.LX_TONEMAP_REINHARD, F.LX_TONEMAP_AGX, F.LX_TONEMAP_ACES):
        got = r.ldr(mode)
        want = _tonemap_reference(acc.reshape(96, 96, 3), spp, mode)
        diff = np.abs(got[..., :3].astype(np.int64) - want)
        print(f"tonemap {mode}: exact {np.mean(diff == 0):.4f}, max diff {diff.max()}", flush=True)
        assert diff.max() <= 1 and np.mean(diff == 0) > 0.99
        assert (got[..., 3] == 0).all()
        assert len(np.unique(got[..., :3])) > 50


def test_compiled_cpp_consumer_renders_the_oracle_image(ctx, oracle, tmp_path):
    """tests/capi_consumer.cpp (the INTEGRATION.md C++ example, compiled
    against include/luxion_mi355x.h by the library's Makefile) renders the
    reference's cornell.glb, 4 iterations with -dmis, in a process of its
    own; its radiance sums equal the oracle's on 2,000 pixels and the ctypes
    path's everywhere."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(here, "_build", "lx_capi_consumer")
    if not os.path.exists(exe):
        pytest.fail("tests/_build/lx_capi_consumer is missing: build with __graft_entry__.build()")
    glb = os.path.join(here, "golden", "scenes", "cornell.glb")
    out = str(tmp_path / "consumer_cornell.bin")
    print("running the compiled consumer on cornell.glb", flush=True)
    subprocess.run([exe, glb, "4", out, "-dmis"], check=True, timeout=120)
    raw = np.fromfile(out, np.uint8)
    w, h, spp = np.frombuffer(raw[:12].tobytes(), np.int32)
    img = np.frombuffer(raw[12:].tobytes(), np.float32).reshape(-1, 3)
    assert spp == 4 and img.shape[0] == w * h
    data = scenes.from_glb(glb)
    pix = np.random.default_rng(3).choice(w * h, 2000, replace=False).astype(np.uint32)
    want, _ = oracle.OracleScene(data).render(1, 4, pixels=pix, dmis=True, threads=16)
    assert np.array_equal(img[pix].view(np.uint32), want.view(np.uint32))
    r = lx.Render(lx.Scene.load_glb(ctx, glb), dmis=True)
    r.iterate(1, 4)
    assert np.array_equal(r.accum()[0].reshape(-1, 3).view(np.uint32), img.view(np.uint32))
