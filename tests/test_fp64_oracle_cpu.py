"""The oracle's double-precision build (-DLXO_FP64): the same restatement and
random streams with double arithmetic and libm's double transcendentals, the
FP64 reference for converged-image error.  Pinned independently of the fp32
build by brute-force traversal, a furnace closed form and the agreement of
two estimators, then compared with the fp32 build."""
import numpy as np
import pytest

from luxion_amd import _ffi as F
from luxion_amd import scenes


def _rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    r[:, 4:7] = rng.normal(size=(n, 3))
    r[:, 7] = 1e16
    return r


def _rel_rmse(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.mean(b))


def test_fp64_traversal_matches_fp64_brute_force(oracle):
    data = scenes.icosphere_hdri(64, 36, level=4, pieces=4, env_size=(64, 32))
    o = oracle.OracleScene(data, fp64=True)
    r = _rays(4000, [-1.5, -1.2, -1.5], [1.5, 1.5, 1.5], 4)
    assert np.array_equal(o.trace(r, mode=0), o.trace(r, mode=1))
    assert np.array_equal(o.trace(r, mode=0, occluded=True), o.trace(r, mode=1, occluded=True))


def test_fp64_hits_agree_with_fp32(oracle):
    """The same triangles are hit except where a ray grazes an edge shared by
    two triangles; distances agree to a few fp32 ulps."""
    data = scenes.icosphere_hdri(64, 36, level=4, pieces=4, env_size=(64, 32))
    r = _rays(20000, [-1.5, -1.2, -1.5], [1.5, 1.5, 1.5], 5)
    a = oracle.OracleScene(data).trace(r)
    b = oracle.OracleScene(data, fp64=True).trace(r)
    same = (a["instance"] == b["instance"]) & (a["prim"] == b["prim"])
    assert same.mean() > 0.999
    hit = same & (a["t"] > 0)
    assert hit.sum() > 5000
    # rounding of the fp32 test scales with the operands (scene size ~1), not with t
    np.testing.assert_allclose(a["t"][hit], b["t"][hit], rtol=2e-6, atol=2e-7)


def test_fp64_legacy_diffuse_furnace_is_the_albedo(oracle):
    """A Lambertian sphere of albedo a in a constant environment of radiance 1
    (legacy diffuse shading: cosine sampling, f cos / pdf = a) returns a on
    every pixel that sees it and 1 elsewhere; in double precision to the
    rounding of the fp32 output."""
    a = (0.5, 0.25, 0.75)
    s = scenes.SceneData()
    m = s.add_material(scenes.material(a, material_type=F.LX_MATERIAL_DIFFUSE))
    v, f = scenes.icosphere(2)
    s.add_instance(s.add_mesh(v, f, None, scenes.sphere_uv(v)), m)
    s.env = np.ones((16, 32, 4), np.float32)
    s.camera = scenes.look_at_camera((0.0, 0.0, 3.2), (0, 0, 0), (0, 1, 0), 45.0, 48, 48)
    s.trace_depth = 8
    img, _ = oracle.OracleScene(s, fp64=True).render(1, 1, shading_model=F.LX_SHADING_LEGACY, threads=8)
    sphere = np.any(img != 1.0, axis=1)
    assert sphere.sum() > 200 and (~sphere).sum() > 200
    np.testing.assert_allclose(img[sphere], np.broadcast_to(np.float64(a), (sphere.sum(), 3)), rtol=1e-7)


def test_fp64_mis_estimators_agree(oracle):
    """-dmis and plain path tracing estimate the same Cornell image."""
    o = oracle.OracleScene(scenes.cornell(24, 24), fp64=True)
    a, _ = o.render(1, 96, dmis=True, threads=8, max_depth=16)
    b, _ = o.render(1, 96, dmis=False, threads=8, max_depth=16)
    assert a.mean() / 96 > 0.05
    assert abs(a.mean() - b.mean()) / b.mean() < 0.04


CASES = {
    # name: (scene, render flags, bound on the same-seed relative RMSE fp32 vs FP64)
    "cornell": (lambda: scenes.cornell(96, 96), dict(dmis=True), 1e-3),
    "icosphere_hdri": (lambda: scenes.icosphere_hdri(96, 54, level=5, pieces=16, env_size=(256, 128)),
                       dict(emis=True), 1e-3),
    "instanced_interior": (lambda: scenes.instanced_interior(64, 36, n_meshes=6, n_instances=60,
                                                             tri_range=(80, 1300), n_emissive=2, tex_size=64),
                           dict(dmis=True), 1e-2),
    "displaced_grid": (lambda: scenes.displaced_grid(quads=64, width=64, height=36), dict(emis=True), 5e-3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_oracle_against_fp64(oracle, name):
    """Same random streams: the fp32 and FP64 images differ only by rounding
    (and the rare path whose random decision falls the other way), far below
    the Monte Carlo noise; against a converged FP64 estimate the fp32 image's
    error equals the FP64 image's own."""
    make, kw, bound = CASES[name]
    data = make()
    spp = 8
    a, _ = oracle.OracleScene(data).render(1, spp, threads=8, **kw)
    o64 = oracle.OracleScene(data, fp64=True)
    b, _ = o64.render(1, spp, threads=8, **kw)
    other, _ = o64.render(1, spp, threads=8, seed=777, **kw)
    conv, _ = o64.render(1 + spp, 8 * spp, threads=8, **kw)
    conv = conv / 8.0
    same = _rel_rmse(a, b)
    noise = _rel_rmse(other, b)
    print(f"{name}: same seeds {same:.2e}, other seeds {noise:.2e}")
    assert same < bound, same
    assert same < noise / 50, (same, noise)
    ea, eb = _rel_rmse(a, conv), _rel_rmse(b, conv)
    assert abs(ea - eb) <= 0.01 * eb, (ea, eb)
