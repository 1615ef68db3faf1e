"""Oracle pins for the material debug views and the legacy (UBER_SHADER 0)
shaders, by closed forms the mathematics fixes:

  - the glass shader's Fresnel: ((1 - eta) / (1 + eta))^2 at normal
    incidence from either side, 1 at grazing incidence and beyond the
    critical angle, and F(theta_i, n1 -> n2) = F(theta_t, n2 -> n1);
  - Lambert: the pdf integrates to 1 over the sphere, samples are
    cosine-distributed (cos^2 uniform), and the throughput factor is the
    albedo;
  - mirror: the reflection law; glass: the reflected fraction equals F, the
    refracted directions obey Snell's law, and the throughput factors are
    (eta_i / eta_t)^2 * mix(1, colour, T) and 1;
  - whole renders in a constant environment: a faceted convex diffuse
    object returns albedo * L per pixel, a mirror colour * L, and the debug
    views return the queried constants (albedo, roughness, metallic, alpha,
    (0.5, 0.5, 1) without a normal map, unit-length normals);
  - the glTF loader's legacy material classes for each clause of the
    reference's isGlass rule.
"""
import math
import os

import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import _ffi as F
from luxion_amd import scenes
from tests import gltf_util

ETAS = (1.33, 1.5, 2.4)


def _f3(v):
    return np.ascontiguousarray(v, np.float32)


def test_glass_fresnel_closed_forms(oracle):
    L = oracle.lib()
    for eta in ETAS:
        f0 = ((1 - eta) / (1 + eta)) ** 2
        assert abs(L.lxo_glass_fresnel(1.0, eta) - f0) < 1e-6
        assert abs(L.lxo_glass_fresnel(-1.0, eta) - f0) < 1e-6  # from inside
        assert L.lxo_glass_fresnel(0.0, eta) == 1.0               # grazing
        crit = math.sqrt(1 - 1 / eta ** 2)                        # cos of the critical angle
        assert L.lxo_glass_fresnel(-0.9 * crit, eta) == 1.0       # total internal reflection
        for ci in np.linspace(0.05, 0.95, 10):
            st = math.sqrt(1 - ci * ci) / eta
            ct = math.sqrt(1 - st * st)
            # reciprocity: entering at theta_i = leaving at theta_t
            assert abs(L.lxo_glass_fresnel(ci, eta) - L.lxo_glass_fresnel(-ct, eta)) < 1e-5
            # textbook unpolarised Fresnel in double precision
            rs = (ci - eta * ct) / (ci + eta * ct)
            rp = (eta * ci - ct) / (eta * ci + ct)
            assert abs(L.lxo_glass_fresnel(ci, eta) - 0.5 * (rs * rs + rp * rp)) < 1e-5


def test_lambert_pdf_normalised_and_floored(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    n = _f3([0.3, 0.8, -0.52] / np.linalg.norm([0.3, 0.8, -0.52]))
    d = rng.normal(size=(40000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dd = _f3(d)  # one array held for the calls (a temporary's buffer may be freed before the call)
    p = np.array([L.lxo_lambert_pdf(dd[i].ctypes.data, n.ctypes.data) for i in range(len(dd))])
    assert abs(np.mean(p) * 4 * math.pi - 1.0) < 0.02  # uniform-sphere estimate of the integral
    below = (d @ n) < 0
    assert np.all(p[below] == np.float32(1e-6))


def test_legacy_diffuse_samples_are_cosine_distributed(oracle):
    L = oracle.lib()
    n = _f3([0.0, 0.6, 0.8])
    d = _f3([0.1, -0.9, 0.2])
    out = np.zeros(3, np.float32)
    cos2 = []
    for i in range(4000):
        k = L.lxo_legacy_sample(F.LX_MATERIAL_DIFFUSE, d.ctypes.data, n.ctypes.data, 1.5, 1, i, 1, 0,
                                out.ctypes.data)
        assert k == 2
        c = float(out @ n)
        assert c >= -1e-6 and abs(np.linalg.norm(out) - 1) < 1e-5
        cos2.append(c * c)
    cos2 = np.sort(cos2)
    ks = np.max(np.abs(cos2 - (np.arange(1, len(cos2) + 1) / len(cos2))))
    assert ks < 1.63 / math.sqrt(len(cos2))  # KS test at 1 %


def test_legacy_mirror_reflection_law(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(5)
    out = np.zeros(3, np.float32)
    for _ in range(200):
        n = _f3(rng.normal(size=3))
        n /= np.linalg.norm(n)
        n = _f3(n)
        d = _f3(rng.normal(size=3))
        d /= np.linalg.norm(d)
        d = _f3(d)
        k = L.lxo_legacy_sample(F.LX_MATERIAL_SPECULAR, d.ctypes.data, n.ctypes.data, 1.5, 1, 0, 1, 0,
                                out.ctypes.data)
        assert k == 0
        want = d.astype(np.float64) - 2 * np.dot(d, n) * n.astype(np.float64)
        assert np.allclose(out, want, atol=1e-6)


@pytest.mark.parametrize("eta", ETAS)
def test_legacy_glass_fresnel_split_and_snell(oracle, eta):
    L = oracle.lib()
    n = _f3([0.0, 1.0, 0.0])
    for ci in (0.95, 0.5, 0.15):
        d = _f3([math.sqrt(1 - ci * ci), -ci, 0.0])  # entering from outside
        F_ = L.lxo_glass_fresnel(ci, eta)
        out = np.zeros(3, np.float32)
        refl = 0
        N = 6000
        for i in range(N):
            k = L.lxo_legacy_sample(F.LX_MATERIAL_GLASS, d.ctypes.data, n.ctypes.data, eta, 7, i, 2, 0,
                                    out.ctypes.data)
            assert k == 1
            if out[1] > 0:
                refl += 1
                assert abs(out[1] - ci) < 1e-6  # mirror about n
            else:
                s_t = math.hypot(out[0], out[2])
                assert abs(s_t - math.sqrt(1 - ci * ci) / eta) < 2e-6  # Snell
        sd = math.sqrt(F_ * (1 - F_) / N)
        assert abs(refl / N - F_) < 4.5 * sd + 1e-3


def test_legacy_update_factors(oracle):
    L = oracle.lib()
    n = _f3([0.0, 0.0, 1.0])
    d = _f3([0.0, 0.6, -0.8])
    col = _f3([0.7, 0.4, 0.2])
    # Lambert: (albedo / pi * cos) / (cos / pi) = albedo
    wi = _f3([0.3, 0.1, math.sqrt(1 - 0.1)])
    thr = _f3([1, 1, 1])
    pdf = L.lxo_legacy_update(F.LX_MATERIAL_DIFFUSE, d.ctypes.data, n.ctypes.data, wi.ctypes.data, col.ctypes.data,
                              1.5, 0.0, thr.ctypes.data)
    assert np.allclose(thr, col, rtol=2e-6)
    assert abs(pdf - wi[2] * 0.3183098) < 1e-6
    # mirror: colour, pdf 1
    thr = _f3([1, 1, 1])
    assert L.lxo_legacy_update(F.LX_MATERIAL_SPECULAR, d.ctypes.data, n.ctypes.data, wi.ctypes.data,
                               col.ctypes.data, 1.5, 0.0, thr.ctypes.data) == 1.0
    assert np.array_equal(thr, col)
    # glass: refraction into ior 1.5 scales by (1/1.5)^2 * mix(1, colour, T); reflection by 1
    T = 0.4
    down = _f3([0.0, 0.3, -math.sqrt(1 - 0.09)])
    thr = _f3([1, 1, 1])
    L.lxo_legacy_update(F.LX_MATERIAL_GLASS, d.ctypes.data, n.ctypes.data, down.ctypes.data, col.ctypes.data, 1.5,
                        T, thr.ctypes.data)
    assert np.allclose(thr, (1 / 1.5) ** 2 * (1 + T * (col - 1)), rtol=2e-6)
    thr = _f3([1, 1, 1])
    L.lxo_legacy_update(F.LX_MATERIAL_GLASS, d.ctypes.data, n.ctypes.data, wi.ctypes.data, col.ctypes.data, 1.5, T,
                        thr.ctypes.data)
    assert np.array_equal(thr, _f3([1, 1, 1]))
    # leaving the glass: (1.5 / 1)^2
    dn = _f3([0.0, 0.6, 0.8])  # travelling along +n: inside a surface whose normal points out
    thr = _f3([1, 1, 1])
    L.lxo_legacy_update(F.LX_MATERIAL_GLASS, dn.ctypes.data, n.ctypes.data, wi.ctypes.data, col.ctypes.data, 1.5,
                        0.0, thr.ctypes.data)
    assert np.allclose(thr, 2.25, rtol=2e-6)


def _furnace(material, flat=True, width=48, height=48, level=2):
    """A faceted unit sphere (outward winding, no vertex normals) in a
    constant environment of radiance 1."""
    s = scenes.SceneData()
    m = s.add_material(material)
    v, f = scenes.icosphere(level)
    s.add_instance(s.add_mesh(v, f, None if flat else v, scenes.sphere_uv(v)), m)
    env = np.ones((16, 32, 4), np.float32)
    s.env = env
    s.camera = scenes.look_at_camera((0.0, 0.0, 3.2), (0, 0, 0), (0, 1, 0), 45.0, width, height)
    s.trace_depth = 8
    return s


def _hit_mask(oracle, data, depth=0):
    """Pixels whose iteration-1 camera ray hits the object.  The camera's RNG
    stream is keyed on the trace depth, so `depth` must be the render's (1
    for a debug view)."""
    ref = oracle.OracleScene(data)
    depth = depth or data.trace_depth
    _, rec, _ = ref.debug_bounce(1, max_depth=depth)
    hit = np.zeros(data.camera["width"] * data.camera["height"], bool)
    r = rec[0]
    hit[r[r[:, 0] > 0, 7].astype(np.int64)] = True
    return ref, hit


def test_legacy_diffuse_furnace_returns_albedo(oracle):
    a = (0.5, 0.25, 0.75)
    data = _furnace(scenes.material(a, material_type=F.LX_MATERIAL_DIFFUSE))
    ref, hit = _hit_mask(oracle, data)
    img, _ = ref.render(1, 1, shading_model=F.LX_SHADING_LEGACY, threads=8)
    assert hit.sum() > 200 and (~hit).sum() > 200
    assert np.array_equal(img[~hit], np.ones(((~hit).sum(), 3), np.float32))
    np.testing.assert_allclose(img[hit], np.broadcast_to(np.float32(a), (hit.sum(), 3)), rtol=4e-6)
    # the uber shader on the same material is not a pure Lambertian: the rule is legacy-specific
    img_u, _ = ref.render(1, 1, threads=8)
    assert not np.allclose(img_u[hit], np.float32(a), rtol=1e-3)


def test_legacy_mirror_furnace_returns_colour(oracle):
    c = (0.9, 0.6, 0.3)
    data = _furnace(scenes.material(c, roughness=0.3, material_type=F.LX_MATERIAL_SPECULAR))
    ref, hit = _hit_mask(oracle, data)
    img, _ = ref.render(1, 1, shading_model=F.LX_SHADING_LEGACY, threads=8)
    assert np.array_equal(img[hit], np.broadcast_to(np.float32(c), (hit.sum(), 3)))
    assert np.all(img[~hit] == 1.0)


def test_legacy_glass_furnace_mean_is_one(oracle):
    """Clear glass (T = 0) in a white furnace: every path that leaves
    carries (1/eta^2)(eta^2) = 1; Russian roulette keeps the mean."""
    data = _furnace(scenes.material((1, 1, 1), roughness=0.0, ior=1.5, material_type=F.LX_MATERIAL_GLASS),
                    flat=False, width=64, height=64, level=3)
    data.trace_depth = 16
    ref, hit = _hit_mask(oracle, data)
    img, _ = ref.render(1, 8, shading_model=F.LX_SHADING_LEGACY, threads=8)
    v = img[hit][:, 0] / 8
    assert np.mean(np.abs(v - 1.0) < 1e-4) > 0.5  # most pixels: every path left after two refractions / a reflection
    se = np.std(v) / math.sqrt(len(v))
    assert abs(np.mean(v) - 1.0) < 4 * se + 2e-3


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6])
def test_debug_views_closed_forms(oracle, mode):
    col = (0.3, 0.6, 0.9)
    mat = scenes.material(col, roughness=0.35, metallic=0.75, alpha=0.4, alpha_mode=F.LX_ALPHA_BLEND)
    data = _furnace(mat)
    data.env = np.full((16, 32, 4), 0.5, np.float32)
    ref, hit = _hit_mask(oracle, data, depth=1)
    img, st = ref.render(1, 1, debug_mode=mode, threads=8, dmis=False)
    assert st["path_bounces"] == data.camera["width"] * data.camera["height"]  # one bounce
    assert np.all(img[~hit] == np.float32(0.5))  # misses: environment, no MIS weight
    h = img[hit]
    want = {1: col, 3: (0.5, 0.5, 1.0), 4: (0.35,) * 3, 5: (0.75,) * 3}
    if mode in want:
        assert np.array_equal(h, np.broadcast_to(np.float32(want[mode]), h.shape))
    elif mode == 2:  # faceted sphere: every pixel shows the geometric normal of some face
        n = (h.astype(np.float64) - 0.5) / 0.5
        assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=3e-6)
        v, f = scenes.icosphere(2)
        fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
        fn /= np.linalg.norm(fn, axis=1, keepdims=True)
        assert np.all(np.max(n @ fn.T, axis=1) > 1 - 1e-6)  # neighbouring faces differ by ~0.04
    else:  # BLEND alpha factor
        assert np.array_equal(h, np.broadcast_to(np.float32(0.4), h.shape))


def test_debug_alpha_mask_is_binary(oracle):
    data = scenes.material_zoo(96, 54, level=2)
    ref = oracle.OracleScene(data)
    img, _ = ref.render(1, 1, debug_mode=F.LX_DEBUG_ALPHA, threads=8)
    sky = data.env[..., 0]
    # hit pixels of opaque / MASK materials show 1, BLEND pixels 0.45; misses the sky radiance
    vals = np.unique(img[:, 0])
    assert np.any(vals == 1.0) and np.any(vals == np.float32(0.45))
    assert np.all((vals == 1.0) | (vals == np.float32(0.45)) | ((vals >= sky.min()) & (vals <= sky.max())))


def test_gltf_legacy_material_classes(tmp_path):
    """Each clause of isGlass (include/scene/gltf_material_utils.h), the
    emissive-first order and the metallic / roughness gate."""
    clear = {"baseColorFactor": [1, 1, 1, 1], "metallicFactor": 0.0, "roughnessFactor": 0.0}
    mats = [
        ({"pbrMetallicRoughness": clear, "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1}}},
         F.LX_MATERIAL_GLASS),
        ({"pbrMetallicRoughness": clear, "extensions": {"KHR_materials_volume": {}}}, F.LX_MATERIAL_GLASS),
        ({"pbrMetallicRoughness": clear, "extensions": {"KHR_materials_ior": {"ior": 1.3}}}, F.LX_MATERIAL_GLASS),
        ({"pbrMetallicRoughness": clear, "extensions": {"KHR_materials_ior": {"ior": 1.005}}},
         F.LX_MATERIAL_MICROFACET),
        ({"pbrMetallicRoughness": dict(clear, baseColorFactor=[1, 1, 1, 0.5]), "alphaMode": "MASK"},
         F.LX_MATERIAL_GLASS),
        ({"pbrMetallicRoughness": dict(clear, baseColorFactor=[1, 1, 1, 0.995]), "alphaMode": "BLEND"},
         F.LX_MATERIAL_MICROFACET),
        ({"pbrMetallicRoughness": clear,
          "extensions": {"KHR_materials_pbrSpecularGlossiness": {"diffuseFactor": [1, 1, 1, 0.5]}}},
         F.LX_MATERIAL_GLASS),
        ({"pbrMetallicRoughness": dict(clear, roughnessFactor=0.5),
          "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1}}}, F.LX_MATERIAL_MICROFACET),
        ({"pbrMetallicRoughness": clear, "emissiveFactor": [1, 1, 1],
          "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1}}}, F.LX_MATERIAL_EMISSIVE),
        ({"pbrMetallicRoughness": clear, "emissiveFactor": [1, 1, 1],
          "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 0.0}}}, F.LX_MATERIAL_MICROFACET),
        ({}, F.LX_MATERIAL_MICROFACET),
    ]
    p = str(tmp_path / "classes.glb")
    gltf_util.write_glb(p, extra_materials=[m for m, _ in mats])
    got = lx.parse_glb(p)["materials"]
    assert got[0]["material_type"] == F.LX_MATERIAL_MICROFACET  # the default material
    types = [m["material_type"] for m in got[-len(mats):]]  # the file's extra materials come last
    assert types == [t for _, t in mats]
    from oracle import gltf_ref
    ref = gltf_ref.load(p)["materials"]
    assert [m["material_type"] for m in ref] == [m["material_type"] for m in got]

