"""Pins of the oracle's transmission, environment-sampling, emitter-sampling
and normal-mapping functions by what the mathematics fixes (closed forms,
changes of variables, reciprocity, limits, chi-squared tests against
distributions computed independently in numpy).  Each test names the
reference lines the pinned oracle function restates.

Why not E[1/pdf] = 4 pi for the transmissive lobe: the reference's refraction
sampler (src/shaders/cook_torrance.cu:215-222) reaches only the directions
that some microfacet refracts wo into (at normal incidence into glass of
IOR 1.5, a 48-degree cone), so E[1/pdf] over its samples is that solid
angle, not 2 pi per side; the checks below pin the same functions through
quantities the reference's formulas satisfy exactly.
"""
import math

import numpy as np
import pytest

from luxion_amd import scenes
from oracle import lxo

IOR = 1.5
N_UP = np.array([0.0, 1.0, 0.0])
GREY = [0.8, 0.8, 0.8]


# ----------------------------------------------------------- numpy closed forms
def _fresnel(c, ei, et):
    """Unpolarised dielectric Fresnel reflectance (textbook form)."""
    c = min(max(c, 0.0), 1.0)
    st = ei / et * math.sqrt(max(0.0, 1.0 - c * c))
    if st >= 1.0:
        return 1.0
    ct = math.sqrt(max(0.0, 1.0 - st * st))
    rp = (et * c - ei * ct) / (et * c + ei * ct)
    rs = (ei * c - et * ct) / (ei * c + et * ct)
    return 0.5 * (rp * rp + rs * rs)


def _frame(axis):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    t = np.cross([0.0, 0.0, 1.0], a) if abs(a[2]) < 0.9 else np.cross([1.0, 0.0, 0.0], a)
    t /= np.linalg.norm(t)
    return t, np.cross(a, t), a


def _snell(wo, nf, eta):
    """Refraction of -wo through a surface with normal nf (facing wo), eta =
    eta_incident / eta_transmitted; None under total internal reflection."""
    c = float(np.dot(nf, wo))
    k = 1.0 - eta * eta * (1.0 - c * c)
    if k < 0.0:
        return None
    return -eta * np.asarray(wo, np.float64) + (eta * c - math.sqrt(k)) * np.asarray(nf, np.float64)


def _wo(side, cz):
    return np.array([math.sqrt(1.0 - cz * cz), cz if side == "enter" else -cz, 0.0], np.float32)


def _etas(side):
    return (1.0, IOR) if side == "enter" else (IOR, 1.0)


def _cap_grid(axis, nu, nphi, power):
    """Directions over the whole sphere in polar coordinates about `axis`,
    with the polar angle psi = pi u^power (nodes crowd towards the axis), and
    their solid-angle weights."""
    t, b, a = _frame(axis)
    u = (np.arange(nu) + 0.5) / nu
    psi = math.pi * u ** power
    dpsi = math.pi * power * u ** (power - 1) / nu
    phi = (np.arange(nphi) + 0.5) / nphi * 2.0 * math.pi
    PS, PH = np.meshgrid(psi, phi, indexing="ij")
    d = (np.sin(PS) * np.cos(PH))[..., None] * t + (np.sin(PS) * np.sin(PH))[..., None] * b + \
        np.cos(PS)[..., None] * a
    w = np.sin(PS) * dpsi[:, None] * (2.0 * math.pi / nphi)
    return d.reshape(-1, 3), w.reshape(-1), PS.reshape(-1), PH.reshape(-1)


def _transmission_side(side, d):
    return (d @ N_UP) < 0 if side == "enter" else (d @ N_UP) > 0


def _ggx_grid(alpha, n):
    """Half-vectors about +y from a midpoint grid over the GGX inverse CDF
    (Walter et al. 2007, eq. 35-36): each node carries D(h)(n.h) d omega_h
    mass 1 / n^2."""
    x = (np.arange(n) + 0.5) / n
    X1, X2 = np.meshgrid(x, x, indexing="ij")
    th = np.arctan(alpha * np.sqrt(X1) / np.sqrt(1.0 - X1))
    ph = 2.0 * math.pi * X2
    t, b, a = _frame(N_UP)
    h = (np.sin(th) * np.cos(ph))[..., None] * t + (np.sin(th) * np.sin(ph))[..., None] * b + \
        np.cos(th)[..., None] * a
    return h.reshape(-1, 3)


def _preimage_mass(side, wo, roughness, n=1000):
    """Integral of D(h)(n.h) over half-vectors h (n.h > 0), each counted once
    per direction wi on the transmission side whose generalised half-vector
    -(eta_o wo + eta_i wi)/|.| is h.  For eta_i wi = -eta_o wo - lam h with
    |wi| = 1, lam solves lam^2 + 2 lam eta_o (wo.h) + eta_o^2 - eta_i^2 = 0
    (lam > 0).  By the change of variables wi -> h this is the integral of
    the transmission PDF over the transmission side."""
    alpha = max(min(roughness, 1.0), 1e-4) ** 2
    eo, ei = _etas(side)
    h = _ggx_grid(alpha, n)
    w = np.asarray(wo, np.float64)
    c = h @ w
    disc = eo * eo * c * c - eo * eo + ei * ei
    root = np.sqrt(np.maximum(disc, 0.0))
    count = np.zeros(len(h))
    for sgn in (1.0, -1.0):
        lam = -eo * c + sgn * root
        wi = -(eo * w + lam[:, None] * h) / ei
        count += (disc >= 0) & (lam > 0) & _transmission_side(side, wi) & ((wi @ N_UP) != 0)
    return count.mean()


# ------------------------------------------------------------ transmission
@pytest.mark.parametrize("side", ["enter", "exit"])
@pytest.mark.parametrize("cz", [0.95, 0.8])
@pytest.mark.parametrize("roughness", [0.1, 0.4, 0.8])
def test_transmission_pdf_is_the_refraction_change_of_variables(oracle, side, cz, roughness):
    """CookTorrance::PDF on the refraction side (cook_torrance.cu:260-295 ->
    PDF_GGX_Transmission :239-258, lobe probability :275-284) integrates to
    probRefract times the GGX mass of the half-vectors that some transmitted
    direction maps to (closed-form preimage count above): exactly
    probRefract when entering glass, less when leaving it (total internal
    reflection).  probRefract = T (1 - m)(1 - F) with the textbook Fresnel.
    A wrong Jacobian (eta_o^2 in place of eta_i^2, a missing |wi.h|), a
    half-vector built with swapped etas or D taken about the facing normal
    changes the integral."""
    wo = _wo(side, cz)
    eo, ei = _etas(side)
    nf = N_UP if side == "enter" else -N_UP
    axis = _snell(wo, nf, eo / ei)
    d, w, _, _ = _cap_grid(axis, 2000, 1000, 3)
    keep = _transmission_side(side, d)
    pdf = lxo.ct_eval(0, wo, d[keep], N_UP, roughness, 0.0, GREY, IOR, 1.0).astype(np.float64)
    integral = float((pdf * w[keep]).sum())
    p_refr = 1.0 - _fresnel(cz, eo, ei)
    mass = _preimage_mass(side, wo, roughness)
    if side == "enter":
        assert mass == 1.0
    assert abs(integral / p_refr - mass) < 2e-3, (integral / p_refr, mass)


@pytest.mark.parametrize("side", ["enter", "exit"])
@pytest.mark.parametrize("cz", [0.95, 0.8, 0.6])
def test_transmission_sampler_snell_limit_and_lobe_probability(oracle, side, cz):
    """sampleCookTorrance (cook_torrance.cu:156-224) at roughness 1e-4
    (alpha 1e-8): every transmitted sample is the Snell refraction of -wo
    (glm::refract with eta_I / eta_T, :215), the transmitted fraction is
    T (1 - m)(1 - F(cos)) within 5 sigma (:177-179), and past the critical
    angle (leaving glass at cos 0.6) nothing is transmitted."""
    wo = _wo(side, cz)
    eo, ei = _etas(side)
    nf = N_UP if side == "enter" else -N_UP
    n = 20000
    d = lxo.ct_sample(wo, N_UP, 1e-4, 0.0, GREY, IOR, 1.0, n, iter=3, depth=2)
    tr = _transmission_side(side, d) & ((d * d).sum(1) > 0)
    p = 1.0 - _fresnel(cz, eo, ei)
    s = _snell(wo, nf, eo / ei)
    if s is None:
        assert p == 0.0 and not tr.any()
        return
    assert np.abs(d[tr] - s).max() < 1e-5
    assert abs(tr.mean() - p) < 5.0 * math.sqrt(p * (1 - p) / n) + 1e-4


@pytest.mark.parametrize("side", ["enter", "exit"])
def test_transmission_samples_follow_the_pdf(oracle, side):
    """Chi-squared of 400,000 sampleCookTorrance directions (roughness 0.4,
    wo at cos 0.9, T = 1) in 48 bins about the Snell direction against
    N * integral of CookTorrance::PDF over each bin (quadrature), on the
    transmission side: the sampled refraction lobe has the density the PDF
    reports.  The bins cover the lobe; the remaining transmission side is one
    more bin.  The expected counts take the PDF only where the sampler can
    put a direction: wi such that glm::refract(-wo, h, eta_I / eta_T) (the
    textbook formula, evaluated in numpy) returns wi again for the
    half-vector h = +-normalize(eta_I wo + eta_T wi) on the sampled side
    (leaving glass, the PDF also has mass at pairs that no half-vector
    refracts into)."""
    roughness, cz, n = 0.4, 0.9, 400000
    wo = _wo(side, cz)
    eo, ei = _etas(side)
    nf = N_UP if side == "enter" else -N_UP
    axis = _snell(wo, nf, eo / ei)
    t, b, a = _frame(axis)
    psi_edges = np.array([0.0, 0.05, 0.1, 0.15, 0.2, 0.3, 0.45, 0.7])
    nphi = 6

    def bin_of(dirs):
        dl = np.stack([dirs @ t, dirs @ b, dirs @ a], 1)
        psi = np.arccos(np.clip(dl[:, 2], -1, 1))
        phi = np.mod(np.arctan2(dl[:, 1], dl[:, 0]), 2 * math.pi)
        k = np.searchsorted(psi_edges, psi, side="right") - 1
        k = np.where(psi >= psi_edges[-1], len(psi_edges) - 1, k)
        j = np.minimum((phi / (2 * math.pi) * nphi).astype(int), nphi - 1)
        return np.where(k == len(psi_edges) - 1, (len(psi_edges) - 1) * nphi, k * nphi + j)

    nbins = (len(psi_edges) - 1) * nphi + 1
    d, w, _, _ = _cap_grid(axis, 2000, 1200, 3)
    w64 = wo.astype(np.float64)
    h = eo * w64 + ei * d
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    h *= np.sign(h @ nf)[:, None]
    c = -(h @ w64)  # dot(N, I) with I = -wo
    eta = eo / ei
    k = 1.0 - eta * eta * (1.0 - c * c)
    back = -eta * w64 - (eta * c + np.sqrt(np.maximum(k, 0.0)))[:, None] * h
    reach = (k >= 0) & (np.abs(back - d).max(1) < 1e-6)
    keep = _transmission_side(side, d) & reach
    pdf = lxo.ct_eval(0, wo, d[keep], N_UP, roughness, 0.0, GREY, IOR, 1.0).astype(np.float64)
    expect = np.bincount(bin_of(d[keep]), weights=pdf * w[keep], minlength=nbins) * n
    s = lxo.ct_sample(wo, N_UP, roughness, 0.0, GREY, IOR, 1.0, n, iter=5, depth=3)
    tr = _transmission_side(side, s) & ((s * s).sum(1) > 0)
    got = np.bincount(bin_of(s[tr].astype(np.float64)), minlength=nbins)
    use = expect > 20
    chi2 = float((((got - expect) ** 2) / expect)[use].sum())
    dof = int(use.sum()) - 1
    assert dof > 30
    assert chi2 < dof + 6.0 * math.sqrt(2.0 * dof), (chi2, dof)


@pytest.mark.parametrize("roughness", [0.2, 0.5, 0.9])
def test_btdf_generalised_reciprocity(oracle, roughness):
    """BTDF (cook_torrance.cu:112-142) is Walter et al.'s microfacet BTDF, which
    satisfies f_t(wo, wi) eta_wi^2 = f_t(wi, wo) eta_wo^2 (the eta_I^2
    numerator of :123 belongs to wo's side; the Fresnel factor of :132 is
    reciprocal across the interface).  Random pairs across the surface: the
    median relative difference is fp32 rounding; a few pairs at the
    visibility and total-reflection borders round apart."""
    rng = np.random.default_rng(int(roughness * 100))
    k = 20000
    a = rng.normal(size=(k, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    a[:, 1] = np.abs(a[:, 1])
    b = rng.normal(size=(k, 3))
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    b[:, 1] = -np.abs(b[:, 1])
    a, b = a.astype(np.float32), b.astype(np.float32)
    lhs = np.array([lxo.ct_eval(1, a[i], b[i:i + 1], N_UP, roughness, 0.0, [1, 1, 1], IOR, 1.0)[0, 0]
                    for i in range(k)], np.float64) * IOR ** 2  # wo outside (eta 1), wi inside (eta 1.5)
    rhs = np.array([lxo.ct_eval(1, b[i], a[i:i + 1], N_UP, roughness, 0.0, [1, 1, 1], IOR, 1.0)[0, 0]
                    for i in range(k)], np.float64)
    big = np.maximum(lhs, rhs)
    both = (lhs > 1e-4 * big.max()) & (rhs > 1e-4 * big.max())
    assert both.sum() > 500
    rel = np.abs(lhs - rhs)[both] / big[both]
    assert np.median(rel) < 1e-5 and np.quantile(rel, 0.95) < 1e-3, (np.median(rel), np.quantile(rel, 0.95))


@pytest.mark.parametrize("side", ["enter", "exit"])
@pytest.mark.parametrize("cz", [0.95, 0.8])
@pytest.mark.parametrize("roughness,tol", [(0.05, 0.003), (0.1, 0.012)])
def test_btdf_smooth_limit_energy(oracle, side, cz, roughness, tol):
    """As roughness -> 0 the BTDF (cook_torrance.cu:112-142) tends to a
    specular transmission whose cosine-weighted integral over the
    transmission side is (eta_o / eta_i)^2 (1 - F(cos theta_o)) (radiance
    scales by the squared IOR ratio across the interface; Walter et al.
    2007, eq. 21 with D -> delta, G -> 1), times T (1 - m) albedo.  Quadrature
    about the Snell direction; the bound allows for roughness 0.05 / 0.1."""
    wo = _wo(side, cz)
    eo, ei = _etas(side)
    nf = N_UP if side == "enter" else -N_UP
    axis = _snell(wo, nf, eo / ei)
    d, w, _, _ = _cap_grid(axis, 4000, 800, 4)
    keep = _transmission_side(side, d)
    f = lxo.ct_eval(1, wo, d[keep], N_UP, roughness, 0.0, [1, 1, 1], IOR, 1.0)[:, 0].astype(np.float64)
    got = float((f * np.abs(d[keep] @ N_UP) * w[keep]).sum())
    want = (eo / ei) ** 2 * (1.0 - _fresnel(cz, eo, ei))
    assert abs(got / want - 1.0) < tol, (got, want)


# ----------------------------------------------------------- environment map
def _env_scene(W=64, H=32, seed=5):
    """A small scene whose environment map has varying colour AND varying
    alpha (0.1 .. 3), so weights built from |rgb| and from |rgba| differ."""
    data = scenes.icosphere_hdri(32, 18, level=1, pieces=2, env_size=(W, H))
    rng = np.random.default_rng(seed)
    env = np.empty((H, W, 4), np.float32)
    env[..., 0] = rng.uniform(0.0, 2.0, (H, W))
    env[..., 1] = rng.uniform(0.0, 1.0, (H, W))
    env[..., 2] = rng.uniform(0.0, 0.5, (H, W))
    env[..., 3] = rng.uniform(0.1, 3.0, (H, W))
    env[3, :, :] = 0.0  # an all-zero row: uniform conditional
    env[10, 17, :3] = 40.0  # a hot texel
    data.env = env
    return data


def _env_weights(env):
    """Texel weights |rgba| sin(theta_y), theta_y = pi (y + 0.5) / H (float64)."""
    H = env.shape[0]
    th = math.pi * (np.arange(H) + 0.5) / H
    return np.linalg.norm(env.astype(np.float64), axis=2) * np.sin(th)[:, None]


def test_env_cdfs_use_vec4_texel_length(oracle):
    """precompute_hdri_emission (scene_precompute.cpp:76-144): conditional CDF
    per row over length(vec4 texel) sin(theta) (alpha included, :88-89),
    uniform (x + 1) / W for a zero row, last entry 1; marginal over row
    totals.  Against a float64 numpy construction (FP64 CDF mode)."""
    data = _env_scene()
    H, W = data.env.shape[:2]
    o = lxo.OracleScene(data)
    marg, cond = o.env_cdfs()
    wgt = _env_weights(data.env)
    rows = wgt.sum(1)
    want_c = np.cumsum(wgt, 1) / np.where(rows > 0, rows, 1.0)[:, None]
    want_c[rows == 0] = (np.arange(W) + 1.0) / W
    want_c[:, -1] = 1.0
    want_m = np.cumsum(rows) / rows.sum()
    want_m[-1] = 1.0
    np.testing.assert_allclose(cond.reshape(H, W), want_c, rtol=0, atol=2e-7)
    np.testing.assert_allclose(marg, want_m, rtol=0, atol=2e-7)
    rgb_only = np.linalg.norm(data.env[..., :3].astype(np.float64), axis=2)
    alt = np.cumsum(rgb_only * np.sin(math.pi * (np.arange(H) + 0.5) / H)[:, None], 1)
    alt /= alt[:, -1:].clip(1e-30)
    assert np.abs(alt - want_c)[rows > 0].max() > 1e-2  # the alpha term is visible in this map


def test_envmap_pdf_integrates_to_one(oracle):
    """envmap_pdf (shade_path.cu:27-42) over the sphere: 12 x 12 midpoint
    nodes per texel in (theta, phi), d omega = sin(theta) d theta d phi.  Per
    texel the integrand is p_marginal p_conditional / (2 pi^2) so the exact
    integral is 1; a missing factor H or W, an off-by-one texel difference or
    a wrong sin(theta) fails."""
    data = _env_scene()
    H, W = data.env.shape[:2]
    o = lxo.OracleScene(data)
    k = 12
    th = (np.arange(H * k) + 0.5) / (H * k) * math.pi
    ph = (np.arange(W * k) + 0.5) / (W * k) * 2.0 * math.pi - math.pi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(TH) * np.cos(PH), np.cos(TH), np.sin(TH) * np.sin(PH)], -1).reshape(-1, 3)
    p = o.envmap_pdf(d).astype(np.float64)
    integral = float((p * np.sin(TH).reshape(-1)).sum() * (math.pi / (H * k)) * (2.0 * math.pi / (W * k)))
    assert abs(integral - 1.0) < 2e-4, integral


def test_envmap_pdf_at_texel_centres_matches_texel_probability(oracle):
    """At a texel centre, envmap_pdf = P(row, col) W H / (2 pi^2 sin theta),
    where P is the texel's share of all |rgba| sin(theta) weight (computed in
    numpy): the density that sampleDirectLight's texel choice
    (sample_direct_light.cu:68-100) induces over the texel's solid angle."""
    data = _env_scene()
    H, W = data.env.shape[:2]
    o = lxo.OracleScene(data)
    wgt = _env_weights(data.env)
    P = wgt / wgt.sum()
    rows = wgt.sum(1)
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    th = math.pi * (yy + 0.5) / H
    ph = 2.0 * math.pi * (xx + 0.5) / W - math.pi
    d = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], -1).reshape(-1, 3)
    got = o.envmap_pdf(d).reshape(H, W).astype(np.float64)
    want = P * W * H / (2.0 * math.pi ** 2 * np.sin(th))
    live = rows > 0
    np.testing.assert_allclose(got[live], want[live], rtol=2e-4, atol=1e-6)
    assert np.all(got[~live] == 0.0)


def test_env_samples_chi2_against_vec4_weights(oracle):
    """sampleDirectLight's environment branch (sample_direct_light.cu:68-100):
    400,000 directions, each at a texel centre, counted per texel; chi-squared
    against N P(row, col) from |rgba| sin(theta) (numpy).  Directions are
    exactly the texel centres."""
    data = _env_scene()
    H, W = data.env.shape[:2]
    o = lxo.OracleScene(data)
    n = 400000
    d = o.sample_env(n, iter=2, depth=1).astype(np.float64)
    theta = np.arccos(np.clip(d[:, 1], -1, 1))
    phi = np.arctan2(d[:, 2], d[:, 0])
    row = np.floor(theta / math.pi * H).astype(int)
    col = np.floor((phi + math.pi) / (2 * math.pi) * W).astype(int) % W
    np.testing.assert_allclose(theta, math.pi * (row + 0.5) / H, atol=2e-6)
    dphi = np.angle(np.exp(1j * (phi - (2.0 * math.pi * (col + 0.5) / W - math.pi))))
    assert np.abs(dphi).max() < 2e-6
    got = np.bincount(row * W + col, minlength=H * W)
    wgt = _env_weights(data.env)
    expect = (wgt / wgt.sum()).reshape(-1) * n
    assert got[expect == 0].sum() == 0
    use = expect > 10
    chi2 = float((((got - expect) ** 2)[use] / expect[use]).sum())
    dof = int(use.sum()) - 1
    assert chi2 < dof + 6.0 * math.sqrt(2.0 * dof), (chi2, dof)


# ------------------------------------------------------------ emitter points
def test_light_points_chi2_by_world_area_and_uniform_in_triangle(oracle):
    """sampleDirectLight's emitter branch (sample_direct_light.cu:32-66) over
    the area CDFs of precompute_emissive_mesh_area (scene_precompute.cpp:22-74):
    three emissive instances (scaled and rotated transforms, triangles of very
    different areas) and a non-emissive one.  Each sampled point lies in one
    emissive world-space triangle; per-triangle counts follow world area
    (chi-squared), and inside each triangle the four midpoint sub-triangles
    each take a quarter (uniform barycentric sampling u = 1 - sqrt(r1),
    v = r2 sqrt(r1))."""
    s = scenes.SceneData()
    s.name = "light_pins"
    tri_a = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [2, 0, 0], [4, 0, 0], [2, 0, 3]], np.float32)
    ia = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    tri_b = np.array([[0, 0, 0], [0.2, 0, 0], [0, 0, 0.3], [1, 0, 1], [1.1, 0, 1], [1, 0, 1.05], [3, 0, 3],
                      [5, 0, 3], [3, 0, 6]], np.float32)
    ib = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]], np.uint32)
    ma = s.add_mesh(tri_a, ia)
    mb = s.add_mesh(tri_b, ib)
    lit = s.add_material(scenes.material(emission=(1.0, 0.9, 0.8), strength=5.0))
    dark = s.add_material(scenes.material())
    xf = [scenes.affine((0, 2, 0), rot_y=0.3, scale=(1.0, 1.0, 2.0)),
          scenes.affine((10, 0, 0), rot_x=0.7, scale=(3.0, 1.0, 1.0)),
          scenes.affine((-8, 1, 4), rot_z=1.1)]
    s.add_instance(ma, lit, xf[0])
    s.add_instance(mb, dark, scenes.IDENTITY)
    s.add_instance(mb, lit, xf[1])
    s.add_instance(ma, lit, xf[2])
    s.camera = scenes.look_at_camera((0, 5, 10), (0, 0, 0), (0, 1, 0), 40, 16, 16)
    o = lxo.OracleScene(s)
    tris = []
    for (mesh, mat, m) in s.instances:
        if mat != lit:
            continue
        M = np.asarray(m, np.float64).reshape(3, 4)
        P = s.meshes[mesh]["pos"].astype(np.float64) @ M[:, :3].T + M[:, 3]
        for t in s.meshes[mesh]["idx"]:
            tris.append(P[t])
    tris = np.array(tris)
    area = 0.5 * np.linalg.norm(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]), axis=1)
    n = 300000
    pts = o.sample_light(n, iter=4, depth=2).astype(np.float64)
    # barycentric coordinates of every point in every triangle (least squares
    # on the triangle's plane); the owner is the one with in-plane distance ~0
    owner = np.full(n, -1)
    bary = np.zeros((n, 2))
    for k, T in enumerate(tris):
        e1, e2 = T[1] - T[0], T[2] - T[0]
        G = np.array([[e1 @ e1, e1 @ e2], [e1 @ e2, e2 @ e2]])
        r = pts - T[0]
        rhs = np.stack([r @ e1, r @ e2], 1)
        uv = np.linalg.solve(G, rhs.T).T
        resid = np.linalg.norm(r - uv[:, :1] * e1 - uv[:, 1:] * e2, axis=1)
        inside = (resid < 1e-4 * (1 + np.linalg.norm(T[0]))) & (uv >= -1e-4).all(1) & (uv.sum(1) <= 1 + 1e-4)
        sel = inside & (owner < 0)
        owner[sel] = k
        bary[sel] = uv[sel]
    assert (owner >= 0).all()
    got = np.bincount(owner, minlength=len(tris))
    expect = area / area.sum() * n
    chi2 = float(((got - expect) ** 2 / expect).sum())
    dof = len(tris) - 1
    assert chi2 < dof + 6.0 * math.sqrt(2.0 * dof) + 10.0, (chi2, got, expect)
    # midpoint sub-triangles: corner at vertex 1, corner at vertex 2, corner
    # at vertex 0, and the middle one
    b1, b2 = bary[:, 0], bary[:, 1]
    sub = np.where(b1 >= 0.5, 0, np.where(b2 >= 0.5, 1, np.where(b1 + b2 <= 0.5, 2, 3)))
    q = np.bincount(sub, minlength=4)
    chi2q = float(((q - n / 4) ** 2 / (n / 4)).sum())
    assert chi2q < 3 + 6.0 * math.sqrt(6.0), q


# -------------------------------------------------------------- normal maps
def test_get_normal_identities(oracle):
    """get_normal (material_queries.cu:33-89) with RGBA32F normal maps (1 x 1,
    so the bilinear wrap lookup returns the texel exactly):
      flat texel (0.5, 0.5, 1)        -> N (TBN (0, 0, 1) = N);
      tilted texel (0.5 + s/2, 0.5, 0.5 + c/2), s^2 + c^2 = 1
                                       -> normalize(s T' + c N), T' = the
                                          tangent made orthogonal to N;
      zero texel (0.5, 0.5, 0.5)       -> N (local length guard);
      texel (0.5, 0.5, 0), mapped -N   -> N (facing guard, dot > 1e-4);
      tangent parallel to N            -> N (tangent guard);
      no normal map                    -> N."""
    s = scenes.SceneData()
    s.name = "normal_pins"
    sang, cang = math.sin(0.6), math.cos(0.6)
    texels = [(0.5, 0.5, 1.0), (0.5 + sang / 2, 0.5, 0.5 + cang / 2), (0.5, 0.5, 0.5), (0.5, 0.5, 0.0)]
    mats = []
    for tx in texels:
        s.textures.append(np.array([[[tx[0], tx[1], tx[2], 1.0]]], np.float32))
        mats.append(s.add_material(scenes.material(normal_tex=len(s.textures) - 1)))
    plain = s.add_material(scenes.material())
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    m0 = s.add_mesh(pos, np.array([[0, 1, 2]], np.uint32), uv=np.array([[0, 0], [1, 0], [0, 1]], np.float32))
    for mt in mats + [plain]:
        s.add_instance(m0, mt)
    s.camera = scenes.look_at_camera((0, 5, 10), (0, 0, 0), (0, 1, 0), 40, 16, 16)
    o = lxo.OracleScene(s)
    N = np.array([0.0, 0.6, 0.8], np.float32)
    T = np.array([1.0, 0.3, -0.1], np.float32)  # not orthogonal to N
    uv = [0.3, 0.7]
    f32 = np.float32
    assert np.allclose(o.get_normal(mats[0], N, T, uv), N, atol=1e-7)
    Tp = T.astype(np.float64) - (T.astype(np.float64) @ N.astype(np.float64)) * N.astype(np.float64)
    Tp /= np.linalg.norm(Tp)
    want = sang * Tp + cang * N.astype(np.float64)
    want /= np.linalg.norm(want)
    np.testing.assert_allclose(o.get_normal(mats[1], N, T, uv), want, atol=2e-6)
    assert np.array_equal(o.get_normal(mats[2], N, T, uv), N)
    assert np.array_equal(o.get_normal(mats[3], N, T, uv), N)
    assert np.array_equal(o.get_normal(mats[1], N, f32(2.5) * N, uv), N)
    assert np.array_equal(o.get_normal(plain, N, T, uv), N)
