"""Pins the oracle's OpenPBR lobes that the 1M-triangle benchmark scene uses
(F82-tinted metals, coated diffuse, rough glass) against mathematics that the
oracle does not share:

  - F82-tint metal Fresnel (openPbrBsdf.wgsl.ts:504-514): F(1) = f0, F(0) = 1,
    F(1/7) = tint * Schlick(1/7); the correction term mu (1 - mu)^6 peaks at
    1/7; the curve equals the published F82-tint model (Hoffman, "Generalization
    of Adobe's Fresnel Model", 2023) written here from the paper; its
    hemispherical average (openPbrBsdf.wgsl.ts:780-785) equals 2 int F mu dmu
    taken by quadrature of the oracle's own F;
  - smooth and rough tinted metals in a white furnace: the render equals
    F82(mu) (smooth) and the GGX x F82 directional albedo by quadrature plus
    the average-Fresnel-squared multiple-scattering term (rough);
  - coat (openPbrBsdf.wgsl.ts:787-830): white coat over a white base passes the
    furnace; a coat over a black base reflects the GGX dielectric directional
    albedo by quadrature with the exact Fresnel equations; coat_ior = 1 leaves
    the base unchanged; a smooth coloured coat over a mirror attenuates by
    F + (1 - F) c^(1 / cos theta_t);
  - sample <-> pdf consistency (openPbrBsdf.wgsl.ts:1456-1704): a chi-square
    test of the sampler's directions against the evaluated pdf integrated over
    solid-angle bins, for metal, rough glass (reflection and refraction) and
    coated diffuse, and the pdf's total mass against the acceptance rate.

Each test names the planted mistake it catches.
"""
import math

import numpy as np
import pytest
from scipy import stats

from mutracer_amd import scene as S

MU_EDGE = 1.0 / 7.0


# ------------------------------------------------------------ closed forms
def schlick(f0, mu):
    return f0 + (1.0 - f0) * (1.0 - mu) ** 5


def f82_published(mu, f0, tint):
    """F82-tint: F = r + (1 - r)(1 - mu)^5 - a mu (1 - mu)^6 with a chosen so
    that F(1/7) = tint * Schlick(1/7) (Hoffman 2023, eqs. for the F82 model)."""
    f0, tint = np.asarray(f0, np.float64), np.asarray(tint, np.float64)
    a = schlick(f0, MU_EDGE) * (1.0 - tint) / (MU_EDGE * (1.0 - MU_EDGE) ** 6)
    mu = np.asarray(mu, np.float64)[..., None]
    return np.clip(schlick(f0, mu) - a * mu * (1.0 - mu) ** 6, 0.0, 1.0)


def fresnel_dielectric(c, n):
    """Unpolarized Fresnel reflectance at incidence cosine c for relative index
    n = n_t / n_i (Walter et al. 2007, eq. 22)."""
    g2 = n * n - 1.0 + c * c
    out = np.ones_like(c)
    ok = g2 > 0
    g = np.sqrt(np.where(ok, g2, 0.0))
    a = (g - c) / (g + c)
    b = (c * (g + c) - 1.0) / (c * (g - c) + 1.0)
    out[ok] = (0.5 * a * a * (1.0 + b * b))[ok]
    return out


def ggx_albedo(alpha, mu_v, fresnel, steps=768):
    """Directional albedo int D G1(v) G1(l) F(v.h) / (4 mu_v) dw_l of an
    isotropic GGX reflection lobe; fresnel(v.h) -> (..., k) array."""
    a2 = alpha * alpha
    g1 = lambda m: 2.0 / (1.0 + np.sqrt(1.0 + a2 * (1.0 - m * m) / (m * m)))  # noqa: E731
    mu = (np.arange(steps) + 0.5) / steps
    phi = (np.arange(steps) + 0.5) * 2.0 * math.pi / steps
    M, P = np.meshgrid(mu, phi, indexing="ij")
    sv = math.sqrt(1.0 - mu_v * mu_v)
    lx, ly, lz = np.sqrt(1.0 - M * M) * np.cos(P), np.sqrt(1.0 - M * M) * np.sin(P), M
    hx, hy, hz = lx + sv, ly, lz + mu_v
    hn = np.sqrt(hx * hx + hy * hy + hz * hz)
    hz = hz / hn
    vh = (sv * hx + mu_v * (lz + mu_v)) / hn
    den = 1.0 + hz * hz * (a2 - 1.0)
    D = a2 / (math.pi * den * den)
    w = D * g1(mu_v) * g1(M) / (4.0 * mu_v)
    F = fresnel(vh)
    if F.ndim == 2:
        F = F[..., None]
    return np.sum(w[..., None] * F, axis=(0, 1)) * 2.0 * math.pi / steps ** 2


# ------------------------------------------------------------------ F82 Fresnel
F82_CASES = [((0.9, 0.6, 0.3), (0.95, 0.8, 0.6)),   # the benchmark scene's metal
             ((0.95, 0.9, 0.8), (0.2, 0.5, 0.9)),
             ((0.04, 0.5, 0.99), (1.0, 0.3, 0.7))]


@pytest.mark.parametrize("f0,tint", F82_CASES)
def test_f82_anchor_values(orc, f0, tint):
    """F(1) = f0, F(0) = 1, F(1/7) = tint * Schlick(1/7). A correction term
    missing its mu factor fails F(0); a tint applied at the wrong place fails
    F(1/7)."""
    f0a, ta = np.array(f0), np.array(tint)
    np.testing.assert_allclose(orc.metal_fresnel(1.0, f0, tint)[0], f0a, atol=2e-6)
    np.testing.assert_allclose(orc.metal_fresnel(0.0, f0, tint)[0], 1.0, atol=2e-6)
    want = np.clip(ta * schlick(f0a, MU_EDGE), 0, 1)
    np.testing.assert_allclose(orc.metal_fresnel(MU_EDGE, f0, tint)[0], want, atol=3e-6)


@pytest.mark.parametrize("f0,tint", F82_CASES[:2])
def test_f82_correction_peaks_at_one_seventh(orc, f0, tint):
    """The F82 correction is proportional to mu (1 - mu)^6, whose maximum is at
    mu = 1/7: Schlick - F is largest there. Exponent 5 (peak at 1/6) fails."""
    mus = np.linspace(0.02, 0.4, 381)
    gap = np.array([schlick(np.array(f0), m) - orc.metal_fresnel(m, f0, tint)[0] for m in mus])
    for c in range(3):
        if abs(1.0 - tint[c]) < 1e-6:
            continue
        peak = mus[np.argmax(np.abs(gap[:, c]))]
        assert abs(peak - MU_EDGE) < 0.002, (c, peak)


@pytest.mark.parametrize("f0,tint", F82_CASES)
def test_f82_matches_published_model(orc, f0, tint):
    mus = np.linspace(0.0, 1.0, 201)
    got = np.array([orc.metal_fresnel(m, f0, tint)[0] for m in mus])
    np.testing.assert_allclose(got, f82_published(mus, f0, tint), atol=4e-6)


@pytest.mark.parametrize("f0,tint", F82_CASES[:2])
def test_f82_average_is_hemispherical_mean(orc, f0, tint):
    """metal_average_fresnel = 2 int_0^1 F(mu) mu dmu of the oracle's own F
    (the closed form f0 + (1 - f0)/21 - b/126 comes from the Beta integrals
    2 B(2, 6) = 1/21 and 2 B(3, 7) = 1/126). A wrong 1/126 fails."""
    n = 4000
    mus = (np.arange(n) + 0.5) / n
    F = np.array([orc.metal_fresnel(m, f0, tint)[0] for m in mus], np.float64)
    mean = 2.0 * np.sum(F * mus[:, None], axis=0) / n
    np.testing.assert_allclose(orc.metal_fresnel(0.5, f0, tint)[1], mean, atol=2e-4)


# ------------------------------------------------------------------ furnaces
def _plane_view(orc, material, mu, spp=256, w=2, h=2, bounces=12, candidates=1):
    """Mean radiance of a huge plane under a uniform white environment, seen
    along a narrow cone at view cosine mu."""
    s = S.Scene()
    s.materials = [material]
    p, i, n = S.quad((-500, 0, 500), (500, 0, 500), (500, 0, -500), (-500, 0, -500))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    r = 3.0
    up = (0.0, 0.0, -1.0) if mu > 0.99 else (0.0, 1.0, 0.0)
    s.camera = S.Camera(eye=(0.0, mu * r, math.sqrt(1.0 - mu * mu) * r), target=(0.0, 0.0, 0.0), up=up, fov=0.02)
    env = np.ones((32, 64, 4), np.float32)
    s.env = env
    out = orc.OracleRenderer(s, w, h, bounces=bounces, candidates=candidates, max_samples=spp).render(0, spp)
    return out[..., 0:3].sum(axis=(0, 1)) / out[..., 11].sum()


@pytest.mark.parametrize("mu", [1.0, 0.5, MU_EDGE])
def test_smooth_tinted_metal_furnace(orc, mu):
    """A mirror-smooth F82 metal returns F82(mu) of the published model. The
    early 'tint == 1' return taken for every tint (no correction) is 4-30%
    off at mu = 1/7 and fails."""
    f0, tint = (0.9, 0.6, 0.3), (0.95, 0.8, 0.6)
    m = S.Material(base_color=f0, base_metalness=1.0, specular_color=tint, specular_roughness=0.0)
    got = _plane_view(orc, m, mu, spp=32)
    np.testing.assert_allclose(got, f82_published(mu, f0, tint), rtol=2e-3)


@pytest.mark.parametrize("mu", [0.8, 0.35])
def test_rough_tinted_metal_furnace(orc, mu):
    """Rough tinted metal: the GGX x F82 single-scattering albedo by quadrature
    plus the multiple-scattering term F_avg^2 (1 - E(mu)) the reference adds
    (openPbrBsdf.wgsl.ts:1016-1022,1240-1244), E the white GGX albedo by
    quadrature. Dropping the multiple-scattering term, or F_avg instead of
    F_avg^2, misses by more than the 2% bound."""
    f0, tint, rough = (0.95, 0.9, 0.8), (0.2, 0.5, 0.9), 0.5
    alpha = rough * rough
    single = ggx_albedo(alpha, mu, lambda vh: f82_published(vh, f0, tint))
    white = ggx_albedo(alpha, mu, lambda vh: np.ones_like(vh))[0]
    f0a, ta = np.array(f0), np.array(tint)
    b = schlick(f0a, MU_EDGE) * (1.0 - ta) / (MU_EDGE * (1.0 - MU_EDGE) ** 6)
    favg = np.clip(f0a + (1.0 - f0a) / 21.0 - b / 126.0, 0, 1)
    want = single + favg * favg * (1.0 - white)
    m = S.Material(base_color=f0, base_metalness=1.0, specular_color=tint, specular_roughness=rough)
    got = _plane_view(orc, m, mu, spp=4096, bounces=1)
    np.testing.assert_allclose(got, want, rtol=0.02)


@pytest.mark.parametrize("mu", [0.9, 0.35])
def test_white_coat_over_white_base_furnace(orc, mu):
    """White coat over a white opaque base reflects everything: the coat takes
    E_coat(mu), the base gets the rest. Leaving out the base's (1 - E_coat)
    factor returns 1 + E_coat (about 1.05 at mu = 0.9, 1.1 at 0.35)."""
    m = S.Material(base_color=(1.0, 1.0, 1.0), coat_weight=1.0, coat_color=(1.0, 1.0, 1.0), coat_roughness=0.3,
                   coat_ior=1.5)
    got = _plane_view(orc, m, mu, spp=2048)
    np.testing.assert_allclose(got, 1.0, rtol=0.015)


@pytest.mark.parametrize("mu,rough", [(0.9, 0.3), (0.4, 0.3), (0.6, 0.6)])
def test_coat_over_black_base(orc, mu, rough):
    """Over a black base (no diffuse, no base specular) only the coat lobe
    reflects: its GGX dielectric directional albedo with the exact Fresnel
    equations at n = coat_ior. Using coat_ior instead of 1/coat_ior in the
    coat's Fresnel (total internal reflection), or alpha = roughness, fails."""
    ior = 1.5
    m = S.Material(base_color=(0.0, 0.0, 0.0), specular_weight=0.0, coat_weight=1.0, coat_color=(1.0, 1.0, 1.0),
                   coat_roughness=rough, coat_ior=ior)
    want = ggx_albedo(rough * rough, mu, lambda vh: fresnel_dielectric(vh, ior))[0]
    got = _plane_view(orc, m, mu, spp=4096, bounces=1)
    np.testing.assert_allclose(got, want, rtol=0.03)


def test_coat_with_ior_one_leaves_base_unchanged(orc):
    """An index-matched white coat neither reflects nor tints nor darkens:
    the render equals the uncoated one (same sample streams). A coat that
    still reflects Schlick's F0 = 0.04, or darkens by (1 - k), changes it."""
    base = dict(base_color=(0.6, 0.5, 0.3), specular_roughness=0.4)
    plain = _plane_view(orc, S.Material(**base), 0.6, spp=512)
    coated = _plane_view(orc, S.Material(coat_weight=1.0, coat_ior=1.0, coat_roughness=0.2, **base), 0.6, spp=512)
    np.testing.assert_allclose(coated, plain, rtol=0.01)


@pytest.mark.parametrize("mu", [1.0, 0.6, 0.3])
def test_smooth_coloured_coat_over_mirror(orc, mu):
    """Smooth coat of colour c over a white mirror: R = F + (1 - F) c^(1 / mu_t)
    with mu_t the refracted cosine (both passages, view and light side, each
    c^(1 / (2 mu_t))) and F the exact Fresnel reflectance. Attenuating by
    c^(1 / mu) (unrefracted cosine) or by one passage only fails."""
    ior, c = 1.5, np.array([0.8, 0.5, 0.2])
    m = S.Material(base_color=(1.0, 1.0, 1.0), base_metalness=1.0, specular_color=(1.0, 1.0, 1.0),
                   specular_roughness=0.0, coat_weight=1.0, coat_color=tuple(c), coat_roughness=0.0, coat_ior=ior)
    F = fresnel_dielectric(np.array([mu]), ior)[0]
    mu_t = math.sqrt(1.0 - (1.0 - mu * mu) / (ior * ior))
    want = F + (1.0 - F) * c ** (1.0 / mu_t)
    got = _plane_view(orc, m, mu, spp=64)
    np.testing.assert_allclose(got, want, rtol=4e-3)


# ----------------------------------------------------- sample <-> pdf (chi^2)
CHI2_CASES = {
    "tinted_metal": (S.Material(base_color=(0.9, 0.6, 0.3), base_metalness=1.0, specular_color=(0.95, 0.8, 0.6),
                                specular_roughness=0.5), 0.7),
    "rough_glass": (S.Material(base_color=(1.0, 1.0, 1.0), transmission_weight=1.0, specular_ior=1.5,
                               specular_roughness=0.45), 0.6),
    "coated_diffuse": (S.Material(base_color=(0.7, 0.4, 0.2), coat_weight=0.3, coat_roughness=0.4), 0.5),
    "metal_grazing": (S.Material(base_color=(0.95, 0.9, 0.8), base_metalness=1.0, specular_color=(0.2, 0.5, 0.9),
                                 specular_roughness=0.6), 0.25),
}


def _bin_mass(orc, m, view, nz=24, nphi=32, sub=12):
    """Integral of the evaluated pdf over (cos theta, phi) bins of the sphere."""
    zc = (np.arange(nz * sub) + 0.5) / (nz * sub) * 2.0 - 1.0
    pc = (np.arange(nphi * sub) + 0.5) / (nphi * sub) * 2.0 * math.pi
    Z, P = np.meshgrid(zc, pc, indexing="ij")
    r = np.sqrt(1.0 - Z * Z)
    dirs = np.stack([r * np.cos(P), r * np.sin(P), Z], -1).reshape(-1, 3)
    _, pdf = orc.bsdf_eval(m, view, dirs)
    dw = (2.0 / (nz * sub)) * (2.0 * math.pi / (nphi * sub))
    mass = pdf.astype(np.float64).reshape(nz, sub, nphi, sub).sum(axis=(1, 3)) * dw
    return mass


@pytest.mark.parametrize("case", sorted(CHI2_CASES))
def test_sampler_follows_pdf_chi2(orc, case):
    """Directions drawn by the lobe sampler follow the pdf the evaluator
    returns for them (the pdf the integrator divides by): chi-square over
    768 solid-angle bins, and the pdf's total mass equals the sampler's
    acceptance rate. A VNDF sample with the wrong alpha, a refraction pdf
    missing its Jacobian, or lobe probabilities that do not sum to the pdf's
    mixture weights fail."""
    m, mu = CHI2_CASES[case]
    view = np.array([math.sqrt(1.0 - mu * mu), 0.0, mu])
    n = 400_000
    rng = np.random.default_rng(1234)
    u = rng.random((n, 3), dtype=np.float32)
    u = np.minimum(u, np.float32(0.99999994))
    d, ok = orc.bsdf_sample(m, view, u)
    d = d[ok].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nz, nphi = 24, 32
    zi = np.clip(((d[:, 2] + 1.0) * 0.5 * nz).astype(int), 0, nz - 1)
    phi = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2.0 * math.pi)
    pi_ = np.clip((phi / (2.0 * math.pi) * nphi).astype(int), 0, nphi - 1)
    obs = np.zeros((nz, nphi))
    np.add.at(obs, (zi, pi_), 1.0)
    mass = _bin_mass(orc, m, view, nz, nphi)
    total = mass.sum()
    accept = ok.mean()
    assert abs(total - accept) < 0.01, (total, accept)
    exp = mass * n
    sel = exp >= 10.0
    chi2 = float(np.sum((obs[sel] - exp[sel]) ** 2 / exp[sel]))
    dof = int(sel.sum()) - 1
    p = stats.chi2.sf(chi2, dof)
    assert p > 1e-4, f"{case}: chi2 {chi2:.1f} over {dof} dof (p = {p:.2e})"
    assert obs[~sel].sum() <= max(50.0, 3.0 * exp[~sel].sum()), "samples where the pdf is ~0"
