"""Oracle pins for the display composition and the JNLM denoiser
(oracle/orc_display.cpp), against mathematics independent of its code:

  - tone curves: linear = clamp, Reinhard = c / (1 + c) exactly (same f32
    operations), the ACES fit (Stephen Hill's RRT/ODT fit with its published
    input/output matrices) and AgX (Troy Sobotka's, three.js constants)
    evaluated here in f64 from their published formulas; sRGB encoding from
    IEC 61966-2-1;
  - JNLM: a constant image (zero variance) is a fixed point; primary misses
    pass through untouched; with identical guides and a huge variance every
    neighbour gets weight 1, so a pass is the plain window mean (the border
    shrinks the window; with a constant albedo the regression adds nothing).
A planted error (a dropped term of the AgX polynomial, the variance of the
sum instead of the mean, a missing border check) fails one of these.
"""
import numpy as np
import pytest

import oracle


def test_linear_and_reinhard_curves():
    rng = np.random.default_rng(1)
    c = rng.uniform(0, 4, (256, 3)).astype(np.float32)
    assert np.array_equal(oracle.tonemap(c, "linear"), np.clip(c, 0, 1))
    r = (c / (np.float32(1) + c)).astype(np.float32)
    assert np.array_equal(oracle.tonemap(c, "reinhard"), np.clip(r, 0, 1))


def test_aces_fit_matches_published_form():
    rng = np.random.default_rng(2)
    c = rng.uniform(0, 3, (512, 3))
    a_in = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]])
    a_out = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])
    v = (a_in @ (c / 0.6).T).T
    fit = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    ref = np.clip((a_out @ fit.T).T, 0, 1)
    assert np.abs(oracle.tonemap(c, "aces") - ref).max() < 2e-6


def test_agx_matches_published_form():
    rng = np.random.default_rng(3)
    c = rng.uniform(0, 8, (512, 3))
    to2020 = np.array([[0.6274, 0.3293, 0.0433], [0.0691, 0.9195, 0.0113], [0.0164, 0.0880, 0.8956]])
    to_srgb = np.array([[1.6605, -0.5876, -0.0728], [-0.1246, 1.1329, -0.0083], [-0.0182, -0.1006, 1.1187]])
    inset = np.array([[0.856627153315983, 0.0951212405381588, 0.0482516061458583],
                      [0.137318972929847, 0.761241990602591, 0.101439036467562],
                      [0.11189821299995, 0.0767994186031903, 0.811302368396859]])
    outset = np.array([[1.1271005818144368, -0.11060664309660323, -0.016493938717834573],
                       [-0.1413297634984383, 1.157823702216272, -0.016493938717834257],
                       [-0.14132976349843826, -0.11060664309660294, 1.2519364065950405]])
    v = (inset @ (to2020 @ c.T)).T
    v = np.clip((np.log2(np.maximum(v, 1e-10)) + 12.47393) / 16.5, 0, 1)
    poly = (15.5 * v ** 6 - 40.14 * v ** 5 + 31.96 * v ** 4 - 6.868 * v ** 3 + 0.4298 * v ** 2 + 0.1191 * v
            - 0.00232)
    v = np.maximum((outset @ poly.T).T, 0) ** 2.2
    ref = np.clip((to_srgb @ v.T).T, 0, 1)
    assert np.abs(oracle.tonemap(c, "agx") - ref).max() < 2e-5


def test_srgb_encoding():
    x = np.linspace(0, 1.2, 3001).astype(np.float32)
    c = np.stack([x, x, x], 1)
    ref = np.where(x <= 0.0031308, x * 12.92, 1.055 * np.power(x.astype(np.float64), 1 / 2.4) - 0.055)
    assert np.abs(oracle.srgb_encode(c)[:, 0] - ref).max() < 3e-7


def _planes(h, w, color, n, lum2=None, albedo=(0.5, 0.5, 0.5), normal=(0.0, 1.0, 0.0), hit=True):
    """FP64 sums of n identical samples per pixel (bottom row first)."""
    s = np.zeros((h, w, 13))
    col = np.broadcast_to(np.asarray(color, np.float64), (h, w, 3))
    s[..., 0:3] = col * n
    s[..., 3] = n
    s[..., 4:7] = np.asarray(albedo) * n
    s[..., 7] = n if hit else 0
    s[..., 8:11] = np.asarray(normal) * n
    s[..., 11] = n
    lum = 0.2126 * col[..., 0] + 0.7152 * col[..., 1] + 0.0722 * col[..., 2]
    s[..., 12] = lum * lum * n if lum2 is None else lum2
    return s


def test_jnlm_constant_image_is_fixed_point():
    s = _planes(20, 24, (0.3, 0.2, 0.1), 16)
    lin, enc = oracle.jnlm(s, dict(tone_mapping="linear"))
    comp = oracle.compose(s, dict(tone_mapping="linear"))
    assert np.abs(lin[..., :3] - lin[0, 0, :3]).max() < 1e-6
    assert np.array_equal(enc, comp)


def test_jnlm_misses_pass_through():
    s = _planes(16, 16, (0.4, 0.4, 0.4), 8)
    rng = np.random.default_rng(5)
    s[..., 0:3] += rng.uniform(0, 2, (16, 16, 3))
    s[4:8, 4:8, 8:11] = 0.0  # no primary surface: normal validity 0
    lin0, _ = oracle.jnlm(s, dict(tone_mapping="linear"), iterations=1)
    lin_ref, _ = oracle.jnlm(s, dict(tone_mapping="linear"), iterations=1, search_radius=1)
    # rows are flipped between accumulators (bottom first) and images (top first)
    assert np.array_equal(lin0[8:12, 4:8], lin_ref[8:12, 4:8])
    comp_lin = oracle.jnlm(s, dict(tone_mapping="linear"), strength=1e-6, iterations=1)[0]
    assert np.array_equal(lin0[8:12, 4:8], comp_lin[8:12, 4:8])


def test_jnlm_pass_is_window_mean_under_huge_variance():
    """Identical albedo/normal guides, a huge per-pixel variance: every colour
    distance is negative, so every neighbour weighs exp(0) = 1 and the centre
    weighs max(weight) = 1; the albedo regression has zero feature variance.
    One pass is the mean over the window clipped to the image."""
    h, w, n = 12, 14, 4
    rng = np.random.default_rng(6)
    s = _planes(h, w, (0.0, 0.0, 0.0), n, lum2=1e12)
    col = rng.uniform(0.05, 0.9, (h, w))
    s[..., 0:3] = col[..., None] * n
    lin, _ = oracle.jnlm(s, dict(tone_mapping="linear"), iterations=1, search_radius=2)
    # the composed display-linear input: linear curve of col, sRGB encoded and decoded
    enc = oracle.srgb_encode(np.repeat(np.clip(col, 0, 1)[..., None], 3, 2).reshape(-1, 3))[:, 0].reshape(h, w)
    dec = np.where(enc <= 0.04045, enc / 12.92, ((enc + 0.055) / 1.055) ** 2.4)
    comp = dec[::-1]  # image rows, top first
    ref = np.zeros_like(comp)
    for y in range(h):
        for x in range(w):
            ref[y, x] = comp[max(0, y - 2):y + 3, max(0, x - 2):x + 3].mean()
    assert np.abs(lin[..., 0] - ref).max() < 2e-6
    assert np.abs(lin[..., 3] - 1.0).max() < 1e-6
