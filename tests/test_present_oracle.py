"""Closed-form pins of the presentation oracle (oracle/present.py): the tone
map of shaders/postprocess/tonemap.wgsl:37-58 and the sRGB encode. CPU only."""
import numpy as np

import oracle.present as PR


def test_aces_closed_forms():
    assert PR.aces(0.0) == 0.0
    assert abs(PR.aces(1.0) - 2.54 / 3.16) < 1e-15
    # the fit's limit is 2.51 / 2.43 > 1: saturated
    assert PR.aces(1e6) == 1.0
    x = np.linspace(0.0, 20.0, 2001)
    y = PR.aces(x)
    assert np.all(np.diff(y) >= 0.0) and y[-1] == 1.0


def test_exposure():
    assert abs(PR.exposure(0.0) - 1.0 / 1.2) < 1e-15
    assert abs(PR.exposure(14.0) * 1.2 * 16384.0 - 1.0) < 1e-15


def test_srgb_encode():
    assert PR.srgb8(0.0) == 0 and PR.srgb8(1.0) == 255 and PR.srgb8(2.0) == 255
    # the segments meet at 0.0031308 (12.92 c = 0.04045)
    j = 0.0031308
    assert abs(12.92 * j - (1.055 * j ** (1 / 2.4) - 0.055)) < 1e-6
    assert PR.srgb8(j) == 10
    # 0.5 encodes to 0.735357 -> 187.52 -> 188; linear 0.2158605 is sRGB 128/255
    assert PR.srgb8(0.5) == 188
    assert PR.srgb8(0.2158605) == 128 and PR.srgb8(0.2140) == 127


def test_present_composes_surface_under_clouds():
    h, w = 2, 3
    comp = np.full((h, w, 4), 7.0)
    lum = np.full((h, w, 4), 0.5)
    trn = np.full((h, w, 4), 0.25)
    depth = np.array([[0.0, 0.3, 0.0], [0.5, 0.0, 0.1]])
    lit = np.full((h, w, 4), 4.0)
    ap_rgb = np.full((h, w, 3), 1.0)
    ap_a = np.full((h, w), 0.5)
    hdr, _ = PR.present(comp, lum, trn, depth, lit, ap_rgb, ap_a, 0.0)
    assert np.all(hdr[depth == 0] == 7.0)
    assert np.allclose(hdr[depth > 0], 0.5 + (4.0 * 0.5 + 1.0) * 0.25)
