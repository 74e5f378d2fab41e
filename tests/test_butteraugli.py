"""CPU: the Butteraugli-structured distance (tests/butteraugli.py), pinned by
what its construction fixes: 0 for identical images, growth with the
amplitude of an added pattern, masking (the same error costs more on a flat
background than on a busy one), and the XYB -> linear RGB inverse against
the oracle's forward XYB."""

from __future__ import annotations

import numpy as np

import butteraugli as ba
from cujpegxl_amd.synth import synthetic_nv12


def test_identical_images_score_zero():
    rng = np.random.default_rng(1)
    img = rng.uniform(0.05, 0.9, (3, 64, 80))
    assert ba.distance(img, img) == (0.0, 0.0)


def test_distance_grows_with_the_error():
    rng = np.random.default_rng(2)
    img = np.clip(0.4 + 0.1 * rng.standard_normal((3, 96, 96)), 0, 1)
    pattern = rng.standard_normal((3, 96, 96))
    d = [ba.distance(img, np.clip(img + a * pattern, 0, 1))[0] for a in (0.002, 0.005, 0.01, 0.02)]
    assert all(x < y for x, y in zip(d, d[1:])), d


def test_masking_by_busy_background():
    rng = np.random.default_rng(3)
    flat = np.full((3, 96, 96), 0.4)
    busy = np.clip(0.4 + 0.15 * rng.standard_normal((3, 96, 96)), 0, 1)
    blob = np.zeros((3, 96, 96))
    blob[:, 40:56, 40:56] = 0.02
    assert ba.distance(flat, flat + blob)[0] > 1.5 * ba.distance(busy, busy + blob)[0]


def test_xyb_inverse_against_the_oracle(jo):
    luma, chroma = synthetic_nv12(64, 48, 4)
    xyb = jo.nv12_to_xyb_f64(luma, chroma, 64, 48).astype(np.float64)
    rgb = ba.xyb_to_linear_rgb(xyb)
    # Forward again in float64: opsin mix + bias, cube root, opponent pair.
    m = np.tensordot(ba._M, rgb, axes=1) + ba._BIAS
    q = np.cbrt(m) - np.cbrt(ba._BIAS)
    back = np.stack([0.5 * (q[0] - q[1]), 0.5 * (q[0] + q[1]), q[2]])
    np.testing.assert_allclose(back, xyb, atol=1e-9)
