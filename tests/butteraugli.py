"""A Butteraugli-structured perceptual distance, written here (test
infrastructure; libjxl's Butteraugli, which the reference's quality tools
call through tools/quality_benchmark/pylibjxl.cc:311-317, is not available
offline, so this is not libjxl's number: parity with it is unpinned).

It follows the published structure of Butteraugli (Alakuijala et al.,
"Guetzli: Perceptually Guided JPEG Encoder", 2017, and the libjxl sources'
description): both images in linear light at an 80-nit intensity target go
through an opsin-dynamics transform (a blurred copy sets a per-pixel
sensitivity, so the response adapts to the local light level), are split
into four frequency bands by Gaussian blurs (low, medium, high, ultra-high),
and the band differences are weighted per channel and masked by the local
high-frequency energy of the images; the diffmap is the square root of the
masked weighted sum, and the distance is its maximum (with a 3-norm beside
it).  The constants are this file's: it is used to compare two encodes of
the same frame at matched settings, never as an absolute score.

Fixed by construction and pinned by tests: 0 for identical images;
increasing with the amplitude of an added pattern; larger for the same
error in a flat area than in a busy one (masking).
"""

from __future__ import annotations

import numpy as np
from scipy.ndimage import gaussian_filter

# JPEG XL's XYB: opsin mixing (xyb_impl.cuh:17-35 in the reference) and bias.
_M = np.array([[0.30, 1.0 - 0.078 - 0.30, 0.078],
               [0.23, 1.0 - 0.078 - 0.23, 0.078],
               [0.24342268924547819, 0.20476744424496821, 1.0 - 0.24342268924547819 - 0.20476744424496821]])
_BIAS = 0.0037930732552754493


def xyb_to_linear_rgb(xyb: np.ndarray) -> np.ndarray:
    """[3][h][w] XYB (the encoder's convention) -> [3][h][w] linear RGB."""
    cb = np.cbrt(_BIAS)
    q0 = xyb[1] + xyb[0] + cb
    q1 = xyb[1] - xyb[0] + cb
    q2 = xyb[2] + cb
    mixed = np.stack([q0 ** 3, q1 ** 3, q2 ** 3]) - _BIAS
    rgb = np.tensordot(np.linalg.inv(_M), mixed, axes=1)
    return np.clip(rgb, 0.0, 1.0)


# Opsin dynamics: sensitivity from a blurred copy (sigma 1.2 px), a
# logarithmic response, then an opponent space.
_SIGMA_SENS = 1.2
_MIX = np.array([[0.2996, 0.6337, 0.0777],
                 [0.2216, 0.6939, 0.0987],
                 [0.0200, 0.0200, 0.2048]])
_MIX_BIAS = np.array([1.7557, 1.7557, 12.2265])


def _gamma(v: np.ndarray) -> np.ndarray:
    return np.float32(19.245) * np.log(v + np.float32(9.9711)) - np.float32(44.274)


def opsin_dynamics(rgb: np.ndarray) -> np.ndarray:
    """[3][h][w] linear RGB in nits -> Butteraugli-style X, Y, B (float32)."""
    rgb = rgb.astype(np.float32)
    mix = _MIX.astype(np.float32)
    bias = _MIX_BIAS.astype(np.float32)[:, None, None]
    blurred = np.stack([gaussian_filter(c, _SIGMA_SENS, mode="nearest") for c in rgb])
    pre = np.maximum(np.tensordot(mix, blurred, axes=1) + bias, np.float32(1e-4))
    sens = np.maximum(_gamma(pre) / pre, np.float32(1e-4))
    cur = (np.tensordot(mix, rgb, axes=1) + bias) * sens
    return np.stack([cur[0] - cur[1], cur[0] + cur[1], cur[2]])


_SIGMA_LF, _SIGMA_MF, _SIGMA_HF = 7.15, 3.22, 1.56


def separate_frequencies(xyb: np.ndarray) -> dict:
    blur = lambda a, s: np.stack([gaussian_filter(c, s, mode="nearest", truncate=3.0) for c in a])
    lf = blur(xyb, _SIGMA_LF)
    rest = xyb - lf
    mf = blur(rest, _SIGMA_MF)
    rest -= mf
    hf = blur(rest, _SIGMA_HF)
    uhf = rest - hf
    return {"lf": lf, "mf": mf, "hf": hf, "uhf": uhf}


# Per band and channel (X, Y, B) weights of the squared differences.
_W = {"lf": (30.0, 3.0, 1.0), "mf": (60.0, 5.5, 0.6), "hf": (40.0, 8.0, 0.0), "uhf": (20.0, 9.0, 0.0)}
_MASK_SIGMA, _MASK_MUL, _MASK_OFFSET = 2.7, 2.5, 0.25


class Reference:
    """The reference image's bands, computed once for many candidates."""

    def __init__(self, rgb0: np.ndarray, nits: float = 80.0):
        self.nits = nits
        self.bands = separate_frequencies(opsin_dynamics(rgb0 * nits))
        self.act = np.abs(self.bands["hf"][1]) + np.abs(self.bands["uhf"][1])

    def diffmap(self, rgb1: np.ndarray) -> np.ndarray:
        """Per-pixel distance of rgb1 ([3][h][w] linear RGB in [0, 1])."""
        p0 = self.bands
        p1 = separate_frequencies(opsin_dynamics(rgb1 * self.nits))
        # Masking from the reference's and the candidate's high-frequency Y energy.
        act = 0.5 * (self.act + np.abs(p1["hf"][1]) + np.abs(p1["uhf"][1]))
        act = gaussian_filter(act, _MASK_SIGMA, mode="nearest", truncate=3.0)
        mask = 1.0 / (_MASK_OFFSET + _MASK_MUL * act)
        ac = np.zeros_like(mask)
        dc = np.zeros_like(mask)
        for band, w in _W.items():
            for c in range(3):
                if w[c] == 0.0:
                    continue
                d = p0[band][c] - p1[band][c]
                d *= d
                d *= np.float32(w[c])
                if band == "lf":
                    dc += d
                else:
                    ac += d
        return np.sqrt(ac * mask + dc * np.sqrt(mask))

    def distance(self, rgb1: np.ndarray) -> tuple[float, float]:
        """(max, 3-norm) of the diffmap."""
        d = self.diffmap(rgb1).astype(np.float64)
        return float(d.max()), float(np.mean(d ** 3) ** (1.0 / 3.0))


def diffmap(rgb0: np.ndarray, rgb1: np.ndarray, nits: float = 80.0) -> np.ndarray:
    return Reference(rgb0, nits).diffmap(rgb1)


def distance(rgb0: np.ndarray, rgb1: np.ndarray) -> tuple[float, float]:
    """(max, 3-norm) of the diffmap of rgb1 against the reference rgb0."""
    return Reference(rgb0).distance(rgb1)
