"""Comparison of a decoded file (oracle.decode) with the frame it came from."""

from __future__ import annotations

import numpy as np

import patterns


def llf_slots(golden, acs: np.ndarray, bw: int, bh: int) -> np.ndarray:
    """Covered-layout slots of the varblocks' lowest frequencies, which the
    codestream carries in DC instead of AC (the reference's natural orders)."""
    mask = np.zeros(bw * bh * 64, bool)
    orders = {n: golden[f"order{n}"] for n in (8, 16, 32)}
    for by in range(bh):
        for bx in range(bw):
            n = int(acs[by * bw + bx])
            if n == 0:
                continue
            s = n // 8
            for e in orders[n][: s * s]:
                mask[patterns.covered_slot(n, bx, by, bw, int(e))] = True
    return mask


def check_decoded(golden, dec, w, h, gs, qdc, acs, qf, dc, ac, ytox, ytob):
    bw, bh = w // 8, h // 8
    assert (dec["width"], dec["height"]) == (w, h)
    assert (dec["global_scale"], dec["quant_dc"]) == (gs, qdc)
    np.testing.assert_array_equal(dec["acs"], acs, err_msg="transform map")
    first = acs != 0
    np.testing.assert_array_equal(dec["qf"][first], np.asarray(qf).ravel()[first], err_msg="quant field")
    np.testing.assert_array_equal(dec["dc"], np.asarray(dc).ravel(), err_msg="DC")
    tiles = ((bw + 7) // 8) * ((bh + 7) // 8)
    zero = np.zeros(tiles, np.int8)
    np.testing.assert_array_equal(dec["ytox"], zero if ytox is None else ytox, err_msg="YtoX")
    np.testing.assert_array_equal(dec["ytob"], zero if ytob is None else ytob, err_msg="YtoB")
    want = np.asarray(ac).reshape(3, -1).astype(np.int16).copy()
    want[:, llf_slots(golden, acs, bw, bh)] = 0
    np.testing.assert_array_equal(dec["ac"], want.ravel(), err_msg="AC")
