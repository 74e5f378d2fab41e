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


DC_INV_QUANT = (4096.0, 512.0, 256.0)
GAB_W1, GAB_W2 = 0.115169525, 0.061248592  # decoder gaborish, libjxl LoopFilter defaults


def _llf_transform(m: int) -> np.ndarray:
    """T[f][o] of the LLF <-> DC relation DC^T = m T^T L T (vardct_layout.h:88-111)."""
    sc = {2: (1.0, 0.901764195028874394),
          4: (1.0, 0.974886821136879522, 0.901764195028874394, 0.787054918159101335)}[m]
    t = np.zeros((m, m))
    for f in range(m):
        norm = np.sqrt(1.0 / m) if f == 0 else np.sqrt(2.0 / m)
        for o in range(m):
            t[f, o] = norm * np.cos(np.pi * (o + 0.5) * f / m) * sc[f]
    return t


def reconstruct_xyb(jo, dec: dict, gaborish: bool = True) -> np.ndarray:
    """Pixels of a decoded frame (float64 XYB planes): dequantization
    (q * w / (qf * global_scale / 2^16)), colour correlation (X += ytox/84 * Y,
    B += (1 + ytob/84) * Y on AC; B += Y on DC), the lowest frequencies of
    DCT16/32 from the DC image, the inverse of the path's DCT basis, and the
    decoder's 3x3 gaborish.  No edge-preserving filter (the default frame
    header enables one pass of it, so this is an approximation of a
    conforming decoder's output)."""
    w, h = dec["width"], dec["height"]
    bw, bh = w // 8, h // 8
    nblk = bw * bh
    cmw = (bw + 7) // 8
    gsf = dec["global_scale"] / 65536.0
    dc_scale = np.float64(np.float32(gsf) * np.float32(dec["quant_dc"]))
    weights = {n: wt.astype(np.float64) for n, wt in zip((8, 16, 32), jo.dequant_weights())}
    ac = dec["ac"].reshape(3, nblk * 64).astype(np.float64)
    dcq = dec["dc"].reshape(3, nblk).astype(np.float64)
    dcv = np.empty_like(dcq)
    dcv[1] = dcq[1] / (DC_INV_QUANT[1] * dc_scale)
    dcv[0] = dcq[0] / (DC_INV_QUANT[0] * dc_scale)
    dcv[2] = dcq[2] / (DC_INV_QUANT[2] * dc_scale) + dcv[1]
    out = np.zeros((3, h, w))
    acs = dec["acs"]
    for n in (8, 16, 32):
        first = np.flatnonzero(acs == n)
        if first.size == 0:
            continue
        s = n // 8
        bx, by = first % bw, first // bw
        e = np.arange(n * n)
        fx, fy = e // n, e % n
        q = e // 64
        slots = (((by[:, None] + q[None, :] // s) * bw + bx[:, None] + q[None, :] % s) * 64 + e[None, :] % 64)
        inv = 1.0 / (dec["qf"][first].astype(np.float64) * np.float64(np.float32(gsf)))
        tile = (by // 8) * cmw + bx // 8
        kx = dec["ytox"][tile].astype(np.float64) / 84.0
        kb = 1.0 + dec["ytob"][tile].astype(np.float64) / 84.0
        cy = ac[1][slots] * weights[n][1][None, :] * inv[:, None]
        cx = ac[0][slots] * weights[n][0][None, :] * inv[:, None] + kx[:, None] * cy
        cb = ac[2][slots] * weights[n][2][None, :] * inv[:, None] + kb[:, None] * cy
        coef = np.stack([cx, cy, cb])  # [3][blocks][n*n], element e = fx*n + fy
        llf = (fx < s) & (fy < s)
        if s == 1:
            coef[:, :, 0] = dcv[:, first]
        else:
            t = _llf_transform(s)
            ti = np.linalg.inv(t)
            for c in range(3):
                d = np.stack([dcv[c][(by + oy) * bw + bx + ox] for oy in range(s) for ox in range(s)], axis=1)
                d = d.reshape(-1, s, s)  # [blocks][oy][ox]
                lm = ti.T @ (np.transpose(d, (0, 2, 1)) / s) @ ti  # L[fx][fy]
                coef[c][:, llf] = lm.reshape(-1, s * s)[:, (fx[llf] * s + fy[llf])]
        a = np.array(jo.dct_basis(n), np.float64).reshape(n, n)
        ai = np.linalg.inv(a)
        cm = coef.reshape(3, -1, n, n)  # [c][block][fx][fy]
        pix = ai[None, None] @ np.swapaxes(cm, 2, 3) @ ai.T[None, None]  # [c][block][row][col]
        for i, (x0, y0) in enumerate(zip(bx * 8, by * 8)):
            out[:, y0 : y0 + n, x0 : x0 + n] = pix[:, i]
    if gaborish:
        norm = 1.0 + 4.0 * GAB_W1 + 4.0 * GAB_W2
        p = np.pad(out, ((0, 0), (1, 1), (1, 1)), mode="symmetric")
        side = p[:, :-2, 1:-1] + p[:, 2:, 1:-1] + p[:, 1:-1, :-2] + p[:, 1:-1, 2:]
        diag = p[:, :-2, :-2] + p[:, :-2, 2:] + p[:, 2:, :-2] + p[:, 2:, 2:]
        out = (out + GAB_W1 * side + GAB_W2 * diag) / norm
    return out


def psnr(got: np.ndarray, want: np.ndarray) -> float:
    peak = want.max() - want.min()
    mse = np.mean((got - want) ** 2)
    return float(10.0 * np.log10(peak * peak / mse)) if mse > 0 else float("inf")
