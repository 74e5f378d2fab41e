"""GPU: the FP16-operand matrix-core DCT form (CJX_MODE_MFMA16).

The region kernel's DCT8/16/32 run as v_mfma_f32_16x16x16_f16 products with
FP32 accumulation and hi + lo FP16 operands.  Its coefficients are no longer
the oracle's bit for bit, so parity has two parts:

- the transform: the device's FP16 coefficients against the oracle's FP64 DCT
  (jo_forward_dct_f64, dct.h:11-16) of the same sharpened planes with the
  device's own transform map, per transform size, on frames forced to all
  DCT8, all DCT16, all DCT32 and on the rate-distortion mix; the errors are
  those of the FP16 storage (one rounding of the exact value) up to a small
  transform error, bounded below;
- everything after the transform: given the device's coefficients, transform
  map and quant field, the CfL maps, quantized AC and DC and the file equal
  the oracle's bit for bit.

The plain path with a forced size is checked against the oracle bit for bit
(the forcing itself is a test mode of the selection step).
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

import cujpegxl_amd as cj
from cujpegxl_amd.synth import synthetic_nv12

pytestmark = pytest.mark.gpu

FORCES = {8: cj.MODE_FORCE8, 16: cj.MODE_FORCE16, 32: cj.MODE_FORCE32}


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def encode(w, h, d, luma, chroma, mode):
    enc = cj.Encoder(w, h, d, stages=True)
    enc.set_mode(mode)
    nv12 = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    data = enc.encode_device(nv12.ptr, nv12.ptr + w * h, w, w)
    out = dict(data=data, acs=enc.fetch("acs", np.int8).copy(), qf=enc.fetch("quant_field", np.int32).copy(),
               c16=enc.fetch("coeffs", np.uint16).copy(), ac=enc.fetch("ac", np.int16).copy(),
               dc=enc.fetch("dc", np.int32).copy(), ytox=enc.fetch("ytox", np.int8).copy(),
               ytob=enc.fetch("ytob", np.int8).copy())
    nv12.close()
    enc.close()
    return out


def slot_classes(acs: np.ndarray, bw: int, bh: int) -> tuple[np.ndarray, np.ndarray]:
    """Per coefficient slot: the size of its block's owner transform, and
    whether the slot holds one of the owner's (n/8)^2 lowest frequencies."""
    a = acs.reshape(bh, bw)
    size = np.zeros((bh, bw), np.int32)
    q = np.zeros((bh, bw), np.int32)
    for n in (8, 16, 32):
        m = n // 8
        ys, xs = np.nonzero(a == n)
        for dy in range(m):
            for dx in range(m):
                size[ys + dy, xs + dx] = n
                q[ys + dy, xs + dx] = dy * m + dx
    assert (size > 0).all()
    raw = q[..., None] * 64 + np.arange(64)
    n = size[..., None]
    llf = ((raw // n) < n // 8) & ((raw % n) < n // 8)
    size3 = np.broadcast_to(np.repeat(size[..., None], 64, axis=2), (bh, bw, 64)).ravel()
    return np.tile(size3, 3), np.tile(llf.ravel(), 3)


def error_table(jo, sharp, dev, w, h):
    """Max and RMS error of the device's FP16 coefficients against the FP64
    DCT, relative to the largest |coefficient| of that transform size, and
    the same for the FP16 rounding of the exact values (the storage floor)."""
    ref = jo.forward_dct_f64(sharp, dev["acs"])
    got = dev["c16"].view(np.float16).astype(np.float64)
    floor = ref.astype(np.float16).astype(np.float64)
    sizes, llf = slot_classes(dev["acs"], w // 8, h // 8)
    rows = {}
    for n in (8, 16, 32):
        sel = sizes == n
        if not sel.any():
            continue
        scale = np.abs(ref[sel]).max()
        e = np.abs(got[sel] - ref[sel]) / scale
        f = np.abs(floor[sel] - ref[sel]) / scale
        sl = sel & llf
        el = np.abs(got[sl] - ref[sl]) / np.abs(ref[sl]).max()
        fl = np.abs(floor[sl] - ref[sl]) / np.abs(ref[sl]).max()
        # Distance from the correctly rounded FP16 value beyond one FP16 step,
        # relative to the largest coefficient (0 when the device value is the
        # rounding of the exact value or a neighbour of it).
        step = np.spacing(np.abs(floor[sel]).astype(np.float16)).astype(np.float64)
        excess = np.maximum(np.abs(got[sel] - floor[sel]) - step, 0.0) / scale
        rows[n] = dict(blocks_frac=float(sel.mean()), max=float(e.max()), rms=float(np.sqrt((e ** 2).mean())),
                       floor_max=float(f.max()), floor_rms=float(np.sqrt((f ** 2).mean())),
                       llf_max=float(el.max()), llf_floor_max=float(fl.max()),
                       off_rounding=float((got[sel] != floor[sel]).mean()), excess=float(excess.max()))
    return rows


def sharpened(jo, luma, chroma, w, h):
    return jo.gaborish(jo.nv12_to_xyb(luma, chroma, w, h))


@pytest.mark.parametrize("n", [8, 16, 32])
def test_plain_forced_size_matches_oracle(jo, n):
    """The plain path with every region forced to one size: coefficients and
    file equal the oracle's for that transform map."""
    w, h, d = 1024, 576, 1.0
    luma, chroma = synthetic_nv12(w, h, 11)
    say(f"plain forced DCT{n} {w}x{h}")
    dev = encode(w, h, d, luma, chroma, FORCES[n])
    acs = dev["acs"]
    assert (acs[acs > 0] == n).all()
    sharp = sharpened(jo, luma, chroma, w, h)
    want = jo.to_f16(jo.forward_dct(sharp, acs))
    np.testing.assert_array_equal(dev["c16"], want)
    check_downstream(jo, dev, w, h, d)


def check_downstream(jo, dev, w, h, d):
    """CfL maps, quantized values and the file from the device's coefficients,
    transform map and quant field, by the oracle."""
    part, ytox, ytob = jo.cfl(dev["c16"], dev["acs"], w, h)
    np.testing.assert_array_equal(dev["ytox"], ytox)
    np.testing.assert_array_equal(dev["ytob"], ytob)
    ac, dc, _ = jo.quantize(dev["c16"], dev["acs"], ytox, ytob, dev["qf"], w, h, d)
    np.testing.assert_array_equal(dev["ac"], ac, err_msg="quantized AC")
    np.testing.assert_array_equal(dev["dc"], dc, err_msg="quantized DC")
    cal = jo.calibrate(d)
    cs = jo.write_codestream(jo.make_frame(w, h, cal.global_scale, cal.quant_dc, dc, ac, dev["acs"], dev["qf"],
                                           ytox, ytob))
    data = dev["data"]
    assert data[40:] == cs or data[48:] == cs, "file differs from the oracle's codestream of the same values"


@pytest.mark.parametrize("w,h,d,seed", [(3840, 2160, 1.0, 0), (1024, 576, 3.0, 5)])
def test_mfma16_error_per_size_and_exact_downstream(jo, w, h, d, seed):
    luma, chroma = synthetic_nv12(w, h, seed)
    sharp = sharpened(jo, luma, chroma, w, h)
    plain = encode(w, h, d, luma, chroma, 0)
    report = {}
    for label, force in (("all-8", cj.MODE_FORCE8), ("all-16", cj.MODE_FORCE16), ("all-32", cj.MODE_FORCE32),
                         ("mix", 0)):
        say(f"mfma16 {label} {w}x{h} d={d}")
        dev = encode(w, h, d, luma, chroma, cj.MODE_MFMA16 | force)
        check_downstream(jo, dev, w, h, d)
        rows = error_table(jo, sharp, dev, w, h)
        for n, r in rows.items():
            say(f"  {label} DCT{n}: blocks {r['blocks_frac']:.3f} max {r['max']:.2e} (floor {r['floor_max']:.2e}) "
                f"rms {r['rms']:.2e} (floor {r['floor_rms']:.2e}) llf max {r['llf_max']:.2e} "
                f"(floor {r['llf_floor_max']:.2e}) off-rounding {r['off_rounding']:.4f} "
                f"excess {r['excess']:.1e}")
            # Every device value is the FP16 rounding of the exact value or a
            # neighbour of it, up to 2e-6 of the largest coefficient (the
            # transform's own error); the errors stay at the storage floor.
            assert r["excess"] <= 2e-6, (label, n)
            assert r["max"] <= 3.0 * r["floor_max"] + 2e-6, (label, n)
            assert r["rms"] <= 1.2 * r["floor_rms"] + 1e-7, (label, n)
            assert r["llf_max"] <= 3.0 * r["llf_floor_max"] + 2e-6, (label, n)
        report[label] = rows
        if label == "mix":
            # The RD choice on matrix-core coefficients agrees with the plain
            # path's almost everywhere, and the rate stays within 1 %.
            agree = float((dev["acs"] == plain["acs"]).mean())
            say(f"  transform map agreement with the plain path {agree:.4f}; "
                f"bytes {len(dev['data'])} vs {len(plain['data'])}")
            assert agree >= 0.97
            assert abs(len(dev["data"]) - len(plain["data"])) <= 0.01 * len(plain["data"])
    assert set(report["all-8"]) == {8}
    assert 16 in report["all-16"] and 32 in report["all-32"]
