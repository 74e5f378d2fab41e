"""GPU: the learned transform-selection / AQ network.

- Training kernels: the device's loss and gradient (cjx_mlp_step) against the
  float64 oracle (oracle/jxl_mlp.c, itself pinned by finite differences) on
  the same parameters and batch; the device forward against the oracle's;
  Adam steps lower the loss.
- Feature export of the plain path: the 96 features of every region against
  a numpy restatement on the oracle's sharpened planes and the oracle's DCT8
  costs, and the labels against the oracle's transform map and quant field.
- Learned mode (CJX_MODE_LEARNED): with an unreachable threshold every region
  falls back to the heuristic (files equal the plain path's); with the
  shipped network, the transform map is the network's where it is confident
  (checked against the oracle's float64 forward on the exported features,
  away from the decision boundary) and everything downstream of the map is
  the oracle's bit for bit.
"""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import cujpegxl_amd as cj
from cujpegxl_amd.synth import synthetic_nv12

pytestmark = pytest.mark.gpu

IN, HID, OUT, LAB = 96, 64, 7, 12


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def random_batch(jo, n, seed):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(jo.MLP_PARAMS) * 0.15).astype(np.float32)
    x = rng.standard_normal((n, IN)).astype(np.float32)
    lab = np.zeros((n, LAB), np.float32)
    lab[:, 0] = rng.integers(0, 2, n)
    lab[:, 1:5] = rng.integers(0, 2, (n, 4))
    lab[:, 5] = rng.standard_normal(n) + 4.0
    lab[:, 6] = rng.integers(0, 2, n)
    lab[:, 7:11] = rng.integers(0, 2, (n, 4))
    return params, x, lab


def device_step(params, x, lab, lam, lr=0.0, steps=1):
    L = cj.stages_lib()
    n = x.shape[0]
    bufs = dict(x=cj.DeviceBuffer.from_array(x), lab=cj.DeviceBuffer.from_array(lab),
                p=cj.DeviceBuffer.from_array(params), m=cj.DeviceBuffer.from_array(np.zeros_like(params)),
                v=cj.DeviceBuffer.from_array(np.zeros_like(params)), g=cj.DeviceBuffer.from_array(np.zeros_like(params)),
                loss=cj.DeviceBuffer(4), part=cj.DeviceBuffer(4 * L.cjx_mlp_partial_floats(n)))
    losses = []
    try:
        for k in range(steps):
            assert L.cjx_mlp_step(bufs["x"].ptr, bufs["lab"].ptr, n, lam, bufs["part"].ptr, bufs["g"].ptr,
                                  bufs["loss"].ptr, bufs["p"].ptr, bufs["m"].ptr, bufs["v"].ptr, k + 1, lr, 0) == cj.OK
            L.cjx_device_synchronize()
            losses.append(float(bufs["loss"].to_host(np.float32)[0]))
        return losses, bufs["g"].to_host(np.float32).copy(), bufs["p"].to_host(np.float32).copy()
    finally:
        for b in bufs.values():
            b.close()


def test_gradient_matches_oracle(jo):
    """Loss and gradient of one batch (n = 1,000, 32 regions per workgroup
    with a ragged last one) against the float64 oracle."""
    params, x, lab = random_batch(jo, 1000, 1)
    lam = 0.25
    say("mlp gradient check, 1000 regions")
    (loss,), grad, _ = device_step(params, x, lab, lam)
    want_loss, want = jo.mlp_loss_grad(params, x, lab, lam)
    assert abs(loss - want_loss) <= 2e-5 * abs(want_loss)
    err = np.abs(grad - want)
    scale = np.abs(want).max()
    say(f"  loss {loss:.6f} vs {want_loss:.6f}; max gradient error {err.max() / scale:.2e} of the largest entry")
    assert err.max() <= 1e-5 * scale
    # Each block of parameters separately, relative to its own size.
    for a, b in ((0, HID * IN), (HID * IN, HID * IN + HID), (HID * IN + HID, HID * IN + HID + OUT * HID),
                 (HID * IN + HID + OUT * HID, jo.MLP_PARAMS)):
        assert err[a:b].max() <= 1e-4 * np.abs(want[a:b]).max() + 1e-7


def test_forward_matches_oracle(jo):
    params, x, _ = random_batch(jo, 700, 2)
    L = cj.stages_lib()
    dp, dx = cj.DeviceBuffer.from_array(params), cj.DeviceBuffer.from_array(x)
    dz = cj.DeviceBuffer(4 * OUT * x.shape[0])
    assert L.cjx_mlp_forward(dp.ptr, dx.ptr, x.shape[0], dz.ptr, 0) == cj.OK
    L.cjx_device_synchronize()
    z = dz.to_host(np.float32).reshape(-1, OUT)
    for b in (dp, dx, dz):
        b.close()
    want = jo.mlp_forward(params, x)
    np.testing.assert_allclose(z, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


def test_adam_steps_lower_the_loss(jo):
    """A learnable problem (labels from a fixed teacher network): 60 Adam
    steps on the device take the loss down by more than half, and the
    oracle agrees on the final parameters' loss."""
    rng = np.random.default_rng(3)
    teacher, x, lab = random_batch(jo, 2048, 4)
    z = jo.mlp_forward(teacher, x)
    lab[:, 0] = (z[:, 1] > z[:, 0]).astype(np.float32)
    lab[:, 1:5] = (z[:, 2:6] > 0).astype(np.float32)
    lab[:, 5] = z[:, 6]
    lab[:, 6] = 1.0
    lab[:, 7:11] = 1.0
    start = (rng.standard_normal(jo.MLP_PARAMS) * 0.05).astype(np.float32)
    losses, _, final = device_step(start, x, lab, 0.1, lr=3e-3, steps=60)
    say(f"  loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.5 * losses[0]
    want, _ = jo.mlp_loss_grad(final, x, lab, 0.1, want_grad=False)
    # The device's last loss is the one before the last step.
    assert want < losses[-2]


def features_numpy(jo, luma, chroma, w, h, d):
    """The region features restated in numpy (float64) from the oracle's
    sharpened planes in the pixel domain (variance, mean squared adjacent
    difference, mean) and the oracle's DCT8 costs and their parts."""
    sharp = jo.gaborish(jo.nv12_to_xyb(luma, chroma, w, h)).astype(np.float64)
    bw, bh = w // 8, h // 8
    rw, rh = (bw + 3) // 4, (bh + 3) // 4
    cost8, dist8, rate8 = (a.reshape(bh, bw).astype(np.float64)
                           for a in jo.block_costs8(sharp[1].astype(np.float32), d, parts=True))
    y = sharp[1].reshape(bh, 8, bw, 8).transpose(0, 2, 1, 3)  # [by][bx][8][8]
    my = y.mean((2, 3))
    vy = np.maximum((y ** 2).mean((2, 3)) - my ** 2, 0)
    gy = ((np.diff(y, axis=3) ** 2).sum((2, 3)) + (np.diff(y, axis=2) ** 2).sum((2, 3))) / 112.0
    f = np.stack([np.log2(1 + 4096 * vy), np.log2(1 + 1024 * gy), np.log2(1 + 64 * np.maximum(cost8, 0)),
                  np.log2(1 + 262144 * dist8), np.log2(1 + rate8), 4 * my], axis=-1)  # [bh][bw][6]
    out = np.zeros((rh, rw, 16, 6))
    for ly in range(4):
        for lx in range(4):
            sub = f[ly::4, lx::4]
            out[: sub.shape[0], : sub.shape[1], ly * 4 + lx] = sub
    return out.reshape(rh * rw, 96)


def device_run(w, h, d, luma, chroma, mode=0, export=False):
    enc = cj.Encoder(w, h, d, stages=True)
    enc.set_mode(mode)
    if export:
        enc.export_features(True)
    buf = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    data = enc.encode_device(buf.ptr, buf.ptr + w * h, w, w)
    out = dict(data=data, acs=enc.fetch("acs", np.int8).copy(), qf=enc.fetch("quant_field", np.int32).copy())
    if export:
        out["rows"] = enc.fetch("features", np.float32).reshape(-1, IN + LAB).copy()
    buf.close()
    enc.close()
    return out


@pytest.mark.parametrize("w,h,d,seed", [(1024, 576, 1.0, 21), (776, 456, 2.0, 22)])
def test_feature_export_matches_restatement(jo, w, h, d, seed):
    luma, chroma = synthetic_nv12(w, h, seed)
    say(f"feature export {w}x{h} d={d}")
    dev = device_run(w, h, d, luma, chroma, export=True)
    rows = dev["rows"]
    want = features_numpy(jo, luma, chroma, w, h, d)
    np.testing.assert_allclose(rows[:, :IN], want, atol=2e-3, rtol=0)
    # Labels: the heuristic's map, the quant field, the masks.
    bw, bh = w // 8, h // 8
    rw, rh = (bw + 3) // 4, (bh + 3) // 4
    acs = dev["acs"].reshape(bh, bw)
    qf = dev["qf"].reshape(bh, bw)
    assert np.array_equal(dev["acs"], jo.select_transforms(jo.gaborish(jo.nv12_to_xyb(luma, chroma, w, h))[1], d))
    for r in range(rw * rh):
        rx, ry = r % rw, r // rw
        lab = rows[r, IN:]
        bx0, by0 = 4 * rx, 4 * ry
        fit32 = bx0 + 4 <= bw and by0 + 4 <= bh
        assert lab[6] == float(fit32)
        assert lab[0] == float(acs[by0, bx0] == 32)
        for q in range(4):
            qx, qy = bx0 + 2 * (q & 1), by0 + 2 * (q >> 1)
            fit16 = qx + 2 <= bw and qy + 2 <= bh
            assert lab[7 + q] == float(fit16)
            if fit16 and acs[by0, bx0] != 32:
                assert lab[1 + q] == float(acs[qy, qx] == 16)
        blk = qf[by0:by0 + 4, bx0:bx0 + 4]
        assert abs(lab[5] - np.log2(blk.astype(np.float64)).mean()) <= 1e-5


def oracle_file_for_map(jo, luma, chroma, w, h, d, acs):
    """The oracle's file for a given transform map (everything else as the
    plain path)."""
    xyb = jo.nv12_to_xyb(luma, chroma, w, h)
    qf = jo.adaptive_quant(xyb, d)
    sharp = jo.gaborish(xyb)
    c16 = jo.to_f16(jo.forward_dct(sharp, acs))
    _, ytox, ytob = jo.cfl(c16, acs, w, h)
    ac, dc, _ = jo.quantize(c16, acs, ytox, ytob, qf, w, h, d)
    cal = jo.calibrate(d)
    return jo.write_codestream(jo.make_frame(w, h, cal.global_scale, cal.quant_dc, dc, ac, acs, qf, ytox, ytob))


def set_weights(params, threshold):
    w1 = params[: HID * IN].astype(np.float16).view(np.uint16)
    b1 = params[HID * IN: HID * IN + HID].astype(np.float32)
    o = HID * IN + HID
    w2 = params[o: o + OUT * HID].astype(np.float16).view(np.uint16)
    b2 = params[o + OUT * HID:].astype(np.float32)
    assert cj.stages_lib().cjx_mlp_set_weights(0, w1.ctypes.data, b1.ctypes.data, w2.ctypes.data, b2.ctypes.data,
                                               threshold) == cj.OK


def test_learned_mode_decisions_and_downstream(jo):
    """Random weights with a small threshold (so that most regions are
    confident): the device's map is the network's on the exported features
    wherever the float64 oracle's logits are clear of the threshold and of
    zero, the heuristic's elsewhere, and the file is the oracle's for that
    map.  An unreachable threshold gives the plain path's file."""
    w, h, d = 1024, 576, 1.0
    luma, chroma = synthetic_nv12(w, h, 23)
    plain = device_run(w, h, d, luma, chroma, export=True)
    rng = np.random.default_rng(9)
    params = (rng.standard_normal(jo.MLP_PARAMS) * 0.1).astype(np.float32)
    try:
        set_weights(params, 1.0e30)
        assert device_run(w, h, d, luma, chroma, mode=cj.MODE_LEARNED)["data"] == plain["data"]
        th = 0.05
        set_weights(params, th)
        say("learned mode 1024x576, random network")
        dev = device_run(w, h, d, luma, chroma, mode=cj.MODE_LEARNED)
    finally:
        assert cj.stages_lib().cjx_mlp_default_weights(0) == cj.OK
    # The region kernel runs FP16 weights: the oracle on the same rounded values.
    pr = params.copy()
    pr[: HID * IN] = pr[: HID * IN].astype(np.float16).astype(np.float32)
    o = HID * IN + HID
    pr[o: o + OUT * HID] = pr[o: o + OUT * HID].astype(np.float16).astype(np.float32)
    z = jo.mlp_forward(pr, plain["rows"][:, :IN])
    bw, bh = w // 8, h // 8
    rw, rh = (bw + 3) // 4, (bh + 3) // 4
    acs = dev["acs"].reshape(bh, bw)
    pacs = plain["acs"].reshape(bh, bw)
    checked = confident = 0
    for r in range(rw * rh):
        bx0, by0 = 4 * (r % rw), 4 * (r // rw)
        lab = plain["rows"][r, IN:]
        fit32 = lab[6] > 0.5
        margins = [abs(z[r, 1] - z[r, 0])] if fit32 else []
        use32 = fit32 and z[r, 1] > z[r, 0]
        if not use32:
            margins += [abs(z[r, 2 + q]) for q in range(4) if lab[7 + q] > 0.5]
        if margins and min(abs(m - th) for m in margins) < 1e-3:
            continue  # too close to the threshold to call
        checked += 1
        ok = all(m >= th for m in margins)
        region = acs[by0:by0 + 4, bx0:bx0 + 4]
        if not ok:
            assert np.array_equal(region, pacs[by0:by0 + 4, bx0:bx0 + 4])
            continue
        confident += 1
        if use32:
            assert region[0, 0] == 32
        else:
            for q in range(4):
                qx, qy = 2 * (q & 1), 2 * (q >> 1)
                if lab[7 + q] > 0.5:
                    assert region[qy, qx] == (16 if z[r, 2 + q] > 0 else 8)
    say(f"  checked {checked} regions, {confident} decided by the network")
    assert checked >= 0.9 * rw * rh and confident >= 0.3 * checked and confident < checked
    want = oracle_file_for_map(jo, luma, chroma, w, h, d, dev["acs"])
    assert dev["data"][40:] == want or dev["data"][48:] == want
