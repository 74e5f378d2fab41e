"""GPU: failure paths and limits of the encoder: a configure whose allocation
fails leaves an encoder that refuses to encode until a configure succeeds,
and the largest frame the C ABI accepts (16384x8192: 2,048 AC groups, whose
table of contents is the head's worst case) encodes to the oracle's bytes."""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import cujpegxl_amd as cj
from cujpegxl_amd.synth import synthetic_nv12

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_failed_configure_leaves_no_stale_buffers(jo):
    w, h, d = 512, 264, 1.0
    say(f"configure {w}x{h} -> 1024x576 with a failing allocation, then recover")
    L = cj.stages_lib()
    luma, chroma = synthetic_nv12(w, h, 1)
    want = jo.encode_nv12(luma, chroma, w, h, d)
    enc = cj.Encoder(w, h, d, stages=True)
    dev = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    assert enc.encode_device(dev.ptr, dev.ptr + w * h, w, w) == want
    big = cj.Config(1024, 576, d, 0)
    try:
        L.cjx_stage_fail_allocation_after(5)  # after the streams and events exist
        assert L.cujpegxl_encoder_configure(enc.handle, C.byref(big)) == cj.INTERNAL_ERROR
    finally:
        L.cjx_stage_fail_allocation_after(0)
    # Unconfigured: every encode is refused (no launch over freed buffers).
    inp = cj.Nv12Input(dev.ptr, dev.ptr + w * h, w, w)
    out = cj.Output(enc.out.ctypes.data, 0, enc.capacity)
    assert L.cujpegxl_encoder_encode(enc.handle, C.byref(inp), C.byref(out)) == cj.INTERNAL_ERROR
    # The same failing shape again really reallocates (no early "same shape"
    # success), and the old shape works again after that.
    l2, c2 = synthetic_nv12(1024, 576, 3)
    dev2 = cj.DeviceBuffer.from_array(np.concatenate([l2.ravel(), c2.ravel()]))
    assert L.cujpegxl_encoder_configure(enc.handle, C.byref(big)) == cj.OK
    inp2 = cj.Nv12Input(dev2.ptr, dev2.ptr + 1024 * 576, 1024, 1024)
    cap = 3 * 128 * 72 * 384 + (1 << 20)
    buf = np.empty(cap, np.uint8)
    out2 = cj.Output(buf.ctypes.data, 0, cap)
    assert L.cujpegxl_encoder_encode(enc.handle, C.byref(inp2), C.byref(out2)) == cj.OK
    assert buf[: out2.size].tobytes() == jo.encode_nv12(l2, c2, 1024, 576, d)
    assert L.cujpegxl_encoder_configure(enc.handle, C.byref(cj.Config(w, h, d, 0))) == cj.OK
    out = cj.Output(enc.out.ctypes.data, 0, enc.capacity)
    assert L.cujpegxl_encoder_encode(enc.handle, C.byref(inp), C.byref(out)) == cj.OK
    assert enc.out[: out.size].tobytes() == want
    dev.close()
    dev2.close()
    enc.close()


def test_largest_frame_table_of_contents(jo):
    """16384x8192 coefficients with one small non-zero per block and channel:
    2,048 AC groups of 10-25 KB each, so every TOC entry takes 16 or 24 bits
    (past the 32,768 bits the head once had room for)."""
    w, h = 16384, 8192
    say(f"encode_coefficients {w}x{h}")
    bw, bh = w // 8, h // 8
    nblk = bw * bh
    rng = np.random.default_rng(5)
    acs = np.full(nblk, 8, np.int8)
    ac = np.zeros((3, nblk, 64), np.int16)
    pos = rng.integers(1, 64, size=(3, nblk))
    val = rng.integers(1, 4, size=(3, nblk)) * rng.choice([-1, 1], size=(3, nblk))
    np.put_along_axis(ac, pos[:, :, None], val[:, :, None].astype(np.int16), axis=2)
    dc = rng.integers(-50, 50, size=3 * nblk).astype(np.int32)
    qf = rng.integers(1, 64, size=nblk).astype(np.int32)
    gs, qdc = 2048, 300
    enc = cj.Encoder(w, h, 1.0, stages=True)
    data = enc.encode_coefficients(ac.ravel(), dc, acs, qf, None, None, gs, qdc)
    groups = enc.fetch("ac_group_bytes", np.uint32)
    assert groups.size == 2048 and groups.min() >= 1024
    enc.close()
    say("oracle codestream 16384x8192")
    cs = jo.write_codestream(jo.make_frame(w, h, gs, qdc, dc, ac.ravel(), acs, qf, None, None))
    assert data[40:] == cs or data[48:] == cs
