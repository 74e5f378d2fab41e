"""GPU: the extension entry points (host-memory input, asynchronous encode)
and page-locked outputs deliver the same file as the synchronous
device-input encode, whose bytes equal the oracle's; the synchronous encode
is ordered after the caller's work on the null stream."""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import cujpegxl_amd as cj
from cujpegxl_amd.synth import synthetic_nv12

pytestmark = pytest.mark.gpu

W, H, D, SEED = 640, 384, 1.0, 7


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def frame(jo):
    luma, chroma = synthetic_nv12(W, H, SEED)
    nv12 = np.concatenate([luma.ravel(), chroma.ravel()])
    want = jo.encode_nv12(luma, chroma, W, H, D)
    return luma, chroma, nv12, want


def test_host_input_matches_oracle(frame):
    luma, chroma, nv12, want = frame
    say(f"encode_host {W}x{H} contiguous, padded and pinned planes")
    enc = cj.Encoder(W, H, D)
    assert enc.encode_host(nv12) == want
    # Padded pitches: the pad bytes must not reach the file.
    pad = 72
    lp = np.full((H, W + pad), 255, np.uint8)
    lp[:, :W] = luma
    cp = np.full((H // 2, W + pad), 7, np.uint8)
    cp[:, :W] = chroma.reshape(H // 2, W)
    inp = cj.Nv12HostInput(lp.ctypes.data, cp.ctypes.data, W + pad, W + pad)
    out = cj.Output(enc.out.ctypes.data, 0, enc.capacity)
    assert enc.lib.cujpegxl_encoder_encode_host(enc.handle, C.byref(inp), C.byref(out)) == cj.OK
    assert enc.out[: out.size].tobytes() == want
    # Page-locked planes take the direct-DMA branch.
    pin = cj.PinnedBuffer(nv12.size)
    pin.array[:] = nv12
    inp = cj.Nv12HostInput(pin.ptr, pin.ptr + W * H, W, W)
    out = cj.Output(enc.out.ctypes.data, 0, enc.capacity)
    assert enc.lib.cujpegxl_encoder_encode_host(enc.handle, C.byref(inp), C.byref(out)) == cj.OK
    assert enc.out[: out.size].tobytes() == want
    # Argument checks of the new entry point.
    short = cj.Nv12HostInput(lp.ctypes.data, cp.ctypes.data, W - 8, W + pad)
    assert enc.lib.cujpegxl_encoder_encode_host(enc.handle, C.byref(short), C.byref(out)) == \
        cj.INVALID_ARGUMENT
    pin.close()
    enc.close()


def test_pinned_output_matches_oracle(frame):
    _luma, _chroma, nv12, want = frame
    say("encode into a page-locked output buffer, twice, then one too small")
    enc = cj.Encoder(W, H, D)
    dev = cj.DeviceBuffer.from_array(nv12)
    pin = cj.PinnedBuffer(enc.capacity)
    for _ in range(2):
        pin.array[:64] = 0xAA
        n = enc.encode_into(dev.ptr, dev.ptr + W * H, W, W, pin.ptr, enc.capacity)
        assert pin.array[:n].tobytes() == want
    inp = cj.Nv12Input(dev.ptr, dev.ptr + W * H, W, W)
    out = cj.Output(pin.ptr, 0, 100)
    assert enc.lib.cujpegxl_encoder_encode(enc.handle, C.byref(inp), C.byref(out)) == cj.BUFFER_TOO_SMALL
    assert out.size == len(want)
    # Word-aligned interior addresses (the device writes the file straight
    # into them) and an unaligned one (copied by DMA), with room for exactly
    # the file or for a prefix that ends inside a word: nothing outside
    # [off, off + cap) changes.
    for off, cap in ((8, len(want)), (1, len(want)), (8, 101), (4, len(want) + 3)):
        pin.array[:] = 0x5C
        out = cj.Output(pin.ptr + off, 0, cap)
        status = enc.lib.cujpegxl_encoder_encode(enc.handle, C.byref(inp), C.byref(out))
        assert out.size == len(want)
        if cap >= len(want):
            assert status == cj.OK and pin.array[off: off + len(want)].tobytes() == want
        else:
            assert status == cj.BUFFER_TOO_SMALL
        assert (pin.array[:off] == 0x5C).all() and (pin.array[off + cap: off + cap + 64] == 0x5C).all()
    pin.close()
    dev.close()
    enc.close()


def test_async_encode_then_finish(frame):
    _luma, _chroma, nv12, want = frame
    say("encode_async on the null stream, pending rules, BUFFER_TOO_SMALL from finish")
    enc = cj.Encoder(W, H, D)
    lib = enc.lib
    dev = cj.DeviceBuffer.from_array(nv12)
    out = cj.Output(enc.out.ctypes.data, 0, enc.capacity)
    assert lib.cujpegxl_encoder_finish(enc.handle, C.byref(out)) == cj.INVALID_ARGUMENT
    assert enc.encode_async(dev.ptr, dev.ptr + W * H, W, W, 0) == cj.OK
    assert enc.encode_async(dev.ptr, dev.ptr + W * H, W, W, 0) == cj.INVALID_ARGUMENT
    host = np.ascontiguousarray(nv12)
    hin = cj.Nv12HostInput(host.ctypes.data, host.ctypes.data + W * H, W, W)
    assert lib.cujpegxl_encoder_encode_host(enc.handle, C.byref(hin), C.byref(out)) == cj.INVALID_ARGUMENT
    # Neither the synchronous encode nor a reconfiguration may drop the
    # pending frame: both are refused, and finish still delivers it.
    din = cj.Nv12Input(dev.ptr, dev.ptr + W * H, W, W)
    assert lib.cujpegxl_encoder_encode(enc.handle, C.byref(din), C.byref(out)) == cj.INVALID_ARGUMENT
    for w2, h2 in ((W, H), (W * 2, H)):
        cfg = cj.Config(w2, h2, D, 0)
        assert lib.cujpegxl_encoder_configure(enc.handle, C.byref(cfg)) == cj.INVALID_ARGUMENT
    assert enc.finish() == want
    assert enc.encode_async(dev.ptr, dev.ptr + W * H, W, W, 0) == cj.OK
    small = cj.Output(enc.out.ctypes.data, 0, 32)
    assert lib.cujpegxl_encoder_finish(enc.handle, C.byref(small)) == cj.BUFFER_TOO_SMALL
    assert small.size == len(want)
    assert lib.cujpegxl_encoder_finish(enc.handle, C.byref(out)) == cj.INVALID_ARGUMENT
    # The synchronous entry point still works after the asynchronous ones.
    assert enc.encode_device(dev.ptr, dev.ptr + W * H, W, W) == want
    dev.close()
    enc.close()


def test_async_encode_is_ordered_on_the_caller_stream(frame):
    _luma, _chroma, nv12, want = frame
    say("encode_async on a caller stream: planes written before, cleared after")
    enc = cj.Encoder(W, H, D)
    L = cj.stages_lib()
    src = cj.DeviceBuffer.from_array(nv12)
    planes = cj.DeviceBuffer(nv12.size)
    s = L.cjx_stream_create()
    assert s != 0
    try:
        for _ in range(2):
            assert L.cjx_fill_async(planes.ptr, 0, nv12.size, s) == cj.OK
            assert L.cjx_copy_async(planes.ptr, src.ptr, nv12.size, s) == cj.OK
            assert enc.encode_async(planes.ptr, planes.ptr + W * H, W, W, s) == cj.OK
            assert L.cjx_fill_async(planes.ptr, 0, nv12.size, s) == cj.OK
            assert enc.finish() == want
        assert L.cjx_stream_synchronize(s) == cj.OK
    finally:
        L.cjx_stream_destroy(s)
    planes.close()
    src.close()
    enc.close()


def test_native_repeat_loop_matches_oracle(frame):
    _luma, _chroma, nv12, want = frame
    say("cjx_stage_encode_repeat: 5 encodes into a page-locked and a pageable output")
    enc = cj.Encoder(W, H, D, stages=True)
    dev = cj.DeviceBuffer.from_array(nv12)
    pin = cj.PinnedBuffer(enc.capacity)
    sec, size = enc.encode_repeat(dev.ptr, dev.ptr + W * H, W, W, pin.ptr, enc.capacity, 5)
    assert sec > 0.0 and pin.array[:size].tobytes() == want
    sec, size = enc.encode_repeat(dev.ptr, dev.ptr + W * H, W, W, enc.out.ctypes.data, enc.capacity, 5)
    assert enc.out[:size].tobytes() == want
    # A count below one and a too-small output are refused.
    inp = cj.Nv12Input(dev.ptr, dev.ptr + W * H, W, W)
    out = cj.Output(pin.ptr, 0, enc.capacity)
    s = C.c_double(0.0)
    assert enc.lib.cjx_stage_encode_repeat(enc.handle, C.byref(inp), C.byref(out), 0, C.byref(s)) == \
        cj.INVALID_ARGUMENT
    small = cj.Output(pin.ptr, 0, 64)
    assert enc.lib.cjx_stage_encode_repeat(enc.handle, C.byref(inp), C.byref(small), 3, C.byref(s)) == \
        cj.BUFFER_TOO_SMALL
    pin.close()
    dev.close()
    enc.close()


def test_sync_encode_is_ordered_after_the_null_stream(jo, frame):
    """A caller that clears and fills the planes on the null stream, as
    callers of the reference's default-stream encoder did, then calls the
    synchronous encode: the frame reads the filled planes.  Two frames
    alternate so a read of stale planes would show."""
    _luma, _chroma, nv12, want = frame
    say("null-stream fill then cujpegxl_encoder_encode, two frames alternating")
    luma2, chroma2 = synthetic_nv12(W, H, SEED + 1)
    nv12b = np.concatenate([luma2.ravel(), chroma2.ravel()])
    want2 = jo.encode_nv12(luma2, chroma2, W, H, D)
    L = cj.stages_lib()
    enc = cj.Encoder(W, H, D)
    srcs = [cj.DeviceBuffer.from_array(nv12), cj.DeviceBuffer.from_array(nv12b)]
    planes = cj.DeviceBuffer(nv12.size)
    for i in range(4):
        src, expect = srcs[i % 2], (want, want2)[i % 2]
        assert L.cjx_fill_async(planes.ptr, 0, nv12.size, 0) == cj.OK
        assert L.cjx_copy_async(planes.ptr, src.ptr, nv12.size, 0) == cj.OK
        assert enc.encode_device(planes.ptr, planes.ptr + W * H, W, W) == expect, i
    assert L.cjx_device_synchronize() == cj.OK
    for b in srcs + [planes]:
        b.close()
    enc.close()


def test_public_form_switch_matches_stage_library(frame):
    """cujpegxl_encoder_set_mode (public library) gives the files of the stage
    library's switch in the same mode; mode 0 is the reference's file again;
    bits outside the public forms are refused."""
    luma, chroma, nv12, want = frame
    say(f"public encoder forms {W}x{H}")
    buf = cj.DeviceBuffer.from_array(nv12)
    pub = cj.Encoder(W, H, D)
    stg = cj.Encoder(W, H, D, stages=True)
    for mode in (cj.MODE_MFMA16, cj.MODE_LEARNED, cj.MODE_MFMA16 | cj.MODE_LEARNED):
        pub.set_mode(mode)
        stg.set_mode(mode)
        assert pub.encode_device(buf.ptr, buf.ptr + W * H, W, W) == stg.encode_device(buf.ptr, buf.ptr + W * H, W, W)
    pub.set_mode(0)
    assert pub.encode_device(buf.ptr, buf.ptr + W * H, W, W) == want
    with pytest.raises(ValueError):
        pub.set_mode(cj.MODE_FORCE16)
    pub.close()
    stg.close()
    buf.close()
