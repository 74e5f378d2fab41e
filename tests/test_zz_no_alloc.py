"""GPU: cujpegxl_encode_noalloc (the allocation-free whole-frame encode, its
buffers in the library's device scratch array) writes the same files as
cujpegxl_encoder_encode, whose bytes equal the oracle's: two calls in a row
on one stream, the second with a smaller frame and another distance, then
the first shape again; plus the limits and argument checks."""

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


def reference_file(w: int, h: int, d: float, nv12: np.ndarray) -> bytes:
    """The same frame through cujpegxl_encoder_encode."""
    enc = cj.Encoder(w, h, d)
    dev = cj.DeviceBuffer.from_array(nv12)
    try:
        return enc.encode_device(dev.ptr, dev.ptr + w * h, w, w)
    finally:
        dev.close()
        enc.close()


def noalloc_file(L, w: int, h: int, d: float, dev: cj.DeviceBuffer, out: cj.DeviceBuffer,
                 size: cj.DeviceBuffer, stream: int) -> tuple[int, bytes]:
    cfg = cj.Config(w, h, d, 0)
    inp = cj.Nv12Input(dev.ptr, dev.ptr + w * h, w, w)
    st = L.cujpegxl_encode_noalloc(C.byref(cfg), C.byref(inp), out.ptr, out.nbytes, size.ptr, stream)
    if st != cj.OK:
        return st, b""
    cj.stages_lib().cjx_stream_synchronize(stream)
    n = int(size.to_host(np.uint64)[0])
    assert 0 < n <= out.nbytes
    return st, out.to_host()[:n].tobytes()


def test_noalloc_matches_encoder_encode(jo):
    L = cj.public_lib()
    scratch = (104 << 20)
    assert L.cujpegxl_noalloc_workspace_bytes(1920, 1088) <= scratch
    frames = [(1920, 1080, 1.0, 3), (640, 360, 2.0, 4), (1920, 1080, 1.0, 3)]
    out = cj.DeviceBuffer(cj.max_output_size(1920, 1080))
    size = cj.DeviceBuffer(8)
    stream = cj.stages_lib().cjx_stream_create()
    try:
        for w, h, d, seed in frames:
            luma, chroma = synthetic_nv12(w, h, seed)
            nv12 = np.concatenate([luma.ravel(), chroma.ravel()])
            say(f"noalloc {w}x{h} d={d} seed={seed}")
            want = reference_file(w, h, d, nv12)
            dev = cj.DeviceBuffer.from_array(nv12)
            st, got = noalloc_file(L, w, h, d, dev, out, size, stream)
            dev.close()
            assert st == cj.OK
            assert got == want, f"{len(got)} vs {len(want)} bytes"
        # The smaller frame on the null stream, straight after the larger.
        w, h, d = 640, 360, 2.0
        luma, chroma = synthetic_nv12(w, h, 4)
        nv12 = np.concatenate([luma.ravel(), chroma.ravel()])
        dev = cj.DeviceBuffer.from_array(nv12)
        say("noalloc 640x360 on the null stream")
        st, got = noalloc_file(L, w, h, d, dev, out, size, 0)
        assert st == cj.OK and got == jo.encode_nv12(luma, chroma, w, h, d)
        # A capacity below the file: the size word still holds the full size.
        cfg = cj.Config(w, h, d, 0)
        inp = cj.Nv12Input(dev.ptr, dev.ptr + w * h, w, w)
        assert L.cujpegxl_encode_noalloc(C.byref(cfg), C.byref(inp), out.ptr, 64, size.ptr, stream) == cj.OK
        cj.stages_lib().cjx_stream_synchronize(stream)
        assert int(size.to_host(np.uint64)[0]) == len(got)
        assert out.to_host()[:64].tobytes() == got[:64]
        # Limits and argument checks.
        big = cj.Config(3840, 2160, 1.0, 0)
        assert L.cujpegxl_encode_noalloc(C.byref(big), C.byref(inp), out.ptr, out.nbytes, size.ptr, stream) == \
            cj.UNSUPPORTED_RESOLUTION
        bad = cj.Config(w, h, -1.0, 0)
        assert L.cujpegxl_encode_noalloc(C.byref(bad), C.byref(inp), out.ptr, out.nbytes, size.ptr, stream) == \
            cj.INVALID_ARGUMENT
        assert L.cujpegxl_encode_noalloc(C.byref(cfg), C.byref(inp), out.ptr, out.nbytes, size.ptr + 4, stream) == \
            cj.INVALID_ARGUMENT
        short = cj.Nv12Input(dev.ptr, dev.ptr + w * h, w - 8, w)
        assert L.cujpegxl_encode_noalloc(C.byref(cfg), C.byref(short), out.ptr, out.nbytes, size.ptr, stream) == \
            cj.INVALID_ARGUMENT
        dev.close()
    finally:
        cj.stages_lib().cjx_stream_synchronize(stream)
        cj.stages_lib().cjx_stream_destroy(stream)
        out.close()
        size.close()
