"""GPU: the HIP path against the oracle, bit for bit.

Each case runs the real device pipeline through the C ABI (or the stage
harness for intermediate buffers and coefficient-only inputs), then compares
with the oracle on the same seeded input: the quant field, transform map, FP16
coefficients, CfL partials and maps, quantized AC/DC, non-zero grid,
histograms and the final `.jxl` bytes.
"""

from __future__ import annotations

import hashlib
import sys

import numpy as np
import pytest

import cujpegxl_amd as cj
import patterns
from cujpegxl_amd.synth import synthetic_nv12
from decoded import check_decoded, psnr, reconstruct_xyb

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def oracle_frontend(jo, luma, chroma, w, h, d):
    xyb = jo.nv12_to_xyb(luma, chroma, w, h)
    qf = jo.adaptive_quant(xyb, d)
    sharp = jo.gaborish(xyb)
    acs = jo.select_transforms(sharp[1], d)
    c16 = jo.to_f16(jo.forward_dct(sharp, acs))
    part, ytox, ytob = jo.cfl(c16, acs, w, h)
    ac, dc, nz = jo.quantize(c16, acs, ytox, ytob, qf, w, h, d)
    return dict(qf=qf, acs=acs, c16=c16, part=part, ytox=ytox, ytob=ytob, ac=ac, dc=dc, nz=nz)


def device_encode(w, h, d, luma, chroma, any_size=False):
    enc = cj.Encoder(w, h, d, stages=True, any_size=any_size)
    nv12 = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    data = enc.encode_device(nv12.ptr, nv12.ptr + w * h, w, w)
    return enc, data, nv12


@pytest.mark.parametrize("w,h,d,seed", [
    (512, 264, 1.0, 1),
    (264, 520, 3.0, 2),
    (1024, 576, 1.0, 3),
    (776, 456, 2.0, 4),
])
def test_pipeline_matches_oracle(jo, w, h, d, seed):
    say(f"pipeline {w}x{h} d={d} seed={seed}")
    luma, chroma = synthetic_nv12(w, h, seed)
    o = oracle_frontend(jo, luma, chroma, w, h, d)
    enc, data, _buf = device_encode(w, h, d, luma, chroma)
    np.testing.assert_array_equal(enc.fetch("quant_field", np.int32), o["qf"], err_msg="AQ field")
    np.testing.assert_array_equal(enc.fetch("acs", np.int8), o["acs"], err_msg="transform map")
    np.testing.assert_array_equal(enc.fetch("coeffs", np.uint16), o["c16"], err_msg="coefficients")
    np.testing.assert_array_equal(enc.fetch("cfl_partials", np.float32), o["part"], err_msg="CfL partials")
    np.testing.assert_array_equal(enc.fetch("ytox", np.int8), o["ytox"])
    np.testing.assert_array_equal(enc.fetch("ytob", np.int8), o["ytob"])
    np.testing.assert_array_equal(enc.fetch("ac", np.int16), o["ac"], err_msg="quantized AC")
    np.testing.assert_array_equal(enc.fetch("dc", np.int32), o["dc"], err_msg="quantized DC")
    nzg = enc.fetch("nz_grid", np.uint8)
    np.testing.assert_array_equal(nzg, o["nz"], err_msg="non-zero grid")
    want = jo.encode_nv12(luma, chroma, w, h, d)
    assert data == want, f"file differs: {len(data)} vs {len(want)} bytes"


@pytest.mark.parametrize("name,maker", [
    ("ac_mixed_512x384", lambda: patterns.mixed_ac_frame(512, 384)),
    ("cs_asm_2560x256", lambda: patterns.assembly_frame(2560, 256)),
    ("cs_asm_640x384", lambda: patterns.assembly_frame(640, 384)),
    ("cs_mixeddc_2560x2304", lambda: patterns.mixed_dc_frame(2560, 2304)),
    ("cs_asm_256", lambda: patterns.assembly_frame(256, 256)),
])
def test_entropy_on_reference_test_frames(jo, golden, name, maker):
    """The reference's entropy/assembly test frames: the device codestream is
    the reference host writer's (golden SHA-256), inside the container."""
    say(f"entropy {name}")
    fc = maker()
    acs = fc.acs if fc.acs is not None else np.full(fc.bw * fc.bh, 8, np.int8)
    enc = cj.Encoder(fc.w, fc.h, 1.0, any_size=True)
    data = enc.encode_coefficients(fc.ac.ravel().astype(np.int16), fc.dc.ravel(), acs, fc.qf,
                                   fc.ytox, fc.ytob, fc.gs, fc.qdc)
    cs = jo.write_codestream(jo.make_frame(fc.w, fc.h, fc.gs, fc.qdc, fc.dc.ravel(), fc.ac.ravel(),
                                           acs, fc.qf, fc.ytox, fc.ytob))
    assert data[40:] == cs
    if name.startswith("cs_"):  # whole-codestream vectors (write_vardct_codestream)
        assert hashlib.sha256(data[40:]).digest() == golden[f"{name}_sha"].tobytes()
    assert data[:12] == b"\x00\x00\x00\x0cJXL \x0d\x0a\x87\x0a"


def test_histograms_and_codes_match_oracle(jo):
    say("histograms 512x384 mixed")
    fc = patterns.mixed_dc_frame(512, 384)
    ac = patterns.mixed_ac_frame(512, 384).ac
    enc = cj.Encoder(512, 384, 1.0, stages=True)
    enc.encode_coefficients(ac.ravel().astype(np.int16), fc.dc.ravel(), fc.acs, fc.qf, fc.ytox,
                            fc.ytob, fc.gs, fc.qdc)
    f = jo.make_frame(512, 384, fc.gs, fc.qdc, fc.dc.ravel(), ac.ravel(), fc.acs, fc.qf, fc.ytox,
                      fc.ytob)
    import ctypes as C
    hist = np.zeros(32 * 256, np.uint32)
    jo.lib().jo_ac_histogram(C.byref(f[0]), jo.ptr(hist))
    dch = np.zeros(256, np.uint32)
    mth = np.zeros(256, np.uint32)
    jo.lib().jo_dc_histograms(C.byref(f[0]), jo.ptr(dch), jo.ptr(mth))
    got = enc.fetch("histograms", np.uint32).reshape(-1, 256)
    np.testing.assert_array_equal(got[:32].ravel(), hist)
    np.testing.assert_array_equal(got[32], dch)
    np.testing.assert_array_equal(got[33], mth)
    depth = enc.fetch("code_depth", np.uint8).reshape(-1, 256)
    bits = enc.fetch("code_bits", np.uint16).reshape(-1, 256)
    for row in range(34):
        h = got[row]
        nzs = np.nonzero(h)[0]
        alpha = int(nzs[-1]) + 1 if nzs.size else 1
        if alpha > 1:
            d, b, _, _ = _huff(jo, np.ascontiguousarray(h[:alpha]))
            np.testing.assert_array_equal(depth[row, :alpha], d, err_msg=f"row {row}")
            np.testing.assert_array_equal(bits[row, :alpha], b, err_msg=f"row {row}")


def _huff(jo, h):
    import ctypes as C
    d = np.zeros(h.size, np.uint8)
    b = np.zeros(h.size, np.uint16)
    out = np.zeros(4096, np.uint8)
    nb = C.c_size_t()
    jo.lib().jo_huffman(jo.ptr(h), h.size, jo.ptr(d), jo.ptr(b), jo.ptr(out), out.size, C.byref(nb))
    return d, b, out, nb.value


def test_single_group_frame(jo):
    """256x256 (one AC group, one combined section) through the harness."""
    say("single group 256x256")
    luma, chroma = synthetic_nv12(256, 256, 5)
    enc, data, _buf = device_encode(256, 256, 1.0, luma, chroma, any_size=True)
    assert data == jo.encode_nv12(luma, chroma, 256, 256, 1.0)


def test_c_abi_behaviour():
    say("C ABI")
    assert cj.query_backend(0) == "fp16-tensor"
    with pytest.raises(RuntimeError):
        cj.query_backend(9999)
    luma, chroma = synthetic_nv12(512, 512, 6)
    nv12 = np.concatenate([luma.ravel(), chroma.ravel()])
    a = cj.encode(nv12, 512, 512)
    b = cj.encode(nv12, 512, 512)
    assert a == b and a.startswith(b"\x00\x00\x00\x0cJXL ")
    data, stages = cj.encode_with_stats(nv12, 512, 512)
    assert data == a and [s["name"] for s in stages] == ["frontend", "entropy", "assembly"]
    enc = cj.Encoder(512, 512, 1.0)
    dev = cj.DeviceBuffer.from_array(nv12)
    s, size = enc.encode_status(dev.ptr, dev.ptr + 512 * 512, 512, 512, 16)
    assert s == cj.BUFFER_TOO_SMALL and size == len(a)
    # Pitched input: rows padded to 640 bytes encode to the same file.
    pl = np.zeros((512, 640), np.uint8)
    pl[:, :512] = luma
    pc = np.zeros((256, 640), np.uint8)
    pc[:, :512] = chroma
    dl, dcb = cj.DeviceBuffer.from_array(pl), cj.DeviceBuffer.from_array(pc)
    assert enc.encode_device(dl.ptr, dcb.ptr, 640, 640) == a


@pytest.mark.parametrize("d,seed", [(1.0, 0), (3.0, 1)])
def test_4k_frame_matches_oracle(jo, d, seed):
    """The plain 4K workloads (3840x2160 at d = 1.0 and 3.0): the file equals
    the oracle's byte for byte, and a second encode of the same planes
    through the same encoder gives the same file."""
    say(f"4K d={d} seed={seed}")
    luma, chroma = synthetic_nv12(3840, 2160, seed)
    enc, data, _buf = device_encode(3840, 2160, d, luma, chroma)
    want = jo.encode_nv12(luma, chroma, 3840, 2160, d)
    assert data == want, f"{len(data)} vs {len(want)}"
    again = enc.encode_device(_buf.ptr, _buf.ptr + 3840 * 2160, 3840, 3840)
    assert again == data


def test_dense_coefficients_match_oracle(jo):
    """Large random coefficients in every slot: some block rows of an AC group
    exceed the emitter's LDS range and take the global-atomic path, the rest
    the staged path; the codestream is the oracle's either way."""
    say("dense coefficients 512x264")
    w, h = 512, 264
    rng = np.random.default_rng(11)
    nblk = (w // 8) * (h // 8)
    acs = np.full(nblk, 8, np.int8)
    ac = rng.integers(-3000, 3000, size=3 * nblk * 64).astype(np.int16)
    ac.reshape(3, nblk, 64)[:, :, 0] = 0
    ac.reshape(3, nblk, 64)[:, : nblk // 2, 8:] = 0  # sparse top half: staged rows
    dc = rng.integers(-500, 500, size=3 * nblk).astype(np.int32)
    qf = rng.integers(1, 256, size=nblk).astype(np.int32)
    tiles = ((w // 8 + 7) // 8) * ((h // 8 + 7) // 8)
    ytox = rng.integers(-128, 128, size=tiles).astype(np.int8)
    ytob = rng.integers(-128, 128, size=tiles).astype(np.int8)
    gs, qdc = 2048, 300
    enc = cj.Encoder(w, h, 1.0, any_size=True)
    data = enc.encode_coefficients(ac, dc, acs, qf, ytox, ytob, gs, qdc)
    cs = jo.write_codestream(jo.make_frame(w, h, gs, qdc, dc, ac, acs, qf, ytox, ytob))
    assert len(cs) > 300_000
    assert data[40:] == cs


def _huffman_cases():
    rng = np.random.default_rng(21)
    cases = []
    fib = [1, 1]
    while len(fib) < 40:
        fib.append(fib[-1] + fib[-2])
    for lo in (0, 30, 200):  # depth limit retries, with and without > 64 symbols
        h = np.zeros(256, np.uint32)
        h[lo:lo + 40] = fib
        cases.append(h)
    while len(fib) < 45:
        fib.append(fib[-1] + fib[-2])
    h = np.zeros(256, np.uint32)
    h[:45] = fib  # totals near 3e9: the count floor climbs past 2^15
    cases.append(h)
    for k in (1, 2, 3, 4, 5, 6, 17, 64, 65, 200, 256):
        h = np.zeros(256, np.uint32)
        idx = rng.choice(256, size=k, replace=False)
        h[idx] = rng.integers(1, 1000, size=k)
        cases.append(h)
    h = np.zeros(256, np.uint32)
    h[:64] = 7  # all ties
    cases.append(h)
    h = np.zeros(256, np.uint32)
    h[0] = 5  # alphabet of one
    cases.append(h)
    cases.append(np.zeros(256, np.uint32))
    h = np.zeros(256, np.uint32)
    h[255] = 3  # single symbol at the top
    cases.append(h)
    for _ in range(40):
        k = int(rng.integers(2, 257))
        h = np.zeros(256, np.uint32)
        h[:k] = rng.geometric(rng.uniform(0.02, 0.6), size=k) * rng.integers(0, 2, size=k)
        cases.append(h)
    return np.stack(cases)


def test_prefix_codes_match_oracle(jo):
    """The device code kernel on adversarial histograms (depth-limit retries,
    ties, 1-6 symbols, alphabets past 64) against the oracle's builder."""
    say("prefix codes")
    hists = _huffman_cases()
    got = cj.build_codes(hists)
    for i, h in enumerate(hists):
        nzs = np.nonzero(h)[0]
        alpha = int(nzs[-1]) + 1 if nzs.size else 1
        assert got[i]["alpha"] == alpha, i
        if alpha <= 1:
            assert got[i]["header_bits"] == 0 and not got[i]["depth"].any()
            continue
        d, b, out, nb = _huff(jo, np.ascontiguousarray(h[:alpha]))
        np.testing.assert_array_equal(got[i]["depth"][:alpha], d, err_msg=f"case {i}")
        np.testing.assert_array_equal(got[i]["bits"][:alpha], b, err_msg=f"case {i}")
        assert got[i]["header_bits"] == nb, (i, got[i]["header_bits"], nb)
        assert got[i]["header"] == out[: (nb + 7) // 8].tobytes(), i


def test_single_symbol_histogram_has_zero_depths(jo):
    """A histogram with one present symbol takes the simple header and keeps
    every depth at 0 (prefix_code.cc build_and_store_huffman_tree)."""
    say("single-symbol histogram")
    h = np.zeros((2, 256), np.uint32)
    h[0, 16] = 1
    h[1, 200] = 12345
    got = cj.build_codes(h)
    for i in range(2):
        a = int(np.nonzero(h[i])[0][-1]) + 1
        d, b, out, nb = _huff(jo, np.ascontiguousarray(h[i, :a]))
        assert not got[i]["depth"].any() and not d.any()
        assert got[i]["header_bits"] == nb and got[i]["header"] == out[: (nb + 7) // 8].tobytes()


@pytest.mark.parametrize("w,h,d,seed", [(512, 264, 1.0, 1), (1024, 576, 3.0, 3), (3840, 2160, 1.0, 0)])
def test_device_files_decode_to_device_buffers(jo, golden, w, h, d, seed):
    """The subset decoder reads every device file back to the device's own
    quantized frame (every section ends exactly at its TOC size)."""
    say(f"decode device file {w}x{h} d={d} seed={seed}")
    luma, chroma = synthetic_nv12(w, h, seed)
    enc, data, _buf = device_encode(w, h, d, luma, chroma)
    cal = jo.calibrate(d)
    check_decoded(golden, jo.decode(data), w, h, cal.global_scale, cal.quant_dc, enc.fetch("acs", np.int8),
                  enc.fetch("quant_field", np.int32), enc.fetch("dc", np.int32), enc.fetch("ac", np.int16),
                  enc.fetch("ytox", np.int8), enc.fetch("ytob", np.int8))


@pytest.mark.parametrize("w,h,seed", [(256, 256, 5), (640, 384, 6), (1024, 576, 7)])
def test_dct8_uniform_mode_matches_oracle(jo, w, h, seed):
    """Config-1 semantics on the device (DCT8 everywhere, uniform quant field,
    quant.cu's quantizer): the file equals the oracle's and decodes to a
    DCT8-only, uniform-field frame."""
    say(f"dct8 uniform {w}x{h} seed={seed}")
    luma, chroma = synthetic_nv12(w, h, seed)
    enc = cj.Encoder(w, h, 1.0, any_size=True)
    nv12 = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    data = enc.encode_dct8_uniform(nv12.ptr, nv12.ptr + w * h, w, w)
    want = jo.encode_dct8_uniform(luma, chroma, w, h, 1.0)
    assert data == want, f"file differs: {len(data)} vs {len(want)} bytes"
    dec = jo.decode(data)
    assert (dec["acs"] == 8).all()
    assert (dec["qf"] == jo.calibrate(1.0).raw_quant_field).all()
    nv12.close()
    enc.close()


def test_device_4k_file_decodes_to_the_source(jo):
    """The 4K device file, decoded and rebuilt to pixels (no EPF), against
    the source frame's XYB."""
    say("decode 4K device file to pixels")
    w, h = 3840, 2160
    luma, chroma = synthetic_nv12(w, h, 0)
    _enc, data, _buf = device_encode(w, h, 1.0, luma, chroma)
    rec = reconstruct_xyb(jo, jo.decode(data))
    xyb = jo.nv12_to_xyb(luma, chroma, w, h).astype(np.float64)
    got = [psnr(rec[c], xyb[c]) for c in range(3)]
    assert all(g >= f for g, f in zip(got, (38.0, 34.0, 35.0))), got


def test_mixed_resolution_frames_match_oracle(jo):
    """Frames of the mixed-resolution batch's sizes (640x360 ... 2560x1440,
    distances in [0.8, 3.0]): heights that are not multiples of 32 or 64,
    partial regions, tiles and groups.  Each encoder takes two seeds in a row
    (its buffers are reused); every file equals the oracle's and decodes."""
    cases = [(640, 360, 0.8), (1280, 720, 2.2), (1920, 1080, 3.0), (2560, 1440, 1.4)]
    for w, h, d in cases:
        enc = cj.Encoder(w, h, d)
        for seed in (11, 12):
            say(f"mixed batch {w}x{h} d={d} seed={seed}")
            luma, chroma = synthetic_nv12(w, h, seed)
            buf = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
            data = enc.encode_device(buf.ptr, buf.ptr + w * h, w, w)
            buf.close()
            want = jo.encode_nv12(luma, chroma, w, h, d)
            assert data == want, (w, h, d, seed, len(data), len(want))
            dec = jo.decode(data)
            assert dec["acs"].size == (w // 8) * (h // 8)
        enc.close()


@pytest.mark.parametrize("kind", ["flat", "noise"])
def test_extreme_content_matches_oracle(jo, kind):
    """A flat frame (every block's AC is zero: the tokenizer's list of items
    with non-zeros is empty and every first block sits in the zero-count
    list) and a full-range noise frame (nearly every item has non-zeros)."""
    w, h, d = 512, 264, 1.0
    rng = np.random.default_rng(21)
    if kind == "flat":
        luma = np.full((h, w), 117, np.uint8)
        chroma = np.full((h // 2, w), 131, np.uint8)
    else:
        luma = rng.integers(0, 256, (h, w), dtype=np.uint8)
        chroma = rng.integers(0, 256, (h // 2, w), dtype=np.uint8)
    say(f"extreme content {kind} {w}x{h}")
    enc, data, _buf = device_encode(w, h, d, luma, chroma)
    want = jo.encode_nv12(luma, chroma, w, h, d)
    assert data == want, (kind, len(data), len(want))
    dec = jo.decode(data)
    if kind == "flat":
        assert not dec["ac"].any()
