"""CPU: the subset decoder (oracle/jxl_decode.c).

It reads the reference's own codestreams -- the oracle writer's bytes for the
reference tests' frames, whose SHA-256 equal the reference host writer's
golden hashes -- back to those frames; it round-trips the oracle's full
encodes, multi-group, single-group and DCT8-only; and it rejects damaged
files (truncation, a trailing byte the TOC does not account for, a bad
signature, flipped bits).  Pixels rebuilt from decoded files match the source
frame to the PSNR the distance implies."""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

import patterns
from cujpegxl_amd.synth import synthetic_nv12
from decoded import check_decoded, psnr, reconstruct_xyb


@pytest.mark.parametrize("name,maker", [
    ("cs_asm_512", lambda: patterns.assembly_frame(512, 512)),
    ("cs_asm_2560x256", lambda: patterns.assembly_frame(2560, 256)),
    ("cs_asm_640x384", lambda: patterns.assembly_frame(640, 384)),
    ("cs_asm_256", lambda: patterns.assembly_frame(256, 256)),
    ("cs_mixed_512x384", lambda: patterns.mixed_ac_frame(512, 384)),
    ("cs_mixeddc_2560x2304", lambda: patterns.mixed_dc_frame(2560, 2304)),
])
def test_decodes_reference_codestreams(jo, golden, name, maker):
    fc = maker()
    acs = fc.acs if fc.acs is not None else np.full(fc.bw * fc.bh, 8, np.int8)
    cs = jo.write_codestream(jo.make_frame(fc.w, fc.h, fc.gs, fc.qdc, fc.dc.ravel(), fc.ac.ravel(), acs,
                                           fc.qf, fc.ytox, fc.ytob))
    assert hashlib.sha256(cs).digest() == golden[f"{name}_sha"].tobytes()  # the reference's bytes
    dec = jo.decode(cs)
    check_decoded(golden, dec, fc.w, fc.h, fc.gs, fc.qdc, acs, fc.qf, fc.dc, fc.ac, fc.ytox, fc.ytob)


def oracle_frame(jo, luma, chroma, w, h, d):
    xyb = jo.nv12_to_xyb(luma, chroma, w, h)
    qf = jo.adaptive_quant(xyb, d)
    sharp = jo.gaborish(xyb)
    acs = jo.select_transforms(sharp[1], d)
    c16 = jo.to_f16(jo.forward_dct(sharp, acs))
    _part, ytox, ytob = jo.cfl(c16, acs, w, h)
    ac, dc, _nz = jo.quantize(c16, acs, ytox, ytob, qf, w, h, d)
    cal = jo.calibrate(d)
    return cal, acs, qf, dc, ac, ytox, ytob


@pytest.mark.parametrize("w,h,d,seed", [(512, 264, 1.0, 1), (256, 256, 1.0, 2), (1024, 576, 2.0, 3),
                                        (776, 456, 3.0, 4)])
def test_round_trips_oracle_encodes(jo, golden, w, h, d, seed):
    luma, chroma = synthetic_nv12(w, h, seed)
    data = jo.encode_nv12(luma, chroma, w, h, d)
    cal, acs, qf, dc, ac, ytox, ytob = oracle_frame(jo, luma, chroma, w, h, d)
    check_decoded(golden, jo.decode(data), w, h, cal.global_scale, cal.quant_dc, acs, qf, dc, ac, ytox, ytob)


def test_round_trips_dct8_uniform_encode(jo):
    w, h = 256, 256
    luma, chroma = synthetic_nv12(w, h, 5)
    dec = jo.decode(jo.encode_dct8_uniform(luma, chroma, w, h, 1.0))
    assert (dec["acs"] == 8).all()
    assert len(set(dec["qf"].tolist())) == 1  # uniform quant field


def test_rejects_damaged_files(jo):
    w, h = 512, 264
    luma, chroma = synthetic_nv12(w, h, 1)
    data = bytearray(jo.encode_nv12(luma, chroma, w, h, 1.0))
    good = jo.decode(bytes(data))
    with pytest.raises(jo.DecodeError):
        jo.decode(bytes(data[:-1]))
    with pytest.raises(jo.DecodeError):
        jo.decode(b"\x00\x00\x00\x0cJXL \x0d\x0a\x87\x0b" + bytes(data[12:]))
    # One more codestream byte than the sections' TOC sizes add up to.
    longer = bytearray(data) + b"\x00"
    longer[32:36] = (int.from_bytes(data[32:36], "big") + 1).to_bytes(4, "big")
    with pytest.raises(jo.DecodeError, match="after the last section"):
        jo.decode(bytes(longer))
    # Flipping any one of these bits inside the sections either breaks the
    # decode or changes the decoded frame: nothing passes silently.
    rng = np.random.default_rng(0)
    body = len(data) // 2
    for pos in rng.integers(body - 2000, len(data) - 1, 24):
        bad = bytearray(data)
        bad[pos] ^= 1 << int(rng.integers(0, 8))
        try:
            dec = jo.decode(bytes(bad))
        except jo.DecodeError:
            continue
        assert any(not np.array_equal(dec[k], good[k]) for k in ("ac", "dc", "acs", "qf", "ytox", "ytob"))


@pytest.mark.parametrize("w,h,d,seed,floor", [(512, 264, 1.0, 1, (38.0, 34.0, 35.0)),
                                              (256, 256, 1.0, 2, (38.0, 34.0, 35.0)),
                                              (512, 264, 3.0, 1, (33.0, 30.0, 31.0))])
def test_decoded_pixels_match_the_source(jo, w, h, d, seed, floor):
    """Pixels rebuilt from the decoded file (dequantization, colour
    correlation, DC-derived lowest frequencies, inverse DCT, the decoder's
    gaborish; no EPF) against the source frame's XYB: per-channel PSNR floors,
    and the decoder's gaborish undoes the encoder's sharpening (it raises the
    PSNR of Y)."""
    luma, chroma = synthetic_nv12(w, h, seed)
    dec = jo.decode(jo.encode_nv12(luma, chroma, w, h, d))
    xyb = jo.nv12_to_xyb(luma, chroma, w, h).astype(np.float64)
    rec = reconstruct_xyb(jo, dec)
    got = [psnr(rec[c], xyb[c]) for c in range(3)]
    assert all(g >= f for g, f in zip(got, floor)), got
    assert psnr(reconstruct_xyb(jo, dec, gaborish=False)[1], xyb[1]) < got[1]


def test_largest_frame_head_round_trips(jo):
    """16384x8192 with one small non-zero per block and channel: 2,048 AC
    groups of at least 1 KiB, so the TOC alone takes more than 4 KB; the
    oracle's file decodes back to the frame and every section ends at its
    TOC size."""
    w, h = 16384, 8192
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
    cs = jo.write_codestream(jo.make_frame(w, h, 2048, 300, dc, ac.ravel(), acs, qf, None, None))
    dec = jo.decode(cs)
    assert dec["width"] == w and dec["height"] == h
    assert np.array_equal(dec["dc"].ravel(), dc)
    assert np.array_equal(dec["ac"].ravel(), ac.ravel())
    assert np.array_equal(dec["qf"].ravel(), qf)
