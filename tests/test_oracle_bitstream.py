"""CPU: the oracle's entropy coder and codestream writer against the golden
vectors produced by the reference's own host code (tests/golden), and against
oracle/_ref directly where it is built."""

from __future__ import annotations

import ctypes as C
import hashlib

import numpy as np
import pytest

import patterns


def sha(b: bytes) -> bytes:
    return hashlib.sha256(b).digest()


def test_natural_orders(jo, golden):
    for n in (8, 16, 32):
        o = np.zeros(n * n, np.uint32)
        jo.lib().jo_natural_order(n, jo.ptr(o))
        np.testing.assert_array_equal(o, golden[f"order{n}"])
        assert sorted(o.tolist()) == list(range(n * n))  # a permutation
        m = n // 8
        # The LLF corner comes first, in raster order of the corner.
        assert sorted(o[: m * m].tolist()) == sorted(y * n + x for y in range(m) for x in range(m))


def test_calibration(jo, golden):
    for i, d in enumerate(golden["cal_d"]):
        c = jo.calibrate(float(d))
        assert (c.global_scale, c.quant_dc, c.raw_quant_field) == tuple(golden["cal_i"][i])
        assert np.float32(c.global_scale_float) == golden["cal_f"][i][0]
        assert np.float32(c.inv_global_scale) == golden["cal_f"][i][1]


def test_ac_cluster_table_and_known_answers(jo, golden):
    got = np.array([jo.lib().jo_ac_cluster(i) for i in range(7425)])
    np.testing.assert_array_equal(got, golden["ac_cluster"])
    # ac_context_test.cc:53-64
    cc = 15 * 37
    L = jo.lib()
    assert [L.jo_ac_cluster(x) for x in (0, 60, 180, 360)] == [0, 1, 2, 3]
    assert [L.jo_ac_cluster(cc + x) for x in (0, 32, 424, 7 * 458)] == [8, 9, 19, 20]


def _huff(jo, h):
    d = np.zeros(h.size, np.uint8)
    b = np.zeros(h.size, np.uint16)
    out = np.zeros(4096, np.uint8)
    nb = C.c_size_t()
    n = jo.lib().jo_huffman(jo.ptr(h), h.size, jo.ptr(d), jo.ptr(b), jo.ptr(out), out.size,
                            C.byref(nb))
    return d, b, out[:n], nb.value


def test_huffman_codes_and_headers(jo, golden):
    off = hoff = 0
    for i, n in enumerate(golden["huff_len"]):
        h = golden["huff_hist"][off : off + n].copy()
        d, b, hdr, nb = _huff(jo, h)
        np.testing.assert_array_equal(d, golden["huff_depth"][off : off + n], err_msg=f"#{i}")
        np.testing.assert_array_equal(b, golden["huff_bits"][off : off + n], err_msg=f"#{i}")
        hl = golden["huff_hdr_len"][i]
        np.testing.assert_array_equal(hdr, golden["huff_hdr"][hoff : hoff + hl], err_msg=f"#{i}")
        assert nb == golden["huff_hdr_bits"][i]
        off += n
        hoff += hl


def test_huffman_is_a_valid_prefix_code(jo):
    """Kraft equality, 15-bit limit and canonical ordering, on skewed inputs that
    force the count-floor retry."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(5, 257))
        h = (2.0 ** rng.uniform(0, 30, n)).astype(np.uint32) * rng.integers(0, 2, n).astype(np.uint32)
        h[:5] = np.maximum(h[:5], 1)
        d, b, _, _ = _huff(jo, h)
        used = h > 0
        assert d[used].max() <= 15 and (d[~used] == 0).all()
        assert abs(sum(2.0 ** -int(x) for x in d[used]) - 1.0) < 1e-12
        codes = {(int(x), int(y)) for x, y in zip(d[used], b[used])}
        assert len(codes) == int(used.sum())  # distinct code words


def test_context_map_and_mtf(jo, golden):
    cmap = golden["ac_cluster"].astype(np.uint8)
    out = np.zeros(65536, np.uint8)
    nb, enc = C.c_size_t(), C.c_int()
    n = jo.lib().jo_best_context_map(jo.ptr(cmap), cmap.size, 32, jo.ptr(out), out.size,
                                     C.byref(nb), C.byref(enc))
    np.testing.assert_array_equal(out[:n], golden["cmap_ac_bytes"])
    assert [nb.value, enc.value] == golden["cmap_ac_bits"].tolist()
    # context_map_test.cc:80-85
    src = np.array([0, 1, 1, 0, 2, 1, 2, 0], np.uint8)
    dst = np.zeros(8, np.uint8)
    jo.lib().jo_move_to_front(jo.ptr(src), 8, jo.ptr(dst))
    assert dst.tolist() == [0, 1, 0, 1, 2, 2, 1, 2] == golden["mtf_out"].tolist()
    rep = np.full(7425, 173, np.uint8)
    out2 = np.zeros(7425, np.uint8)
    jo.lib().jo_move_to_front(jo.ptr(rep), rep.size, jo.ptr(out2))
    assert out2[0] == 173 and not out2[1:].any()


def _frame(jo, fc, qf=None):
    return jo.make_frame(fc.w, fc.h, fc.gs, fc.qdc, fc.dc.ravel(), fc.ac.ravel(), fc.acs,
                         fc.qf if qf is None else qf, fc.ytox, fc.ytob)


@pytest.mark.parametrize("name,maker", [
    ("ac_dct8_256", lambda: patterns.dct8_ac_frame(256, 256)),
    ("ac_mixed_256", lambda: patterns.mixed_ac_frame(256, 256)),
    ("ac_mixed_512x384", lambda: patterns.mixed_ac_frame(512, 384)),
])
def test_ac_entropy_matches_reference(jo, golden, name, maker):
    """entropy_gpu_test.cc:113-164 expected values (reference_ac_encode)."""
    f, keep = _frame(jo, maker())
    hist = np.zeros(32 * 256, np.uint32)
    jo.lib().jo_ac_histogram(C.byref(f), jo.ptr(hist))
    np.testing.assert_array_equal(hist, golden[f"{name}_hist"])
    depth = np.zeros(32 * 256, np.uint8)
    bits = np.zeros(32 * 256, np.uint16)
    out = np.zeros(1 << 16, np.uint8)
    jo.lib().jo_ac_global(jo.ptr(hist), 2, jo.ptr(depth), jo.ptr(bits), jo.ptr(out), out.size)
    np.testing.assert_array_equal(depth, golden[f"{name}_depth"])
    np.testing.assert_array_equal(bits, golden[f"{name}_bits"])
    sizes = np.zeros(golden[f"{name}_sizes"].size, np.uint32)
    buf = np.zeros(8 << 20, np.uint8)
    n = jo.lib().jo_ac_groups(C.byref(f), jo.ptr(depth), jo.ptr(bits), jo.ptr(sizes),
                              jo.ptr(buf), buf.size)
    np.testing.assert_array_equal(sizes, golden[f"{name}_sizes"])
    assert sha(buf[:n].tobytes()) == golden[f"{name}_sha"].tobytes()


@pytest.mark.parametrize("name,maker", [
    ("dc_dct8_256", lambda: patterns.dct8_dc_frame(256, 256)),
    ("dc_mixed_256", lambda: patterns.mixed_dc_frame(256, 256)),
    ("dc_mixed_2560x256", lambda: patterns.mixed_dc_frame(2560, 256)),
    ("dc_mixed_2560x2304", lambda: patterns.mixed_dc_frame(2560, 2304)),
])
def test_dc_entropy_matches_reference(jo, golden, name, maker):
    """dc_entropy_gpu_test.cc:130-167 expected sections (reference_dc_encode)."""
    f, keep = _frame(jo, maker())
    sizes = np.zeros(golden[f"{name}_sizes"].size, np.uint32)
    buf = np.zeros(16 << 20, np.uint8)
    n = jo.lib().jo_dc_groups(C.byref(f), jo.ptr(sizes), jo.ptr(buf), buf.size)
    np.testing.assert_array_equal(sizes, golden[f"{name}_sizes"])
    assert sha(buf[:n].tobytes()) == golden[f"{name}_sha"].tobytes()


@pytest.mark.parametrize("name,maker", [
    ("cs_asm_512", lambda: patterns.assembly_frame(512, 512)),
    ("cs_asm_2560x256", lambda: patterns.assembly_frame(2560, 256)),
    ("cs_asm_640x384", lambda: patterns.assembly_frame(640, 384)),
    ("cs_asm_256", lambda: patterns.assembly_frame(256, 256)),
    ("cs_mixed_512x384", lambda: patterns.mixed_ac_frame(512, 384)),
    ("cs_mixeddc_2560x2304", lambda: patterns.mixed_dc_frame(2560, 2304)),
])
def test_codestream_matches_reference(jo, golden, name, maker):
    """frame_assembly_test.cc:120-164 / vardct_frame.cc write_vardct_codestream."""
    cs = jo.write_codestream(_frame(jo, maker()))
    assert len(cs) == int(golden[f"{name}_len"][0])
    assert sha(cs) == golden[f"{name}_sha"].tobytes()
    if f"{name}_bytes" in golden:
        assert cs == golden[f"{name}_bytes"].tobytes()


def test_globals_head_container(jo, golden):
    out = np.zeros(64, np.uint8)
    n = jo.lib().jo_dc_global(4096, 32, jo.ptr(out), 64)
    np.testing.assert_array_equal(out[:n], golden["dc_global_4096_32"])
    sizes = golden["head_sizes"]
    out = np.zeros(4096, np.uint8)
    n = jo.lib().jo_codestream_head(3840, 2160, jo.ptr(sizes), sizes.size, jo.ptr(out), out.size)
    np.testing.assert_array_equal(out[:n], golden["head_3840x2160"])
    n = jo.lib().jo_codestream_head(256, 256, jo.ptr(sizes), 1, jo.ptr(out), out.size)
    np.testing.assert_array_equal(out[:n], golden["head_256"])
    assert bytes(golden["head_3840x2160"][:2]) == b"\xff\x0a"
    for cs in (1234, 0x100000000 - 8):
        out = np.zeros(64, np.uint8)
        n = jo.lib().jo_container_framing(cs, jo.ptr(out), 64)
        np.testing.assert_array_equal(out[:n], golden[f"container_{cs}"])
    # container_test.cc:84-141 known bytes
    out = np.zeros(64, np.uint8)
    n = jo.lib().jo_container_framing(1234, jo.ptr(out), 64)
    assert n == 40 and bytes(out[:12]) == b"\x00\x00\x00\x0cJXL \x0d\x0a\x87\x0a"
    assert bytes(out[12:32]) == b"\x00\x00\x00\x14ftypjxl \x00\x00\x00\x00jxl "
    assert int.from_bytes(bytes(out[32:36]), "big") == 1242 and bytes(out[36:40]) == b"jxlc"


def test_oracle_against_live_reference_build(jo):
    """Where oracle/_ref is built, random coefficient frames (not just the fixed
    patterns) must give byte-identical codestreams."""
    R = jo.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(11)
    for (w, h) in ((264, 256), (512, 520), (1040, 264)):
        fc = patterns.Coeffs(w, h)
        fc.acs = np.full(fc.bw * fc.bh, 8, np.int8)
        for by in range(0, fc.bh - 3, 4):
            for bx in range(0, fc.bw - 3, 4):
                r = rng.integers(0, 3)
                if r == 1:
                    fc.set_first_block(32, bx, by)
                elif r == 2:
                    fc.set_first_block(16, bx, by)
        mag = rng.integers(0, 3, fc.ac.shape)
        fc.ac[:] = (rng.integers(-300, 300, fc.ac.shape) * (mag == 0)
                    + rng.integers(-3, 4, fc.ac.shape) * (mag == 1)).astype(np.int32)
        fc.dc[:] = rng.integers(-2000, 2000, fc.dc.shape)
        tiles = ((fc.bw + 7) // 8) * ((fc.bh + 7) // 8)
        fc.ytox = rng.integers(-128, 128, tiles).astype(np.int8)
        fc.ytob = rng.integers(-128, 128, tiles).astype(np.int8)
        ours = jo.write_codestream(_frame(jo, fc))
        buf = np.zeros(64 << 20, np.uint8)
        P = jo.ptr
        dc, ac = np.ascontiguousarray(fc.dc.ravel()), np.ascontiguousarray(fc.ac.ravel())
        n = R.ref_write_codestream(w, h, fc.gs, fc.qdc, fc.rqf, P(dc), P(ac), P(fc.acs),
                                   P(fc.ytox), P(fc.ytob), 1, P(buf), buf.size)
        assert ours == buf[:n].tobytes(), (w, h)


def test_threaded_oracle_writes_the_same_file(jo):
    """The oracle's front-end loops split rows, tiles, regions and blocks
    over threads with disjoint outputs: any thread count gives the same file
    (the CPU baseline's all-cores leg relies on it)."""
    from cujpegxl_amd.synth import synthetic_nv12

    w, h = 776, 456
    luma, chroma = synthetic_nv12(w, h, 31)
    one = jo.encode_nv12(luma, chroma, w, h, 1.0)
    try:
        jo.set_threads(8)
        many = jo.encode_nv12(luma, chroma, w, h, 1.0)
    finally:
        jo.set_threads(1)
    assert many == one
