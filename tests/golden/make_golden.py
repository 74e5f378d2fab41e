"""Writes the golden fixtures under tests/golden/ from oracle/_ref.

oracle/_ref is the reference's own host C++ (bitstream writer, host VarDCT
encoder, selection/CfL/calibration headers, dequant tables) compiled from its
sources by `make -C oracle ref`.  This script runs it on the input patterns of
the reference's tests (tests/patterns.py) and on seeded inputs, and stores
inputs' parameters with expected outputs: the dequant tables as raw float32,
everything else in ref_vectors.npz (numeric arrays only).

Run from the repo root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import ctypes as C
import hashlib
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
import patterns  # noqa: E402

OUT = ROOT / "tests" / "golden"


def sha(b: bytes) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(b).digest(), np.uint8).copy()


def main() -> None:
    oracle.build(with_ref=True)
    R = oracle.ref()
    assert R is not None, "oracle/_ref did not build"
    P = oracle.ptr
    v: dict[str, np.ndarray] = {}

    # Dequant matrices (quant_weights_dct{8,16,32}.h).
    w8 = np.zeros(3 * 64, np.float32)
    w16 = np.zeros(3 * 256, np.float32)
    w32 = np.zeros(3 * 1024, np.float32)
    R.ref_dequant_tables(P(w8), P(w16), P(w32))
    np.concatenate([w8, w16, w32]).astype("<f4").tofile(OUT / "dequant_weights.bin")

    # Natural coefficient orders (coeff_order.cc).
    for n in (8, 16, 32):
        o = np.zeros(n * n, np.uint32)
        R.ref_natural_coeff_order(n, P(o))
        v[f"order{n}"] = o

    # Calibration (quant_calibration.h) over a distance sweep.
    dists = np.array([0.1, 0.5, 0.8, 1.0, 1.5, 2.0, 2.5, 3.0, 5.0, 8.0, 15.0, 25.0], np.float32)
    cal_i = np.zeros((dists.size, 3), np.uint32)
    cal_f = np.zeros((dists.size, 2), np.float32)
    for i, d in enumerate(dists):
        R.ref_calibrate(float(d), P(cal_i[i]), P(cal_f[i]))
    v["cal_d"], v["cal_i"], v["cal_f"] = dists, cal_i, cal_f

    # AC context clustering (ac_context.h) over the whole context space.
    v["ac_cluster"] = np.array([R.ref_ac_cluster(i) for i in range(7425)], np.int32)

    # Prefix codes (prefix_code.cc): seeded histograms of several shapes.
    rng = np.random.default_rng(20261202)
    hists = []
    for shape in ("geo", "flat", "one", "two", "three", "four", "five", "skew", "wide"):
        for _ in range(4):
            n = int(rng.integers(2, 257))
            h = np.zeros(n, np.uint32)
            if shape == "geo":
                h[:] = (1e6 * 0.7 ** np.arange(n)).astype(np.uint32)
            elif shape == "flat":
                h[:] = rng.integers(1, 50, n)
            elif shape in ("one", "two", "three", "four", "five"):
                k = ("one", "two", "three", "four", "five").index(shape) + 1
                idx = rng.choice(n, size=min(k, n), replace=False)
                h[idx] = rng.integers(1, 1000, idx.size)
            elif shape == "skew":
                h[:] = rng.integers(0, 3, n)
                h[0] = 10_000_000
            else:
                h[:] = (rng.pareto(0.6, n) * 10).astype(np.uint32)
            if not h.any():
                h[0] = 1
            hists.append(h)
    hl = np.array([h.size for h in hists], np.uint32)
    hcat = np.concatenate(hists)
    depth_all, bits_all, hdr_all, hdr_bits = [], [], [], []
    for h in hists:
        d = np.zeros(h.size, np.uint8)
        b = np.zeros(h.size, np.uint16)
        out = np.zeros(4096, np.uint8)
        nb = C.c_size_t()
        n = R.ref_huffman(P(h), h.size, P(d), P(b), P(out), out.size, C.byref(nb))
        depth_all.append(d)
        bits_all.append(b)
        hdr_all.append(out[:n])
        hdr_bits.append(nb.value)
    v["huff_len"], v["huff_hist"] = hl, hcat
    v["huff_depth"], v["huff_bits"] = np.concatenate(depth_all), np.concatenate(bits_all)
    v["huff_hdr_len"] = np.array([a.size for a in hdr_all], np.uint32)
    v["huff_hdr"] = np.concatenate(hdr_all)
    v["huff_hdr_bits"] = np.array(hdr_bits, np.uint64)

    # Context maps (context_map.cc): the fixed AC map and an MTF known answer.
    cmap = v["ac_cluster"].astype(np.uint8)
    out = np.zeros(65536, np.uint8)
    nb = C.c_size_t()
    enc = C.c_int()
    n = R.ref_best_context_map(P(cmap), cmap.size, 32, P(out), out.size, C.byref(nb),
                               C.byref(enc))
    v["cmap_ac_bytes"], v["cmap_ac_bits"] = out[:n].copy(), np.array([nb.value, enc.value])
    mtf_in = np.array([0, 1, 1, 0, 2, 1, 2, 0], np.uint8)
    mtf_out = np.zeros(8, np.uint8)
    R.ref_move_to_front(P(mtf_in), 8, P(mtf_out))
    v["mtf_out"] = mtf_out

    # Entropy sections on the reference tests' frames.
    def fc_arrays(fc):
        return (np.ascontiguousarray(fc.dc.ravel()), np.ascontiguousarray(fc.ac.ravel()),
                fc.acs, fc.ytox, fc.ytob)

    for name, fc in (("ac_dct8_256", patterns.dct8_ac_frame(256, 256)),
                     ("ac_mixed_256", patterns.mixed_ac_frame(256, 256)),
                     ("ac_mixed_512x384", patterns.mixed_ac_frame(512, 384))):
        dc, ac, acs, _, _ = fc_arrays(fc)
        hist = np.zeros(32 * 256, np.uint32)
        depth = np.zeros(32 * 256, np.uint8)
        bits = np.zeros(32 * 256, np.uint16)
        ng = ((fc.bw + 31) // 32) * ((fc.bh + 31) // 32)
        sizes = np.zeros(ng, np.uint32)
        buf = np.zeros(8 << 20, np.uint8)
        n = R.ref_ac_encode(fc.w, fc.h, P(dc), P(ac), P(acs), P(hist), P(depth), P(bits),
                            P(sizes), P(buf), buf.size)
        v[f"{name}_hist"], v[f"{name}_depth"], v[f"{name}_bits"] = hist, depth, bits
        v[f"{name}_sizes"], v[f"{name}_sha"] = sizes, sha(buf[:n].tobytes())

    for name, fc in (("dc_dct8_256", patterns.dct8_dc_frame(256, 256)),
                     ("dc_mixed_256", patterns.mixed_dc_frame(256, 256)),
                     ("dc_mixed_2560x256", patterns.mixed_dc_frame(2560, 256)),
                     ("dc_mixed_2560x2304", patterns.mixed_dc_frame(2560, 2304))):
        dc, _, acs, ytox, ytob = fc_arrays(fc)
        ng = ((fc.bw + 255) // 256) * ((fc.bh + 255) // 256)
        sizes = np.zeros(ng, np.uint32)
        buf = np.zeros(16 << 20, np.uint8)
        n = R.ref_dc_encode(fc.w, fc.h, fc.rqf, P(dc), P(acs), P(ytox), P(ytob), P(sizes),
                            P(buf), buf.size)
        v[f"{name}_sizes"], v[f"{name}_sha"] = sizes, sha(buf[:n].tobytes())

    # Whole codestreams (vardct_frame.cc write_vardct_codestream, clustered).
    for name, fc in (("cs_asm_512", patterns.assembly_frame(512, 512)),
                     ("cs_asm_2560x256", patterns.assembly_frame(2560, 256)),
                     ("cs_asm_640x384", patterns.assembly_frame(640, 384)),
                     ("cs_asm_256", patterns.assembly_frame(256, 256)),
                     ("cs_mixed_512x384", patterns.mixed_ac_frame(512, 384)),
                     ("cs_mixeddc_2560x2304", patterns.mixed_dc_frame(2560, 2304))):
        dc, ac, acs, ytox, ytob = fc_arrays(fc)
        buf = np.zeros(64 << 20, np.uint8)
        n = R.ref_write_codestream(fc.w, fc.h, fc.gs, fc.qdc, fc.rqf, P(dc), P(ac), P(acs),
                                   P(ytox), P(ytob), 1, P(buf), buf.size)
        v[f"{name}_len"], v[f"{name}_sha"] = np.array([n], np.uint64), sha(buf[:n].tobytes())
        if n < 4096:
            v[f"{name}_bytes"] = buf[:n].copy()

    # Globals, head, container.
    out = np.zeros(64, np.uint8)
    n = R.ref_dc_global(4096, 32, P(out), 64)
    v["dc_global_4096_32"] = out[:n].copy()
    sizes = np.array([5, 1000, 2000, 70000, 30, 123456, 5000000], np.uint32)
    out = np.zeros(4096, np.uint8)
    n = R.ref_codestream_head(3840, 2160, P(sizes), sizes.size, P(out), out.size)
    v["head_sizes"], v["head_3840x2160"] = sizes, out[:n].copy()
    n = R.ref_codestream_head(256, 256, P(sizes[:1]), 1, P(out), out.size)
    v["head_256"] = out[:n].copy()
    for cs in (1234, 0x100000000 - 8):
        out = np.zeros(64, np.uint8)
        n = R.ref_container_framing(cs, P(out), 64)
        v[f"container_{cs}"] = out[:n].copy()

    # Selection (decide_region) on the transform_select_gpu_test.cc planes.
    def sel(y, w, h, d=1.0):
        acs = np.full((w // 8) * (h // 8), -1, np.int8)
        R.ref_select_transforms(P(y), w, h, d, P(acs))
        return acs

    py, px = np.mgrid[0:64, 0:96]
    y = np.where(px < 48, 0.2 + 0.3 * ((px + py) / (96 + 64)),
                 0.5 + 0.4 * np.where((px + py) & 1, 1.0, -1.0)).astype(np.float32)
    v["sel_texture_96x64"] = sel(np.ascontiguousarray(y), 96, 64)
    v["sel_flat_64"] = sel(np.full(64 * 64, 0.37, np.float32), 64, 64)
    i = np.arange(88 * 72)
    y = (0.3 + 0.2 * ((i * 7) % 13) / 13.0).astype(np.float32)
    v["sel_edge_88x72"] = sel(y, 88, 72)
    # ... and on the sharpened Y of seeded synthetic frames (the oracle's
    # NV12 -> XYB and gaborish, themselves pinned by closed forms).
    from cujpegxl_amd.synth import synthetic_nv12

    for w, h, seed, d in ((264, 200, 3, 1.0), (264, 200, 3, 3.0), (776, 456, 21, 1.0), (776, 456, 22, 3.0),
                          (1280, 720, 23, 1.0)):
        luma, chroma = synthetic_nv12(w, h, seed)
        yp = np.ascontiguousarray(oracle.gaborish(oracle.nv12_to_xyb(luma, chroma, w, h))[1])
        v[f"sel_synth_{w}x{h}_s{seed}_d{d:g}"] = sel(yp, w, h, d)

    # CfL map quantization (cfl.h) and LLF -> DC (vardct_layout.h).
    sums = rng.normal(0, 1, (64, 3))
    sums[:, 1] = np.abs(sums[:, 1]) + 1e-3
    sums[0] = (0.0, 0.0, 0.0)
    maps = np.zeros((64, 2), np.int32)
    for k in range(64):
        mx, mb = C.c_int(), C.c_int()
        R.ref_cfl_maps_from_sums(*map(float, sums[k]), C.byref(mx), C.byref(mb))
        maps[k] = (mx.value, mb.value)
    v["cfl_sums"], v["cfl_maps"] = sums, maps
    for m in (2, 4):
        llf = rng.uniform(-1, 1, (32, m * m)).astype(np.float32)
        dc = np.zeros((32, m * m), np.float32)
        for k in range(32):
            R.ref_dc_from_llf(m, P(llf[k]), P(dc[k]))
        v[f"llf{m}_in"], v[f"llf{m}_out"] = llf, dc

    np.savez(OUT / "ref_vectors.npz", **v)
    print(f"wrote {len(v)} vectors to {OUT / 'ref_vectors.npz'}")


if __name__ == "__main__":
    main()
