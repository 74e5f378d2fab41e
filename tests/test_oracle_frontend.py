"""CPU: the oracle's float front end against closed forms, FP64 restatements,
the reference's own host code (golden vectors / oracle/_ref) and the
invariants the mathematics fixes."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cujpegxl_amd.synth import synthetic_nv12


def _edges(lo_exp: int, hi_exp: int) -> np.ndarray:
    """Both sides of every table segment edge (16 per binade) in
    [2^lo_exp, 2^hi_exp)."""
    out = []
    for e in range(lo_exp, hi_exp):
        for j in range(16):
            v = np.float32(np.ldexp(1.0 + j / 16.0, e))
            out += [v, np.nextafter(v, np.float32(np.inf)), np.nextafter(v, np.float32(0))]
    out = np.array(out, dtype=np.float32)
    return out[(out >= np.float32(2.0 ** lo_exp)) & (out < np.float32(2.0 ** hi_exp))]


def test_numeric_primitives_error_bounds(jo):
    L = jo.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0.0037, 1.2, 4000), rng.uniform(1e-3, 50, 2000)])
    for x in xs.astype(np.float32):
        assert abs(L.jo_log2(float(x)) - np.log2(float(x))) <= 2e-7 * max(1.0, abs(np.log2(float(x))))
    for x in rng.uniform(-20, 20, 4000).astype(np.float32):
        assert abs(L.jo_exp2(float(x)) - 2.0 ** float(x)) <= 3e-7 * 2.0 ** float(x)
    # The table-driven powers over their whole domains: the XYB conversion
    # calls x^2.4 on [0.0905, 1] and x^(1/3) on [0.0038, 1.0038].
    cb = np.concatenate([np.exp2(rng.uniform(-9, 1, 6000)).astype(np.float32), _edges(-9, 1),
                         np.float32([0.0037930732552754493, 1.0037930732552754, 1.0])])
    for x in cb:
        want = np.cbrt(float(x))
        assert abs(L.jo_cbrt(float(x)) - want) <= 5e-7 * want, x
    pw = np.concatenate([np.exp2(rng.uniform(-4, 1, 6000)).astype(np.float32), _edges(-4, 1),
                         np.float32([(0.04045 + 0.055) / 1.055, 1.0])])
    for x in pw:
        want = float(x) ** 2.4
        assert abs(L.jo_pow24(float(x)) - want) <= 8e-7 * want, x


def test_power_table_is_rounded_powers_of_segment_centres(jo):
    """The 272-float table: centres c_j = 1 + (j + 1/2)/16, RN(1/c_j), then
    RN((2^k c_j)^2.4) for k = -4..0 and RN((2^k c_j)^(1/3)) for k = -9..0."""
    t = np.zeros(272, np.float32)
    jo.lib().jo_xyb_table(t.ctypes.data_as(C.POINTER(C.c_float)))
    c = 1.0 + (np.arange(16) + 0.5) / 16.0
    assert np.array_equal(t[0:32:2], c.astype(np.float32))
    assert np.array_equal(t[1:32:2], (1.0 / c).astype(np.float32))
    pw = np.array([[(c[j] * 2.0 ** k) ** 2.4 for j in range(16)] for k in range(-4, 1)])
    cb = np.array([[np.cbrt(c[j] * 2.0 ** k) for j in range(16)] for k in range(-9, 1)])
    assert np.allclose(t[32:112], pw.ravel().astype(np.float32), rtol=6e-8, atol=0)
    assert np.allclose(t[112:272], cb.ravel().astype(np.float32), rtol=6e-8, atol=0)


def test_power_series_argument_from_the_integer_mantissa(jo):
    """The device forms the series argument t = (m - c_j) r_j of the oracle's
    powers as (x - 2^18) (r_j 2^-23), x the low 19 mantissa bits: the same
    real product, so the same float for every one of the 2^23 mantissas."""
    t = np.zeros(272, np.float32)
    jo.lib().jo_xyb_table(t.ctypes.data_as(C.POINTER(C.c_float)))
    centre, recip = t[0:32:2], t[1:32:2]
    u = np.arange(1 << 23, dtype=np.uint32) | np.uint32(0x3F800000)
    j = (u >> np.uint32(19)) & np.uint32(15)
    m = u.view(np.float32)
    oracle_t = (m - centre[j]) * recip[j]
    scaled = np.ldexp(recip, -23).astype(np.float32)
    assert np.array_equal(np.ldexp(scaled.astype(np.float64), 23).astype(np.float32), recip)
    device_t = ((u & np.uint32(0x7FFFF)).astype(np.float32) - np.float32(262144.0)) * scaled[j]
    assert np.array_equal(oracle_t.view(np.uint32), device_t.view(np.uint32))


def test_f16_conversion_matches_ieee_rne(jo):
    L = jo.lib()
    rng = np.random.default_rng(2)
    vals = np.concatenate([
        rng.normal(0, 1e-5, 3000), rng.normal(0, 1e-3, 3000), rng.normal(0, 1, 3000),
        np.array([0.0, -0.0, 6.1035156e-05, 5.9604645e-08, 2.98e-8, 65504.0, 65519.0, 1e-9]),
    ]).astype(np.float32)
    for v in vals:
        h = L.jo_f32_to_f16(float(v))
        assert h == int(np.float32(v).astype(np.float16).view(np.uint16)), v
        assert L.jo_f16_to_f32(h) == float(np.float16(v))


def _nv12(w, h, seed):
    luma, chroma = synthetic_nv12(w, h, seed)
    return np.ascontiguousarray(luma), np.ascontiguousarray(chroma)


def test_xyb_against_fp64_and_flat_tile(jo):
    luma, chroma = _nv12(128, 96, 7)
    got = jo.nv12_to_xyb(luma, chroma, 128, 96)
    want = jo.nv12_to_xyb_f64(luma, chroma, 128, 96)
    # 1e-5 relative on the magnitude scale of each channel's values.
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
    # xyb_conformance_test.py:60-64 flat tile: the chroma upsample is exact.
    luma = np.full((64, 64), 140, np.uint8)
    chroma = np.empty((32, 64), np.uint8)
    chroma[:, 0::2], chroma[:, 1::2] = 100, 170
    got = jo.nv12_to_xyb(luma, chroma, 64, 64)
    assert np.ptp(got.reshape(3, -1), axis=1).max() == 0.0
    want = jo.nv12_to_xyb_f64(luma, chroma, 64, 64)
    assert np.max(np.abs(got - want)) < 3e-7
    # Grey (Cb = Cr = 128) maps to X = 0 (opsin rows 0 and 1 are equal for grey).
    chroma[:] = 128
    assert np.abs(jo.nv12_to_xyb(luma, chroma, 64, 64)[0]).max() < 1e-7


def _xyb_constants():
    import json
    import pathlib

    path = pathlib.Path(__file__).resolve().parent / "golden" / "xyb_constants.json"
    consts = json.loads(path.read_text())["constants"]
    return {k: float(np.float32(v["f32"])) for k, v in consts.items() if "f32" in v}


def _bilinear_f64(plane, w, h):
    """tests/reference.py:23-46: centred 4:2:0 bilinear with clamped taps."""
    ch, cw = plane.shape
    xs = (np.arange(w) + 0.5) / 2.0 - 0.5
    ys = (np.arange(h) + 0.5) / 2.0 - 0.5
    x0, y0 = np.floor(xs), np.floor(ys)
    fx, fy = xs - x0, ys - y0
    x0i, x1i = np.clip(x0.astype(int), 0, cw - 1), np.clip(x0.astype(int) + 1, 0, cw - 1)
    y0i, y1i = np.clip(y0.astype(int), 0, ch - 1), np.clip(y0.astype(int) + 1, 0, ch - 1)
    p = plane.astype(np.float64)
    top = p[y0i][:, x0i] * (1 - fx) + p[y0i][:, x1i] * fx
    bot = p[y1i][:, x0i] * (1 - fx) + p[y1i][:, x1i] * fx
    return top * (1 - fy)[:, None] + bot * fy[:, None]


def _xyb_f64_from_constants(luma, chroma, w, h, k):
    """NV12 -> XYB in float64 with the reference's constants `k`
    (xyb_impl.cuh:17-70): BT.709 full-range inverse, the sRGB EOTF, the opsin
    matrix with bias, cube root minus cbrt(bias)."""
    cb = _bilinear_f64(chroma[:, 0::2], w, h) / 255.0 - 128.0 / 255.0
    cr = _bilinear_f64(chroma[:, 1::2], w, h) / 255.0 - 128.0 / 255.0
    y = luma.astype(np.float64) / 255.0
    rgb = (y + k["CR_TO_R"] * cr, y - k["CB_TO_G"] * cb - k["CR_TO_G"] * cr, y + k["CB_TO_B"] * cb)
    lin = []
    for e in rgb:
        e = np.clip(e, 0.0, 1.0)
        lin.append(np.clip(np.where(e <= 0.04045, e / 12.92, ((e + 0.055) / 1.055) ** 2.4), 0.0, 1.0))
    rows = ((k["M00"], 1.0 - k["M02"] - k["M00"], k["M02"]), (k["M10"], 1.0 - k["M12"] - k["M10"], k["M12"]),
            (k["M20"], k["M21"], 1.0 - k["M20"] - k["M21"]))
    bias = k["OPSIN_BIAS"]
    q = [np.cbrt(np.maximum(a * lin[0] + b * lin[1] + c * lin[2] + bias, 0.0)) - np.cbrt(bias) for a, b, c in rows]
    return np.stack([0.5 * (q[0] - q[1]), 0.5 * (q[0] + q[1]), q[2]])


def _xyb_error(got, want):
    """Max error per channel relative to the channel's largest magnitude."""
    return [float(np.abs(got[c] - want[c]).max() / np.abs(want[c]).max()) for c in range(3)]


_XYB_BOUND = (1e-5, 1e-6, 1e-6)


def _xyb_within(err):
    return all(e <= b for e, b in zip(err, _XYB_BOUND))


def test_xyb_against_reference_constants(jo):
    """The oracle's NV12 -> XYB against a float64 recomputation from the
    constants written in the reference's source (tests/golden/xyb_constants.json,
    extracted from src/xyb_impl.cuh:17-35 by extract_constants.py), on a
    synthetic frame and on uniformly random NV12 (saturated colours reach every
    clamp).  Bounds, relative to each channel's largest magnitude: X 1e-5
    (measured 5e-6 on the synthetic frame, where |X| < 0.01), Y and B 1e-6
    (measured 2.5e-7).  Every constant matters at these bounds: perturbing any
    one of them by 1e-4 relative in the float64 recomputation breaks the
    agreement on the random frame."""
    k = _xyb_constants()
    rng = np.random.default_rng(0)
    cases = [_nv12(128, 96, 7) + (128, 96)]
    luma = rng.integers(0, 256, (48, 64), dtype=np.uint8)
    chroma = rng.integers(0, 256, (24, 64), dtype=np.uint8)
    cases.append((luma, chroma, 64, 48))
    for luma, chroma, w, h in cases:
        got = jo.nv12_to_xyb(luma, chroma, w, h).reshape(3, h, w).astype(np.float64)
        want = _xyb_f64_from_constants(luma, chroma, w, h, k)
        assert _xyb_within(_xyb_error(got, want)), _xyb_error(got, want)
    luma, chroma, w, h = cases[1]
    got = jo.nv12_to_xyb(luma, chroma, w, h).reshape(3, h, w).astype(np.float64)
    for name in k:
        bad = dict(k)
        bad[name] *= 1.0 + 1e-4
        assert not _xyb_within(_xyb_error(got, _xyb_f64_from_constants(luma, chroma, w, h, bad))), name


def test_bt709_constants_from_kr_kb():
    """The reference's BT.709 inverse coefficients (xyb_impl.cuh:27-33) are
    the full-range inverse for Kr = 0.2126, Kb = 0.0722 (its comment):
    R = Y + 2(1-Kr) Cr, B = Y + 2(1-Kb) Cb, G = Y - 2Kb(1-Kb)/Kg Cb - 2Kr(1-Kr)/Kg Cr."""
    k = _xyb_constants()
    kr, kb = 0.2126, 0.0722
    kg = 1.0 - kr - kb
    derived = {"CR_TO_R": 2 * (1 - kr), "CB_TO_B": 2 * (1 - kb), "CB_TO_G": 2 * kb * (1 - kb) / kg,
               "CR_TO_G": 2 * kr * (1 - kr) / kg}
    for name, v in derived.items():
        assert k[name] == float(np.float32(v)), (name, k[name], v)
    # Forward BT.709 then this inverse returns the RGB it started from.
    rgb = np.random.default_rng(3).uniform(0, 1, (3, 100))
    yy = kr * rgb[0] + kg * rgb[1] + kb * rgb[2]
    cb = (rgb[2] - yy) / (2 * (1 - kb))
    cr = (rgb[0] - yy) / (2 * (1 - kr))
    back = np.stack([yy + derived["CR_TO_R"] * cr, yy - derived["CB_TO_G"] * cb - derived["CR_TO_G"] * cr,
                     yy + derived["CB_TO_B"] * cb])
    assert np.abs(back - rgb).max() < 1e-12


def test_chroma_upsample_is_the_exact_bilinear(jo):
    """tests/reference.py:23-46: centred bilinear 4:2:0 -> 4:4:4, each tap
    clamped to the edge independently, evaluated here in float64."""
    rng = np.random.default_rng(6)
    w, h = 22, 14
    chroma = rng.integers(0, 256, (h // 2, w), dtype=np.uint8)
    cb = np.zeros((h, w), np.float32)
    cr = np.zeros((h, w), np.float32)
    jo.lib().jo_chroma_upsample(jo.ptr(chroma), w, w, h, jo.ptr(cb), jo.ptr(cr))

    def upsample(plane):
        ch, cw = plane.shape
        xs = (np.arange(w) + 0.5) / 2.0 - 0.5
        ys = (np.arange(h) + 0.5) / 2.0 - 0.5
        x0, y0 = np.floor(xs), np.floor(ys)
        fx, fy = xs - x0, ys - y0
        x0i, x1i = np.clip(x0.astype(int), 0, cw - 1), np.clip(x0.astype(int) + 1, 0, cw - 1)
        y0i, y1i = np.clip(y0.astype(int), 0, ch - 1), np.clip(y0.astype(int) + 1, 0, ch - 1)
        p = plane.astype(np.float64)
        top = p[y0i][:, x0i] * (1 - fx) + p[y0i][:, x1i] * fx
        bot = p[y1i][:, x0i] * (1 - fx) + p[y1i][:, x1i] * fx
        return top * (1 - fy)[:, None] + bot * fy[:, None]

    np.testing.assert_array_equal(cb, upsample(chroma[:, 0::2]))  # exact in float
    np.testing.assert_array_equal(cr, upsample(chroma[:, 1::2]))


def test_chroma_integer_form_gives_the_float_bits(jo):
    """The device's form of the bilinear: 9 s_kk + 3 s_kn + 3 s_nk + s_nn over
    the pixel's own chroma cell k = x >> 1 and its clamped neighbour n towards
    the pixel, in each direction, then the 8-bit scale with the 1/16 folded
    in.  Checked against the oracle's float bilinear on odd and even sizes
    and every clamped edge, and the scale over every possible sum."""
    rng = np.random.default_rng(61)
    for w, h in ((22, 14), (8, 8), (40, 18)):
        chroma = rng.integers(0, 256, (h // 2, w), dtype=np.uint8)
        chroma[0, :4] = 255
        cb = np.zeros((h, w), np.float32)
        cr = np.zeros((h, w), np.float32)
        jo.lib().jo_chroma_upsample(jo.ptr(chroma), w, w, h, jo.ptr(cb), jo.ptr(cr))

        def taps(n, size):
            x = np.arange(n)
            k = x >> 1
            other = np.where(x & 1, np.minimum(k + 1, size - 1), np.maximum(k - 1, 0))
            return k, other

        kx, nx = taps(w, w // 2)
        ky, ny = taps(h, h // 2)
        for plane, out in ((chroma[:, 0::2], cb), (chroma[:, 1::2], cr)):
            s = plane.astype(np.int64)
            v = (9 * s[ky][:, kx] + 3 * s[ky][:, nx] + 3 * s[ny][:, kx] + s[ny][:, nx])
            np.testing.assert_array_equal(v.astype(np.float32) * np.float32(0.0625), out)
    k = np.float32(1.0) / np.float32(255.0)
    c = np.float32(128.0) / np.float32(255.0)
    v = np.arange(0, 16 * 255 + 1, dtype=np.float32)
    folded = v * np.float32(k * np.float32(0.0625)) - c
    direct = (v * np.float32(0.0625)) * k - c
    np.testing.assert_array_equal(folded.view(np.uint32), direct.view(np.uint32))


def test_gaborish_against_bruteforce(jo):
    rng = np.random.default_rng(5)
    x = rng.uniform(-0.1, 0.9, (3, 40, 56)).astype(np.float32)
    got = jo.gaborish(x)
    k = [-0.09495815671340026, -0.041031725066768575, 0.013710004822696948,
         0.006510206083837737, -0.0014789063378272242]
    wgt = np.zeros((5, 5))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            a, b = sorted((abs(dx), abs(dy)))
            wgt[dy + 2, dx + 2] = {(0, 0): 1, (0, 1): k[0], (0, 2): k[2], (1, 1): k[1],
                                   (1, 2): k[3], (2, 2): k[4]}[(a, b)]
    wgt /= wgt.sum()
    pad = np.pad(x.astype(np.float64), ((0, 0), (2, 2), (2, 2)), mode="edge")
    want = np.zeros_like(pad[:, 2:-2, 2:-2])
    for dy in range(5):
        for dx in range(5):
            want += wgt[dy, dx] * pad[:, dy : dy + 40, dx : dx + 56]
    assert np.max(np.abs(got - want)) < 2e-6
    flat = np.full((3, 16, 16), 0.25, np.float32)
    assert np.max(np.abs(jo.gaborish(flat) - 0.25)) < 1e-6  # weights sum to one


def _aq_block_f64(xyb, bx, by, w, h, d):
    """FP64 restatement of block (bx, by)'s AQ integer from
    frontend_fused.cu:54-204: the block's 32x32 tile (4x4 blocks) with a
    one-pixel halo taken from the frame (coordinates clamped to the frame
    only), the tile's 8x8 pre-erosion cells, erosion over the 3x3 cell
    neighbourhood clamped to the tile, then the block's modulation terms."""
    import math

    def ratio(v, inv):
        v = max(v, 0.0)
        mul, mul2, lg = 226.77216153508914, 1 / 73.377132366608819, 0.693147181
        ret = mul2 * 18.6580932135 * lg
        num = ret * 3 * mul * v * v + 1e-2
        den = lg * mul * v ** 3 + 7.7825991679894591 * lg + 1e-2
        return num / den if inv else den / num

    tx0, ty0 = (bx // 4) * 32, (by // 4) * 32
    xs = np.clip(np.arange(tx0 - 1, tx0 + 33), 0, w - 1)
    ys = np.clip(np.arange(ty0 - 1, ty0 + 33), 0, h - 1)
    ty = xyb[1].astype(np.float64)[ys][:, xs]
    pre = np.zeros((8, 8))
    for cy in range(8):
        for cx in range(8):
            s = 0.0
            for dy in range(4):
                for dx in range(4):
                    lx, ly = cx * 4 + dx + 1, cy * 4 + dy + 1
                    yv = ty[ly, lx]
                    base = 0.25 * (ty[ly + 1, lx] + ty[ly - 1, lx] + ty[ly, lx - 1] + ty[ly, lx + 1])
                    diff = min((ratio(yv + 0.019, False) * (yv - base)) ** 2, 0.2)
                    s += 0.25 * math.sqrt(diff * math.sqrt(211.66567973503678e8) + 27.505837037000106)
            pre[cy, cx] = 0.25 * s
    efm = (2 - d) * 0.5 if d < 2 else 0.0
    km = np.array([0.125, 0.10 - efm * 0.10, 0.09 - efm * 0.09, 0.06 - efm * 0.06])
    km *= 0.29959705784054957 / km.sum()
    er = 0.0
    lbx, lby = bx % 4, by % 4
    for cy in (2 * lby, 2 * lby + 1):
        for cx in (2 * lbx, 2 * lbx + 1):
            nb = sorted(pre[min(max(cy + a, 0), 7), min(max(cx + b, 0), 7)]
                        for a in (-1, 0, 1) for b in (-1, 0, 1))
            er += float(np.dot(km, nb[:4]))
    yv = xyb[1, by * 8 : by * 8 + 8, bx * 8 : bx * 8 + 8].astype(np.float64)
    xv = xyb[0, by * 8 : by * 8 + 8, bx * 8 : bx * 8 + 8].astype(np.float64)
    bv = xyb[2, by * 8 : by * 8 + 8, bx * 8 : bx * 8 + 8].astype(np.float64)
    hf = np.minimum(0.0206, np.abs(np.diff(yv, axis=1))).sum() + \
        np.minimum(0.0206, np.abs(np.diff(yv, axis=0))).sum()
    gam = sum(ratio(a + 0.16 - b, True) + ratio(a + 0.16 + b, True)
              for a, b in zip(yv.ravel(), xv.ravel())) * 0.5 / 64
    pye = yv + 0.084381641171960495 + np.abs(xv)
    lim = 0.027121074570634722
    blue = np.where(bv > pye, np.minimum(bv - pye, lim), 0).sum()
    if blue >= 32 * lim:
        blue = 64 * lim - blue
    blue = min(blue, 15.398788439047934 * lim)
    v1 = max(er * 0.80061762862741759, 1e-3)
    mask = (-0.7647 + 9.4708735624378946 / (v1 * v1 + 0.25 * 3.7179635626140772)
            + 17.35036561631863 / (v1 + 302.59587815579727)
            + 6.7943250517376494 / (v1 * v1 + 3.7179635626140772))
    ov = mask + hf * -0.38 + 0.42 + 0.1005613337192697 * math.log2(gam) + blue * 0.14207000358439159
    scale = 0.725 / d
    damp = 1.0 if d < 2 else max(0.0, 1 - (d - 2) / 12)
    q = math.exp(ov) * scale * damp + (1 - damp) * 0.48 * scale
    return q


@pytest.mark.parametrize("d,w,h,seed", [(1.0, 32, 32, 9), (3.0, 32, 32, 9), (1.0, 96, 64, 12),
                                        (3.0, 96, 64, 13), (1.5, 80, 56, 14)])
def test_adaptive_quant_against_fp64(jo, d, w, h, seed):
    """Every block of the frame, so interior tiles (whose halo pixels come
    from neighbouring tiles) and partial tiles at the right and bottom edges
    (80x56: 10x7 blocks) are covered, besides the all-clamped single tile."""
    luma, chroma = _nv12(w, h, seed)
    xyb = jo.nv12_to_xyb(luma, chroma, w, h).reshape(3, h, w)
    qf = jo.adaptive_quant(np.ascontiguousarray(xyb), d)
    inv = jo.calibrate(d).inv_global_scale
    bw, bh = w // 8, h // 8
    checked = 0
    for by in range(bh):
        for bx in range(bw):
            want = _aq_block_f64(xyb, bx, by, w, h, d) * inv + 0.5
            got = int(qf[by * bw + bx])
            if abs(want - round(want)) > 1e-3:  # not at a rounding tie
                assert got == min(256, max(1, int(want))), (bx, by, want)
                checked += 1
    assert checked >= bw * bh - 2


def test_adaptive_quant_flat_and_range(jo):
    luma = np.full((64, 96), 100, np.uint8)
    chroma = np.full((32, 96), 128, np.uint8)
    xyb = jo.nv12_to_xyb(luma, chroma, 96, 64)
    qf = jo.adaptive_quant(xyb, 1.0)
    assert np.ptp(qf) == 0  # a flat frame modulates every block alike
    luma, chroma = _nv12(256, 128, 4)
    qf = jo.adaptive_quant(jo.nv12_to_xyb(luma, chroma, 256, 128), 1.0)
    assert qf.min() >= 1 and qf.max() <= 256 and np.ptp(qf) > 0


def test_selection_matches_reference(jo, golden):
    """transform_select_gpu_test.cc:67-118 planes against the reference's host
    decide_region; the flat plane picks all DCT32."""
    py, px = np.mgrid[0:64, 0:96]
    y = np.where(px < 48, 0.2 + 0.3 * ((px + py) / (96 + 64)),
                 0.5 + 0.4 * np.where((px + py) & 1, 1.0, -1.0)).astype(np.float32)
    np.testing.assert_array_equal(jo.select_transforms(y, 1.0), golden["sel_texture_96x64"])
    flat = np.full((64, 64), 0.37, np.float32)
    acs = jo.select_transforms(flat, 1.0)
    np.testing.assert_array_equal(acs, golden["sel_flat_64"])
    assert acs.reshape(8, 8)[::4, ::4].tolist() == [[32, 32], [32, 32]]
    i = np.arange(88 * 72)
    y = (0.3 + 0.2 * ((i * 7) % 13) / 13.0).astype(np.float32).reshape(72, 88)
    np.testing.assert_array_equal(jo.select_transforms(y, 1.0), golden["sel_edge_88x72"])


def _valid_tiling(acs, bw, bh):
    owner = np.full(bw * bh, -1)
    for by in range(bh):
        for bx in range(bw):
            s = int(acs[by * bw + bx])
            if s == 0:
                continue
            assert s in (8, 16, 32)
            k = s // 8
            assert bx + k <= bw and by + k <= bh
            for dy in range(k):
                for dx in range(k):
                    assert owner[(by + dy) * bw + bx + dx] == -1
                    owner[(by + dy) * bw + bx + dx] = by * bw + bx
    assert (owner >= 0).all()


def test_selection_on_synthetic_frames(jo, golden):
    """Real content: valid tilings, all three sizes used, and every decision
    equal to the reference's host decide_region (transform_select_impl.cuh:130-201,
    compiled from its sources into oracle/_ref) on the same sharpened Y plane:
    the golden vectors hold its output for five seeded frames at d = 1 and 3,
    and where _ref is built it is run again on the first frame."""
    cases = ((264, 200, 3, 1.0), (264, 200, 3, 3.0), (776, 456, 21, 1.0), (776, 456, 22, 3.0),
             (1280, 720, 23, 1.0))
    for w, h, seed, d in cases:
        luma, chroma = _nv12(w, h, seed)
        y = np.ascontiguousarray(jo.gaborish(jo.nv12_to_xyb(luma, chroma, w, h))[1])
        acs = jo.select_transforms(y, d)
        _valid_tiling(acs, w // 8, h // 8)
        assert {8, 16, 32} <= set(np.unique(acs).tolist())
        np.testing.assert_array_equal(acs, golden[f"sel_synth_{w}x{h}_s{seed}_d{d:g}"])
    R = jo.ref()
    if R is not None:
        luma, chroma = _nv12(264, 200, 3)
        y = np.ascontiguousarray(jo.gaborish(jo.nv12_to_xyb(luma, chroma, 264, 200))[1])
        want = np.full(y.size // 64, -1, np.int8)
        R.ref_select_transforms(jo.ptr(y), 264, 200, 1.0, jo.ptr(want))
        np.testing.assert_array_equal(jo.select_transforms(y, 1.0), want)


def test_dct_against_fp64(jo):
    rng = np.random.default_rng(8)
    x = rng.uniform(-0.2, 1.0, (3, 64, 64)).astype(np.float32)
    for acs_side in (8, 16, 32):
        acs = np.zeros(64, np.int8)
        for by in range(0, 8, acs_side // 8):
            for bx in range(0, 8, acs_side // 8):
                acs[by * 8 + bx] = acs_side
        got = jo.forward_dct(x, acs)
        want = jo.forward_dct_f64(x, acs)
        err = np.abs(got - want)
        assert err.max() <= 1e-5 * max(1.0, np.abs(want).max()), acs_side
    # Orthogonality: the DC of a constant block is the constant (1/N scaling).
    c = jo.forward_dct(np.full((3, 32, 32), 0.42, np.float32), np.array(
        [32] + [0] * 15, np.int8))
    assert abs(c[0] - 0.42) < 1e-6 and np.abs(c[1:1024]).max() < 1e-6


def test_cfl_maps_and_llf(jo, golden):
    R = jo.ref()
    sums = golden["cfl_sums"]
    # jo_cfl_maps reduces per-region partials; feed each sum as one region.
    for k in range(sums.shape[0]):
        part = np.zeros((1, 3), np.float32)
        part[0] = sums[k]
        ytox, ytob = np.zeros(1, np.int8), np.zeros(1, np.int8)
        jo.lib().jo_cfl_maps(jo.ptr(part), 32, 32, jo.ptr(ytox), jo.ptr(ytob))
        if R is not None:
            mx, mb = C.c_int(), C.c_int()
            R.ref_cfl_maps_from_sums(*map(float, part[0]), C.byref(mx), C.byref(mb))
            assert (ytox[0], ytob[0]) == (mx.value, mb.value)
    assert (ytox[0], ytob[0]) == (golden["cfl_maps"][-1][0], golden["cfl_maps"][-1][1]) or R
    # Degenerate tile maps to the base (cfl_gpu_test.cc:115-124).
    part = np.zeros((1, 3), np.float32)
    ytox, ytob = np.ones(1, np.int8), np.ones(1, np.int8)
    jo.lib().jo_cfl_maps(jo.ptr(part), 32, 32, jo.ptr(ytox), jo.ptr(ytob))
    assert (ytox[0], ytob[0]) == (0, 0)
    for m in (2, 4):
        for k in range(golden[f"llf{m}_in"].shape[0]):
            llf = np.ascontiguousarray(golden[f"llf{m}_in"][k])
            dc = np.zeros(m * m, np.float32)
            jo.lib().jo_dc_from_llf(m, jo.ptr(llf), jo.ptr(dc))
            want = golden[f"llf{m}_out"][k]
            assert np.all(np.abs(dc - want) <= np.spacing(np.abs(want)) * 1.01)
    # llf_dc_conformance_test.py:74-82: a constant block's DCs equal the constant.
    for n in (16, 32):
        m = n // 8
        acs = np.zeros(m * m, np.int8)
        acs[0] = n
        coef = jo.forward_dct(np.full((3, n, n), 0.42, np.float32), acs)
        llf = np.array([coef[(e // 64) * 64 + e % 64] for e in
                        [fx * n + fy for fx in range(m) for fy in range(m)]], np.float32)
        dc = np.zeros(m * m, np.float32)
        jo.lib().jo_dc_from_llf(m, jo.ptr(llf), jo.ptr(dc))
        assert np.max(np.abs(dc - 0.42)) < 2e-6


def test_quantization_is_rint_of_coeff_over_step(jo):
    """quant_selfconsistency_test.py:127-179: every quantized integer equals
    rint(residual / step) in float64 with the calibrated step (ties excepted),
    and the CfL / base-correlation residuals are the decoder's."""
    w8, w16, w32 = jo.dequant_weights()
    for seed, d, (w, h) in ((4, 1.0, (64, 64)), (5, 3.0, (96, 64))):
        luma, chroma = _nv12(w, h, seed)
        xyb = jo.gaborish(jo.nv12_to_xyb(luma, chroma, w, h))
        acs = jo.select_transforms(xyb[1], d)
        c16 = jo.to_f16(jo.forward_dct(xyb, acs))
        qf = jo.adaptive_quant(jo.nv12_to_xyb(luma, chroma, w, h), d)
        _, ytox, ytob = jo.cfl(c16, acs, w, h)
        ac, dc, _ = jo.quantize(c16, acs, ytox, ytob, qf, w, h, d)
        cal = jo.calibrate(d)
        coef = c16.view(np.float16).astype(np.float64).reshape(3, -1)
        bw = w // 8
        ties = checked = 0
        for blk in range(acs.size):
            n = int(acs[blk])
            if n == 0:
                continue
            bx, by = blk % bw, blk // bw
            qgsf = float(qf[blk]) * cal.global_scale_float
            t = (by // 8) * ((bw + 7) // 8) + bx // 8
            yx, yb = ytox[t] / 84.0, 1.0 + ytob[t] / 84.0
            W = {8: w8, 16: w16, 32: w32}[n]
            m = n // 8
            for e in range(n * n):
                if e // n < m and e % n < m:
                    continue
                q = (e // 64)
                slot = ((by + q // m) * bw + bx + q % m) * 64 + e % 64
                yv = coef[1, slot] * qgsf / W[1, e]
                qy = np.rint(yv)
                yr = qy * W[1, e] / qgsf
                xv = (coef[0, slot] - yx * yr) * qgsf / W[0, e]
                bv = (coef[2, slot] - yb * yr) * qgsf / W[2, e]
                for val, got in ((yv, ac.reshape(3, -1)[1, slot]), (xv, ac.reshape(3, -1)[0, slot]),
                                 (bv, ac.reshape(3, -1)[2, slot])):
                    if abs(abs(val - np.floor(val)) - 0.5) < 1e-4:
                        ties += 1
                        continue
                    assert got == np.rint(val), (blk, e, val, got)
                    checked += 1
        assert checked > 1000 and ties < checked // 100


def test_dct8_uniform_quantizer_is_rint_of_coeff_over_step(jo):
    """quant.cu:30-97 semantics (DCT8 everywhere, uniform quant field, base
    Y-to-B correlation 1.0), checked as quant_selfconsistency_test.py:41-102
    and :148-179 do: every integer is rint(residual / step) in float64 with
    step = w / (raw_quant_field * global_scale_float) for AC and
    1 / (DC_INV_QUANT[c] * global_scale_float * quant_dc) for DC."""
    import ctypes as C

    w8, _, _ = jo.dequant_weights()
    inv_dc = (4096.0, 512.0, 256.0)
    for seed, d, (w, h) in ((4, 1.0, (64, 64)), (5, 3.0, (96, 64)), (6, 0.5, (64, 32))):
        luma, chroma = _nv12(w, h, seed)
        xyb = jo.gaborish(jo.nv12_to_xyb(luma, chroma, w, h))
        coef = jo.forward_dct(xyb, np.full((w // 8) * (h // 8), 8, np.int8))
        nblk = (w // 8) * (h // 8)
        ac = np.zeros(3 * nblk * 64, np.int32)
        dc = np.zeros(3 * nblk, np.int32)
        jo.lib().jo_quantize_dct8_uniform(jo.ptr(coef), w, h, C.c_float(d), jo.ptr(ac), jo.ptr(dc))
        cal = jo.calibrate(d)
        gsf, rq, qdc = float(cal.global_scale_float), float(cal.raw_quant_field), float(cal.quant_dc)
        v = coef.astype(np.float64).reshape(3, nblk, 64)
        a = ac.reshape(3, nblk, 64)
        dq = dc.reshape(3, nblk)
        ties = checked = 0

        def check(val, got):
            nonlocal ties, checked
            frac = np.abs(np.abs(val - np.floor(val)) - 0.5)
            ok = frac >= 1e-4
            ties += int(np.count_nonzero(~ok))
            np.testing.assert_array_equal(got[ok], np.rint(val[ok]))
            checked += int(np.count_nonzero(ok))

        step = w8 / (rq * gsf)
        qy = np.rint(v[1, :, 1:] / step[1, 1:])
        check(v[1, :, 1:] / step[1, 1:], a[1, :, 1:])
        check(v[0, :, 1:] / step[0, 1:], a[0, :, 1:])
        check((v[2, :, 1:] - qy * step[1, 1:]) / step[2, 1:], a[2, :, 1:])
        dstep = [1.0 / (inv_dc[c] * gsf * qdc) for c in range(3)]
        qdy = np.rint(v[1, :, 0] / dstep[1])
        check(v[1, :, 0] / dstep[1], dq[1])
        check(v[0, :, 0] / dstep[0], dq[0])
        check((v[2, :, 0] - qdy * dstep[1]) / dstep[2], dq[2])
        assert np.all(a[:, :, 0] == 0)
        assert checked > 3 * nblk * 60 and ties < checked // 100


def test_dct_error_report_per_size(jo):
    """Error of the path's DCT against an FP64 DCT of the same sharpened
    planes, per transform size, on a synthetic frame with every block forced
    to that size: the FP32 transform (what the device computes, bit for bit,
    on the VALU and the matrix cores) and the FP16 values it stores.  The FP32
    error stays within 1e-6 of the largest coefficient magnitude; the FP16
    value is the FP32 one rounded to nearest (within half an FP16 ulp,
    2^-11 relative, or 2^-25 absolute in the subnormal range)."""
    w = h = 256
    luma, chroma = synthetic_nv12(w, h, 11)
    sharp = jo.gaborish(jo.nv12_to_xyb(luma, chroma, w, h))
    report = {}
    for n in (8, 16, 32):
        acs = np.zeros((h // 8) * (w // 8), np.int8)
        s = n // 8
        for by in range(0, h // 8, s):
            for bx in range(0, w // 8, s):
                acs[by * (w // 8) + bx] = n
        c32 = jo.forward_dct(sharp, acs).astype(np.float64)
        c64 = jo.forward_dct_f64(sharp, acs)
        c16 = jo.to_f16(c32.astype(np.float32)).view(np.float16).astype(np.float64)
        scale = np.abs(c64).max()
        e32, e16 = c32 - c64, c16 - c64
        report[n] = (np.abs(e32).max() / scale, np.sqrt(np.mean(e32 ** 2)) / scale,
                     np.abs(e16).max() / scale, np.sqrt(np.mean(e16 ** 2)) / scale)
        assert report[n][0] < 1e-6, n
        assert (np.abs(c16 - c32) <= np.abs(c32) * 2.0 ** -11 + 2.0 ** -25).all(), n
    print("DCT error vs FP64 (relative to max |coef|): " + "; ".join(
        f"N={n}: fp32 max {r[0]:.2e} rms {r[1]:.2e}, fp16 max {r[2]:.2e} rms {r[3]:.2e}"
        for n, r in report.items()))
