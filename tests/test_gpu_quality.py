"""GPU: what the learned and matrix-core forms cost in rate and quality, and
what the trained network gains over the heuristic alone.

On held-out seeded frames (seeds the network was not trained or tuned on;
one 1920x1080 frame by default, two 3840x2160 frames with CJX_QUALITY_FULL
set), each form encodes at distances 0.9, 1.0 and 1.1; every file is
decoded by the oracle's subset decoder, rebuilt to pixels (tests/decoded.py)
and scored against the source with the Butteraugli-structured distance of
tests/butteraugli.py (3-norm of its diffmap; not libjxl's Butteraugli) and
PSNR.  The rate of a form at the plain path's d = 1.0 score is interpolated
from its three (score, bpp) points.  The ablation: the trained network
decides most regions itself (the rest fall back to the heuristic), its
decisions agree with the heuristic's on most of them, the region kernel is
faster than the plain one (timed on a held-out 3840x2160 frame), and rate
and quality stay within 2 % of the plain path's.  With CJX_QUALITY_OUT set, the summary is written there as JSON.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

import butteraugli as ba
import cujpegxl_amd as cj
from cujpegxl_amd import mlp_export
from cujpegxl_amd.synth import synthetic_nv12
from decoded import psnr, reconstruct_xyb

pytestmark = pytest.mark.gpu

FULL = bool(os.environ.get("CJX_QUALITY_FULL"))
W, H = (3840, 2160) if FULL else (1920, 1080)
SEEDS = (7000, 7001) if FULL else (7000,)
DISTANCES = (0.9, 1.0, 1.1)
FORMS = {"plain": 0, "mfma16": cj.MODE_MFMA16, "learned": cj.MODE_LEARNED,
         "learned+mfma16": cj.MODE_LEARNED | cj.MODE_MFMA16}


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def encode_all(luma, chroma):
    """files[form][d] and the transform maps at d = 1.0."""
    nv12 = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    files, maps = {}, {}
    for form, mode in FORMS.items():
        files[form] = {}
        for d in DISTANCES:
            enc = cj.Encoder(W, H, d, stages=True)
            enc.set_mode(mode)
            files[form][d] = enc.encode_device(nv12.ptr, nv12.ptr + W * H, W, W)
            if d == 1.0:
                maps[form] = enc.fetch("acs", np.int8).copy()
                if form == "plain":
                    enc.export_features(True)
                    enc.encode_device(nv12.ptr, nv12.ptr + W * H, W, W)
                    maps["rows"] = enc.fetch("features", np.float32).reshape(-1, 108).copy()
                    enc.export_features(False)
            enc.close()
    nv12.close()
    return files, maps


def region_times(seed: int = 7100, reps: int = 24) -> dict:
    """Median region-kernel time per form on a held-out 3840x2160 frame at
    d = 1.0 (the bench's workload; at smaller sizes the kernel's last partial
    round of workgroups dominates)."""
    w, h = 3840, 2160
    luma, chroma = synthetic_nv12(w, h, seed)
    nv12 = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    out = {}
    enc = cj.Encoder(w, h, 1.0, stages=True)
    for form, mode in FORMS.items():
        enc.set_mode(mode)
        for _ in range(4):
            enc.encode_device(nv12.ptr, nv12.ptr + w * h, w, w)
        enc.set_profiling(True, region_only=True)
        t = []
        for _ in range(reps):
            enc.encode_device(nv12.ptr, nv12.ptr + w * h, w, w)
            t.append(enc.region_ms())
        enc.set_profiling(False)
        out[form] = float(np.median(t))
    enc.close()
    nv12.close()
    return out


def score(jo, ref, src_xyb, data):
    dec = jo.decode(data)
    xyb = reconstruct_xyb(jo, dec)
    rgb = ba.xyb_to_linear_rgb(xyb)
    bmax, bnorm = ref.distance(rgb)
    return {"bytes": len(data), "bpp": 8.0 * len(data) / (W * H), "butteraugli_max": bmax,
            "butteraugli_3norm": bnorm, "psnr_xyb_y": psnr(xyb[1], src_xyb[1])}


def confident_share(jo, rows: np.ndarray) -> float:
    """Share of regions the shipped network decides itself (the region
    kernel's rule on its float64 forward; the rest fall back)."""
    params, th = mlp_export.read_inc()
    z = jo.mlp_forward(params, rows[:, :96])
    lab = rows[:, 96:]
    fit32 = lab[:, 6] > 0.5
    root = z[:, 1] - z[:, 0]
    use32 = fit32 & (root > 0)
    fit16 = lab[:, 7:11] > 0.5
    conf = ~fit32 | (np.abs(root) >= th)
    conf &= use32 | np.all(~fit16 | (np.abs(z[:, 2:6]) >= th), axis=1)
    return float(conf.mean())


def bpp_at(points: list[dict], target: float) -> float:
    """bpp at a 3-norm score by linear interpolation over the distance
    sweep (scores rise with the distance)."""
    s = np.array([p["butteraugli_3norm"] for p in points])
    b = np.array([p["bpp"] for p in points])
    order = np.argsort(s)
    return float(np.interp(target, s[order], b[order]))


def test_rate_quality_and_ablation(jo):
    summary = {"frames": [], "size": f"{W}x{H}", "distances": list(DISTANCES),
               "metric": "Butteraugli-structured 3-norm (tests/butteraugli.py, not libjxl's)"}
    for seed in SEEDS:
        luma, chroma = synthetic_nv12(W, H, seed)
        src_xyb = jo.nv12_to_xyb_f64(luma, chroma, W, H).astype(np.float64)
        ref = ba.Reference(ba.xyb_to_linear_rgb(src_xyb))
        say(f"quality study seed {seed}: encoding {len(FORMS)} forms x {len(DISTANCES)} distances")
        files, maps = encode_all(luma, chroma)
        rows = {}
        for form in FORMS:
            rows[form] = []
            for d in DISTANCES:
                r = score(jo, ref, src_xyb, files[form][d])
                r["distance"] = d
                rows[form].append(r)
                say(f"  {form:15s} d={d}: {r['bpp']:.4f} bpp, score {r['butteraugli_3norm']:.4f} "
                    f"(max {r['butteraugli_max']:.3f}), PSNR Y {r['psnr_xyb_y']:.2f} dB")
        target = rows["plain"][1]["butteraugli_3norm"]
        frame = {"seed": seed, "rows": rows, "plain_score_d1": target,
                 "bpp_at_plain_score": {f: bpp_at(rows[f], target) for f in FORMS}}
        frame["bpp_increase"] = {f: frame["bpp_at_plain_score"][f] / frame["bpp_at_plain_score"]["plain"] - 1.0
                                 for f in FORMS}
        agree = float((maps["learned"] == maps["plain"]).mean())
        frame["learned_map_agreement"] = agree
        frame["network_decided"] = confident_share(jo, maps["rows"])
        say(f"  regions the network decides: {frame['network_decided']:.3f}")
        say(f"  bpp increase at equal score {json.dumps({k: round(v, 4) for k, v in frame['bpp_increase'].items()})}"
            f"; learned map agreement {agree:.4f}")
        summary["frames"].append(frame)
        # Matrix-core and learned forms: rate and quality within 2 % of plain.
        for form in ("mfma16", "learned", "learned+mfma16"):
            assert frame["bpp_increase"][form] <= 0.02, form
            r, p = rows[form][1], rows["plain"][1]
            assert r["butteraugli_3norm"] <= 1.02 * p["butteraugli_3norm"], form
        assert agree >= 0.9
        assert frame["network_decided"] >= 0.5
    keys = ("mfma16", "learned", "learned+mfma16")
    summary["mean_bpp_increase"] = {k: float(np.mean([f["bpp_increase"][k] for f in summary["frames"]]))
                                    for k in keys}
    summary["region_ms_4k"] = region_times()
    say(f"region kernel ms at 3840x2160 {json.dumps({k: round(v, 4) for k, v in summary['region_ms_4k'].items()})}")
    out = os.environ.get("CJX_QUALITY_OUT")
    if out:
        with open(out, "w") as f:
            json.dump(summary, f, indent=1)
    # The network's gain: a faster region kernel than the heuristic alone,
    # on the general and on the matrix-core DCTs.
    t = summary["region_ms_4k"]
    assert t["learned"] < t["plain"]
    assert t["learned+mfma16"] < t["mfma16"]
