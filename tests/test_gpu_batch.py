"""GPU: batches of frames of uneven sizes and distances, every stage one
launch over the whole batch, replayed from a captured hipGraph.

Each frame's file equals the oracle's (plain path) or the single-frame
encoder's in the same mode (matrix-core / learned forms) byte for byte, on
the first replay and on later ones, after new inputs are set, and after the
autotuner has changed the launch shapes."""

from __future__ import annotations

import sys

import numpy as np
import pytest

import cujpegxl_amd as cj
from cujpegxl_amd.synth import synthetic_nv12

pytestmark = pytest.mark.gpu

FRAMES = [(640, 360, 0.8), (1280, 720, 1.7), (264, 520, 3.0), (1920, 1080, 1.0), (776, 456, 2.2),
          (640, 360, 2.9)]


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def frames_for(seed0):
    host = [synthetic_nv12(w, h, seed0 + i) for i, (w, h, _) in enumerate(FRAMES)]
    devs = [cj.DeviceBuffer.from_array(np.concatenate([l.ravel(), c.ravel()])) for l, c in host]
    inputs = [(d.ptr, d.ptr + w * h, w, w) for d, (w, h, _) in zip(devs, FRAMES)]
    return host, devs, inputs


def test_batch_plain_matches_oracle(jo):
    say(f"batch of {len(FRAMES)} frames, plain")
    host, devs, inputs = frames_for(40)
    b = cj.Batch([(w, h) for w, h, _ in FRAMES], [d for _, _, d in FRAMES])
    b.set_inputs(inputs)
    b.run(1)
    want = [jo.encode_nv12(l, c, w, h, d) for (l, c), (w, h, d) in zip(host, FRAMES)]
    assert b.files() == want
    b.run(3)
    assert b.files() == want
    # New inputs, same graph.
    host2, devs2, inputs2 = frames_for(60)
    b.set_inputs(inputs2)
    b.run(1)
    assert b.files() == [jo.encode_nv12(l, c, w, h, d) for (l, c), (w, h, d) in zip(host2, FRAMES)]
    # The autotuner's choice changes launch shapes, not files.
    report = b.tune(2)
    say("  " + report.strip().splitlines()[-1])
    assert report.count("batch_ms") >= 12
    b.set_inputs(inputs)
    b.run(1)
    assert b.files() == want
    b.close()
    for d in devs + devs2:
        d.close()


@pytest.mark.parametrize("mode", [cj.MODE_MFMA16, cj.MODE_LEARNED | cj.MODE_MFMA16])
def test_batch_modes_match_single_frames(mode):
    say(f"batch of {len(FRAMES)} frames, mode {mode}")
    host, devs, inputs = frames_for(80)
    b = cj.Batch([(w, h) for w, h, _ in FRAMES], [d for _, _, d in FRAMES], mode=mode)
    b.set_inputs(inputs)
    b.run(2)
    got = b.files()
    for i, ((w, h, d), inp) in enumerate(zip(FRAMES, inputs)):
        enc = cj.Encoder(w, h, d, stages=True)
        enc.set_mode(mode)
        assert got[i] == enc.encode_device(*inp), i
        enc.close()
    b.close()
    for d in devs:
        d.close()


def test_two_batches_in_flight(jo):
    """Two batches queued on their own streams, alternately, without waits
    between their replays: each one's files still equal the oracle's."""
    say("two batches in flight")
    sizes = [(w, h) for w, h, _ in FRAMES[:3]]
    dists = [d for _, _, d in FRAMES[:3]]
    pair, devs_all, want = [], [], []
    for seed0 in (100, 120):
        host = [synthetic_nv12(w, h, seed0 + i) for i, (w, h) in enumerate(sizes)]
        devs = [cj.DeviceBuffer.from_array(np.concatenate([l.ravel(), c.ravel()])) for l, c in host]
        b = cj.Batch(sizes, dists)
        b.set_inputs([(d.ptr, d.ptr + w * h, w, w) for d, (w, h) in zip(devs, sizes)])
        pair.append(b)
        devs_all += devs
        want.append([jo.encode_nv12(l, c, w, h, d) for (l, c), (w, h), d in zip(host, sizes, dists)])
    for _ in range(3):
        for b in pair:
            b.launch(1)
    for b in pair:
        b.wait()
    for b, w in zip(pair, want):
        assert b.files() == w
        b.close()
    for d in devs_all:
        d.close()


@pytest.mark.parametrize("mode", [0, cj.MODE_MFMA16 | cj.MODE_LEARNED])
def test_persistent_region_grid(jo, mode):
    """The region stage as a persistent grid (workgroups claiming regions
    through per-XCD counters) gives the files of one workgroup per region,
    for grids smaller and larger than the batch's region count."""
    say(f"persistent region grid, mode {mode}")
    host, devs, inputs = frames_for(140)
    b = cj.Batch([(w, h) for w, h, _ in FRAMES], [d for _, _, d in FRAMES], mode=mode)
    b.set_inputs(inputs)
    b.run(1)
    want = b.files()
    if mode == 0:
        assert want == [jo.encode_nv12(l, c, w, h, d) for (l, c), (w, h, d) in zip(host, FRAMES)]
    for grid in (8, 1280, 4096):
        b.set_shape(512, 16, grid)
        b.run(2)
        assert b.files() == want, grid
    with pytest.raises(ValueError):
        b.set_shape(512, 16, 12)
    b.close()
    for d in devs:
        d.close()
