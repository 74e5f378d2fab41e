"""CPU: the built libraries load, export exactly the declared C ABI, and their
host-side tables and argument checks match the reference without a device."""

from __future__ import annotations

import ctypes as C
import pathlib
import re
import subprocess

import numpy as np
import pytest

import cujpegxl_amd as cj

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared(header: str, prefix: str) -> set[str]:
    text = (ROOT / "include" / "cujpegxl" / header).read_text()
    return set(re.findall(rf"\b({prefix}\w+)\s*\(", text))


def exported(lib: pathlib.Path) -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1].split("@")[0] for line in out.splitlines() if " T " in line}


def test_public_library_exports_exactly_the_abi():
    want = declared("cujpegxl.h", "cujpegxl_")
    # cujpegxl.h:66-82, plus the extension version query, the host-input
    # and asynchronous entry points, the allocation-free encode with its
    # workspace query, and the encoder-form switch.
    assert len(want) == 16
    assert exported(cj.PUBLIC_LIB) == want
    lib = cj.public_lib()
    for name in want:
        assert hasattr(lib, name)


def test_stage_library_exports_abi_and_harness():
    want = declared("cujpegxl.h", "cujpegxl_") | declared("cujpegxl_stages.h", "cjx_")
    assert exported(cj.STAGES_LIB) == want


def test_status_and_backend_strings():
    # cujpegxl.cc:69-99
    names = ["ok", "not-implemented", "invalid-argument", "unsupported-resolution", "no-device",
             "internal-error", "buffer-too-small"]
    assert [cj.status_string(i) for i in range(7)] == names
    lib = cj.public_lib()
    assert [lib.cujpegxl_backend_string(i).decode() for i in range(3)] == ["unknown", "fp32-simt",
                                                                            "fp16-tensor"]
    assert cj.api_version() == 3 and cj.ext_version() == 2


def test_max_output_size_and_validation():
    # cujpegxl.cc:50-59 bound, and the resolution / distance rules (:36-48).
    bw, bh = 480, 270
    assert cj.max_output_size(3840, 2160) == 3 * bw * bh * 384 + 48 * bw * bh + (135 + 4) * 2048 + 16384
    lib = cj.public_lib()
    n = C.c_size_t()
    for (w, h, d, want) in ((100, 100, 1.0, cj.UNSUPPORTED_RESOLUTION),
                            (256, 256, 1.0, cj.UNSUPPORTED_RESOLUTION),  # one AC group
                            (512, 512, 0.0, cj.INVALID_ARGUMENT),
                            (512, 512, float("nan"), cj.INVALID_ARGUMENT),
                            (512, 520, 1.0, cj.OK)):
        assert lib.cujpegxl_max_output_size(C.byref(cj.Config(w, h, d, 0)), C.byref(n)) == want
    assert lib.cujpegxl_max_output_size(None, C.byref(n)) == cj.INVALID_ARGUMENT
    h = C.c_void_p()
    assert lib.cujpegxl_encoder_create(C.byref(cj.Config(100, 100, 1.0, 0)), C.byref(h)) == \
        cj.UNSUPPORTED_RESOLUTION and not h.value
    assert lib.cujpegxl_encoder_encode(None, None, None) == cj.INVALID_ARGUMENT


def test_host_tables_match_reference_fixtures(jo, golden):
    L = cj.stages_lib()
    w = np.zeros(3 * (64 + 256 + 1024), np.float32)
    assert L.cjx_host_table(0, w.ctypes.data, w.nbytes) == w.nbytes
    np.testing.assert_array_equal(w, np.fromfile(ROOT / "tests/golden/dequant_weights.bin", "<f4"))
    o = np.zeros(64 + 256 + 1024, np.uint16)
    L.cjx_host_table(2, o.ctypes.data, o.nbytes)
    np.testing.assert_array_equal(o, np.concatenate([golden["order8"], golden["order16"],
                                                     golden["order32"]]))
    b = np.zeros(64 + 256 + 1024, np.float32)
    L.cjx_host_table(1, b.ctypes.data, b.nbytes)
    for n, off in ((8, 0), (16, 64), (32, 320)):
        ob = np.zeros(n * n, np.float32)
        jo.lib().jo_dct_basis(n, jo.ptr(ob))
        np.testing.assert_array_equal(b[off : off + n * n], ob)


def test_python_binding_argument_errors():
    with pytest.raises(ValueError):
        cj.encode(np.zeros(10, np.uint8), 512, 512)
    with pytest.raises(ValueError):
        cj.encode(np.zeros((512 * 512 * 3 // 2) * 2, np.uint8)[::2], 512, 512)


def test_symbol_versions_keep_the_reference_node():
    """The reference's nine entry points stay under CUJPEGXL_1
    (cujpegxl.lds:4-9); the extensions sit under CUJPEGXL_2, the
    allocation-free encode under CUJPEGXL_3, the encoder-form switch under
    CUJPEGXL_4."""
    out = subprocess.run(["objdump", "-T", str(cj.PUBLIC_LIB)], capture_output=True, text=True,
                         check=True).stdout
    nodes = {}
    for line in out.splitlines():
        m = re.search(r"\s(CUJPEGXL_\d)\s+(cujpegxl_\w+)$", line)
        if m:
            nodes[m.group(2)] = m.group(1)
    ext = {"cujpegxl_ext_version", "cujpegxl_encoder_encode_host", "cujpegxl_encoder_encode_async",
           "cujpegxl_encoder_finish"}
    noalloc = {"cujpegxl_noalloc_workspace_bytes", "cujpegxl_encode_noalloc"}
    assert len(nodes) == 16
    for name, node in nodes.items():
        want = "CUJPEGXL_2" if name in ext else ("CUJPEGXL_3" if name in noalloc else "CUJPEGXL_1")
        if name == "cujpegxl_encoder_set_mode":
            want = "CUJPEGXL_4"
        assert node == want, name


def test_pycujpegxl_module_as_the_reference_tests_it():
    """python/pycujpegxl_test.py:10-27 of the reference, run against the
    `pycujpegxl` module here (no device: query_backend raises)."""
    import pycujpegxl

    assert pycujpegxl.api_version() == 3
    assert pycujpegxl.API_VERSION == 3
    assert callable(pycujpegxl.query_backend)
    assert callable(pycujpegxl.encode)
    assert callable(pycujpegxl.encode_with_stats)
    with pytest.raises(RuntimeError):
        pycujpegxl.query_backend(9999)
    with pytest.raises(ValueError):
        pycujpegxl.encode(np.zeros(10, np.uint8), 512, 512)


def test_noalloc_scratch_holds_every_accepted_shape():
    """The allocation-free encode's scratch array (CUJPEGXL_NOALLOC_SCRATCH_BYTES)
    holds the buffers of every shape it accepts: for each width, the tallest
    accepted height (the workspace grows with the height)."""
    text = (ROOT / "include" / "cujpegxl" / "cujpegxl.h").read_text()
    scratch = int(re.search(r"NOALLOC_SCRATCH_BYTES \((\d+)ull << 20\)", text).group(1)) << 20
    max_side = int(re.search(r"NOALLOC_MAX_WIDTH (\d+)u", text).group(1))
    a, b = re.search(r"NOALLOC_MAX_PIXELS \((\d+)u \* (\d+)u\)", text).groups()
    max_px = int(a) * int(b)
    lib = cj.public_lib()
    worst = 0
    for w in range(8, max_side + 1, 8):
        h = min(max_side, max_px // w // 8 * 8)
        if h >= 8:
            worst = max(worst, lib.cujpegxl_noalloc_workspace_bytes(w, h))
            worst = max(worst, lib.cujpegxl_noalloc_workspace_bytes(h, w))
    assert 0 < worst <= scratch
    # 4K needs far more than the array: such frames are refused, not truncated.
    assert lib.cujpegxl_noalloc_workspace_bytes(3840, 2160) > scratch
    cfg = cj.Config(3840, 2160, 1.0, 0)
    inp = cj.Nv12Input(1 << 20, 2 << 20, 3840, 3840)
    assert lib.cujpegxl_encode_noalloc(C.byref(cfg), C.byref(inp), 1 << 24, 1 << 20, 1 << 12, 0) == \
        cj.UNSUPPORTED_RESOLUTION
