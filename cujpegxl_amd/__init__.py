"""cujpegxl for AMD Instinct MI355X: Python access to the C ABI.

Mirrors the reference's `pycujpegxl` module (python/pycujpegxl.cc:31-90):
`encode`, `encode_with_stats`, `query_backend`, `api_version`, `API_VERSION`,
with the same argument meaning and errors (ValueError for a wrongly sized or
non-contiguous array, RuntimeError for an encode or device failure).  Unlike
the reference, encoding goes through the C ABI (libcujpegxl.so.1): host NV12
through cujpegxl_encoder_encode_host, which uploads the planes itself, so no
framework is needed for device memory.
"""

from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent
# CJX_REGION_CUT=1 selects the profiling build (lib_cut/, made by
# `python -m cujpegxl_amd.build --region-cut`) for cujpegxl_amd/region_phases.py.
LIB_DIR = PKG / ("lib_cut" if os.environ.get("CJX_REGION_CUT") == "1" else "lib")
PUBLIC_LIB = LIB_DIR / "libcujpegxl.so.1"
STAGES_LIB = LIB_DIR / "libcujpegxl_stages.so"

API_VERSION = 3  # cujpegxl.h's CUJPEGXL_API_VERSION, the reference's value
EXT_VERSION = 2  # CUJPEGXL_EXT_VERSION: host-input, asynchronous and encoder-form entry points

OK, NOT_IMPLEMENTED, INVALID_ARGUMENT, UNSUPPORTED_RESOLUTION, NO_DEVICE, INTERNAL_ERROR, \
    BUFFER_TOO_SMALL = range(7)

MODE_MFMA16 = 1  # CJX_MODE_MFMA16: FP16-operand matrix-core DCTs
MODE_LEARNED = 2  # CJX_MODE_LEARNED: the selection network where it is confident
MODE_FORCE8, MODE_FORCE16, MODE_FORCE32 = 0x10, 0x20, 0x30  # test modes: one transform size

STAGE_NAMES = ("region", "quantize", "histograms", "codes", "sizes", "emit", "readback")
BUFFERS = {
    "coeffs": 0, "acs": 1, "quant_field": 2, "cfl_partials": 3, "ytox": 4, "ytob": 5, "ac": 6,
    "dc": 7, "nz_grid": 8, "nz_count": 9, "histograms": 10, "ac_group_bytes": 11,
    "dc_group_bytes": 12, "code_depth": 13, "code_bits": 14, "features": 15,
}


class Config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("distance", C.c_float),
                ("device_ordinal", C.c_int32)]


class Nv12HostInput(C.Structure):
    _fields_ = [("luma", C.c_void_p), ("chroma", C.c_void_p), ("luma_pitch", C.c_size_t),
                ("chroma_pitch", C.c_size_t)]


class Nv12Input(C.Structure):
    _fields_ = [("luma", C.c_size_t), ("chroma", C.c_size_t), ("luma_pitch", C.c_size_t),
                ("chroma_pitch", C.c_size_t)]


class Output(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t), ("capacity", C.c_size_t)]


_public = None
_stages = None


def _bind_public(L: C.CDLL) -> C.CDLL:
    L.cujpegxl_api_version.restype = C.c_uint32
    L.cujpegxl_ext_version.restype = C.c_uint32
    L.cujpegxl_status_string.restype = C.c_char_p
    L.cujpegxl_status_string.argtypes = [C.c_int]
    L.cujpegxl_backend_string.restype = C.c_char_p
    L.cujpegxl_backend_string.argtypes = [C.c_int]
    L.cujpegxl_query_backend.argtypes = [C.c_int32, C.POINTER(C.c_int)]
    L.cujpegxl_max_output_size.argtypes = [C.POINTER(Config), C.POINTER(C.c_size_t)]
    L.cujpegxl_encoder_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.cujpegxl_encoder_configure.argtypes = [C.c_void_p, C.POINTER(Config)]
    L.cujpegxl_encoder_set_mode.argtypes = [C.c_void_p, C.c_uint32]
    L.cujpegxl_encoder_encode.argtypes = [C.c_void_p, C.POINTER(Nv12Input), C.POINTER(Output)]
    L.cujpegxl_encoder_destroy.argtypes = [C.c_void_p]
    L.cujpegxl_encoder_destroy.restype = None
    L.cujpegxl_encoder_encode_host.argtypes = [C.c_void_p, C.POINTER(Nv12HostInput), C.POINTER(Output)]
    L.cujpegxl_encoder_encode_async.argtypes = [C.c_void_p, C.POINTER(Nv12Input), C.c_size_t]
    L.cujpegxl_encoder_finish.argtypes = [C.c_void_p, C.POINTER(Output)]
    L.cujpegxl_noalloc_workspace_bytes.restype = C.c_size_t
    L.cujpegxl_noalloc_workspace_bytes.argtypes = [C.c_uint32, C.c_uint32]
    L.cujpegxl_encode_noalloc.argtypes = [C.POINTER(Config), C.POINTER(Nv12Input), C.c_size_t, C.c_size_t,
                                          C.c_size_t, C.c_size_t]
    return L


def public_lib() -> C.CDLL:
    """libcujpegxl.so.1; raises if it has not been built."""
    global _public
    if _public is None:
        if not PUBLIC_LIB.exists():
            raise RuntimeError(f"{PUBLIC_LIB} is missing: run cujpegxl_amd/build.py")
        _public = _bind_public(C.CDLL(str(PUBLIC_LIB)))
    return _public


def stages_lib() -> C.CDLL:
    """libcujpegxl_stages.so (C ABI plus the cjx_* stage harness)."""
    global _stages
    if _stages is None:
        if not STAGES_LIB.exists():
            raise RuntimeError(f"{STAGES_LIB} is missing: run cujpegxl_amd/build.py")
        L = _bind_public(C.CDLL(str(STAGES_LIB)))
        P = C.c_void_p
        L.cjx_stage_buffer_bytes.restype = C.c_size_t
        L.cjx_stage_buffer_bytes.argtypes = [P, C.c_int]
        L.cjx_stage_fetch.argtypes = [P, C.c_int, P, C.c_size_t]
        L.cjx_stage_encode_coefficients.argtypes = [P, P, P, P, P, P, P, C.c_uint32, C.c_uint32, P,
                                                    C.c_size_t, C.POINTER(C.c_size_t)]
        L.cjx_stage_encoder_create_any.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
        L.cjx_stage_encode_nv12.argtypes = [P, C.POINTER(Nv12Input), C.POINTER(Output)]
        L.cjx_stage_set_profiling.argtypes = [P, C.c_int]
        L.cjx_stage_set_mode.argtypes = [P, C.c_uint32]
        L.cjx_stage_export_features.argtypes = [P, C.c_int]
        L.cjx_mlp_set_weights.argtypes = [C.c_int32, P, P, P, P, C.c_float]
        L.cjx_mlp_default_weights.argtypes = [C.c_int32]
        L.cjx_mlp_partial_floats.restype = C.c_size_t
        L.cjx_mlp_partial_floats.argtypes = [C.c_uint32]
        L.cjx_mlp_step.argtypes = [P, P, C.c_uint32, C.c_float, P, P, P, P, P, P, C.c_uint32, C.c_float, C.c_size_t]
        L.cjx_mlp_forward.argtypes = [P, P, C.c_uint32, P, C.c_size_t]
        L.cjx_batch_create.argtypes = [C.c_int32, C.c_uint32, P, P, P, C.c_uint32, C.POINTER(C.c_void_p)]
        L.cjx_batch_set_inputs.argtypes = [P, P]
        L.cjx_batch_run.argtypes = [P, C.c_int32, C.POINTER(C.c_double)]
        L.cjx_batch_launch.argtypes = [P, C.c_int32]
        L.cjx_batch_wait.argtypes = [P]
        L.cjx_batch_result.argtypes = [P, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        L.cjx_batch_tune.argtypes = [P, C.c_int32, C.c_char_p, C.c_size_t]
        L.cjx_batch_set_shape.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
        L.cjx_batch_destroy.argtypes = [P]
        L.cjx_batch_destroy.restype = None
        L.cjx_stage_encode_repeat.argtypes = [P, C.POINTER(Nv12Input), C.POINTER(Output), C.c_int,
                                              C.POINTER(C.c_double)]
        L.cjx_stage_encode_cycle.argtypes = [P, C.POINTER(Nv12Input), C.c_int, C.POINTER(Output), C.c_int,
                                             C.POINTER(C.c_double)]
        L.cjx_stage_encode_dct8_uniform.argtypes = [P, C.POINTER(Nv12Input), C.POINTER(Output)]
        L.cjx_stage_timings.argtypes = [P, C.POINTER(C.c_float), C.c_int]
        L.cjx_device_alloc.restype = P
        L.cjx_stage_build_codes.argtypes = [C.c_int32, P, C.c_uint32, P, P, P, P, P]
        L.cjx_device_alloc.argtypes = [C.c_size_t]
        L.cjx_stream_create.restype = C.c_size_t
        L.cjx_stream_destroy.argtypes = [C.c_size_t]
        L.cjx_stream_destroy.restype = None
        L.cjx_stream_synchronize.argtypes = [C.c_size_t]
        L.cjx_copy_async.argtypes = [P, P, C.c_size_t, C.c_size_t]
        L.cjx_fill_async.argtypes = [P, C.c_int, C.c_size_t, C.c_size_t]
        L.cjx_host_alloc.restype = P
        L.cjx_host_alloc.argtypes = [C.c_size_t]
        L.cjx_host_free.argtypes = [P]
        L.cjx_host_free.restype = None
        L.cjx_device_free.argtypes = [P]
        L.cjx_device_free.restype = None
        L.cjx_copy_to_device.argtypes = [P, P, C.c_size_t]
        L.cjx_copy_to_host.argtypes = [P, P, C.c_size_t]
        L.cjx_device_copy.argtypes = [P, P, C.c_size_t]
        L.cjx_stage_fail_allocation_after.argtypes = [C.c_int]
        L.cjx_stage_fail_allocation_after.restype = None
        L.cjx_host_table.restype = C.c_size_t
        L.cjx_host_table.argtypes = [C.c_int, P, C.c_size_t]
        _stages = L
    return _stages


def status_string(status: int) -> str:
    return public_lib().cujpegxl_status_string(status).decode()


def api_version() -> int:
    return int(public_lib().cujpegxl_api_version())


def ext_version() -> int:
    return int(public_lib().cujpegxl_ext_version())


def query_backend(device_ordinal: int = 0) -> str:
    b = C.c_int(0)
    s = public_lib().cujpegxl_query_backend(device_ordinal, C.byref(b))
    if s != OK:
        raise RuntimeError(f"query_backend failed: {status_string(s)}")
    return public_lib().cujpegxl_backend_string(b.value).decode()


def max_output_size(width: int, height: int, distance: float = 1.0) -> int:
    n = C.c_size_t(0)
    s = public_lib().cujpegxl_max_output_size(C.byref(Config(width, height, distance, 0)), C.byref(n))
    if s != OK:
        raise ValueError(status_string(s))
    return n.value


def build_codes(hists: np.ndarray, device_ordinal: int = 0) -> list[dict]:
    """Prefix codes of up to 96 histograms (rows of 256 uint32 bins) built by
    the device code kernel: alphabet size, depths, codes and header bytes."""
    h = np.ascontiguousarray(hists, dtype=np.uint32).reshape(-1, 256)
    n = h.shape[0]
    if not 0 < n <= 96:
        raise ValueError("between 1 and 96 histograms")
    alpha = np.zeros(n, np.uint32)
    depth = np.zeros((n, 256), np.uint8)
    bits = np.zeros((n, 256), np.uint16)
    hdr = np.zeros((n, 64), np.uint32)
    nbits = np.zeros(n, np.uint32)
    s = stages_lib().cjx_stage_build_codes(device_ordinal, h.ctypes.data, n, alpha.ctypes.data,
                                           depth.ctypes.data, bits.ctypes.data, hdr.ctypes.data,
                                           nbits.ctypes.data)
    if s != OK:
        raise RuntimeError(f"build_codes failed: {status_string(s)}")
    out = []
    for i in range(n):
        nb = int(nbits[i])
        out.append({"alpha": int(alpha[i]), "depth": depth[i], "bits": bits[i],
                    "header": hdr[i].view(np.uint8)[: (nb + 7) // 8].tobytes(), "header_bits": nb})
    return out


class DeviceBuffer:
    """A device allocation from the stage library (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        self.ptr = stages_lib().cjx_device_alloc(max(nbytes, 1))
        if not self.ptr:
            raise RuntimeError("device allocation failed")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        if stages_lib().cjx_copy_to_device(buf.ptr, a.ctypes.data, a.nbytes) != OK:
            raise RuntimeError("host to device copy failed")
        return buf

    def to_host(self, dtype=np.uint8) -> np.ndarray:
        out = np.empty(self.nbytes, np.uint8)
        if stages_lib().cjx_copy_to_host(out.ctypes.data, self.ptr, self.nbytes) != OK:
            raise RuntimeError("device to host copy failed")
        return out.view(dtype)

    def close(self) -> None:
        if self.ptr:
            stages_lib().cjx_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory from the runtime (hipHostMalloc), viewable as a
    numpy array.  The library reads inputs from and writes files into such
    buffers directly (DMA, or the assembly kernel's own stores), skipping its
    staging copy."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        self.ptr = stages_lib().cjx_host_alloc(max(nbytes, 1))
        if not self.ptr:
            raise RuntimeError("pinned host allocation failed")
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(nbytes, 1)).from_address(self.ptr))[:nbytes]

    def close(self) -> None:
        if self.ptr:
            self.array = None
            stages_lib().cjx_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Encoder:
    """An encoder of the C ABI (or of the stage library with `stages=True`,
    which also exposes intermediate buffers and per-stage timings)."""

    def __init__(self, width: int, height: int, distance: float = 1.0, device_ordinal: int = 0,
                 stages: bool = False, any_size: bool = False):
        self.lib = stages_lib() if (stages or any_size) else public_lib()
        self.width, self.height, self.distance = width, height, distance
        self.device_ordinal = device_ordinal
        h = C.c_void_p()
        cfg = Config(width, height, distance, device_ordinal)
        if any_size:
            s = self.lib.cjx_stage_encoder_create_any(C.byref(cfg), C.byref(h))
        else:
            s = self.lib.cujpegxl_encoder_create(C.byref(cfg), C.byref(h))
        if s != OK:
            raise RuntimeError(f"encoder_create failed: {status_string(s)}")
        self.handle = h
        self._encode = self.lib.cjx_stage_encode_nv12 if any_size else self.lib.cujpegxl_encoder_encode
        bw, bh = width // 8, height // 8
        self.capacity = 3 * bw * bh * 384 + 48 * bw * bh + (1 << 20)
        self.out = np.empty(self.capacity, np.uint8)

    def configure(self, distance: float) -> None:
        """A new distance for the same frame size (cujpegxl_encoder_configure:
        no reallocation)."""
        cfg = Config(self.width, self.height, distance, self.device_ordinal)
        s = self.lib.cujpegxl_encoder_configure(self.handle, C.byref(cfg))
        if s != OK:
            raise RuntimeError(f"encoder_configure failed: {status_string(s)}")
        self.distance = distance

    def encode_device(self, luma_ptr: int, chroma_ptr: int, luma_pitch: int,
                      chroma_pitch: int) -> bytes:
        inp = Nv12Input(luma_ptr, chroma_ptr, luma_pitch, chroma_pitch)
        out = Output(self.out.ctypes.data, 0, self.capacity)
        s = self._encode(self.handle, C.byref(inp), C.byref(out))
        if s != OK:
            raise RuntimeError(f"encode failed: {status_string(s)}")
        return self.out[: out.size].tobytes()

    def encode_into(self, luma_ptr: int, chroma_ptr: int, luma_pitch: int, chroma_pitch: int,
                    out_ptr: int, capacity: int) -> int:
        """Device planes -> the file into caller memory at `out_ptr`; returns
        its size."""
        inp = Nv12Input(luma_ptr, chroma_ptr, luma_pitch, chroma_pitch)
        out = Output(out_ptr, 0, capacity)
        s = self._encode(self.handle, C.byref(inp), C.byref(out))
        if s != OK:
            raise RuntimeError(f"encode failed: {status_string(s)}")
        return out.size

    def encode_repeat(self, luma_ptr: int, chroma_ptr: int, luma_pitch: int, chroma_pitch: int,
                      out_ptr: int, capacity: int, count: int) -> tuple[float, int]:
        """`count` synchronous encodes of one frame into caller memory, back
        to back in native code (stage library); returns the loop's wall
        seconds and the last file's size."""
        inp = Nv12Input(luma_ptr, chroma_ptr, luma_pitch, chroma_pitch)
        out = Output(out_ptr, 0, capacity)
        sec = C.c_double(0.0)
        s = stages_lib().cjx_stage_encode_repeat(self.handle, C.byref(inp), C.byref(out), count, C.byref(sec))
        if s != OK:
            raise RuntimeError(f"encode_repeat failed: {status_string(s)}")
        return sec.value, out.size

    def encode_cycle(self, inputs: list[tuple[int, int, int, int]], out_ptr: int, capacity: int,
                     count: int) -> tuple[float, int]:
        """`count` synchronous encodes cycling through `inputs` (luma, chroma,
        luma pitch, chroma pitch) in native code (stage library); returns the
        loop's wall seconds and the last file's size."""
        arr = (Nv12Input * len(inputs))(*(Nv12Input(*i) for i in inputs))
        out = Output(out_ptr, 0, capacity)
        sec = C.c_double(0.0)
        s = stages_lib().cjx_stage_encode_cycle(self.handle, arr, len(inputs), C.byref(out), count, C.byref(sec))
        if s != OK:
            raise RuntimeError(f"encode_cycle failed: {status_string(s)}")
        return sec.value, out.size

    def encode_dct8_uniform(self, luma_ptr: int, chroma_ptr: int, luma_pitch: int,
                            chroma_pitch: int) -> bytes:
        """Config-1 semantics on the device (stage library): DCT8 everywhere,
        uniform quant field, no CfL maps."""
        inp = Nv12Input(luma_ptr, chroma_ptr, luma_pitch, chroma_pitch)
        out = Output(self.out.ctypes.data, 0, self.capacity)
        s = stages_lib().cjx_stage_encode_dct8_uniform(self.handle, C.byref(inp), C.byref(out))
        if s != OK:
            raise RuntimeError(f"encode_dct8_uniform failed: {status_string(s)}")
        return self.out[: out.size].tobytes()

    def encode_host(self, nv12: np.ndarray) -> bytes:
        """Host NV12 array (luma rows, then interleaved CbCr rows) -> bytes,
        through cujpegxl_encoder_encode_host."""
        a = _planes(nv12, self.width, self.height)
        base = a.ctypes.data
        inp = Nv12HostInput(base, base + self.width * self.height, self.width, self.width)
        out = Output(self.out.ctypes.data, 0, self.capacity)
        s = self.lib.cujpegxl_encoder_encode_host(self.handle, C.byref(inp), C.byref(out))
        if s != OK:
            raise RuntimeError(f"encode failed: {status_string(s)}")
        return self.out[: out.size].tobytes()

    def encode_async(self, luma_ptr: int, chroma_ptr: int, luma_pitch: int, chroma_pitch: int,
                     stream: int = 0) -> int:
        """Queues a frame after the work on `stream`; returns the status."""
        inp = Nv12Input(luma_ptr, chroma_ptr, luma_pitch, chroma_pitch)
        return self.lib.cujpegxl_encoder_encode_async(self.handle, C.byref(inp), stream)

    def finish(self) -> bytes:
        out = Output(self.out.ctypes.data, 0, self.capacity)
        s = self.lib.cujpegxl_encoder_finish(self.handle, C.byref(out))
        if s != OK:
            raise RuntimeError(f"finish failed: {status_string(s)}")
        return self.out[: out.size].tobytes()

    def encode_status(self, luma_ptr: int, chroma_ptr: int, luma_pitch: int, chroma_pitch: int,
                      capacity: int) -> tuple[int, int]:
        inp = Nv12Input(luma_ptr, chroma_ptr, luma_pitch, chroma_pitch)
        out = Output(self.out.ctypes.data, 0, capacity)
        s = self.lib.cujpegxl_encoder_encode(self.handle, C.byref(inp), C.byref(out))
        return s, out.size

    def fetch(self, name: str, dtype) -> np.ndarray:
        which = BUFFERS[name]
        n = self.lib.cjx_stage_buffer_bytes(self.handle, which)
        out = np.empty(n, np.uint8)
        if self.lib.cjx_stage_fetch(self.handle, which, out.ctypes.data, n) != OK:
            raise RuntimeError(f"fetch {name} failed")
        return out.view(dtype)

    def encode_coefficients(self, ac, dc, acs, qf, ytox, ytob, global_scale: int,
                            quant_dc: int) -> bytes:
        arrs = [np.ascontiguousarray(ac, np.int16), np.ascontiguousarray(dc, np.int32),
                np.ascontiguousarray(acs, np.int8), np.ascontiguousarray(qf, np.int32),
                None if ytox is None else np.ascontiguousarray(ytox, np.int8),
                None if ytob is None else np.ascontiguousarray(ytob, np.int8)]
        size = C.c_size_t(0)
        s = self.lib.cjx_stage_encode_coefficients(
            self.handle, *(None if a is None else a.ctypes.data for a in arrs), global_scale,
            quant_dc, self.out.ctypes.data, self.capacity, C.byref(size))
        if s != OK:
            raise RuntimeError(f"encode_coefficients failed: {status_string(s)}")
        return self.out[: size.value].tobytes()

    def set_profiling(self, on: bool, region_only: bool = False) -> None:
        """HIP-event timing of the following encodes: every stage, or with
        `region_only` the region kernel alone (fewer events on the stream)."""
        self.lib.cjx_stage_set_profiling(self.handle, (2 if region_only else 1) if on else 0)

    def set_mode(self, mode: int) -> None:
        """Encoder mode bits (MODE_MFMA16 | MODE_LEARNED; 0 for the plain
        path), from the next encode: the public library's
        cujpegxl_encoder_set_mode, or the stage library's switch, which also
        takes the forced-size test modes."""
        st = (self.lib.cjx_stage_set_mode(self.handle, mode) if self.lib is stages_lib()
              else self.lib.cujpegxl_encoder_set_mode(self.handle, mode))
        if st != OK:
            raise ValueError(f"bad mode {mode}")

    def export_features(self, on: bool = True) -> None:
        """From the next encode, the per-region network features and the
        heuristic's labels ([regions][108] floats, fetch("features"))."""
        if self.lib.cjx_stage_export_features(self.handle, int(on)) != OK:
            raise RuntimeError("feature export failed")

    def region_ms(self) -> float:
        """The region kernel's device time in the last encode (profiling on)."""
        ms = C.c_float(0.0)
        self.lib.cjx_stage_timings(self.handle, C.byref(ms), 1)
        return float(ms.value)

    def timings(self) -> dict[str, float]:
        ms = (C.c_float * len(STAGE_NAMES))()
        self.lib.cjx_stage_timings(self.handle, ms, len(STAGE_NAMES))
        return dict(zip(STAGE_NAMES, (float(x) for x in ms)))

    def close(self) -> None:
        if self.handle:
            self.lib.cujpegxl_encoder_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _planes(nv12: np.ndarray, width: int, height: int) -> np.ndarray:
    if not isinstance(nv12, np.ndarray) or nv12.dtype != np.uint8:
        raise ValueError("nv12 must be a uint8 numpy array")
    if nv12.size != width * height * 3 // 2:
        raise ValueError("nv12 array must hold width*height*3/2 bytes")
    if not nv12.flags.c_contiguous:
        raise ValueError("nv12 array must be C-contiguous")
    return nv12


def _encode(nv12, width, height, distance, device_ordinal, stats: bool):
    a = _planes(nv12, width, height)
    enc = Encoder(width, height, distance, device_ordinal, stages=stats)
    try:
        if stats:
            enc.set_profiling(True)
        data = enc.encode_host(a)
        t = enc.timings() if stats else None
    finally:
        enc.close()
    return data, t


def encode(nv12: np.ndarray, width: int, height: int, distance: float = 1.0,
           device_ordinal: int = 0) -> bytes:
    """Host NV12 (luma rows then interleaved CbCr rows) -> `.jxl` bytes."""
    return _encode(nv12, width, height, distance, device_ordinal, False)[0]


def encode_with_stats(nv12: np.ndarray, width: int, height: int, distance: float = 1.0,
                      device_ordinal: int = 0) -> tuple[bytes, list[dict]]:
    """As `encode`, plus stage records {name, bytes_moved, gpu_us, cpu_us} for
    the front end, the entropy coder and the assembly (pycujpegxl.cc:56-77);
    gpu_us is device time from HIP events, bytes_moved the stage's
    compulsory HBM traffic."""
    data, t = _encode(nv12, width, height, distance, device_ordinal, True)
    px = width * height
    fe = (t["region"] + t["quantize"]) * 1e3
    ent = (t["histograms"] + t["codes"] + t["sizes"] + t["emit"]) * 1e3
    asm = t["readback"] * 1e3
    return data, [
        {"name": "frontend", "bytes_moved": px * 3 // 2, "gpu_us": fe, "cpu_us": 0.0},
        {"name": "entropy", "bytes_moved": len(data), "gpu_us": ent, "cpu_us": 0.0},
        {"name": "assembly", "bytes_moved": len(data), "gpu_us": asm, "cpu_us": 0.0},
    ]


class Batch:
    """Frames of uneven sizes and distances encoded together (stage library):
    one launch per stage over all frames, captured once as a hipGraph."""

    def __init__(self, sizes: list[tuple[int, int]], distances: list[float], mode: int = 0,
                 device_ordinal: int = 0):
        self.lib = stages_lib()
        n = len(sizes)
        w = np.array([s[0] for s in sizes], np.uint32)
        h = np.array([s[1] for s in sizes], np.uint32)
        d = np.array(distances, np.float32)
        self.handle = C.c_void_p()
        st = self.lib.cjx_batch_create(device_ordinal, n, w.ctypes.data, h.ctypes.data, d.ctypes.data, mode,
                                       C.byref(self.handle))
        if st != OK:
            raise RuntimeError(f"batch create failed: {status_string(st)}")
        self.sizes = list(sizes)

    def set_inputs(self, inputs: list[tuple[int, int, int, int]]) -> None:
        """Per frame (luma, chroma, luma pitch, chroma pitch) device planes."""
        arr = (Nv12Input * len(inputs))(*(Nv12Input(*i) for i in inputs))
        if self.lib.cjx_batch_set_inputs(self.handle, arr) != OK:
            raise ValueError("bad batch inputs")

    def run(self, count: int = 1) -> float:
        sec = C.c_double(0.0)
        st = self.lib.cjx_batch_run(self.handle, count, C.byref(sec))
        if st != OK:
            raise RuntimeError(f"batch run failed: {status_string(st)}")
        return sec.value

    def launch(self, count: int = 1) -> None:
        """Queues `count` replays without waiting (see wait())."""
        if self.lib.cjx_batch_launch(self.handle, count) != OK:
            raise RuntimeError("batch launch failed")

    def wait(self) -> None:
        if self.lib.cjx_batch_wait(self.handle) != OK:
            raise RuntimeError("batch wait failed")

    def files(self) -> list[bytes]:
        out = []
        for i in range(len(self.sizes)):
            p, n = C.c_void_p(), C.c_uint64()
            if self.lib.cjx_batch_result(self.handle, i, C.byref(p), C.byref(n)) != OK:
                raise RuntimeError("batch result failed")
            out.append(C.string_at(p.value, n.value))
        return out

    def tune(self, reps: int = 5) -> str:
        buf = C.create_string_buffer(8192)
        if self.lib.cjx_batch_tune(self.handle, reps, buf, len(buf)) != OK:
            raise RuntimeError("batch tune failed")
        return buf.value.decode()

    def set_shape(self, token_cap: int, assemble_words: int, region_wgs: int = 0) -> None:
        """Launch shape: tokenizer workgroups per frame, output words per
        assembly thread, and 0 (one workgroup per region) or the size of a
        persistent region grid."""
        if self.lib.cjx_batch_set_shape(self.handle, token_cap, assemble_words, region_wgs) != OK:
            raise ValueError("bad batch shape")

    def close(self) -> None:
        if self.handle:
            self.lib.cjx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
