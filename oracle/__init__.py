"""Test infrastructure only: ctypes access to the CPU oracle and to oracle/_ref.

`lib()` is the C restatement (oracle/build/libjxloracle.so); `ref()` is the
reference's own host code compiled from source (oracle/_ref/libjxlref.so),
present only where the reference tree was available to build it.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use this package.
"""

from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent
LIB_PATH = HERE / "build" / "libjxloracle.so"
REF_PATH = HERE / "_ref" / "libjxlref.so"
WEIGHTS_PATH = REPO / "tests" / "golden" / "dequant_weights.bin"

_P = C.c_void_p
_SZ = C.c_size_t
_U32 = C.c_uint32
_F = C.c_float

_lib = None
_ref = None


def build(with_ref: bool = False) -> None:
    """Compiles the oracle (and `_ref` when asked and the reference exists)."""
    target = ["ref"] if with_ref else ["all"]
    subprocess.run(["make", "-s", "-C", str(HERE), *target], check=True)


def set_threads(n: int) -> None:
    """Threads for the oracle's front-end loops (1 by default)."""
    lib().jo_set_threads(int(n))


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def dequant_weights() -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """libjxl default dequant matrices [3][n*n] for n = 8, 16, 32 (fixture)."""
    raw = np.fromfile(WEIGHTS_PATH, dtype="<f4")
    return raw[:192].reshape(3, 64), raw[192:960].reshape(3, 256), raw[960:].reshape(3, 1024)


class Cal(C.Structure):
    _fields_ = [
        ("global_scale", _U32),
        ("quant_dc", _U32),
        ("raw_quant_field", _U32),
        ("global_scale_float", _F),
        ("inv_global_scale", _F),
    ]


class Frame(C.Structure):
    _fields_ = [
        ("width", _U32),
        ("height", _U32),
        ("global_scale", _U32),
        ("quant_dc", _U32),
        ("dc", _P),
        ("ac", _P),
        ("acs", _P),
        ("qf", _P),
        ("ytox", _P),
        ("ytob", _P),
    ]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name in ("jo_cbrt", "jo_log2", "jo_exp2", "jo_pow24", "jo_exp"):
            getattr(L, name).restype = _F
            getattr(L, name).argtypes = [_F]
        L.jo_f32_to_f16.restype = C.c_uint16
        L.jo_f32_to_f16.argtypes = [_F]
        L.jo_f16_to_f32.restype = _F
        L.jo_f16_to_f32.argtypes = [C.c_uint16]
        L.jo_ac_cluster.restype = C.c_int
        L.jo_ac_cluster.argtypes = [_U32]
        L.jo_calibrate.argtypes = [_F, C.POINTER(Cal)]
        for name in (
            "jo_huffman",
            "jo_best_context_map",
            "jo_ac_groups",
            "jo_dc_groups",
            "jo_dc_global",
            "jo_ac_global",
            "jo_codestream_head",
            "jo_container_framing",
            "jo_write_codestream",
            "jo_encode_nv12",
            "jo_encode_dct8_uniform",
            "jo_max_output_size",
        ):
            getattr(L, name).restype = _SZ
        L.jo_nv12_to_xyb.argtypes = [_P, _SZ, _P, _SZ, _U32, _U32, _P]
        L.jo_chroma_upsample.argtypes = [_P, _SZ, _U32, _U32, _P, _P]
        L.jo_nv12_to_xyb_f64.argtypes = [_P, _SZ, _P, _SZ, _U32, _U32, _P]
        L.jo_gaborish.argtypes = [_P, _U32, _U32, _P]
        L.jo_adaptive_quant.argtypes = [_P, _U32, _U32, _F, _P]
        L.jo_select_transforms.argtypes = [_P, _U32, _U32, _F, _P]
        L.jo_forward_dct.argtypes = [_P, _U32, _U32, _P, _P]
        L.jo_forward_dct_f64.argtypes = [_P, _U32, _U32, _P, _P]
        L.jo_mlp_forward.argtypes = [_P, _P, _U32, _P]
        L.jo_mlp_loss_grad.restype = C.c_double
        L.jo_mlp_loss_grad.argtypes = [_P, _P, _P, _U32, _F, _P]
        L.jo_block_costs8.argtypes = [_P, _U32, _U32, _F, _P, _P, _P]
        L.jo_cfl_region_partials.argtypes = [_P, _P, _U32, _U32, _P]
        L.jo_cfl_maps.argtypes = [_P, _U32, _U32, _P, _P]
        L.jo_quantize.argtypes = [_P, _P, _P, _P, _P, _U32, _U32, _F, _P, _P, _P]
        L.jo_quantize_dct8_uniform.argtypes = [_P, _U32, _U32, _F, _P, _P]
        L.jo_dc_from_llf.argtypes = [_U32, _P, _P]
        L.jo_natural_order.argtypes = [_U32, _P]
        L.jo_dct_basis.argtypes = [_U32, _P]
        L.jo_set_dequant.argtypes = [_P, _P, _P]
        L.jo_set_threads.argtypes = [C.c_int]
        L.jo_set_threads.restype = None
        L.jo_huffman.argtypes = [_P, _U32, _P, _P, _P, _SZ, C.POINTER(_SZ)]
        L.jo_best_context_map.argtypes = [_P, _U32, _U32, _P, _SZ, C.POINTER(_SZ),
                                          C.POINTER(C.c_int)]
        L.jo_move_to_front.argtypes = [_P, _U32, _P]
        L.jo_ac_histogram.argtypes = [C.POINTER(Frame), _P]
        L.jo_dc_histograms.argtypes = [C.POINTER(Frame), _P, _P]
        L.jo_ac_groups.argtypes = [C.POINTER(Frame), _P, _P, _P, _P, _SZ]
        L.jo_dc_groups.argtypes = [C.POINTER(Frame), _P, _P, _SZ]
        L.jo_dc_global.argtypes = [_U32, _U32, _P, _SZ]
        L.jo_ac_global.argtypes = [_P, _U32, _P, _P, _P, _SZ]
        L.jo_codestream_head.argtypes = [_U32, _U32, _P, _U32, _P, _SZ]
        L.jo_container_framing.argtypes = [C.c_uint64, _P, _SZ]
        L.jo_write_codestream.argtypes = [C.POINTER(Frame), _P, _SZ]
        L.jo_encode_nv12.argtypes = [_P, _SZ, _P, _SZ, _U32, _U32, _F, _P, _SZ]
        L.jo_encode_dct8_uniform.argtypes = [_P, _SZ, _P, _SZ, _U32, _U32, _F, _P, _SZ]
        L.jo_max_output_size.argtypes = [_U32, _U32]
        L.jd_probe.argtypes = [_P, _SZ, C.POINTER(_U32), C.POINTER(_U32)]
        L.jd_decode.argtypes = [_P, _SZ, _U32, _U32, C.POINTER(_U32), C.POINTER(_U32), _P, _P, _P, _P,
                                _P, _P]
        L.jd_error.restype = C.c_char_p
        w8, w16, w32 = (np.ascontiguousarray(w, dtype=np.float32) for w in dequant_weights())
        L.jo_set_dequant(ptr(w8), ptr(w16), ptr(w32))
        _lib = L
    return _lib


def ref() -> C.CDLL | None:
    """The reference's own host code, or None where it was not built."""
    global _ref
    if _ref is None:
        if not REF_PATH.exists():
            return None
        R = C.CDLL(str(REF_PATH))
        for name in (
            "ref_huffman",
            "ref_best_context_map",
            "ref_write_codestream",
            "ref_ac_encode",
            "ref_dc_encode",
            "ref_dc_global",
            "ref_ac_global",
            "ref_codestream_head",
            "ref_container_framing",
        ):
            getattr(R, name).restype = _SZ
        R.ref_natural_coeff_order.argtypes = [_U32, _P]
        R.ref_calibrate.argtypes = [_F, _P, _P]
        R.ref_dequant_tables.argtypes = [_P, _P, _P]
        R.ref_select_transforms.argtypes = [_P, _U32, _U32, _F, _P]
        R.ref_block_ac_cost.restype = _F
        R.ref_block_ac_cost.argtypes = [_P, _U32, _U32, _U32, C.c_int, _F]
        R.ref_cfl_maps_from_sums.argtypes = [C.c_double, C.c_double, C.c_double,
                                             C.POINTER(C.c_int), C.POINTER(C.c_int)]
        R.ref_dc_from_llf.argtypes = [_U32, _P, _P]
        R.ref_ac_cluster.restype = C.c_int
        R.ref_ac_cluster.argtypes = [_U32]
        R.ref_zero_density_context.restype = _U32
        R.ref_zero_density_context.argtypes = [_U32] * 5
        R.ref_nonzero_context.restype = _U32
        R.ref_nonzero_context.argtypes = [_U32, C.c_int]
        R.ref_block_context.restype = C.c_int
        R.ref_block_context.argtypes = [C.c_int, C.c_int]
        R.ref_huffman.argtypes = [_P, _U32, _P, _P, _P, _SZ, C.POINTER(_SZ)]
        R.ref_best_context_map.argtypes = [_P, _U32, _U32, _P, _SZ, C.POINTER(_SZ),
                                           C.POINTER(C.c_int)]
        R.ref_move_to_front.argtypes = [_P, _U32, _P]
        R.ref_write_codestream.argtypes = [_U32, _U32, _U32, _U32, _U32, _P, _P, _P, _P, _P,
                                           C.c_int, _P, _SZ]
        R.ref_ac_encode.argtypes = [_U32, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _SZ]
        R.ref_dc_encode.argtypes = [_U32, _U32, _U32, _P, _P, _P, _P, _P, _P, _SZ]
        R.ref_dc_global.argtypes = [_U32, _U32, _P, _SZ]
        R.ref_ac_global.argtypes = [_P, _U32, _P, _P, _P, _SZ]
        R.ref_codestream_head.argtypes = [_U32, _U32, _P, _U32, _P, _SZ]
        R.ref_container_framing.argtypes = [C.c_uint64, _P, _SZ]
        _ref = R
    return _ref


# ---- numpy-level helpers over lib() ----------------------------------------


def calibrate(distance: float) -> Cal:
    c = Cal()
    lib().jo_calibrate(distance, C.byref(c))
    return c


def nv12_to_xyb(luma: np.ndarray, chroma: np.ndarray, w: int, h: int) -> np.ndarray:
    out = np.empty((3, h, w), np.float32)
    lib().jo_nv12_to_xyb(ptr(luma), w, ptr(chroma), w, w, h, ptr(out))
    return out


def nv12_to_xyb_f64(luma: np.ndarray, chroma: np.ndarray, w: int, h: int) -> np.ndarray:
    out = np.empty((3, h, w), np.float64)
    lib().jo_nv12_to_xyb_f64(ptr(luma), w, ptr(chroma), w, w, h, ptr(out))
    return out


def gaborish(xyb: np.ndarray) -> np.ndarray:
    _, h, w = xyb.shape
    out = np.empty_like(xyb)
    lib().jo_gaborish(ptr(xyb), w, h, ptr(out))
    return out


def adaptive_quant(xyb: np.ndarray, distance: float) -> np.ndarray:
    _, h, w = xyb.shape
    out = np.empty((h // 8) * (w // 8), np.int32)
    lib().jo_adaptive_quant(ptr(xyb), w, h, distance, ptr(out))
    return out


def select_transforms(y: np.ndarray, distance: float) -> np.ndarray:
    h, w = y.shape
    out = np.full((h // 8) * (w // 8), -1, np.int8)
    y = np.ascontiguousarray(y, np.float32)
    lib().jo_select_transforms(ptr(y), w, h, distance, ptr(out))
    return out


def forward_dct(planes: np.ndarray, acs: np.ndarray | None) -> np.ndarray:
    _, h, w = planes.shape
    out = np.zeros(3 * (h // 8) * (w // 8) * 64, np.float32)
    lib().jo_forward_dct(ptr(planes), w, h, ptr(acs), ptr(out))
    return out


def forward_dct_f64(planes: np.ndarray, acs: np.ndarray | None) -> np.ndarray:
    _, h, w = planes.shape
    out = np.zeros(3 * (h // 8) * (w // 8) * 64, np.float64)
    planes = np.ascontiguousarray(planes, np.float64)
    lib().jo_forward_dct_f64(ptr(planes), w, h, ptr(acs), ptr(out))
    return out


def dct_basis(n: int) -> np.ndarray:
    """A[k][t] = g_k cos(pi (t + 1/2) k / n) as the path uses it (float32)."""
    out = np.zeros(n * n, np.float32)
    lib().jo_dct_basis(n, ptr(out))
    return out


def to_f16(a: np.ndarray) -> np.ndarray:
    """Binary16 bits, round to nearest even with subnormals (numpy's cast, which
    tests pin against the oracle's own jo_f32_to_f16)."""
    return np.ascontiguousarray(a, np.float32).astype(np.float16).view(np.uint16).ravel()


def cfl(c16: np.ndarray, acs: np.ndarray, w: int, h: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    bw, bh = w // 8, h // 8
    rw, rh = (bw + 3) // 4, (bh + 3) // 4
    part = np.zeros(rw * rh * 3, np.float32)
    lib().jo_cfl_region_partials(ptr(c16), ptr(acs), w, h, ptr(part))
    tiles = ((bw + 7) // 8) * ((bh + 7) // 8)
    ytox = np.zeros(tiles, np.int8)
    ytob = np.zeros(tiles, np.int8)
    lib().jo_cfl_maps(ptr(part), w, h, ptr(ytox), ptr(ytob))
    return part, ytox, ytob


def quantize(c16, acs, ytox, ytob, qf, w, h, distance):
    nblk = (w // 8) * (h // 8)
    ac = np.zeros(3 * nblk * 64, np.int16)
    dc = np.zeros(3 * nblk, np.int32)
    nz = np.zeros(3 * nblk, np.uint8)
    lib().jo_quantize(ptr(c16), ptr(acs), ptr(ytox), ptr(ytob), ptr(qf), w, h, distance,
                      ptr(ac), ptr(dc), ptr(nz))
    return ac, dc, nz


def make_frame(w, h, gs, qdc, dc, ac, acs, qf, ytox=None, ytob=None):
    """A jo_frame over numpy arrays (kept alive by the returned tuple)."""
    keep = [np.ascontiguousarray(dc, np.int32), np.ascontiguousarray(ac, np.int32),
            None if acs is None else np.ascontiguousarray(acs, np.int8),
            np.ascontiguousarray(qf, np.int32),
            None if ytox is None else np.ascontiguousarray(ytox, np.int8),
            None if ytob is None else np.ascontiguousarray(ytob, np.int8)]
    f = Frame(w, h, gs, qdc, *(ptr(a) for a in keep))
    return f, keep


def write_codestream(frame) -> bytes:
    f, _keep = frame
    cap = lib().jo_max_output_size(f.width, f.height) + 65536
    out = np.zeros(cap, np.uint8)
    n = lib().jo_write_codestream(C.byref(f), ptr(out), cap)
    return out[:n].tobytes()


def encode_nv12(luma: np.ndarray, chroma: np.ndarray, w: int, h: int, distance: float) -> bytes:
    cap = lib().jo_max_output_size(w, h) + 65536
    out = np.zeros(cap, np.uint8)
    n = lib().jo_encode_nv12(ptr(luma), w, ptr(chroma), w, w, h, distance, ptr(out), cap)
    return out[:n].tobytes()


def encode_dct8_uniform(luma, chroma, w, h, distance) -> bytes:
    cap = lib().jo_max_output_size(w, h) + 65536
    out = np.zeros(cap, np.uint8)
    n = lib().jo_encode_dct8_uniform(ptr(luma), w, ptr(chroma), w, w, h, distance, ptr(out), cap)
    return out[:n].tobytes()


class DecodeError(ValueError):
    """The file is not a codestream of the subset the encoder writes."""


def decode(data: bytes) -> dict:
    """Subset decoder (jxl_decode.c): `.jxl` bytes -> the quantized frame
    {width, height, global_scale, quant_dc, acs, qf, dc, ac, ytox, ytob} with
    the encoder's layouts (qf at first blocks, zero elsewhere)."""
    buf = np.frombuffer(data, np.uint8)
    w, h = _U32(), _U32()
    if lib().jd_probe(ptr(buf), buf.size, C.byref(w), C.byref(h)) != 0:
        raise DecodeError(lib().jd_error().decode())
    w, h = w.value, h.value
    nblk = (w // 8) * (h // 8)
    tiles = ((w // 8 + 7) // 8) * ((h // 8 + 7) // 8)
    out = {"acs": np.zeros(nblk, np.int8), "qf": np.zeros(nblk, np.int32),
           "dc": np.zeros(3 * nblk, np.int32), "ac": np.zeros(3 * nblk * 64, np.int16),
           "ytox": np.zeros(tiles, np.int8), "ytob": np.zeros(tiles, np.int8)}
    gs, qdc = _U32(), _U32()
    if lib().jd_decode(ptr(buf), buf.size, w, h, C.byref(gs), C.byref(qdc), ptr(out["acs"]),
                       ptr(out["qf"]), ptr(out["dc"]), ptr(out["ac"]), ptr(out["ytox"]),
                       ptr(out["ytob"])) != 0:
        raise DecodeError(lib().jd_error().decode())
    out.update(width=w, height=h, global_scale=gs.value, quant_dc=qdc.value)
    return out


MLP_IN, MLP_HID, MLP_OUT, MLP_LABELS = 96, 64, 7, 12
MLP_PARAMS = MLP_HID * MLP_IN + MLP_HID + MLP_OUT * MLP_HID + MLP_OUT


def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Network outputs [n][7] in float64 (jxl_mlp.c)."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
    z = np.zeros((x.shape[0], MLP_OUT), np.float64)
    lib().jo_mlp_forward(ptr(params), ptr(x), x.shape[0], ptr(z))
    return z


def mlp_loss_grad(params: np.ndarray, x: np.ndarray, labels: np.ndarray, lam: float,
                  want_grad: bool = True) -> tuple[float, np.ndarray | None]:
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
    labels = np.ascontiguousarray(labels, np.float32).reshape(-1, MLP_LABELS)
    grad = np.zeros(MLP_PARAMS, np.float64) if want_grad else None
    loss = lib().jo_mlp_loss_grad(ptr(params), ptr(x), ptr(labels), x.shape[0], lam, ptr(grad))
    return float(loss), grad


def block_costs8(y: np.ndarray, distance: float, parts: bool = False):
    """Per-block DCT8 RD costs of the selection; with `parts`, also the
    squared-error and rate sums they are made of."""
    h, w = y.shape
    n = (h // 8) * (w // 8)
    out, dist, rate = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    lib().jo_block_costs8(ptr(y), w, h, distance, ptr(out), ptr(dist), ptr(rate))
    return (out, dist, rate) if parts else out
