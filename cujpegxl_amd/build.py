"""Builds the gfx950 libraries in-tree with hipcc.

    lib/libcujpegxl.so.1        the cujpegxl C ABI (exports cujpegxl_* only)
    lib/libcujpegxl_stages.so   the same plus the cjx_* stage harness

Every translation unit is compiled as HIP for gfx950 with contraction off,
correctly rounded f32 division and square root, and f32 denormals kept: the
float results then match the oracle's bit for bit.
"""

from __future__ import annotations

import concurrent.futures
import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
OBJ = PKG / "build" / "obj"
LIB = PKG / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CJX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["device.hip", "tables.cc", "encoder.cc", "capi.cc", "noalloc.cc", "batch.cc", "stages.cc"]

CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{SRC}",
]


def _compile(src: str) -> pathlib.Path:
    out = OBJ / (src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(SRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    warn = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if warn:
        sys.stderr.write("\n".join(warn) + "\n")
    return out


def build(jobs: int = 8) -> list[pathlib.Path]:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    by_name = {o.stem: o for o in objs}
    core = [by_name[n] for n in ("device", "tables", "encoder")]
    outs = []
    for name, extra, script, soname in (
        ("libcujpegxl.so.1", [by_name["capi"], by_name["noalloc"]], "libcujpegxl.lds", "libcujpegxl.so.1"),
        ("libcujpegxl_stages.so", [by_name["capi"], by_name["noalloc"], by_name["batch"], by_name["stages"]], "libcujpegxl_stages.lds",
         "libcujpegxl_stages.so"),
    ):
        target = LIB / name
        tmp = target.with_suffix(target.suffix + ".tmp")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, core),
               *map(str, extra), f"-Wl,--version-script={SRC / script}", f"-Wl,-soname,{soname}"]
        subprocess.run(cmd, check=True)
        os.replace(tmp, target)
        outs.append(target)
    return outs


if __name__ == "__main__":
    for p in build():
        print(p)
