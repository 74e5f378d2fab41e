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
OBJ_CUT = PKG / "build" / "obj_cut"
LIB_CUT = PKG / "lib_cut"
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


def _compile(src: str, obj_dir: pathlib.Path = OBJ, extra_flags: list[str] | None = None) -> pathlib.Path:
    out = obj_dir / (src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, *(extra_flags or []), "-c", str(SRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    warn = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if warn:
        sys.stderr.write("\n".join(warn) + "\n")
    return out


def build(jobs: int = 8, region_cut: bool = False) -> list[pathlib.Path]:
    """Builds both libraries into lib/.  region_cut=True instead builds the
    profiling variant (CJX_REGION_CUT: the region kernel's cut switch and the
    region-only stage entry) into lib_cut/, which only
    cujpegxl_amd/region_phases.py loads."""
    obj_dir, lib_dir = (OBJ_CUT, LIB_CUT) if region_cut else (OBJ, LIB)
    obj_dir.mkdir(parents=True, exist_ok=True)
    lib_dir.mkdir(parents=True, exist_ok=True)
    extra_flags = ["-DCJX_REGION_CUT"] if region_cut else []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, obj_dir, extra_flags), SOURCES))
    by_name = {o.stem: o for o in objs}
    core = [by_name[n] for n in ("device", "tables", "encoder")]
    outs = []
    for name, extra, script, soname in (
        ("libcujpegxl.so.1", [by_name["capi"], by_name["noalloc"]], "libcujpegxl.lds", "libcujpegxl.so.1"),
        ("libcujpegxl_stages.so", [by_name["capi"], by_name["noalloc"], by_name["batch"], by_name["stages"]], "libcujpegxl_stages.lds",
         "libcujpegxl_stages.so"),
    ):
        target = lib_dir / name
        tmp = target.with_suffix(target.suffix + ".tmp")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, core),
               *map(str, extra), f"-Wl,--version-script={SRC / script}", f"-Wl,-soname,{soname}"]
        subprocess.run(cmd, check=True)
        os.replace(tmp, target)
        outs.append(target)
    return outs


if __name__ == "__main__":
    for p in build(region_cut="--region-cut" in sys.argv[1:]):
        print(p)
