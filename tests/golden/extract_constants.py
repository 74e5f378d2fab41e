"""Writes tests/golden/xyb_constants.json: the literal constants of the
reference's NV12 -> XYB transform, read from its source text.

Source: src/xyb_impl.cuh:17-35 of the reference (libjxl's opsin absorbance
matrix and bias, the BT.709 full-range inverse coefficients and the chroma
centre).  The values are stored as the decimal strings written there, plus
the float32 value each names, so a test can recompute XYB in float64 from
constants that do not come from this repository's code.

Run from the repo root, in a tree where the reference is present:
    python tests/golden/extract_constants.py
"""

from __future__ import annotations

import json
import pathlib
import re

import numpy as np

REF = pathlib.Path("/root/reference/src/xyb_impl.cuh")
OUT = pathlib.Path(__file__).resolve().parent / "xyb_constants.json"

NAMES = ("M00", "M02", "M10", "M12", "M20", "M21", "OPSIN_BIAS", "CR_TO_R", "CB_TO_G", "CR_TO_G",
         "CB_TO_B")
DERIVED = {"M01": "1.0f - M02 - M00", "M11": "1.0f - M12 - M10", "M22": "1.0f - M20 - M21",
           "CHROMA_CENTER": "128.0f / 255.0f"}


def main() -> None:
    text = REF.read_text()
    consts: dict[str, dict] = {}
    for name in NAMES:
        m = re.search(rf"constexpr\s+float\s+{name}\s*\{{\s*([0-9.eE+-]+)f\s*\}}", text)
        if m is None:
            raise SystemExit(f"{name} not found as a float literal in {REF}")
        lit = m.group(1)
        consts[name] = {"literal": lit, "f32": float(np.float32(float(lit)))}
    for name, expr in DERIVED.items():
        m = re.search(rf"constexpr\s+float\s+{name}\s*\{{\s*([^}}]+)\}}", text)
        if m is None or m.group(1).replace(" ", "") != expr.replace(" ", ""):
            raise SystemExit(f"{name}: expected '{expr}' in {REF}")
        consts[name] = {"expression": expr}
    OUT.write_text(json.dumps({"source": "src/xyb_impl.cuh:17-35", "constants": consts}, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
