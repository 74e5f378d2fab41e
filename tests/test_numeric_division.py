"""Host check of the reciprocal-based division the quantizer uses.

The quantizer's y_round = qy * w / qgsf is computed as
div_by_recip(RN(qy * w), qgsf, RN(1 / qgsf)) (cujpegxl_amd/csrc/numeric.h):
two Markstein corrections after the reciprocal product, which give the
correctly rounded quotient when the quotient and the remainders stay normal.
A small host program built from the same header (contraction off, as the
library) compares it with the IEEE quotient over every qy in
[-2048, 2048], a logarithmic spread of weights and quant-field scales
(qf in 1..255 times global scales k / 65536), and the operands' signs and
zeros.  The oracle computes the plain quotient, so a mismatch here would
be a parity break.
"""

from __future__ import annotations

import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <cstring>

#include "numeric.h"

int main() {
    unsigned long long checked = 0, bad = 0;
    float weights[256];
    for (int i = 0; i < 256; ++i) {
        // 2^-12 .. 2^8, with a non-trivial mantissa at every step.
        weights[i] = std::ldexp(1.0f + static_cast<float>((i * 37) % 256) / 256.0f, -12 + (i * 20) / 256);
    }
    float scales[96];
    for (int i = 0; i < 96; ++i) {
        const int qf = 1 + (i * 53) % 255;
        const int k = 1 + (i * 7919) % 65535;
        scales[i] = static_cast<float>(qf) * (static_cast<float>(k) * (1.0f / 65536.0f));
    }
    for (float q : scales) {
        const float r = 1.0f / q;
        for (float w : weights) {
            for (int qy = -2048; qy <= 2048; ++qy) {
                const float a = static_cast<float>(qy) * w;
                const float want = a / q;
                const float got = cjx::div_by_recip(a, q, r);
                ++checked;
                // Equal values (a zero quotient may differ in sign only).
                if (!(got == want)) {
                    if (++bad <= 5) {
                        std::printf("mismatch a=%a q=%a got=%a want=%a\n", a, q, got, want);
                    }
                }
            }
        }
    }
    std::printf("checked %llu bad %llu\n", checked, bad);
    return bad == 0 ? 0 : 1;
}
"""


@pytest.mark.timeout(300)
def test_div_by_recip_is_the_correctly_rounded_quotient(tmp_path):
    if not pathlib.Path(HIPCC).exists():
        pytest.skip("hipcc is not available")
    src = tmp_path / "div_check.cpp"
    exe = tmp_path / "div_check"
    src.write_text(PROGRAM)
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-fno-fast-math", f"-I{ROOT / 'cujpegxl_amd' / 'csrc'}", str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True, timeout=240)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " bad 0" in out.stdout, out.stdout
