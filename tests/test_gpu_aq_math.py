"""The region kernel's AQ terms use the compiler's correctly rounded division
and square root without their operand scaling and special-case steps
(identities for the terms' moderate operands).  The stage library's check
evaluates both forms on the GPU for every float input of the terms' domains
(ratio_cbrt_gamma: v in [0, 4), both directions; masking_sqrt: v in
[0, 0.25)) and counts the results whose bits differ."""

from __future__ import annotations

import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_aq_division_and_square_root_equal_the_correctly_rounded_forms():
    import cujpegxl_amd as cj

    L = cj.stages_lib()
    L.cjx_stage_check_aq_math.argtypes = [C.c_int32, C.POINTER(C.c_uint64)]
    L.cjx_stage_check_aq_math.restype = C.c_int
    n = C.c_uint64(1 << 40)
    print("[gpu] AQ math check over 1,082,130,432 inputs", flush=True)
    assert L.cjx_stage_check_aq_math(0, C.byref(n)) == 0
    assert n.value == 0
