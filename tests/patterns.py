"""Deterministic coefficient frames from the reference's own tests.

Restates the input patterns of entropy_gpu_test.cc:36-111,
dc_entropy_gpu_test.cc:25-95 and frame_assembly_test.cc:20-46 so the same
frames drive the oracle, oracle/_ref and the device path.
"""

from __future__ import annotations

import numpy as np


def covered_slot(side: int, bx: int, by: int, bw: int, raw: int) -> int:
    """vardct_layout.h:54-62."""
    s = side // 8
    q = raw // 64
    return ((by + q // s) * bw + (bx + q % s)) * 64 + raw % 64


class Coeffs:
    def __init__(self, w: int, h: int, rqf: int = 32):
        self.w, self.h = w, h
        self.bw, self.bh = w // 8, h // 8
        self.gs, self.qdc, self.rqf = 4096, 32, rqf
        blocks = self.bw * self.bh
        self.dc = np.zeros((3, blocks), np.int32)
        self.ac = np.zeros((3, blocks * 64), np.int32)
        self.acs = None
        self.ytox = None
        self.ytob = None

    @property
    def qf(self) -> np.ndarray:
        return np.full(self.bw * self.bh, self.rqf, np.int32)

    def set_first_block(self, side: int, bx: int, by: int) -> None:
        s = side // 8
        for dy in range(s):
            for dx in range(s):
                self.acs[(by + dy) * self.bw + bx + dx] = side if dx == 0 and dy == 0 else 0


def ac_pattern(fc: Coeffs) -> None:
    """entropy_gpu_test.cc:53-69."""
    for by in range(fc.bh):
        for bx in range(fc.bw):
            b = (by * fc.bw + bx) * 64
            for c in range(3):
                fc.ac[c, b + 1] = (bx + c) % 7 - 3
                fc.ac[c, b + 8] = (by * 2 + c) % 11 - 5
                fc.ac[c, b + 9] = (bx + by) % 5 - 2
                fc.ac[c, b + 16] = (bx * 3 + by) % 260 - 130
                fc.ac[c, b + 2] = (by + c) % 3 - 1


def mixed_acs(fc: Coeffs) -> None:
    fc.acs = np.full(fc.bw * fc.bh, 8, np.int8)
    fc.set_first_block(32, 0, 0)
    fc.set_first_block(16, 4, 0)
    fc.set_first_block(16, 0, 4)


def mixed_ac_frame(w: int, h: int) -> Coeffs:
    """entropy_gpu_test.cc:87-111."""
    fc = Coeffs(w, h)
    mixed_acs(fc)
    for by in range(fc.bh):
        for bx in range(fc.bw):
            side = int(fc.acs[by * fc.bw + bx])
            if side == 0:
                continue
            for raw in range(1, side * side, 3):
                slot = covered_slot(side, bx, by, fc.bw, raw)
                for c in range(3):
                    fc.ac[c, slot] = (raw + bx + by + c) % 9 - 4
    return fc


def dct8_ac_frame(w: int, h: int) -> Coeffs:
    fc = Coeffs(w, h)
    ac_pattern(fc)
    fc.acs = np.full(fc.bw * fc.bh, 8, np.int8)
    return fc


def dc_pattern(fc: Coeffs) -> None:
    """dc_entropy_gpu_test.cc:42-53."""
    by, bx = np.mgrid[0 : fc.bh, 0 : fc.bw]
    fc.dc[0] = ((bx * 3 + by) % 517 - 258).ravel()
    fc.dc[1] = ((bx + by * 2) % 41 - 20).ravel()
    fc.dc[2] = ((bx + by) % 7 - 3).ravel()


def mixed_dc_frame(w: int, h: int) -> Coeffs:
    """dc_entropy_gpu_test.cc:76-95."""
    fc = Coeffs(w, h)
    mixed_acs(fc)
    dc_pattern(fc)
    tiles = ((fc.bw + 7) // 8) * ((fc.bh + 7) // 8)
    i = np.arange(tiles)
    fc.ytox = (i % 13 - 6).astype(np.int8)
    fc.ytob = (i % 9 - 4).astype(np.int8)
    return fc


def dct8_dc_frame(w: int, h: int) -> Coeffs:
    fc = Coeffs(w, h)
    dc_pattern(fc)
    fc.acs = np.full(fc.bw * fc.bh, 8, np.int8)
    return fc


def assembly_frame(w: int, h: int) -> Coeffs:
    """frame_assembly_test.cc:20-46 (all DCT8, no CfL maps)."""
    fc = Coeffs(w, h)
    for by in range(fc.bh):
        for bx in range(fc.bw):
            b = by * fc.bw + bx
            for c in range(3):
                fc.dc[c, b] = (bx * 3 + by + c) % 517 - 258
                fc.ac[c, b * 64 + 1] = (bx + c) % 7 - 3
                fc.ac[c, b * 64 + 8] = (by + c) % 5 - 2
                fc.ac[c, b * 64 + 9] = (bx + by) % 11 - 5
    return fc
