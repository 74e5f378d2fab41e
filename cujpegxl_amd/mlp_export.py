"""Writes the selection network's weights as the C++ include the library is
built with (csrc/mlp_weights.inc): FP16 W1 [64][96] and W2 [7][64] bit
patterns, FP32 biases and the confidence threshold (a logit magnitude)."""

from __future__ import annotations

import pathlib

import numpy as np

HID, IN, OUT = 64, 96, 7
INC = pathlib.Path(__file__).resolve().parent / "csrc" / "mlp_weights.inc"


def _rows(vals, per_line: int, fmt) -> str:
    items = [fmt(v) for v in vals]
    return ",\n".join("    " + ", ".join(items[i:i + per_line]) for i in range(0, len(items), per_line))


def write_inc(w1: np.ndarray, b1: np.ndarray, w2: np.ndarray, b2: np.ndarray, threshold: float,
              note: str, path: pathlib.Path = INC) -> None:
    w1h = np.asarray(w1, np.float32).reshape(HID, IN).astype(np.float16).view(np.uint16).ravel()
    w2h = np.asarray(w2, np.float32).reshape(OUT, HID).astype(np.float16).view(np.uint16).ravel()
    b1 = np.asarray(b1, np.float32).ravel()
    b2 = np.asarray(b2, np.float32).ravel()
    f32 = lambda v: f"{float(np.float32(v))!r}f" if np.isfinite(v) else "1.0e30f"
    text = (
        f"// The selection network's weights ({note}).\n"
        "// Written by cujpegxl_amd/mlp_export.py; FP16 bit patterns for W1 [64][96] and W2 [7][64].\n"
        f"constexpr std::uint16_t kTrainedW1[{HID * IN}] = {{\n{_rows(w1h, 12, lambda v: f'0x{int(v):04x}')}\n}};\n"
        f"constexpr float kTrainedB1[{HID}] = {{\n{_rows(b1, 6, f32)}\n}};\n"
        f"constexpr std::uint16_t kTrainedW2[{OUT * HID}] = {{\n{_rows(w2h, 12, lambda v: f'0x{int(v):04x}')}\n}};\n"
        f"constexpr float kTrainedB2[{OUT}] = {{\n{_rows(b2, 7, f32)}\n}};\n"
        f"constexpr float kTrainedThreshold = {f32(threshold)};\n")
    path.write_text(text)


def read_inc(path: pathlib.Path = INC) -> tuple[np.ndarray, float]:
    """The include's weights as one float32 parameter vector (W1, b1, W2,
    b2; the matrices as their FP16 values) and the threshold."""
    import re

    text = path.read_text()

    def block(name: str) -> list[str]:
        body = re.search(name + r"\[\d+\] = \{(.*?)\};", text, re.S).group(1)
        return [t.strip() for t in body.replace("\n", " ").split(",") if t.strip()]

    w1 = np.array([int(t, 16) for t in block("kTrainedW1")], np.uint16).view(np.float16).astype(np.float32)
    w2 = np.array([int(t, 16) for t in block("kTrainedW2")], np.uint16).view(np.float16).astype(np.float32)
    b1 = np.array([float(t.rstrip("f")) for t in block("kTrainedB1")], np.float32)
    b2 = np.array([float(t.rstrip("f")) for t in block("kTrainedB2")], np.float32)
    th = float(re.search(r"kTrainedThreshold = ([^;]+?)f?;", text).group(1))
    return np.concatenate([w1, b1, w2, b2]), th


if __name__ == "__main__":
    # Untrained: zero weights and an unreachable threshold (every region
    # falls back to the heuristic).
    write_inc(np.zeros(HID * IN), np.zeros(HID), np.zeros(OUT * HID), np.zeros(OUT), float("inf"),
              "untrained: every region falls back to the heuristic")
