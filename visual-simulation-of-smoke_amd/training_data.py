"""Reader of the training-pair files written by
smoke_simulation_ext_write_training_crops ("SMOKETD1").

Each record holds one crop of a pressure solve: r = b / scale and
z = x / scale as bfloat16 (x fastest), the operator's flag byte per cell
(bits 0-5: fluid neighbour x-, x+, y-, y+, z-, z+; bit 6: identity row), the
crop's lower corner, the step it was taken at and `scale` (the RMS of b over
the crop).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

MAGIC = b"SMOKETD1"
HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("crop", "<u4", 3), ("grid", "<u4", 3), ("record_bytes", "<u4"),
                   ("reserved", "<u4", 6)])


def record_dtype(cells: int) -> np.dtype:
    return np.dtype([("origin", "<i4", 3), ("step", "<u4"), ("scale", "<f8"), ("reserved", "<u4", 2),
                     ("r", "<u2", cells), ("z", "<u2", cells), ("flags", "u1", cells)])


def bf16_to_float32(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def float32_to_bf16(values: np.ndarray) -> np.ndarray:
    """Round to nearest even, NaN kept quiet -- the device packer's rule."""
    u = np.ascontiguousarray(values, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return np.where(nan, (u >> 16) | 0x40, rounded).astype(np.uint16)


def read(path) -> dict:
    """All records of a file: origin (n, 3), step (n,), scale (n,), and r, z
    (float32) and flags (uint8) shaped (n, cz, cy, cx)."""
    raw = Path(path).read_bytes()
    h = np.frombuffer(raw, HEADER, count=1)[0]
    if h["magic"] != MAGIC or h["version"] != 1:
        raise ValueError(f"{path}: not a SMOKETD1 file")
    crop = tuple(int(v) for v in h["crop"])
    cells = crop[0] * crop[1] * crop[2]
    rec = record_dtype(cells)
    if rec.itemsize != int(h["record_bytes"]):
        raise ValueError(f"{path}: record size {h['record_bytes']} != {rec.itemsize}")
    body = raw[HEADER.itemsize:]
    if len(body) % rec.itemsize:
        raise ValueError(f"{path}: truncated record")
    recs = np.frombuffer(body, rec)
    shape = (len(recs), crop[2], crop[1], crop[0])
    return {
        "crop": crop, "grid": tuple(int(v) for v in h["grid"]),
        "origin": recs["origin"].copy(), "step": recs["step"].copy(), "scale": recs["scale"].copy(),
        "r_bits": recs["r"].reshape(shape).copy(), "z_bits": recs["z"].reshape(shape).copy(),
        "r": bf16_to_float32(recs["r"]).reshape(shape), "z": bf16_to_float32(recs["z"]).reshape(shape),
        "flags": recs["flags"].reshape(shape).copy(),
    }


def read_manifest(path) -> list[dict]:
    return [json.loads(line) for line in Path(str(path) + ".jsonl").read_text().splitlines() if line.strip()]
