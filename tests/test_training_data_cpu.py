"""The training-pair file format (SMOKETD1) and its bfloat16 rule, on the CPU.

The reader is pinned on a file assembled here byte by byte from the layout
documented with smoke_simulation_ext_write_training_crops, and the
round-to-nearest-even conversion on closed-form cases (exact values, the two
tie directions, overflow to infinity, signed zero, NaN).
"""
import json
import struct

import numpy as np

import training_data as td


def test_bf16_round_to_nearest_even_cases():
    f = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0, -0.0, np.inf,
                  3.4028235e38, np.nan], np.float32)
    got = td.float32_to_bf16(f)
    # 1 + 2^-8 ties to the even 1.0; 1 + 3*2^-8 ties to the even 1 + 2^-6;
    # just above a tie rounds up; the largest float overflows to +inf
    want = [0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x0000, 0x8000, 0x7F80, 0x7F80]
    assert list(got[:9]) == want
    assert np.isnan(td.bf16_to_float32(got[9:]))[0]
    # bf16 values convert exactly both ways
    exact = np.array([0.5, -3.0, 1.0 + 2.0 ** -7, 2.0 ** -126], np.float32)
    np.testing.assert_array_equal(td.bf16_to_float32(td.float32_to_bf16(exact)), exact)
    # relative rounding error at most 2^-8 for normal numbers
    x = np.random.default_rng(0).normal(size=10000).astype(np.float32)
    back = td.bf16_to_float32(td.float32_to_bf16(x))
    assert np.max(np.abs(back - x) / np.abs(x)) <= 2.0 ** -8


def test_reader_on_a_hand_assembled_file(tmp_path):
    crop, grid = (3, 2, 2), (8, 6, 5)
    cells = 12
    path = tmp_path / "pairs.bin"
    header = b"SMOKETD1" + struct.pack("<I3I3II6I", 1, *crop, *grid, 32 + 5 * cells, *([0] * 6))
    assert len(header) == 64
    rng = np.random.default_rng(1)
    recs, expect = b"", []
    for k in range(2):
        r = rng.normal(size=cells).astype(np.float32)
        z = rng.normal(size=cells).astype(np.float32)
        fl = rng.integers(0, 128, size=cells).astype(np.uint8)
        origin, step, scale = (k, 2 * k, 1), 7 + k, 0.25 * (k + 1)
        recs += struct.pack("<3iId2I", *origin, step, scale, 0, 0)
        recs += td.float32_to_bf16(r).tobytes() + td.float32_to_bf16(z).tobytes() + fl.tobytes()
        expect.append((origin, step, scale, r, z, fl))
    path.write_bytes(header + recs)
    (tmp_path / "pairs.bin.jsonl").write_text(json.dumps({"offset": 64}) + "\n")
    d = td.read(path)
    assert d["crop"] == crop and d["grid"] == grid and d["r"].shape == (2, 2, 2, 3)
    for k, (origin, step, scale, r, z, fl) in enumerate(expect):
        assert tuple(d["origin"][k]) == origin and d["step"][k] == step and d["scale"][k] == scale
        np.testing.assert_array_equal(d["r_bits"][k].ravel(), td.float32_to_bf16(r))
        np.testing.assert_array_equal(d["z"][k].ravel(), td.bf16_to_float32(td.float32_to_bf16(z)))
        np.testing.assert_array_equal(d["flags"][k].ravel(), fl)
    assert td.read_manifest(path) == [{"offset": 64}]
