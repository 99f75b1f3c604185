"""The SMOKECK1 checkpoint format on the CPU: the library's host-side
validator (smoke_simulation_ext_inspect_checkpoint, the same parser
load_checkpoint runs before it touches the device) against files assembled
byte by byte here from the documented layout (tests/checkpoint_format.py), and
against every way such a file can be wrong."""
import struct

import numpy as np
import pytest

import checkpoint_format as ckf
import smoke_abi as abi


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not abi.LIB_PATH.exists():
        pytest.skip("library not built (run __graft_entry__.build())")
    abi.load_library()


def _write(tmp_path, blob, name="state.ck"):
    p = tmp_path / name
    p.write_bytes(blob)
    return p


def test_hand_assembled_file_is_accepted(tmp_path):
    dims = (5, 4, 3)
    st = ckf.random_state(*dims, seed=1)
    blob = ckf.assemble(*dims, st, solves_done=2)
    # header 28 B, 18 field headers of 32 B, the data, the trailing count
    data = sum(a.nbytes for a in st.values())
    assert len(blob) == 28 + 18 * 32 + data + 4
    rc, info = abi.inspect_checkpoint(_write(tmp_path, blob))
    assert rc == 0, info.error
    assert info.version == 2 and tuple(info.dims) == dims and info.field_count == 18 and info.solves_done == 2
    # the independent reader reads back what was assembled
    got_dims, got, solves = ckf.parse(blob)
    assert got_dims == dims and solves == 2
    for k in st:
        np.testing.assert_array_equal(got[k], st[k])


def _corrupt_cases(dims, st):
    good = ckf.assemble(*dims, st, solves_done=1)
    off_first_name = 28
    cases = {
        "truncated in the last field": good[:-10],
        "truncated header": good[:20],
        "missing warm-start count": good[:-4],
        "trailing byte": good + b"\0",
        "wrong magic": b"SMOKECK2" + good[8:],
        "wrong version": good[:8] + struct.pack("<I", 1) + good[12:],
        "zero dimension": good[:12] + struct.pack("<i", 0) + good[16:],
        "wrong count": good[:24] + struct.pack("<I", 17) + good[28:],
        "wrong first name": good[:off_first_name] + b"velocity_q".ljust(24, b"\0") + good[off_first_name + 24:],
        "wrong first size": good[:off_first_name + 24] + struct.pack("<Q", 12) + good[off_first_name + 32:],
        "bad warm-start count": good[:-4] + struct.pack("<i", 3),
    }
    # a file of other dimensions whose fields were written for these ones
    cases["dimensions disagree with the fields"] = good[:12] + struct.pack("<3i", dims[0] + 1, dims[1], dims[2]) + good[24:]
    return cases


@pytest.mark.parametrize("case", ["truncated in the last field", "truncated header", "missing warm-start count", "trailing byte",
                                  "wrong magic", "wrong version", "zero dimension", "wrong count", "wrong first name",
                                  "wrong first size", "bad warm-start count", "dimensions disagree with the fields"])
def test_malformed_files_are_rejected(tmp_path, case):
    dims = (4, 3, 2)
    blob = _corrupt_cases(dims, ckf.random_state(*dims, seed=2))[case]
    rc, info = abi.inspect_checkpoint(_write(tmp_path, blob))
    assert rc == 2, case     # SMOKE_SIMULATION_RESULT_BACKEND_FAILURE
    assert info.error.decode(), case
    with pytest.raises(ValueError):
        ckf.parse(blob)


def test_fields_out_of_order_are_rejected(tmp_path):
    dims = (3, 3, 3)
    st = ckf.random_state(*dims, seed=3)
    blob = bytearray(ckf.assemble(*dims, st))
    # swap the names of two equal-sized fields: sizes still match, names do not
    i = blob.index(b"density".ljust(24, b"\0"))
    j = blob.index(b"temperature".ljust(24, b"\0"))
    blob[i:i + 24], blob[j:j + 24] = b"temperature".ljust(24, b"\0"), b"density".ljust(24, b"\0")
    rc, info = abi.inspect_checkpoint(_write(tmp_path, bytes(blob)))
    assert rc == 2 and b"density" in info.error


def test_missing_file_is_rejected(tmp_path):
    rc, info = abi.inspect_checkpoint(tmp_path / "absent.ck")
    assert rc == 2 and info.error
