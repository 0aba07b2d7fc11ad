"""Text edge lists: scd_edges_read_text follows the reference convert's
reader (src/convert.cpp:74-136).  Each crafted file is checked against the
edges its rules give (written out by hand below) and, where the reference's
own ./convert is built (oracle/_ref/convert), against that program: the
same output file bytes after conversion, or the same error."""
import os
import subprocess

import numpy as np
import pytest

import scd_amd
from scd_amd import _lib

REF_CONVERT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "convert")

LONG_OK = "2 3 " + "x" * (1023 - 4)          # exactly 1023 characters: read whole
LONG_BAD = "3 4 " + "9" * (1024 - 4)         # 1024 characters: the reader's buffer overflows

# (name, file bytes, expected edges or the expected error)
CASES = [
    ("plain", b"0 1\n1 2\n2 0\n", [(0, 1), (1, 2), (2, 0)]),
    ("comments_blank_ws", b"# header\n\n   \n  1 2\n2\t3 trailing words\n#x 9 9\n3 1\r\n\x0b\x0c4 5",
     [(1, 2), (2, 3), (3, 1), (4, 5)]),
    ("no_newline_at_end", b"5 6\n6 7\n7 5", [(5, 6), (6, 7), (7, 5)]),
    ("wrapping_ids", b"4294967297 5\n5 6\n6 4294967297\n", [(1, 5), (5, 6), (6, 1)]),
    ("loops_dups", b"0 0\n1 2\n2 1\n1 2\n2 3\n3 1\n3 3\n", [(0, 0), (1, 2), (2, 1), (1, 2), (2, 3), (3, 1), (3, 3)]),
    ("nul_line", b"1 2\n\x003 4\n2 3\n3 1\n", [(1, 2), (2, 3), (3, 1)]),
    ("long_line_ok", (LONG_OK + "\n3 1\n1 2\n").encode(), [(2, 3), (3, 1), (1, 2)]),
    ("empty", b"", []),
    ("corrupt_first_number", b"1 2\nx 3\n", ("corrupt", 2)),
    ("corrupt_second_number", b"1 2\n\n# c\n4 y\n", ("corrupt", 4)),
    ("corrupt_missing_second", b"1 2\n3 \n", ("corrupt", 2)),
    ("corrupt_sign", b"1 2\n2 3\n-1 2\n", ("corrupt", 3)),
    ("long_line_bad", ("1 2\n" + LONG_BAD + "\n2 3\n").encode(), ("read", 2)),
]


def _write(tmp_path, name, data):
    p = tmp_path / f"{name}.txt"
    p.write_bytes(data)
    return p


@pytest.mark.parametrize("name,data,expect", CASES, ids=[c[0] for c in CASES])
def test_reader_rules(tmp_path, name, data, expect):
    p = _write(tmp_path, name, data)
    if isinstance(expect, tuple):
        kind, line = expect
        with pytest.raises(scd_amd.ScdError) as ei:
            scd_amd.Edges.read_text(p)
        if kind == "corrupt":
            assert ei.value.code == _lib.SCD_EINVAL
            assert ei.value.message == f"line {line} is corrupt"
        else:
            assert ei.value.code == _lib.SCD_EIO
            assert ei.value.message.startswith("an error occured while reading input file")
        return
    e = scd_amd.Edges.read_text(p)
    assert list(zip(e.u.tolist(), e.v.tolist())) == expect


@pytest.mark.skipif(not os.path.exists(REF_CONVERT), reason="the reference's ./convert is not built here")
@pytest.mark.parametrize("name,data,expect", CASES, ids=[c[0] for c in CASES])
def test_reader_matches_reference_convert(tmp_path, name, data, expect):
    p = _write(tmp_path, name, data)
    out = tmp_path / "ref.bin"
    r = subprocess.run([REF_CONVERT, str(p), str(out)], capture_output=True, text=True, timeout=60,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
    if isinstance(expect, tuple):
        assert r.returncode != 0
        kind, line = expect
        msg = f"line {line} is corrupt" if kind == "corrupt" else "an error occured while reading input file"
        assert msg in r.stderr
        return
    assert r.returncode == 0, r.stderr
    g = scd_amd.Edges.read_text(p).convert()
    assert g.to_bytes() == out.read_bytes()


def test_round_trip_large(tmp_path):
    """An R-MAT edge list written and read back (multi-chunk parallel parse)."""
    e = scd_amd.rmat(16, 16, 0.57, 0.19, 0.19, 3)
    p = tmp_path / "rmat16.txt"
    e.write_text(p)
    r = scd_amd.Edges.read_text(p)
    assert np.array_equal(r.u, e.u) and np.array_equal(r.v, e.v)


def test_corrupt_line_number_across_chunks(tmp_path):
    """The first corrupt line is reported by its number even when the parse
    runs in parallel chunks and a later chunk also holds a corrupt line."""
    lines = [f"{i} {i + 1}" for i in range(300000)]
    lines[123456] = "12 z"
    lines[250000] = "q"
    p = tmp_path / "bad.txt"
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(scd_amd.ScdError) as ei:
        scd_amd.Edges.read_text(p)
    assert ei.value.message == "line 123457 is corrupt"
