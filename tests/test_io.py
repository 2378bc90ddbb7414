"""File entry points of the C-ABI (host only, no GPU): the reference's text
input format (main.cu:286-323), the binary CSR format, pair files and the
pair-set comparison that replaces compare() (main.cu:143-171)."""
import numpy as np
import pytest

import apss


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def rows_of(rp, col, val):
    return [dict(zip(col[rp[r]:rp[r + 1]].tolist(), val[rp[r]:rp[r + 1]].tolist()))
            for r in range(len(rp) - 1)]


def test_reference_layout(tmp_path):
    # main.cu reads one character, then "d v" pairs, each followed by one
    # separator character, until that character is the newline; a line with
    # just a newline is a row without features.
    text = " 3 0.5 1 0.25\n\n 7 1\n 0 0.125 2 0.75 9 0.5\n"
    rp, col, val, d = apss.load_text(write(tmp_path, "a.txt", text))
    assert d == 10
    assert rp.tolist() == [0, 2, 2, 3, 6]
    assert col.tolist() == [1, 3, 7, 0, 2, 9]          # ids sorted within a row
    assert val.tolist() == [0.25, 0.5, 1.0, 0.125, 0.75, 0.5]


def test_lenient_forms(tmp_path):
    # Tab separators, CRLF line ends, trailing blanks, a digit as the leading
    # separator (main.cu:301 consumes the first character of a line whatever it
    # is, so "15 0.1" is the row {5: 0.1}), and a last line without a newline.
    text = "\t2 0.5  4 0.5 \r\n15 0.1\n \r\n 6 2"
    rp, col, val, d = apss.load_text(write(tmp_path, "b.txt", text), n_cols=8)
    assert d == 8
    assert rows_of(rp, col, val) == [{2: 0.5, 4: 0.5}, {5: np.float32(0.1)}, {}, {6: 2.0}]


@pytest.mark.parametrize("text,what", [
    (" 3 0.5 3 0.25\n", "repeated"),
    (" 3 0.5 4\n", "value"),
    (" -1 0.5\n", "id"),
    (" 3,0.5\n", "id value"),
    (" 3 0.5x 4 1\n", "value"),
])
def test_text_errors(tmp_path, text, what):
    with pytest.raises(apss.ApssError) as e:
        apss.load_text(write(tmp_path, "c.txt", text))
    assert e.value.code == apss.APSS_E_INVALID and what in str(e.value)


def test_text_beyond_n_cols(tmp_path):
    with pytest.raises(apss.ApssError):
        apss.load_text(write(tmp_path, "d.txt", " 3 0.5 12 1\n"), n_cols=12)


def test_missing_file(tmp_path):
    with pytest.raises(apss.ApssError):
        apss.load_text(str(tmp_path / "nope.txt"))
    with pytest.raises(apss.ApssError):
        apss.load_binary(str(tmp_path / "nope.bin"))


@pytest.mark.parametrize("name", ["small", "large"])
def test_round_trips(tmp_path, name):
    rp, col, val, d = apss.workload(name, n=2000)
    apss.save_text(str(tmp_path / "m.txt"), rp, col, val, d)
    r2, c2, v2, d2 = apss.load_text(str(tmp_path / "m.txt"), n_cols=d)
    assert d2 == d and np.array_equal(r2, rp) and np.array_equal(c2, col)
    assert np.array_equal(v2.view(np.uint32), val.view(np.uint32))   # %.9g round-trips FP32
    apss.save_binary(str(tmp_path / "m.bin"), rp, col, val, d)
    r3, c3, v3, d3 = apss.load_binary(str(tmp_path / "m.bin"))
    assert d3 == d and np.array_equal(r3, rp) and np.array_equal(c3, col)
    assert np.array_equal(v3.view(np.uint32), val.view(np.uint32))


def test_empty_matrix(tmp_path):
    p = write(tmp_path, "e.txt", "")
    rp, col, val, d = apss.load_text(p)
    assert rp.tolist() == [0] and len(col) == 0 and d == 0
    apss.save_binary(str(tmp_path / "e.bin"), rp, col, val, 5)
    r2, c2, v2, d2 = apss.load_binary(str(tmp_path / "e.bin"))
    assert r2.tolist() == [0] and len(c2) == 0 and d2 == 5


def test_binary_rejects_bad_files(tmp_path):
    rp = np.array([0, 2, 3], np.int64)
    col = np.array([1, 4, 2], np.int32)
    val = np.array([0.6, 0.8, 1.0], np.float32)
    good = tmp_path / "g.bin"
    apss.save_binary(str(good), rp, col, val, 5)
    blob = good.read_bytes()
    cases = {
        "magic": b"XPSSCSR1" + blob[8:],
        "truncated": blob[:-4],
        "trailing": blob + b"\0",
        "header": blob[:8] + np.int64(-1).tobytes() + blob[16:],
    }
    # Column ids not increasing in row 0 (swap 1 and 4).
    off = 8 + 24 + 8 * 3
    cases["order"] = blob[:off] + np.array([4, 1, 2], np.int32).tobytes() + blob[off + 12:]
    # Column id beyond n_cols.
    cases["range"] = blob[:off] + np.array([1, 5, 2], np.int32).tobytes() + blob[off + 12:]
    # row_ptr decreasing.
    cases["row_ptr"] = blob[:32] + np.array([0, 4, 3], np.int64).tobytes() + blob[56:]
    for name, data in cases.items():
        p = tmp_path / f"{name}.bin"
        p.write_bytes(data)
        with pytest.raises(apss.ApssError, match="apss error"):
            apss.load_binary(str(p))
    # Saving checks the structure too.
    with pytest.raises(apss.ApssError):
        apss.save_binary(str(tmp_path / "x.bin"), rp, np.array([4, 1, 2], np.int32), val, 5)


def test_pair_files(tmp_path):
    pairs = np.zeros(4, apss.PAIR_DTYPE)
    pairs["i"] = [0, 0, 3, 7]
    pairs["j"] = [1, 5, 4, 9]
    pairs["score"] = [0.5, 0.75, 0.9, 1.0]
    a, b, c = (str(tmp_path / n) for n in ("a.txt", "b.txt", "c.txt"))
    apss.write_pairs(a, pairs)
    lines = open(a).read().splitlines()
    assert lines[0] == "0 1 0.5" and len(lines) == 4
    # Same set in another order and with (j, i) swapped.
    sw = pairs[::-1].copy()
    sw["i"], sw["j"] = pairs[::-1]["j"], pairs[::-1]["i"]
    apss.write_pairs(b, sw)
    assert apss.compare_pair_files(a, b) == (0, 0)
    # One pair dropped, one added.
    other = pairs.copy()
    other["j"][2] = 8
    apss.write_pairs(c, other[1:])
    assert apss.compare_pair_files(a, c) == (2, 1)
    apss.write_pairs(c, pairs[:0])
    assert apss.compare_pair_files(a, c) == (4, 0)
