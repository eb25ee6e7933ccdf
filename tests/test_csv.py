"""The host driver's CSV reader (skycube-templates_amd/host/skycube.hpp:
read_dims / read_data, the reference's read_dims / read_data / split,
src/utilities/utilities.h:81-102,140-178,213-227) against an independent
reading with C strtof (correctly rounded, like the reference's
istringstream >> float), on the reference's own dataset and on variants of
its line format: CRLF endings, trailing commas, no final newline, spaces."""
import gzip
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NBA = os.path.join(HERE, "golden", "nba-U-8-17264.csv.gz")


@pytest.fixture(scope="module")
def reader(tmp_path_factory):
    exe = tmp_path_factory.mktemp("bin") / "csv_rows"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "helpers", "csv_rows.cpp")], check=True)
    return str(exe)


def _read(reader, path):
    r = subprocess.run([reader, path], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr
    head, _, body = r.stdout.partition(b"\n")
    n, d = (int(v) for v in head.split())
    return np.frombuffer(body, dtype="<f4").reshape(n, d)


def _nba_text():
    with gzip.open(NBA, "rb") as f:
        return f.read()


def _strtof_rows(text: bytes):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden import read_csv_strtof
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".csv", delete=False) as f:
        f.write(text)
        p = f.name
    try:
        return read_csv_strtof(p)
    finally:
        os.unlink(p)


@pytest.mark.parametrize("variant", ["plain", "crlf", "trailing_comma", "crlf_trailing_comma",
                                     "no_final_newline", "spaces"])
def test_reader_matches_strtof(reader, tmp_path, variant):
    text = _nba_text()
    lines = text.split(b"\n")
    if lines and lines[-1] == b"":
        lines = lines[:-1]
    base = [ln.rstrip(b"\r").rstrip(b",") for ln in lines]
    if variant == "crlf":
        out = b"\r\n".join(base) + b"\r\n"
    elif variant == "trailing_comma":
        out = b"\n".join(ln + b"," for ln in base) + b"\n"
    elif variant == "crlf_trailing_comma":
        out = b"\r\n".join(ln + b"," for ln in base) + b"\r\n"
    elif variant == "no_final_newline":
        out = b"\n".join(base)
    elif variant == "spaces":
        out = b"\n".join(b", ".join(ln.split(b",")) for ln in base) + b"\n"
    else:
        out = text
    p = tmp_path / f"{variant}.csv"
    p.write_bytes(out)
    got = _read(reader, str(p))
    want = _strtof_rows(b"\n".join(base) + b"\n")
    assert got.shape == (17264, 8) and want.shape == (17264, 8)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_reader_rounds_like_strtof(reader, tmp_path):
    """Decimal strings that are not exact binary fractions, halfway cases,
    exponents and signs: every value is the correctly rounded float32."""
    vals = ["0.1", "0.2", "0.30000001192092896", "1e-45", "3.4028234e38", "-0.0", "16777217",
            "16777219", "1.00000006", "2.5e-3", "-7.25", "123456789"]
    p = tmp_path / "round.csv"
    p.write_bytes((",".join(vals) + "\n" + ",".join(reversed(vals)) + "\n").encode())
    got = _read(reader, str(p))
    want = np.array([[np.float32(v) for v in vals], [np.float32(v) for v in reversed(vals)]],
                    dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_reader_rejects_ragged_rows(reader, tmp_path):
    p = tmp_path / "ragged.csv"
    p.write_bytes(b"1,2,3\n4,5\n")
    r = subprocess.run([reader, str(p)], capture_output=True, timeout=60)
    assert r.returncode == 2 and b"ragged" in r.stderr
