"""The `skycube` command (host C++ driver over the C ABI): the reference's
flags and output format (src/main.cu:66-177, src/main.h:152-356)."""
import gzip
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "skycube-templates_amd", "skycube")
NBA = os.path.join(ROOT, "tests", "golden", "nba-U-8-17264.csv.gz")


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.dirname(BIN)], check=True)


def run(args, **kw):
    return subprocess.run([BIN] + args, capture_output=True, text=True, timeout=300, **kw)


@pytest.fixture(scope="module")
def nba_csv(tmp_path_factory):
    p = tmp_path_factory.mktemp("csv") / "nba-U-8-17264.csv"
    with gzip.open(NBA, "rb") as f:
        p.write_bytes(f.read())
    return str(p)


def test_usage_and_bad_algorithm(nba_csv):
    assert os.path.exists(BIN)
    r = run([])
    assert r.returncode == 1 and "USAGE" in r.stdout
    r = run(["-f", nba_csv, "-s", "stsc"])
    assert r.returncode == 1 and "not provided" in r.stderr
    r = run(["-x"])
    assert r.returncode == 1


def test_missing_file():
    r = run(["-f", "/nonexistent.csv", "-s", "mdmc"])
    assert r.returncode == 2 and "No such file" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("algo,m", [("mdmc", None), ("mdmc", 3), ("sdsc", 4)])
def test_driver_runs_and_checks(nba_csv, algo, m):
    args = ["-f", nba_csv, "-s", algo, "-c", "-g", "0"] + (["-m", str(m)] if m else [])
    print("driver", " ".join(args), flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    # the check prints one line per cuboid mismatch between MDMC and SDSC
    body = out.split("checking\n", 1)[1].split("done checking\n", 1)
    assert body[0] == "", body[0][:500]
    assert re.fullmatch(r"\d+\t", body[1])
