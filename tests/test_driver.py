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


def _read_result(path):
    """Parses the -w result file (see host/main.cpp, write_hashcube /
    write_lattice)."""
    import numpy as np
    a = np.fromfile(path, dtype="<u4")
    assert a[:1].tobytes() == b"SKYC"
    d, kind = int(a[1]), int(a[2])
    pos = 3
    if kind == 0:
        maps = []
        words = int(a[pos]); pos += 1
        for j in range(words):
            ng = int(a[pos]); pos += 1
            groups = {}
            for _ in range(ng):
                v, c = int(a[pos]), int(a[pos + 1]); pos += 2
                groups[v] = a[pos:pos + c].copy(); pos += c
            maps.append(groups)
        assert pos == a.size
        return d, kind, maps
    lat = {}
    for s in range(1, 1 << d):
        c = int(a[pos]); pos += 1
        sky = a[pos:pos + c].copy(); pos += c
        c = int(a[pos]); pos += 1
        ext = a[pos:pos + c].copy(); pos += c
        lat[s] = (sky, ext)
    assert pos == a.size
    return d, kind, lat


@pytest.mark.gpu
def test_driver_mdmc_result_is_golden(nba_csv, tmp_path):
    """The whole command-line path -- CSV parse, MDMC on the device, HashCube
    into the reference's map layout -- rebuilt into per-point bitmaps gives
    the golden digest of the NBA skycube (written by the oracle)."""
    import hashlib
    import json
    import numpy as np
    out = str(tmp_path / "nba.hc")
    print("driver -s mdmc -w", flush=True)
    r = run(["-f", nba_csv, "-s", "mdmc", "-w", out])
    assert r.returncode == 0, r.stderr
    d, kind, maps = _read_result(out)
    assert d == 8 and kind == 0 and len(maps) == 8 + 1
    # the extra always-zero word holds every extended-skyline point
    ext = np.sort(np.concatenate(list(maps[-1].values())))
    assert list(maps[-1]) == [0]
    pos = {int(p): i for i, p in enumerate(ext)}
    dom = np.full((ext.size, 8), 0xFFFFFFFF, dtype=np.uint32)
    for j, groups in enumerate(maps[:-1]):
        for v, ids in groups.items():
            assert v != 0xFFFFFFFF
            for p in ids:
                dom[pos[int(p)], j] = v
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))["nba"]
    h = hashlib.sha256(ext.astype("<u4").tobytes() + dom.astype("<u4").tobytes()).hexdigest()
    assert ext.size == gold["n_ext"] and h == gold["sha256"]


@pytest.mark.gpu
def test_driver_sdsc_result_is_golden(nba_csv, tmp_path):
    """The SDSC lattice written by the command: each cuboid's skyline equals
    the golden NBA bitmaps' members, and sky + ext of the full space is the
    golden extended skyline."""
    import numpy as np
    gold = np.load(os.path.join(ROOT, "tests", "golden", "nba_mdmc.npz"))
    out = str(tmp_path / "nba.lat")
    print("driver -s sdsc -w", flush=True)
    r = run(["-f", nba_csv, "-s", "sdsc", "-w", out])
    assert r.returncode == 0, r.stderr
    d, kind, lat = _read_result(out)
    assert d == 8 and kind == 1
    ext, dom = gold["ext"], gold["dom"]
    for s in range(1, 256):
        w, b = divmod(s, 32)
        want = ext[((dom[:, w] >> np.uint32(b)) & 1) == 0]
        assert np.array_equal(np.sort(lat[s][0]), np.sort(want)), s
    full = np.sort(np.concatenate(lat[255]))
    assert np.array_equal(full, ext)
