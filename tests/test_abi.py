"""CPU checks of the C ABI: the product library loads and exports every
function include/skycube_amd.h declares (no device work without a GPU)."""
import ctypes
import os
import re

import pytest

import skycube_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "skycube_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sky_[a-z_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(skycube_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(skycube_amd.library_path())
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_words_per_point():
    L = skycube_amd.load()
    for d in range(1, 17):
        assert L.sky_words_per_point(d) == skycube_amd.words_per_point(d) == max(1, 2 ** d // 32)


def test_scope_mask():
    m = skycube_amd.in_scope_mask(6, 2)
    bits = [(int(m[s >> 5]) >> (s & 31)) & 1 for s in range(64)]
    assert bits == [int(1 <= s and bin(s).count("1") <= 2) for s in range(64)]


def test_product_does_not_load_the_oracle():
    """The product library must not depend on the oracle library."""
    import subprocess
    out = subprocess.run(["ldd", skycube_amd.library_path()], capture_output=True, text=True).stdout
    assert "skyorc" not in out
    src = open(os.path.join(ROOT, "skycube-templates_amd", "skycube_amd", "__init__.py")).read()
    assert "oracle" not in src.replace("no fallback", "")
