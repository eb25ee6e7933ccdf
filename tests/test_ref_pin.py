"""Pins the CPU oracle against the reference's own CPU skyline code.

oracle/_ref/libskyref.so is built from the reference's sources where they lie
(src/hybrid/mommyq.cpp: MommyQ, the per-cuboid skyline of its STSC, with the
DominateLeft_array tests of src/utilities/dt_avx2.h; src/utilities/
lattice_utilities.cpp) by `make -C oracle ref`, and walks the lattice as
HybridCube::Execute does (src/stsc/HybridCube.cpp:30-136).  Every cuboid's
sky and ext lists of the oracle's STSC walk and of its per-point closure
must equal the reference's, for d = 1..16, with ties, continuous values and
-m caps.  Skipped where the reference build is absent.
"""
import numpy as np
import pytest

import npref
from oracle import skyorc, skyref

pytestmark = pytest.mark.skipif(not skyref.available(),
                                reason="oracle/_ref not built (make -C oracle ref)")


def _oracle_lists(x, max_d):
    """Per cuboid (sky ids, ext-not-sky ids) from the oracle's STSC walk."""
    n, d = x.shape
    S = 1 << d
    dom, sdom, _, _ = skyorc.stsc(x, max_d)
    D = npref.unpack_bits(dom, S)
    SD = npref.unpack_bits(sdom, S)
    return D, SD


def _check(x, max_d):
    n, d = x.shape
    S = 1 << d
    so, sky, eo, ext, _ = skyref.stsc(x, max_d, threads=4)
    D, SD = _oracle_lists(x, max_d)
    scope = npref.scope(d, max_d)
    scope[S - 1] = True  # the full space is always computed
    for s in np.flatnonzero(scope):
        rs = sky[so[s]:so[s + 1]]
        re_ = ext[eo[s]:eo[s + 1]]
        assert np.array_equal(rs, np.flatnonzero(~D[:, s])), (d, max_d, s, "sky")
        assert np.array_equal(re_, np.flatnonzero(D[:, s] & ~SD[:, s])), (d, max_d, s, "ext")
    # the oracle's per-point closure (the algorithm the GPU results are
    # compared with) gives the same cuboid skylines
    e, mdom = skyorc.mdmc(x, max_d)
    M = np.ones((n, S), dtype=bool)
    M[e] = npref.unpack_bits(mdom, S)
    for s in np.flatnonzero(npref.scope(d, max_d)):
        assert np.array_equal(np.flatnonzero(~M[:, s]), sky[so[s]:so[s + 1]]), (d, max_d, s)
    # the full-space extended skyline is the closure's point set
    assert np.array_equal(e, np.sort(np.concatenate([sky[so[S - 1]:so[S]],
                                                     ext[eo[S - 1]:eo[S]]])))


@pytest.mark.parametrize("d", list(range(1, 17)))
def test_ties_match_reference(d):
    rng = np.random.default_rng(900 + d)
    n = 300 if d <= 8 else (120 if d <= 12 else 48)
    _check(npref.tie_heavy(rng, n, d, 3 if d > 8 else 4), d)


@pytest.mark.parametrize("d", [2, 5, 8, 10, 12, 14, 16])
def test_continuous_match_reference(d):
    rng = np.random.default_rng(950 + d)
    n = 400 if d <= 8 else (150 if d <= 12 else 60)
    _check(rng.random((n, d), dtype=np.float32), d)


@pytest.mark.parametrize("d,max_d", [(6, 2), (9, 4), (12, 3), (13, 6), (15, 5), (16, 8), (16, 3)])
def test_caps_match_reference(d, max_d):
    rng = np.random.default_rng(980 + d + max_d)
    n = 200 if d <= 12 else 80
    _check(rng.integers(0, 6, size=(n, d)).astype(np.float32), max_d)


def test_workload_samples_match_reference():
    """Small instances of the three plain workloads (more than 50 candidates
    per cuboid, so MommyQ runs for the cuboids too)."""
    from skycube_amd import datagen
    _check(datagen.workload("independent_distinct", 3000), 8)
    _check(datagen.workload("anticorrelated_plane", 600)[:, :9], 9)
    _check(datagen.workload("integer_ties", 700), 8)


def test_nba_matches_reference():
    """The reference's own dataset (workloads/nba-U-8-17264.csv)."""
    from golden.make_golden import NBA_GZ, read_csv_strtof
    _check(read_csv_strtof(NBA_GZ), 8)
