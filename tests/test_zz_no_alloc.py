"""sky_mdmc_device_noalloc (the MDMC for callers that cannot allocate device
memory) against sky_mdmc_device, through the C ABI: identical ids and
bitmaps, bit for bit, on calls in a row with shrinking inputs (the scratch
array keeps the last call's contents), with no device memory taken by the
context, and a clear error past the size limit."""
import numpy as np
import pytest
import torch

import skycube_amd
from skycube_amd import datagen

pytestmark = pytest.mark.gpu


def _noalloc(c, x, max_d):
    n, d = x.shape
    W = skycube_amd.words_per_point(d)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to("cuda")
    ext = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    dom = torch.full((max(n * W, 1),), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    before = c.bytes_in_use()
    ne = c.mdmc_device_noalloc(xd.data_ptr(), n, d, max_d, ext.data_ptr(), dom.data_ptr(), n * W)
    assert c.bytes_in_use() == before  # nothing was allocated on the context
    e = ext[:ne].cpu().numpy().view(np.uint32).copy()
    dm = dom[: ne * W].cpu().numpy().view(np.uint32).reshape(ne, W).copy()
    return e, dm


def _existing(x, max_d):
    n, d = x.shape
    xd = torch.from_numpy(np.ascontiguousarray(x)).to("cuda")
    torch.cuda.synchronize()
    with skycube_amd.Context(0) as c:
        ne = c.mdmc_device(xd.data_ptr(), n, d, max_d)
        return c.mdmc_fetch(ne, d)


CASES = [
    ("integer_ties d=16, 150k points", lambda: datagen.workload("integer_ties", 150_000), 8),
    ("anticorrelated d=12, 20k points", lambda: datagen.workload("anticorrelated_plane", 20_000), 12),
    ("independent d=8, 3k points", lambda: datagen.workload("independent_distinct", 3000), 8),
    ("ties d=6, 500 points", lambda: np.random.default_rng(3).integers(0, 4, (500, 6)).astype(np.float32), 3),
]


def test_noalloc_matches_existing_entry_point():
    with skycube_amd.Context(0) as c:
        for tag, make, max_d in CASES:  # each call smaller than the one before
            print(f"noalloc: {tag}", flush=True)
            x = make()
            got = _noalloc(c, x, max_d)
            want = _existing(x, max_d)
            assert np.array_equal(got[0], want[0]), tag
            assert np.array_equal(got[1], want[1]), tag
        # the context's own results are not disturbed by, nor kept after, the call
        with pytest.raises(skycube_amd.SkyError) as ei:
            c.mdmc_fetch(1, 6)
        assert ei.value.code == -6


def test_noalloc_rejects_sizes_past_the_limit():
    with skycube_amd.Context(0) as c:
        d, n = 16, 250_000
        x = torch.zeros((n, d), dtype=torch.float32, device="cuda")
        out = torch.zeros((n * skycube_amd.words_per_point(d),), dtype=torch.int32, device="cuda")
        ids = torch.zeros((n,), dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        with pytest.raises(skycube_amd.SkyError) as ei:
            c.mdmc_device_noalloc(x.data_ptr(), n, d, 8, ids.data_ptr(), out.data_ptr(), out.numel())
        assert ei.value.code == -5
        # an output buffer too small is rejected too
        with pytest.raises(skycube_amd.SkyError) as ei:
            c.mdmc_device_noalloc(x.data_ptr(), 1000, d, 8, ids.data_ptr(), out.data_ptr(), 10)
        assert ei.value.code == -5
