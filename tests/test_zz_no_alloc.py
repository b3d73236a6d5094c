"""bh_dataset_variance_noalloc through the C-ABI: its results equal
bh_dataset_variance's bit for bit on the same device inputs, on two calls in a
row (the second smaller, so it runs on what the first left in the shared
device array), and it refuses what it cannot take without allocating."""
import numpy as np
import pytest
import torch

from bayadera_amd import _lib, engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fac():
    f = engine.HipBayaderaFactory(0, 256)
    yield f
    f.release()


def _variance(fn, ctx, a, m, n, ld):
    out = torch.full((m,), float("nan"), dtype=torch.float32, device="cuda")
    _lib.call(fn, ctx, a.data_ptr(), m, n, ld, out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_variance_equals_variance_on_two_calls(fac):
    # (rows, columns, leading dimension): the largest row count the entry
    # point takes over a column count that fills every partial-sum block,
    # then a small strided block
    for m, n, ld, seed in ((256, 70001, 259, 1), (5, 1000, 7, 2)):
        print(f"noalloc variance: m={m} n={n} ld={ld} seed={seed}", flush=True)
        rng = np.random.default_rng(seed)
        host = (rng.standard_normal((n, ld)) * 3 + 100).astype(np.float32)
        a = torch.from_numpy(host).cuda()
        want = _variance("bh_dataset_variance", fac.handle, a, m, n, ld)
        got = _variance("bh_dataset_variance_noalloc", fac.handle, a, m, n, ld)
        assert np.array_equal(got, want)
        ref = host[:, :m].astype(np.float64).var(0)
        assert np.allclose(got, ref, rtol=1e-5, atol=0)


def test_noalloc_variance_refuses_what_needs_allocation(fac):
    a = torch.zeros((10, 257), dtype=torch.float32, device="cuda")
    out = torch.zeros(257, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.ArgumentError, match="at most 256"):
        _lib.call("bh_dataset_variance_noalloc", fac.handle, a.data_ptr(), 257, 10, 257,
                  out.data_ptr())
    host = np.zeros((10, 4), np.float32)
    with pytest.raises(_lib.ArgumentError, match="device memory"):
        _lib.call("bh_dataset_variance_noalloc", fac.handle, host.ctypes.data, 4, 10, 4,
                  out.data_ptr())
