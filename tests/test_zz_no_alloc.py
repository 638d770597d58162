"""atm_twostream_noalloc (the general two-stream columns with the library's
static scratch instead of an allocated status buffer) against
atm_twostream: the same net flux, flux components and diagnostics, bit for
bit, on two calls in a row with the second smaller (the scratch keeps the
first call's contents); the column limit is refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(cols, prec):
    from atmos.twostream import to_device, twostream
    d = to_device(cols)
    a, fa, sa = twostream(d, prec, fluxes=True)
    b, fb, sb = twostream(d, prec, fluxes=True, noalloc=True)
    return (a.cpu().numpy(), fa.cpu().numpy(), sa), (b.cpu().numpy(), fb.cpu().numpy(), sb)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_matches_twostream(gpu_available, prec):
    from atmos.twostream import synthetic_columns
    print(f"noalloc vs twostream {prec}", flush=True)
    for ncol, nlay, seed in ((9000, 64, 1), (257, 40, 2)):  # the second call smaller
        cols = synthetic_columns(ncol, nlay=nlay, n_sw=6, n_lw=5, seed=seed)
        (a, fa, sa), (b, fb, sb) = _both(cols, prec)
        assert np.array_equal(a, b) and np.array_equal(fa, fb)
        assert sa == sb and sa["bad_columns"] == 0


def test_noalloc_async_and_limits(gpu_available):
    import torch
    from atmos.twostream import _lib, _struct, stream, synchronize, synthetic_columns, to_device, twostream
    print("noalloc async and limits", flush=True)
    d = to_device(synthetic_columns(300, nlay=16, n_sw=3, n_lw=3, seed=3))
    ref, _, _ = twostream(d, "fp64")
    out, _, st = twostream(d, "fp64", noalloc=True, stats=False, wait=False)
    synchronize()
    assert st is None and torch.equal(out, ref)
    s = _struct(d, 1e-4)
    s.ncol = (1 << 20) + 1  # refused on the host, before any launch
    net = torch.empty((300, 17), dtype=torch.float64, device="cuda")
    L = _lib()
    assert L.atm_twostream_noalloc(C.byref(s), 0, net.data_ptr(), None, None, C.c_void_p(stream())) == 1
    assert "exceeds the scratch" in L.atm_twostream_last_error().decode()
