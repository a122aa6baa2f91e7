"""mgpu_thomasz_noalloc (THOMASZ_NewType2 with its scratch in a static device
array) against mgpu_thomasz in device mode, through the C ABI.

Two calls in a row on device arrays, the second on a smaller Size, each
bit-identical to mgpu_thomasz on the same inputs; a Size beyond the scratch's
capacity is refused with MGPU_ERR_BADARG before any work is enqueued.
"""
import ctypes as ct

import numpy as np
import pytest

from mohid_amd import _lib as L

pytestmark = pytest.mark.gpu


def _system(rng, ni, nj, nk):
    shape = (ni + 2, nj + 2, nk + 2)
    D = -rng.uniform(0, 1, shape)
    F = -rng.uniform(0, 1, shape)
    E = 1.0 + rng.uniform(0, 1, shape) - D - F
    TI = rng.normal(0, 3, shape)
    WP = (rng.uniform(size=shape) < 0.8).astype(np.int32)
    res = rng.normal(size=shape)
    return shape, [np.asfortranarray(a) for a in (D, E, F, TI)], np.asfortranarray(WP), \
        np.asfortranarray(res)


def _dev(torch, a):
    # Fortran order on the host -> the same bytes on the device
    flat = np.ravel(a, order="F")
    return torch.from_numpy(flat.copy()).to("cuda")


def _solve(torch, fn_noalloc, size, ws, arrays, wp, res0):
    """Runs one entry point on device copies; returns res as a host array."""
    lib = L.lib()
    dD, dE, dF, dT = (_dev(torch, a) for a in arrays)
    dW = _dev(torch, wp)
    dR = _dev(torch, res0)
    s, w = L.Size3D(*size), L.Size3D(*ws)
    ptrs = [ct.c_void_p(t.data_ptr()) for t in (dD, dE, dF, dT, dW, dR)]
    stream = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    if fn_noalloc:
        rc = lib.mgpu_thomasz_noalloc(ct.byref(s), ct.byref(w), *ptrs, stream)
    else:
        rc = lib.mgpu_thomasz(ct.byref(s), ct.byref(w), *ptrs, 1, stream)
    L.check("thomasz_noalloc" if fn_noalloc else "thomasz", rc)
    torch.cuda.synchronize()
    return dR.cpu().numpy()


def test_noalloc_thomasz_matches_thomasz_twice_then_smaller():
    import torch
    torch.cuda.set_device(0)
    rng = np.random.default_rng(11)
    for ni, nj, nk in ((61, 47, 33), (19, 23, 9)):
        size = (0, ni + 1, 0, nj + 1, 0, nk + 1)
        ws = (1, ni, 1, nj, 1, nk)
        print(f"thomasz_noalloc: WorkSize {ni} x {nj} x {nk}", flush=True)
        _, arrays, wp, res0 = _system(rng, ni, nj, nk)
        got = _solve(torch, True, size, ws, arrays, wp, res0)
        want = _solve(torch, False, size, ws, arrays, wp, res0)
        assert np.array_equal(got, want), (ni, nj, nk)
        # the solved cells really were solved
        assert not np.array_equal(got, np.ravel(res0, order="F"))


def test_noalloc_thomasz_rejects_sizes_beyond_its_scratch():
    import torch
    torch.cuda.set_device(0)
    lib = L.lib()
    n = 256
    nk = L.MGPU_NOALLOC_MAX_CELLS // (n * n)        # (n+2)^2 (nk+2) cells exceed the cap
    s = L.Size3D(0, n + 1, 0, n + 1, 0, nk + 1)
    w = L.Size3D(1, n, 1, n, 1, nk)
    assert (n + 2) * (n + 2) * (nk + 2) > L.MGPU_NOALLOC_MAX_CELLS
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = ct.c_void_p(dummy.data_ptr())
    rc = lib.mgpu_thomasz_noalloc(ct.byref(s), ct.byref(w), p, p, p, p, p, p, None)
    assert rc == L.MGPU_ERR_BADARG
    assert b"scratch" in lib.mgpu_last_error()
