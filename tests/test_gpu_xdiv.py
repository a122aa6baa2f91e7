"""The shared-reciprocal division the FP64 kernels use (xdiv in
mohid_amd/csrc/face.h) against the IEEE division, bit for bit, on the GPU.

The kernels divide several numerators by one denominator -- a cell volume
(the dt/V terms of AD:3993-4001, 4560-4576, 5185-5351), a stencil spacing
(MF:10642-10699) or a Thomas pivot (MF:4090-4099) -- and reuse the
denominator's refined reciprocal for all of them.  The parity tests show the
transport stays bit-identical to the oracle; this checks the identity itself
on 2^26 seeded pairs, random and adversarial (all-ones significands, exponents
at the ends of the safe range, zeros, subnormals, infinities, NaNs).
"""
import ctypes as ct

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 20260613])
def test_xdiv_matches_ieee_division(seed):
    from mohid_amd import _lib as L
    lib = L.lib()
    n = 1 << 26
    bad, fast = ct.c_uint64(0), ct.c_uint64(0)
    L.check("mgpu_selfcheck_xdiv", lib.mgpu_selfcheck_xdiv(n, seed, ct.byref(bad),
                                                              ct.byref(fast)))
    print(f"xdiv seed {seed}: {n} pairs, {fast.value} on the shared-reciprocal path, "
          f"{bad.value} differ from a / b")
    assert bad.value == 0
    # most pairs exercise the fast path, and the rest the fallback
    assert 0.5 * n < fast.value < n
