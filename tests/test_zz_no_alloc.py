"""GPU: serenx_diagnostics_noalloc, the allocation-free variant of
serenx_diagnostics, called through the C ABI, writes the same 20 FP64
diagnostics bit for bit (same kernels, same partial-sum order) into a
caller-owned device buffer.  Two calls in a row on one context, the second
after loading a smaller cloud, check that the shared scratch array carries
nothing from one call into the next."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud  # noqa: E402

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself links (already loaded with it)."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


def test_noalloc_diagnostics_equal_the_allocating_entry_point():
    from seren_amd.engine import ForceEngine, _check
    eng = ForceEngine(multipole="quadrupole", mac="geometric", precision="fp64",
                      thetamaxsqd=0.25, sound_speed=CS)
    hip = _hip()
    dbuf = C.c_void_p()
    assert hip.hipMalloc(C.byref(dbuf), 8 * 20) == 0
    try:
        results = []
        for n, prof, seed in ((6000, "r2", 51), (2100, "plummer", 52)):
            print(f"case no-alloc diagnostics: {prof} cloud of {n}", flush=True)
            pos, vel, mass, h = cloud(n, prof, seed)
            eng.load(pos, vel, mass, h)
            eng.force_eval(hguess=True)
            ref = np.zeros(20)
            _check(eng.lib.serenx_diagnostics(eng.ctx, ref.ctypes.data_as(C.POINTER(C.c_double))))
            assert hip.hipMemset(dbuf, 0xFF, 8 * 20) == 0      # NaN bytes before the call
            _check(eng.lib.serenx_diagnostics_noalloc(eng.ctx, dbuf))
            _check(eng.lib.serenx_sync(eng.ctx))
            got = np.zeros(20)
            assert hip.hipMemcpy(got.ctypes.data, dbuf, 8 * 20, 2) == 0   # device to host
            assert np.array_equal(got, ref), (n, got, ref)
            assert got[0] == pytest.approx(mass.sum(), rel=1e-12)
            results.append(got)
        assert results[0][0] != results[1][0]
    finally:
        hip.hipFree(dbuf)
        eng.close()


def test_noalloc_diagnostics_reject_bad_arguments():
    from seren_amd.engine import ForceEngine
    eng = ForceEngine(sound_speed=CS)
    try:
        assert eng.lib.serenx_diagnostics_noalloc(eng.ctx, None) != 0       # null output
        assert eng.lib.serenx_diagnostics_noalloc(eng.ctx, C.c_void_p(1)) != 0   # nothing loaded
    finally:
        eng.close()
