"""The allocation-free FP64 dot product (brtm_dot_noalloc) against
brtm_vec_dot, the entry point it varies: the same kernels on the library's
static scratch instead of context buffers, so the same bits."""
import ctypes as C
import sys

import numpy as np
import pytest

from sep_code_amd import _lib
from sep_code_amd._lib import call

pytestmark = pytest.mark.gpu


def test_dot_noalloc_matches_vec_dot():
    from sep_code_amd.device import BornDevice
    from sep_code_amd.operator import DeviceVector
    lib = _lib.load()
    r = np.random.default_rng(5)
    with BornDevice(0) as dev:
        # two calls in a row on one context, the second smaller; then a
        # second context on the same GPU shares the same scratch
        for n in (3_000_017, 70_001, 0):
            print(f"dot_noalloc n={n}", file=sys.stderr, flush=True)
            a = r.standard_normal(max(n, 1)).astype(np.float32)[:n]
            b = r.standard_normal(max(n, 1)).astype(np.float32)[:n]
            out = C.c_double(123.0)
            if n == 0:
                call("brtm_dot_noalloc", dev.h, None, None, 0, C.byref(out))
                assert out.value == 0.0
                continue
            va, vb = DeviceVector.from_array(dev, a), DeviceVector.from_array(dev, b)
            ref = va.dot(vb)   # brtm_vec_dot
            call("brtm_dot_noalloc", dev.h, va.data_ptr, vb.data_ptr, n, C.byref(out))
            assert out.value == ref, (out.value, ref)
            exact = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
            assert abs(out.value - exact) <= 1e-12 * float(np.abs(a * b.astype(np.float64)).sum())
            va.close()
            vb.close()
        # arguments it refuses
        assert lib.brtm_dot_noalloc(dev.h, None, None, -1, C.byref(out)) == -1
        assert lib.brtm_dot_noalloc(dev.h, None, None, 5, C.byref(out)) == -1
    with BornDevice(0) as dev2:
        a = r.standard_normal(1000).astype(np.float32)
        va = DeviceVector.from_array(dev2, a)
        out = C.c_double()
        call("brtm_dot_noalloc", dev2.h, va.data_ptr, va.data_ptr, 1000, C.byref(out))
        assert out.value == va.dot(va)
        va.close()
