"""Object lifetime across the C ABI: vectors, operators, solvers and networks
may be destroyed after the context they were created on."""
import ctypes as C
import sys

import numpy as np
import pytest

from sep_code_amd import _lib
from sep_code_amd import operator as _op  # noqa: F401  (registers the operator ABI)
from sep_code_amd import unet as _unet  # noqa: F401  (registers the network ABI)
from sep_code_amd._lib import call

pytestmark = pytest.mark.gpu

BRTM_ESTATE = -4


def test_children_destroyed_after_their_context():
    """brtm_destroy before brtm_vec_destroy / brtm_linop_destroy /
    brtm_unet_destroy: calls on the children in between fail cleanly with
    BRTM_ESTATE, and their destroy calls then free them (no use of freed
    memory: the context's host record outlives brtm_destroy)."""
    print("lifetime: children after context", file=sys.stderr, flush=True)
    _unet._reg()  # declares the vector, operator and network signatures
    lib = _lib.load()
    h = C.c_void_p()
    call("brtm_create", 0, C.byref(h))
    v, w = C.c_void_p(), C.c_void_p()
    call("brtm_vec_create", h, 4096, C.byref(v))
    call("brtm_vec_create", h, 4096, C.byref(w))
    a = np.arange(4096, dtype=np.float32)
    call("brtm_vec_upload", v, a.ctypes.data)
    call("brtm_vec_upload", w, a.ctypes.data)
    out = C.c_double()
    call("brtm_vec_dot", v, w, C.byref(out))
    assert out.value == float((a.astype(np.float64) ** 2).sum())
    ident = C.c_void_p()
    call("brtm_linop_identity", h, 4096, C.byref(ident))
    net = C.c_void_p()
    call("brtm_unet_create", h, 1, 64, C.c_ulonglong(3), C.byref(net))
    call("brtm_destroy", h)
    # the children report the destroyed context instead of touching it
    assert lib.brtm_vec_zero(v) == BRTM_ESTATE
    assert lib.brtm_vec_dot(v, w, C.byref(out)) == BRTM_ESTATE
    assert lib.brtm_linop_forward(ident, 0, v, w) == BRTM_ESTATE
    host = np.zeros(4096, np.float32)
    assert lib.brtm_vec_download(v, host.ctypes.data) == BRTM_ESTATE
    # destroying the context twice is refused while children hold it
    assert lib.brtm_destroy(h) == BRTM_ESTATE
    call("brtm_unet_destroy", net)
    call("brtm_linop_destroy", ident)
    call("brtm_vec_destroy", w)
    call("brtm_vec_destroy", v)  # the last holder: the context record goes here


def test_python_objects_outlive_their_device():
    """The Python mirror: a vector closed after its device no longer skips
    the C destroy call (nothing leaks), in either order."""
    print("lifetime: python objects", file=sys.stderr, flush=True)
    from sep_code_amd.device import BornDevice
    from sep_code_amd.operator import DeviceVector
    d = BornDevice(0)
    x = DeviceVector.from_array(d, np.ones(1000, np.float32))
    y = DeviceVector.from_array(d, np.ones(1000, np.float32))
    call("brtm_destroy", d.h)   # behind the mirror's back: x and y still open
    d.h = None
    x.close()
    y.close()
    assert x.h is None and y.h is None
    d2 = BornDevice(0)
    z = DeviceVector.from_array(d2, np.ones(10, np.float32))
    d2.close()                  # closes z first
    assert z.h is None
