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


def test_operators_sharing_a_context_keep_their_own_stores(dev):
    """Operators on one context share its per-shot source-wavefield stores.
    Destroying an older operator (explicitly, or late, by garbage
    collection) leaves the stores a newer one prepared, and an operator
    whose stores another one has replaced refuses to run until it prepares
    them again, instead of imaging with the other operator's fields."""
    print("lifetime: operators sharing a context", file=sys.stderr, flush=True)
    from sep_code_amd import configs
    from sep_code_amd.operator import RtmZeroOp
    from sep_code_amd._lib import BrtmError
    w = configs.lsrtm(seed=5, n_side=32, nt=40, nshots=(2, 1), rec_step=3, store="f32")
    dev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    dev.set_velocity(w.V)
    dev.set_imaging_scale(w.cimg)
    n1, n2, n3 = w.n
    nb = w.nbound
    mbox = (n1 - 2 * nb, n2 - 2 * nb, n3 - 2 * nb, nb, nb, nb)
    old = RtmZeroOp(dev, w, store="f32", model_box=mbox)
    old.prepare()
    new = RtmZeroOp(dev, w, store="f32", model_box=mbox)
    new.prepare()
    m = new.model_vector(np.random.default_rng(5).standard_normal(new.model_size))
    d1 = new.data_vector()
    new.forward(False, m, d1)
    ref = d1.numpy()
    assert np.abs(ref).max() > 0
    # the older operator's shots now hold the newer one's fields
    with pytest.raises(BrtmError) as e:
        old.forward(False, m, old.data_vector())
    assert e.value.code == BRTM_ESTATE and "another operator" in str(e.value)
    old.close()  # must not release the newer operator's stores
    d2 = new.data_vector()
    new.forward(False, m, d2)
    np.testing.assert_array_equal(d2.numpy(), ref)
    new.close()
